// Training of the proposal MLP (dmk_proposal.hpp) on the GPU: maximum
// likelihood of displacements under q(delta | phi), weighted per sample
// (accepted moves of a chain weigh 1, rejected 0), Adam on minibatches.
// FP32 forward and backward; gradients per minibatch are reduced from
// per-block partial sums (no atomics on the parameters).
#include <hip/hip_runtime.h>

#include "dmk_proposal.hpp"
#include "mlp_kernels.hpp"

namespace hpdmk_amd {

namespace {

constexpr int MTB = 128;  // samples per block
constexpr int kVec = kFeat + 4 * kHid + kOut;  // per-sample vectors kept in shared memory

struct Offs {
    static constexpr int W1 = 0, B1 = W1 + kFeat * kHid, W2 = B1 + kHid, B2 = W2 + kHid * kHid, W3 = B2 + kHid,
                         B3 = W3 + kHid * kOut;
};

// forward pass of one sample (FP32); h1, h2, o written to the given arrays
__device__ __forceinline__ void forward(const float* w, const float* phi, float* h1, float* h2, float* o) {
    for (int h = 0; h < kHid; ++h) {
        float s = w[Offs::B1 + h];
        for (int f = 0; f < kFeat; ++f) s += w[Offs::W1 + f * kHid + h] * phi[f];
        h1[h] = tanhf(s);
    }
    for (int h = 0; h < kHid; ++h) {
        float s = w[Offs::B2 + h];
        for (int i = 0; i < kHid; ++i) s += w[Offs::W2 + i * kHid + h] * h1[i];
        h2[h] = tanhf(s);
    }
    for (int k = 0; k < kOut; ++k) {
        float s = w[Offs::B3 + k];
        for (int h = 0; h < kHid; ++h) s += w[Offs::W3 + h * kOut + k] * h2[h];
        o[k] = s;
    }
}

// per-block gradient partials of sum_s wt_s NLL_s over the samples sel[b0 .. b0 + n)
__global__ void __launch_bounds__(MTB) k_mlp_grad(MlpData d, const int* sel, int b0, int n, const float* w,
                                                  float* part, float* loss_part) {
    __shared__ float V[MTB][kVec];  // phi | h1 | g1 | h2 | g2 | do
    __shared__ float lred[MTB / 64];
    const int t = threadIdx.x;
    const int k = blockIdx.x * MTB + t;
    const bool on = k < n;
    float* v = V[t];
    float* phi = v;
    float* h1 = v + kFeat;
    float* g1 = h1 + kHid;
    float* h2 = g1 + kHid;
    float* g2 = h2 + kHid;
    float* dout = g2 + kHid;
    float L = 0.0f;
    if (on) {
        const int s = sel ? sel[b0 + k] : b0 + k;
        const float wt = d.wt[s];
        for (int f = 0; f < kFeat; ++f) phi[f] = d.X[(size_t)kFeat * s + f];
        float o[kOut];
        forward(w, phi, h1, h2, o);
        for (int a = 0; a < 3; ++a) {
            const float th = tanhf(o[a]);
            const float mu = d.max_mean * th;
            const float sg = 1.0f / (1.0f + expf(-o[3 + a]));
            const float sig = d.smin + (d.smax - d.smin) * sg;
            const float y = (d.D[3 * (size_t)s + a] - mu) / sig;
            L += wt * (logf(sig) + 0.5f * y * y);
            const float dmu = -wt * y / sig;
            const float dsig = wt * (1.0f - y * y) / sig;
            dout[a] = dmu * d.max_mean * (1.0f - th * th);
            dout[3 + a] = dsig * (d.smax - d.smin) * sg * (1.0f - sg);
        }
        for (int h = 0; h < kHid; ++h) {
            float s2 = 0.0f;
            for (int q = 0; q < kOut; ++q) s2 += w[Offs::W3 + h * kOut + q] * dout[q];
            g2[h] = s2 * (1.0f - h2[h] * h2[h]);
        }
        for (int i = 0; i < kHid; ++i) {
            float s1 = 0.0f;
            for (int h = 0; h < kHid; ++h) s1 += w[Offs::W2 + i * kHid + h] * g2[h];
            g1[i] = s1 * (1.0f - h1[i] * h1[i]);
        }
    } else {
        for (int c = 0; c < kVec; ++c) v[c] = 0.0f;
    }
    __syncthreads();
    // parameter gradients: sums over the block's samples of outer products
    float* out = part + (size_t)blockIdx.x * kMlpParams;
    for (int p = t; p < kMlpParams; p += MTB) {
        float acc = 0.0f;
        if (p < Offs::B1) {
            const int f = p / kHid, h = p % kHid;
            for (int s = 0; s < MTB; ++s) acc += V[s][f] * V[s][kFeat + kHid + h];
        } else if (p < Offs::W2) {
            const int h = p - Offs::B1;
            for (int s = 0; s < MTB; ++s) acc += V[s][kFeat + kHid + h];
        } else if (p < Offs::B2) {
            const int i = (p - Offs::W2) / kHid, h = (p - Offs::W2) % kHid;
            for (int s = 0; s < MTB; ++s) acc += V[s][kFeat + i] * V[s][kFeat + 3 * kHid + h];
        } else if (p < Offs::W3) {
            const int h = p - Offs::B2;
            for (int s = 0; s < MTB; ++s) acc += V[s][kFeat + 3 * kHid + h];
        } else if (p < Offs::B3) {
            const int h = (p - Offs::W3) / kOut, q = (p - Offs::W3) % kOut;
            for (int s = 0; s < MTB; ++s) acc += V[s][kFeat + 2 * kHid + h] * V[s][kFeat + 4 * kHid + q];
        } else {
            const int q = p - Offs::B3;
            for (int s = 0; s < MTB; ++s) acc += V[s][kFeat + 4 * kHid + q];
        }
        out[p] = acc;
    }
    for (int o2 = 32; o2 > 0; o2 >>= 1) L += __shfl_xor(L, o2);
    if ((t & 63) == 0) lred[t >> 6] = L;
    __syncthreads();
    if (t == 0) loss_part[blockIdx.x] = lred[0] + lred[1];
}

// sum of the block partials (scaled by 1/wsum) into grad, then an Adam step
__global__ void k_mlp_adam(float* w, float* m, float* v, const float* part, int nblk, float* grad, float scale,
                           AdamStep st) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float g = 0.0f;
    for (int b = 0; b < nblk; ++b) g += part[(size_t)b * kMlpParams + p];
    g *= scale;
    if (grad) grad[p] = g;
    if (st.lr <= 0.0f) return;
    m[p] = st.b1 * m[p] + (1.0f - st.b1) * g;
    v[p] = st.b2 * v[p] + (1.0f - st.b2) * g * g;
    const float mh = m[p] / st.c1, vh = v[p] / st.c2;
    w[p] -= st.lr * mh / (sqrtf(vh) + st.eps);
}

// outputs mu, sigma per sample (FP32 forward), for checks
__global__ void k_mlp_out(MlpData d, int n, const float* w, float* out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float phi[kFeat], h1[kHid], h2[kHid], o[kOut];
    for (int f = 0; f < kFeat; ++f) phi[f] = d.X[(size_t)kFeat * s + f];
    forward(w, phi, h1, h2, o);
    for (int a = 0; a < 3; ++a) {
        out[6 * (size_t)s + a] = d.max_mean * tanhf(o[a]);
        out[6 * (size_t)s + 3 + a] = d.smin + (d.smax - d.smin) / (1.0f + expf(-o[3 + a]));
    }
}

}  // namespace

int mlp_blocks(int n) { return (n + MTB - 1) / MTB; }

void launch_mlp_grad(const MlpData& d, const int* sel, int b0, int n, const float* w, float* part, float* loss_part,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(mlp_blocks(n)), dim3(MTB), 0, st, d, sel, b0, n, w, part, loss_part);
}

void launch_mlp_adam(float* w, float* m, float* v, const float* part, int nblk, float* grad, float scale,
                     const AdamStep& s, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, w, m, v, part, nblk, grad, scale,
                       s);
}

void launch_mlp_out(const MlpData& d, int n, const float* w, float* out, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_out, dim3((n + 127) / 128), dim3(128), 0, st, d, n, w, out);
}

}  // namespace hpdmk_amd
