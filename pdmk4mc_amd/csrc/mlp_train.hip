// Training of the proposal MLP (dmk_proposal.hpp) on the GPU: the weighted
// Gaussian NLL of displacements under q(delta | phi) (weights 1/0 for the
// likelihood of accepted moves; signed rewards minus a baseline for the
// policy gradient of the accepted squared jump), Adam on minibatches.  FP32
// forward and backward; gradients per minibatch are reduced from per-block
// partial sums (no atomics on the parameters).  The deployed BF16 inference
// is k_mlp_infer.
#include <hip/hip_runtime.h>

#include "dmk_proposal.hpp"
#include "mlp_kernels.hpp"

namespace hpdmk_amd {

namespace {

constexpr int MTB = 64;                            // samples per block
constexpr int kVec = kFeat + 6 * kHid + kOut;      // per-sample vectors in shared memory
// phi | h1 | g1 | h2 | g2 | h3 | g3 | dout
constexpr int vH1 = kFeat, vG1 = vH1 + kHid, vH2 = vG1 + kHid, vG2 = vH2 + kHid, vH3 = vG2 + kHid, vG3 = vH3 + kHid,
              vDO = vG3 + kHid;

// FP32 forward pass of one sample into its shared vector v (phi set); o out
__device__ __forceinline__ void forward(const float* w, float* v, float* o) {
    const float* phi = v;
    for (int h = 0; h < kHid; ++h) {
        float s = w[kB1 + h];
        for (int f = 0; f < kFeat; ++f) s += w[kW1 + f * kHid + h] * phi[f];
        v[vH1 + h] = tanhf(s);
    }
    for (int h = 0; h < kHid; ++h) {
        float s = w[kB2 + h];
        for (int i = 0; i < kHid; ++i) s += w[kW2 + i * kHid + h] * v[vH1 + i];
        v[vH2 + h] = tanhf(s);
    }
    for (int h = 0; h < kHid; ++h) {
        float s = w[kB3 + h];
        for (int i = 0; i < kHid; ++i) s += w[kW3 + i * kHid + h] * v[vH2 + i];
        v[vH3 + h] = tanhf(s);
    }
    for (int k = 0; k < kOut; ++k) {
        float s = w[kB4 + k];
        for (int h = 0; h < kHid; ++h) s += w[kW4 + h * kOut + k] * v[vH3 + h];
        o[k] = s;
    }
}

// per-block gradient partials of sum_s wt_s NLL_s over the samples sel[b0 .. b0 + n)
// (dynamic shared memory: MTB x kVec floats)
__global__ void __launch_bounds__(MTB) k_mlp_grad(MlpData d, const int* sel, int b0, int n, const float* w,
                                                  float* part, float* loss_part) {
    extern __shared__ float Vs[];
    const int t = threadIdx.x;
    const int k = blockIdx.x * MTB + t;
    const bool on = k < n;
    float* v = Vs + (size_t)t * kVec;
    float L = 0.0f;
    if (on) {
        const int s = sel ? sel[b0 + k] : b0 + k;
        const float wt = d.wt[s];
        for (int f = 0; f < kFeat; ++f) v[f] = d.X[(size_t)kFeat * s + f];
        float o[kOut];
        forward(w, v, o);
        for (int a = 0; a < 3; ++a) {
            const float th = tanhf(o[a]);
            const float mu = d.max_mean * th;
            const float sg = 1.0f / (1.0f + expf(-o[3 + a]));
            const float sig = d.smin + (d.smax - d.smin) * sg;
            const float y = (d.D[3 * (size_t)s + a] - mu) / sig;
            L += wt * (logf(sig) + 0.5f * y * y);
            const float dmu = -wt * y / sig;
            const float dsig = wt * (1.0f - y * y) / sig;
            v[vDO + a] = dmu * d.max_mean * (1.0f - th * th);
            v[vDO + 3 + a] = dsig * (d.smax - d.smin) * sg * (1.0f - sg);
        }
        for (int h = 0; h < kHid; ++h) {
            float s3 = 0.0f;
            for (int q = 0; q < kOut; ++q) s3 += w[kW4 + h * kOut + q] * v[vDO + q];
            v[vG3 + h] = s3 * (1.0f - v[vH3 + h] * v[vH3 + h]);
        }
        for (int i = 0; i < kHid; ++i) {
            float s2 = 0.0f;
            for (int h = 0; h < kHid; ++h) s2 += w[kW3 + i * kHid + h] * v[vG3 + h];
            v[vG2 + i] = s2 * (1.0f - v[vH2 + i] * v[vH2 + i]);
        }
        for (int i = 0; i < kHid; ++i) {
            float s1 = 0.0f;
            for (int h = 0; h < kHid; ++h) s1 += w[kW2 + i * kHid + h] * v[vG2 + h];
            v[vG1 + i] = s1 * (1.0f - v[vH1 + i] * v[vH1 + i]);
        }
    } else {
        for (int c = 0; c < kVec; ++c) v[c] = 0.0f;
    }
    __syncthreads();
    // parameter gradients: sums over the block's samples of outer products
    // (input side x output-side gradient) per layer, biases the gradient sums
    float* out = part + (size_t)blockIdx.x * kMlpParams;
    for (int p = t; p < kMlpParams; p += MTB) {
        int xi, gi;  // vector offsets of the input side (-1: bias) and the gradient side
        if (p < kB1) {
            xi = (p - kW1) / kHid;
            gi = vG1 + (p - kW1) % kHid;
        } else if (p < kW2) {
            xi = -1;
            gi = vG1 + p - kB1;
        } else if (p < kB2) {
            xi = vH1 + (p - kW2) / kHid;
            gi = vG2 + (p - kW2) % kHid;
        } else if (p < kW3) {
            xi = -1;
            gi = vG2 + p - kB2;
        } else if (p < kB3) {
            xi = vH2 + (p - kW3) / kHid;
            gi = vG3 + (p - kW3) % kHid;
        } else if (p < kW4) {
            xi = -1;
            gi = vG3 + p - kB3;
        } else if (p < kB4) {
            xi = vH3 + (p - kW4) / kOut;
            gi = vDO + (p - kW4) % kOut;
        } else {
            xi = -1;
            gi = vDO + p - kB4;
        }
        float acc = 0.0f;
        if (xi < 0)
            for (int s = 0; s < MTB; ++s) acc += Vs[(size_t)s * kVec + gi];
        else
            for (int s = 0; s < MTB; ++s) acc += Vs[(size_t)s * kVec + xi] * Vs[(size_t)s * kVec + gi];
        out[p] = acc;
    }
    for (int o2 = 32; o2 > 0; o2 >>= 1) L += __shfl_xor(L, o2);
    if (t == 0) loss_part[blockIdx.x] = L;
}

// sum of the block partials (scaled) into grad, then an Adam step
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

// the deployed inference (BF16 operands on the matrix cores, FP32
// accumulation) of one sample per block of 64 threads: mu, sigma
__global__ void __launch_bounds__(64) k_mlp_infer(Proposal p, const float* X, int n, float* out) {
    __shared__ double phi[kFeat], mu[3], sig[3];
    const int s = blockIdx.x;
    if (s >= n) return;
    if (threadIdx.x < kFeat) phi[threadIdx.x] = X[(size_t)kFeat * s + threadIdx.x];
    __syncthreads();
    mlp_forward_mfma(p, phi, mu, sig);
    __syncthreads();
    if (threadIdx.x < 3) {
        out[6 * (size_t)s + threadIdx.x] = (float)mu[threadIdx.x];
        out[6 * (size_t)s + 3 + threadIdx.x] = (float)sig[threadIdx.x];
    }
}

}  // namespace

int mlp_blocks(int n) { return (n + MTB - 1) / MTB; }

void launch_mlp_grad(const MlpData& d, const int* sel, int b0, int n, const float* w, float* part, float* loss_part,
                     hipStream_t st) {
    const size_t sh = sizeof(float) * (size_t)MTB * kVec;
    (void)hipFuncSetAttribute((const void*)k_mlp_grad, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    hipLaunchKernelGGL(k_mlp_grad, dim3(mlp_blocks(n)), dim3(MTB), sh, st, d, sel, b0, n, w, part, loss_part);
}

void launch_mlp_adam(float* w, float* m, float* v, const float* part, int nblk, float* grad, float scale,
                     const AdamStep& s, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, w, m, v, part, nblk, grad, scale,
                       s);
}

void launch_mlp_infer(const Proposal& p, const float* X, int n, float* out, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_infer, dim3(n), dim3(64), 0, st, p, X, n, out);
}

}  // namespace hpdmk_amd
