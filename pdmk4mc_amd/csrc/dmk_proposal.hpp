// Learned proposal of the Monte Carlo chain (no reference counterpart).
//
// A trial move of particle i draws its displacement from a Gaussian whose
// mean and diagonal scale an MLP computes from features of i's neighbourhood:
//     delta = mu(phi(o)) + sigma(phi(o)) z,     z ~ N(0, I_3)
// and the Metropolis-Hastings decision adds the exact proposal ratio
//     H = log q(-delta | phi(t)) - log q(delta | phi(o)),
// accept iff lB dE - H / beta < -ln(u) / beta  (and no hard-sphere overlap).
// phi(x) is evaluated with particle i removed, so phi(t) is the feature of
// the reverse move.  Periodic images of the Gaussian beyond the nearest are
// dropped (their weight is below exp(-L^2 / (2 smax^2))).
//
// Features of a point x for particle i, over the other particles j within the
// feature radius rf (minimum image v = x - x_j, r = |v|, 0 < r < rf):
//     s1 = sum q_j v / (r^2 + a^2)^{3/2}      (softened local field, a = 0.5)
//     s2 = number of neighbours,  s3 = sum w v / r,  w = (1 - r/rf)^2
//     phi = [q_i, tanh(q_i s1 / 1), tanh(s2 / 8), tanh(s3 / 2)]   (8 values)
// MLP 8 -> 64 -> 64 -> 64 -> 6 with tanh hidden layers; outputs o give
//     mu = max_mean tanh(o[0:3]),  sigma = smin + (smax - smin) sigmoid(o[3:6]).
// Weights (float, row-major): W1[8][64] b1[64] W2[64][64] b2[64] W3[64][64]
// b3[64] W4[64][6] b4[6].  Training is FP32 (mlp_train.hip); inference runs
// on the matrix cores with BF16 operands (features, weights, hidden
// activations rounded to nearest even) and FP32 accumulation, biases and tanh
// in FP32 (mlp_forward_mfma).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace hpdmk_amd {

constexpr int kFeat = 8;
constexpr int kHid = 64;
constexpr int kOut = 6;
constexpr int kMlpParams = kFeat * kHid + kHid + 2 * (kHid * kHid + kHid) + kHid * kOut + kOut;
// offsets of the parameter blocks
constexpr int kW1 = 0, kB1 = kW1 + kFeat * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid, kW3 = kB2 + kHid,
              kB3 = kW3 + kHid * kHid, kW4 = kB3 + kHid, kB4 = kW4 + kHid * kOut;
constexpr int kFeatSums = 7;
constexpr double kFeatSoft = 0.5;
constexpr double kFeatScale[3] = {1.0, 8.0, 2.0};
// BF16 weights in matrix-core operand order (mlp_pack_weights): per layer,
// per (k-step s, output tile t), 64 lanes x 4 values, lane l holding
// W[16 s + 4 (l / 16) + j][16 t + l % 16], j = 0..3 (zero outside the layer)
constexpr int kPackL1 = 1 * 4 * 256, kPackL2 = 4 * 4 * 256, kPackL3 = 4 * 4 * 256, kPackL4 = 4 * 1 * 256;
constexpr int kPackOff1 = 0, kPackOff2 = kPackOff1 + kPackL1, kPackOff3 = kPackOff2 + kPackL2,
              kPackOff4 = kPackOff3 + kPackL3, kPackSize = kPackOff4 + kPackL4;
static_assert(kPackSize % 8 == 0, "the packed weights copy as int4");

struct Proposal {
    int learned;       // 0: displacements as loaded; 1: drawn from the MLP (the list holds z)
    const float* w;    // kMlpParams (learned): the biases are read from here
    const unsigned short* wp;  // kPackSize BF16 weights in operand order (learned)
    const double* z;   // per move: N(0,1) triple (learned)
    double rf;         // feature radius (0: no features)
    double max_mean, smin, smax;
    double* dprop;     // per move: displacement used (learned)
    double* logq;      // per move: log q(delta | phi(o)) (learned)
    double* feat;      // per move: phi(o) (kFeat), when rf > 0
    double* feat_t;    // per move: phi(t) (kFeat), learned
    double* H;         // per move: proposal log-ratio (learned)
};

// adds the pair (v = x - x_j, r2 = |v|^2 in (0, rf^2), charge qj) to the sums s[7]
__host__ __device__ inline void feature_pair(double vx, double vy, double vz, double r2, double qj, double rf,
                                             double* s) {
    const double r = sqrt(r2);
    const double a2 = kFeatSoft * kFeatSoft;
    const double inv = 1.0 / ((r2 + a2) * sqrt(r2 + a2));
    s[0] += qj * vx * inv;
    s[1] += qj * vy * inv;
    s[2] += qj * vz * inv;
    const double ir = 1.0 / r, x = 1.0 - r / rf;
    const double wr = x * x * ir;
    s[3] += 1.0;
    s[4] += wr * vx;
    s[5] += wr * vy;
    s[6] += wr * vz;
}

__host__ __device__ inline void features_from_sums(double q, const double* s, double* phi) {
    phi[0] = q;
    for (int a = 0; a < 3; ++a) phi[1 + a] = tanh(q * s[a] / kFeatScale[0]);
    phi[4] = tanh(s[3] / kFeatScale[1]);
    for (int a = 0; a < 3; ++a) phi[5 + a] = tanh(s[4 + a] / kFeatScale[2]);
}
// feature f of features_from_sums alone (the same expression)
__host__ __device__ inline double feature_from_sums(int f, double q, const double* s) {
    if (f == 0) return q;
    if (f < 4) return tanh(q * s[f - 1] / kFeatScale[0]);
    if (f == 4) return tanh(s[3] / kFeatScale[1]);
    return tanh(s[f - 1] / kFeatScale[2]);
}

// log N(delta; mu, diag sigma^2)
__host__ __device__ inline double gauss_logq(const double* delta, const double* mu, const double* sig) {
    double v = -1.5 * log(2.0 * M_PI);
    for (int a = 0; a < 3; ++a) {
        const double y = (delta[a] - mu[a]) / sig[a];
        v -= log(sig[a]) + 0.5 * y * y;
    }
    return v;
}

// round-to-nearest-even FP32 -> BF16 bits
__host__ __device__ inline unsigned short bf16_bits(float x) {
    union {
        float f;
        unsigned int u;
    } v;
    v.f = x;
    const unsigned int lsb = (v.u >> 16) & 1u;
    return (unsigned short)((v.u + 0x7fffu + lsb) >> 16);
}

// The packed BF16 weights (kPackSize values) from the float parameters.
inline void mlp_pack_weights(const float* w, unsigned short* wp) {
    const int off[4] = {kPackOff1, kPackOff2, kPackOff3, kPackOff4};
    const int wofs[4] = {kW1, kW2, kW3, kW4};
    const int nin[4] = {kFeat, kHid, kHid, kHid}, nout[4] = {kHid, kHid, kHid, kOut};
    for (int L = 0; L < 4; ++L) {
        const int ns = (nin[L] + 15) / 16, nt = (nout[L] + 15) / 16;
        for (int s = 0; s < ns; ++s)
            for (int t = 0; t < nt; ++t)
                for (int l = 0; l < 64; ++l)
                    for (int j = 0; j < 4; ++j) {
                        const int k = 16 * s + 4 * (l / 16) + j, n = 16 * t + l % 16;
                        const float v = (k < nin[L] && n < nout[L]) ? w[wofs[L] + k * nout[L] + n] : 0.0f;
                        wp[off[L] + ((s * nt + t) * 64 + l) * 4 + j] = bf16_bits(v);
                    }
    }
}

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

// One layer on v_mfma_f32_16x16x16bf16_1k for one sample: the input vector is
// row 0 of the A operand (lanes l % 16 == 0 hold a[s] = x[16 s + 4 (l/16) + j];
// the other lanes zeros), the packed weights the B operand; on return lane
// l < 16 holds unit 16 t + l of the FP32 sums in acc[t][0].
template <int NS, int NT>
__device__ inline void mlp_layer(const mlp_s4* a, const unsigned short* wp, mlp_f4* acc) {
    const int l = threadIdx.x & 63;
    mlp_s4 b[NS][NT];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int t = 0; t < NT; ++t) b[s][t] = *reinterpret_cast<const mlp_s4*>(wp + ((s * NT + t) * 64 + l) * 4);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        acc[t] = mlp_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[s], b[s][t], acc[t], 0, 0, 0);
    }
}

// tanh(sum + bias) of the 64 hidden units, rounded to BF16 and moved into the
// A-operand layout of the next layer by lane shuffles
__device__ inline void mlp_hidden(const mlp_f4* acc, const float* bias, mlp_s4* a) {
    const int l = threadIdx.x & 63;
    short h[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) h[t] = l < 16 ? (short)bf16_bits(tanhf(acc[t][0] + bias[16 * t + l])) : (short)0;
    const int g = l >> 4;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        mlp_s4 v;
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = (short)__shfl((int)h[s], 4 * g + j);
        a[s] = (l & 15) == 0 ? v : mlp_s4{0, 0, 0, 0};
    }
}

// the biases b1 b2 b3 b4 one after the other (a block's LDS copy)
constexpr int kBiasPack = 3 * kHid + kOut;

// A block's LDS copy of the packed BF16 weights and the biases, loaded by
// every thread of the block (no barrier: the caller syncs before the MLP).
__device__ inline void mlp_stage(const Proposal& p, unsigned short* wp_s, float* b_s) {
    const int4* src = reinterpret_cast<const int4*>(p.wp);
    int4* dst = reinterpret_cast<int4*>(wp_s);
    for (int k = threadIdx.x; k < kPackSize / 8; k += blockDim.x) dst[k] = src[k];
    for (int k = threadIdx.x; k < kBiasPack; k += blockDim.x)
        b_s[k] = k < 3 * kHid ? p.w[kB1 + (k / kHid) * (kHid * kHid + kHid) + k % kHid] : p.w[kB4 + k - 3 * kHid];
}

// BF16 forward pass by one wave (wave 0 of the block; the others idle).
// phi (kFeat doubles, shared) in; mu/sig (3 each, shared) out, formed in
// FP64 from the FP32 outputs.  The caller syncs the block around the call.
// wp_s / b_s: the block's LDS copy of the weights (mlp_stage), else p's.
__device__ inline void mlp_forward_mfma(const Proposal& p, const double* phi, double* mu, double* sig,
                                        const unsigned short* wp_s = nullptr, const float* b_s = nullptr) {
    if (threadIdx.x >= 64) return;
    const unsigned short* wp = wp_s ? wp_s : p.wp;
    const float* b1 = b_s ? b_s : p.w + kB1;
    const float* b2 = b_s ? b_s + kHid : p.w + kB2;
    const float* b3 = b_s ? b_s + 2 * kHid : p.w + kB3;
    const float* b4 = b_s ? b_s + 3 * kHid : p.w + kB4;
    const int l = threadIdx.x, g = l >> 4;
    mlp_s4 a[4];
    {  // features: K = 8, one k-step (lanes with g < 2 carry values)
        mlp_s4 v;
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = (short)bf16_bits((float)(4 * g + j < kFeat ? phi[4 * g + j] : 0.0));
        a[0] = (l & 15) == 0 && g < 2 ? v : mlp_s4{0, 0, 0, 0};
    }
    mlp_f4 acc[4];
    mlp_layer<1, 4>(a, wp + kPackOff1, acc);
    mlp_hidden(acc, b1, a);
    mlp_layer<4, 4>(a, wp + kPackOff2, acc);
    mlp_hidden(acc, b2, a);
    mlp_layer<4, 4>(a, wp + kPackOff3, acc);
    mlp_hidden(acc, b3, a);
    mlp_layer<4, 1>(a, wp + kPackOff4, acc);
    if (l < 3) {
        const double om = (double)(acc[0][0] + b4[l]);
        mu[l] = p.max_mean * tanh(om);
    } else if (l < 6) {
        const double os = (double)(acc[0][0] + b4[l]);
        sig[l - 3] = p.smin + (p.smax - p.smin) / (1.0 + exp(-os));
    }
}

}  // namespace hpdmk_amd
