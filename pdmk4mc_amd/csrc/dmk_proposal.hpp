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
//     s1 = sum q_j v / (r^2 + a^2)^{3/2}      (softened field, a = 0.5)
//     s2 = sum w,  s3 = sum w v / r,           w = (1 - r/rf)^2
//     phi = [q_i, tanh(q_i s1 / 1), tanh(s2 / 4), tanh(s3 / 2)]   (8 values)
// MLP 8 -> 32 -> 32 -> 6 with tanh hidden layers; outputs o give
//     mu = max_mean tanh(o[0:3]),  sigma = smin + (smax - smin) sigmoid(o[3:6]).
// Weights (float, row-major): W1[8][32] b1[32] W2[32][32] b2[32] W3[32][6] b3[6].
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace hpdmk_amd {

constexpr int kFeat = 8;
constexpr int kHid = 32;
constexpr int kOut = 6;
constexpr int kMlpParams = kFeat * kHid + kHid + kHid * kHid + kHid + kHid * kOut + kOut;
constexpr int kFeatSums = 7;
constexpr double kFeatSoft = 0.5;
constexpr double kFeatScale[3] = {1.0, 4.0, 2.0};

struct Proposal {
    int learned;       // 0: displacements as loaded; 1: drawn from the MLP (the list holds z)
    const float* w;    // kMlpParams (learned)
    const double* z;   // per move: N(0,1) triple (learned)
    double rf;         // feature radius (0: no features)
    double max_mean, smin, smax;
    double* dprop;     // per move: displacement used (learned)
    double* logq;      // per move: log q(delta | phi(o)) (learned)
    double* feat;      // per move: phi(o) (kFeat), when rf > 0
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
    const double w = (1.0 - r / rf) * (1.0 - r / rf);
    s[3] += w;
    s[4] += w * vx / r;
    s[5] += w * vy / r;
    s[6] += w * vz / r;
}

__host__ __device__ inline void features_from_sums(double q, const double* s, double* phi) {
    phi[0] = q;
    for (int a = 0; a < 3; ++a) phi[1 + a] = tanh(q * s[a] / kFeatScale[0]);
    phi[4] = tanh(s[3] / kFeatScale[1]);
    for (int a = 0; a < 3; ++a) phi[5 + a] = tanh(s[4 + a] / kFeatScale[2]);
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

// FP64 forward pass by all threads of a block (at least 64): the weights are
// staged in shared memory (wsh, kMlpParams floats), every unit is computed by
// 8 lanes (partial dot products joined with shuffles).  phi (8), hbuf
// (2 kHid) and o (kOut) are shared; on return mu/sig (3 each, shared) are set.
__device__ inline void mlp_forward_block(const float* w, const Proposal& p, const double* phi, double* hbuf,
                                         double* o, double* mu, double* sig, float* wsh) {
    const int t = threadIdx.x;
    for (int k = t; k < kMlpParams; k += blockDim.x) wsh[k] = w[k];
    const float* W1 = wsh;
    const float* b1 = W1 + kFeat * kHid;
    const float* W2 = b1 + kHid;
    const float* b2 = W2 + kHid * kHid;
    const float* W3 = b2 + kHid;
    const float* b3 = W3 + kHid * kOut;
    const int part = t & 7, nu = blockDim.x >> 3;  // 8 lanes per unit, units strided by the block
    __syncthreads();
    for (int unit = t >> 3; unit < kHid; unit += nu) {
        double s = (double)W1[part * kHid + unit] * phi[part];  // kFeat = 8: one input per lane
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        if (part == 0) hbuf[unit] = tanh(s + (double)b1[unit]);
    }
    __syncthreads();
    for (int unit = t >> 3; unit < kHid; unit += nu) {
        double s = 0.0;
        for (int h = part; h < kHid; h += 8) s += (double)W2[h * kHid + unit] * hbuf[h];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        if (part == 0) hbuf[kHid + unit] = tanh(s + (double)b2[unit]);
    }
    __syncthreads();
    const int unit = t >> 3;
    if (unit < kOut) {
        double s = 0.0;
        for (int h = part; h < kHid; h += 8) s += (double)W3[h * kOut + unit] * hbuf[kHid + h];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        if (part == 0) o[unit] = s + (double)b3[unit];
    }
    __syncthreads();
    if (t < 3) {
        mu[t] = p.max_mean * tanh(o[t]);
        sig[t] = p.smin + (p.smax - p.smin) / (1.0 + exp(-o[3 + t]));
    }
    __syncthreads();
}

}  // namespace hpdmk_amd
