// Device helpers shared by the build and Monte Carlo kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <limits>

#include "dmk_common.hpp"
#include "dmk_kernels.hpp"

namespace hpdmk_amd {

constexpr int TB = 256;   // threads per block (4 waves of 64)
constexpr int kRow = 136; // longest 1-D phase row set: 2(2n+1) + n + 1 for n <= 26

// ------------------------------------------------------------------ helpers
template <typename T>
__device__ __forceinline__ Cplx<T> cmul(Cplx<T> a, Cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// block-wide FP64 sum; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    __syncthreads();
    return s;
}

// N block-wide sums at once (one barrier): v[] holds this thread's values;
// on return thread 0's v[] holds the block totals.  red: (blockDim / 64) * N.
template <int N>
__device__ __forceinline__ void block_sum_n(double* v, double* red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < N; ++c) {
        v[c] = wave_sum(v[c]);
        if (lane == 0) red[w * N + c] = v[c];
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < N; ++c) {
            double s = 0;
            for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k * N + c];
            v[c] = s;
        }
    __syncthreads();
}

// my_mod of include/utils.hpp:96-105 (repeated add/subtract, bounded)
__device__ __forceinline__ double wrap(double x, double L) {
    int guard = 0;
    while (x < 0 && guard++ < 1024) x += L;
    while (x >= L && guard++ < 2048) x -= L;
    if (x < 0 || x >= L) {
        x = fmod(x, L);
        if (x < 0) x += L;
        if (x >= L) x -= L;
    }
    return x;
}

__device__ __forceinline__ int cell_at(double x, double L, int l) {
    const int n = 1 << l;
    long c = (long)floor(x / L * (double)n);
    c = c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
    return (int)c;
}

// descend from the root to the leaf containing (x,y,z); path[l] = node at level l.
// With the dense per-level cell tables the D lookups are independent loads
// (one memory round trip); otherwise a chain of child lookups.
template <typename T>
__device__ int locate(const DevTree<T>& t, double x, double y, double z, int* path) {
    if (t.node_at) {
        int nodes[kNodeAtLevels];
#pragma unroll
        for (int l = 0; l < kNodeAtLevels; ++l) {
            nodes[l] = -1;
            if (l < t.max_depth) {
                const int off = ((1 << (3 * l)) - 1) / 7;
                const int cx = cell_at(x, t.L, l), cy = cell_at(y, t.L, l), cz = cell_at(z, t.L, l);
                nodes[l] = t.node_at[off + cx + (cy << l) + (cz << (2 * l))];
            }
        }
        int lam = 0;
#pragma unroll
        for (int l = 0; l < kNodeAtLevels; ++l)
            if (nodes[l] >= 0) {
                path[l] = nodes[l];
                lam = l;
            }
        return lam;
    }
    int nd = 0, l = 0;
    path[0] = 0;
    while (!t.is_leaf[nd]) {
        const int l1 = l + 1;
        const int bits = (cell_at(x, t.L, l1) & 1) | ((cell_at(y, t.L, l1) & 1) << 1) | ((cell_at(z, t.L, l1) & 1) << 2);
        nd = t.child[8 * nd + bits];
        path[++l] = nd;
    }
    return l;  // leaf level
}

// residual kernel R_b(r) = P(2r/b - 1)/r for 0 < r < b (include/kernels.hpp:51-58)
template <typename R>
__device__ __forceinline__ R residual_poly_over_r(R r, R inv_b, const double* poly, int n_poly) {
    const R u = (r * inv_b - R(0.5)) * R(2);
    if (u >= R(1) || u <= R(-1)) return R(0);
    R v = R(poly[n_poly - 1]);
    for (int k = n_poly - 2; k >= 0; --k) v = v * u + R(poly[k]);
    return v / r;
}

// 1-D phase rows e^{-i (m-n) dk c} (FP64 argument, stored as T); charge folded into X
template <typename T>
__device__ void phase_rows(Cplx<T>* X, Cplx<T>* Y, Cplx<T>* Z, int n, double dk, double x, double y, double z,
                           double q, int tid, int nthr) {
    const int d = 2 * n + 1;
    for (int t = tid; t < 2 * d + n + 1; t += nthr) {
        double c, m;
        double scale = 1.0;
        Cplx<T>* dst;
        if (t < d) {
            c = x;
            m = t - n;
            dst = X + t;
            scale = q;
        } else if (t < 2 * d) {
            c = y;
            m = t - d - n;
            dst = Y + (t - d);
        } else {
            c = z;
            m = t - 2 * d - n;
            dst = Z + (t - 2 * d);
        }
        double s, co;
        sincos(m * (dk * c), &s, &co);
        dst->x = (T)(co * scale);
        dst->y = (T)(-s * scale);
    }
}

template <typename T>
__device__ __forceinline__ Cplx<T> phase_at(const Cplx<T>* X, const Cplx<T>* Y, const Cplx<T>* Z, int d, int e) {
    const int i = e % d, r = e / d, j = r % d, k = r / d;
    return cmul(cmul(X[i], Y[j]), Z[k]);
}

// ------------------------------------------------------------------ trial set-up
template <typename T>
__device__ void trial_setup(const DevTree<T>& t, int i, const double* d, TrialInfo& ti) {
    const int s = t.slot_of[i];
    ti.i = i;
    ti.q = t.pq[s];
    ti.o[0] = t.px[s];
    ti.o[1] = t.py[s];
    ti.o[2] = t.pz[s];
    for (int a = 0; a < 3; ++a) {
        ti.dv[a] = d[a];
        ti.t[a] = wrap(ti.o[a] + d[a], t.L);
    }
    ti.lam_o = locate(t, ti.o[0], ti.o[1], ti.o[2], ti.path_o);
    ti.lam_t = locate(t, ti.t[0], ti.t[1], ti.t[2], ti.path_t);
}

// self-pair correction q^2 R_{b_2}(d_min) of src/energy_shift.cpp:33-41, :58
// The nearest of the 27 images of o (shifts -1, 0, 1 per axis) is the
// per-axis nearest one, so its distance comes from three per-axis minima
// (the same sum, hence the same value, as the minimum over all 27).
template <typename T>
__device__ double self_term(const DevTree<T>& t, const TrialInfo& ti) {
    double dr = sqrt(ti.dv[0] * ti.dv[0] + ti.dv[1] * ti.dv[1] + ti.dv[2] * ti.dv[2]);
    double sq[3];
    for (int c = 0; c < 3; ++c) {
        sq[c] = INFINITY;
        for (int k = -1; k <= 1; ++k) {
            const double e = ti.t[c] - (ti.o[c] + k * t.L);
            sq[c] = fmin(sq[c], e * e);
        }
    }
    dr = fmin(dr, sqrt(sq[0] + sq[1] + sq[2]));
    if (dr == 0) return 0.0;
    const double b2 = t.boxsize[2];
    const double u = (dr / b2 - 0.5) * 2.0;
    if (fabs(u) >= 1.0) return 0.0;
    double v = t.poly[t.n_poly - 1];
    for (int k = t.n_poly - 2; k >= 0; --k) v = v * u + t.poly[k];
    return v / dr * ti.q * ti.q;
}

}  // namespace hpdmk_amd
