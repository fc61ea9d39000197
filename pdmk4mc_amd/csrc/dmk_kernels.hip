// HIP kernels (gfx950, wave64) for the DMK Monte Carlo hot path:
//   * outgoing plane waves (root grid and per box), incoming assembly
//   * total energy: window, difference and residual terms
//   * trial-move energy change, Metropolis decision, accept/update
//   * batched trials: pairwise corrections and the sequential decision scan
//
// Precision: T is the storage/arithmetic type of the bulk (double = FP64 mode,
// float = mixed mode).  Phase arguments k.x are always formed and reduced in
// FP64 (1-D tables e^{-i k x} in LDS), and every reduction accumulates in FP64.
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

// descend from the root to the leaf containing (x,y,z); path[l] = node at level l
template <typename T>
__device__ int locate(const DevTree<T>& t, double x, double y, double z, int* path) {
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

// ------------------------------------------------------------------ trial info
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
template <typename T>
__device__ double self_term(const DevTree<T>& t, const TrialInfo& ti) {
    double dr = sqrt(ti.dv[0] * ti.dv[0] + ti.dv[1] * ti.dv[1] + ti.dv[2] * ti.dv[2]);
    for (int mx = -1; mx <= 1; ++mx)
        for (int my = -1; my <= 1; ++my)
            for (int mz = -1; mz <= 1; ++mz) {
                const double ax = ti.t[0] - (ti.o[0] + mx * t.L), ay = ti.t[1] - (ti.o[1] + my * t.L),
                             az = ti.t[2] - (ti.o[2] + mz * t.L);
                dr = fmin(dr, sqrt(ax * ax + ay * ay + az * az));
            }
    if (dr == 0) return 0.0;
    const double b2 = t.boxsize[2];
    const double u = (dr / b2 - 0.5) * 2.0;
    if (fabs(u) >= 1.0) return 0.0;
    double v = t.poly[t.n_poly - 1];
    for (int k = t.n_poly - 2; k >= 0; --k) v = v * u + t.poly[k];
    return v / dr * ti.q * ti.q;
}

// ------------------------------------------------------------------ trial kernel
// grid: (blocks_per_trial, n_trials).  Block roles inside a trial:
//   [0, nWb)                          window chunk
//   [nWb, nWb + 2*nL*nDb)             difference chunk (endpoint, level, chunk)
//   [.., + 2*max_near)                near-field entry (endpoint, entry)
// The last block of a trial to finish assembles dE (src/energy_shift.cpp:67)
// and, when requested, takes the Metropolis decision.
template <typename T>
__global__ void __launch_bounds__(TB) k_trial(DevTree<T> t, TrialArgs a) {
    __shared__ TrialInfo ti;
    __shared__ double red[TB / 64];
    __shared__ Cplx<T> tab[4 * 64];
    __shared__ int last;
    const int tr = blockIdx.y;
    const int b = blockIdx.x;
    if (*t.err) return;  // sticky device error: the run is void, do nothing
    if (threadIdx.x == 0) trial_setup(t, a.idx[a.move0 + tr], a.d + 3 * (a.move0 + tr), ti);
    __syncthreads();

    const int nWb = (t.Mw + TB - 1) / TB;
    const int nDb = (t.Md + TB - 1) / TB;
    const int nL = t.max_depth >= 4 ? t.max_depth - 3 : 0;
    double part = 0.0;
    int slot_acc = -1, nscan = 0;
    if (b < nWb) {
        // window: 2 cW sum W Re[(tau - omega) conj(rho - omega)] (include/vecops.hpp:118-145)
        const int nw = t.n_window, dw = t.dw, len = 2 * dw + nw + 1;
        Cplx<T>* Xt = tab;
        Cplx<T>* Yt = Xt + dw;
        Cplx<T>* Zt = Yt + dw;
        Cplx<T>* Xo = tab + len;
        Cplx<T>* Yo = Xo + dw;
        Cplx<T>* Zo = Yo + dw;
        phase_rows(Xt, Yt, Zt, nw, t.delta_k[0], ti.t[0], ti.t[1], ti.t[2], ti.q, threadIdx.x, blockDim.x);
        phase_rows(Xo, Yo, Zo, nw, t.delta_k[0], ti.o[0], ti.o[1], ti.o[2], ti.q, threadIdx.x, blockDim.x);
        __syncthreads();
        const int e = b * TB + threadIdx.x;
        if (e < t.Mw) {
            const Cplx<T> ta = phase_at(Xt, Yt, Zt, dw, e), om = phase_at(Xo, Yo, Zo, dw, e);
            const Cplx<T> g = t.root[e];
            const T ar = ta.x - om.x, ai = ta.y - om.y, gr = g.x - om.x, gi = g.y - om.y;
            part = (double)((ar * gr + ai * gi) * t.Wmat[e]);
        }
        part *= 2.0 * t.cW;
        slot_acc = 0;
    } else if (b < nWb + 2 * nL * nDb) {
        // difference levels strictly above the endpoint's leaf (src/energy_shift.cpp:92-140)
        const int r = b - nWb;
        const int ep = r / (nL * nDb);  // 0 target, 1 origin
        const int li = (r / nDb) % nL;
        const int ch = r % nDb;
        const int l = 2 + li;
        const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
        slot_acc = 1;
        if (l <= lam - 1) {
            const int node = ep == 0 ? ti.path_t[l] : ti.path_o[l];
            const Cplx<T>* in = t.inc + (size_t)t.pw_slot[node] * t.Md;
            const T* D = t.Dmat + (size_t)li * t.Md;
            const double* c = ep == 0 ? ti.t : ti.o;
            const int nd = t.n_diff, dd = t.dd;
            Cplx<T>* X = tab;
            Cplx<T>* Y = X + dd;
            Cplx<T>* Z = Y + dd;
            phase_rows(X, Y, Z, nd, t.delta_k[l], c[0], c[1], c[2], ti.q, threadIdx.x, blockDim.x);
            __syncthreads();
            const int e = ch * TB + threadIdx.x;
            if (e < t.Md) {
                const Cplx<T> w = phase_at(X, Y, Z, dd, e);
                const Cplx<T> v = in[e];
                T re = w.x * v.x - w.y * v.y;
                if (ep == 1) re -= w.x * w.x + w.y * w.y;  // origin excludes its own wave (vecops.hpp:161-163)
                part = (double)(re * D[e]);
            }
            part *= 2.0 * t.cD[l] * (ep == 0 ? 1.0 : -1.0);
        }
    } else {
        // near field of the endpoint's leaf (src/energy_shift.cpp:233-302, direct_eval.hpp:309-403)
        const int r = b - nWb - 2 * nL * nDb;
        const int ep = r / t.max_near;
        const int k = r % t.max_near;
        const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
        const int leaf = ep == 0 ? ti.path_t[lam] : ti.path_o[lam];
        slot_acc = 2 + ep;
        const int off = t.near_off[leaf], cnt = t.near_off[leaf + 1] - off;
        if (k < cnt) {
            const int nb = t.near_ids[off + k];
            const int code = t.near_img[off + k];
            const double sx = (code % 3 - 1) * t.L, sy = ((code / 3) % 3 - 1) * t.L, sz = (code / 9 - 1) * t.L;
            const double* c = ep == 0 ? ti.t : ti.o;
            const double bb = t.boxsize[lam];
            const T b2 = (T)(bb * bb), inv_b = (T)(1.0 / bb);
            const T eps = std::numeric_limits<T>::epsilon();
            const int s0 = t.slot_start[nb], n = t.slot_count[nb];
            nscan = n;
            double acc = 0;
            for (int p = threadIdx.x; p < n; p += blockDim.x) {
                const int s = s0 + p;
                const T dx = (T)(c[0] - (t.px[s] + sx)), dy = (T)(c[1] - (t.py[s] + sy)),
                        dz = (T)(c[2] - (t.pz[s] + sz));
                const T r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > eps && r2 < b2) acc += (double)((T)t.pq[s] * residual_poly_over_r<T>(sqrt(r2), inv_b, t.poly, t.n_poly));
            }
            part = acc * ti.q;
        }
    }
    const double s = block_sum(part, red);
    double* acc = a.acc + kAccPerTrial * tr;
    if (threadIdx.x == 0) {
        if (s != 0.0) atomicAdd(acc + slot_acc, s);
        if (nscan) atomicAdd(acc + 2 + slot_acc, (double)nscan);
        __threadfence();
        const unsigned prev = atomicAdd(a.done + tr, 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    __threadfence();
    double v[kAccPerTrial];
    for (int k = 0; k < kAccPerTrial; ++k)
        v[k] = __hip_atomic_load(acc + k, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const double self = self_term(t, ti);
    const double dE = v[0] + v[1] + v[2] - v[3] - self;
    const int m = a.move0 + tr;
    a.dE[m] = dE;
    if (a.parts) {
        a.parts[5 * m + 0] = v[0];
        a.parts[5 * m + 1] = v[1];
        a.parts[5 * m + 2] = v[2];
        a.parts[5 * m + 3] = v[3];
        a.parts[5 * m + 4] = self;
    }
    TrialInfo& out = a.info[tr];
    out = ti;
    int acc_flag = 0;
    if (a.mode == kTrialMetropolis) {
        acc_flag = a.u[m] < exp(-a.beta * dE);
        if (acc_flag) {
            const int lt = ti.path_t[ti.lam_t], lo = ti.path_o[ti.lam_o];
            if (lt != lo && t.slot_count[lt] >= t.slot_cap[lt]) {
                atomicOr(t.err, (int)kErrLeafOverflow);
                acc_flag = 0;
            }
        }
    } else if (a.mode == kTrialForceAccept) {
        acc_flag = 1;
        const int lt = ti.path_t[ti.lam_t], lo = ti.path_o[ti.lam_o];
        if (lt != lo && t.slot_count[lt] >= t.slot_cap[lt]) {
            atomicOr(t.err, (int)kErrLeafOverflow);
            acc_flag = 0;
        }
    }
    a.accept[tr] = acc_flag;
    if (a.accept_out) a.accept_out[m] = acc_flag;
    if (a.stats) {
        int* st = a.stats + kStatsPerMove * m;
        st[0] = ti.lam_t - 2 > 0 ? ti.lam_t - 2 : 0;
        st[1] = ti.lam_o - 2 > 0 ? ti.lam_o - 2 : 0;
        st[2] = (int)v[4];
        st[3] = (int)v[5];
        for (int ep = 0; ep < 2; ++ep) {  // arrays an accepted move rewrites (src/tree.cpp:484-579)
            const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
            int arrays = 0;
            for (int l = 2; l <= lam && l <= t.max_depth - 2; ++l) {
                const int node = ep == 0 ? ti.path_t[l] : ti.path_o[l];
                arrays += 2;
                for (int c = 0; c < 26; ++c) {
                    const int C = t.colleague[26 * node + c];
                    arrays += (C >= 0 && !t.is_leaf[C]);
                }
            }
            st[4 + ep] = arrays;
        }
    }
    for (int k = 0; k < kAccPerTrial; ++k) acc[k] = 0.0;  // ready for the next launch (no memset node needed)
    a.done[tr] = 0;
}

// ------------------------------------------------------------------ update kernel
// Applies accepted moves (src/tree.cpp:447-597): root += tau - omega; at each
// path level l in [2, leaf level] (l != D-1): out[B] += tau, in[B] += conj tau,
// in[C] += conj(tau s_p) for non-leaf colleagues C; minus for the origin.
// Every element index is owned by one thread across all arrays and moves, so
// moves that touch the same arrays never race.  One extra block relocates the
// particles between leaf slots.
template <typename T>
__device__ void apply_level(const DevTree<T>& t, const TrialInfo& ti, int ep, int l, int li, int e,
                            const Cplx<T>* tab) {
    const int nd = t.n_diff, dd = t.dd;
    // 9 rows: axis a, shift s in {-1,0,1} at tab + (3*a + s + 1) * dd
    const Cplx<T>* row[3][3];
    for (int ax = 0; ax < 3; ++ax)
        for (int s = 0; s < 3; ++s) row[ax][s] = tab + (3 * ax + s) * dd;
    const int i = e % dd, rr = e / dd, j = rr % dd, k = rr / dd;
    (void)nd;
    const T sgn = ep == 0 ? T(1) : T(-1);
    const int node = ep == 0 ? ti.path_t[l] : ti.path_o[l];
    Cplx<T> tau = cmul(cmul(row[0][1][i], row[1][1][j]), row[2][1][k]);
    tau.x *= sgn;
    tau.y *= sgn;
    const size_t base = (size_t)t.pw_slot[node] * t.Md + e;
    Cplx<T> o = t.out[base];
    o.x += tau.x;
    o.y += tau.y;
    t.out[base] = o;
    Cplx<T> in = t.inc[base];
    in.x += tau.x;
    in.y -= tau.y;
    t.inc[base] = in;
    const int nl = 1 << l;
    const int* cl = t.colleague + 26 * node;
    int sidx = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!dx && !dy && !dz) continue;
                const int C = cl[sidx++];
                if (C < 0 || t.is_leaf[C]) continue;
                // image of the moving particle adjacent to C: opposite of C's wrap seen from B
                const int cx = t.cell[3 * node] + dx, cy = t.cell[3 * node + 1] + dy, cz = t.cell[3 * node + 2] + dz;
                const int px = cx < 0 ? 1 : (cx >= nl ? -1 : 0);
                const int py = cy < 0 ? 1 : (cy >= nl ? -1 : 0);
                const int pz = cz < 0 ? 1 : (cz >= nl ? -1 : 0);
                Cplx<T> ts = cmul(cmul(row[0][px + 1][i], row[1][py + 1][j]), row[2][pz + 1][k]);
                const size_t cb = (size_t)t.pw_slot[C] * t.Md + e;
                Cplx<T> v = t.inc[cb];
                v.x += sgn * ts.x;
                v.y -= sgn * ts.y;
                t.inc[cb] = v;
            }
}

template <typename T>
__global__ void __launch_bounds__(TB) k_update(DevTree<T> t, UpdateArgs a) {
    __shared__ Cplx<T> tab[9 * 64];
    if (*t.err) return;
    const int nWb = (t.Mw + TB - 1) / TB;
    const int nDb = (t.Md + TB - 1) / TB;
    const int nL = t.max_depth >= 4 ? t.max_depth - 3 : 0;
    const int b = blockIdx.x;
    if (b < nWb) {
        const int nw = t.n_window, dw = t.dw;
        const int e = b * TB + threadIdx.x;
        Cplx<T> acc = {T(0), T(0)};
        for (int m = 0; m < a.n_moves; ++m) {
            if (!a.accept[m]) continue;
            const TrialInfo& ti = a.info[m];
            __syncthreads();
            phase_rows(tab, tab + dw, tab + 2 * dw, nw, t.delta_k[0], ti.t[0], ti.t[1], ti.t[2], ti.q, threadIdx.x,
                       blockDim.x);
            Cplx<T>* o = tab + 3 * dw;
            phase_rows(o, o + dw, o + 2 * dw, nw, t.delta_k[0], ti.o[0], ti.o[1], ti.o[2], ti.q, threadIdx.x,
                       blockDim.x);
            __syncthreads();
            if (e < t.Mw) {
                const Cplx<T> ta = phase_at(tab, tab + dw, tab + 2 * dw, dw, e);
                const Cplx<T> om = phase_at(o, o + dw, o + 2 * dw, dw, e);
                acc.x += ta.x - om.x;
                acc.y += ta.y - om.y;
            }
        }
        if (e < t.Mw) {
            Cplx<T> r = t.root[e];
            r.x += acc.x;
            r.y += acc.y;
            t.root[e] = r;
        }
        return;
    }
    if (b < nWb + nL * nDb) {
        const int r = b - nWb;
        const int li = r / nDb, ch = r % nDb;
        const int l = 2 + li;
        const int e = ch * TB + threadIdx.x;
        const int dd = t.dd, nd = t.n_diff;
        for (int m = 0; m < a.n_moves; ++m) {
            if (!a.accept[m]) continue;
            const TrialInfo& ti = a.info[m];
            for (int ep = 0; ep < 2; ++ep) {
                const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
                if (l > lam) continue;  // no box of this path at level l
                const double* c = ep == 0 ? ti.t : ti.o;
                __syncthreads();
                // rows for positions shifted by -L, 0, +L on each axis
                for (int s = 0; s < 3; ++s) {
                    const double sh = (s - 1) * t.L;
                    Cplx<T>* X = tab + (0 * 3 + s) * dd;
                    Cplx<T>* Y = tab + (1 * 3 + s) * dd;
                    Cplx<T>* Z = tab + (2 * 3 + s) * dd;
                    // phase_rows writes Z with n+1 entries; X carries the charge
                    phase_rows(X, Y, Z, nd, t.delta_k[l], c[0] + sh, c[1] + sh, c[2] + sh, ti.q, threadIdx.x,
                               blockDim.x);
                }
                __syncthreads();
                if (e < t.Md) apply_level(t, ti, ep, l, li, e, tab);
            }
        }
        return;
    }
    // particle relocation (src/tree.cpp:454-463, 583-596), exact membership
    if (threadIdx.x != 0) return;
    for (int m = 0; m < a.n_moves; ++m) {
        if (!a.accept[m]) continue;
        const TrialInfo& ti = a.info[m];
        const int i = ti.i;
        const int s = t.slot_of[i];
        const int lo = t.leaf_of[i];
        const int lt = ti.path_t[ti.lam_t];
        if (lo == lt) {
            t.px[s] = ti.t[0];
            t.py[s] = ti.t[1];
            t.pz[s] = ti.t[2];
            continue;
        }
        const int c = t.slot_count[lt];
        const int ns = t.slot_start[lt] + c;
        t.px[ns] = ti.t[0];
        t.py[ns] = ti.t[1];
        t.pz[ns] = ti.t[2];
        t.pq[ns] = t.pq[s];
        t.porig[ns] = i;
        t.slot_count[lt] = c + 1;
        t.slot_of[i] = ns;
        t.leaf_of[i] = lt;
        const int last = t.slot_start[lo] + t.slot_count[lo] - 1;
        if (last != s) {
            t.px[s] = t.px[last];
            t.py[s] = t.py[last];
            t.pz[s] = t.pz[last];
            t.pq[s] = t.pq[last];
            const int j = t.porig[last];
            t.porig[s] = j;
            t.slot_of[j] = s;
        }
        t.px[last] = 0;
        t.py[last] = 0;
        t.pz[last] = 0;
        t.pq[last] = 0;
        t.porig[last] = -1;
        t.slot_count[lo] -= 1;
    }
}

// ------------------------------------------------------------------ batched trials
// Effect of accepted move i on trial j (i < j, distinct particles), exactly the
// change the update of move i makes to the quantities trial j reads:
//   window: 2 cW sum W Re[(tau_j - omega_j) conj(tau_i - omega_i)]
//   difference level l of j's endpoint box B: in[B] changes by
//           conj(tau_i s) if box(t_i, l) is B or its neighbour, minus the same for o_i
//   near field: +- q_i q_j R_b(|x_j - y_i|) for y_i in {t_i, o_i} (b of x_j's leaf)
template <typename T>
__global__ void __launch_bounds__(TB) k_pair_corr(DevTree<T> t, CorrArgs a) {
    __shared__ Cplx<T> tab[12 * 64];
    __shared__ double red[TB / 64];
    // pair index -> (i, j), i < j
    const int K = a.K;
    int p = blockIdx.x, i = 0;
    while (p >= K - 1 - i) {
        p -= K - 1 - i;
        ++i;
    }
    const int j = i + 1 + p;
    if (*t.err) return;
    const TrialInfo& A = a.info[i];
    const TrialInfo& Bj = a.info[j];
    double total = 0;
    {   // window
        const int nw = t.n_window, dw = t.dw, len = 2 * dw + nw + 1;
        Cplx<T>* T0 = tab;
        Cplx<T>* T1 = tab + len;
        Cplx<T>* T2 = tab + 2 * len;
        Cplx<T>* T3 = tab + 3 * len;
        phase_rows(T0, T0 + dw, T0 + 2 * dw, nw, t.delta_k[0], Bj.t[0], Bj.t[1], Bj.t[2], Bj.q, threadIdx.x, blockDim.x);
        phase_rows(T1, T1 + dw, T1 + 2 * dw, nw, t.delta_k[0], Bj.o[0], Bj.o[1], Bj.o[2], Bj.q, threadIdx.x, blockDim.x);
        phase_rows(T2, T2 + dw, T2 + 2 * dw, nw, t.delta_k[0], A.t[0], A.t[1], A.t[2], A.q, threadIdx.x, blockDim.x);
        phase_rows(T3, T3 + dw, T3 + 2 * dw, nw, t.delta_k[0], A.o[0], A.o[1], A.o[2], A.q, threadIdx.x, blockDim.x);
        __syncthreads();
        double part = 0;
        for (int e = threadIdx.x; e < t.Mw; e += blockDim.x) {
            const Cplx<T> tj = phase_at(T0, T0 + dw, T0 + 2 * dw, dw, e), oj = phase_at(T1, T1 + dw, T1 + 2 * dw, dw, e);
            const Cplx<T> ti = phase_at(T2, T2 + dw, T2 + 2 * dw, dw, e), oi = phase_at(T3, T3 + dw, T3 + 2 * dw, dw, e);
            const T ar = tj.x - oj.x, ai = tj.y - oj.y, br = ti.x - oi.x, bi = ti.y - oi.y;
            part += (double)((ar * br + ai * bi) * t.Wmat[e]);
        }
        total += block_sum(part, red) * 2.0 * t.cW;
    }
    // difference levels
    for (int ep = 0; ep < 2; ++ep) {
        const int lam = ep == 0 ? Bj.lam_t : Bj.lam_o;
        const double* xj = ep == 0 ? Bj.t : Bj.o;
        for (int l = 2; l <= lam - 1; ++l) {
            const int nl = 1 << l;
            const int node = ep == 0 ? Bj.path_t[l] : Bj.path_o[l];
            int shift[2][3];
            bool on[2];
            for (int w = 0; w < 2; ++w) {  // w = 0: i's target (+), 1: i's origin (-)
                const int lami = w == 0 ? A.lam_t : A.lam_o;
                const double* y = w == 0 ? A.t : A.o;
                on[w] = lami >= l;
                for (int ax = 0; ax < 3 && on[w]; ++ax) {
                    const int ci = cell_at(y[ax], t.L, l);
                    int dlt = ci - t.cell[3 * node + ax];
                    // periodic neighbour test; shift = image of y adjacent to B
                    if (dlt == 0 || dlt == 1 || dlt == -1) {
                        shift[w][ax] = 0;
                    } else if (dlt == nl - 1) {
                        shift[w][ax] = -1;
                    } else if (dlt == -(nl - 1)) {
                        shift[w][ax] = 1;
                    } else {
                        on[w] = false;
                    }
                }
            }
            if (!on[0] && !on[1]) continue;  // uniform across the block
            const int dd = t.dd, nd = t.n_diff, len = 2 * dd + nd + 1;
            Cplx<T>* Tj = tab;
            Cplx<T>* Ti = tab + len;
            Cplx<T>* Oi = tab + 2 * len;
            __syncthreads();
            phase_rows(Tj, Tj + dd, Tj + 2 * dd, nd, t.delta_k[l], xj[0], xj[1], xj[2], Bj.q, threadIdx.x, blockDim.x);
            if (on[0])
                phase_rows(Ti, Ti + dd, Ti + 2 * dd, nd, t.delta_k[l], A.t[0] + shift[0][0] * t.L,
                           A.t[1] + shift[0][1] * t.L, A.t[2] + shift[0][2] * t.L, A.q, threadIdx.x, blockDim.x);
            if (on[1])
                phase_rows(Oi, Oi + dd, Oi + 2 * dd, nd, t.delta_k[l], A.o[0] + shift[1][0] * t.L,
                           A.o[1] + shift[1][1] * t.L, A.o[2] + shift[1][2] * t.L, A.q, threadIdx.x, blockDim.x);
            __syncthreads();
            const T* D = t.Dmat + (size_t)(l - 2) * t.Md;
            double part = 0;
            for (int e = threadIdx.x; e < t.Md; e += blockDim.x) {
                const Cplx<T> w = phase_at(Tj, Tj + dd, Tj + 2 * dd, dd, e);
                // delta in = conj(tau_i) - conj(omega_i)
                T dr = 0, di = 0;
                if (on[0]) {
                    const Cplx<T> v = phase_at(Ti, Ti + dd, Ti + 2 * dd, dd, e);
                    dr += v.x;
                    di -= v.y;
                }
                if (on[1]) {
                    const Cplx<T> v = phase_at(Oi, Oi + dd, Oi + 2 * dd, dd, e);
                    dr -= v.x;
                    di += v.y;
                }
                part += (double)((w.x * dr - w.y * di) * D[e]);
            }
            total += block_sum(part, red) * 2.0 * t.cD[l] * (ep == 0 ? 1.0 : -1.0);
        }
    }
    if (threadIdx.x == 0) {
        // near field
        const T eps = std::numeric_limits<T>::epsilon();
        for (int ep = 0; ep < 2; ++ep) {
            const int lam = ep == 0 ? Bj.lam_t : Bj.lam_o;
            const double* xj = ep == 0 ? Bj.t : Bj.o;
            const double bb = t.boxsize[lam];
            const T b2 = (T)(bb * bb), inv_b = (T)(1.0 / bb);
            double s = 0;
            for (int w = 0; w < 2; ++w) {
                const double* y = w == 0 ? A.t : A.o;
                double dx = xj[0] - y[0], dy = xj[1] - y[1], dz = xj[2] - y[2];
                dx -= t.L * rint(dx / t.L);
                dy -= t.L * rint(dy / t.L);
                dz -= t.L * rint(dz / t.L);
                const T fx = (T)dx, fy = (T)dy, fz = (T)dz;
                const T r2 = fx * fx + fy * fy + fz * fz;
                if (r2 > eps && r2 < b2)
                    s += (w == 0 ? 1.0 : -1.0) * (double)residual_poly_over_r<T>(sqrt(r2), inv_b, t.poly, t.n_poly);
            }
            total += (ep == 0 ? 1.0 : -1.0) * s * A.q * Bj.q;
        }
        a.corr[i * K + j] = total;
    }
}

// Sequential Metropolis decisions of a batch: dE_j = stale_j + sum over
// accepted i < j of corr[i][j]; the leaf-capacity check follows the counts the
// earlier accepted moves of the batch leave behind.
template <typename T>
__global__ void __launch_bounds__(TB) k_scan(DevTree<T> t, ScanArgs a) {
    __shared__ double red[TB / 64];
    __shared__ int accs[kMaxBatch];
    __shared__ int delta;
    if (*t.err) return;
    const int K = a.K;
    for (int j = 0; j < K; ++j) {
        double part = 0;
        int dcount = 0;
        const TrialInfo& tj = a.info[j];
        const int lt = tj.path_t[tj.lam_t];
        for (int i = threadIdx.x; i < j; i += blockDim.x)
            if (accs[i]) {
                part += a.corr[i * K + j];
                const TrialInfo& ti = a.info[i];
                dcount += (ti.path_t[ti.lam_t] == lt) - (ti.path_o[ti.lam_o] == lt);
            }
        if (threadIdx.x == 0) delta = 0;
        __syncthreads();
        if (dcount) atomicAdd(&delta, dcount);
        const double s = block_sum(part, red);
        if (threadIdx.x == 0) {
            const int m = a.move0 + j;
            const double dE = a.dE_stale[m] + s;
            a.dE[m] = dE;
            int acc = a.u[m] < exp(-a.beta * dE);
            if (acc) {
                const int lo = tj.path_o[tj.lam_o];
                if (lt != lo && t.slot_count[lt] + delta >= t.slot_cap[lt]) {
                    atomicOr(t.err, (int)kErrLeafOverflow);
                    acc = 0;
                }
            }
            accs[j] = acc;
            a.accept[j] = acc;
            if (a.accept_out) a.accept_out[m] = acc;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ expansions
// outgoing root expansion, split over particle chunks; FP64 atomics into acc[2*Mw]
template <typename T>
__global__ void __launch_bounds__(TB) k_pw_root(DevTree<T> t, double* acc, int chunk) {
    constexpr int TILE = 16;
    __shared__ Cplx<T> tab[TILE][kRow];
    const int nw = t.n_window, dw = t.dw;
    const int e = blockIdx.x * TB + threadIdx.x;
    const int s0 = blockIdx.y * chunk, s1 = min(s0 + chunk, t.total_slots);
    double sr = 0, si = 0;
    for (int base = s0; base < s1; base += TILE) {
        __syncthreads();
        for (int p = threadIdx.x / 64; p < TILE; p += blockDim.x / 64) {
            const int s = base + p;
            const bool ok = s < s1;
            phase_rows(tab[p], tab[p] + dw, tab[p] + 2 * dw, nw, t.delta_k[0], ok ? t.px[s] : 0.0,
                       ok ? t.py[s] : 0.0, ok ? t.pz[s] : 0.0, ok ? t.pq[s] : 0.0, threadIdx.x & 63, 64);
        }
        __syncthreads();
        if (e < t.Mw) {
            const int i = e % dw, r = e / dw, j = r % dw, k = r / dw;
            for (int p = 0; p < TILE; ++p) {
                const Cplx<T> v = cmul(cmul(tab[p][i], tab[p][dw + j]), tab[p][2 * dw + k]);
                sr += (double)v.x;
                si += (double)v.y;
            }
        }
    }
    if (e < t.Mw) {
        atomicAdd(acc + 2 * e, sr);
        atomicAdd(acc + 2 * e + 1, si);
    }
}

template <typename T>
__global__ void k_store_root(DevTree<T> t, const double* acc) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < t.Mw) t.root[e] = {(T)acc[2 * e], (T)acc[2 * e + 1]};
}

// outgoing expansion of every box of levels 2..D-2 (src/planewave.cpp:48-84):
// grid (ceil(Md/TB), n_pw_nodes); the box's particles are the slot range of
// its descendant leaves (unused slots carry q = 0)
template <typename T>
__global__ void __launch_bounds__(TB) k_pw_level(DevTree<T> t, PwNodes pn) {
    constexpr int TILE = 16;
    __shared__ Cplx<T> tab[TILE][kRow];
    const int slot = blockIdx.y;
    const int l = pn.level[slot];
    const int nd = t.n_diff, dd = t.dd;
    const int e = blockIdx.x * TB + threadIdx.x;
    const int s0 = pn.lo[slot], s1 = pn.hi[slot];
    double sr = 0, si = 0;
    for (int base = s0; base < s1; base += TILE) {
        __syncthreads();
        for (int p = threadIdx.x / 64; p < TILE; p += blockDim.x / 64) {
            const int s = base + p;
            const bool ok = s < s1;
            phase_rows(tab[p], tab[p] + dd, tab[p] + 2 * dd, nd, t.delta_k[l], ok ? t.px[s] : 0.0,
                       ok ? t.py[s] : 0.0, ok ? t.pz[s] : 0.0, ok ? t.pq[s] : 0.0, threadIdx.x & 63, 64);
        }
        __syncthreads();
        if (e < t.Md) {
            const int i = e % dd, r = e / dd, j = r % dd, k = r / dd;
            for (int p = 0; p < TILE; ++p) {
                const Cplx<T> v = cmul(cmul(tab[p][i], tab[p][dd + j]), tab[p][2 * dd + k]);
                sr += (double)v.x;
                si += (double)v.y;
            }
        }
    }
    if (e < t.Md) t.out[(size_t)slot * t.Md + e] = {(T)sr, (T)si};
}

// incoming expansion of non-leaf boxes (src/planewave.cpp:99-155):
// in_B = conj(sum over B and its 26 colleagues of out_C s_p), p = image of C adjacent to B
template <typename T>
__global__ void __launch_bounds__(TB) k_incoming(DevTree<T> t, PwNodes pn) {
    __shared__ Cplx<T> sh[9 * 64];
    const int slot = blockIdx.y;
    const int node = pn.node[slot];
    if (t.is_leaf[node]) return;
    const int l = pn.level[slot];
    const int nd = t.n_diff, dd = t.dd;
    // sh[(3*axis + s) * dd + m] = e^{-i (m - n) dk_l (s-1) L}
    for (int q = threadIdx.x; q < 9 * dd; q += blockDim.x) {
        const int ax = q / (3 * dd), s = (q / dd) % 3, m = q % dd;
        (void)ax;
        double sn, cs;
        sincos((double)(m - nd) * (t.delta_k[l] * (s - 1) * t.L), &sn, &cs);
        sh[q] = {(T)cs, (T)(-sn)};
    }
    __syncthreads();
    const int e = blockIdx.x * TB + threadIdx.x;
    if (e >= t.Md) return;
    const int i = e % dd, r = e / dd, j = r % dd, k = r / dd;
    const Cplx<T> self = t.out[(size_t)slot * t.Md + e];
    double ar = (double)self.x, ai = (double)self.y;
    const int nl = 1 << l;
    int sidx = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!dx && !dy && !dz) continue;
                const int C = t.colleague[26 * node + sidx++];
                if (C < 0) continue;
                const int cx = t.cell[3 * node] + dx, cy = t.cell[3 * node + 1] + dy, cz = t.cell[3 * node + 2] + dz;
                const int px = cx < 0 ? -1 : (cx >= nl ? 1 : 0);
                const int py = cy < 0 ? -1 : (cy >= nl ? 1 : 0);
                const int pz = cz < 0 ? -1 : (cz >= nl ? 1 : 0);
                const Cplx<T> s = cmul(cmul(sh[(0 + px + 1) * dd + i], sh[(3 + py + 1) * dd + j]), sh[(6 + pz + 1) * dd + k]);
                const Cplx<T> v = cmul(t.out[(size_t)t.pw_slot[C] * t.Md + e], s);
                ar += (double)v.x;
                ai += (double)v.y;
            }
    t.inc[(size_t)slot * t.Md + e] = {(T)ar, (T)(-ai)};
}

// ------------------------------------------------------------------ energies
template <typename T>
__global__ void __launch_bounds__(TB) k_energy_window(DevTree<T> t, double* res) {
    __shared__ double red[TB / 64];
    const int e = blockIdx.x * TB + threadIdx.x;
    double v = 0;
    if (e < t.Mw) {
        const Cplx<T> g = t.root[e];
        v = ((double)g.x * g.x + (double)g.y * g.y) * (double)t.Wmat[e];
    }
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) atomicAdd(res + 0, s * t.cW);
}

template <typename T>
__global__ void __launch_bounds__(TB) k_energy_diff(DevTree<T> t, PwNodes pn, double* res) {
    __shared__ double red[TB / 64];
    const int slot = blockIdx.y;
    const int node = pn.node[slot];
    if (t.is_leaf[node]) return;
    const int l = pn.level[slot];
    const int e = blockIdx.x * TB + threadIdx.x;
    double v = 0;
    if (e < t.Md) {
        const Cplx<T> o = t.out[(size_t)slot * t.Md + e], in = t.inc[(size_t)slot * t.Md + e];
        v = ((double)o.x * in.x - (double)o.y * in.y) * (double)t.Dmat[(size_t)(l - 2) * t.Md + e] * t.cD[l];
    }
    if (blockIdx.x == 0)  // self-interaction term -Q_B S_l (src/energy.cpp:67-71)
        for (int s = pn.lo[slot] + threadIdx.x; s < pn.hi[slot]; s += blockDim.x) v -= t.pq[s] * t.pq[s] * t.self_diff[l];
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) atomicAdd(res + 1, s);
}

// residual energy (src/energy.cpp:82-173): grid (max_near, n_leaf_list);
// block = (leaf B, near entry); pairs counted from B's side with weight 1/2
template <typename T>
__global__ void __launch_bounds__(TB) k_energy_res(DevTree<T> t, const int* leaves, double* res) {
    __shared__ double red[TB / 64];
    const int leaf = leaves[blockIdx.y];
    const int k = blockIdx.x;
    const int off = t.near_off[leaf], cnt = t.near_off[leaf + 1] - off;
    if (k >= cnt) return;
    const int nb = t.near_ids[off + k];
    const int code = t.near_img[off + k];
    const double sx = (code % 3 - 1) * t.L, sy = ((code / 3) % 3 - 1) * t.L, sz = (code / 9 - 1) * t.L;
    const double bb = t.boxsize[t.level[leaf]];
    const T b2 = (T)(bb * bb), inv_b = (T)(1.0 / bb);
    const int a0 = t.slot_start[leaf], na = t.slot_count[leaf];
    const int c0 = t.slot_start[nb], nc = t.slot_count[nb];
    double v = 0;
    for (int p = threadIdx.x; p < na; p += blockDim.x) {
        const int s = a0 + p;
        const double x = t.px[s], y = t.py[s], z = t.pz[s];
        double acc = 0;
        for (int q = 0; q < nc; ++q) {
            const int s2 = c0 + q;
            if (s2 == s) continue;
            const T dx = (T)(x - (t.px[s2] + sx)), dy = (T)(y - (t.py[s2] + sy)), dz = (T)(z - (t.pz[s2] + sz));
            const T r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > T(0) && r2 < b2) acc += (double)((T)t.pq[s2] * residual_poly_over_r<T>(sqrt(r2), inv_b, t.poly, t.n_poly));
        }
        v += 0.5 * acc * t.pq[s];
    }
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) atomicAdd(res + 2, s);
}

// ------------------------------------------------------------------ launchers
// zero fill used instead of hipMemsetAsync, so that any of these sequences can
// be captured into a HIP graph without memset nodes
__global__ void k_zero(double* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

template <typename T>
void launch_trial(const DevTree<T>& t, const TrialArgs& a, int n_trials, hipStream_t st) {
    const int nWb = (t.Mw + TB - 1) / TB, nDb = (t.Md + TB - 1) / TB;
    const int nL = t.max_depth >= 4 ? t.max_depth - 3 : 0;
    const int per = nWb + 2 * nL * nDb + 2 * t.max_near;
    hipLaunchKernelGGL(k_trial<T>, dim3(per, n_trials), dim3(TB), 0, st, t, a);
}

template <typename T>
void launch_update(const DevTree<T>& t, const UpdateArgs& a, hipStream_t st) {
    const int nWb = (t.Mw + TB - 1) / TB, nDb = (t.Md + TB - 1) / TB;
    const int nL = t.max_depth >= 4 ? t.max_depth - 3 : 0;
    hipLaunchKernelGGL(k_update<T>, dim3(nWb + nL * nDb + 1), dim3(TB), 0, st, t, a);
}

template <typename T>
void launch_pair_corr(const DevTree<T>& t, const CorrArgs& a, hipStream_t st) {
    const int np = a.K * (a.K - 1) / 2;
    if (np > 0) hipLaunchKernelGGL(k_pair_corr<T>, dim3(np), dim3(TB), 0, st, t, a);
}

template <typename T>
void launch_scan(const DevTree<T>& t, const ScanArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_scan<T>, dim3(1), dim3(TB), 0, st, t, a);
}

template <typename T>
void launch_outgoing(const DevTree<T>& t, const PwNodes& pn, double* acc_root, hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3((2 * t.Mw + TB - 1) / TB), dim3(TB), 0, st, acc_root, 2 * t.Mw);
    const int chunk = 2048;
    const int nch = (t.total_slots + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_pw_root<T>, dim3((t.Mw + TB - 1) / TB, nch), dim3(TB), 0, st, t, acc_root, chunk);
    hipLaunchKernelGGL(k_store_root<T>, dim3((t.Mw + TB - 1) / TB), dim3(TB), 0, st, t, (const double*)acc_root);
    if (t.n_pw_nodes > 0)
        hipLaunchKernelGGL(k_pw_level<T>, dim3((t.Md + TB - 1) / TB, t.n_pw_nodes), dim3(TB), 0, st, t, pn);
}

template <typename T>
void launch_incoming(const DevTree<T>& t, const PwNodes& pn, hipStream_t st) {
    if (t.n_pw_nodes > 0)
        hipLaunchKernelGGL(k_incoming<T>, dim3((t.Md + TB - 1) / TB, t.n_pw_nodes), dim3(TB), 0, st, t, pn);
}

template <typename T>
void launch_energy(const DevTree<T>& t, const PwNodes& pn, const int* leaves, int n_leaves, double* res,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(TB), 0, st, res, 3);
    hipLaunchKernelGGL(k_energy_window<T>, dim3((t.Mw + TB - 1) / TB), dim3(TB), 0, st, t, res);
    if (t.n_pw_nodes > 0)
        hipLaunchKernelGGL(k_energy_diff<T>, dim3((t.Md + TB - 1) / TB, t.n_pw_nodes), dim3(TB), 0, st, t, pn, res);
    if (n_leaves > 0)
        hipLaunchKernelGGL(k_energy_res<T>, dim3(t.max_near, n_leaves), dim3(TB), 0, st, t, leaves, res);
}

#define INSTANTIATE(T)                                                                                     \
    template void launch_trial<T>(const DevTree<T>&, const TrialArgs&, int, hipStream_t);                \
    template void launch_update<T>(const DevTree<T>&, const UpdateArgs&, hipStream_t);                   \
    template void launch_pair_corr<T>(const DevTree<T>&, const CorrArgs&, hipStream_t);                  \
    template void launch_scan<T>(const DevTree<T>&, const ScanArgs&, hipStream_t);                       \
    template void launch_outgoing<T>(const DevTree<T>&, const PwNodes&, double*, hipStream_t);           \
    template void launch_incoming<T>(const DevTree<T>&, const PwNodes&, hipStream_t);                    \
    template void launch_energy<T>(const DevTree<T>&, const PwNodes&, const int*, int, double*, hipStream_t);
INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace hpdmk_amd
