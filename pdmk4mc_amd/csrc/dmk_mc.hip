// HIP kernels (gfx950, wave64) of the Monte Carlo step:
//   k_setup     per trial: endpoints, root-to-leaf paths, FP64 phase rows
//   k_trial     per trial: dE = window + difference + near field - self
//               (src/energy_shift.cpp:18-302); the last block assembles dE,
//               takes the Metropolis decision and (single move) plans the update
//   k_pair_corr per pair of trials of a batch: exact effect of an accepted
//               earlier move on a later trial
//   k_scan      one block: sequential decisions of a batch, update plan
//   k_apply     accept/update (src/tree.cpp:447-597) over the compact list of
//               rewritten arrays; every array element is read and written once
// T is the storage/arithmetic type of the bulk (double or float); phase
// arguments are formed in FP64 and every reduction accumulates in FP64.
#include "dmk_device.hpp"
#include "dmk_kernels.hpp"

namespace hpdmk_amd {

template <typename T>
__device__ __forceinline__ Cplx<T>* tabs_of(const McScratch& s, int stride, int tr) {
    return static_cast<Cplx<T>*>(s.tabs) + (size_t)tr * stride;
}

__device__ __forceinline__ int n_levels_pw(int D) { return D >= 4 ? D - 3 : 0; }

// periodic shift (-1, 0, 1) of a particle in box a so that it sits next to box
// b (same level, cells ca, cb); returns 2 when the boxes are not neighbours
__device__ __forceinline__ int image_shift(int ca, int cb, int nl) {
    const int dl = ca - cb;
    if (dl >= -1 && dl <= 1) return 0;
    if (dl == nl - 1) return -1;
    if (dl == -(nl - 1)) return 1;
    return 2;
}

// ------------------------------------------------------------------ setup
template <typename T>
__global__ void __launch_bounds__(TB) k_setup(DevTree<T> t, McArgs a) {
    __shared__ TrialInfo ti;
    if (*t.err) return;
    const int tr = blockIdx.x, m = a.move0 + tr;
    if (threadIdx.x == 0) {
        trial_setup(t, a.ml.idx[m], a.ml.d + 3 * m, ti);
        a.s.info[tr] = ti;
    }
    __syncthreads();
    const int D = t.max_depth, nL = n_levels_pw(D);
    const int stride = table_stride(D);
    Cplx<T>* tb = tabs_of<T>(a.s, stride, tr);
    const int rows = 6 + 2 * nL * 9;
    for (int w = threadIdx.x; w < rows * kRowStride; w += blockDim.x) {
        const int r = w / kRowStride, mm = w % kRowStride;
        int ep, axis, n, len;
        double c, dk;
        if (r < 6) {
            ep = r / 3;
            axis = r % 3;
            n = t.n_window;
            dk = t.delta_k[0];
            c = ep == 0 ? ti.t[axis] : ti.o[axis];
        } else {
            const int rr = r - 6;
            const int s1 = rr % 3;
            axis = (rr / 3) % 3;
            const int li = (rr / 9) % nL;
            ep = rr / (9 * nL);
            const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
            if (2 + li > lam) continue;  // this endpoint has no box at that level
            n = t.n_diff;
            dk = t.delta_k[2 + li];
            c = (ep == 0 ? ti.t[axis] : ti.o[axis]) + (s1 - 1) * t.L;
        }
        len = axis == 2 ? n + 1 : 2 * n + 1;
        if (mm >= len) continue;
        double sn, cs;
        sincos((double)(mm - n) * (dk * c), &sn, &cs);
        const double scale = axis == 0 ? ti.q : 1.0;
        tb[r * kRowStride + mm] = {(T)(cs * scale), (T)(-sn * scale)};
    }
}

// ------------------------------------------------------------------ update plan
// Marks every expansion array the accepted moves rewrite (out of each path box
// of levels 2..min(leaf, D-2), in of the non-leaf ones and of their non-leaf
// colleagues), compacts the marks into s.dest in slot order, and relocates
// the moved particles between leaf slots (thread 0, moves in order).  Run by
// one whole block.
template <typename T>
__device__ void plan_update(const DevTree<T>& t, const McArgs& a, int K) {
    __shared__ int wtot[TB / 64];
    __shared__ int base_sh;
    const int D = t.max_depth;
    for (int w = threadIdx.x; w < 2 * K; w += blockDim.x) {
        const int m = w >> 1, ep = w & 1;
        if (!a.s.accept[m]) continue;
        const TrialInfo& ti = a.s.info[m];
        const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
        const int* path = ep == 0 ? ti.path_t : ti.path_o;
        for (int l = 2; l <= lam && l <= D - 2; ++l) {
            const int node = path[l];
            const int sl = t.pw_slot[node];
            a.s.mark[2 * sl] = 1;
            if (!t.is_leaf[node]) a.s.mark[2 * sl + 1] = 1;
            for (int c = 0; c < 26; ++c) {
                const int C = t.colleague[26 * node + c];
                if (C >= 0 && !t.is_leaf[C]) a.s.mark[2 * t.pw_slot[C] + 1] = 1;
            }
        }
    }
    __syncthreads();
    const int n2 = 2 * t.n_pw_nodes;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (threadIdx.x == 0) base_sh = 0;
    __syncthreads();
    for (int c0 = 0; c0 < n2; c0 += blockDim.x) {
        const int idx = c0 + threadIdx.x;
        const bool f = idx < n2 && a.s.mark[idx];
        const unsigned long long bal = __ballot(f);
        const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const int pre = __popcll(bal & lt);
        if (lane == 0) wtot[wv] = __popcll(bal);
        __syncthreads();
        int off = base_sh;
        for (int k = 0; k < wv; ++k) off += wtot[k];
        if (f) {
            a.s.dest[off + pre] = idx;
            a.s.mark[idx] = 0;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < nw; ++k) base_sh += wtot[k];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    int any = 0;
    for (int m = 0; m < K; ++m) any |= a.s.accept[m];
    a.s.plan[0] = base_sh;
    a.s.plan[1] = any;
    // particle relocation (src/tree.cpp:454-463, 583-596), exact membership
    for (int m = 0; m < K; ++m) {
        if (!a.s.accept[m]) continue;
        const TrialInfo& ti = a.s.info[m];
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

// capacity check of an accepted move given the count change `delta` that
// earlier acceptances of the same batch leave in the target leaf
template <typename T>
__device__ int capacity_ok(const DevTree<T>& t, const TrialInfo& ti, int delta) {
    const int lt = ti.path_t[ti.lam_t], lo = ti.path_o[ti.lam_o];
    if (lt == lo) return 1;
    if (t.slot_count[lt] + delta < t.slot_cap[lt]) return 1;
    atomicOr(t.err, (int)kErrLeafOverflow);
    return 0;
}

// ------------------------------------------------------------------ trial
// grid (blocks per trial, K).  Block roles inside a trial:
//   [0, nWb)                          window chunk
//   [nWb, nWb + 2*nL*nDb)             difference chunk (endpoint, level, chunk)
//   [.., + 2*max_near)                near-field entry (endpoint, entry)
template <typename T>
__global__ void __launch_bounds__(TB) k_trial(DevTree<T> t, McArgs a) {
    __shared__ double red[TB / 64];
    __shared__ int last;
    if (*t.err) return;
    const int tr = blockIdx.y, b = blockIdx.x, m = a.move0 + tr;
    const TrialInfo& ti = a.s.info[tr];
    const int D = t.max_depth, nL = n_levels_pw(D);
    const Cplx<T>* tb = tabs_of<T>(a.s, table_stride(D), tr);
    const int nWb = (t.Mw + TB - 1) / TB;
    const int nDb = (t.Md + TB - 1) / TB;
    double part = 0.0;
    int slot_acc = 0, nscan = 0;
    if (b < nWb) {
        // window: 2 cW sum W Re[(tau - omega) conj(rho - omega)] (include/vecops.hpp:118-145)
        const int e = b * TB + threadIdx.x;
        if (e < t.Mw) {
            const int dw = t.dw;
            const Cplx<T> ta = phase_at(tb + root_row(0, 0), tb + root_row(0, 1), tb + root_row(0, 2), dw, e);
            const Cplx<T> om = phase_at(tb + root_row(1, 0), tb + root_row(1, 1), tb + root_row(1, 2), dw, e);
            const Cplx<T> g = t.root[e];
            const T ar = ta.x - om.x, ai = ta.y - om.y, gr = g.x - om.x, gi = g.y - om.y;
            part = (double)((ar * gr + ai * gi) * t.Wmat[e]) * (2.0 * t.cW);
        }
    } else if (b < nWb + 2 * nL * nDb) {
        // difference levels strictly above the endpoint's leaf (src/energy_shift.cpp:92-140)
        const int r = b - nWb;
        const int ep = r / (nL * nDb);  // 0 target, 1 origin
        const int li = (r / nDb) % nL;
        const int l = 2 + li;
        const int e = (r % nDb) * TB + threadIdx.x;
        const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
        slot_acc = 1;
        if (l <= lam - 1 && e < t.Md) {
            const int node = ep == 0 ? ti.path_t[l] : ti.path_o[l];
            const Cplx<T> v = t.inc[(size_t)t.pw_slot[node] * t.Md + e];
            const Cplx<T> w = phase_at(tb + level_row(nL, ep, li, 0, 0), tb + level_row(nL, ep, li, 1, 0),
                                       tb + level_row(nL, ep, li, 2, 0), t.dd, e);
            T re = w.x * v.x - w.y * v.y;
            if (ep == 1) re -= w.x * w.x + w.y * w.y;  // origin drops its own wave (vecops.hpp:161-163)
            part = (double)(re * t.Dmat[(size_t)li * t.Md + e]) * (2.0 * t.cD[l] * (ep == 0 ? 1.0 : -1.0));
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
            const double cx = ep == 0 ? ti.t[0] : ti.o[0], cy = ep == 0 ? ti.t[1] : ti.o[1],
                         cz = ep == 0 ? ti.t[2] : ti.o[2];
            const double bb = t.boxsize[lam];
            const T b2 = (T)(bb * bb), inv_b = (T)(1.0 / bb);
            const T eps = std::numeric_limits<T>::epsilon();
            const int s0 = t.slot_start[nb], n = t.slot_count[nb];
            nscan = n;
            double acc = 0;
            for (int p = threadIdx.x; p < n; p += blockDim.x) {
                const int s = s0 + p;
                const T dx = (T)(cx - (t.px[s] + sx)), dy = (T)(cy - (t.py[s] + sy)), dz = (T)(cz - (t.pz[s] + sz));
                const T r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > eps && r2 < b2)
                    acc += (double)((T)t.pq[s] * residual_poly_over_r<T>(sqrt(r2), inv_b, t.poly, t.n_poly));
            }
            part = acc * ti.q;
        }
    }
    const double s = block_sum(part, red);
    double* acc = a.s.acc + kAccPerTrial * tr;
    if (threadIdx.x == 0) {
        if (s != 0.0) atomicAdd(acc + slot_acc, s);
        if (nscan) atomicAdd(acc + 2 + slot_acc, (double)nscan);
        __threadfence();
        const unsigned prev = atomicAdd(a.s.done + tr, 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    if (threadIdx.x == 0) {
        __threadfence();
        double v[kAccPerTrial];
        for (int k = 0; k < kAccPerTrial; ++k)
            v[k] = __hip_atomic_load(acc + k, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const double self = self_term(t, ti);
        const double dE = v[0] + v[1] + v[2] - v[3] - self;
        double* dEout = a.mode == kTrialOnly ? a.ml.dE_stale : a.ml.dE;
        dEout[m] = dE;
        if (a.mode == kTrialOnly && a.ml.dE) a.ml.dE[m] = dE;
        if (a.ml.parts) {
            double* p = a.ml.parts + 5 * m;
            p[0] = v[0];
            p[1] = v[1];
            p[2] = v[2];
            p[3] = v[3];
            p[4] = self;
        }
        if (a.ml.stats) {
            int* st = a.ml.stats + kStatsPerMove * m;
            st[0] = ti.lam_t - 2 > 0 ? ti.lam_t - 2 : 0;
            st[1] = ti.lam_o - 2 > 0 ? ti.lam_o - 2 : 0;
            st[2] = (int)v[4];
            st[3] = (int)v[5];
            for (int ep = 0; ep < 2; ++ep) {  // arrays an acceptance rewrites (src/tree.cpp:484-579)
                const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
                int arrays = 0;
                for (int l = 2; l <= lam && l <= D - 2; ++l) {
                    const int node = ep == 0 ? ti.path_t[l] : ti.path_o[l];
                    arrays += t.is_leaf[node] ? 1 : 2;
                    for (int c = 0; c < 26; ++c) {
                        const int C = t.colleague[26 * node + c];
                        arrays += (C >= 0 && !t.is_leaf[C]);
                    }
                }
                st[4 + ep] = arrays;
            }
        }
        int ok = 0;
        if (a.mode == kTrialMetropolis) ok = a.ml.u[m] < exp(-a.beta * dE);
        if (a.mode == kTrialForceAccept) ok = 1;
        if (ok) ok = capacity_ok(t, ti, 0);
        a.s.accept[tr] = ok;
        if (a.ml.accept_out) a.ml.accept_out[m] = ok;
        for (int k = 0; k < kAccPerTrial; ++k) acc[k] = 0.0;  // ready for the next launch
        a.s.done[tr] = 0;
    }
    if (a.mode == kTrialOnly) return;
    __syncthreads();
    plan_update(t, a, 1);  // single-move step: this block plans the update
}

// ------------------------------------------------------------------ batched trials
// Effect of accepted move i on trial j (i < j, distinct particles): exactly
// what the update of i changes in the quantities trial j reads
//   window:     2 cW sum W Re[(tau_j - omega_j) conj(tau_i - omega_i)]
//   difference: in[B] of j's path box B at level l changes by conj(tau_i s)
//               when i's target box at level l is B or a neighbour of B
//               (s = phase of the image of t_i next to B), minus the same for o_i
//   near field: +-q_i q_j R_b(|x_j - y|) for y in {t_i, o_i}, b of x_j's leaf
template <typename T>
__global__ void __launch_bounds__(TB) k_pair_corr(DevTree<T> t, McArgs a) {
    __shared__ double red[TB / 64];
    const int K = a.K;
    int p = blockIdx.x, i = 0;
    while (p >= K - 1 - i) {
        p -= K - 1 - i;
        ++i;
    }
    const int j = i + 1 + p;
    if (*t.err) return;
    const TrialInfo& A = a.s.info[i];
    const TrialInfo& Bj = a.s.info[j];
    const int D = t.max_depth, nL = n_levels_pw(D), stride = table_stride(D);
    const Cplx<T>* ta = tabs_of<T>(a.s, stride, i);
    const Cplx<T>* tj = tabs_of<T>(a.s, stride, j);
    double total = 0;
    {
        const int dw = t.dw;
        double part = 0;
        for (int e = threadIdx.x; e < t.Mw; e += blockDim.x) {
            const Cplx<T> x1 = phase_at(tj + root_row(0, 0), tj + root_row(0, 1), tj + root_row(0, 2), dw, e);
            const Cplx<T> x2 = phase_at(tj + root_row(1, 0), tj + root_row(1, 1), tj + root_row(1, 2), dw, e);
            const Cplx<T> y1 = phase_at(ta + root_row(0, 0), ta + root_row(0, 1), ta + root_row(0, 2), dw, e);
            const Cplx<T> y2 = phase_at(ta + root_row(1, 0), ta + root_row(1, 1), ta + root_row(1, 2), dw, e);
            const T ar = x1.x - x2.x, ai = x1.y - x2.y, br = y1.x - y2.x, bi = y1.y - y2.y;
            part += (double)((ar * br + ai * bi) * t.Wmat[e]);
        }
        total += block_sum(part, red) * 2.0 * t.cW;
    }
    for (int ep = 0; ep < 2; ++ep) {
        const int lam = ep == 0 ? Bj.lam_t : Bj.lam_o;
        for (int l = 2; l <= lam - 1; ++l) {
            const int li = l - 2, nl = 1 << l;
            const int node = ep == 0 ? Bj.path_t[l] : Bj.path_o[l];
            int sh[2][3];
            bool on[2];
            for (int w = 0; w < 2; ++w) {  // w = 0: i's target (+), 1: i's origin (-)
                const int lami = w == 0 ? A.lam_t : A.lam_o;
                on[w] = lami >= l;
                if (!on[w]) continue;
                const int bi = w == 0 ? A.path_t[l] : A.path_o[l];
                for (int ax = 0; ax < 3 && on[w]; ++ax) {
                    sh[w][ax] = image_shift(t.cell[3 * bi + ax], t.cell[3 * node + ax], nl);
                    if (sh[w][ax] == 2) on[w] = false;
                }
            }
            if (!on[0] && !on[1]) continue;  // uniform across the block
            const T* Dl = t.Dmat + (size_t)li * t.Md;
            const int dd = t.dd;
            double part = 0;
            for (int e = threadIdx.x; e < t.Md; e += blockDim.x) {
                const Cplx<T> w = phase_at(tj + level_row(nL, ep, li, 0, 0), tj + level_row(nL, ep, li, 1, 0),
                                           tj + level_row(nL, ep, li, 2, 0), dd, e);
                T dr = 0, di = 0;  // delta in = conj(tau_i s) - conj(omega_i s')
                for (int wv = 0; wv < 2; ++wv) {
                    if (!on[wv]) continue;
                    const Cplx<T> v = phase_at(ta + level_row(nL, wv, li, 0, sh[wv][0]),
                                               ta + level_row(nL, wv, li, 1, sh[wv][1]),
                                               ta + level_row(nL, wv, li, 2, sh[wv][2]), dd, e);
                    const T sg = wv == 0 ? T(1) : T(-1);
                    dr += sg * v.x;
                    di -= sg * v.y;
                }
                part += (double)((w.x * dr - w.y * di) * Dl[e]);
            }
            total += block_sum(part, red) * 2.0 * t.cD[l] * (ep == 0 ? 1.0 : -1.0);
        }
    }
    if (threadIdx.x == 0) {
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
        a.s.corr[i * K + j] = total;
    }
}

// Sequential Metropolis decisions of a batch: dE_j = stale_j + sum over the
// accepted i < j of corr[i][j]; then the update plan of the whole batch.
template <typename T>
__global__ void __launch_bounds__(TB) k_scan(DevTree<T> t, McArgs a) {
    __shared__ double red[TB / 64];
    __shared__ int delta;
    if (*t.err) return;
    const int K = a.K;
    for (int j = 0; j < K; ++j) {
        double part = 0;
        int dcount = 0;
        const TrialInfo& tj = a.s.info[j];
        const int lt = tj.path_t[tj.lam_t];
        for (int i = threadIdx.x; i < j; i += blockDim.x)
            if (a.s.accept[i]) {
                part += a.s.corr[i * K + j];
                const TrialInfo& ti = a.s.info[i];
                dcount += (ti.path_t[ti.lam_t] == lt) - (ti.path_o[ti.lam_o] == lt);
            }
        if (threadIdx.x == 0) delta = 0;
        __syncthreads();
        if (dcount) atomicAdd(&delta, dcount);
        const double s = block_sum(part, red);
        if (threadIdx.x == 0) {
            const int m = a.move0 + j;
            const double dE = a.ml.dE_stale[m] + s;
            a.ml.dE[m] = dE;
            int ok = a.ml.u[m] < exp(-a.beta * dE);
            if (ok) ok = capacity_ok(t, tj, delta);
            a.s.accept[j] = ok;
            if (a.ml.accept_out) a.ml.accept_out[m] = ok;
        }
        __syncthreads();
    }
    plan_update(t, a, K);
}

// ------------------------------------------------------------------ apply
// Work items: root chunks, then (rewritten array, element chunk).  For an
// array the block first lists the (move, endpoint, image) contributions in move
// order, then every thread adds them to its element and writes it back once.
template <typename T>
__global__ void __launch_bounds__(TB) k_apply(DevTree<T> t, McArgs a) {
    __shared__ int mlist[2 * kMaxBatch];
    __shared__ int wtot[TB / 64];
    __shared__ int mcount;
    if (*t.err) return;
    const int nD = a.s.plan[0];
    if (!a.s.plan[1]) return;
    const int K = a.K, D = t.max_depth, nL = n_levels_pw(D), stride = table_stride(D);
    const int nWb = (t.Mw + TB - 1) / TB, nDb = (t.Md + TB - 1) / TB;
    const int total = nWb + nD * nDb;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        if (item < nWb) {  // root += tau - omega (src/tree.cpp:479-480)
            const int e = item * TB + threadIdx.x;
            if (e >= t.Mw) continue;
            T ar = 0, ai = 0;
            for (int m = 0; m < K; ++m) {
                if (!a.s.accept[m]) continue;
                const Cplx<T>* tb = tabs_of<T>(a.s, stride, m);
                const Cplx<T> ta = phase_at(tb + root_row(0, 0), tb + root_row(0, 1), tb + root_row(0, 2), t.dw, e);
                const Cplx<T> om = phase_at(tb + root_row(1, 0), tb + root_row(1, 1), tb + root_row(1, 2), t.dw, e);
                ar += ta.x - om.x;
                ai += ta.y - om.y;
            }
            Cplx<T> g = t.root[e];
            g.x += ar;
            g.y += ai;
            t.root[e] = g;
            continue;
        }
        const int r = item - nWb;
        const int key = a.s.dest[r / nDb];
        const int slot = key >> 1, kind = key & 1;
        const int node = a.pn.node[slot];
        const int l = a.pn.level[slot], li = l - 2, nl = 1 << l;
        __syncthreads();  // mlist of the previous item fully consumed
        if (threadIdx.x == 0) mcount = 0;
        __syncthreads();
        for (int c0 = 0; c0 < 2 * K; c0 += blockDim.x) {
            const int w = c0 + threadIdx.x;
            int code = -1;
            if (w < 2 * K) {
                const int m = w >> 1, ep = w & 1;
                if (a.s.accept[m]) {
                    const TrialInfo& ti = a.s.info[m];
                    const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
                    if (l <= lam) {
                        const int B = ep == 0 ? ti.path_t[l] : ti.path_o[l];
                        if (B == node) {
                            code = 13;
                        } else if (kind == 1) {
                            int px = image_shift(t.cell[3 * B], t.cell[3 * node], nl);
                            int py = image_shift(t.cell[3 * B + 1], t.cell[3 * node + 1], nl);
                            int pz = image_shift(t.cell[3 * B + 2], t.cell[3 * node + 2], nl);
                            if (px != 2 && py != 2 && pz != 2) code = (px + 1) + 3 * (py + 1) + 9 * (pz + 1);
                        }
                    }
                }
                if (code >= 0) code |= (m << 8) | (ep << 7);
            }
            const bool f = code >= 0;
            const unsigned long long bal = __ballot(f);
            const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
            if (lane == 0) wtot[wv] = __popcll(bal);
            __syncthreads();
            int off = mcount;
            for (int k = 0; k < wv; ++k) off += wtot[k];
            if (f) mlist[off + __popcll(bal & lt)] = code;
            __syncthreads();
            if (threadIdx.x == 0)
                for (int k = 0; k < (int)(blockDim.x >> 6); ++k) mcount += wtot[k];
            __syncthreads();
        }
        const int e = (r % nDb) * TB + threadIdx.x;
        if (e >= t.Md) continue;
        T ar = 0, ai = 0;
        const int nm = mcount;
        for (int k = 0; k < nm; ++k) {
            const int c = mlist[k];
            const int m = c >> 8, ep = (c >> 7) & 1, code = c & 127;
            const int px = code % 3 - 1, py = (code / 3) % 3 - 1, pz = code / 9 - 1;
            const Cplx<T>* tb = tabs_of<T>(a.s, stride, m);
            const Cplx<T> v = phase_at(tb + level_row(nL, ep, li, 0, px), tb + level_row(nL, ep, li, 1, py),
                                       tb + level_row(nL, ep, li, 2, pz), t.dd, e);
            const T sg = ep == 0 ? T(1) : T(-1);
            ar += sg * v.x;
            ai += sg * v.y;
        }
        const size_t at = (size_t)slot * t.Md + e;
        if (kind == 0) {  // out[B] += tau
            Cplx<T> o = t.out[at];
            o.x += ar;
            o.y += ai;
            t.out[at] = o;
        } else {  // in[C] += conj(tau s)
            Cplx<T> o = t.inc[at];
            o.x += ar;
            o.y -= ai;
            t.inc[at] = o;
        }
    }
}

// ------------------------------------------------------------------ launchers
template <typename T>
void launch_setup(const DevTree<T>& t, const McArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_setup<T>, dim3(a.K), dim3(TB), 0, st, t, a);
}

template <typename T>
void launch_trial(const DevTree<T>& t, const McArgs& a, hipStream_t st) {
    const int nWb = (t.Mw + TB - 1) / TB, nDb = (t.Md + TB - 1) / TB;
    const int nL = t.max_depth >= 4 ? t.max_depth - 3 : 0;
    const int per = nWb + 2 * nL * nDb + 2 * t.max_near;
    hipLaunchKernelGGL(k_trial<T>, dim3(per, a.K), dim3(TB), 0, st, t, a);
}

template <typename T>
void launch_pair_corr(const DevTree<T>& t, const McArgs& a, hipStream_t st) {
    const int np = a.K * (a.K - 1) / 2;
    if (np > 0) hipLaunchKernelGGL(k_pair_corr<T>, dim3(np), dim3(TB), 0, st, t, a);
}

template <typename T>
void launch_scan(const DevTree<T>& t, const McArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_scan<T>, dim3(1), dim3(TB), 0, st, t, a);
}

template <typename T>
void launch_apply(const DevTree<T>& t, const McArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_apply<T>, dim3(1024), dim3(TB), 0, st, t, a);
}

#define INSTANTIATE(T)                                                                \
    template void launch_setup<T>(const DevTree<T>&, const McArgs&, hipStream_t);     \
    template void launch_trial<T>(const DevTree<T>&, const McArgs&, hipStream_t);     \
    template void launch_pair_corr<T>(const DevTree<T>&, const McArgs&, hipStream_t); \
    template void launch_scan<T>(const DevTree<T>&, const McArgs&, hipStream_t);      \
    template void launch_apply<T>(const DevTree<T>&, const McArgs&, hipStream_t);
INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace hpdmk_amd
