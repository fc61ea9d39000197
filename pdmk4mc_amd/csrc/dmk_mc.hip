// HIP kernels (gfx950, wave64) of the Monte Carlo step:
//   k_setup     per trial: endpoints, root-to-leaf paths, FP64 phase rows
//   k_propose_setup  the same after the proposal: features at the origin and,
//               learned, the MLP displacement
//   k_trial     per trial: partial sums of dE = window + difference + near
//               field - self (src/energy_shift.cpp:18-302), one per block;
//               for a batch also the window pair terms (Gram blocks, MFMA);
//               for a single move also the update plan's marking (last row)
//   k_decide    single move: assembles dE, decides, moves the particle
//   k_pair_corr one block per trial assembles its record (dE, threshold with
//               the learned proposal's reverse-move log-ratio, leaf room);
//               then per pair of trials of a batch: exact effect of an
//               accepted earlier move on a later trial
//   k_scan      one block: sequential decisions of a batch, update plan
//   k_apply     accept/update (src/tree.cpp:447-597) over the compact list of
//               rewritten arrays; every array element is read and written
//               once (FP32 sums on the matrix cores); for a batch, block 0
//               moves the particles
//   k_compact   periodic squeeze of the holes moves leave in leaf slots
// T is the storage/arithmetic type of the bulk (double or float); phase
// arguments are formed in FP64 and every reduction accumulates in FP64.
#include <hip/hip_ext.h>

#include <stdexcept>

#include "dmk_device.hpp"
#include "dmk_kernels.hpp"

namespace hpdmk_amd {


template <typename T>
__device__ __forceinline__ Cplx<T>* tabs_of(const McScratch& s, int stride, int tr) {
    return static_cast<Cplx<T>*>(s.tabs) + (size_t)tr * stride;
}

__device__ __forceinline__ int n_levels_pw(int D) { return D >= 4 ? D - 3 : 0; }

// Metropolis test u < exp(-beta dE) written as dE < -ln(u) / beta, one
// comparison shared by the single-move and the batched decisions
__device__ __forceinline__ double mh_threshold(double u, double beta) {
    if (beta <= 0) return INFINITY;
    return u > 0 ? -log(u) / beta : INFINITY;
}

__device__ __forceinline__ void probe_add(const McScratch& s, int k, unsigned long long& t0) {
    if (!s.probe) return;
    const unsigned long long t1 = wall_clock64();
    atomicAdd(s.probe + k, t1 - t0);
    atomicAdd(s.probe + kProbes + k, 1ull);
    t0 = t1;
}

// periodic shift (-1, 0, 1) of a particle in box a so that it sits next to box
// b (same level, cells ca, cb); returns 2 when the boxes are not neighbours
__device__ __forceinline__ int image_shift(int ca, int cb, int nl) {
    const int dl = ca - cb;
    if (dl >= -1 && dl <= 1) return 0;
    if (dl == nl - 1) return -1;
    if (dl == -(nl - 1)) return 1;
    return 2;
}

// a box cell below 1024 per axis in one int (x | y << 10 | z << 20)
__device__ __forceinline__ int pack_cell(int cx, int cy, int cz) { return cx | (cy << 10) | (cz << 20); }

// first move of this launch: the step counter when one is attached
__device__ __forceinline__ int move_base(const McArgs& a) { return (a.move_ctr ? *a.move_ctr : 0) + a.move0; }
// trials of this launch: with a step counter, the moves left (at most K)
__device__ __forceinline__ int keff(const McArgs& a) {
    return a.move_ctr ? min(a.K, a.n_moves - move_base(a)) : a.K;
}
// scratch slot (TrialInfo, phase rows) of trial tr of move m: the ring slot
// m & 1 in the fused single-move chain, else the trial's index
__device__ __forceinline__ int trial_slot(const McArgs& a, int tr, int m) { return a.ring ? (m & 1) : tr; }
// the fused chain's ring slot of this launch's move (no counter load)
__device__ __forceinline__ int ring_slot(const McArgs& a) { return a.move0 & 1; }

// ------------------------------------------------------------------ setup
// FP64-formed phase rows of a trial into tb, by the whole block.  Row r of
// grid half-width n holds scale e^{-i (k - n) dk c} for k = 0..2n (the z row
// stops at k_z = 0; the x row carries the charge).  Two phases: one FP64
// sincos per row gives its base e^{-i dk c} (in LDS); then one thread per
// (row, run of kRowRun powers) raises the base to the run's first power by
// squaring and multiplies on (FP64 complex arithmetic throughout).
constexpr int kRowRun = 4;
constexpr int kMaxRows = 6 + 2 * (kMaxLevels - 3) * 9;
template <typename T>
__device__ void form_rows(const DevTree<T>& t, const TrialInfo& ti, Cplx<T>* tb) {
    __shared__ double bre[kMaxRows], bim[kMaxRows], bsc[kMaxRows];
    __shared__ int bn[kMaxRows];
    const int D = t.max_depth, nL = n_levels_pw(D);
    const int rows = 6 + 2 * nL * 9;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
        int ep, axis, n;
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
            n = 2 + li > lam ? -1 : t.n_diff;  // -1: this endpoint has no box at that level
            dk = t.delta_k[2 + li];
            c = (ep == 0 ? ti.t[axis] : ti.o[axis]) + (s1 - 1) * t.L;
        }
        double sn, cs;
        sincos(dk * c, &sn, &cs);
        bre[r] = cs;
        bim[r] = -sn;  // e^{-i dk c}
        bsc[r] = axis == 0 ? ti.q : 1.0;
        bn[r] = n;
    }
    __syncthreads();
    const int runs = (kMaxGridD / 2 + 1 + kRowRun - 1) / kRowRun;  // powers 0..n of the widest grid
    for (int w = threadIdx.x; w < rows * runs; w += blockDim.x) {
        const int r = w / runs, k0 = (w % runs) * kRowRun, n = bn[r];
        if (k0 > n) continue;
        const double br = bre[r], bi = bim[r];
        // base^k0 by squaring
        double pr = 1.0, pi = 0.0, qr = br, qi = bi;
        for (int e = k0; e > 0; e >>= 1) {
            if (e & 1) {
                const double tr = pr * qr - pi * qi;
                pi = pr * qi + pi * qr;
                pr = tr;
            }
            const double sr = qr * qr - qi * qi;
            qi = 2.0 * qr * qi;
            qr = sr;
        }
        double wr = bsc[r] * pr, wi = bsc[r] * pi;
        Cplx<T>* row = tb + r * kRowStride;
        const bool zrow = r < 6 ? r % 3 == 2 : ((r - 6) / 3) % 3 == 2;
        for (int k = k0; k < k0 + kRowRun && k <= n; ++k) {
            if (k == 0) {
                row[n] = {(T)wr, (T)0};
            } else {
                if (!zrow) row[n + k] = {(T)wr, (T)wi};  // the z row stops at k_z = 0
                row[n - k] = {(T)wr, (T)(-wi)};
            }
            const double tr = wr * br - wi * bi;
            wi = wr * bi + wi * br;
            wr = tr;
        }
    }
}

// trial state of move m (paths, endpoints) into slot sl, by thread 0 of the
// block; the caller syncs, then form_rows
template <typename T>
__device__ __forceinline__ void setup_move(const DevTree<T>& t, const McArgs& a, int m, int sl, TrialInfo& ti) {
    trial_setup(t, a.ml.idx[m], a.ml.d + 3 * m, ti);
    a.s.info[sl] = ti;
    if (a.ml.stats) {  // filled by k_pair_corr (pair terms) and the update planner (batch arrays)
        a.ml.stats[kStatsPerMove * m + 6] = 0;
        a.ml.stats[kStatsPerMove * m + 7] = 0;
    }
}

// Fused single-move chain: the whole trial state of move m in ring slot sl,
// by one block: paths, the path boxes' expansion slots, the endpoints' near
// lists and the phase rows.  prev (may be null): the particle was moved by
// the previous move and sits at prev->t (its leaf path is prev's target path).
// nm (may be null): the move's particle as prefetched (position after the
// previous move), so that only the paths remain to be found.
// rows: also the self-pair term, the threshold and the phase rows (else the
// update kernel's row block forms them: upd1_rows)
template <typename T>
__device__ void prep_body(const DevTree<T>& t, const McArgs& a, int m, int sl, const TrialInfo* prev,
                          const NextMove* nm = nullptr, bool rows = true) {
    __shared__ TrialInfo ti;
    const int D = t.max_depth;
    if (threadIdx.x == 0) {
        if (nm && !prev) {
            ti.i = nm->i;
            ti.q = nm->q;
            for (int c = 0; c < 3; ++c) {
                ti.o[c] = nm->o[c];
                ti.dv[c] = nm->dv[c];
                ti.t[c] = wrap(ti.o[c] + ti.dv[c], t.L);
            }
            ti.lam_o = locate(t, ti.o[0], ti.o[1], ti.o[2], ti.path_o);
            ti.lam_t = locate(t, ti.t[0], ti.t[1], ti.t[2], ti.path_t);
        } else if (prev) {
            ti.i = prev->i;
            ti.q = prev->q;
            for (int c = 0; c < 3; ++c) {
                ti.o[c] = prev->t[c];
                ti.dv[c] = a.ml.d[3 * m + c];
                ti.t[c] = wrap(ti.o[c] + ti.dv[c], t.L);
            }
            ti.lam_o = prev->lam_t;
            for (int l = 0; l < kMaxLevels; ++l) ti.path_o[l] = prev->path_t[l];
            ti.lam_t = locate(t, ti.t[0], ti.t[1], ti.t[2], ti.path_t);
        } else {
            trial_setup(t, a.ml.idx[m], a.ml.d + 3 * m, ti);
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * kMaxLevels) {
        const int ep = threadIdx.x / kMaxLevels, l = threadIdx.x % kMaxLevels;
        const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
        ti.pws[ep][l] = (l >= 2 && l <= lam && l <= D - 2) ? t.pw_slot[ep == 0 ? ti.path_t[l] : ti.path_o[l]] : -1;
    } else if (threadIdx.x == 64 || threadIdx.x == 65) {
        const int ep = threadIdx.x - 64;
        const int leaf = ep == 0 ? ti.path_t[ti.lam_t] : ti.path_o[ti.lam_o];
        const int o0 = t.near_off[leaf];
        ti.noff[ep] = o0;
        ti.ncnt[ep] = t.near_off[leaf + 1] - o0;
    }
    __syncthreads();
    // every thread knows the rows slot itself (thread 0's writes below are
    // not waited for before the rows are formed)
    const int rslot = prev ? 2 + (m & 1) : (m & 1);  // a redo keeps off the slot the row block may be writing
    if (threadIdx.x == 0) {
        ti.m = m;
        ti.u = nm && !prev ? nm->u : a.ml.u[m];
        ti.rs = rslot;
        if (rows) a.s.st[rslot] = make_double2(self_term(t, ti), mh_threshold(ti.u, a.beta));
        a.s.info[sl] = ti;
        if (a.ml.stats) {
            a.ml.stats[kStatsPerMove * m + 6] = 0;
            a.ml.stats[kStatsPerMove * m + 7] = 0;
        }
    }
    if (rows) form_rows(t, ti, tabs_of<T>(a.s, table_stride(D), rslot));
}

// Fused chain: the near-field entries of the move in ring slot sl (its
// TrialInfo written), by a whole block: slot range and image code of every
// near-list leaf of both endpoints.  Counts are read now; with moved >= 0 the
// caller has just appended a particle to leaf `moved` (count + 1) in the
// same kernel.
// Layout of each (slot, endpoint) list: max_near + 1 int4, the first holding
// the entry count (so a reader needs no other state to find its entries).
template <typename T>
__device__ void fill_near(const DevTree<T>& t, const McArgs& a, const TrialInfo& ti, int sl, int moved, int moved_cnt) {
    const int mx = t.max_near;
    if (threadIdx.x < 2) a.s.near[(size_t)(2 * sl + threadIdx.x) * (mx + 1)] = make_int4(ti.ncnt[threadIdx.x], 0, 0, 0);
    for (int w = threadIdx.x; w < 2 * mx; w += blockDim.x) {
        const int ep = w / mx, k = w % mx;
        if (k >= ti.ncnt[ep]) continue;
        const int nb = t.near_ids[ti.noff[ep] + k];
        const int code = t.near_img[ti.noff[ep] + k];
        const int cnt = nb == moved ? moved_cnt : t.slot_count[nb];
        a.s.near[(size_t)(2 * sl + ep) * (mx + 1) + 1 + k] = make_int4(t.slot_start[nb], cnt, code, nb);
    }
}

// the record of move m (its particle from the current state) into nm slot sl
template <typename T>
__device__ __forceinline__ void load_next_move(const DevTree<T>& t, const McArgs& a, int m, NextMove& r) {
    const int i = a.ml.idx[m], s = t.slot_of[i];
    r.i = i;
    r.m = m;
    r.q = t.pq[s];
    r.o[0] = t.px[s];
    r.o[1] = t.py[s];
    r.o[2] = t.pz[s];
    for (int c = 0; c < 3; ++c) r.dv[c] = a.ml.d[3 * m + c];
    r.u = a.ml.u[m];
}

template <typename T>
__device__ __forceinline__ void setup_body(const DevTree<T>& t, const McArgs& a, int tr) {
    __shared__ TrialInfo ti;
    if (*t.err) return;
    const int m = move_base(a) + tr;
    if (tr >= keff(a)) return;
    if (a.ring) {  // a run's first move (or the next after a compaction) and its successor's particle
        if (threadIdx.x == 64 && m + 1 < a.n_moves) {
            NextMove r;
            load_next_move(t, a, m + 1, r);
            a.s.nm[(m + 1) & 1] = r;
        }
        prep_body(t, a, m, m & 1, nullptr);
        __syncthreads();
        fill_near(t, a, a.s.info[m & 1], m & 1, -1, 0);
        return;
    }
    const int sl = trial_slot(a, tr, m);
    if (threadIdx.x == 0) setup_move(t, a, m, sl, ti);
    __syncthreads();
    form_rows(t, ti, tabs_of<T>(a.s, table_stride(t.max_depth), sl));
}

template <typename T>
__global__ void __launch_bounds__(TB) k_setup(DevTree<T> t, McArgs a) {
    setup_body(t, a, blockIdx.x);
}

// one trial per replica: replica blockIdx.z
template <typename T>
__global__ void __launch_bounds__(TB) k_setup_rep(const DevTree<T>* __restrict__ ts, const McArgs* __restrict__ as) {
    const DevTree<T>& t = ts[blockIdx.z];
    const McArgs& a = as[blockIdx.z];
    setup_body(t, a, blockIdx.x);
}

// ------------------------------------------------------------------ propose
// LDS of the proposal (propose_body's big, in bytes): the packed weights,
// the biases, then the near leaves' image shifts of the neighbour pass, whose
// space the feature partial sums take over for the reduction
constexpr int kProposeWpsBytes = 2 * kPackSize;
constexpr int kProposeBshBytes = (4 * kBiasPack + 15) / 16 * 16;
constexpr int kProposeLds = kProposeWpsBytes + kProposeBshBytes + 8 * TB * (kFeatSums > 3 ? kFeatSums : 3);
static_assert(kProposeWpsBytes % 16 == 0, "16-byte aligned regions");

// Features of each trial's particle at its origin (particle removed) and, for
// the learned proposal, its displacement delta = mu + sigma z and log q;
// into ti (shared, by thread 0) the trial's particle, origin and
// displacement, and (by thread 64, while the features are summed) the
// origin's root-to-leaf path.  One block per trial; returns false (uniformly)
// when the block has no trial (tr: the trial of this block).
template <typename T>
__device__ __forceinline__ bool propose_body(const DevTree<T>& t, const McArgs& a, TrialInfo& ti, int tr, char* big) {
    __shared__ double phi[kFeat], mu[3], sig[3];
    // the learned network's packed weights and biases, staged while the
    // features are summed, and the near leaves' image shifts, whose space the
    // feature partial sums take over after the neighbour pass (kProposeLds bytes of big)
    unsigned short* wps = reinterpret_cast<unsigned short*>(big);
    float* bsh = reinterpret_cast<float*>(big + kProposeWpsBytes);
    double (*rsh)[3] = reinterpret_cast<double (*)[3]>(big + kProposeWpsBytes + kProposeBshBytes);
    double (*fpart)[TB] = reinterpret_cast<double (*)[TB]>(big + kProposeWpsBytes + kProposeBshBytes);
    if (*t.err) return false;
    const Proposal& P = a.prop;
    const int m = move_base(a) + tr;
    if (tr >= keff(a)) return false;
    const int i = a.ml.idx[m];
    const int own = t.slot_of[i], leaf = t.leaf_of[i];
    const double ox = t.px[own], oy = t.py[own], oz = t.pz[own], q = t.pq[own];
    if (threadIdx.x == 0) {
        ti.i = i;
        ti.q = q;
        ti.o[0] = ox;
        ti.o[1] = oy;
        ti.o[2] = oz;
        if (!P.learned)
            for (int c = 0; c < 3; ++c) ti.dv[c] = a.ml.d[3 * m + c];
    } else if (threadIdx.x == 64) {
        ti.lam_o = locate(t, ox, oy, oz, ti.path_o);
    }
    const double rf2 = P.rf * P.rf;
    double fs[kFeatSums] = {0, 0, 0, 0, 0, 0, 0};
    const int off = t.near_off[leaf], cnt = t.near_off[leaf + 1] - off;
    // near-list leaves whose box (at its image) reaches into the ball of
    // radius rf: tested TB entries at a time in parallel, the reaching ones
    // compacted in list order; then their particles in one flat parallel loop
    __shared__ int rs0[TB], rn[TB], rpre[TB + 1], wcnt[TB / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int base = 0; base < cnt; base += TB) {
        const int k = base + threadIdx.x;
        bool reach = false;
        int s0 = 0, n = 0;
        double sx = 0, sy = 0, sz = 0;
        if (k < cnt) {
            const int nb = t.near_ids[off + k];
            const int code = t.near_img[off + k];
            sx = (code % 3 - 1) * t.L;
            sy = ((code / 3) % 3 - 1) * t.L;
            sz = (code / 9 - 1) * t.L;
            // the leaf's slots are loaded with its level and cell (used only
            // when it reaches the ball), and its size is L / 2^level exactly
            // (the boxsize table's value), so one round of loads decides
            const int lv = t.level[nb], ss = t.slot_start[nb], sn = t.slot_count[nb];
            const double bs = ldexp(t.L, -lv);
            const double lo[3] = {t.cell[3 * nb] * bs + sx, t.cell[3 * nb + 1] * bs + sy, t.cell[3 * nb + 2] * bs + sz};
            const double x[3] = {ox, oy, oz};
            double g2 = 0;
            for (int c = 0; c < 3; ++c) {
                const double e = fmax(fmax(lo[c] - x[c], x[c] - (lo[c] + bs)), 0.0);
                g2 += e * e;
            }
            reach = g2 < rf2;
            if (reach) {
                s0 = ss;
                n = sn;
            }
        }
        const unsigned long long bal = __ballot(reach);
        const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        if (lane == 0) wcnt[wv] = __popcll(bal);
        __syncthreads();
        int pos = __popcll(bal & lt);
        for (int w = 0; w < wv; ++w) pos += wcnt[w];
        if (reach) {
            rs0[pos] = s0;
            rn[pos] = n;
            rsh[pos][0] = sx;
            rsh[pos][1] = sy;
            rsh[pos][2] = sz;
        }
        __syncthreads();
        int nr = 0;
        for (int w = 0; w < TB / 64; ++w) nr += wcnt[w];
        if (threadIdx.x == 0) {
            rpre[0] = 0;
            for (int r = 0; r < nr; ++r) rpre[r + 1] = rpre[r] + rn[r];
        }
        __syncthreads();
        const int total = rpre[nr];
        for (int u = threadIdx.x; u < total; u += blockDim.x) {
            int r = 0;
            while (rpre[r + 1] <= u) ++r;  // nr is small (the leaves around one point)
            const int s = rs0[r] + (u - rpre[r]);
            const double ex = ox - (t.px[s] + rsh[r][0]), ey = oy - (t.py[s] + rsh[r][1]),
                         ez = oz - (t.pz[s] + rsh[r][2]);
            const double r2 = ex * ex + ey * ey + ez * ez;
            if (r2 > 0.0 && r2 < rf2 && s != own && t.porig[s] >= 0) feature_pair(ex, ey, ez, r2, t.pq[s], P.rf, fs);
        }
        __syncthreads();  // the lists are consumed before the next pass
    }
    // staged now: the weights' loads land while the features are reduced
    // (issued earlier, they would delay the feature loads, which wait in order)
    if (P.learned) mlp_stage(P, wps, bsh);
    // the feature sums: every thread's partial sums through LDS, kFeatSums x 8
    // threads add 32 each, kFeatSums threads the eight; then one thread per feature
    __shared__ double fp8[kFeatSums][8], fsum[kFeatSums];
#pragma unroll
    for (int c = 0; c < kFeatSums; ++c) fpart[c][threadIdx.x] = fs[c];
    __syncthreads();
    if (threadIdx.x < kFeatSums * 8) {
        const int c = threadIdx.x / 8, k = threadIdx.x % 8;
        double v = 0.0;
        for (int u = 0; u < TB / 8; ++u) v += fpart[c][k * (TB / 8) + u];
        fp8[c][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kFeatSums) {
        double v = 0.0;
        for (int k = 0; k < 8; ++k) v += fp8[threadIdx.x][k];
        fsum[threadIdx.x] = v;
    }
    __syncthreads();
    if (threadIdx.x < kFeat) {
        phi[threadIdx.x] = feature_from_sums(threadIdx.x, q, fsum);
        if (P.feat) P.feat[(size_t)kFeat * m + threadIdx.x] = phi[threadIdx.x];
    }
    if (!P.learned) return true;  // uniform across the grid
    __syncthreads();
    mlp_forward_mfma(P, phi, mu, sig, wps, bsh);
    __syncthreads();
    if (threadIdx.x == 0) {
        double dl[3];
        for (int c = 0; c < 3; ++c) {
            dl[c] = mu[c] + sig[c] * P.z[3 * m + c];
            P.dprop[3 * m + c] = dl[c];
            ti.dv[c] = dl[c];
        }
        P.logq[m] = gauss_logq(dl, mu, sig);
    }
    return true;
}

// the proposal (or the features) and the set-up of each trial in one launch:
// the set-up completes the trial state the proposal began (the target and its
// path), then forms the phase rows (src/energy_shift.cpp:18-51)
template <typename T>
__device__ __forceinline__ void propose_setup_body(const DevTree<T>& t, const McArgs& a, int tr, char* big) {
    __shared__ TrialInfo ti;
    if (!propose_body(t, a, ti, tr, big)) return;
    __syncthreads();
    const int m = move_base(a) + tr, sl = trial_slot(a, tr, m);
    if (threadIdx.x == 0) {
        for (int c = 0; c < 3; ++c) ti.t[c] = wrap(ti.o[c] + ti.dv[c], t.L);
        ti.lam_t = locate(t, ti.t[0], ti.t[1], ti.t[2], ti.path_t);
        a.s.info[sl] = ti;
        if (a.ml.stats) {  // filled by k_pair_corr (pair terms) and the update planner (batch arrays)
            a.ml.stats[kStatsPerMove * m + 6] = 0;
            a.ml.stats[kStatsPerMove * m + 7] = 0;
        }
    }
    __syncthreads();
    form_rows(t, ti, tabs_of<T>(a.s, table_stride(t.max_depth), sl));
}

template <typename T>
__global__ void __launch_bounds__(TB) k_propose_setup(DevTree<T> t, McArgs a) {
    __shared__ __align__(16) char big[kProposeLds];
    propose_setup_body(t, a, blockIdx.x, big);
}

// ------------------------------------------------------------------ update plan
// Moves an accepted particle whose leaf changes: it is appended to the
// target leaf at slot_count + rank (rank = earlier movers of the batch into the
// same leaf) and its old slot becomes a hole (q = 0, porig = -1) until the
// next compaction (k_compact).
template <typename T>
__device__ void relocate_to(const DevTree<T>& t, const TrialInfo& ti, int lt, int ns) {
    const int i = ti.i;
    const int s = t.slot_of[i];
    t.px[ns] = ti.t[0];
    t.py[ns] = ti.t[1];
    t.pz[ns] = ti.t[2];
    t.pq[ns] = t.pq[s];
    t.porig[ns] = i;
    t.slot_of[i] = ns;
    t.leaf_of[i] = lt;
    t.px[s] = 0;
    t.py[s] = 0;
    t.pz[s] = 0;
    t.pq[s] = 0;
    t.porig[s] = -1;
}
template <typename T>
__device__ void relocate(const DevTree<T>& t, const TrialInfo& ti, int rank) {
    const int lt = ti.path_t[ti.lam_t];
    relocate_to(t, ti, lt, t.slot_start[lt] + t.slot_count[lt] + rank);
}

// Particle moves of an accepted batch of at most 64 trials (src/tree.cpp:454-463,
// 583-596), run by one wave: lane m takes trial m; a move inside its leaf
// rewrites the position in place, a mover gets the next free slot of its
// target leaf in move order (its rank among the batch's movers into that
// leaf, by readlane), and the first mover into a leaf adds the count.
template <typename T>
__device__ void relocate_batch(const DevTree<T>& t, const McArgs& a) {
    const int m = threadIdx.x;
    if (m >= 64) return;
    const bool in = m < a.K && a.s.accept[m];
    const TrialInfo& ti = a.s.info[in ? m : 0];
    const int lo = in ? ti.path_o[ti.lam_o] : -1, lt = in ? ti.path_t[ti.lam_t] : -1;
    const bool mover = in && lo != lt;
    const int key = mover ? lt : -1;
    int rank = 0, total = 0;
#pragma unroll 2  // few scalar registers live at once (the update kernel's loop keeps its own)
    for (int k = 0; k < 64; ++k) {
        const int kk = __builtin_amdgcn_readlane(key, k);
        if (mover && kk == key) {
            rank += k < m;
            ++total;
        }
    }
    if (in && !mover) {
        const int s = t.slot_of[ti.i];
        t.px[s] = ti.t[0];
        t.py[s] = ti.t[1];
        t.pz[s] = ti.t[2];
    }
    if (mover) {
        const int cnt = t.slot_count[lt];  // every mover's load precedes the count's store (data dependence)
        relocate_to(t, ti, lt, t.slot_start[lt] + cnt + rank);
        if (rank == 0) t.slot_count[lt] = cnt + total;
    }
}

// Marks the arrays one accepted endpoint rewrites at level l (out of its box,
// in of the box when it is not a leaf and of its non-leaf colleagues) in the
// shared-memory bitmap; the first marker of an array appends it to s.dest: the
// incoming arrays of level 2 (each neighbours a quarter of the level's boxes,
// so a batch's moves pile up there) from the front, the others from the back.
// The order inside each group is immaterial: every array is rewritten by its
// own work items.
template <typename T>
__device__ __forceinline__ void mark_arrays(const DevTree<T>& t, const McArgs& a, const TrialInfo& ti, int ep, int l,
                                            unsigned* bits, int* nfront, int* nback, int n2) {
    if (l > (ep == 0 ? ti.lam_t : ti.lam_o)) return;
    const int node = ep == 0 ? ti.path_t[l] : ti.path_o[l];
    const int own = t.pw_slot[node];
    const bool front = l == 2;
    // the 28 candidate arrays (own out, own in, colleagues' in): every bit is
    // set first (independent LDS atomics in flight together), then the first
    // markers append
    int ix[28];
    ix[0] = 2 * own;
    ix[1] = t.is_leaf[node] ? -1 : 2 * own + 1;
#pragma unroll
    for (int c = 0; c < 26; ++c) {
        const int sc = t.col_in_slot[26 * node + c];
        ix[2 + c] = sc >= 0 ? 2 * sc + 1 : -1;
    }
    unsigned old[28];
#pragma unroll
    for (int c = 0; c < 28; ++c) old[c] = ix[c] >= 0 ? atomicOr(&bits[ix[c] >> 5], 1u << (ix[c] & 31)) : ~0u;
#pragma unroll
    for (int c = 0; c < 28; ++c) {
        if (ix[c] < 0 || (old[c] & (1u << (ix[c] & 31)))) continue;
        if (c > 0 && front)
            a.s.dest[atomicAdd(nfront, 1)] = ix[c];
        else
            a.s.dest[n2 - 1 - atomicAdd(nback, 1)] = ix[c];
    }
}

// Single-move steps (premark): the update plan's arrays, marked by the trial
// kernel's last row before the decision as if the move were accepted (they
// depend on the two paths only); a rejection leaves the accepted count at
// zero.  One block; the bitmap is the launch's dynamic shared memory,
// (2 n_pw + 31) / 32 words, so that the trial blocks keep their occupancy.
template <typename T>
__device__ void mark_single(const DevTree<T>& t, const McArgs& a) {
    extern __shared__ unsigned mark_bits[];
    __shared__ int nfront, nback;
    const int nL = n_levels_pw(t.max_depth), n2 = 2 * t.n_pw_nodes;
    unsigned long long tp = wall_clock64();
    for (int w = threadIdx.x; w < (n2 + 31) / 32; w += blockDim.x) mark_bits[w] = 0u;
    if (threadIdx.x == 0) nfront = nback = 0;
    __syncthreads();
    // one thread per (endpoint, level, candidate array): few registers in the
    // trial kernel, every candidate's loads in flight together
    const TrialInfo& ti = a.s.info[a.ring ? ring_slot(a) : 0];
    if (threadIdx.x == 0) {  // room in the target leaf (read before the update kernel moves anything)
        const int lt = ti.path_t[ti.lam_t], lo = ti.path_o[ti.lam_o];
        a.s.plan[3] = lt == lo || t.slot_count[lt] < t.slot_cap[lt];
    }
    for (int w = threadIdx.x; w < 2 * nL * 28; w += blockDim.x) {
        const int c = w % 28, el = w / 28, ep = el / nL, l = 2 + el % nL;
        if (l > (ep == 0 ? ti.lam_t : ti.lam_o)) continue;
        const int node = ep == 0 ? ti.path_t[l] : ti.path_o[l];
        int ix;
        if (c == 0) {
            ix = 2 * t.pw_slot[node];
        } else if (c == 1) {
            ix = t.is_leaf[node] ? -1 : 2 * t.pw_slot[node] + 1;
        } else {
            const int sc = t.col_in_slot[26 * node + c - 2];
            ix = sc >= 0 ? 2 * sc + 1 : -1;
        }
        if (ix < 0 || (atomicOr(&mark_bits[ix >> 5], 1u << (ix & 31)) & (1u << (ix & 31)))) continue;
        if (a.ring) {
            // fused chain: one list in marking order, with the array box's
            // level and cell (the endpoint's cell at level l, shifted by the
            // colleague offset and wrapped) so that the update kernel finds
            // every contribution without table loads
            const double* x = ep == 0 ? ti.t : ti.o;
            const int nl = 1 << l;
            int cc[3] = {cell_at(x[0], t.L, l), cell_at(x[1], t.L, l), cell_at(x[2], t.L, l)};
            if (c >= 2) {
                const int j = c - 2 < 13 ? c - 2 : c - 1;  // offset (dx+1) + 3(dy+1) + 9(dz+1), 13 skipped
                cc[0] = (cc[0] + j % 3 - 1 + nl) % nl;
                cc[1] = (cc[1] + (j / 3) % 3 - 1 + nl) % nl;
                cc[2] = (cc[2] + j / 9 - 1 + nl) % nl;
            }
            const int pos = atomicAdd(&nfront, 1);
            a.s.dest[pos] = ix;
            a.s.dinfo[pos] = l | (cc[0] << 4) | (cc[1] << 13) | (cc[2] << 22);
        } else if (c > 0 && l == 2) {
            a.s.dest[atomicAdd(&nfront, 1)] = ix;
        } else {
            a.s.dest[n2 - 1 - atomicAdd(&nback, 1)] = ix;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.s.plan[0] = nfront + nback;
        a.s.plan[2] = nfront;
        probe_add(a.s, kPrPlanMark, tp);
    }
}

// The update plan of a step: marks every array the accepted moves rewrite
// (mark_arrays) and relocates the moved particles between leaf slots.  Run by
// one whole block.  MARKED: mark_single already marked the arrays (and wrote
// their counts); this block only moves the particle and writes the accepted
// count.
template <typename T, bool MARKED = false>
__device__ void plan_update(const DevTree<T>& t, const McArgs& a, int K, bool reloc = true) {
    __shared__ unsigned bits[MARKED ? 1 : kMarkWords];
    __shared__ int nfront, nback;
    __shared__ int lo_s[kMaxBatch], lt_s[kMaxBatch], ac_s[kMaxBatch], conflict[kMaxBatch];
    const int D = t.max_depth, nL = n_levels_pw(D);
    const int n2 = 2 * t.n_pw_nodes;  // at most 32 kMarkWords (checked at set-up)
    unsigned long long tp = wall_clock64();
    if (!MARKED)
        for (int w = threadIdx.x; w < (n2 + 31) / 32; w += blockDim.x) bits[w] = 0u;
    if (threadIdx.x == 0) nfront = nback = 0;
    for (int m = threadIdx.x; m < K; m += blockDim.x) {
        const TrialInfo& ti = a.s.info[m];
        ac_s[m] = a.s.accept[m];
        lo_s[m] = ti.path_o[ti.lam_o];
        lt_s[m] = ti.path_t[ti.lam_t];
    }
    __syncthreads();
    // one thread per (move, endpoint, level): own out, own in, colleagues' in
    for (int w = threadIdx.x; w < (MARKED ? 0 : 2 * K * nL); w += blockDim.x) {
        const int li = w % nL, me = w / nL, m = me >> 1, ep = me & 1;
        if (ac_s[m]) mark_arrays(t, a, a.s.info[m], ep, 2 + li, bits, &nfront, &nback, n2);
    }
    // particles that stay in their leaf move in place, all at once
    for (int m = threadIdx.x; m < K && reloc; m += blockDim.x) {
        if (!ac_s[m] || lt_s[m] != lo_s[m]) continue;
        const TrialInfo& ti = a.s.info[m];
        const int s = t.slot_of[ti.i];
        t.px[s] = ti.t[0];
        t.py[s] = ti.t[1];
        t.pz[s] = ti.t[2];
    }
    __syncthreads();
    if (threadIdx.x == 0 && !MARKED) probe_add(a.s, kPrPlanMark, tp);
    const int base_sh = MARKED ? a.s.plan[2] : nfront;
    const int base2_sh = MARKED ? a.s.plan[0] - base_sh : nback;
    if (threadIdx.x == 0) probe_add(a.s, kPrPlanCompact, tp);
    // leaf changes (src/tree.cpp:454-463, 583-596), all in parallel: each mover
    // gets the next free slots of its target leaf in move order (unless the
    // update kernel moves the particles: relocate_batch)
    if (!reloc) {
    } else if (K > 8 && K <= 64) {  // one wave: lane m compares its target leaf with every lane's by readlane
        if (threadIdx.x < 64) {
            const int m = threadIdx.x;
            const bool mover = m < K && ac_s[m] && lo_s[m] != lt_s[m];
            const int key = mover ? lt_s[m] : -1;
            int rank = 0, total = 0;
#pragma unroll
            for (int k = 0; k < 64; ++k) {
                const int kk = __builtin_amdgcn_readlane(key, k);
                if (mover && kk == key) {
                    rank += k < m;
                    ++total;
                }
            }
            if (m < K) conflict[m] = mover ? (rank == 0 ? total : 0) : -1;  // first mover carries the count
            if (mover) relocate(t, a.s.info[m], rank);
        }
    } else {
        for (int m = threadIdx.x; m < K; m += blockDim.x) {
            int rank = 0, total = 0;
            const bool mover = ac_s[m] && lo_s[m] != lt_s[m];
            if (mover)
                for (int k = 0; k < K; ++k)
                    if (ac_s[k] && lo_s[k] != lt_s[k] && lt_s[k] == lt_s[m]) {
                        rank += k < m;
                        ++total;
                    }
            conflict[m] = mover ? (rank == 0 ? total : 0) : -1;  // first mover into a leaf carries the count
            if (mover) relocate(t, a.s.info[m], rank);
        }
    }
    __syncthreads();
    for (int m = threadIdx.x; m < K && reloc; m += blockDim.x)
        if (conflict[m] > 0) t.slot_count[lt_s[m]] += conflict[m];
    if (threadIdx.x != 0) return;
    int nacc = 0;
    for (int m = 0; m < K; ++m) nacc += ac_s[m] != 0;
    if (!MARKED) {
        a.s.plan[0] = base_sh + base2_sh;
        a.s.plan[2] = base_sh;
    }
    a.s.plan[1] = nacc;
    if (a.ml.stats && nacc) a.ml.stats[kStatsPerMove * move_base(a) + 7] = base_sh + base2_sh;
    probe_add(a.s, kPrPlanRelocate, tp);
}

// capacity check of an accepted move given the count change `delta` that
// earlier acceptances of the same batch leave in the target leaf
template <typename T>
__device__ int capacity_ok(const DevTree<T>& t, const TrialInfo& ti, int delta) {
    const int lt = ti.path_t[ti.lam_t], lo = ti.path_o[ti.lam_o];
    if (lt == lo) return 1;
    if (t.slot_count[lt] + delta < t.slot_cap[lt]) return 1;  // holes keep their slots until compaction
    atomicOr(t.err, (int)kErrLeafOverflow);
    return 0;
}

// ------------------------------------------------------------------ trial
// Sums of the per-block partials of trial tr: window, difference, res_t,
// res_o, particles scanned at t and o, hard-sphere overlaps at t.  The gs
// lanes of a contiguous lane group (g0 = rank in the group, gs a power of two
// <= 64) take the blocks g0, g0 + gs, ... and join their sums with shuffles:
// every lane of the group ends with the totals.  Callers make the partials
// visible first (kernel boundary); plain loads then see them.
template <typename T>
__device__ void sum_trial_partials(const DevTree<T>& t, const McArgs& a, int tr, int g0, int gs, double* v) {
    const Tiles& g = a.tiles;
    const int D = t.max_depth, nL = n_levels_pw(D), per = trial_blocks(D, g);
    const double* p = a.s.part + kPartStride * (size_t)tr * per;
    const int d0 = g.win, d1 = d0 + 2 * nL * g.diff, d2 = d1 + g.near;
#pragma unroll
    for (int c = 0; c < kAccPerTrial; ++c) v[c] = 0.0;
    constexpr int U = 8;  // blocks whose loads are in flight together
    for (int b0 = g0; b0 < per; b0 += U * gs) {
        double q0[U], q1[U], q2[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + u * gs;
            q0[u] = q1[u] = q2[u] = 0.0;
            if (b < per) {
                const double* q = p + kPartStride * b;
                q0[u] = q[0];
                q1[u] = q[1];
                q2[u] = q[2];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + u * gs;
            if (b >= per) continue;
            if (b < d0) {
                v[0] += q0[u];
            } else if (b < d1) {
                v[1] += q0[u];
            } else if (b < d2) {
                v[2] += q0[u];
                v[4] += q1[u];
                v[6] += q2[u];
            } else {
                v[3] += q0[u];
                v[5] += q1[u];
            }
        }
    }
    for (int o = 1; o < gs; o <<= 1)
#pragma unroll
        for (int c = 0; c < kAccPerTrial; ++c) v[c] += __shfl_xor(v[c], o);
}

// feature sums of the target's near field (learned proposal) from the
// partials, by a lane group as above
template <typename T>
__device__ void sum_target_features(const DevTree<T>& t, const McArgs& a, int tr, int g0, int gs, double* fs) {
    const Tiles& g = a.tiles;
    const int nL = n_levels_pw(t.max_depth), per = trial_blocks(t.max_depth, g);
    const double* p = a.s.part + kPartStride * (size_t)tr * per;
    const int d1 = g.win + 2 * nL * g.diff;
#pragma unroll
    for (int c = 0; c < kFeatSums; ++c) fs[c] = 0.0;
    for (int b = g0; b < g.near; b += gs)
#pragma unroll
        for (int c = 0; c < kFeatSums; ++c) fs[c] += p[kPartStride * (d1 + b) + 3 + c];
    for (int o = 1; o < gs; o <<= 1)
#pragma unroll
        for (int c = 0; c < kFeatSums; ++c) fs[c] += __shfl_xor(fs[c], o);
}

// dE of trial tr from its partial sums v (src/energy_shift.cpp:67) plus the
// optional per-move parts and counts; *overlaps = particles within sigma of
// the target
template <typename T>
__device__ double assemble_trial(const DevTree<T>& t, const McArgs& a, int tr, int m, const double* v,
                                 int* overlaps) {
    const TrialInfo& ti = a.s.info[tr];
    *overlaps = (int)v[6];
    const double self = self_term(t, ti);
    const double dE = v[0] + v[1] + v[2] - v[3] - self;
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
            for (int l = 2; l <= lam && l <= t.max_depth - 2; ++l) {
                const int node = ep == 0 ? ti.path_t[l] : ti.path_o[l];
                arrays += t.is_leaf[node] ? 1 : 2;
                for (int c = 0; c < 26; ++c) arrays += t.col_in_slot[26 * node + c] >= 0;
            }
            st[4 + ep] = arrays;
        }
    }
    return dE;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Window part of the pair corrections of a batch on the matrix cores:
//   G[i][j] = sum_k W_k Re(Phi_j(k) conj Phi_i(k)),
// Phi = tau - omega of a trial (its root-level phases at the target minus
// those at the origin, charge in the x-row), over the Mw window modes: the
// window term of trial j (include/vecops.hpp:118-145) taken against the change
// an accepted move i makes to the root expansion.  Block g takes modes [kGramModes g, kGramModes
// (g + 1)) in chunks of kGramSW: each chunk's Phi of every trial goes to LDS
// as real rows (index r = 2 m + part), and every 16 x 16 tile i <= j is a
// chain of v_mfma_f32_16x16x4f32 with A[i][r] = P_i(r), B[r][j] = W P_j(r);
// each chunk's FP32 tile is added into FP64.  gram[(i G + g) K + j] holds
// block g's share of pair i < j (a tile row's 16 columns are one 128-byte
// store); the pair blocks add the G shares in order.
__device__ void gram_body(const DevTree<float>& t, const McArgs& a, int g) {
    constexpr int RS = 2 * kGramSW + 4;  // LDS row stride in floats: 16 rows land 4 banks apart
    __shared__ float P[kGramMaxK * RS];
    __shared__ float Wc[kGramSW];
    const int G = gram_blocks(t.Mw), K = a.K, Ke = keff(a);
    if (g >= G || Ke < 2) return;  // uniform; no pair without two trials
    const int stride = table_stride(t.max_depth), dw = t.dw;
    const int nt = (K + 15) / 16, ntile = nt * (nt + 1) / 2;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, li = lane & 15, kq = lane >> 4;
    // this wave's tiles: wv, wv + 4, wv + 8 of the upper triangle (it <= jt)
    int tI[3], tJ[3];
    bool tv[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        int idx = wv + 4 * u, it = 0;
        tv[u] = idx < ntile;
        while (idx >= nt - it && it < nt) {
            idx -= nt - it;
            ++it;
        }
        tI[u] = it;
        tJ[u] = it + idx;
    }
    double sum[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
    for (int c = 0; c < kGramChunks; ++c) {
        const int mb = g * kGramModes + c * kGramSW;
        if (mb >= t.Mw) break;  // uniform
        __syncthreads();        // the previous chunk is consumed
        // every load of the chunk in flight at once: indices clamped into
        // range, out-of-range entries zeroed after the loads
#pragma unroll
        for (int it = 0; it < kGramMaxK * kGramSW / TB; ++it) {
            const int e = threadIdx.x + it * TB;
            const int tr = e / kGramSW, mm = e % kGramSW, m = mb + mm;
            const Cplx<float>* tb = tabs_of<float>(a.s, stride, min(tr, Ke - 1));
            const int mc = min(m, t.Mw - 1);
            const Cplx<float> ta = phase_at(tb + root_row(0, 0), tb + root_row(0, 1), tb + root_row(0, 2), dw, mc);
            const Cplx<float> om = phase_at(tb + root_row(1, 0), tb + root_row(1, 1), tb + root_row(1, 2), dw, mc);
            const bool ok = tr < Ke && m < t.Mw;
            if (tr < nt * 16) {
                P[tr * RS + 2 * mm] = ok ? ta.x - om.x : 0.f;
                P[tr * RS + 2 * mm + 1] = ok ? ta.y - om.y : 0.f;
            }
        }
        if (threadIdx.x < kGramSW) {
            const int m = mb + threadIdx.x;
            Wc[threadIdx.x] = m < t.Mw ? t.Wmat[m] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            if (!tv[u]) continue;  // uniform per wave
            f32x4 d = {0.f, 0.f, 0.f, 0.f};
            const float* pa = P + (tI[u] * 16 + li) * RS;
            const float* pb = P + (tJ[u] * 16 + li) * RS;
#pragma unroll
            for (int ks = 0; ks < 2 * kGramSW / 4; ++ks) {
                const int r = 4 * ks + kq;
                d = __builtin_amdgcn_mfma_f32_16x16x4f32(pa[r], pb[r] * Wc[r >> 1], d, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) sum[u][r] += (double)d[r];
        }
    }
    // D layout of v_mfma_f32_16x16x4f32: lane holds rows 4 kq + r of column li
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        if (!tv[u]) continue;
        const int gj = tJ[u] * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gi = tI[u] * 16 + 4 * kq + r;
            if (gi < gj && gj < Ke) a.s.gram[((size_t)gi * G + g) * K + gj] = sum[u][r];
        }
    }
}

// grid (trial_blocks(D, tiles), gram_rows + K).  With use_gram the first
// gram_rows rows are Gram blocks (window pair terms, above); then per trial,
// block roles (each strides over its share):
//   [0, win)                               window
//   [.., + 2*nL*diff)                      difference (endpoint, level, share)
//   [.., + 2*near)                         near field (endpoint, share of the near list)
// Every block stores its partial sums; k_decide (single moves) or k_scan
// (batches) assembles them after the kernel ends.
template <typename T, bool PM = false>
__device__ __forceinline__ void trial_body(const DevTree<T>& t, const McArgs& a) {
    __shared__ double red[TB / 64];
    if (*t.err) return;
    const int gy = gram_rows(t, a);
    if ((int)blockIdx.y < gy) {
        if constexpr (std::is_same<T, float>::value) gram_body(t, a, blockIdx.y * gridDim.x + blockIdx.x);
        return;
    }
    if constexpr (PM) {  // premark(a): the update plan's marking, before the decision
        if ((int)blockIdx.y == gy + a.K) {
            if (blockIdx.x == 0 && keff(a) > 0) mark_single(t, a);
            return;
        }
        // fused chain: the next move's trial state from the current state (the
        // update kernel redoes it when this move relocates the same particle)
        if ((int)blockIdx.y == gy + a.K + 1) {
            const int mn = a.s.info[ring_slot(a)].m + 1;
            if (blockIdx.x == 0 && mn < a.n_moves) prep_body(t, a, mn, mn & 1, nullptr, &a.s.nm[mn & 1], false);
            return;
        }
    }
    const int tr = blockIdx.y - gy, b = blockIdx.x;
    // the fused chain's graphs hold valid moves only, and its slot needs no counter
    if (!a.ring && tr >= keff(a)) return;
    const int sl = a.ring ? ring_slot(a) : tr;
    const TrialInfo& ti = a.s.info[sl];
    const int D = t.max_depth, nL = n_levels_pw(D), per = trial_blocks(D, a.tiles);
    const Cplx<T>* tb = tabs_of<T>(a.s, table_stride(D), a.ring ? ti.rs : sl);
    double part = 0.0;
    int nscan = 0, novl = 0;
    double fsum[kFeatSums] = {0, 0, 0, 0, 0, 0, 0};
    const int WB = a.tiles.win, DB = a.tiles.diff, NB = a.tiles.near;
    if (b < WB) {
        // window: 2 cW sum W Re[(tau - omega) conj(rho - omega)] (include/vecops.hpp:118-145)
        const int dw = t.dw;
#pragma unroll 4
        for (int e = b * TB + threadIdx.x; e < t.Mw; e += WB * TB) {
            const Cplx<T> ta = phase_at(tb + root_row(0, 0), tb + root_row(0, 1), tb + root_row(0, 2), dw, e);
            const Cplx<T> om = phase_at(tb + root_row(1, 0), tb + root_row(1, 1), tb + root_row(1, 2), dw, e);
            const Cplx<T> g = t.root[e];
            const T ar = ta.x - om.x, ai = ta.y - om.y, gr = g.x - om.x, gi = g.y - om.y;
            part += (double)((ar * gr + ai * gi) * t.Wmat[e]);
        }
        part *= 2.0 * t.cW;
    } else if (b < WB + 2 * nL * DB) {
        // difference levels strictly above the endpoint's leaf (src/energy_shift.cpp:92-140)
        const int r = b - WB;
        const int ep = r / (nL * DB);  // 0 target, 1 origin
        const int li = (r / DB) % nL;
        const int l = 2 + li;
        const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
        if (l <= lam - 1) {
            const int node = ep == 0 ? ti.path_t[l] : ti.path_o[l];
            const int ps = a.ring ? ti.pws[ep][l] : t.pw_slot[node];
            const Cplx<T>* in = t.inc + (size_t)ps * t.Md;
            const T* Dl = t.Dmat + (size_t)li * t.Md;
            const Cplx<T>* X = tb + level_row(nL, ep, li, 0, 0);
            const Cplx<T>* Y = tb + level_row(nL, ep, li, 1, 0);
            const Cplx<T>* Z = tb + level_row(nL, ep, li, 2, 0);
#pragma unroll 4
            for (int e = (r % DB) * TB + threadIdx.x; e < t.Md; e += DB * TB) {
                const Cplx<T> v = in[e];
                const Cplx<T> w = phase_at(X, Y, Z, t.dd, e);
                T re = w.x * v.x - w.y * v.y;
                if (ep == 1) re -= w.x * w.x + w.y * w.y;  // origin drops its own wave (vecops.hpp:161-163)
                part += (double)(re * Dl[e]);
            }
            part *= 2.0 * t.cD[l] * (ep == 0 ? 1.0 : -1.0);
        }
    } else {
        // near field of the endpoint's leaf (src/energy_shift.cpp:233-302, direct_eval.hpp:309-403)
        const int r = b - WB - 2 * nL * DB;
        const int ep = r / NB;
        const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
        const int leaf = ep == 0 ? ti.path_t[lam] : ti.path_o[lam];
        // fused chain: the prepared entries (count first); this block's first
        // 64 entries load together with the count
        const int4* ent = a.s.near + (size_t)(2 * sl + ep) * (t.max_near + 1) + 1;
        int4 pre = make_int4(0, 0, 0, 0);
        if (a.ring && threadIdx.x < 64 && r % NB + NB * (int)threadIdx.x < t.max_near)
            pre = ent[r % NB + NB * threadIdx.x];
        const int off = a.ring ? 0 : t.near_off[leaf];
        const int cnt = a.ring ? ent[-1].x : t.near_off[leaf + 1] - off;
        const double cx = ep == 0 ? ti.t[0] : ti.o[0], cy = ep == 0 ? ti.t[1] : ti.o[1],
                     cz = ep == 0 ? ti.t[2] : ti.o[2];
        const double bb = t.boxsize[lam];
        const T b2 = (T)(bb * bb), inv_b = (T)(1.0 / bb);
        const T eps = std::numeric_limits<T>::epsilon();
        // hard spheres: particles other than the mover (its old slot) within
        // sigma of the target; sigma never exceeds the finest box size, so the
        // near list covers them
        const bool hs = ep == 0 && a.sigma2 > 0.0;
        const bool ft = ep == 0 && a.prop.learned;  // reverse-move features at the target
        const int own = (hs || ft) ? t.slot_of[ti.i] : -1;
        const double rf2 = a.prop.rf * a.prop.rf;
        double acc = 0;
        // this block's near-list entries k = r % NB + NB j, 64 at a time: their
        // slot ranges are loaded in parallel, then all their particles form
        // one flat loop (no chain of dependent loads per entry)
        __shared__ int es0[64], epre[65];
        __shared__ double esh[64][3];
        const int nk = cnt > r % NB ? (cnt - r % NB + NB - 1) / NB : 0;
        for (int j0 = 0; j0 < nk; j0 += 64) {
            const int ne = min(64, nk - j0);
            if (threadIdx.x < ne) {
                const int k = r % NB + NB * (j0 + threadIdx.x);
                int code;
                if (a.ring) {
                    const int4 e4 = j0 == 0 ? pre : ent[k];
                    es0[threadIdx.x] = e4.x;
                    epre[threadIdx.x + 1] = e4.y;
                    code = e4.z;
                } else {
                    const int nb = t.near_ids[off + k];
                    code = t.near_img[off + k];
                    es0[threadIdx.x] = t.slot_start[nb];
                    epre[threadIdx.x + 1] = t.slot_count[nb];
                }
                esh[threadIdx.x][0] = (code % 3 - 1) * t.L;
                esh[threadIdx.x][1] = ((code / 3) % 3 - 1) * t.L;
                esh[threadIdx.x][2] = (code / 9 - 1) * t.L;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                epre[0] = 0;
                for (int e = 1; e <= ne; ++e) epre[e] += epre[e - 1];
            }
            __syncthreads();
            const int total = epre[ne];
            nscan += total;
            // two particles per pass (u and u + TB): both loads in flight together
            int e = 0;
            for (int u0 = threadIdx.x; u0 < total; u0 += 2 * TB) {
                int sl2[2];
                double px2[2], py2[2], pz2[2], pq2[2], shx[2], shy[2], shz[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int u = u0 + h * TB;
                    const bool in = u < total;
                    while (in && epre[e + 1] <= u) ++e;  // u increases: the entry index only moves forward
                    sl2[h] = in ? es0[e] + (u - epre[e]) : -1;
                    const int sv = in ? sl2[h] : es0[0];
                    px2[h] = t.px[sv];
                    py2[h] = t.py[sv];
                    pz2[h] = t.pz[sv];
                    pq2[h] = t.pq[sv];
                    shx[h] = esh[e][0];
                    shy[h] = esh[e][1];
                    shz[h] = esh[e][2];
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int s = sl2[h];
                    if (s < 0) continue;
                    const double ex = cx - (px2[h] + shx[h]), ey = cy - (py2[h] + shy[h]), ez = cz - (pz2[h] + shz[h]);
                    const T dx = (T)ex, dy = (T)ey, dz = (T)ez;
                    const T r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 > eps && r2 < b2)
                        acc += (double)((T)pq2[h] * residual_poly_over_r<T>(sqrt(r2), inv_b, t.poly, t.n_poly));
                    // distances first: the hole test (porig) loads only for close particles
                    const double r2d = ex * ex + ey * ey + ez * ez;
                    if (hs && r2d < a.sigma2 && s != own && t.porig[s] >= 0) ++novl;
                    if (ft && r2d > 0.0 && r2d < rf2 && s != own && t.porig[s] >= 0)
                        feature_pair(ex, ey, ez, r2d, pq2[h], a.prop.rf, fsum);
                }
            }
            __syncthreads();  // the entry lists are consumed before the next pass
        }
        part = acc * ti.q;
    }
    const double s = block_sum(part, red);
    const double ovs = a.sigma2 > 0.0 ? block_sum((double)novl, red) : 0.0;  // uniform across the grid
    const bool target_near = b >= WB + 2 * nL * DB && b < WB + 2 * nL * DB + NB;
    if (a.prop.learned && target_near) {  // uniform across the block
        __shared__ double redn[(TB / 64) * kFeatSums];
        block_sum_n<kFeatSums>(fsum, redn);
    }
    if (threadIdx.x == 0) {
        double* p = a.s.part + kPartStride * ((size_t)tr * per + b);
        __hip_atomic_store(p, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(p + 1, (double)nscan, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(p + 2, ovs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int c = 0; c < kFeatSums; ++c)
            __hip_atomic_store(p + 3 + c, target_near ? fsum[c] : 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Decision of one trial (single-move modes; trial-only batches): assembles dE
// from the partials of its trial blocks (kernel boundary: no cross-block
// fences), adds the learned proposal's log-ratio, takes the Metropolis
// decision and, for a single move, plans the update.  One block per trial.
template <typename T, bool PM>
__device__ __forceinline__ void decide_body(const DevTree<T>& t, const McArgs& a) {
    __shared__ double sh_dE, phi[kFeat], mu[3], sig[3];
    __shared__ int sh_ovl;
    if (*t.err) return;
    const int tr = blockIdx.x, m = move_base(a) + tr;
    if (tr >= keff(a)) return;
    const TrialInfo& ti = a.s.info[tr];
    unsigned long long tp = wall_clock64();
    const bool learned = a.prop.learned && a.mode == kTrialMetropolis;
    if (threadIdx.x < 64) {  // wave 0 sums the partials, lane 0 assembles
        double v[kAccPerTrial], fs[kFeatSums];
        sum_trial_partials(t, a, tr, threadIdx.x, 64, v);
        if (learned) sum_target_features(t, a, tr, threadIdx.x, 64, fs);
        if (threadIdx.x == 0) {
            int ovl = 0;
            sh_dE = assemble_trial(t, a, tr, m, v, &ovl);
            sh_ovl = ovl;
            if (learned) {
                features_from_sums(ti.q, fs, phi);
                if (a.prop.feat_t)
                    for (int f = 0; f < kFeat; ++f) a.prop.feat_t[(size_t)kFeat * m + f] = phi[f];
            }
        }
    }
    double H = 0.0;
    if (learned) {  // proposal log-ratio: reverse move from the target
        __syncthreads();
        mlp_forward_mfma(a.prop, phi, mu, sig);
        __syncthreads();
        if (threadIdx.x == 0) {
            const double back[3] = {-a.prop.dprop[3 * m], -a.prop.dprop[3 * m + 1], -a.prop.dprop[3 * m + 2]};
            H = gauss_logq(back, mu, sig) - a.prop.logq[m];
            a.prop.H[m] = H;
        }
    }
    if (threadIdx.x == 0) {
        double dE = sh_dE;
        const int ovl = sh_ovl;
        if (a.mode == kTrialOnly && a.ml.dE_stale) a.ml.dE_stale[m] = dE;
        int ok = 0;
        if (a.mode == kTrialMetropolis) {
            dE = ovl > 0 ? (double)INFINITY : a.lB * dE;  // a hard-sphere overlap is an infinite energy
            ok = ovl == 0 && dE - H / a.beta < mh_threshold(a.ml.u[m], a.beta);
        }
        a.ml.dE[m] = dE;
        if (a.mode == kTrialForceAccept) ok = 1;
        if (ok) ok = capacity_ok(t, ti, 0);
        a.s.accept[tr] = ok;
        if (a.ml.accept_out) a.ml.accept_out[m] = ok;
        probe_add(a.s, kPrTrialFinal, tp);
    }
    if (a.mode == kTrialOnly) return;
    __syncthreads();
    // single-move step: this block plans the update; with PM (launch_trial
    // added the marking row) the trial kernel marked the arrays and this block
    // moves the particle and writes the accepted count
    if (PM && premark(a))
        plan_update<T, true>(t, a, 1);
    else
        plan_update(t, a, 1);
}

template <typename T>
__global__ void __launch_bounds__(TB) k_decide(DevTree<T> t, McArgs a) {
    decide_body<T, true>(t, a);
}

// The replica's decision is its step's last reader of the move counter (the
// update reads the plan), so thread 0 advances the counter when the block is
// done: every other thread read it before the block's first barrier, or
// returned early with it (no move left, or an error) where a later value
// changes nothing.
template <typename T>
__global__ void __launch_bounds__(TB) k_decide_rep(const DevTree<T>* __restrict__ ts,
                                                   const McArgs* __restrict__ as) {
    const McArgs& a = as[blockIdx.z];
    decide_body<T, false>(ts[blockIdx.z], a);
    if (threadIdx.x == 0) *a.move_ctr += 1;
}

// six waves per SIMD (at most 80 registers): the role blocks are chains of
// dependent loads, more of them resident beats more registers each
template <typename T, bool PM>  // PM: with the marking row (single-move steps)
__global__ void __launch_bounds__(TB, 6) k_trial(DevTree<T> t, McArgs a) {
    trial_body<T, PM>(t, a);
}

// replica blockIdx.z; the grid is sized for the deepest replica and the
// blocks a shallower replica does not use return at once
template <typename T>
__global__ void __launch_bounds__(TB) k_trial_rep(const DevTree<T>* __restrict__ ts, const McArgs* __restrict__ as) {
    const DevTree<T>& t = ts[blockIdx.z];
    const McArgs& a = as[blockIdx.z];
    if ((int)blockIdx.x >= trial_blocks(t.max_depth, a.tiles)) return;
    trial_body(t, a);
}

// |x - y| below sqrt(r2max) under the minimum image
__device__ __forceinline__ bool within(const double* x, const double* y, double L, double r2max) {
    double s = 0;
    for (int c = 0; c < 3; ++c) {
        double v = x[c] - y[c];
        v -= L * rint(v / L);
        s += v * v;
    }
    return s < r2max;
}

// Learned proposal in a batch: trial tr's proposal log-ratio
// H = log q(-delta | phi(t)) - log q(delta | phi(o)) from the target feature
// sums of its trial blocks (read after the trial kernel's boundary: no
// cross-block fences), by one block (k_pair_corr's record block of the
// trial).  Called by every thread of the block; H is stored and returned in
// thread 0.
template <typename T>
__device__ double reverse_logratio(const DevTree<T>& t, const McArgs& a, int tr, int m) {
    __shared__ double phi[kFeat], mu[3], sig[3];
    if (threadIdx.x < 64) {
        double fs[kFeatSums];
        sum_target_features(t, a, tr, threadIdx.x, 64, fs);
        if (threadIdx.x == 0) {
            features_from_sums(a.s.info[tr].q, fs, phi);
            if (a.prop.feat_t)
                for (int f = 0; f < kFeat; ++f) a.prop.feat_t[(size_t)kFeat * m + f] = phi[f];
        }
    }
    __syncthreads();
    mlp_forward_mfma(a.prop, phi, mu, sig);
    __syncthreads();
    double H = 0.0;
    if (threadIdx.x == 0) {
        const double back[3] = {-a.prop.dprop[3 * m], -a.prop.dprop[3 * m + 1], -a.prop.dprop[3 * m + 2]};
        H = gauss_logq(back, mu, sig) - a.prop.logq[m];
        a.prop.H[m] = H;
    }
    return H;
}


// ------------------------------------------------------------------ batched trials
// Effect of accepted move i on trial j (i < j, distinct particles): exactly
// what the update of i changes in the quantities trial j reads
//   window:     2 cW sum W Re[(tau_j - omega_j) conj(tau_i - omega_i)]
//   difference: in[B] of j's path box B at level l changes by conj(tau_i s)
//               when i's target box at level l is B or a neighbour of B
//               (s = phase of the image of t_i next to B), minus the same for o_i
//   near field: +-q_i q_j R_b(|x_j - y|) for y in {t_i, o_i}, b of x_j's leaf
// Weighted half-grid sums of products of two separable waves, NC of them
// against the same table W in one pass over it,
//   sum_c sg[c] Re sum_k W[k] a_c[k] conj(b_c[k]),  a_c[k] = ax_c[x] ay_c[y] az_c[z],
//   k = x + d y + d^2 z,
//   = sum_c sg[c] Re sum_{y,z} (ay conj by)_c[y] (az conj bz)_c[z] sum_x W[k] (ax conj bx)_c[x]:
// d^2 nz real-times-complex products per term instead of forming every phase,
// and every element of W is loaded once for all NC terms.  rows[6c..6c+5] are
// term c's ax, ay, az, bx, by, bz.  The factor rows go to shared A/B/C (NC
// rows of kRowStride each); each thread takes (y, z) rows and returns its
// share (the caller reduces over the block).  dot_rows is the second phase
// (A/B/C formed, kRowStride apart per term); separable_dots forms the rows
// from the six row pointers per term and calls it.
template <typename T, int NC>
__device__ __forceinline__ double dot_rows(const T* W, int d, int nz, const Cplx<T>* A, const Cplx<T>* B,
                                           const Cplx<T>* C, const double* sg) {
    double acc = 0.0;
    for (int rr = threadIdx.x; rr < d * nz; rr += blockDim.x) {
        const int y = rr % d, z = rr / d;
        const T* w = W + (size_t)d * y + (size_t)d * d * z;
        T sr[NC], si[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) sr[c] = si[c] = T(0);
        int x = 0;
        for (; x + 8 <= d; x += 8) {  // eight independent loads in flight
            T wv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) wv[u] = w[x + u];
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const Cplx<T> av = A[c * kRowStride + x + u];
                    sr[c] += wv[u] * av.x;
                    si[c] += wv[u] * av.y;
                }
        }
        for (; x < d; ++x) {
            const T wv = w[x];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const Cplx<T> av = A[c * kRowStride + x];
                sr[c] += wv * av.x;
                si[c] += wv * av.y;
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const Cplx<T> bc = cmul(B[c * kRowStride + y], C[c * kRowStride + z]);
            acc += sg[c] * (double)(bc.x * sr[c] - bc.y * si[c]);
        }
    }
    return acc;
}

template <typename T, int NC>
__device__ __forceinline__ double separable_dots(const T* W, int d, int nz, const Cplx<T>* const* rows, const double* sg,
                                 Cplx<T>* A, Cplx<T>* B, Cplx<T>* C) {
    __syncthreads();  // the previous call's A/B/C are consumed
    const int per = 2 * d + nz;
#pragma unroll
    for (int c = 0; c < NC; ++c)  // compile-time term index: the row pointers stay in registers
        for (int x = threadIdx.x; x < per; x += blockDim.x) {
            const Cplx<T> u = x < d ? rows[6 * c][x] : x < 2 * d ? rows[6 * c + 1][x - d] : rows[6 * c + 2][x - 2 * d];
            const Cplx<T> v =
                x < d ? rows[6 * c + 3][x] : x < 2 * d ? rows[6 * c + 4][x - d] : rows[6 * c + 5][x - 2 * d];
            const Cplx<T> r = {u.x * v.x + u.y * v.y, u.y * v.x - u.x * v.y};  // u conj(v)
            if (x < d)
                A[c * kRowStride + x] = r;
            else if (x < 2 * d)
                B[c * kRowStride + x - d] = r;
            else
                C[c * kRowStride + x - 2 * d] = r;
        }
    __syncthreads();
    return dot_rows<T, NC>(W, d, nz, A, B, C, sg);
}

// One trial's record for the decisions (wave 0 of an extra block of
// k_pair_corr): partial sums, dE, threshold (H: the learned proposal's
// log-ratio), target-leaf room.
template <typename T>
__device__ void assemble_record(const DevTree<T>& t, const McArgs& a, int j, double H) {
    double v[kAccPerTrial];
    sum_trial_partials(t, a, j, threadIdx.x, 64, v);
    if (threadIdx.x != 0) return;
    const int m = move_base(a) + j;
    TrialRec r;
    r.dE = assemble_trial(t, a, j, m, v, &r.ov);
    a.ml.dE_stale[m] = r.dE;
    const TrialInfo& tj = a.s.info[j];
    r.leaf_t = tj.path_t[tj.lam_t];
    r.leaf_o = tj.path_o[tj.lam_o];
    r.room = r.leaf_t == r.leaf_o ? (1 << 30) : t.slot_cap[r.leaf_t] - t.slot_count[r.leaf_t];
    // learned proposal: lB dE - H / beta < -ln(u) / beta
    r.thr = mh_threshold(a.ml.u[m], a.beta) + (a.prop.learned ? H / a.beta : 0.0);
    r.pid = tj.i;
    r.pad = 0;
    a.s.rec[j] = r;
}

// Batches of at most 64 trials in trees of at most kScanLv expansion levels:
// k_pair_corr's record blocks list the arrays each trial's update would
// rewrite (as if accepted) and k_scan unites the lists of the accepted ones.
template <typename T>
__device__ __forceinline__ bool scan_fastmark(const DevTree<T>& t, const McArgs& a) {
    return a.s.cand != nullptr && a.K <= 64 && n_levels_pw(t.max_depth) <= kScanLv &&
           2 * t.n_pw_nodes <= 4 * kMarkWords;  // a byte mark per array (and 16-bit array ids)
}

// Candidate arrays of trial j's update into s.cand (entry ((ep nL + li) 64 +
// j) 28 + c: endpoint ep at level 2 + li, c = 0 own out, 1 own in (not a
// leaf), 2 + k the k-th colleague's in (non-leaf colleagues), 0xffff where
// there is none), by the threads th, th + nth, ... of a block.
template <typename T>
__device__ void list_candidates(const DevTree<T>& t, const McArgs& a, int j, int th, int nth) {
    const TrialInfo& ti = a.s.info[j];
    const int nL = n_levels_pw(t.max_depth);
    for (int w = th; w < 2 * nL * 28; w += nth) {
        const int c = w % 28, el = w / 28, ep = el / nL, l = 2 + el % nL;
        int ix = 0xffff;
        if (l <= (ep == 0 ? ti.lam_t : ti.lam_o)) {
            const int node = ep == 0 ? ti.path_t[l] : ti.path_o[l];
            if (c == 0) {
                ix = 2 * t.pw_slot[node];
            } else if (c == 1) {
                ix = t.is_leaf[node] ? -1 : 2 * t.pw_slot[node] + 1;
            } else {
                const int sc = t.col_in_slot[26 * node + c - 2];
                ix = sc >= 0 ? 2 * sc + 1 : 0xffff;
            }
            if (ix < 0) ix = 0xffff;
        }
        a.s.cand[(el * 64 + j) * 28 + c] = (unsigned short)ix;
    }
}

constexpr int kPairTB = 128;  // threads per pair-correction block (two waves: twice the resident blocks of four)

// Blocks 0 .. K-1 assemble the trial records (first, so that their chains of
// dependent loads overlap the pair blocks); the K(K-1)/2 blocks after them
// form the pair corrections.
template <typename T>
__global__ void __launch_bounds__(TB) k_pair_corr(DevTree<T> t, McArgs a) {
    __shared__ double red[TB / 64];
    __shared__ Cplx<T> A[8 * kRowStride], B[8 * kRowStride], C[8 * kRowStride];
    const int K = a.K;
    if ((int)blockIdx.x < K) {
        const int j = blockIdx.x;
        if (*t.err || j >= keff(a)) return;  // uniform
        __shared__ double Hs;
        if (a.prop.learned) {  // the reverse-move log-ratio first (the whole block)
            const double H = reverse_logratio(t, a, j, move_base(a) + j);
            if (threadIdx.x == 0) Hs = H;
            __syncthreads();
        }
        if (threadIdx.x >= 64) {  // the other waves list the update's candidate arrays
            if (scan_fastmark(t, a)) list_candidates(t, a, j, threadIdx.x - 64, blockDim.x - 64);
            return;
        }
        assemble_record(t, a, j, a.prop.learned ? Hs : 0.0);
        return;
    }
    int p = blockIdx.x - K, i = 0;
    while (p >= K - 1 - i) {
        p -= K - 1 - i;
        ++i;
    }
    const int j = i + 1 + p;
    // the error flag and the step counter are read with the trial records
    // below and tested after them (their slots are valid memory either way)
    const bool skip = *t.err || j >= keff(a);
    const TrialInfo& A_ = a.s.info[i];
    const TrialInfo& Bj = a.s.info[j];
    const int D = t.max_depth, nL = n_levels_pw(D), stride = table_stride(D);
    const Cplx<T>* ta = tabs_of<T>(a.s, stride, i);
    const Cplx<T>* tj = tabs_of<T>(a.s, stride, j);
    // window: (tau_j - omega_j) against (tau_i - omega_i): the Gram blocks'
    // shares (a thread's first two loaded here and added at the end, so that
    // no wait for them precedes the level terms' loads), or four separable
    // terms
    double part = 0.0;
    const int G = use_gram<T>(a) ? gram_blocks(t.Mw) : 0;
    const double* gsh = a.s.gram + (size_t)i * G * K + j;
    const double gs0 = (int)threadIdx.x < G ? gsh[(size_t)threadIdx.x * K] : 0.0;
    const double gs1 = (int)(threadIdx.x + blockDim.x) < G ? gsh[(size_t)(threadIdx.x + blockDim.x) * K] : 0.0;
    // difference levels of j's paths next to i's boxes.  Group (ep, li): j's
    // endpoint ep (0 target +, 1 origin -) at level l = 2 + li; slot w of the
    // group: i's endpoint w (0 target +, 1 origin -); on when l is above j's
    // leaf level, i's box exists at level l and (as the periodic image
    // shifted by sh) is B or a neighbour of j's box B.  The image shifts of
    // every (group, slot, axis) first, one thread each, from the endpoints'
    // cells (computed from the positions, as locate finds the path boxes).
    __shared__ int shf[2 * kMaxLevels * 6];  // ((ep nL + li) 2 + w) 3 + axis: image shift, 2 = off
    for (int tt = threadIdx.x; tt < 12 * nL; tt += blockDim.x) {
        const int ax = tt % 3, w = (tt / 3) & 1, gg = tt / 6;
        const int ep = gg / nL, l = 2 + gg % nL;
        const int lam = ep == 0 ? Bj.lam_t : Bj.lam_o;
        const int lami = w == 0 ? A_.lam_t : A_.lam_o;
        int v = 2;
        if (l <= lam - 1 && lami >= l) {
            const double xj = ep == 0 ? Bj.t[ax] : Bj.o[ax], xi = w == 0 ? A_.t[ax] : A_.o[ax];
            v = image_shift(cell_at(xi, t.L, l), cell_at(xj, t.L, l), 1 << l);
        }
        shf[tt] = v;
    }
    if (skip) return;  // uniform
    if (!G) {
        const Cplx<T>* rows[24];
        double sg[4];
        for (int ej = 0; ej < 2; ++ej)
            for (int ei = 0; ei < 2; ++ei) {
                const int c = 2 * ej + ei;
                sg[c] = (ej == ei) ? 1.0 : -1.0;
                for (int ax = 0; ax < 3; ++ax) {
                    rows[6 * c + ax] = tj + root_row(ej, ax);
                    rows[6 * c + 3 + ax] = ta + root_row(ei, ax);
                }
            }
        part = separable_dots<T, 4>(t.Wmat, t.dw, t.n_window + 1, rows, sg, A, B, C) * (2.0 * t.cW);
    }
    // Passes over pairs of levels (2 pp, 2 pp + 1), four groups each
    // (gq: ep = gq & 1, li = 2 pp + gq / 2), slot q = 2 gq + w of A/B/C:
    // every slot's factor rows are formed at once, then the dots taken level
    // by level, the four slots of a level in one pass over its table (two
    // barriers per pass).
    int ndots = 0;  // level terms formed (counted for the profiled runs)
    // over the folded tables (Dfold): x and y rows folded, A'[x'] = A[n + x'] + A[n - x']
    const int nd = t.n_diff, n1 = nd + 1, nzd = nd + 1, perd = 2 * n1 + nzd;
    for (int pp = 0; 2 * pp < nL; ++pp) {
        const int ng = 2 * pp + 1 < nL ? 4 : 2;
        __syncthreads();  // the shifts written; the previous pass's (or the window's) LDS consumed
        unsigned on = 0;  // the same in every thread
        for (int q = 0; q < 2 * ng; ++q) {
            const int gq = q >> 1, w = q & 1, gg = (gq & 1) * nL + 2 * pp + (gq >> 1);
            const int* sh3 = shf + (2 * gg + w) * 3;
            if (sh3[0] != 2 && sh3[1] != 2 && sh3[2] != 2) on |= 1u << q;
        }
        if (!on) continue;
        // factor rows u conj(v) of the on slots (u = j's row, v = i's shifted
        // row; zero rows for the off slots), the x and y rows folded at the
        // grid centre n: a thread's entries FR at a time, all their loads in
        // flight together (indices clamped into the rows; the unfolded z rows
        // add nothing)
        constexpr int FR = 4;
        const int nfe = 2 * ng * perd;
        for (int e0 = threadIdx.x; e0 < nfe; e0 += FR * blockDim.x) {
            Cplx<T> u1[FR], v1[FR], u2[FR], v2[FR];
#pragma unroll
            for (int k = 0; k < FR; ++k) {
                const int e = min(e0 + k * (int)blockDim.x, nfe - 1), q = e / perd, x = e % perd;
                const int gq = q >> 1, w = q & 1, ep = gq & 1, li = 2 * pp + (gq >> 1);
                const int ax = x < n1 ? 0 : x < 2 * n1 ? 1 : 2, xi = x - ax * n1;
                const int sh = shf[(2 * (ep * nL + li) + w) * 3 + ax];
                const Cplx<T>* ru = tj + level_row(nL, ep, li, ax, 0);
                const Cplx<T>* rv = ta + level_row(nL, w, li, ax, sh == 2 ? 0 : sh);
                const int x1 = ax == 2 ? xi : nd + xi, x2 = ax == 2 ? xi : nd - xi;
                u1[k] = ru[x1];
                v1[k] = rv[x1];
                u2[k] = ru[x2];
                v2[k] = rv[x2];
            }
#pragma unroll
            for (int k = 0; k < FR; ++k) {
                const int e = e0 + k * (int)blockDim.x;
                if (e >= nfe) continue;
                const int q = e / perd, x = e % perd;
                const int ax = x < n1 ? 0 : x < 2 * n1 ? 1 : 2, xi = x - ax * n1;
                Cplx<T> r = {u1[k].x * v1[k].x + u1[k].y * v1[k].y, u1[k].y * v1[k].x - u1[k].x * v1[k].y};
                if (ax < 2 && xi > 0) {  // the mirrored entry of an x or y row
                    r.x += u2[k].x * v2[k].x + u2[k].y * v2[k].y;
                    r.y += u2[k].y * v2[k].x - u2[k].x * v2[k].y;
                }
                if (!((on >> q) & 1u)) r = {T(0), T(0)};  // an off slot adds nothing to its level's dot
                Cplx<T>* dst = ax == 0 ? A : ax == 1 ? B : C;
                dst[q * kRowStride + xi] = r;
            }
        }
        __syncthreads();
        for (int lq = 0; 2 * lq < ng; ++lq) {
            // slots 4 lq + c, c = 2 ep + w: sign + when ep and w agree
            const unsigned m4 = (on >> (4 * lq)) & 15u;
            if (!m4) continue;  // uniform
            const int li = 2 * pp + lq, o = 4 * lq * kRowStride;
            const T* Dl = t.Dfold + (size_t)li * n1 * n1 * nzd;
            const double sg[4] = {1.0, -1.0, -1.0, 1.0};
            const double lv = dot_rows<T, 4>(Dl, n1, nzd, A + o, B + o, C + o, sg);
            ndots += __popc(m4);
            part += lv * 2.0 * t.cD[2 + li];
        }
    }
    // near field: four threads of the last wave take one (ep, w) pair each
    // (+-q_i q_j R_b(|x_j - y|), b of x_j's leaf box), a fifth the flags; their
    // work overlaps the other waves' dots
    const int q = threadIdx.x - ((int)blockDim.x - 64);
    if (q >= 0 && q < 4) {
        const int ep = q >> 1, w = q & 1;
        const double* xj = ep == 0 ? Bj.t : Bj.o;
        const double* y = w == 0 ? A_.t : A_.o;
        const double bb = t.boxsize[ep == 0 ? Bj.lam_t : Bj.lam_o];
        const T b2 = (T)(bb * bb), inv_b = (T)(1.0 / bb);
        double dx = xj[0] - y[0], dy = xj[1] - y[1], dz = xj[2] - y[2];
        dx -= t.L * rint(dx / t.L);
        dy -= t.L * rint(dy / t.L);
        dz -= t.L * rint(dz / t.L);
        const T fx = (T)dx, fy = (T)dy, fz = (T)dz;
        const T r2 = fx * fx + fy * fy + fz * fz;
        if (r2 > std::numeric_limits<T>::epsilon() && r2 < b2)
            part += ((ep == w) ? 1.0 : -1.0) * A_.q * Bj.q *
                    (double)residual_poly_over_r<T>(sqrt(r2), inv_b, t.poly, t.n_poly);
    } else if (q == 4 && (a.sigma2 > 0.0 || a.prop.learned)) {
        // hard spheres: 1 t_j within sigma of o_i (counted, now gone), 2 of t_i (new);
        // learned proposal: 4 j's proposal is stale once i is accepted
        int flags = 0;
        if (a.sigma2 > 0.0) {
            if (within(Bj.t, A_.o, t.L, a.sigma2)) flags |= 1;
            if (within(Bj.t, A_.t, t.L, a.sigma2)) flags |= 2;
        }
        if (a.prop.learned) {
            const double rf2 = a.prop.rf * a.prop.rf;
            if (A_.i == Bj.i || within(A_.o, Bj.o, t.L, rf2) || within(A_.t, Bj.o, t.L, rf2) ||
                within(A_.o, Bj.t, t.L, rf2) || within(A_.t, Bj.t, t.L, rf2))
                flags |= 4;
        }
        a.s.ovl[i * K + j] = (unsigned char)flags;
    }
    if (G) {
        double gw = gs0 + gs1;
        for (int g = threadIdx.x + 2 * blockDim.x; g < G; g += blockDim.x) gw += gsh[(size_t)g * K];
        part += gw * (2.0 * t.cW);
    }
    const double total = block_sum(part, red);
    if (threadIdx.x == 0) {
        a.s.corr[i * K + j] = total;
        if (a.ml.stats && ndots) atomicAdd(a.ml.stats + kStatsPerMove * (move_base(a) + j) + 6, ndots);
    }
}

// r0 plus the entries of c whose bit is set in mask, in a fixed order (four
// interleaved partial sums, then joined): the batch walk forms every dE it
// compares and reports with this one function.
__device__ __forceinline__ double masked_sum(double r0, const double* c, unsigned long long mask) {
    double p[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 64; ++i) p[i & 3] += ((mask >> i) & 1ull) ? c[i] : 0.0;
    return r0 + ((p[0] + p[1]) + (p[2] + p[3]));
}

// Batch decisions: dE_j = stale_j + sum over the accepted i < j of corr[i][j],
// taken in order by one wave (lane l owns trials l, l+64, ...: it keeps their
// running corrected dE and the appends into their target leaves); then the
// update plan of the whole batch.
//
// With a device move counter (learned proposal) the step decides the trials
// up to the first one whose proposal an earlier accepted move of the batch
// made stale: the same particle, or an origin or target within the feature
// radius of its origin or target (its features, hence its proposal density,
// changed).  That trial starts the next step; the counter advances by the
// number decided, so the chain is the plain chain.
// HS: hard spheres may be on; LR: a learned proposal or a step counter may
// be attached.  The walk of the uniform chain without hard spheres compiles
// without the flag, overlap and staleness bookkeeping.
template <typename T, bool HS, bool LR>
__global__ void __launch_bounds__(TB) k_scan(DevTree<T> t, McArgs a) {
    __shared__ double dEs[kMaxBatch], thr[kMaxBatch], dEf[kMaxBatch];
    __shared__ int leafT[kMaxBatch], leafO[kMaxBatch], accs[kMaxBatch], room[kMaxBatch], ov0[kMaxBatch],
        pids[kMaxBatch];
    __shared__ double C[64 * 64];
    __shared__ unsigned Fw[64 * 64 / 4];  // pair flags, read as words
    unsigned char* F = reinterpret_cast<unsigned char*>(Fw);
    __shared__ int cons;
    // update plan from the trials' candidate lists (scan_fastmark)
    __shared__ __align__(16) unsigned short cl[kScanCand];
    __shared__ unsigned mbits[kMarkWords];
    __shared__ int accf[64], wtf[TB / 64], wtb[TB / 64];
    __shared__ unsigned short pmask[4][4][64];  // (mask kind, quarter of the earlier trials, trial)
    __shared__ double pabs[4][64];
    // particle moves made here (scan_reloc, K <= 64): per trial the target,
    // the particle's slot and charge and its target leaf's first slot and
    // count, loaded by the second wave while the first walks the batch
    __shared__ double rv_t[64][3], rv_q[64];
    __shared__ int rv_s[64], rv_start[64], rv_cnt[64];
    if (*t.err) return;
    const int K = a.K;             // stride of the pair tables
    const int base = move_base(a);
    const int Ke = keff(a);        // trials of this step
    const bool hs = HS && a.sigma2 > 0.0;
    const bool dyn = LR && a.move_ctr != nullptr;
    const bool flg = hs || (LR && a.prop.learned);  // pair flags present
    const int nL = n_levels_pw(t.max_depth), n2 = 2 * t.n_pw_nodes;
    const bool fm = scan_fastmark(t, a);  // implies K <= 64
    unsigned long long tp = wall_clock64();
    if (fm)  // the byte marks of the update plan
        for (int w = threadIdx.x; w < (n2 + 3) / 4; w += blockDim.x) mbits[w] = 0u;
    auto put_rec = [&](int j, const TrialRec& r) {
        dEs[j] = r.dE;
        thr[j] = r.thr;
        room[j] = r.room;
        leafT[j] = r.leaf_t;
        leafO[j] = r.leaf_o;
        ov0[j] = r.ov;
        pids[j] = r.pid;
    };
    const bool in_lds = K <= 64;
    if (!in_lds)
        for (int j = threadIdx.x; j < Ke; j += blockDim.x) put_rec(j, a.s.rec[j]);  // records from k_pair_corr
    if (in_lds) {
        // the records (one per thread), the corrections, (four bytes per
        // load) the pair flags and the candidate arrays: every load in flight
        // together, then the LDS stores
        TrialRec rec;
        if ((int)threadIdx.x < Ke) rec = a.s.rec[threadIdx.x];
        const unsigned* ow = reinterpret_cast<const unsigned*>(a.s.ovl);
        const int nw = flg ? (K * K + 3) / 4 : 0;  // <= 1024; the last word stays inside the kMaxBatch^2 flags
        double v[16];
        unsigned fv[4];
        // and the batch's candidate arrays (k_pair_corr's record blocks), int4
        // at a time: the levels in use, all 64 trial rows of each (a trial
        // past the step's end is never read)
        constexpr int CL = kScanCand * 2 / 16 / TB;
        static_assert(kScanCand * 2 == CL * TB * 16, "whole int4 rounds");
        // unconditional loads (no branch between them): from a buffer of at
        // least CL TB int4 whether or not the candidates are listed
        const int n4 = fm ? 2 * nL * 64 * 28 * 2 / 16 : 0;
        const int4* csrc = fm ? reinterpret_cast<const int4*>(a.s.cand) : reinterpret_cast<const int4*>(a.s.corr);
        int4 cv[CL];
#pragma unroll
        for (int r = 0; r < CL; ++r) cv[r] = csrc[threadIdx.x + r * TB];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int q = threadIdx.x + r * TB;
            if (q < K * K) v[r] = a.s.corr[q];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int q = threadIdx.x + r * TB;
            if (q < nw) fv[r] = ow[q];
        }
        asm volatile("" ::: "memory");  // the loads above issue before any store below
        if ((int)threadIdx.x < Ke) put_rec(threadIdx.x, rec);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int q = threadIdx.x + r * TB;
            if (q < K * K) C[q] = v[r];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int q = threadIdx.x + r * TB;
            if (q < nw) Fw[q] = fv[r];
        }
        int4* cdst = reinterpret_cast<int4*>(cl);
#pragma unroll
        for (int r = 0; r < CL; ++r)
            if ((int)threadIdx.x + r * TB < n4) cdst[threadIdx.x + r * TB] = cv[r];
    }
    if (threadIdx.x == 0) cons = Ke > 0 ? Ke : 0;
    __syncthreads();
    if (threadIdx.x == 0) probe_add(a.s, kPrScanAssemble, tp);
    if (in_lds) {
        // for lane j of the walk, over the earlier trials i < j (thread (j, q)
        // takes i = 16 q .. 16 q + 15): i's acceptance removes an overlap
        // (flag 1) or adds one (flag 2), makes j's proposal stale (flag 4, or
        // the same particle), moves a particle into j's target leaf; and the
        // sum of the corrections' magnitudes
        const int j = threadIdx.x & 63, q = threadIdx.x >> 6;
        const bool in = j < Ke;
        const int lt_j = in ? leafT[j] : -1, pid_j = in ? pids[j] : -1;
        unsigned r_rm = 0u, r_add = 0u, r_st = 0u, r_lf = 0u;
        double sa = 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int i = 16 * q + u;  // i K + j < 64 * 64: every read is in bounds, used when i < j < Ke
            const bool e = in && i < j;
            const int fc = flg ? (int)F[i * K + j] : 0;
            const int lti = leafT[i], loi = leafO[i], pdi = pids[i];
            const double cv = C[i * K + j];
            const unsigned b = e ? 1u << u : 0u;
            r_rm |= (fc & 1) ? b : 0u;
            r_add |= (fc & 2) ? b : 0u;
            r_st |= ((fc & 4) || pdi == pid_j) ? b : 0u;
            r_lf |= (loi != lti && lti == lt_j) ? b : 0u;
            sa += e ? fabs(cv) : 0.0;
        }
        pmask[0][q][j] = (unsigned short)r_rm;
        pmask[1][q][j] = (unsigned short)r_add;
        pmask[2][q][j] = (unsigned short)r_st;
        pmask[3][q][j] = (unsigned short)r_lf;
        pabs[q][j] = sa;
        __syncthreads();
    }
    const bool rl = a.scan_reloc && K <= 64 && a.mode == kTrialBatch;
    if (rl && threadIdx.x >= 64 && threadIdx.x < 128) {
        const int j = threadIdx.x - 64;
        if (j < Ke) {
            const TrialInfo& ti = a.s.info[j];
            const int lt = leafT[j], s = t.slot_of[pids[j]];
            rv_t[j][0] = ti.t[0];
            rv_t[j][1] = ti.t[1];
            rv_t[j][2] = ti.t[2];
            rv_start[j] = t.slot_start[lt];
            rv_cnt[j] = t.slot_count[lt];
            rv_s[j] = s;
            rv_q[j] = t.pq[s];
        }
    }
    if (threadIdx.x < 64 && in_lds) {
        // Lane j owns trial j and holds its column of pair corrections in
        // registers.  Trial j's dE given the accepted set S is its stale dE
        // plus the corrections of the accepted trials before it (masked_sum:
        // one fixed order); the counts an acceptance changes come from the
        // masks (hard-sphere overlaps removed or added, movers into the target
        // leaf, stale proposals).  A trial whose outcome no set of earlier
        // acceptances can change (dE farther from its threshold than the sum
        // of its corrections' magnitudes, no flag, room for every earlier
        // mover) is known before the walk, which visits the others in move
        // order: the same outcomes as deciding the trials one by one.
        const int j = threadIdx.x;
        const bool in = j < Ke;
        double col[64];
        // unconditional loads (i K + j < 64 * 64 for K <= 64), so that they
        // issue back to back; entries outside the column are zeroed after
#pragma unroll
        for (int i = 0; i < 64; ++i) col[i] = C[i * K + j];
#pragma unroll
        for (int i = 0; i < 64; ++i) col[i] = (in && i < j) ? col[i] : 0.0;
        unsigned long long m_rm = 0ull, m_add = 0ull, m_st = 0ull, m_lf = 0ull;
        double amax = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            m_rm |= (unsigned long long)pmask[0][q][j] << (16 * q);
            m_add |= (unsigned long long)pmask[1][q][j] << (16 * q);
            m_st |= (unsigned long long)pmask[2][q][j] << (16 * q);
            m_lf |= (unsigned long long)pmask[3][q][j] << (16 * q);
            amax += pabs[q][j];
        }
        const double run0 = in ? dEs[j] : 0.0, thr_j = in ? thr[j] : 0.0, lB = a.lB;
        const int ov_j = in ? ov0[j] : 0, room_j = in ? room[j] : 0;
        // every sum of a subset of the corrections (each addition rounded)
        // stays within dl of the stale dE
        const double dl = amax + 1e-13 * (fabs(run0) + amax);
        const double hi = lB * (run0 + dl), lo = lB * (run0 - dl);
        const bool e_acc = hi + 1e-15 * fabs(hi) < thr_j, e_rej = lo - 1e-15 * fabs(lo) >= thr_j;
        const bool fixed_counts = (!hs || (m_add | m_rm) == 0ull) && __popcll(m_lf) < room_j && (!dyn || m_st == 0ull);
        const bool known = fixed_counts && (ov_j > 0 || e_acc || e_rej);
        const bool known_ok = ov_j == 0 && e_acc;
        unsigned long long tw = wall_clock64();
        if (j == 0) probe_add(a.s, kPrScanPrep, tp);
        unsigned long long U = __ballot(in && !known);        // uniform: decided in the walk
        unsigned long long S = __ballot(in && known && known_ok);  // uniform: accepted (known ones from the start)
        int nc = Ke > 0 ? Ke : 0;  // uniform: trials decided (a stale proposal ends the step)
        int my_ok = known_ok ? 1 : 0;
        while (U) {
            const int i = (int)__builtin_ctzll(U);  // uniform
            U &= U - 1ull;
            const unsigned long long Sb = S & ((1ull << i) - 1ull);
            if (dyn && __builtin_amdgcn_readlane((int)((m_st & Sb) != 0ull), i)) {
                nc = i;
                break;
            }
            const int ovr = ov_j + (hs ? __popcll(m_add & Sb) - __popcll(m_rm & Sb) : 0);
            const int dcr = __popcll(m_lf & Sb);
            const double dE = lB * masked_sum(run0, col, Sb);
            const int okl = (ovr == 0 && dE < thr_j) ? (dcr < room_j ? 1 : 2) : 0;  // 2: the target leaf is full
            if (j == i) my_ok = okl;
            if (__builtin_amdgcn_readlane(okl, i) == 1) S |= 1ull << i;
        }
        if (nc < 64) S &= (1ull << nc) - 1ull;
        if (in) {
            const int ovr = ov_j + (hs ? __popcll(m_add & S) - __popcll(m_rm & S) : 0);
            dEf[j] = ovr > 0 ? (double)INFINITY : lB * masked_sum(run0, col, S);  // an overlap is an infinite energy
            accs[j] = my_ok;
        }
        if (j == 0) {
            cons = nc;
            probe_add(a.s, kPrScanWalk, tw);
        }
    } else if (threadIdx.x < 64 && Ke <= 64) {
        // Lane j owns trial j and keeps what the walk reads of it in
        // registers; the owner's decision reaches the other lanes by
        // readlane, and the pair entries of row i + 1 are read while row i
        // is decided.
        const int j = threadIdx.x;
        const bool in = j < Ke;
        double run = in ? dEs[j] : 0.0;
        const double thr_j = in ? thr[j] : 0.0;
        int ovr = in ? ov0[j] : 0, dcr = 0, stl = 0;
        const int room_j = in ? room[j] : 0, lt_j = in ? leafT[j] : -1, lo_j = in ? leafO[j] : -1;
        const int pid_j = in ? pids[j] : -1;
        const double lB = a.lB;
        double cn = 0.0;
        int fn = 0;
        if (0 < j && in) {
            cn = in_lds ? C[j] : a.s.corr[j];
            if (flg) fn = in_lds ? F[j] : a.s.ovl[j];
        }
        for (int i = 0; i < Ke; ++i) {
            const double cc = cn;
            const int fc = fn;
            if (i + 1 < j && in) {
                cn = in_lds ? C[(i + 1) * K + j] : a.s.corr[(i + 1) * K + j];
                if (flg) fn = in_lds ? F[(i + 1) * K + j] : a.s.ovl[(i + 1) * K + j];
            }
            if (dyn && __builtin_amdgcn_readlane(stl, i)) {  // a stale proposal ends the step
                if (j == 0) cons = i;
                break;
            }
            const double dE = lB * run;
            const int okl = (ovr == 0 && dE < thr_j) ? (dcr < room_j ? 1 : 2) : 0;  // 2: the target leaf is full
            if (j == i) {
                dEf[i] = ovr > 0 ? (double)INFINITY : dE;  // a hard-sphere overlap is an infinite energy
                accs[i] = okl;
            }
            if (__builtin_amdgcn_readlane(okl, i) == 1 && j > i && in) {
                const int lti = __builtin_amdgcn_readlane(lt_j, i);
                run += cc;
                dcr += (__builtin_amdgcn_readlane(lo_j, i) != lti) && (lti == lt_j);
                if (hs) ovr += ((fc >> 1) & 1) - (fc & 1);
                stl |= (fc >> 2) & 1;
                if (dyn && pid_j == __builtin_amdgcn_readlane(pid_j, i)) stl = 1;  // the same particle again
            }
        }
    } else if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        double run[4];
        int dcr[4], ovr[4], stl[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int j = lane + 64 * k;
            run[k] = j < Ke ? dEs[j] : 0.0;
            ovr[k] = j < Ke ? ov0[j] : 0;
            dcr[k] = 0;
            stl[k] = 0;
        }
        const double lB = a.lB;
        for (int i = 0; i < Ke; ++i) {
            const int owner = i & 63, sl = i >> 6;
            if (dyn) {  // a stale proposal ends the step
                const int st = lane == owner ? (sl == 0 ? stl[0] : sl == 1 ? stl[1] : sl == 2 ? stl[2] : stl[3]) : 0;
                if (__builtin_amdgcn_readlane(st, owner)) {
                    if (lane == 0) cons = i;
                    break;
                }
            }
            int ok = 0;
            if (lane == owner) {
                const double dE = lB * (sl == 0 ? run[0] : sl == 1 ? run[1] : sl == 2 ? run[2] : run[3]);
                const int dc = sl == 0 ? dcr[0] : sl == 1 ? dcr[1] : sl == 2 ? dcr[2] : dcr[3];
                const int ov = sl == 0 ? ovr[0] : sl == 1 ? ovr[1] : sl == 2 ? ovr[2] : ovr[3];
                ok = (ov == 0 && dE < thr[i]) ? (dc < room[i] ? 1 : 2) : 0;  // 2: the target leaf is full
                dEf[i] = ov > 0 ? (double)INFINITY : dE;  // a hard-sphere overlap is an infinite energy
                accs[i] = ok;
            }
            ok = __builtin_amdgcn_readlane(ok, owner) == 1;
            if (ok) {
                const int lti = leafT[i], mv = leafO[i] != lti;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int j = lane + 64 * k;
                    if (j > i && j < Ke) {
                        run[k] += in_lds ? C[i * K + j] : a.s.corr[i * K + j];
                        dcr[k] += mv && (lti == leafT[j]);
                        if (flg) {
                            const int f = in_lds ? F[i * K + j] : a.s.ovl[i * K + j];
                            if (hs) ovr[k] += ((f >> 1) & 1) - (f & 1);
                            stl[k] |= (f >> 2) & 1;
                        }
                        if (dyn && pids[j] == pids[i]) stl[k] = 1;  // the same particle again
                    }
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) probe_add(a.s, kPrScanDecide, tp);
    const int nc = cons;
    for (int j = threadIdx.x; j < K; j += blockDim.x) {
        int ok = 0;
        if (j < nc) {
            const int m = base + j;
            ok = accs[j];
            if (ok == 2) {
                atomicOr(t.err, (int)kErrLeafOverflow);
                ok = 0;
            }
            a.ml.dE[m] = dEf[j];
            if (a.ml.accept_out) a.ml.accept_out[m] = ok;
        }
        a.s.accept[j] = ok;  // undecided and unused trials apply nothing
        if (fm) accf[j] = ok;
    }
    __syncthreads();
    if (fm) {
        // the plan's arrays: every accepted trial's candidates marked in a
        // byte per array (2: a level-2 in array, 1: any other; an array
        // always gets the same mark), then one pass over the marks lists them
        // in s.dest, level-2 in arrays from the front and the others from the
        // back
        unsigned char* mk = reinterpret_cast<unsigned char*>(mbits);
        for (int w = threadIdx.x; w < 2 * nc * nL; w += blockDim.x) {
            const int m = w / (2 * nL), el = w % (2 * nL);
            if (!accf[m]) continue;
            const unsigned short* lst = cl + (el * 64 + m) * 28;
            const unsigned char lv2 = el % nL == 0 ? 2 : 1;
            int ix[28];
#pragma unroll
            for (int c = 0; c < 28; ++c) ix[c] = lst[c];
#pragma unroll
            for (int c = 0; c < 28; ++c)
                if (ix[c] != 0xffff) mk[ix[c]] = c > 0 ? lv2 : (unsigned char)1;
        }
        __syncthreads();
        if (threadIdx.x == 0) probe_add(a.s, kPrPlanUnite, tp);
        const int per = (n2 + TB - 1) / TB, e0 = threadIdx.x * per, e1 = min(e0 + per, n2);
        int nf = 0, nb = 0;
        for (int e = e0; e < e1; ++e) {
            nf += mk[e] == 2;
            nb += mk[e] == 1;
        }
        int of = nf, ob = nb;  // inclusive scans over the block
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int uf = __shfl_up(of, o), ub = __shfl_up(ob, o);
            if (lane >= o) {
                of += uf;
                ob += ub;
            }
        }
        if (lane == 63) {
            wtf[wv] = of;
            wtb[wv] = ob;
        }
        __syncthreads();
        int bf = 0, bb = 0, tf = 0, tb = 0;
#pragma unroll
        for (int w = 0; w < TB / 64; ++w) {
            bf += w < wv ? wtf[w] : 0;
            bb += w < wv ? wtb[w] : 0;
            tf += wtf[w];
            tb += wtb[w];
        }
        of += bf - nf;  // exclusive offsets
        ob += bb - nb;
        for (int e = e0; e < e1; ++e) {
            if (mk[e] == 2) a.s.dest[of++] = e;
            else if (mk[e] == 1) a.s.dest[n2 - 1 - ob++] = e;
        }
        if (threadIdx.x == 0) {
            a.s.plan[0] = tf + tb;
            a.s.plan[2] = tf;
            probe_add(a.s, kPrPlanMark, tp);
        }
        if (rl) {
            // the particles move below: the plan needs only the accepted count
            if (threadIdx.x < 64) {
                const unsigned long long acc = __ballot((int)threadIdx.x < nc && accs[threadIdx.x] == 1);
                if (threadIdx.x == 0) {
                    const int nacc = __popcll(acc);
                    a.s.plan[1] = nacc;
                    if (a.ml.stats && nacc) a.ml.stats[kStatsPerMove * base + 7] = tf + tb;
                }
            }
        } else {
            __syncthreads();
            plan_update<T, true>(t, a, K, !apply_relocates(a));
        }
    } else {
        plan_update(t, a, K, !apply_relocates(a) && !rl);
    }
    if (dyn && threadIdx.x == 0) *a.move_ctr = base + nc;
    if (rl && threadIdx.x < 64) {
        // the accepted moves, last, so that no barrier waits for their
        // stores (src/tree.cpp:454-463, 583-596): a move inside its leaf
        // rewrites the position in place, a mover takes the next free slot of
        // its target leaf in move order (its rank among the batch's movers
        // into that leaf) and leaves a hole behind; the first mover into a
        // leaf writes the count
        const int m = threadIdx.x;
        const bool in = m < nc && accs[m] == 1;
        const int lt = in ? leafT[m] : -1;
        const bool mover = in && leafO[m] != lt;
        const int key = mover ? lt : -1;
        int rank = 0, total = 0;
#pragma unroll 4
        for (int k = 0; k < 64; ++k) {
            const int kk = __builtin_amdgcn_readlane(key, k);
            if (mover && kk == key) {
                rank += k < m;
                ++total;
            }
        }
        if (in) {
            const int s = rv_s[m], i = pids[m];
            const int ns = mover ? rv_start[m] + rv_cnt[m] + rank : s;
            t.px[ns] = rv_t[m][0];
            t.py[ns] = rv_t[m][1];
            t.pz[ns] = rv_t[m][2];
            if (mover) {
                t.pq[ns] = rv_q[m];
                t.porig[ns] = i;
                t.slot_of[i] = ns;
                t.leaf_of[i] = lt;
                t.px[s] = 0;
                t.py[s] = 0;
                t.pz[s] = 0;
                t.pq[s] = 0;
                t.porig[s] = -1;
                if (rank == 0) t.slot_count[lt] = rv_cnt[m] + total;
            }
        }
    }
}

// Packs the live particles of every leaf (porig >= 0) to the front of its
// slots, in slot order, after relocations left holes behind.  One block per leaf.
template <typename T>
__device__ __forceinline__ void compact_body(const DevTree<T>& t, const int* leaves) {
    __shared__ int wtot[TB / 64], base;
    if (*t.err) return;
    const int leaf = leaves[blockIdx.x];
    const int s0 = t.slot_start[leaf], n = t.slot_count[leaf];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    // chunks are processed in order; every chunk is read completely before any write
    for (int c0 = 0; c0 < n; c0 += blockDim.x) {
        const int s = s0 + c0 + threadIdx.x;
        const bool in = c0 + threadIdx.x < n;
        const int o = in ? t.porig[s] : -1;
        double x = 0, y = 0, z = 0, q = 0;
        if (o >= 0) {
            x = t.px[s];
            y = t.py[s];
            z = t.pz[s];
            q = t.pq[s];
        }
        const bool live = o >= 0;
        const unsigned long long bal = __ballot(live);
        const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        if (lane == 0) wtot[wv] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int k = 0; k < wv; ++k) off += wtot[k];
        const int dst = s0 + off + __popcll(bal & lt);
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < (int)(blockDim.x >> 6); ++k) base += wtot[k];
        if (live) {
            t.px[dst] = x;
            t.py[dst] = y;
            t.pz[dst] = z;
            t.pq[dst] = q;
            t.porig[dst] = o;
            t.slot_of[o] = dst;
        }
        __syncthreads();
    }
    __syncthreads();
    for (int s = s0 + base + threadIdx.x; s < s0 + n; s += blockDim.x) {
        t.px[s] = 0;
        t.py[s] = 0;
        t.pz[s] = 0;
        t.pq[s] = 0;
        t.porig[s] = -1;
    }
    if (threadIdx.x == 0) t.slot_count[leaf] = base;
}

template <typename T>
__global__ void __launch_bounds__(TB) k_compact(DevTree<T> t, const int* leaves) {
    compact_body(t, leaves);
}

template <typename T>
__global__ void __launch_bounds__(TB) k_compact_rep(const DevTree<T>* __restrict__ ts, const int* const* leaves, const int* n_leaves) {
    if ((int)blockIdx.x >= n_leaves[blockIdx.z]) return;
    const DevTree<T>& t = ts[blockIdx.z];
    compact_body(t, leaves[blockIdx.z]);
}

// next move of every replica (graph-replayable step counter)
__global__ void k_advance(int* move_ctr, int n, int inc) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) move_ctr[r] += inc;
}

// ------------------------------------------------------------------ apply
// Work items: (array, block of grid rows) for the root and every rewritten
// array; a grid row is the d elements sharing (k_y, k_z).  An item first
// loads its part of the array into registers (coalesced; the read is in
// flight meanwhile), lists the (move, endpoint, image) contributions to the
// array in move order and stages their phase rows in LDS a tile at a time.
// Each update is separable, s X[x] Y[y] Z[z]: the lanes of a row form
// s Y[y] Z[z] once per contribution and add X[x] times it to their share of
// the row's x; the sums go through LDS back to the array's order, so the
// write is coalesced too, every element read and written once.  The root (in
// batches with more than kApplyHeavy accepted endpoints) has its contributions
// split over kApplySplit items, whose sums are added atomically instead; other
// arrays are never split (atomic adds over whole arrays cost more than the
// longer contribution loop).
constexpr int kApplyTile = 16;
constexpr int kApplyLv = 4;     // levels of the per-block path table (deeper levels read TrialInfo)
constexpr int kApplySplit = 4;  // items sharing the root
constexpr int kApplyHeavy = 16; // contributions above which the root is split

// lanes per grid row, and the x values one lane holds for grids up to d
template <typename T>
__host__ __device__ constexpr int apply_lpr() { return sizeof(T) == 8 ? 4 : 2; }
template <typename T>
__host__ __device__ constexpr int apply_xmax(int d) { return (d + apply_lpr<T>() - 1) / apply_lpr<T>(); }
// the unguarded reads of a lane end at most kApplyPad elements past the tiles
constexpr int kApplyPad = 64;
// dynamic shared memory of k_apply: two tiles of rows (double buffer) at row
// stride dmax = max(d_w, d) and kApplyPad elements of padding, sharing their
// space with the item's sums (written after the last tile is consumed), then
// the path table of the batch's K moves
template <typename T>
__host__ __device__ inline size_t apply_rows_elems(int dmax) {
    const size_t sums = (size_t)(TB / apply_lpr<T>()) * dmax, tiles = 2 * (size_t)kApplyTile * 3 * dmax + kApplyPad;
    return sums > tiles ? sums : tiles;
}
template <typename T>
__host__ __device__ inline size_t apply_shmem(int dmax, int K) {
    return sizeof(Cplx<T>) * apply_rows_elems<T>(dmax) + sizeof(int) * 2 * (size_t)K * kApplyLv;
}

template <typename T, int XMAX, bool RB = false>
__device__ __forceinline__ void apply_body(const DevTree<T>& t, const McArgs& a, int lead = 0) {
    extern __shared__ char apply_dyn[];
    __shared__ int mlist[2 * kMaxBatch];     // contributions: move << 8 | endpoint << 7 | image code
    __shared__ int mrow[3 * 2 * kMaxBatch];  // their three table rows (x, y, z), offsets into s.tabs
    __shared__ int wtot[TB / 64];
    __shared__ int mcount;
    __shared__ int ac_s[kMaxBatch];
    if (*t.err) return;
    const int nD = a.s.plan[0], nacc = a.s.plan[1], nH = a.s.plan[2];
    if (!nacc) return;
    // RB (a batch, apply_relocates): block 0 moves the particles, the others
    // rewrite arrays; a separate instantiation, so that the single-move update
    // compiles without it
    if constexpr (RB) {
        if ((int)blockIdx.x == lead) {
            relocate_batch(t, a);
            return;
        }
    }
    // blocks before lead have another role (the next step's set-up)
    const int bid = (int)blockIdx.x - lead - (RB ? 1 : 0), nblk = (int)gridDim.x - lead - (RB ? 1 : 0);
    const int K = a.K, D = t.max_depth, nL = n_levels_pw(D), stride = table_stride(D);
    constexpr int LPR = apply_lpr<T>();  // lanes per grid row
    constexpr int RPI = TB / LPR;        // grid rows per work item
    const int dmax = t.dw > t.dd ? t.dw : t.dd;
    Cplx<T>* S = reinterpret_cast<Cplx<T>*>(apply_dyn);  // the item's sums, row-major like the array
    Cplx<T>* rows2 = S;                                   // two tiles of rows at stride dmax (before the sums)
    int* pnode = reinterpret_cast<int*>(S + apply_rows_elems<T>(dmax));  // (move, endpoint, level-2) -> box
    const int rows_w = t.dw * (t.n_window + 1), rows_d = t.dd * (t.n_diff + 1);
    const int nWb = (rows_w + RPI - 1) / RPI, nDb = (rows_d + RPI - 1) / RPI;
    // every accepted endpoint rewrites the root: its contributions are split
    // over sr items.  The level-2 incoming arrays (front of s.dest, the most
    // contributions after the root) come next, so they start first.
    const int sr = 2 * nacc > kApplyHeavy ? kApplySplit : 1;
    const int n_root = nWb * sr;
    const int total = n_root + nD * nDb;
    if (bid >= total) return;  // no work item for this block
    const int n2 = 2 * t.n_pw_nodes;
    for (int m = threadIdx.x; m < K; m += blockDim.x) ac_s[m] = a.s.accept[m];
    for (int w = threadIdx.x; w < 2 * K * kApplyLv; w += blockDim.x) {
        const int li = w % kApplyLv, me = w / kApplyLv, m = me >> 1, ep = me & 1;
        const TrialInfo& ti = a.s.info[m];
        const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
        const int l = 2 + li;
        pnode[w] = (l <= lam && l <= D - 2) ? (ep == 0 ? ti.path_t[l] : ti.path_o[l]) : -1;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int it = bid; it < total; it += nblk) {
        bool is_root = false;
        int chunk, part_c = 0, nparts = 1, key = 0;
        if (it < n_root) {
            is_root = true;
            chunk = it / sr;
            part_c = it % sr;
            nparts = sr;
        } else {
            const int r = it - n_root, arr = r / nDb;
            key = a.s.dest[arr < nH ? arr : n2 - 1 - (arr - nH)];
            chunk = r % nDb;
        }
        const bool heavy = nparts > 1;
        const int slot = key >> 1, kind = is_root ? 2 : key & 1;
        const int node = is_root ? 0 : a.pn.node[slot];
        const int l = is_root ? 0 : a.pn.level[slot], li = l - 2;
        const int nl = 1 << l;
        const int d = is_root ? t.dw : t.dd;
        const int nrows = is_root ? rows_w : rows_d;
        const int e0 = chunk * RPI * d, ne = min(RPI, nrows - chunk * RPI) * d;
        Cplx<T>* dst = (is_root ? t.root : (kind == 0 ? t.out : t.inc) + (size_t)slot * t.Md) + e0;
        const T sy = (is_root || kind == 0) ? T(1) : T(-1);  // in[C] += conj(tau s)
        // the item's elements e = threadIdx.x + TB j (at most XMAX), in flight
        // while the contributions are listed and staged
        Cplx<T> pf[XMAX];
        if (!heavy) {
#pragma unroll
            for (int j = 0; j < XMAX; ++j) {
                const int e = threadIdx.x + TB * j;
                if (e < ne) pf[j] = dst[e];
            }
        }
        unsigned long long tp = wall_clock64();
        __syncthreads();  // previous item's lists, rows and sums fully consumed
        if (threadIdx.x == 0) mcount = 0;
        __syncthreads();
        for (int c0 = 0; c0 < 2 * K; c0 += blockDim.x) {
            const int w = c0 + threadIdx.x;
            int code = -1;
            if (w < 2 * K) {
                const int m = w >> 1, ep = w & 1;
                if (ac_s[m]) {
                    if (is_root) {
                        code = 13;
                    } else {
                        int B;
                        if (li < kApplyLv) {
                            B = pnode[(2 * m + ep) * kApplyLv + li];
                        } else {
                            const TrialInfo& ti = a.s.info[m];
                            B = l <= (ep == 0 ? ti.lam_t : ti.lam_o) ? (ep == 0 ? ti.path_t[l] : ti.path_o[l]) : -1;
                        }
                        if (B >= 0) {
                            if (B == node) {
                                code = 13;
                            } else if (kind == 1) {
                                const int px = image_shift(t.cell[3 * B], t.cell[3 * node], nl);
                                const int py = image_shift(t.cell[3 * B + 1], t.cell[3 * node + 1], nl);
                                const int pz = image_shift(t.cell[3 * B + 2], t.cell[3 * node + 2], nl);
                                if (px != 2 && py != 2 && pz != 2) code = (px + 1) + 3 * (py + 1) + 9 * (pz + 1);
                            }
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
            if (f) {
                const int pos = off + __popcll(bal & lt);
                mlist[pos] = code;
                const int m = code >> 8, ep = (code >> 7) & 1, cc = code & 127;
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    const int sh = ax == 0 ? cc % 3 - 1 : (ax == 1 ? (cc / 3) % 3 - 1 : cc / 9 - 1);
                    mrow[3 * pos + ax] =
                        m * stride + (is_root ? root_row(ep, ax) : level_row(nL, ep, li, ax, sh));
                }
            }
            __syncthreads();
            if (threadIdx.x == 0)
                for (int k = 0; k < (int)(blockDim.x >> 6); ++k) mcount += wtot[k];
            __syncthreads();
        }
        const int nm_all = mcount;
        const int cbeg = part_c * nm_all / nparts, cend = (part_c + 1) * nm_all / nparts;
        const int nm = cend - cbeg;
        const int* mlc = mlist + cbeg;
        if (threadIdx.x == 0) probe_add(a.s, kPrApplyMatch, tp);
        if (heavy && nm == 0) continue;  // uniform across the block
        const int part = threadIdx.x % LPR, lr = threadIdx.x / LPR, row = chunk * RPI + lr;
        const bool on = row < nrows;
        const int y = on ? row % d : 0, z = on ? row / d : 0;
        const int hx = (d + LPR - 1) / LPR, x0 = part * hx;
        const int nx = on ? max(0, min(hx, d - x0)) : 0;
        T ar[XMAX], ai[XMAX];
#pragma unroll
        for (int x = 0; x < XMAX; ++x) ar[x] = ai[x] = T(0);
        // Row tiles of kApplyTile contributions (3 table rows each, RW lanes
        // per row), double-buffered: the global loads of tile n+1 are in
        // flight while tile n is accumulated.
        constexpr int RW = LPR * XMAX <= 32 ? 32 : 64;  // lanes per staged row (d <= LPR XMAX)
        constexpr int RPP = TB / RW;                    // staged rows per pass
        constexpr int R = (3 * kApplyTile + RPP - 1) / RPP;
        const int mm = threadIdx.x % RW, rr0 = threadIdx.x / RW;
        const Cplx<T>* tbase = reinterpret_cast<const Cplx<T>*>(a.s.tabs);
        const int* mrc = mrow + 3 * cbeg;
        T vr[R], vi[R];  // scalar halves: a complex array captured by the lambdas would live in scratch
        auto load_tile = [&](int k0) {
            const int nr = 3 * min(nm - k0, kApplyTile);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int rr = rr0 + r * RPP;
                if (rr < nr && mm < d) {
                    const Cplx<T> c = tbase[mrc[3 * k0 + rr] + mm];
                    vr[r] = c.x;
                    vi[r] = c.y;
                }
            }
        };
        auto store_tile = [&](int k0, Cplx<T>* dstrows) {
            const int nr = 3 * min(nm - k0, kApplyTile);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int rr = rr0 + r * RPP;
                if (rr < nr && mm < d) dstrows[rr * dmax + mm] = {vr[r], vi[r]};
            }
        };
        if (nm > 0) {
            load_tile(0);
            store_tile(0, rows2);
        }
        int buf = 0;
        for (int k0 = 0; k0 < nm; k0 += kApplyTile) {
            const int nt = nm - k0 < kApplyTile ? nm - k0 : kApplyTile;
            const bool more = k0 + kApplyTile < nm;
            if (more) load_tile(k0 + kApplyTile);  // in flight during the accumulation
            __syncthreads();                        // tile k0 stored by every thread
            const Cplx<T>* rows = rows2 + buf * kApplyTile * 3 * dmax;
            // No per-x guard: a lane accumulates all XMAX values (those past its
            // share of the row read on into the following rows or the padding
            // after the tiles, and are never written), so the row's LDS reads
            // are issued together instead of one branch and wait per x.
            for (int k = 0; k < nt; ++k) {
                const Cplx<T>* Rw = rows + k * 3 * dmax;
                const T sg = ((mlc[k0 + k] >> 7) & 1) ? T(-1) : T(1);
                Cplx<T> yz = cmul(Rw[dmax + y], Rw[2 * dmax + z]);
                yz.x *= sg;
                yz.y *= sg;
                Cplx<T> xv[XMAX];
#pragma unroll
                for (int x = 0; x < XMAX; ++x) xv[x] = Rw[x0 + x];
#pragma unroll
                for (int x = 0; x < XMAX; ++x) {
                    ar[x] = fma(xv[x].x, yz.x, ar[x]);
                    ar[x] = fma(-xv[x].y, yz.y, ar[x]);
                    ai[x] = fma(xv[x].x, yz.y, ai[x]);
                    ai[x] = fma(xv[x].y, yz.x, ai[x]);
                }
            }
            if (more) store_tile(k0 + kApplyTile, rows2 + (buf ^ 1) * kApplyTile * 3 * dmax);
            buf ^= 1;
        }
        if (threadIdx.x == 0) probe_add(a.s, kPrApplyTiles, tp);
        __syncthreads();  // the last tile is consumed: its space takes the sums
        // sums to LDS in the array's order, then a coalesced write
#pragma unroll
        for (int x = 0; x < XMAX; ++x)
            if (x < nx) S[lr * d + x0 + x] = {ar[x], ai[x]};
        __syncthreads();
        if (heavy) {
            for (int e = threadIdx.x; e < ne; e += TB) {
                const Cplx<T> v = S[e];
                atomicAdd(&dst[e].x, v.x);
                atomicAdd(&dst[e].y, sy * v.y);
            }
        } else {
#pragma unroll
            for (int j = 0; j < XMAX; ++j) {
                const int e = threadIdx.x + TB * j;
                if (e < ne) {
                    const Cplx<T> v = S[e];
                    dst[e] = {pf[j].x + v.x, pf[j].y + sy * v.y};
                }
            }
        }
        if (threadIdx.x == 0) probe_add(a.s, kPrApplyRmw, tp);
    }
}

// The update on the matrix cores (FP32 grids up to 32 points per axis).
// Work items: the root's row blocks (its contributions split over items when
// many), each row block of the level-2 incoming arrays (front of s.dest: the
// most contributions), and every other rewritten array as one item over all
// its row blocks, so that its contributions are listed and their rows staged
// once.  Per row block of 128 grid rows: the sums S[r][x] = sum_c P_c[r] X_c[x]
// (P_c[r] = s_c Y_c[y(r)] Z_c[z(r)]) as a complex (rows by contributions by x)
// product, three v_mfma_f32_16x16x4f32 per complex 16 x 16 x 4 step, wave wv
// owning row tiles 2 wv and 2 wv + 1 and both column tiles; each lane adds
// its accumulators straight to the array elements it holds in the D layout
// (loaded while the first tile is staged), so the write back needs no LDS
// pass.
constexpr int kApplyMfRows = 128;
__host__ __device__ inline size_t apply_shmem_mf(int dmax, int K) {
    return sizeof(Cplx<float>) * (2 * (size_t)kApplyTile * 3 * dmax + kApplyPad) +
           sizeof(int) * 2 * (size_t)K * kApplyLv;
}
template <bool RB>
__device__ __forceinline__ void apply_body_mf(const DevTree<float>& t, const McArgs& a, int lead = 0) {
    extern __shared__ char apply_dyn[];
    __shared__ int mlist[2 * kMaxBatch];     // contributions: move << 8 | endpoint << 7 | image code
    __shared__ int mrow[3 * 2 * kMaxBatch];  // their three table rows (x, y, z), offsets into s.tabs
    __shared__ int wtot[TB / 64];
    __shared__ int mcount;
    __shared__ int ac_s[kMaxBatch];
    // the error flag and the plan in one round trip
    const int err = *t.err, nD = a.s.plan[0], nacc = a.s.plan[1], nH = a.s.plan[2];
    if (err || !nacc) return;
    if constexpr (RB) {
        if ((int)blockIdx.x == lead) {
            relocate_batch(t, a);
            return;
        }
    }
    // blocks before lead have another role (the next step's set-up)
    const int bid = (int)blockIdx.x - lead - (RB ? 1 : 0), nblk = (int)gridDim.x - lead - (RB ? 1 : 0);
    const int K = a.K, D = t.max_depth, nL = n_levels_pw(D), stride = table_stride(D);
    constexpr int RPI = kApplyMfRows;
    const int dmax = t.dw > t.dd ? t.dw : t.dd;
    Cplx<float>* rows2 = reinterpret_cast<Cplx<float>*>(apply_dyn);  // two tiles of rows at stride dmax
    // (move, endpoint, level - 2) -> packed cell of the endpoint's box, -1 past its leaf
    int* pcell = reinterpret_cast<int*>(rows2 + 2 * kApplyTile * 3 * dmax + kApplyPad);
    const int rows_w = t.dw * (t.n_window + 1), rows_d = t.dd * (t.n_diff + 1);
    const int nWb = (rows_w + RPI - 1) / RPI, nDb = (rows_d + RPI - 1) / RPI;
    const int sr = 2 * nacc > kApplyHeavy ? kApplySplit : 1;
    const int n_root = nWb * sr, nF = nH * nDb;
    const int total = n_root + nF + (nD - nH);
    if (bid >= total) return;  // no work item for this block
    const int n2 = 2 * t.n_pw_nodes;
    for (int m = threadIdx.x; m < K; m += blockDim.x) ac_s[m] = a.s.accept[m];
    // the endpoints' path boxes by their cells (loaded here, while the
    // first item's array and its box are looked up)
    for (int w = threadIdx.x; w < 2 * K * kApplyLv; w += blockDim.x) {
        const int li = w % kApplyLv, me = w / kApplyLv, m = me >> 1, ep = me & 1;
        const TrialInfo& ti = a.s.info[m];
        const int lam = ep == 0 ? ti.lam_t : ti.lam_o;
        const int l = 2 + li;
        const bool on = l <= lam && l <= D - 2;
        const int B = on ? (ep == 0 ? ti.path_t[l] : ti.path_o[l]) : 0;
        const int cx = t.cell[3 * B], cy = t.cell[3 * B + 1], cz = t.cell[3 * B + 2];
        pcell[w] = on ? pack_cell(cx, cy, cz) : -1;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, li16 = lane & 15, kq = lane >> 4;
    for (int it = bid; it < total; it += nblk) {
        bool is_root = false;
        int c_lo, c_hi, part_c = 0, nparts = 1, key = 0;
        if (it < n_root) {
            is_root = true;
            c_lo = it / sr;
            c_hi = c_lo + 1;
            part_c = it % sr;
            nparts = sr;
        } else if (it < n_root + nF) {
            const int r = it - n_root;
            key = a.s.dest[r / nDb];
            c_lo = r % nDb;
            c_hi = c_lo + 1;
        } else {
            key = a.s.dest[n2 - 1 - (it - n_root - nF)];
            c_lo = 0;
            c_hi = nDb;
        }
        const bool heavy = nparts > 1;
        const int slot = key >> 1, kind = is_root ? 2 : key & 1;
        int l = 0, ncx = 0, ncy = 0, ncz = 0;  // the array box's level and cell
        if (!is_root) {
            const int node = a.pn.node[slot];
            l = a.pn.level[slot];
            ncx = t.cell[3 * node];
            ncy = t.cell[3 * node + 1];
            ncz = t.cell[3 * node + 2];
        }
        const int li = l - 2, nl = 1 << l;
        const int d = is_root ? t.dw : t.dd;
        const int nrows = is_root ? rows_w : rows_d;
        Cplx<float>* arr = is_root ? t.root : (kind == 0 ? t.out : t.inc) + (size_t)slot * t.Md;
        const float sy = (is_root || kind == 0) ? 1.f : -1.f;  // in[C] += conj(tau s)
        // the elements this lane updates (D layout) of row blocks c and c + 1,
        // the first two loaded with the first tile's rows: loads complete in
        // issue order, so loads issued before the lists would hold up the
        // lists, which need their data first
        Cplx<float> pf[2][2][4], pn[2][2][4];
        auto load_block = [&](int chunk, Cplx<float> (&dst)[2][2][4]) {
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int row = chunk * RPI + (2 * wv + u) * 16 + 4 * kq + q, x = ct * 16 + li16;
                        const bool ok = row < nrows && x < d;
                        dst[u][ct][q] = arr[ok ? row * d + x : 0];
                    }
        };
        unsigned long long tp = wall_clock64();
        __syncthreads();  // previous item's lists and rows fully consumed
        if (threadIdx.x == 0) mcount = 0;
        __syncthreads();
        for (int c0 = 0; c0 < 2 * K; c0 += blockDim.x) {
            const int w = c0 + threadIdx.x;
            int code = -1;
            if (w < 2 * K) {
                const int m = w >> 1, ep = w & 1;
                if (ac_s[m]) {
                    if (is_root) {
                        code = 13;
                    } else {
                        bool on;  // the endpoint has a box at level l, of cell (bx, by, bz)
                        int bx, by, bz;
                        if (li < kApplyLv) {
                            const int pc = pcell[(2 * m + ep) * kApplyLv + li];
                            on = pc >= 0;
                            bx = pc & 1023;
                            by = (pc >> 10) & 1023;
                            bz = pc >> 20;
                        } else {
                            const TrialInfo& ti = a.s.info[m];
                            on = l <= (ep == 0 ? ti.lam_t : ti.lam_o);
                            const int B = on ? (ep == 0 ? ti.path_t[l] : ti.path_o[l]) : 0;
                            bx = t.cell[3 * B];
                            by = t.cell[3 * B + 1];
                            bz = t.cell[3 * B + 2];
                        }
                        if (on) {
                            if (bx == ncx && by == ncy && bz == ncz) {
                                code = 13;
                            } else if (kind == 1) {
                                const int px = image_shift(bx, ncx, nl);
                                const int py = image_shift(by, ncy, nl);
                                const int pz = image_shift(bz, ncz, nl);
                                if (px != 2 && py != 2 && pz != 2) code = (px + 1) + 3 * (py + 1) + 9 * (pz + 1);
                            }
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
            if (f) {
                const int pos = off + __popcll(bal & lt);
                mlist[pos] = code;
                const int m = code >> 8, ep = (code >> 7) & 1, cc = code & 127;
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    const int sh = ax == 0 ? cc % 3 - 1 : (ax == 1 ? (cc / 3) % 3 - 1 : cc / 9 - 1);
                    mrow[3 * pos + ax] =
                        m * stride + (is_root ? root_row(ep, ax) : level_row(nL, ep, li, ax, sh));
                }
            }
            __syncthreads();
            if (threadIdx.x == 0)
                for (int k = 0; k < (int)(blockDim.x >> 6); ++k) mcount += wtot[k];
            __syncthreads();
        }
        const int nm_all = mcount;
        const int cbeg = part_c * nm_all / nparts, cend = (part_c + 1) * nm_all / nparts;
        const int nm = cend - cbeg;
        const int* mlc = mlist + cbeg;
        if (threadIdx.x == 0) probe_add(a.s, kPrApplyMatch, tp);
        if (heavy && nm == 0) continue;  // uniform across the block
        // row tiles of kApplyTile contributions (3 table rows each, 32 lanes
        // per row), double-buffered: the global loads of tile n+1 are in flight
        // while tile n is accumulated; an item with one tile stages it once for
        // all its row blocks
        constexpr int RW = 32, RPP = TB / RW, R = (3 * kApplyTile + RPP - 1) / RPP;
        const int mm = threadIdx.x % RW, rr0 = threadIdx.x / RW;
        const Cplx<float>* tbase = reinterpret_cast<const Cplx<float>*>(a.s.tabs);
        const int* mrc = mrow + 3 * cbeg;
        float vr[R], vi[R];
        auto load_tile = [&](int k0) {
            const int nr = 3 * min(nm - k0, kApplyTile);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int rr = rr0 + r * RPP;
                if (rr < nr && mm < d) {
                    const Cplx<float> c = tbase[mrc[3 * k0 + rr] + mm];
                    vr[r] = c.x;
                    vi[r] = c.y;
                }
            }
        };
        auto store_tile = [&](int k0, Cplx<float>* dstrows) {
            const int nr = 3 * min(nm - k0, kApplyTile);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int rr = rr0 + r * RPP;
                if (rr < nr && mm < d) dstrows[rr * dmax + mm] = {vr[r], vi[r]};
            }
        };
        const bool one_tile = nm <= kApplyTile;
        for (int chunk = c_lo; chunk < c_hi; ++chunk) {
            if (chunk > c_lo && !one_tile) __syncthreads();  // the previous block's tiles are consumed
            int oy[2], oz[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int rg = chunk * RPI + (2 * wv + u) * 16 + li16;
                oy[u] = rg < nrows ? rg % d : 0;
                oz[u] = rg < nrows ? rg / d : 0;
            }
            f32x4 mr[2][2], mi[2][2], m3[2][2];
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) mr[u][ct] = mi[u][ct] = m3[u][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
            if (chunk == c_lo || !one_tile) {
                load_tile(0);
                if (chunk == c_lo && !heavy) {
                    load_block(c_lo, pf);
                    if (c_lo + 1 < c_hi) load_block(c_lo + 1, pn);
                }
                store_tile(0, rows2);
            }
            int buf = 0;
            for (int k0 = 0; k0 < nm; k0 += kApplyTile) {
                const int nt = nm - k0 < kApplyTile ? nm - k0 : kApplyTile;
                const bool more = k0 + kApplyTile < nm;
                if (more) load_tile(k0 + kApplyTile);  // in flight during the accumulation
                __syncthreads();                        // tile k0 stored by every thread
                const Cplx<float>* rows = rows2 + buf * kApplyTile * 3 * dmax;
                for (int c0 = 0; c0 < nt; c0 += 4) {  // uniform; lane group kq takes contribution c0 + kq
                    const int k = c0 + kq;
                    const bool kv = k < nt;
                    const Cplx<float>* Rw = rows + (kv ? k : 0) * 3 * dmax;
                    const float sg = (kv ? ((mlc[k0 + k] >> 7) & 1) ? -1.f : 1.f : 0.f);  // 0: past the tile's end
                    float pr[2], pim[2], xr[2], xim[2];
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        const Cplx<float> p = cmul(Rw[dmax + oy[u]], Rw[2 * dmax + oz[u]]);
                        pr[u] = sg * p.x;
                        pim[u] = sg * p.y;
                    }
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        const int x = ct * 16 + li16;
                        const Cplx<float> xv = Rw[x < d ? x : 0];
                        xr[ct] = x < d ? xv.x : 0.f;
                        xim[ct] = x < d ? xv.y : 0.f;
                    }
                    // complex products with three real ones (Gauss):
                    // K1 = (pr + pi) xr, K2 = pr (xi - xr), K3 = pi (xr + xi),
                    // re = K1 - K3, im = K1 + K2, each K summed on its own
#pragma unroll
                    for (int u = 0; u < 2; ++u)
#pragma unroll
                        for (int ct = 0; ct < 2; ++ct) {
                            mr[u][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(pr[u] + pim[u], xr[ct], mr[u][ct], 0, 0, 0);
                            mi[u][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(pr[u], xim[ct] - xr[ct], mi[u][ct], 0, 0, 0);
                            m3[u][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(pim[u], xr[ct] + xim[ct], m3[u][ct], 0, 0, 0);
                        }
                }
                if (more) store_tile(k0 + kApplyTile, rows2 + (buf ^ 1) * kApplyTile * 3 * dmax);
                buf ^= 1;
            }
            if (threadIdx.x == 0) probe_add(a.s, kPrApplyTiles, tp);
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {  // mr, mi, m3 hold K1, K2, K3
                    mi[u][ct] = mr[u][ct] + mi[u][ct];
                    mr[u][ct] = mr[u][ct] - m3[u][ct];
                }
            // D layout of v_mfma_f32_16x16x4f32: lane holds rows 4 kq + q of
            // column li16.  A light item adds its elements to the sums, moves
            // the next row block's elements up and issues the loads of the
            // one after before its stores, so that no wait for elements
            // waits for stores as well.
            if (!heavy) {
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            mr[u][ct][q] = pf[u][ct][q].x + mr[u][ct][q];
                            mi[u][ct][q] = pf[u][ct][q].y + sy * mi[u][ct][q];
                        }
                if (chunk + 1 < c_hi)
#pragma unroll
                    for (int u = 0; u < 2; ++u)
#pragma unroll
                        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                            for (int q = 0; q < 4; ++q) pf[u][ct][q] = pn[u][ct][q];
                if (chunk + 2 < c_hi) load_block(chunk + 2, pn);
            }
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int row = chunk * RPI + (2 * wv + u) * 16 + 4 * kq + q, x = ct * 16 + li16;
                        if (row >= nrows || x >= d) continue;
                        Cplx<float>* e = arr + row * d + x;
                        if (heavy) {
                            atomicAdd(&e->x, mr[u][ct][q]);
                            atomicAdd(&e->y, sy * mi[u][ct][q]);
                        } else {
                            *e = {mr[u][ct][q], mi[u][ct][q]};
                        }
                    }
            if (threadIdx.x == 0) probe_add(a.s, kPrApplyRmw, tp);
        }
    }
}

// waves per SIMD the compiler keeps registers for: two for the FP32 grids of
// digits <= 6 (the matrix-core update fits in 184 registers without spills;
// three waves spill and gain under 1 %), the compiler's choice otherwise
// (wider grids and FP64 keep more per thread)
template <typename T, int XMAX>
constexpr int apply_min_waves() { return std::is_same<T, float>::value && XMAX <= 13 ? 2 : 1; }

// the FP32 grids of digits <= 6 take the matrix-core update (apply_body_mf)
template <typename T, int XMAX>
__host__ __device__ constexpr bool apply_uses_mf() { return std::is_same<T, float>::value && apply_lpr<T>() * XMAX <= 32; }

template <typename T, int XMAX, bool RB>
__global__ void __launch_bounds__(TB, (apply_min_waves<T, XMAX>())) k_apply(DevTree<T> t, McArgs a) {
    if constexpr (apply_uses_mf<T, XMAX>())
        apply_body_mf<RB>(t, a);
    else
        apply_body<T, XMAX, RB>(t, a);
}

// The update of batched step b and the set-up of step b + 1 in one launch
// (chains whose scan moved the accepted particles, a.scan_reloc): blocks
// [0, an.K) set up the next step's trials, the others rewrite the arrays of
// step b, the two sets running side by side instead of one after the other.
// The set-up reads the particles and the leaves' slots, which the scan left
// final, and writes the next step's trial state (an.s.info, an.s.tabs: the
// other buffers of the pair) and its moves' records; the update reads step
// b's trial state, its plan and the expansion arrays.  PROP: the proposal's
// set-up (features, the learned network), else the uniform one.
template <typename T, int XMAX, bool PROP>
__global__ void __launch_bounds__(TB, (apply_min_waves<T, XMAX>())) k_apply_next(DevTree<T> t, McArgs a, McArgs an) {
    extern __shared__ __align__(16) char next_dyn[];  // the update's dynamic LDS, or the proposal's
    if ((int)blockIdx.x < an.K) {
        if constexpr (PROP)
            propose_setup_body(t, an, blockIdx.x, next_dyn);
        else
            setup_body(t, an, blockIdx.x);
        return;
    }
    if constexpr (apply_uses_mf<T, XMAX>())
        apply_body_mf<false>(t, a, an.K);
    else
        apply_body<T, XMAX, false>(t, a, an.K);
}

// ------------------------------------------------------------------ fused single-move update
// The update kernel of the fused single-move chain (a.ring; src/tree.cpp:447-597).
// Every block sums the trial kernel's partials in the same order and takes
// the same Metropolis decision (dE < -ln(u) / beta, no overlap, room in the
// target leaf as the marking row found it).  Block 0 writes the move's dE and
// decision, moves the particle and sets up move m + 1 in the other ring slot
// (its particle's data are loaded while the decision is pending; a move of
// the same particle is redone from the target).  The other blocks rewrite one
// chunk of kUpd1J * TB elements of the root or of a marked array each: the
// contributions of the two endpoints follow from their box cells at the
// array's level (computed from the positions) and the array box's cell (from
// the marking row), and the array's elements and the phase-row entries are
// loaded before the decision is known; only the write-back waits for it.
constexpr int kUpd1J = 4;

// wave 0: the partial sums of the move's trial blocks.  Lane b loads blocks
// b, b + 64, ... (issued early, into registers: kUpd1P per lane, enough for
// every launch shape of trees the fused chain runs) and upd1_reduce adds them
// per role in block order; every lane ends with the totals.
struct Upd1Part {
    double q0[kUpd1P], q1[kUpd1P], q2[kUpd1P];
};
template <typename T>
__device__ __forceinline__ void upd1_load(const DevTree<T>& t, const McArgs& a, Upd1Part& P) {
    const int per = trial_blocks(t.max_depth, a.tiles), lane = threadIdx.x & 63;
#pragma unroll
    for (int u = 0; u < kUpd1P; ++u) {
        const int b = lane + 64 * u;
        const double* q = a.s.part + kPartStride * (size_t)(b < per ? b : 0);
        P.q0[u] = b < per ? q[0] : 0.0;
        P.q1[u] = b < per ? q[1] : 0.0;
        P.q2[u] = b < per ? q[2] : 0.0;
    }
}
template <typename T>
__device__ __forceinline__ void upd1_reduce(const DevTree<T>& t, const McArgs& a, const Upd1Part& P, double* v) {
    const Tiles& g = a.tiles;
    const int nL = n_levels_pw(t.max_depth), per = trial_blocks(t.max_depth, g), lane = threadIdx.x & 63;
    const int d0 = g.win, d1 = d0 + 2 * nL * g.diff, d2 = d1 + g.near;
#pragma unroll
    for (int c = 0; c < kAccPerTrial; ++c) v[c] = 0.0;
#pragma unroll
    for (int u = 0; u < kUpd1P; ++u) {
        const int b = lane + 64 * u;
        if (b >= per) continue;
        if (b < d0) {
            v[0] += P.q0[u];
        } else if (b < d1) {
            v[1] += P.q0[u];
        } else if (b < d2) {
            v[2] += P.q0[u];
            v[4] += P.q1[u];
            v[6] += P.q2[u];
        } else {
            v[3] += P.q0[u];
            v[5] += P.q1[u];
        }
    }
    for (int o = 1; o < 64; o <<= 1)
#pragma unroll
        for (int c = 0; c < kAccPerTrial; ++c) v[c] += __shfl_xor(v[c], o);
}
template <typename T>
__device__ __forceinline__ void upd1_partials(const DevTree<T>& t, const McArgs& a, double* v) {
    Upd1Part P;
    upd1_load(t, a, P);
    upd1_reduce(t, a, P, v);
}

// the decision from the partial sums v, the trial state's precomputed
// self-pair term and threshold, and the marking row's room flag
template <typename T>
__device__ __forceinline__ int upd1_decide(const DevTree<T>& t, const McArgs& a, const TrialInfo& ti, int m,
                                           const double* v, int room, bool writer) {
    const int ovl = (int)v[6];
    const double2 st = a.s.st[ti.rs];  // self-pair term, threshold
    double dE = v[0] + v[1] + v[2] - v[3] - st.x;  // src/energy_shift.cpp:67
    if (writer && a.ml.parts) {
        double* p = a.ml.parts + 5 * m;
        p[0] = v[0];
        p[1] = v[1];
        p[2] = v[2];
        p[3] = v[3];
        p[4] = st.x;
    }
    if (writer && a.ml.stats) {
        int o2;
        (void)assemble_trial(t, a, m & 1, m, v, &o2);  // the per-move counts
    }
    dE = ovl > 0 ? (double)INFINITY : a.lB * dE;  // a hard-sphere overlap is an infinite energy
    int ok = ovl == 0 && dE < st.y;
    if (ok && !room) {  // the target leaf is full: the chain stops here
        ok = 0;
        if (writer) atomicOr(t.err, (int)kErrLeafOverflow);
    }
    if (writer) {
        a.ml.dE[m] = dE;
        if (a.ml.accept_out) a.ml.accept_out[m] = ok;
    }
    return ok;
}

template <typename T>
__device__ void upd1_block0(const DevTree<T>& t, const McArgs& a) {
    __shared__ int ok_sh, redo;
    const int sl = ring_slot(a);
    const TrialInfo& ti = a.s.info[sl];
    const int m = ti.m;
    const bool has_next = m + 1 < a.n_moves;
    // loads of the relocation and of the next move's particle, in flight with the partials
    const int lt = ti.path_t[ti.lam_t], lo = ti.path_o[ti.lam_o];
    __shared__ int moved_sh, mcnt_sh;
    int s_i = 0, cnt_t = 0, st_t = 0, nxt_i = -1, room = 0;
    if (threadIdx.x == 0) {
        room = a.s.plan[3];
        s_i = t.slot_of[ti.i];
        st_t = t.slot_start[lt];
        cnt_t = t.slot_count[lt];
        nxt_i = has_next ? a.s.info[sl ^ 1].i : -1;  // set up by the trial kernel
    }
    // move m + 2's particle (for the trial kernel after next), in flight with the decision
    const bool has_nn = m + 2 < a.n_moves;
    NextMove r2;
    if (threadIdx.x == 64 && has_nn) load_next_move(t, a, m + 2, r2);
    // the next move's near-list leaves (ids, codes, slot ranges), in flight with the decision
    const TrialInfo& tn = a.s.info[sl ^ 1];
    const int mx = t.max_near;
    constexpr int NPT = kUpd1Near / TB;  // entries per thread (2 max_near <= kUpd1Near, checked by the host)
    int e_nb[NPT], e_code[NPT], e_st[NPT], e_cnt[NPT];
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
        const int w = threadIdx.x + TB * u, ep = w / mx, k = w % mx;
        e_nb[u] = -1;
        if (has_next && w < 2 * mx && k < tn.ncnt[ep]) {
            e_nb[u] = t.near_ids[tn.noff[ep] + k];
            e_code[u] = t.near_img[tn.noff[ep] + k];
        }
    }
#pragma unroll
    for (int u = 0; u < NPT; ++u)
        if (e_nb[u] >= 0) {
            e_st[u] = t.slot_start[e_nb[u]];
            e_cnt[u] = t.slot_count[e_nb[u]];  // before this kernel's relocation (patched below)
        }
    if (threadIdx.x < 64) {
        double v[kAccPerTrial];
        upd1_partials(t, a, v);
        if (threadIdx.x == 0) {
            const int ok = upd1_decide(t, a, ti, m, v, room, true);
            ok_sh = ok;
            if (ok) {  // src/tree.cpp:454-463, 583-596
                if (lt == lo) {
                    t.px[s_i] = ti.t[0];
                    t.py[s_i] = ti.t[1];
                    t.pz[s_i] = ti.t[2];
                } else {  // appended to the target leaf; its old slot becomes a hole
                    const int ns = st_t + cnt_t;
                    t.px[ns] = ti.t[0];
                    t.py[ns] = ti.t[1];
                    t.pz[ns] = ti.t[2];
                    t.pq[ns] = ti.q;
                    t.porig[ns] = ti.i;
                    t.slot_of[ti.i] = ns;
                    t.leaf_of[ti.i] = lt;
                    t.px[s_i] = 0;
                    t.py[s_i] = 0;
                    t.pz[s_i] = 0;
                    t.pq[s_i] = 0;
                    t.porig[s_i] = -1;
                    t.slot_count[lt] = cnt_t + 1;
                }
            }
            redo = ok && nxt_i == ti.i;
            moved_sh = ok && lt != lo ? lt : -1;
            mcnt_sh = cnt_t + 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 64 && has_nn) {
        if (r2.i == ti.i && ok_sh)  // this move's particle: now at its target
            for (int c = 0; c < 3; ++c) r2.o[c] = ti.t[c];
        a.s.nm[m & 1] = r2;  // (m + 2) & 1
    }
    if (redo) {  // the next move is the same particle's, now at this move's target
        prep_body(t, a, m + 1, sl ^ 1, &ti);
        __syncthreads();
        fill_near(t, a, a.s.info[sl ^ 1], sl ^ 1, moved_sh, mcnt_sh);
        return;
    }
    const int moved = moved_sh, mcnt = mcnt_sh;
    if (has_next && threadIdx.x < 2)
        a.s.near[(size_t)(2 * (sl ^ 1) + threadIdx.x) * (mx + 1)] = make_int4(tn.ncnt[threadIdx.x], 0, 0, 0);
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
        const int w = threadIdx.x + TB * u, ep = w / mx, k = w % mx;
        if (e_nb[u] >= 0)
            a.s.near[(size_t)(2 * (sl ^ 1) + ep) * (mx + 1) + 1 + k] =
                make_int4(e_st[u], e_nb[u] == moved ? mcnt : e_cnt[u], e_code[u], e_nb[u]);
    }
}

// Block 1 of the fused update: the self-pair term, threshold and phase rows
// of move m + 1 (whose paths the trial kernel found), beside the array
// items and without waiting for the decision.  When this move relocates the
// next move's particle, block 0 sets the next move up again into rows slot
// 2 + ((m + 1) & 1), and what this block wrote goes unused.
template <typename T>
__device__ void upd1_rows(const DevTree<T>& t, const McArgs& a) {
    const int sl = ring_slot(a);
    if (a.s.info[sl].m + 1 >= a.n_moves) return;  // uniform
    const TrialInfo& tn = a.s.info[sl ^ 1];
    const int rs = sl ^ 1;  // (m + 1) & 1
    if (threadIdx.x == TB - 1)  // a thread that forms no row base
        a.s.st[rs] = make_double2(self_term(t, tn), mh_threshold(tn.u, a.beta));
    form_rows(t, tn, tabs_of<T>(a.s, table_stride(t.max_depth), rs));
}

template <typename T>
__device__ void upd1_body(const DevTree<T>& t, const McArgs& a) {
    __shared__ int ok_sh;
    if (*t.err) return;
    if (blockIdx.x == 0) {
        upd1_block0(t, a);
        return;
    }
    if (blockIdx.x == 1) {
        upd1_rows(t, a);
        return;
    }
    const int sl = ring_slot(a);
    const TrialInfo& ti = a.s.info[sl];
    const int m = ti.m;
    Upd1Part P;
    if (threadIdx.x < 64) upd1_load(t, a, P);  // in flight with everything below
    const int nD = a.s.plan[0], room = a.s.plan[3];
    constexpr int CH = kUpd1J * TB;  // elements per item
    const int nR = (t.Mw + CH - 1) / CH, nA = (t.Md + CH - 1) / CH;
    const int it = blockIdx.x - 2;
    const bool have = it < nR + nD * nA;  // uniform
    if (!have) return;  // the grid is sized for the most arrays a move can mark
    const int D = t.max_depth, nL = n_levels_pw(D);
    const Cplx<T>* tb = tabs_of<T>(a.s, table_stride(D), ti.rs);
    bool is_root = true;
    int chunk = it, key = 0, l = 0, ccell[3] = {0, 0, 0};
    if (have && it >= nR) {
        const int r = it - nR, arr = r / nA;
        chunk = r % nA;
        key = a.s.dest[arr];
        const int di = a.s.dinfo[arr];
        is_root = false;
        l = di & 15;
        ccell[0] = (di >> 4) & 511;
        ccell[1] = (di >> 13) & 511;
        ccell[2] = (di >> 22) & 511;
    }
    const int slot = key >> 1, kind = is_root ? 2 : key & 1;
    const int d = is_root ? t.dw : t.dd, M = is_root ? t.Mw : t.Md;
    Cplx<T>* dst = is_root ? t.root : (kind == 0 ? t.out : t.inc) + (size_t)slot * t.Md;
    // contribution of endpoint ep (0 target, 1 origin): its rows at the
    // image shift of its box next to the array's box, sign -1 at the origin,
    // 0 when the endpoint does not touch the array (rows then any valid ones)
    const Cplx<T>* rx[2];
    const Cplx<T>* ry[2];
    const Cplx<T>* rz[2];
    T sg[2];
#pragma unroll
    for (int ep = 0; ep < 2; ++ep) {
        rx[ep] = tb + root_row(ep, 0);
        ry[ep] = tb + root_row(ep, 1);
        rz[ep] = tb + root_row(ep, 2);
        sg[ep] = have && is_root ? (ep == 0 ? T(1) : T(-1)) : T(0);
        if (!have || is_root || l > (ep == 0 ? ti.lam_t : ti.lam_o)) continue;
        const double* x = ep == 0 ? ti.t : ti.o;
        const int nl = 1 << l;
        int sh[3];
        bool same = true, adj = true;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int bc = cell_at(x[c], t.L, l);
            same = same && bc == ccell[c];
            sh[c] = image_shift(bc, ccell[c], nl);
            adj = adj && sh[c] != 2;
        }
        if (!(same || (kind == 1 && adj))) continue;
        const int li = l - 2;
        rx[ep] = tb + level_row(nL, ep, li, 0, same ? 0 : sh[0]);
        ry[ep] = tb + level_row(nL, ep, li, 1, same ? 0 : sh[1]);
        rz[ep] = tb + level_row(nL, ep, li, 2, same ? 0 : sh[2]);
        sg[ep] = ep == 0 ? T(1) : T(-1);
    }
    // the chunk's elements and their phase-row entries, all loads in flight
    // together; then the trial's partial sums (wave 0) and the decision
    Cplx<T> pf[kUpd1J], px_[kUpd1J][2], py_[kUpd1J][2], pz_[kUpd1J][2];
#pragma unroll
    for (int j = 0; j < kUpd1J; ++j) {
        const int e = min(chunk * CH + threadIdx.x + TB * j, M - 1);
        const int x = e % d, r = e / d, y = r % d, z = r / d;
        pf[j] = dst[have ? e : 0];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            px_[j][c] = rx[c][x];
            py_[j][c] = ry[c][y];
            pz_[j][c] = rz[c][z];
        }
    }
    double v[kAccPerTrial];
    if (threadIdx.x < 64) upd1_reduce(t, a, P, v);
    Cplx<T> val[kUpd1J];
#pragma unroll
    for (int j = 0; j < kUpd1J; ++j) {
        val[j] = {T(0), T(0)};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (sg[c] == T(0)) continue;  // rows of an absent contribution may be unformed
            Cplx<T> yz = cmul(py_[j][c], pz_[j][c]);
            yz.x *= sg[c];
            yz.y *= sg[c];
            const Cplx<T> xv = px_[j][c];
            val[j].x = fma(xv.x, yz.x, val[j].x);
            val[j].x = fma(-xv.y, yz.y, val[j].x);
            val[j].y = fma(xv.x, yz.y, val[j].y);
            val[j].y = fma(xv.y, yz.x, val[j].y);
        }
    }
    if (threadIdx.x == 0) ok_sh = upd1_decide(t, a, ti, m, v, room, false);
    __syncthreads();
    if (!ok_sh || !have) return;
    const T sy = kind == 1 ? T(-1) : T(1);  // in[C] += conj(tau s)
#pragma unroll
    for (int j = 0; j < kUpd1J; ++j) {
        const int e = chunk * CH + threadIdx.x + TB * j;
        if (e < M) dst[e] = {pf[j].x + val[j].x, pf[j].y + sy * val[j].y};
    }
}

template <typename T>
__global__ void __launch_bounds__(TB) k_update1(DevTree<T> t, McArgs a) {
    upd1_body(t, a);
}

// The update of a replica step: one move per replica (replica blockIdx.z),
// so an element of a rewritten array takes at most two contributions, the
// target's and the origin's, straight from the trial's phase rows.  Work
// items are chunks of kRep1J TB consecutive elements: the root array, then
// the arrays the decision listed (s.dest); block b takes items b, b +
// gridDim.x, ...  Per item the block stages the x rows of both endpoints and
// the products s Y[y] Z[z] of the grid rows the chunk spans in LDS; each
// thread then adds X_t[x] (s Y Z)_t + X_o[x] (s Y Z)_o to its elements
// (conjugated for an incoming array), whose loads were issued first.
constexpr int kRep1J = 8;                      // elements per thread and item
constexpr int kRep1Rows = kRep1J * TB / 4 + 2;  // grid rows a chunk spans (grid rows hold >= 5 elements)
// items of a block whose descriptions are formed together (one thread each)
constexpr int kRep1Pass = 64;

// bx, gx: this block's index and the number of update blocks of the replica.
// The block's items (bx, bx + gx, ...) are described first, a pass of up to
// kRep1Pass at a time, one thread per item (the array, the chunk, and per
// endpoint the table rows at its image shift and the sign), so that the
// chain of dependent lookups (plan entry, box, cells) is paid once per pass;
// then every item stages its rows, and the elements of the next item are
// loaded before the current item's sums are formed and stored.
template <typename T>
__device__ __forceinline__ void apply1_rep_body(const DevTree<T>& t, const McArgs& a, int bx, int gx) {
    __shared__ Cplx<T> sx[2][kMaxGridD], syz[2][kRep1Rows];
    __shared__ int ik[kRep1Pass];       // kind << 28 | slot (kind 2: the root)
    __shared__ int ie0[kRep1Pass];      // first element of the item's chunk
    __shared__ int irow[kRep1Pass][6];  // table rows (x, y, z) of the target, then the origin
    __shared__ int isg[kRep1Pass];      // signs: (sg_t + 1) | (sg_o + 1) << 2
    if (*t.err) return;
    const int nD = a.s.plan[0], nacc = a.s.plan[1], nH = a.s.plan[2];
    if (!nacc) return;  // uniform: rejected, or no move left
    constexpr int CH = kRep1J * TB;  // elements per item
    const int nR = (t.Mw + CH - 1) / CH, nA = (t.Md + CH - 1) / CH;
    const int total = nR + nD * nA;
    const TrialInfo& ti = a.s.info[0];
    const int D = t.max_depth, nL = n_levels_pw(D), n2 = 2 * t.n_pw_nodes;
    const Cplx<T>* tb = tabs_of<T>(a.s, table_stride(D), 0);
    const int nk = total > bx ? (total - bx + gx - 1) / gx : 0;  // items of this block
    auto dst_of = [&](int k) -> Cplx<T>* {
        const int kind = ik[k] >> 28, slot = ik[k] & 0x0fffffff;
        return kind == 2 ? t.root : (kind == 0 ? t.out : t.inc) + (size_t)slot * t.Md;
    };
    for (int p0 = 0; p0 < nk; p0 += kRep1Pass) {
        const int np = min(kRep1Pass, nk - p0);
        __syncthreads();  // the previous pass's descriptions are consumed
        if ((int)threadIdx.x < np) {
            const int it = bx + (p0 + (int)threadIdx.x) * gx;
            const bool is_root = it < nR;
            int chunk = it, key = 0;
            if (!is_root) {
                const int r = it - nR, arr = r / nA;
                chunk = r % nA;
                key = a.s.dest[arr < nH ? arr : n2 - 1 - (arr - nH)];
            }
            const int slot = key >> 1, kind = is_root ? 2 : key & 1;
            const int node = is_root ? 0 : a.pn.node[slot];
            const int l = is_root ? 0 : a.pn.level[slot];
            int rows[6], sgb = 0;
            // rows of endpoint ep (0 target, 1 origin) at the image shift of
            // its box next to the array's box; sign -1 at the origin, 0 when
            // the endpoint does not touch the array
#pragma unroll
            for (int ep = 0; ep < 2; ++ep) {
                int sg = is_root ? (ep == 0 ? 1 : -1) : 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) rows[3 * ep + c] = root_row(ep, c);
                if (!is_root && l <= (ep == 0 ? ti.lam_t : ti.lam_o)) {
                    const double* x = ep == 0 ? ti.t : ti.o;
                    const int nl = 1 << l;
                    int sh[3];
                    bool same = true, adj = true;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const int bc = cell_at(x[c], t.L, l);
                        const int cc = t.cell[3 * node + c];
                        same = same && bc == cc;
                        sh[c] = image_shift(bc, cc, nl);
                        adj = adj && sh[c] != 2;
                    }
                    if (same || (kind == 1 && adj)) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) rows[3 * ep + c] = level_row(nL, ep, l - 2, c, same ? 0 : sh[c]);
                        sg = ep == 0 ? 1 : -1;
                    }
                }
                sgb |= (sg + 1) << (2 * ep);
            }
            ik[threadIdx.x] = kind << 28 | slot;
            ie0[threadIdx.x] = chunk * CH;
            isg[threadIdx.x] = sgb;
#pragma unroll
            for (int c = 0; c < 6; ++c) irow[threadIdx.x][c] = rows[c];
        }
        __syncthreads();
        // the first item's elements
        Cplx<T> pf[kRep1J];
        {
            const Cplx<T>* dst = dst_of(0);
            const int M = (ik[0] >> 28) == 2 ? t.Mw : t.Md, e0 = ie0[0];
#pragma unroll
            for (int j = 0; j < kRep1J; ++j) pf[j] = dst[min(e0 + (int)threadIdx.x + TB * j, M - 1)];
        }
        for (int k = 0; k < np; ++k) {
            const bool is_root = (ik[k] >> 28) == 2;
            const int kind = ik[k] >> 28;
            const int d = is_root ? t.dw : t.dd, M = is_root ? t.Mw : t.Md;
            const int e0 = ie0[k], sgb = isg[k];
            const T sg[2] = {T((sgb & 3) - 1), T(((sgb >> 2) & 3) - 1)};
            Cplx<T>* dst = dst_of(k);
            const int r0 = e0 / d, nr = min((min(e0 + CH, M) - 1) / d - r0 + 1, kRep1Rows);
            __syncthreads();  // the previous item's staged rows are consumed
            for (int q = threadIdx.x; q < 2 * (d + nr); q += TB) {
                const int ep = q & 1, w = q >> 1;
                if (w < d) {
                    sx[ep][w] = tb[irow[k][3 * ep] + w];
                } else {
                    const int row = r0 + w - d, y = row % d, z = row / d;
                    Cplx<T> yz = cmul(tb[irow[k][3 * ep + 1] + y], tb[irow[k][3 * ep + 2] + z]);
                    yz.x *= sg[ep];
                    yz.y *= sg[ep];
                    syz[ep][w - d] = yz;
                }
            }
            __syncthreads();
            // the next item's elements, in flight while this one is summed and stored
            Cplx<T> pn[kRep1J];
            if (k + 1 < np) {
                const Cplx<T>* dn = dst_of(k + 1);
                const int Mn = (ik[k + 1] >> 28) == 2 ? t.Mw : t.Md, en = ie0[k + 1];
#pragma unroll
                for (int j = 0; j < kRep1J; ++j) pn[j] = dn[min(en + (int)threadIdx.x + TB * j, Mn - 1)];
            }
            const T sy = kind == 1 ? T(-1) : T(1);  // in[C] += conj(tau s)
#pragma unroll
            for (int j = 0; j < kRep1J; ++j) {
                const int e = e0 + threadIdx.x + TB * j;
                const int x = e % d, rr = e / d - r0;
                Cplx<T> val = {T(0), T(0)};
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    if (sg[c] == T(0)) continue;
                    const Cplx<T> yz = syz[c][rr < nr ? rr : 0];
                    const Cplx<T> xv = sx[c][x];
                    val.x = fma(xv.x, yz.x, val.x);
                    val.x = fma(-xv.y, yz.y, val.x);
                    val.y = fma(xv.x, yz.y, val.y);
                    val.y = fma(xv.y, yz.x, val.y);
                }
                if (e < M) dst[e] = {pf[j].x + val.x, pf[j].y + sy * val.y};
            }
            if (k + 1 < np) {
#pragma unroll
                for (int j = 0; j < kRep1J; ++j) pf[j] = pn[j];
            }
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(TB) k_apply1_rep(const DevTree<T>* __restrict__ ts, const McArgs* __restrict__ as) {
    apply1_rep_body(ts[blockIdx.z], as[blockIdx.z], blockIdx.x, gridDim.x);
}

// The update of every replica's step and the set-up of its next step (asn:
// the next step's arguments, its trial state in the other buffers of the
// pair) in one launch: block 0 of each replica's row sets up, the others
// update.  The decision kernel moved the particle and advanced the move
// counter, so the set-up reads the next move and final positions.
template <typename T>
__global__ void __launch_bounds__(TB) k_apply1_next_rep(const DevTree<T>* __restrict__ ts,
                                                      const McArgs* __restrict__ as,
                                                      const McArgs* __restrict__ asn) {
    if (blockIdx.x == 0) {
        setup_body(ts[blockIdx.z], asn[blockIdx.z], 0);
        return;
    }
    apply1_rep_body(ts[blockIdx.z], as[blockIdx.z], (int)blockIdx.x - 1, (int)gridDim.x - 1);
}

// grids up to 26 (float) or 28 (double) points per axis use the narrow variant
template <typename T>
constexpr int kApplyXSmall = apply_xmax<T>(sizeof(T) == 8 ? 28 : 26);
template <typename T>
constexpr int kApplyXLarge = apply_xmax<T>(kMaxGridD);
// a lane's reads reach x0 + XMAX <= LPR XMAX elements into its last row
static_assert(apply_lpr<float>() * apply_xmax<float>(kMaxGridD) <= kApplyPad, "apply padding");
static_assert(apply_lpr<double>() * apply_xmax<double>(kMaxGridD) <= kApplyPad, "apply padding");

// ------------------------------------------------------------------ launchers
// dynamic shared memory above the default 64 KB must be allowed per kernel
inline void allow_shmem(const void* kernel, size_t bytes) {
    if (bytes > 65536) (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// e0/e1 (may be null) take the kernel's own start and end timestamps
template <typename T>
void launch_propose_setup(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    hipExtLaunchKernelGGL(k_propose_setup<T>, dim3(a.K), dim3(TB), 0, st, e0, e1, 0, t, a);
}

template <typename T>
void launch_setup(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    hipExtLaunchKernelGGL(k_setup<T>, dim3(a.K), dim3(TB), 0, st, e0, e1, 0, t, a);
}

template <typename T>
void launch_trial(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    // Gram rows, trial rows, the marking row (premark: its bitmap is the dynamic shared memory)
    const bool pm = premark(a);
    const int gy = gram_rows(t, a) + a.K + (pm ? 1 : 0) + (pm && a.ring ? 1 : 0);
    const size_t sh = pm ? sizeof(unsigned) * ((2 * (size_t)t.n_pw_nodes + 31) / 32) : 0;
    const dim3 grid(trial_blocks(t.max_depth, a.tiles), gy);
    if (pm)
        hipExtLaunchKernelGGL((k_trial<T, true>), grid, dim3(TB), sh, st, e0, e1, 0, t, a);
    else
        hipExtLaunchKernelGGL((k_trial<T, false>), grid, dim3(TB), 0, st, e0, e1, 0, t, a);
}

template <typename T>
void launch_decide(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    hipExtLaunchKernelGGL(k_decide<T>, dim3(a.K), dim3(TB), 0, st, e0, e1, 0, t, a);
}

template <typename T>
void launch_pair_corr(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    const int nb = a.K + a.K * (a.K - 1) / 2;  // one record block per trial, then the pairs
    if (nb > 0) hipExtLaunchKernelGGL(k_pair_corr<T>, dim3(nb), dim3(kPairTB), 0, st, e0, e1, 0, t, a);
}

template <typename T>
void launch_scan(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    const bool hs = a.sigma2 > 0.0, lr = a.prop.learned || a.move_ctr != nullptr;
    if (hs && lr)
        hipExtLaunchKernelGGL((k_scan<T, true, true>), dim3(1), dim3(TB), 0, st, e0, e1, 0, t, a);
    else if (hs)
        hipExtLaunchKernelGGL((k_scan<T, true, false>), dim3(1), dim3(TB), 0, st, e0, e1, 0, t, a);
    else if (lr)
        hipExtLaunchKernelGGL((k_scan<T, false, true>), dim3(1), dim3(TB), 0, st, e0, e1, 0, t, a);
    else
        hipExtLaunchKernelGGL((k_scan<T, false, false>), dim3(1), dim3(TB), 0, st, e0, e1, 0, t, a);
}

template <typename T>
void launch_apply(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    const int dmax = t.dw > t.dd ? t.dw : t.dd;
    const bool mf = apply_xmax<T>(dmax) <= kApplyXSmall<T> && apply_uses_mf<T, kApplyXSmall<T>>();
    const size_t sh = mf ? apply_shmem_mf(dmax, a.K) : apply_shmem<T>(dmax, a.K);
    const bool rb = apply_relocates(a);
    const int nb = a.tiles.apply + (rb ? 1 : 0);  // + the particle-move block
    auto go = [&](auto kernel) {
        allow_shmem((const void*)kernel, sh);
        hipExtLaunchKernelGGL(kernel, dim3(nb), dim3(TB), sh, st, e0, e1, 0, t, a);
    };
    if (a.ring) {
        // block 0, the row block, then items: the root's chunks and every
        // marked array's (at most 2 endpoints x levels x 28 arrays)
        constexpr int CH = kUpd1J * TB;
        const int nL = t.max_depth >= 4 ? t.max_depth - 3 : 0;
        const int items = (t.Mw + CH - 1) / CH + 2 * nL * 28 * ((t.Md + CH - 1) / CH);
        hipExtLaunchKernelGGL(k_update1<T>, dim3(2 + items), dim3(TB), 0, st, e0, e1, 0, t, a);
        return;
    }
    if (apply_xmax<T>(dmax) <= kApplyXSmall<T>) {
        if (rb)
            go(k_apply<T, kApplyXSmall<T>, true>);
        else
            go(k_apply<T, kApplyXSmall<T>, false>);
    } else {
        if (rb)
            go(k_apply<T, kApplyXLarge<T>, true>);
        else
            go(k_apply<T, kApplyXLarge<T>, false>);
    }
}

template <typename T>
void launch_apply_next(const DevTree<T>& t, const McArgs& a, const McArgs& an, bool prop, hipStream_t st,
                       hipEvent_t e0, hipEvent_t e1) {
    if (apply_relocates(a) || a.ring || !an.scan_reloc)
        throw std::logic_error("launch_apply_next: the scans of the chain must move the particles");
    const int dmax = t.dw > t.dd ? t.dw : t.dd;
    const bool small = apply_xmax<T>(dmax) <= kApplyXSmall<T>;
    const bool mf = small && apply_uses_mf<T, kApplyXSmall<T>>();
    size_t sh = mf ? apply_shmem_mf(dmax, a.K) : apply_shmem<T>(dmax, a.K);
    if (prop && sh < (size_t)kProposeLds) sh = kProposeLds;
    const int nb = an.K + a.tiles.apply;
    auto go = [&](auto kernel) {
        allow_shmem((const void*)kernel, sh);
        hipExtLaunchKernelGGL(kernel, dim3(nb), dim3(TB), sh, st, e0, e1, 0, t, a, an);
    };
    if (small) {
        if (prop)
            go(k_apply_next<T, kApplyXSmall<T>, true>);
        else
            go(k_apply_next<T, kApplyXSmall<T>, false>);
    } else {
        if (prop)
            go(k_apply_next<T, kApplyXLarge<T>, true>);
        else
            go(k_apply_next<T, kApplyXLarge<T>, false>);
    }
}

template <typename T>
void launch_compact(const DevTree<T>& t, const int* leaves, int n_leaves, hipStream_t st) {
    if (n_leaves > 0) hipLaunchKernelGGL(k_compact<T>, dim3(n_leaves), dim3(TB), 0, st, t, leaves);
}

template <typename T>
void launch_replica_steps(const ReplicaLaunch<T>& r, int k, hipStream_t st) {
    const dim3 rep(1, 1, r.n);
    hipLaunchKernelGGL(k_setup_rep<T>, rep, dim3(TB), 0, st, r.trees, r.args);
    for (int j = 0; j < k; ++j) {
        const McArgs* a = (j & 1) ? r.args_alt : r.args;
        const McArgs* an = (j & 1) ? r.args : r.args_alt;
        // one trial row: with 256 replicas the trial kernel fills the chip, so the
        // decision plans the whole update (a marking row measured slower)
        hipLaunchKernelGGL(k_trial_rep<T>, dim3(trial_blocks(r.max_depth, r.tiles), 1, r.n), dim3(TB), 0, st,
                           r.trees, a);
        hipLaunchKernelGGL(k_decide_rep<T>, rep, dim3(TB), 0, st, r.trees, a);  // advances the move counters
        if (j + 1 < k)
            hipLaunchKernelGGL(k_apply1_next_rep<T>, dim3(1 + r.apply_blocks, 1, r.n), dim3(TB), 0, st, r.trees,
                               a, an);
        else
            hipLaunchKernelGGL(k_apply1_rep<T>, dim3(r.apply_blocks, 1, r.n), dim3(TB), 0, st, r.trees, a);
    }
}

void launch_advance(int* move_ctr, int n, hipStream_t st, int inc) {
    hipLaunchKernelGGL(k_advance, dim3((n + TB - 1) / TB), dim3(TB), 0, st, move_ctr, n, inc);
}

template <typename T>
void launch_replica_compact(const ReplicaLaunch<T>& r, hipStream_t st) {
    hipLaunchKernelGGL(k_compact_rep<T>, dim3(r.max_leaves, 1, r.n), dim3(TB), 0, st, r.trees, r.leaves,
                       r.n_leaves);
}

#define INSTANTIATE(T)                                                                \
    template void launch_compact<T>(const DevTree<T>&, const int*, int, hipStream_t); \
    template void launch_decide<T>(const DevTree<T>&, const McArgs&, hipStream_t, hipEvent_t, hipEvent_t);    \
    template void launch_setup<T>(const DevTree<T>&, const McArgs&, hipStream_t, hipEvent_t, hipEvent_t);     \
    template void launch_propose_setup<T>(const DevTree<T>&, const McArgs&, hipStream_t, hipEvent_t, hipEvent_t); \
    template void launch_trial<T>(const DevTree<T>&, const McArgs&, hipStream_t, hipEvent_t, hipEvent_t);     \
    template void launch_pair_corr<T>(const DevTree<T>&, const McArgs&, hipStream_t, hipEvent_t, hipEvent_t); \
    template void launch_scan<T>(const DevTree<T>&, const McArgs&, hipStream_t, hipEvent_t, hipEvent_t);      \
    template void launch_apply<T>(const DevTree<T>&, const McArgs&, hipStream_t, hipEvent_t, hipEvent_t); \
    template void launch_apply_next<T>(const DevTree<T>&, const McArgs&, const McArgs&, bool, hipStream_t, hipEvent_t, \
                                       hipEvent_t);                                                          \
    template void launch_replica_steps<T>(const ReplicaLaunch<T>&, int, hipStream_t);                      \
    template void launch_replica_compact<T>(const ReplicaLaunch<T>&, hipStream_t);
INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace hpdmk_amd
