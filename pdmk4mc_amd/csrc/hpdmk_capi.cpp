// C ABI of libhpdmk_amd.so for one DMK system (declared in include/hpdmk.h):
// owns the device state of the system and sequences the HIP kernels of
// dmk_mc.hip and dmk_build.hip on one stream.  No exception crosses the
// boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include <chrono>

#include "dmk_octree.hpp"
#include "hpdmk_handle.hpp"

using namespace hpdmk_amd;

namespace hpdmk_amd {
namespace capi {

thread_local std::string g_err;

void check_err(Handle* h) {
    int e = 0;
    HIPX(hipMemcpyAsync(&e, h->err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPX(hipStreamSynchronize(h->stream));
    if (e & kErrLeafOverflow)
        throw HpdmkError(HPDMK_ECAPACITY,
                         "an accepted move would overflow a leaf's slot capacity; rebuild the tree from "
                         "hpdmk_get_positions and continue");
}


hipEvent_t take_event(Handle* h) {
    if (h->ev_pool.empty()) {
        hipEvent_t e;
        HIPX(hipEventCreate(&e));
        return e;
    }
    hipEvent_t e = h->ev_pool.back();
    h->ev_pool.pop_back();
    return e;
}

void collect_events(Handle* h) {
    if (h->ev_used.empty()) return;
    HIPX(hipStreamSynchronize(h->stream));
    for (auto& u : h->ev_used) {
        float ms = 0;
        HIPX(hipEventElapsedTime(&ms, u.second.first, u.second.second));
        h->kms[u.first] += ms;
        h->kcount[u.first] += 1;
        h->ev_pool.push_back(u.second.first);
        h->ev_pool.push_back(u.second.second);
    }
    h->ev_used.clear();
}

namespace {
template <typename T>
void fill_dev(Handle* h, DevTree<T>& d, const std::vector<std::vector<double>>& dtabs,
              const std::vector<double>& wtab) {
    auto& o = h->owned;
    std::vector<T> W(wtab.begin(), wtab.end());
    std::vector<T> D, F;
    for (auto& v : dtabs) D.insert(D.end(), v.begin(), v.end());
    // folded tables: the four entries k_x = +-x' dk, k_y = +-y' dk are equal
    // (checked), so sum_x sum_y D A[x] B[y] = sum_x' sum_y' Dfold A'[x'] B'[y']
    // with A'[x'] = A[n + x'] + A[n - x'] (x' > 0)
    const int n = d.n_diff, dd = 2 * n + 1, n1 = n + 1;
    for (auto& v : dtabs)
        for (int z = 0; z < n1; ++z)
            for (int y = 0; y < n1; ++y)
                for (int x = 0; x < n1; ++x) {
                    auto at = [&](int i, int j) { return v[((size_t)z * dd + j) * dd + i]; };
                    const double e = at(n + x, n + y);
                    if (at(n - x, n + y) != e || at(n + x, n - y) != e || at(n - x, n - y) != e)
                        throw HpdmkError(HPDMK_ERUNTIME, "difference table is not symmetric in k_x, k_y");
                    F.push_back((T)e);
                }
    d.Wmat = dupload(W, o);
    d.Dmat = dupload(D, o);
    d.Dfold = dupload(F, o);
    d.root = dalloc<Cplx<T>>(d.Mw, o);
    d.out = dalloc<Cplx<T>>((size_t)d.n_pw_nodes * d.Md, o);
    d.inc = dalloc<Cplx<T>>((size_t)d.n_pw_nodes * d.Md, o);
    HIPX(hipMemset(d.root, 0, sizeof(Cplx<T>) * d.Mw));
    HIPX(hipMemset(d.out, 0, sizeof(Cplx<T>) * std::max<size_t>(1, (size_t)d.n_pw_nodes * d.Md)));
    HIPX(hipMemset(d.inc, 0, sizeof(Cplx<T>) * std::max<size_t>(1, (size_t)d.n_pw_nodes * d.Md)));
}
}  // namespace

Handle* create_handle(HPDMKParams params, int precision, int n_src, const double* r_src, const double* charge,
                      hipStream_t stream) {
    if (!r_src || !charge) throw HpdmkError(HPDMK_EINVAL, "null argument");
    if (precision != HPDMK_FP64 && precision != HPDMK_MIXED)
        throw HpdmkError(HPDMK_EINVAL, "precision must be HPDMK_FP64 or HPDMK_MIXED");
    if (precision == HPDMK_MIXED && params.digits != 3 && params.digits != 6)
        throw HpdmkError(HPDMK_EINVAL, "float only supports 3 and 6 digits accuracy");  // src/tree.cpp:19-23
    if (params.n_per_leaf < 1) throw HpdmkError(HPDMK_EINVAL, "n_per_leaf must be positive");
    auto h = std::make_unique<Handle>();
    h->params = params;
    h->precision = precision;
    try {
        h->pw = pswf_consts(params.digits);
    } catch (const HpdmkError& e) {
        throw HpdmkError(HPDMK_EINVAL, e.what());
    }
    HIPX(hipGetDevice(&h->device));
    if (stream) {
        h->stream = stream;
        h->own_stream = false;
    } else {
        HIPX(hipStreamCreate(&h->stream));
    }
    // the octree topology on the GPU (the host build only for trees deeper
    // than the device build's pre-order keys hold)
    h->tree = build_octree_device(n_src, r_src, params.L, params.n_per_leaf, h->stream);
    if (h->tree.n_nodes == 0) h->tree = build_octree(n_src, r_src, params.L, params.n_per_leaf);
    h->n_src = n_src;
    const Octree& tr = h->tree;
    const int D = tr.max_depth;
    if (D > kMaxLevels - 1) throw HpdmkError(HPDMK_ERUNTIME, "tree deeper than supported");
    auto& o = h->owned;

    // ---- slots: leaf B owns [start, start + cap)
    const int nn = tr.n_nodes;
    std::vector<int> cap(nn, 0), start(nn, 0), count(nn, 0);
    int total = 0;
    for (int i = 0; i < nn; ++i) {
        if (!tr.is_leaf[i]) continue;
        count[i] = tr.leaf_count[i];
        cap[i] = ((std::max(2 * count[i], count[i] + 64) + 3) / 4) * 4;
        start[i] = total;
        total += cap[i];
        ++h->n_leaves;
    }
    std::vector<double> px(total, 0), py(total, 0), pz(total, 0), pq(total, 0);
    std::vector<int> porig(total, -1), slot_of(n_src), leaf_of(n_src);
    {
        int pos = 0;
        for (int i = 0; i < nn; ++i) {
            if (!tr.is_leaf[i]) continue;
            for (int k = 0; k < count[i]; ++k) {
                const int id = tr.order[pos++];
                const int s = start[i] + k;
                px[s] = r_src[3 * id];
                py[s] = r_src[3 * id + 1];
                pz[s] = r_src[3 * id + 2];
                pq[s] = charge[id];
                porig[s] = id;
                slot_of[id] = s;
                leaf_of[id] = i;
            }
        }
    }
    for (int i = 0; i < n_src; ++i) h->Q += charge[i] * charge[i];  // src/tree.cpp:306-309

    // ---- boxes carrying expansions: levels 2..D-2
    std::vector<int> pw_slot(nn, -1), pn_node, pn_level, pn_lo, pn_hi;
    for (int i = 0; i < nn; ++i)
        if (tr.level[i] >= 2 && tr.level[i] <= D - 2) {
            pw_slot[i] = (int)pn_node.size();
            pn_node.push_back(i);
            pn_level.push_back(tr.level[i]);
            pn_lo.push_back(start[tr.first_leaf[i]]);
            pn_hi.push_back(start[tr.last_leaf[i]] + cap[tr.last_leaf[i]]);
        }
    h->n_pw = (int)pn_node.size();
    std::vector<int> leaves;
    int max_near = 0;
    for (int i = 0; i < nn; ++i)
        if (tr.is_leaf[i]) {
            leaves.push_back(i);
            max_near = std::max(max_near, tr.near_off[i + 1] - tr.near_off[i]);
        }

    // ---- scalars shared by both precisions
    auto set_common = [&](auto& d) {
        d.n_src = n_src;
        d.n_nodes = nn;
        d.max_depth = D;
        d.n_window = h->pw.n_window;
        d.n_diff = h->pw.n_diff;
        d.dw = 2 * d.n_window + 1;
        d.dd = 2 * d.n_diff + 1;
        d.Mw = d.dw * d.dw * (d.n_window + 1);
        d.Md = d.dd * d.dd * (d.n_diff + 1);
        d.n_pw_nodes = h->n_pw;
        d.total_slots = total;
        d.max_near = max_near;
        d.L = params.L;
        for (int l = 0; l <= kMaxLevels; ++l) d.boxsize[l] = params.L / std::ldexp(1.0, l);
        d.delta_k[0] = d.delta_k[1] = 2 * M_PI / params.L;
        for (int l = 2; l <= kMaxLevels; ++l) d.delta_k[l] = (2.0 * M_PI / 3.0) / d.boxsize[l];
        const double tp3 = std::pow(2 * M_PI, 3);
        d.cW = std::pow(d.delta_k[0], 3) / (2 * tp3);
        for (int l = 0; l < kMaxLevels; ++l) {
            d.cD[l] = std::pow(d.delta_k[l], 3) / (2 * tp3);
            d.self_diff[l] = h->pw.psi0_0 * (1.0 / (2 * d.boxsize[l + 1] * h->pw.C0) -
                                             1.0 / (2 * d.boxsize[l] * h->pw.C0));
        }
        d.self_window = h->pw.psi0_0 / (2 * d.boxsize[2] * h->pw.C0);
        d.n_poly = (int)h->pw.poly.size();
        for (int k = 0; k < d.n_poly; ++k) d.poly[k] = h->pw.poly[k];
    };
    std::vector<int> col_in(26 * (size_t)nn, -1);
    for (int i = 0; i < nn; ++i)
        for (int c = 0; c < 26; ++c) {
            const int C = tr.colleague[26 * i + c];
            if (C >= 0 && !tr.is_leaf[C] && pw_slot[C] >= 0) col_in[26 * i + c] = pw_slot[C];
        }
    std::vector<int8_t> leaf8(tr.is_leaf.begin(), tr.is_leaf.end());
    const int* child = dupload(tr.child, o);
    const int* level = dupload(tr.level, o);
    const int8_t* isl = dupload(leaf8, o);
    const int* cellp = dupload(tr.cell, o);
    const int* coll = dupload(tr.colleague, o);
    const int* colin = dupload(col_in, o);
    const int* pws = dupload(pw_slot, o);
    const int* noff = dupload(tr.near_off, o);
    const int* nids = dupload(tr.near_ids, o);
    const int8_t* nimg = dupload(tr.near_img, o);
    const int* sstart = dupload(start, o);
    const int* scap = dupload(cap, o);
    int* scount = dupload(count, o);
    double* dpx = dupload(px, o);
    double* dpy = dupload(py, o);
    double* dpz = dupload(pz, o);
    double* dpq = dupload(pq, o);
    int* dporig = dupload(porig, o);
    int* dslot = dupload(slot_of, o);
    int* dleaf = dupload(leaf_of, o);
    h->err = dalloc<int>(1, o);
    HIPX(hipMemset(h->err, 0, sizeof(int)));
    // dense cell -> node table of every level (one round of independent loads per locate)
    const int* node_at = nullptr;
    if (D <= kNodeAtLevels) {
        std::vector<int> tab((((size_t)1 << (3 * D)) - 1) / 7, -1);
        for (int i = 0; i < nn; ++i) {
            const int l = tr.level[i];
            const size_t off = (((size_t)1 << (3 * l)) - 1) / 7;
            tab[off + tr.cell[3 * i] + ((size_t)tr.cell[3 * i + 1] << l) + ((size_t)tr.cell[3 * i + 2] << (2 * l))] = i;
        }
        node_at = dupload(tab, o);
    }
    auto set_ptrs = [&](auto& d) {
        d.node_at = node_at;
        d.child = child;
        d.level = level;
        d.is_leaf = isl;
        d.cell = cellp;
        d.colleague = coll;
        d.col_in_slot = colin;
        d.pw_slot = pws;
        d.near_off = noff;
        d.near_ids = nids;
        d.near_img = nimg;
        d.slot_start = sstart;
        d.slot_cap = scap;
        d.slot_count = scount;
        d.px = dpx;
        d.py = dpy;
        d.pz = dpz;
        d.pq = dpq;
        d.porig = dporig;
        d.slot_of = dslot;
        d.leaf_of = dleaf;
        d.err = h->err;
    };
    std::vector<double> wtab = window_table(h->pw, params.L);
    std::vector<std::vector<double>> dtabs;
    for (int l = 2; l <= D - 2; ++l) dtabs.push_back(difference_table(h->pw, params.L, l));
    if (precision == HPDMK_FP64) {
        set_common(h->d64);
        set_ptrs(h->d64);
        fill_dev(h.get(), h->d64, dtabs, wtab);
    } else {
        set_common(h->d32);
        set_ptrs(h->d32);
        fill_dev(h.get(), h->d32, dtabs, wtab);
    }
    h->pn.node = dupload(pn_node, o);
    h->pn.level = dupload(pn_level, o);
    h->pn.lo = dupload(pn_lo, o);
    h->pn.hi = dupload(pn_hi, o);
    h->leaves = dupload(leaves, o);
    const int Mw = (2 * h->pw.n_window + 1) * (2 * h->pw.n_window + 1) * (h->pw.n_window + 1);
    h->acc_root = dalloc<double>(2 * Mw, o);
    h->energy = dalloc<double>(4, o);
    h->sc.part = dalloc<double>(kPartStride * (size_t)kMaxBatch * max_trial_blocks(D), o);
    h->sc.rec = dalloc<TrialRec>(kMaxBatch, o);
    h->sc.info = dalloc<TrialInfo>(kMaxBatch, o);
    h->info_alt = dalloc<TrialInfo>(kMaxBatch, o);
    h->sc.accept = dalloc<int>(kMaxBatch, o);
    h->sc.corr = dalloc<double>(kMaxBatch * kMaxBatch, o);
    h->sc.gram = precision == HPDMK_FP64 ? nullptr : dalloc<double>((size_t)kGramMaxK * kGramMaxK * gram_blocks(Mw), o);
    h->sc.ovl = dalloc<unsigned char>(kMaxBatch * kMaxBatch, o);
    const size_t tsz = precision == HPDMK_FP64 ? sizeof(Cplx<double>) : sizeof(Cplx<float>);
    h->sc.tabs = dalloc<char>(tsz * (size_t)table_stride(D) * kMaxBatch, o);
    h->tabs_alt = dalloc<char>(tsz * (size_t)table_stride(D) * kMaxBatch, o);
    if (2 * (size_t)h->n_pw > 32 * (size_t)kMarkWords)
        throw HpdmkError(HPDMK_EINVAL, "tree has more expansion boxes than the update planner's bitmap holds");
    h->sc.dest = dalloc<int>(2 * (size_t)h->n_pw + 2, o);
    h->sc.dinfo = dalloc<int>(2 * (size_t)h->n_pw + 2, o);
    h->sc.near = dalloc<int4>(4 * ((size_t)max_near + 1), o);
    h->sc.nm = dalloc<NextMove>(2, o);
    h->sc.st = dalloc<double2>(4, o);
    h->sc.plan = dalloc<int>(4, o);
    h->sc.cand = dalloc<unsigned short>(kScanCand, o);
    h->probe = dalloc<unsigned long long>(2 * kProbes, o);
    HIPX(hipMemset(h->probe, 0, sizeof(unsigned long long) * 2 * kProbes));
    HIPX(hipMemset(h->sc.plan, 0, 4 * sizeof(int)));
    HIPX(hipDeviceSynchronize());
    return h.release();
}

void ensure_moves(Handle* h, MoveBuf& b, int n) {
    if (n <= b.cap) return;
    auto& o = h->owned;
    const int c = std::max(n, 1024);
    b.idx = dalloc<int>(c, o);
    b.d = dalloc<double>(3 * (size_t)c, o);
    b.u = dalloc<double>(c, o);
    b.dE = dalloc<double>(c, o);
    b.dE_stale = dalloc<double>(c, o);
    b.parts = dalloc<double>(5 * (size_t)c, o);
    b.acc = dalloc<int>(c, o);
    b.stats = dalloc<int>(kStatsPerMove * (size_t)c, o);
    b.dprop = dalloc<double>(3 * (size_t)c, o);
    b.logq = dalloc<double>(c, o);
    b.feat = dalloc<double>(kFeat * (size_t)c, o);
    b.feat_t = dalloc<double>(kFeat * (size_t)c, o);
    b.H = dalloc<double>(c, o);
    b.cap = c;
}

void upload_moves(Handle* h, MoveBuf& b, int n, const long* idx, const double* d, const double* u) {
    for (int k = 0; k < n; ++k)
        if (idx[k] < 0 || idx[k] >= h->n_src) throw HpdmkError(HPDMK_EINVAL, "particle index out of range");
    ensure_moves(h, b, n);
    std::vector<int> i32(idx, idx + n);
    b.h_idx.assign(idx, idx + n);
    HIPX(hipMemcpyAsync(b.idx, i32.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->stream));
    HIPX(hipMemcpyAsync(b.d, d, sizeof(double) * 3 * n, hipMemcpyHostToDevice, h->stream));
    if (u) HIPX(hipMemcpyAsync(b.u, u, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    HIPX(hipStreamSynchronize(h->stream));
    b.n = n;
}

// The decisions compare dE with -ln(u) / beta (and add H / beta for a learned
// proposal): beta must be a finite positive number.
void check_beta(double beta) {
    if (!(beta > 0.0) || !std::isfinite(beta))
        throw HpdmkError(HPDMK_EINVAL, "beta must be finite and > 0");
}

McArgs mc_args(Handle* h, const MoveBuf& b, int move0, int K, int mode, double beta) {
    McArgs a{};
    a.s = h->sc;
    a.s.probe = h->profiling >= 2 ? h->probe : nullptr;
    a.ml.idx = b.idx;
    a.ml.d = b.d;
    a.ml.u = b.u;
    a.ml.dE = b.dE;
    a.ml.dE_stale = b.dE_stale;
    a.ml.parts = b.parts;
    a.ml.accept_out = b.acc;
    a.ml.stats = h->profiling ? b.stats : nullptr;  // counting runs in the finalizer: profiled runs only
    a.pn = h->pn;
    a.move0 = move0;
    a.K = K;
    a.mode = mode;
    a.beta = beta;
    a.tiles = h->tiles;
    const bool chain = mode == kTrialMetropolis || mode == kTrialBatch;
    a.prop = Proposal{};
    if (chain && h->prop_rf > 0.0) {
        a.prop.rf = h->prop_rf;
        a.prop.feat = b.feat;
        if (h->prop_learned) {
            a.prop.learned = 1;
            a.prop.w = h->mlp_w;
            a.prop.wp = h->mlp_wp;
            a.prop.feat_t = b.feat_t;
            a.prop.z = b.d;
            a.prop.max_mean = h->prop_max_mean;
            a.prop.smin = h->prop_smin;
            a.prop.smax = h->prop_smax;
            a.prop.dprop = b.dprop;
            a.prop.logq = b.logq;
            a.prop.H = b.H;
            a.ml.d = b.dprop;  // set-up and trial read the drawn displacements
        }
    }
    a.lB = chain ? h->bjerrum : 1.0;
    a.sigma2 = chain ? h->sigma * h->sigma : 0.0;
    return a;
}

// Runs the loaded move list as a Metropolis chain (mode 0 plain, 1 batched;
// batch 0 = the tuned batch size); returns the GPU time in ms.
// a batched step's arguments with its trial state in buffer pair p (0: the
// scratch's own, 1: the alternate): consecutive steps alternate, so that a
// step's set-up can run beside the previous step's update, whose scan moves
// the particles
static McArgs alt_args(const Handle* h, McArgs a, int p) {
    if (p) {
        a.s.info = h->info_alt;
        a.s.tabs = h->tabs_alt;
    }
    a.scan_reloc = 1;
    return a;
}

double run_chain(Handle* h, double beta, int mode, int batch) {
    check_beta(beta);
    MoveBuf& mb = h->chain;
    const int n = mb.n;
    if (n <= 0) throw HpdmkError(HPDMK_EINVAL, "no moves loaded (hpdmk_mc_load)");
    if (mode != 0 && mode != 1) throw HpdmkError(HPDMK_EINVAL, "mode must be 0 (plain) or 1 (batched)");
    if (mode == 1 && batch == 0) batch = h->tuned_batch;
    if (mode == 1 && (batch < 1 || batch > kMaxBatch)) throw HpdmkError(HPDMK_EINVAL, "batch must be in [0, 256]");
    for (int k = 0; k < kKinds; ++k) {
        h->kms[k] = 0;
        h->kcount[k] = 0;
    }
    if (h->profiling >= 2) HIPX(hipMemsetAsync(h->probe, 0, sizeof(unsigned long long) * 2 * kProbes, h->stream));
    // batches never repeat a particle (the corrections assume distinct particles)
    std::vector<std::pair<int, int>> batches;
    if (mode == 1 && !h->prop_learned) {
        std::vector<int> seen(h->n_src, -1);
        int m0 = 0;
        for (int m = 0; m < n; ++m) {
            const int p = (int)mb.h_idx[m];
            if (m - m0 == batch || seen[p] >= m0) {
                batches.push_back({m0, m - m0});
                m0 = m;
            }
            seen[p] = m;
        }
        batches.push_back({m0, n - m0});
    }
    struct Ev {
        hipEvent_t e = nullptr;
        ~Ev() {
            if (e) (void)hipEventDestroy(e);
        }
    } e0, e1;
    HIPX(hipEventCreate(&e0.e));
    HIPX(hipEventCreate(&e1.e));
    HIPX(hipEventRecord(e0.e, h->stream));
    dispatch(h, [&](auto& dv) {
        // whether the next call of compact_if_due(moved) compacts
        auto compact_due = [&](int moved) { return h->since_compact + moved >= kCompactEvery; };
        auto compact_if_due = [&](int moved) {
            h->since_compact += moved;
            if (h->since_compact >= kCompactEvery) {
                launch_compact(dv, h->leaves, h->n_leaves, h->stream);
                h->since_compact = 0;
            }
        };
        const Tiles ftiles = h->tiles_set ? h->tiles : kPlainTiles;
        if (mode == 0 && !(h->prop_rf > 0.0) && h->tree.max_depth <= kFusedMaxDepth &&
            trial_blocks(h->tree.max_depth, ftiles) <= 64 * kUpd1P && 2 * dv.max_near <= kUpd1Near) {
            // fused single-move chain: per move a trial kernel (with the update
            // plan's marking row) and an update kernel whose blocks take the
            // decision themselves; the update kernel's block 0 moves the
            // particle and sets up the next move, so only the first move needs
            // the set-up kernel.  Runs of kPlainGraphMoves moves are replayed
            // from a HIP graph whose kernels read the run's first move from a
            // device counter (advanced once per run).
            McArgs a = mc_args(h, mb, 0, 1, kTrialMetropolis, beta);
            a.ring = 1;
            a.n_moves = n;
            a.tiles = ftiles;
            timed(h, kKSetup, [&](hipEvent_t x, hipEvent_t y) { launch_setup(dv, a, h->stream, x, y); });
            // a slot compaction moves particles inside their leaves: the next
            // move's prepared near-field entries are formed again after it
            auto compact_fused = [&](int moved, int next) {
                h->since_compact += moved;
                if (h->since_compact < kCompactEvery) return;
                launch_compact(dv, h->leaves, h->n_leaves, h->stream);
                h->since_compact = 0;
                if (next < n) {
                    McArgs s = a;
                    s.move0 = next;
                    launch_setup(dv, s, h->stream);
                }
            };
            int m = 0;
            if (!h->profiling && n >= 2 * kPlainGraphMoves) {
                if (!h->step_ctr) h->step_ctr = dalloc<int>(1, h->owned);
                HIPX(hipMemsetAsync(h->step_ctr, 0, sizeof(int), h->stream));
                McArgs g = a;
                g.move_ctr = h->step_ctr;
                struct Graph {
                    hipGraph_t g = nullptr;
                    hipGraphExec_t x = nullptr;
                    ~Graph() {
                        if (x) (void)hipGraphExecDestroy(x);
                        if (g) (void)hipGraphDestroy(g);
                    }
                } gr;
                HIPX(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                for (int k = 0; k < kPlainGraphMoves; ++k) {
                    g.move0 = k;
                    launch_trial(dv, g, h->stream);
                    launch_apply(dv, g, h->stream);
                }
                launch_advance(h->step_ctr, 1, h->stream, kPlainGraphMoves);
                HIPX(hipStreamEndCapture(h->stream, &gr.g));
                HIPX(hipGraphInstantiate(&gr.x, gr.g, nullptr, nullptr, 0));
                for (; m + kPlainGraphMoves <= n; m += kPlainGraphMoves) {
                    HIPX(hipGraphLaunch(gr.x, h->stream));
                    compact_fused(kPlainGraphMoves, m + kPlainGraphMoves);
                }
            }
            for (; m < n; ++m) {
                a.move0 = m;
                timed(h, kKTrial, [&](hipEvent_t x, hipEvent_t y) { launch_trial(dv, a, h->stream, x, y); });
                timed(h, kKUpdate, [&](hipEvent_t x, hipEvent_t y) { launch_apply(dv, a, h->stream, x, y); });
                compact_fused(1, m + 1);
            }
        } else if (mode == 0 && !h->profiling && n >= 2 * kPlainGraphMoves) {
            // plain chain from a HIP graph of kPlainGraphMoves moves: the
            // kernels read the move index from a device counter that the last
            // node of every move advances
            if (!h->step_ctr) h->step_ctr = dalloc<int>(1, h->owned);
            HIPX(hipMemsetAsync(h->step_ctr, 0, sizeof(int), h->stream));
            McArgs a = mc_args(h, mb, 0, 1, kTrialMetropolis, beta);
            a.move_ctr = h->step_ctr;
            a.n_moves = n;
            auto one_move = [&] {
                if (a.prop.rf > 0.0)
                    launch_propose_setup(dv, a, h->stream);
                else
                    launch_setup(dv, a, h->stream);
                launch_trial(dv, a, h->stream);
                launch_decide(dv, a, h->stream);
                launch_apply(dv, a, h->stream);
                launch_advance(h->step_ctr, 1, h->stream);
            };
            struct Graph {
                hipGraph_t g = nullptr;
                hipGraphExec_t x = nullptr;
                ~Graph() {
                    if (x) (void)hipGraphExecDestroy(x);
                    if (g) (void)hipGraphDestroy(g);
                }
            } gr;
            HIPX(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < kPlainGraphMoves; ++k) one_move();
            HIPX(hipStreamEndCapture(h->stream, &gr.g));
            HIPX(hipGraphInstantiate(&gr.x, gr.g, nullptr, nullptr, 0));
            int m = 0;
            for (; m + kPlainGraphMoves <= n; m += kPlainGraphMoves) {
                HIPX(hipGraphLaunch(gr.x, h->stream));
                compact_if_due(kPlainGraphMoves);
            }
            for (; m < n; ++m) {
                one_move();
                compact_if_due(1);
            }
        } else if (mode == 0) {
            for (int m = 0; m < n; ++m) {
                const McArgs a = mc_args(h, mb, m, 1, kTrialMetropolis, beta);
                if (a.prop.rf > 0.0)
                    timed(h, kKPropose,
                          [&](hipEvent_t x, hipEvent_t y) { launch_propose_setup(dv, a, h->stream, x, y); });
                else
                    timed(h, kKSetup, [&](hipEvent_t x, hipEvent_t y) { launch_setup(dv, a, h->stream, x, y); });
                timed(h, kKTrial, [&](hipEvent_t x, hipEvent_t y) { launch_trial(dv, a, h->stream, x, y); });
                timed(h, kKScan, [&](hipEvent_t x, hipEvent_t y) { launch_decide(dv, a, h->stream, x, y); });
                timed(h, kKUpdate, [&](hipEvent_t x, hipEvent_t y) { launch_apply(dv, a, h->stream, x, y); });
                compact_if_due(1);
            }
        } else if (h->prop_learned) {
            // steps decide a variable number of moves (up to the first stale
            // proposal); the device counter tracks the chain, read back after
            // each round of launches
            if (!h->step_ctr) h->step_ctr = dalloc<int>(1, h->owned);
            HIPX(hipMemsetAsync(h->step_ctr, 0, sizeof(int), h->stream));
            int done = 0, par = 0;
            bool ready = false;  // the step's set-up ran beside the previous update
            while (done < n) {
                const int steps = (n - done + batch - 1) / batch;
                for (int k = 0; k < steps; ++k) {
                    auto step_args = [&](int p) {
                        McArgs s = alt_args(h, mc_args(h, mb, 0, batch, kTrialBatch, beta), p);
                        s.move_ctr = h->step_ctr;
                        s.n_moves = n;
                        return s;
                    };
                    const McArgs a = step_args(par);
                    if (!ready)
                        timed(h, kKPropose,
                              [&](hipEvent_t x, hipEvent_t y) { launch_propose_setup(dv, a, h->stream, x, y); });
                    timed(h, kKTrial, [&](hipEvent_t x, hipEvent_t y) { launch_trial(dv, a, h->stream, x, y); });
                    // record blocks (dE, threshold with the reverse-move log-ratio), then the pairs
                    timed(h, kKCorr, [&](hipEvent_t x, hipEvent_t y) { launch_pair_corr(dv, a, h->stream, x, y); });
                    timed(h, kKScan, [&](hipEvent_t x, hipEvent_t y) { launch_scan(dv, a, h->stream, x, y); });
                    // the next step's proposals beside this update, unless a
                    // slot compaction comes between them (the last step of
                    // the chain sets up no trial: its counter is at the end)
                    ready = !compact_due(batch);
                    if (ready) {
                        const McArgs an = step_args(par ^ 1);
                        timed(h, kKUpdate,
                              [&](hipEvent_t x, hipEvent_t y) { launch_apply_next(dv, a, an, true, h->stream, x, y); });
                    } else {
                        timed(h, kKUpdate, [&](hipEvent_t x, hipEvent_t y) { launch_apply(dv, a, h->stream, x, y); });
                    }
                    compact_if_due(batch);
                    par ^= 1;
                }
                HIPX(hipMemcpyAsync(&done, h->step_ctr, sizeof(int), hipMemcpyDeviceToHost, h->stream));
                HIPX(hipStreamSynchronize(h->stream));
                int err = 0;
                HIPX(hipMemcpy(&err, h->err, sizeof(int), hipMemcpyDeviceToHost));
                if (err) break;  // reported by check_err below
            }
        } else {
            int par = 0;
            bool ready = false;  // the batch's set-up ran beside the previous update
            for (size_t b = 0; b < batches.size(); ++b) {
                const int m0 = batches[b].first, k = batches[b].second;
                const McArgs a = alt_args(h, mc_args(h, mb, m0, k, kTrialBatch, beta), par);
                const bool feat = a.prop.rf > 0.0;  // features only (uniform proposals)
                if (!ready) {
                    if (feat)
                        timed(h, kKPropose,
                              [&](hipEvent_t x, hipEvent_t y) { launch_propose_setup(dv, a, h->stream, x, y); });
                    else
                        timed(h, kKSetup, [&](hipEvent_t x, hipEvent_t y) { launch_setup(dv, a, h->stream, x, y); });
                }
                timed(h, kKTrial, [&](hipEvent_t x, hipEvent_t y) { launch_trial(dv, a, h->stream, x, y); });
                // the record blocks run for a batch of one too (k_scan reads its record)
                timed(h, kKCorr, [&](hipEvent_t x, hipEvent_t y) { launch_pair_corr(dv, a, h->stream, x, y); });
                timed(h, kKScan, [&](hipEvent_t x, hipEvent_t y) { launch_scan(dv, a, h->stream, x, y); });
                // the next batch's set-up beside this update, unless a slot
                // compaction comes between them
                ready = b + 1 < batches.size() && !compact_due(k);
                if (ready) {
                    const McArgs an =
                        alt_args(h, mc_args(h, mb, batches[b + 1].first, batches[b + 1].second, kTrialBatch, beta),
                                 par ^ 1);
                    timed(h, kKUpdate,
                          [&](hipEvent_t x, hipEvent_t y) { launch_apply_next(dv, a, an, feat, h->stream, x, y); });
                } else {
                    timed(h, kKUpdate, [&](hipEvent_t x, hipEvent_t y) { launch_apply(dv, a, h->stream, x, y); });
                }
                compact_if_due(k);
                par ^= 1;
            }
        }
    });
    HIPX(hipGetLastError());
    HIPX(hipEventRecord(e1.e, h->stream));
    HIPX(hipEventSynchronize(e1.e));
    float ms = 0;
    HIPX(hipEventElapsedTime(&ms, e0.e, e1.e));
    collect_events(h);
    check_err(h);
    return ms;
}

// current positions (caller's order, N x 3) and charges of a system
void snapshot(Handle* h, std::vector<double>& r, std::vector<double>& q) {
    const int n = h->n_src;
    r.assign(3 * (size_t)n, 0.0);
    q.assign(n, 0.0);
    HIPX(hipStreamSynchronize(h->stream));
    dispatch(h, [&](auto& dv) {
        const int S = dv.total_slots;
        std::vector<double> x(S), y(S), z(S), c(S);
        std::vector<int> orig(S);
        HIPX(hipMemcpy(x.data(), dv.px, sizeof(double) * S, hipMemcpyDeviceToHost));
        HIPX(hipMemcpy(y.data(), dv.py, sizeof(double) * S, hipMemcpyDeviceToHost));
        HIPX(hipMemcpy(z.data(), dv.pz, sizeof(double) * S, hipMemcpyDeviceToHost));
        HIPX(hipMemcpy(c.data(), dv.pq, sizeof(double) * S, hipMemcpyDeviceToHost));
        HIPX(hipMemcpy(orig.data(), dv.porig, sizeof(int) * S, hipMemcpyDeviceToHost));
        for (int s = 0; s < S; ++s)
            if (orig[s] >= 0) {
                r[3 * orig[s]] = x[s];
                r[3 * orig[s] + 1] = y[s];
                r[3 * orig[s] + 2] = z[s];
                q[orig[s]] = c[s];
            }
    });
}

// PROXY set-up at first use: tables, coefficient storage, the box lists
// (every box of level D-2 and every shallower leaf takes its coefficients from
// its particles; the other boxes of levels 2..D-3 from their children) and the
// scratch of the GEMM passes, sized for chunks of about 48 MB
static void proxy_setup(Handle* h) {
    if (h->proxy.p) return;
    const int p = proxy_order(h->params.digits);
    const Octree& tr = h->tree;
    const int D = tr.max_depth;
    const int n = h->pw.n_diff, d = 2 * n + 1, nz = n + 1;
    const ProxyTables tb = proxy_tables(p, n);
    ProxyArgs& pa = h->proxy;
    auto& o = h->owned;
    const size_t p2 = (size_t)p * p, p3 = p2 * p;
    // stacked real matrices of the GEMM passes
    std::vector<double> az8(8 * p2), px(2 * (size_t)d * p), py(4 * (size_t)d * p), pz(4 * (size_t)nz * p);
    for (int l3 = 0; l3 < p; ++l3)
        for (int ci = 0; ci < 8; ++ci)
            for (int c = 0; c < p; ++c) az8[(size_t)l3 * 8 * p + ci * p + c] = tb.A[((ci >> 2) & 1) * p2 + l3 * p + c];
    // stacked planes of a complex d x p matrix P: x (2d x p, rows re then
    // im) and [[re, -im], [im, re]] of its first `rows` rows (2 rows x 2p)
    auto stack_x = [p](const std::vector<double>& P, int dd, std::vector<double>& m) {
        for (int k = 0; k < dd; ++k)
            for (int a = 0; a < p; ++a) {
                m[(size_t)k * p + a] = P[2 * ((size_t)k * p + a)];
                m[(size_t)(dd + k) * p + a] = P[2 * ((size_t)k * p + a) + 1];
            }
    };
    auto complex_block = [p](const std::vector<double>& P, std::vector<double>& m, int rows) {
        for (int k = 0; k < rows; ++k)
            for (int b = 0; b < p; ++b) {
                const double re = P[2 * ((size_t)k * p + b)], im = P[2 * ((size_t)k * p + b) + 1];
                m[(size_t)k * 2 * p + b] = re;
                m[(size_t)k * 2 * p + p + b] = -im;
                m[(size_t)(rows + k) * 2 * p + b] = im;
                m[(size_t)(rows + k) * 2 * p + p + b] = re;
            }
    };
    stack_x(tb.P, d, px);
    complex_block(tb.P, py, d);
    complex_block(tb.P, pz, nz);
    pa.A = dupload(tb.A, o);
    pa.Az8 = dupload(az8, o);
    pa.Px = dupload(px, o);
    pa.Py = dupload(py, o);
    pa.Pz = dupload(pz, o);
    pa.coef = dalloc<double>((size_t)h->n_pw * p3, o);
    std::vector<int> slot_of(tr.n_nodes, -1), direct;
    int slot = 0;
    for (int i = 0; i < tr.n_nodes; ++i) {
        const int l = tr.level[i];
        if (l < 2 || l > D - 2) continue;
        slot_of[i] = slot;
        if (l == D - 2 || tr.is_leaf[i]) direct.push_back(slot);
        ++slot;
    }
    pa.direct = dupload(direct, o);
    pa.n_direct = (int)direct.size();
    // child -> parent chunks, deepest parent level first
    const size_t budget = (size_t)48 << 20;
    const int per_chunk = (int)std::max<size_t>(1, budget / (8 * p3 * sizeof(double)));
    size_t max_ch = 0;
    pa.c2p.clear();
    for (int l = D - 3; l >= 2; --l) {
        std::vector<int> parents;
        for (int i = 0; i < tr.n_nodes; ++i)
            if (tr.level[i] == l && !tr.is_leaf[i]) parents.push_back(i);
        for (size_t c0 = 0; c0 < parents.size(); c0 += per_chunk) {
            const size_t c1 = std::min(parents.size(), c0 + per_chunk);
            std::vector<int> par, ch, ox, oy, seg;
            for (size_t k = c0; k < c1; ++k) {
                const int nd = parents[k];
                par.push_back(slot_of[nd]);
                for (int ci = 0; ci < 8; ++ci) {
                    const int c = tr.child[8 * nd + ci];
                    if (c < 0 || slot_of[c] < 0) {
                        seg.push_back(-1);
                        continue;
                    }
                    seg.push_back((int)ch.size());
                    ch.push_back(slot_of[c]);
                    ox.push_back(ci & 1);
                    oy.push_back((ci >> 1) & 1);
                }
            }
            ProxyChunk pc;
            pc.n_par = (int)par.size();
            pc.n_ch = (int)ch.size();
            pc.par = dupload(par, o);
            pc.ch = dupload(ch, o);
            pc.ox = dupload(ox, o);
            pc.oy = dupload(oy, o);
            pc.seg = dupload(seg, o);
            pa.c2p.push_back(pc);
            max_ch = std::max(max_ch, ch.size());
        }
    }
    pa.n = n, pa.d = d, pa.nz = nz;
    const size_t y1 = 2 * (size_t)d * p2, y2 = 2 * (size_t)d * d * p, y3 = 2 * (size_t)nz * d * d;
    pa.pw_chunk = (int)std::max<size_t>(1, std::min<size_t>(std::max(h->n_pw, 1), budget / (y1 * sizeof(double))));
    size_t n1 = std::max(max_ch * p3, pa.pw_chunk * y1), n2 = std::max(max_ch * p3, pa.pw_chunk * y2),
           n3 = pa.pw_chunk * y3;
    // the root from the level-2 boxes when every particle lies in one (no
    // leaf above level 2): root wavenumbers (k - nw) 2 pi / L over boxes of
    // size L / 4, i.e. hw = pi / 4
    bool root_ok = D - 2 >= 2;
    for (int i = 0; i < tr.n_nodes && root_ok; ++i)
        if (tr.level[i] < 2 && tr.is_leaf[i]) root_ok = false;
    pa.nw = h->pw.n_window;
    pa.n_root = 0;
    if (root_ok) {
        const int nw = pa.nw, dw = 2 * nw + 1, nzw = nw + 1;
        const ProxyTables tr2 = proxy_tables(p, nw, M_PI / 4.0);
        std::vector<double> prx(2 * (size_t)dw * p), pry(4 * (size_t)dw * p), prz(4 * (size_t)nzw * p);
        stack_x(tr2.P, dw, prx);
        complex_block(tr2.P, pry, dw);
        complex_block(tr2.P, prz, nzw);
        pa.Prx = dupload(prx, o);
        pa.Pry = dupload(pry, o);
        pa.Prz = dupload(prz, o);
        std::vector<int> rs;
        for (int i = 0; i < tr.n_nodes; ++i)
            if (tr.level[i] == 2) rs.push_back(slot_of[i]);
        pa.root_slots = dupload(rs, o);
        pa.n_root = (int)rs.size();
        const size_t nr = rs.size();
        n1 = std::max(n1, nr * 2 * dw * p2);
        n2 = std::max(n2, nr * 2 * (size_t)dw * dw * p);
        n3 = std::max(n3, nr * 2 * (size_t)nzw * dw * dw);
    }
    pa.s1 = dalloc<double>(n1, o);
    pa.s2 = dalloc<double>(n2, o);
    pa.s3 = dalloc<double>(n3, o);
    pa.p = p;
}

// Work items of the MFMA build: every expansion box and the root, each as
// runs of at most kMfmaLeavesPerItem of its descendant leaves; boxes with more
// than one run sum through an FP64 buffer
static void mfma_setup(Handle* h) {
    if (h->mfma_ready) return;
    const Octree& tr = h->tree;
    const int D = tr.max_depth;
    std::vector<int> leaf_pos(tr.n_nodes, -1);
    int nlf = 0;
    for (int i = 0; i < tr.n_nodes; ++i)
        if (tr.is_leaf[i]) leaf_pos[i] = nlf++;
    std::vector<PwWork> work;
    std::vector<int> acc_slots;
    auto add_runs = [&](int slot, int node, int lo, int hi) {
        const int runs = (hi - lo + kMfmaLeavesPerItem - 1) / kMfmaLeavesPerItem;
        int acc = -1;
        if (slot >= 0 && runs > 1) {
            acc = (int)acc_slots.size();
            acc_slots.push_back(slot);
        }
        if (runs == 0) work.push_back({slot, node, lo, lo, acc});
        for (int r = 0; r < runs; ++r)
            work.push_back({slot, node, lo + r * kMfmaLeavesPerItem, std::min(hi, lo + (r + 1) * kMfmaLeavesPerItem), acc});
    };
    add_runs(-1, 0, 0, nlf);  // root
    int slot = 0;
    for (int i = 0; i < tr.n_nodes; ++i) {
        if (tr.level[i] < 2 || tr.level[i] > D - 2) continue;
        add_runs(slot++, i, leaf_pos[tr.first_leaf[i]], leaf_pos[tr.last_leaf[i]] + 1);
    }
    auto& o = h->owned;
    h->mfma.work = dupload(work, o);
    h->mfma.n_work = (int)work.size();
    h->mfma.acc_slots = dupload(acc_slots, o);
    h->mfma.n_acc = (int)acc_slots.size();
    h->mfma.acc_box = dalloc<double>(2 * (size_t)h->d32.Md * std::max<size_t>(acc_slots.size(), 1), o);
    h->mfma_ready = true;
}

void form_outgoing(Handle* h, hipStream_t st) {
    if (h->params.init == PROXY) {
        proxy_setup(h);
        dispatch(h, [&](auto& d) { launch_outgoing_proxy(d, h->pn, h->acc_root, h->proxy, st); });
    } else if (h->precision == HPDMK_MIXED && h->pw.n_diff <= 12 && h->pw.n_window <= 12) {
        mfma_setup(h);  // FP32 MFMA contractions with FP64 sums (digits <= 6)
        launch_outgoing_mfma(h->d32, h->leaves, h->mfma, h->acc_root, st);
    } else {
        dispatch(h, [&](auto& d) { launch_outgoing(d, h->pn, h->acc_root, st); });
    }
}

}  // namespace capi
}  // namespace hpdmk_amd

using namespace hpdmk_amd::capi;

extern "C" {

HPDMKParams hpdmk_default_params(void) {
    HPDMKParams p;
    p.n_per_leaf = 200;  // include/hpdmk.h:12-16
    p.digits = 3;
    p.L = 0.0;
    p.prolate_order = 16;
    p.init = PROXY;
    return p;
}

const char* hpdmk_last_error(void) { return g_err.c_str(); }

int hpdmk_tree_create_ex(HPDMKParams params, int precision, int n_src, const double* r_src, const double* charge,
                         hpdmk_tree* out) {
    return guard([&] {
        if (!out) throw HpdmkError(HPDMK_EINVAL, "null argument");
        *out = nullptr;
        *out = create_handle(params, precision, n_src, r_src, charge, nullptr);
    });
}

int hpdmk_tree_create(HPDMKParams params, int n_src, const double* r_src, const double* charge, hpdmk_tree* out) {
    return hpdmk_tree_create_ex(params, HPDMK_FP64, n_src, r_src, charge, out);
}

int hpdmk_tree_destroy(hpdmk_tree t) {
    return guard([&] {
        Handle* h = H(t);
        if (h->borrowed) throw HpdmkError(HPDMK_EINVAL, "this tree belongs to a replica set (hpdmk_replicas_destroy)");
        (void)hipStreamSynchronize(h->stream);
        delete h;
    });
}

int hpdmk_tree_info(hpdmk_tree t, int* info) {
    return guard([&] {
        Handle* h = H(t);
        dispatch(h, [&](auto& d) {
            info[0] = d.n_nodes;
            info[1] = d.max_depth;
            info[2] = d.n_window;
            info[3] = d.n_diff;
            info[4] = d.Mw;
            info[5] = d.Md;
            info[6] = d.n_pw_nodes;
            info[7] = d.max_near;
            info[8] = h->n_leaves;
            info[9] = d.total_slots;
        });
    });
}

int hpdmk_form_outgoing_pw(hpdmk_tree t) {
    return guard([&] {
        Handle* h = H(t);
        form_outgoing(h, h->stream);
        HIPX(hipGetLastError());
        HIPX(hipStreamSynchronize(h->stream));
        h->formed_out = true;
    });
}

int hpdmk_form_incoming_pw(hpdmk_tree t) {
    return guard([&] {
        Handle* h = H(t);
        dispatch(h, [&](auto& d) { launch_incoming(d, h->pn, h->stream); });
        HIPX(hipGetLastError());
        HIPX(hipStreamSynchronize(h->stream));
        h->formed_in = true;
    });
}

int hpdmk_tree_rebuild(hpdmk_tree t, hpdmk_tree* out) {
    return guard([&] {
        Handle* h = H(t);
        if (!out) throw HpdmkError(HPDMK_EINVAL, "null argument");
        *out = nullptr;
        std::vector<double> r, q;
        snapshot(h, r, q);
        std::unique_ptr<Handle> nh(create_handle(h->params, h->precision, h->n_src, r.data(), q.data(), nullptr));
        nh->tiles = h->tiles;
        nh->tiles_set = h->tiles_set;
        nh->tuned_batch = h->tuned_batch;
        nh->bjerrum = h->bjerrum;
        nh->sigma = h->sigma;
        nh->profiling = h->profiling;
        // the hard-sphere overlap search and the proposal features read the
        // target leaf's near list, which reaches one box of the finest level:
        // a deeper rebuilt tree may no longer cover sigma or rf
        const double finest = h->params.L / (double)(1 << (nh->tree.max_depth - 1));
        if (h->sigma > finest || h->prop_rf > finest)
            throw HpdmkError(HPDMK_EINVAL, "the rebuilt tree's finest box (" + std::to_string(finest) +
                                               ") is smaller than the hard-sphere diameter or the proposal's "
                                               "feature radius; rebuild with a larger n_per_leaf");
        nh->prop_learned = h->prop_learned;
        nh->prop_rf = h->prop_rf;
        nh->prop_max_mean = h->prop_max_mean;
        nh->prop_smin = h->prop_smin;
        nh->prop_smax = h->prop_smax;
        if (h->mlp_w) {
            nh->mlp_w = dalloc<float>(kMlpParams, nh->owned);
            HIPX(hipMemcpy(nh->mlp_w, h->mlp_w, sizeof(float) * kMlpParams, hipMemcpyDeviceToDevice));
            nh->mlp_wp = dalloc<unsigned short>(kPackSize, nh->owned);
            HIPX(hipMemcpy(nh->mlp_wp, h->mlp_wp, sizeof(unsigned short) * kPackSize, hipMemcpyDeviceToDevice));
        }
        if (h->formed_out) {
            form_outgoing(nh.get(), nh->stream);
            nh->formed_out = true;
        }
        if (h->formed_in) {
            dispatch(nh.get(), [&](auto& d) { launch_incoming(d, nh->pn, nh->stream); });
            nh->formed_in = true;
        }
        HIPX(hipGetLastError());
        HIPX(hipStreamSynchronize(nh->stream));
        *out = nh.release();
    });
}

int hpdmk_get_expansion(hpdmk_tree t, int which, int node, double* out, int* count) {
    return guard([&] {
        Handle* h = H(t);
        if (!out || which < 0 || which > 2) throw HpdmkError(HPDMK_EINVAL, "bad expansion selector or null output");
        const Octree& tr = h->tree;
        int slot = -1;
        if (which > 0) {
            if (node < 0 || node >= tr.n_nodes || tr.level[node] < 2 || tr.level[node] > tr.max_depth - 2)
                throw HpdmkError(HPDMK_EINVAL, "node carries no expansion");
            slot = 0;
            for (int i = 0; i < node; ++i) slot += tr.level[i] >= 2 && tr.level[i] <= tr.max_depth - 2;
            if (which == 2 && tr.is_leaf[node]) throw HpdmkError(HPDMK_EINVAL, "leaves carry no incoming expansion");
        }
        HIPX(hipStreamSynchronize(h->stream));
        dispatch(h, [&](auto& d) {
            using T = std::decay_t<decltype(d.root[0].x)>;
            const size_t M = which == 0 ? d.Mw : d.Md;
            const Cplx<T>* src = which == 0 ? d.root : (which == 1 ? d.out : d.inc) + (size_t)slot * d.Md;
            std::vector<Cplx<T>> v(M);
            HIPX(hipMemcpy(v.data(), src, M * sizeof(Cplx<T>), hipMemcpyDeviceToHost));
            for (size_t e = 0; e < M; ++e) {
                out[2 * e] = (double)v[e].x;
                out[2 * e + 1] = (double)v[e].y;
            }
            if (count) *count = (int)M;
        });
    });
}

int hpdmk_eval_energy(hpdmk_tree t, double* energy, double* parts) {
    return guard([&] {
        Handle* h = H(t);
        double e[3];
        dispatch(h, [&](auto& d) { launch_energy(d, h->pn, h->leaves, h->n_leaves, h->energy, h->stream); });
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(e, h->energy, sizeof(e), hipMemcpyDeviceToHost, h->stream));
        HIPX(hipStreamSynchronize(h->stream));
        const double self_w = h->Q * (h->precision == HPDMK_FP64 ? h->d64.self_window : h->d32.self_window);
        e[0] -= self_w;  // src/energy.cpp:33-36
        if (parts) std::memcpy(parts, e, sizeof(e));
        if (energy) *energy = e[0] + e[1] + e[2];
    });
}


int hpdmk_eval_shift_energy(hpdmk_tree t, long i, double dx, double dy, double dz, double* dE, double* parts) {
    return guard([&] {
        Handle* h = H(t);
        const double d[3] = {dx, dy, dz};
        upload_moves(h, h->single, 1, &i, d, nullptr);
        dispatch(h, [&](auto& dv) {
            const McArgs a = mc_args(h, h->single, 0, 1, kTrialOnly, 0.0);
            launch_setup(dv, a, h->stream);
            launch_trial(dv, a, h->stream);
            launch_decide(dv, a, h->stream);
        });
        HIPX(hipGetLastError());
        double p[5], e;
        HIPX(hipMemcpyAsync(&e, h->single.dE, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPX(hipMemcpyAsync(p, h->single.parts, sizeof(p), hipMemcpyDeviceToHost, h->stream));
        check_err(h);
        if (dE) *dE = e;
        if (parts) std::memcpy(parts, p, sizeof(p));
    });
}

int hpdmk_update_shift(hpdmk_tree t, long i, double dx, double dy, double dz) {
    return guard([&] {
        Handle* h = H(t);
        const double d[3] = {dx, dy, dz};
        upload_moves(h, h->single, 1, &i, d, nullptr);
        dispatch(h, [&](auto& dv) {
            const McArgs a = mc_args(h, h->single, 0, 1, kTrialForceAccept, 0.0);
            launch_setup(dv, a, h->stream);
            launch_trial(dv, a, h->stream);
            launch_decide(dv, a, h->stream);
            launch_apply(dv, a, h->stream);
            if (++h->since_compact >= kCompactEvery) {
                launch_compact(dv, h->leaves, h->n_leaves, h->stream);
                h->since_compact = 0;
            }
        });
        HIPX(hipGetLastError());
        check_err(h);
    });
}

int hpdmk_get_positions(hpdmk_tree t, double* r) {
    return guard([&] {
        Handle* h = H(t);
        dispatch(h, [&](auto& d) {
            const int S = d.total_slots;
            std::vector<double> x(S), y(S), z(S);
            std::vector<int> orig(S);
            HIPX(hipMemcpy(x.data(), d.px, sizeof(double) * S, hipMemcpyDeviceToHost));
            HIPX(hipMemcpy(y.data(), d.py, sizeof(double) * S, hipMemcpyDeviceToHost));
            HIPX(hipMemcpy(z.data(), d.pz, sizeof(double) * S, hipMemcpyDeviceToHost));
            HIPX(hipMemcpy(orig.data(), d.porig, sizeof(int) * S, hipMemcpyDeviceToHost));
            for (int s = 0; s < S; ++s)
                if (orig[s] >= 0) {
                    r[3 * orig[s]] = x[s];
                    r[3 * orig[s] + 1] = y[s];
                    r[3 * orig[s] + 2] = z[s];
                }
        });
    });
}

int hpdmk_eval_shift_energy_batch(hpdmk_tree t, int n, const long* idx, const double* d, double* dE) {
    return guard([&] {
        Handle* h = H(t);
        if (n < 0) throw HpdmkError(HPDMK_EINVAL, "n must be non-negative");
        if (n == 0) return;
        upload_moves(h, h->single, n, idx, d, nullptr);
        for (int m0 = 0; m0 < n; m0 += kMaxBatch) {
            const int k = std::min(kMaxBatch, n - m0);
            dispatch(h, [&](auto& dv) {
                const McArgs a = mc_args(h, h->single, m0, k, kTrialOnly, 0.0);
                launch_setup(dv, a, h->stream);
                launch_trial(dv, a, h->stream);
                launch_decide(dv, a, h->stream);
            });
        }
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(dE, h->single.dE, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
        check_err(h);
    });
}

int hpdmk_mc_load(hpdmk_tree t, int n, const long* idx, const double* d, const double* u) {
    return guard([&] {
        Handle* h = H(t);
        if (n <= 0 || !idx || !d || !u) throw HpdmkError(HPDMK_EINVAL, "empty or null move list");
        upload_moves(h, h->chain, n, idx, d, u);
    });
}

int hpdmk_mc_run(hpdmk_tree t, double beta, int mode, int batch, double* gpu_ms) {
    return guard([&] {
        const double ms = run_chain(H(t), beta, mode, batch);
        if (gpu_ms) *gpu_ms = ms;
    });
}

int hpdmk_mc_results(hpdmk_tree t, double* dE, int* accept) {
    return guard([&] {
        Handle* h = H(t);
        const MoveBuf& mb = h->chain;
        if (dE) HIPX(hipMemcpyAsync(dE, mb.dE, sizeof(double) * mb.n, hipMemcpyDeviceToHost, h->stream));
        if (accept) HIPX(hipMemcpyAsync(accept, mb.acc, sizeof(int) * mb.n, hipMemcpyDeviceToHost, h->stream));
        HIPX(hipStreamSynchronize(h->stream));
    });
}

int hpdmk_mc_stats(hpdmk_tree t, int* stats) {
    return guard([&] {
        Handle* h = H(t);
        const MoveBuf& mb = h->chain;
        HIPX(hipMemcpyAsync(stats, mb.stats, sizeof(int) * kStatsPerMove * mb.n, hipMemcpyDeviceToHost, h->stream));
        HIPX(hipStreamSynchronize(h->stream));
    });
}

int hpdmk_set_profiling(hpdmk_tree t, int level) {
    return guard([&] {
        if (level < 0 || level > 2) throw HpdmkError(HPDMK_EINVAL, "profiling level must be 0, 1 or 2");
        H(t)->profiling = level;
    });
}

int hpdmk_kernel_times(hpdmk_tree t, double* ms, int* count) {
    return guard([&] {
        Handle* h = H(t);
        for (int k = 0; k < kKinds; ++k) {
            if (ms) ms[k] = h->kms[k];
            if (count) count[k] = h->kcount[k];
        }
    });
}

int hpdmk_read_particle_info(const char* path, int capacity, double* r, double* q, int* count) {
    return guard([&] {
        if (!path || !count || capacity < 0 || (capacity > 0 && (!r || !q)))
            throw HpdmkError(HPDMK_EINVAL, "null argument or negative capacity");
        std::ifstream in(path);
        if (!in) throw HpdmkError(HPDMK_EINVAL, std::string("cannot open ") + path);
        std::string line;
        std::getline(in, line);  // header
        int n = 0;
        while (std::getline(in, line)) {
            std::istringstream ss(line);
            double x, y, z, c;
            if (!(ss >> x >> y >> z >> c)) continue;
            if (n < capacity) {
                r[3 * n] = x;
                r[3 * n + 1] = y;
                r[3 * n + 2] = z;
                q[n] = c;
            }
            ++n;
        }
        *count = n;
    });
}

namespace {
// the per-node summary of hpdmk_plan_tree / hpdmk_plan_tree_device
void write_plan(const Octree& tr, int* n_nodes, int* nodes, int* particle_leaf) {
        *n_nodes = tr.n_nodes;
        if (nodes) {
            std::vector<int> sub(tr.n_nodes, 0);
            for (int i = tr.n_nodes - 1; i >= 0; --i) {
                sub[i] = tr.is_leaf[i] ? tr.leaf_count[i] : 0;
                if (!tr.is_leaf[i])
                    for (int c = 0; c < 8; ++c) sub[i] += sub[tr.child[8 * i + c]];
            }
            for (int i = 0; i < tr.n_nodes; ++i) {
                int* o = nodes + 9 * i;
                o[0] = tr.level[i];
                o[1] = tr.is_leaf[i];
                o[2] = tr.parent[i];
                o[3] = tr.cell[3 * i];
                o[4] = tr.cell[3 * i + 1];
                o[5] = tr.cell[3 * i + 2];
                o[6] = sub[i];
                int nc = 0;
                if (tr.level[i] >= 2)
                    for (int k = 0; k < 26; ++k) nc += tr.colleague[26 * i + k] >= 0;
                o[7] = nc;
                o[8] = tr.near_off[i + 1] - tr.near_off[i];
            }
        }
        if (particle_leaf) {
            int pos = 0;
            for (int i = 0; i < tr.n_nodes; ++i)
                if (tr.is_leaf[i])
                    for (int k = 0; k < tr.leaf_count[i]; ++k) particle_leaf[tr.order[pos++]] = i;
        }
}
}  // namespace

int hpdmk_plan_tree(HPDMKParams params, int n_src, const double* r_src, int* n_nodes, int* nodes,
                    int* particle_leaf) {
    return guard([&] {
        if (!r_src || !n_nodes) throw HpdmkError(HPDMK_EINVAL, "null argument");
        if (params.n_per_leaf < 1) throw HpdmkError(HPDMK_EINVAL, "n_per_leaf must be positive");
        write_plan(build_octree(n_src, r_src, params.L, params.n_per_leaf), n_nodes, nodes, particle_leaf);
    });
}

int hpdmk_plan_tree_device(HPDMKParams params, int n_src, const double* r_src, int* n_nodes, int* nodes,
                           int* particle_leaf, double* ms) {
    return guard([&] {
        if (!r_src || !n_nodes) throw HpdmkError(HPDMK_EINVAL, "null argument");
        if (params.n_per_leaf < 1) throw HpdmkError(HPDMK_EINVAL, "n_per_leaf must be positive");
        struct S {
            hipStream_t s = nullptr;
            ~S() {
                if (s) (void)hipStreamDestroy(s);
            }
        } st;
        HIPX(hipStreamCreate(&st.s));
        const auto t0 = std::chrono::steady_clock::now();
        const Octree tr = build_octree_device(n_src, r_src, params.L, params.n_per_leaf, st.s);
        if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (tr.n_nodes == 0) throw HpdmkError(HPDMK_ENOTIMPL, "tree deeper than the device octree build supports");
        write_plan(tr, n_nodes, nodes, particle_leaf);
    });
}

int hpdmk_pswf_constants(int digits, double* out) {
    return guard([&] {
        PswfConsts p;
        try {
            p = pswf_consts(digits);
        } catch (const HpdmkError& e) {
            throw HpdmkError(HPDMK_EINVAL, e.what());
        }
        out[0] = p.c;
        out[1] = p.lambda;
        out[2] = p.C0;
        out[3] = p.diff0;
        out[4] = p.psi0_0;
        out[5] = p.n_window;
        out[6] = p.n_diff;
    });
}

int hpdmk_phase_times(hpdmk_tree t, double* us, int* count) {  // us[kProbes], count[kProbes]
    return guard([&] {
        Handle* h = H(t);
        unsigned long long p[2 * kProbes];
        HIPX(hipMemcpy(p, h->probe, sizeof(p), hipMemcpyDeviceToHost));
        int freq_khz = 100000;
        (void)hipDeviceGetAttribute(&freq_khz, hipDeviceAttributeWallClockRate, 0);
        for (int k = 0; k < kProbes; ++k) {
            if (us) us[k] = p[kProbes + k] ? (double)p[k] / p[kProbes + k] / (freq_khz * 1e-3) : 0.0;
            if (count) count[k] = (int)p[kProbes + k];
        }
    });
}

int hpdmk_mc_set_model(hpdmk_tree t, double bjerrum, double sigma) {
    return guard([&] {
        Handle* h = H(t);
        if (!(bjerrum > 0.0) || !std::isfinite(bjerrum)) throw HpdmkError(HPDMK_EINVAL, "Bjerrum length must be > 0");
        if (!(sigma >= 0.0) || !std::isfinite(sigma)) throw HpdmkError(HPDMK_EINVAL, "sigma must be >= 0");
        // the overlap search reads the target leaf's near list, which covers at
        // least one box of the leaf's level in every direction
        const double finest = h->params.L / (double)(1 << (h->tree.max_depth - 1));
        if (sigma > finest)
            throw HpdmkError(HPDMK_EINVAL, "sigma exceeds the finest box size " + std::to_string(finest));
        h->bjerrum = bjerrum;
        h->sigma = sigma;
    });
}

int hpdmk_mc_set_proposal(hpdmk_tree t, int learned, const float* weights, double rf, double max_mean,
                          double sigma_min, double sigma_max) {
    return guard([&] {
        Handle* h = H(t);
        if (learned != 0 && learned != 1) throw HpdmkError(HPDMK_EINVAL, "learned must be 0 or 1");
        if (!(rf >= 0.0) || !std::isfinite(rf)) throw HpdmkError(HPDMK_EINVAL, "feature radius must be >= 0");
        const double finest = h->params.L / (double)(1 << (h->tree.max_depth - 1));
        if (rf > finest)
            throw HpdmkError(HPDMK_EINVAL, "feature radius exceeds the finest box size " + std::to_string(finest));
        if (learned) {
            if (!weights) throw HpdmkError(HPDMK_EINVAL, "learned proposal needs MLP weights");
            if (!(rf > 0.0)) throw HpdmkError(HPDMK_EINVAL, "learned proposal needs a feature radius > 0");
            if (!(sigma_min > 0.0) || !(sigma_max > sigma_min) || !(max_mean >= 0.0))
                throw HpdmkError(HPDMK_EINVAL, "need 0 < sigma_min < sigma_max and max_mean >= 0");
            if (sigma_max + max_mean > 0.25 * h->params.L)
                throw HpdmkError(HPDMK_EINVAL, "proposal scale too large for the box (sigma_max + max_mean > L/4)");
            if (!h->mlp_w) h->mlp_w = dalloc<float>(kMlpParams, h->owned);
            if (!h->mlp_wp) h->mlp_wp = dalloc<unsigned short>(kPackSize, h->owned);
            HIPX(hipMemcpy(h->mlp_w, weights, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
            std::vector<unsigned short> packed(kPackSize);
            mlp_pack_weights(weights, packed.data());
            HIPX(hipMemcpy(h->mlp_wp, packed.data(), sizeof(unsigned short) * kPackSize, hipMemcpyHostToDevice));
            h->prop_max_mean = max_mean;
            h->prop_smin = sigma_min;
            h->prop_smax = sigma_max;
        }
        h->prop_learned = learned;
        h->prop_rf = rf;
    });
}

int hpdmk_mc_proposal_results(hpdmk_tree t, double* delta, double* feat, double* logratio) {
    return guard([&] {
        Handle* h = H(t);
        const MoveBuf& mb = h->chain;
        const size_t n = mb.n;
        if (delta) {
            if (h->prop_learned)
                HIPX(hipMemcpyAsync(delta, mb.dprop, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, h->stream));
            else
                HIPX(hipMemcpyAsync(delta, mb.d, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, h->stream));
        }
        if (feat) HIPX(hipMemcpyAsync(feat, mb.feat, sizeof(double) * kFeat * n, hipMemcpyDeviceToHost, h->stream));
        if (logratio) HIPX(hipMemcpyAsync(logratio, mb.H, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
        HIPX(hipStreamSynchronize(h->stream));
    });
}

int hpdmk_mc_target_features(hpdmk_tree t, double* feat_t) {
    return guard([&] {
        Handle* h = H(t);
        const MoveBuf& mb = h->chain;
        if (!feat_t) throw HpdmkError(HPDMK_EINVAL, "null output");
        if (!h->prop_learned) throw HpdmkError(HPDMK_EINVAL, "target features are recorded by the learned proposal only");
        HIPX(hipMemcpyAsync(feat_t, mb.feat_t, sizeof(double) * kFeat * mb.n, hipMemcpyDeviceToHost, h->stream));
        HIPX(hipStreamSynchronize(h->stream));
    });
}

int hpdmk_set_tiles(hpdmk_tree t, const int* tiles, int batch) {
    return guard([&] {
        Handle* h = H(t);
        if (!tiles) throw HpdmkError(HPDMK_EINVAL, "null tiles");
        for (int k = 0; k < 3; ++k)
            if (tiles[k] < 1 || tiles[k] > kMaxRoleBlocks)
                throw HpdmkError(HPDMK_EINVAL, "blocks per trial role must be in [1, 32]");
        if (tiles[3] < 1 || tiles[3] > 65535) throw HpdmkError(HPDMK_EINVAL, "update blocks must be in [1, 65535]");
        if (batch < 1 || batch > kMaxBatch) throw HpdmkError(HPDMK_EINVAL, "batch must be in [1, 256]");
        h->tiles_set = true;
        h->tiles.win = tiles[0];
        h->tiles.diff = tiles[1];
        h->tiles.near = tiles[2];
        h->tiles.apply = tiles[3];
        h->tuned_batch = batch;
    });
}

int hpdmk_get_tiles(hpdmk_tree t, int* out) {
    return guard([&] {
        Handle* h = H(t);
        out[0] = h->tiles.win;
        out[1] = h->tiles.diff;
        out[2] = h->tiles.near;
        out[3] = h->tiles.apply;
        out[4] = h->tuned_batch;
    });
}

int hpdmk_autotune(hpdmk_tree t, int n_moves, const long* idx, const double* d, const double* u, double beta,
                   int* out, double* best_ms) {
    return guard([&] {
        check_beta(beta);
        Handle* h = H(t);
        if (n_moves < 1 || !idx || !d || !u) throw HpdmkError(HPDMK_EINVAL, "empty or null move list");
        // a twin of the current state takes the timing runs; this system is untouched
        const int n = h->n_src;
        std::vector<double> r, q;
        snapshot(h, r, q);
        std::unique_ptr<Handle> tw(create_handle(h->params, h->precision, n, r.data(), q.data(), nullptr));
        tw->bjerrum = h->bjerrum;
        tw->sigma = h->sigma;
        dispatch(tw.get(), [&](auto& dv) {
            launch_outgoing(dv, tw->pn, tw->acc_root, tw->stream);
            launch_incoming(dv, tw->pn, tw->stream);
        });
        // Every timing run moves other particles (the sample's indices shifted by
        // a multiple of a large odd step, displacements and uniforms kept):
        // rerunning one move list would find its data in cache.
        std::vector<long> shifted(n_moves);
        long round = 0;
        auto load_fresh = [&] {
            const long step = 1000003L % n;
            for (int k = 0; k < n_moves; ++k) shifted[k] = (idx[k] + (round * step) % n) % n;
            ++round;
            upload_moves(tw.get(), tw->chain, n_moves, shifted.data(), d, u);
        };
        // coordinate search from the current shape: batch, then each role in turn
        Tiles best = h->tiles;
        int best_batch = h->tuned_batch;
        auto timeit = [&](const Tiles& g, int batch) {
            tw->tiles = g;
            load_fresh();
            run_chain(tw.get(), beta, 1, batch);  // first run: code objects, queues
            load_fresh();
            return run_chain(tw.get(), beta, 1, batch);
        };
        double best_t = timeit(best, best_batch);
        auto consider = [&](const Tiles& g, int batch) {
            const double tm = timeit(g, batch);
            if (tm < best_t) {
                best_t = tm;
                best = g;
                best_batch = batch;
            }
        };
        for (int b : {16, 32, 48, 64, 96, 128, 192, 256}) consider(best, b);
        for (int v : {4, 8, 12, 16}) {
            Tiles g = best;
            g.near = v;
            consider(g, best_batch);
        }
        for (int v : {2, 4, 8, 16}) {
            Tiles g = best;
            g.diff = v;
            consider(g, best_batch);
        }
        for (int v : {1, 2, 4}) {
            Tiles g = best;
            g.win = v;
            consider(g, best_batch);
        }
        for (int v : {256, 512, 1024, 2048}) {
            Tiles g = best;
            g.apply = v;
            consider(g, best_batch);
        }
        h->tiles = best;
        h->tiles_set = true;
        h->tuned_batch = best_batch;
        if (out) {
            out[0] = best.win;
            out[1] = best.diff;
            out[2] = best.near;
            out[3] = best.apply;
            out[4] = best_batch;
        }
        if (best_ms) *best_ms = best_t;
    });
}

int hpdmk_copy_probe(long bytes, int reps, double* gbs) {
    return guard([&] {
        if (bytes < 16 || reps < 1 || !gbs) throw HpdmkError(HPDMK_EINVAL, "copy probe needs bytes >= 16, reps >= 1");
        const size_t n16 = (size_t)bytes / 16;
        void *a = nullptr, *b = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto release = [&] {
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        };
        try {
            HIPX(hipMalloc(&a, n16 * 16));
            HIPX(hipMalloc(&b, n16 * 16));
            HIPX(hipEventCreate(&e0));
            HIPX(hipEventCreate(&e1));
            launch_copy16(a, b, n16, nullptr);  // warm-up (page mapping)
            HIPX(hipEventRecord(e0, nullptr));
            for (int r = 0; r < reps; ++r) launch_copy16(a, b, n16, nullptr);
            HIPX(hipEventRecord(e1, nullptr));
            HIPX(hipEventSynchronize(e1));
            float ms = 0;
            HIPX(hipEventElapsedTime(&ms, e0, e1));
            *gbs = 2.0 * 16.0 * (double)n16 * reps / (ms * 1e-3) / 1e9;
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}

int hpdmk_clear_error(hpdmk_tree t) {
    return guard([&] {
        Handle* h = H(t);
        HIPX(hipMemsetAsync(h->err, 0, sizeof(int), h->stream));
        HIPX(hipStreamSynchronize(h->stream));
    });
}

}  // extern "C"
