// Device state of one DMK system behind an hpdmk_tree handle, and the host
// helpers shared by the C ABI files (hpdmk_capi.cpp: one system;
// hpdmk_replicas.cpp: a set of independent systems).  Internal header.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hpdmk.h"
#include "dmk_common.hpp"
#include "dmk_host.hpp"
#include "dmk_kernels.hpp"
#include "dmk_nudft_mfma.hpp"
#include "dmk_proxy.hpp"

namespace hpdmk_amd {
namespace capi {

// message of the last failure on the calling thread (hpdmk_last_error)
extern thread_local std::string g_err;

#define HIPX(call)                                                                                         \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            throw HpdmkError(HPDMK_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

template <typename V>
V* dalloc(size_t n, std::vector<void*>& owned) {
    void* p = nullptr;
    HIPX(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(V)));
    owned.push_back(p);
    return static_cast<V*>(p);
}

template <typename V>
V* dupload(const std::vector<V>& h, std::vector<void*>& owned) {
    V* p = dalloc<V>(h.size(), owned);
    if (!h.empty()) HIPX(hipMemcpy(p, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice));
    return p;
}

struct MoveBuf {
    int n = 0, cap = 0;
    std::vector<long> h_idx;
    int* idx = nullptr;
    double* d = nullptr;
    double* u = nullptr;
    double* dE = nullptr;
    double* dE_stale = nullptr;
    double* parts = nullptr;
    int* acc = nullptr;
    int* stats = nullptr;
    // proposal: displacement used, log q, features at the origin and at the
    // target (learned), log-ratio
    double* dprop = nullptr;
    double* logq = nullptr;
    double* feat = nullptr;
    double* feat_t = nullptr;
    double* H = nullptr;
};

// Moved particles leave holes in their old leaf slots; the slots are packed
// again after at most this many moves.  Leaf capacity is twice the initial
// count (at least count + 64), so a leaf would need 64 net arrivals in between
// to run out.
constexpr int kCompactEvery = 2048;
// moves per captured graph of the plain chain
constexpr int kPlainGraphMoves = 64;
static_assert(kPlainGraphMoves % 2 == 0, "the fused chain's ring slot is the graph-local move's parity");
// the fused single-move chain packs box cells in 9 bits (levels <= 9, i.e.
// trees of at most 11 levels); deeper trees run the four-kernel chain
constexpr int kFusedMaxDepth = 11;
// launch shape of the fused single-move chain unless the caller chose one:
// more near-field blocks than a batch wants (one trial fills few CUs)
constexpr Tiles kPlainTiles{4, 8, 32, 1024};

enum KernelKind { kKTrial = 0, kKUpdate = 1, kKCorr = 2, kKScan = 3, kKSetup = 4, kKPropose = 5, kKinds = 6 };

struct Handle {
    HPDMKParams params{};
    int precision = HPDMK_FP64;
    PswfConsts pw;
    Octree tree;
    int n_src = 0, n_leaves = 0, n_pw = 0;
    double Q = 0;
    hipStream_t stream = nullptr;
    int device = 0;          // HIP device current at create (the handle's memory lives there)
    bool own_stream = true;  // false when the stream belongs to a replica set
    bool borrowed = false;   // member of a replica set: destroyed with the set
    std::vector<void*> owned;
    DevTree<double> d64{};
    DevTree<float> d32{};
    PwNodes pn{};
    int* leaves = nullptr;
    double* acc_root = nullptr;
    double* energy = nullptr;
    int* err = nullptr;
    // PROXY initialisation (allocated at its first use): Chebyshev tables,
    // coefficients per expansion box, box lists
    ProxyArgs proxy{};
    // which expansions the caller formed (a rebuild forms the same)
    bool formed_out = false, formed_in = false;
    // mixed-precision DIRECT build on MFMA (allocated at its first use)
    PwMfmaPlan mfma{};
    bool mfma_ready = false;
    // Monte Carlo scratch (kMaxBatch trials)
    McScratch sc{};
    // the other trial-state buffers (TrialInfo, phase rows) of batched
    // chains, whose consecutive steps alternate between the two
    TrialInfo* info_alt = nullptr;
    void* tabs_alt = nullptr;
    unsigned long long* probe = nullptr;
    // move lists in HBM: `single` serves eval/update calls, `chain` the staged Metropolis run
    MoveBuf single, chain;
    int since_compact = 0;  // leaf changes (upper bound) since the last slot compaction
    // launch shape of the trial and update kernels, and the batch size used
    // when hpdmk_mc_run is given batch 0 (hpdmk_set_tiles, hpdmk_autotune)
    Tiles tiles{};
    bool tiles_set = false;  // set by hpdmk_set_tiles / hpdmk_autotune (else the fused chain uses kPlainTiles)
    int tuned_batch = 64;
    // proposal of hpdmk_mc_run (hpdmk_mc_set_proposal)
    int prop_learned = 0;
    int* step_ctr = nullptr;  // device move counter of batched learned runs
    float* mlp_w = nullptr;  // kMlpParams on the device
    unsigned short* mlp_wp = nullptr;  // kPackSize BF16 weights in matrix-core operand order
    double prop_rf = 0.0, prop_max_mean = 0.5, prop_smin = 0.05, prop_smax = 1.0;
    // Monte Carlo model of hpdmk_mc_run (hpdmk_mc_set_model)
    double bjerrum = 1.0;
    double sigma = 0.0;
    // profiling
    int profiling = 0;  // 0 off, 1 kernel events + per-move counts, 2 + phase timers
    double kms[kKinds] = {};
    int kcount[kKinds] = {};
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;

    ~Handle() {
        for (auto& e : ev_pool) (void)hipEventDestroy(e);
        for (void* p : owned) (void)hipFree(p);
        if (stream && own_stream) (void)hipStreamDestroy(stream);
    }
};

// outgoing expansions of a handle's system by its params.init (DIRECT or
// PROXY), enqueued on st (src/planewave.cpp:86-97)
void form_outgoing(Handle* h, hipStream_t st);
// current positions (caller's order) and charges, copied to the host
void snapshot(Handle* h, std::vector<double>& r, std::vector<double>& q);

template <typename F>
int guard(F&& f) {
    try {
        f();
        return HPDMK_OK;
    } catch (const HpdmkError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return HPDMK_ERUNTIME;
    }
}

inline Handle* H(hpdmk_tree t) {
    if (!t) throw HpdmkError(HPDMK_EINVAL, "null hpdmk_tree handle");
    return static_cast<Handle*>(t);
}

template <typename T>
DevTree<T>& dev(Handle* h);
template <>
inline DevTree<double>& dev<double>(Handle* h) {
    return h->d64;
}
template <>
inline DevTree<float>& dev<float>(Handle* h) {
    return h->d32;
}

template <typename F>
void dispatch(Handle* h, F&& f) {
    if (h->precision == HPDMK_FP64)
        f(h->d64);
    else
        f(h->d32);
}

hipEvent_t take_event(Handle* h);
void collect_events(Handle* h);
// throws HPDMK_ECAPACITY when a kernel flagged a leaf overflow
void check_err(Handle* h);

template <typename L>
void timed(Handle* h, int kind, L&& launch) {
    if (!h->profiling) {
        launch(nullptr, nullptr);
        return;
    }
    // the events are attached to the dispatch itself, so they bracket the
    // kernel's execution only (not the queueing around it)
    hipEvent_t a = take_event(h), b = take_event(h);
    launch(a, b);
    h->ev_used.push_back({kind, {a, b}});
}

// Builds the octree, kernel tables and device state of one system.  The
// handle uses `stream` when given (not owned), else creates its own.
Handle* create_handle(HPDMKParams params, int precision, int n_src, const double* r_src, const double* charge,
                      hipStream_t stream);
void ensure_moves(Handle* h, MoveBuf& b, int n);
// throws HPDMK_EINVAL unless beta is finite and > 0
void check_beta(double beta);
void upload_moves(Handle* h, MoveBuf& b, int n, const long* idx, const double* d, const double* u);
McArgs mc_args(Handle* h, const MoveBuf& b, int move0, int K, int mode, double beta);
double run_chain(Handle* h, double beta, int mode, int batch);

}  // namespace capi
}  // namespace hpdmk_amd
