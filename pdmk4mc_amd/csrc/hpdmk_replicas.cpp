// C ABI of a replica set: many independent DMK systems of different sizes
// (ragged) advanced together, one Metropolis move of every replica per step.
// Every replica is a full hpdmk_tree (its own octree, expansions and scratch);
// a step is one launch each of the set-up, trial and update kernels over all
// replicas (replica = blockIdx.z), and runs of steps are captured once into a
// HIP graph and replayed.  No reference counterpart: the reference advances
// one system on one thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "hpdmk_handle.hpp"

using namespace hpdmk_amd;
using namespace hpdmk_amd::capi;

namespace {

// steps between slot compactions of every replica (a leaf has room for at
// least 64 arrivals beyond its initial count)
constexpr int kReplicaCompactEvery = 512;

// blocks per trial role of a replica step (window, difference, near field):
// with hundreds of replicas the trial kernel has blocks enough, and fewer,
// longer role blocks per replica run the step faster (256 ragged replicas at
// digits 6: 168.5 us per step at the single-system default {2, 8, 8},
// 151-156 us over {1-2, 1-4, 2-4})
constexpr Tiles kReplicaTiles{1, 2, 4, 1024};

struct ReplicaSet {
    int precision = HPDMK_FP64;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<Handle>> reps;
    std::vector<void*> owned;
    void* trees = nullptr;  // DevTree<T>[n]
    McArgs* args = nullptr;
    McArgs* args_alt = nullptr;  // the same with each replica's other trial-state buffers
    int* move_ctr = nullptr;
    const int** leaves = nullptr;
    int* n_leaves = nullptr;
    int max_depth = 0, max_leaves = 0, apply_blocks = 0;
    int n_moves = 0;
    // the instantiated whole-run graph and its step count: its kernels take
    // only set-level pointers (the moves and the chain's arguments are read
    // from device memory at run time), so a later run of as many steps
    // launches it again
    hipGraphExec_t run_exec = nullptr;
    int run_steps = 0;

    ~ReplicaSet() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (run_exec) (void)hipGraphExecDestroy(run_exec);
        reps.clear();
        for (void* p : owned) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int n() const { return (int)reps.size(); }
};

ReplicaSet* RS(hpdmk_replicas s) {
    if (!s) throw HpdmkError(HPDMK_EINVAL, "null hpdmk_replicas handle");
    return static_cast<ReplicaSet*>(s);
}

template <typename T>
ReplicaLaunch<T> launch_args(ReplicaSet* s) {
    ReplicaLaunch<T> r{};
    r.trees = static_cast<const DevTree<T>*>(s->trees);
    r.args = s->args;
    r.args_alt = s->args_alt;
    r.move_ctr = s->move_ctr;
    r.leaves = s->leaves;
    r.n_leaves = s->n_leaves;
    r.n = s->n();
    r.max_depth = s->max_depth;
    r.max_leaves = s->max_leaves;
    r.apply_blocks = s->apply_blocks;
    r.dmax = 0;
    for (auto& h : s->reps) {
        const DevTree<T>& d = dev<T>(h.get());
        r.dmax = std::max(r.dmax, std::max(d.dw, d.dd));
    }
    r.tiles = kReplicaTiles;
    return r;
}

template <typename T>
void upload_trees(ReplicaSet* s) {
    std::vector<DevTree<T>> v;
    for (auto& h : s->reps) v.push_back(dev<T>(h.get()));
    HIPX(hipMemcpy(s->trees, v.data(), sizeof(DevTree<T>) * v.size(), hipMemcpyHostToDevice));
}

}  // namespace

extern "C" {

int hpdmk_replicas_create(int n_rep, const HPDMKParams* params, int precision, const int* n_src, const double* r_src,
                          const double* charge, hpdmk_replicas* out) {
    return guard([&] {
        if (!out || !params || !n_src || !r_src || !charge || n_rep < 1)
            throw HpdmkError(HPDMK_EINVAL, "null argument or empty replica set");
        *out = nullptr;
        if (n_rep > 65535) throw HpdmkError(HPDMK_EINVAL, "at most 65535 replicas");
        auto s = std::make_unique<ReplicaSet>();
        s->precision = precision;
        HIPX(hipStreamCreate(&s->stream));
        size_t off = 0;
        std::vector<const int*> lv;
        std::vector<int> nl;
        for (int r = 0; r < n_rep; ++r) {
            if (n_src[r] < 1) throw HpdmkError(HPDMK_EINVAL, "replica " + std::to_string(r) + " is empty");
            Handle* h = nullptr;
            try {
                h = create_handle(params[r], precision, n_src[r], r_src + 3 * off, charge + off, s->stream);
            } catch (const HpdmkError& e) {
                throw HpdmkError(e.code, "replica " + std::to_string(r) + ": " + e.what());
            }
            h->borrowed = true;
            s->reps.emplace_back(h);
            off += n_src[r];
            s->max_depth = std::max(s->max_depth, h->tree.max_depth);
            s->max_leaves = std::max(s->max_leaves, h->n_leaves);
            lv.push_back(h->leaves);
            nl.push_back(h->n_leaves);
        }
        // update blocks per replica (k_apply1_rep, 2048-element work items):
        // the root alone without difference levels, else the root and the
        // rewritten level arrays of the move (32 measured best at digits 6)
        s->apply_blocks = s->max_depth >= 4 ? 32 : 4;
        const size_t tsz = precision == HPDMK_FP64 ? sizeof(DevTree<double>) : sizeof(DevTree<float>);
        s->trees = dalloc<char>(tsz * n_rep, s->owned);
        s->args = dalloc<McArgs>(n_rep, s->owned);
        s->args_alt = dalloc<McArgs>(n_rep, s->owned);
        s->move_ctr = dalloc<int>(n_rep, s->owned);
        s->leaves = (const int**)dalloc<const int*>(n_rep, s->owned);
        s->n_leaves = dalloc<int>(n_rep, s->owned);
        HIPX(hipMemcpy((void*)s->leaves, lv.data(), sizeof(const int*) * n_rep, hipMemcpyHostToDevice));
        HIPX(hipMemcpy(s->n_leaves, nl.data(), sizeof(int) * n_rep, hipMemcpyHostToDevice));
        if (precision == HPDMK_FP64)
            upload_trees<double>(s.get());
        else
            upload_trees<float>(s.get());
        *out = s.release();
    });
}

int hpdmk_replicas_destroy(hpdmk_replicas s) {
    return guard([&] { delete RS(s); });
}

int hpdmk_replicas_count(hpdmk_replicas s, int* n) {
    return guard([&] { *n = RS(s)->n(); });
}

int hpdmk_replicas_tree(hpdmk_replicas s, int r, hpdmk_tree* tree) {
    return guard([&] {
        ReplicaSet* set = RS(s);
        if (r < 0 || r >= set->n()) throw HpdmkError(HPDMK_EINVAL, "replica index out of range");
        *tree = set->reps[r].get();
    });
}

int hpdmk_replicas_form_pw(hpdmk_replicas s) {
    return guard([&] {
        ReplicaSet* set = RS(s);
        for (auto& h : set->reps) {
            form_outgoing(h.get(), set->stream);
            dispatch(h.get(), [&](auto& d) { launch_incoming(d, h->pn, set->stream); });
        }
        HIPX(hipGetLastError());
        HIPX(hipStreamSynchronize(set->stream));
    });
}

int hpdmk_replicas_mc_load(hpdmk_replicas s, int n_moves, const long* idx, const double* d, const double* u) {
    return guard([&] {
        ReplicaSet* set = RS(s);
        if (n_moves <= 0 || !idx || !d || !u) throw HpdmkError(HPDMK_EINVAL, "empty or null move list");
        for (int r = 0; r < set->n(); ++r) {
            const size_t o = (size_t)r * n_moves;
            upload_moves(set->reps[r].get(), set->reps[r]->chain, n_moves, idx + o, d + 3 * o, u + o);
        }
        set->n_moves = n_moves;
    });
}

int hpdmk_replicas_mc_run(hpdmk_replicas s, double beta, int graph_steps, double* gpu_ms) {
    return guard([&] {
        ReplicaSet* set = RS(s);
        const int n = set->n_moves;
        if (n <= 0) throw HpdmkError(HPDMK_EINVAL, "no moves loaded (hpdmk_replicas_mc_load)");
        if (graph_steps < 0) throw HpdmkError(HPDMK_EINVAL, "graph_steps must be >= 0");
        check_beta(beta);
        std::vector<McArgs> args;
        for (auto& h : set->reps)
            if (h->prop_rf > 0.0)
                throw HpdmkError(HPDMK_EINVAL, "replica sets run uniform proposals without features (rf = 0)");
        for (auto& h : set->reps) {
            McArgs a = mc_args(h.get(), h->chain, 0, 1, kTrialMetropolis, beta);
            a.s.probe = nullptr;
            a.ml.stats = nullptr;
            args.push_back(a);
        }
        for (int r = 0; r < set->n(); ++r) {
            args[r].move_ctr = set->move_ctr + r;
            args[r].n_moves = n;
            args[r].tiles = kReplicaTiles;  // the grid of a replica step
        }
        HIPX(hipMemcpyAsync(set->args, args.data(), sizeof(McArgs) * args.size(), hipMemcpyHostToDevice, set->stream));
        // consecutive steps alternate between the two trial-state buffers of
        // each replica (a step's set-up runs in the previous step's update launch)
        std::vector<McArgs> alt(args);
        for (int r = 0; r < set->n(); ++r) {
            alt[r].s.info = set->reps[r]->info_alt;
            alt[r].s.tabs = set->reps[r]->tabs_alt;
        }
        HIPX(hipMemcpyAsync(set->args_alt, alt.data(), sizeof(McArgs) * alt.size(), hipMemcpyHostToDevice,
                            set->stream));
        HIPX(hipMemsetAsync(set->move_ctr, 0, sizeof(int) * set->n(), set->stream));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        auto release = [&] {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        };
        try {
            HIPX(hipEventCreate(&e0));
            HIPX(hipEventCreate(&e1));
            // count consecutive steps, self-contained: a set-up first, the
            // last update without the next step's set-up
            auto step = [&](int count) {
                if (set->precision == HPDMK_FP64)
                    launch_replica_steps(launch_args<double>(set), count, set->stream);
                else
                    launch_replica_steps(launch_args<float>(set), count, set->stream);
            };
            auto compact = [&] {
                if (set->precision == HPDMK_FP64)
                    launch_replica_compact(launch_args<double>(set), set->stream);
                else
                    launch_replica_compact(launch_args<float>(set), set->stream);
            };
            // the run: steps with a slot compaction after every
            // kReplicaCompactEvery; `chunk` issues k consecutive steps
            auto schedule = [&](auto&& chunk) {
                int done = 0, since = 0;
                while (done < n) {
                    const int k = chunk(std::min(n - done, kReplicaCompactEvery - since));
                    done += k;
                    since += k;
                    if (since >= kReplicaCompactEvery) {
                        compact();
                        since = 0;
                    }
                }
            };
            if (graph_steps >= n && n > 1) {
                // the whole run captured into one graph: a single launch
                if (!set->run_exec || set->run_steps != n) {
                    if (set->run_exec) (void)hipGraphExecDestroy(set->run_exec);
                    set->run_exec = nullptr;
                    HIPX(hipStreamBeginCapture(set->stream, hipStreamCaptureModeThreadLocal));
                    schedule([&](int k) {
                        step(k);
                        return k;
                    });
                    HIPX(hipStreamEndCapture(set->stream, &graph));
                    HIPX(hipGraphInstantiate(&set->run_exec, graph, nullptr, nullptr, 0));
                    set->run_steps = n;
                }
                HIPX(hipEventRecord(e0, set->stream));
                HIPX(hipGraphLaunch(set->run_exec, set->stream));
            } else {
                // steps run as replays of one captured graph of G steps; the
                // rest (and every step when graph_steps <= 1) are launched directly
                const int G = std::min(std::min(graph_steps, n), kReplicaCompactEvery);
                if (G > 1) {
                    HIPX(hipStreamBeginCapture(set->stream, hipStreamCaptureModeThreadLocal));
                    step(G);
                    HIPX(hipStreamEndCapture(set->stream, &graph));
                    HIPX(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                }
                HIPX(hipEventRecord(e0, set->stream));
                schedule([&](int k) {
                    if (exec && k >= G) {
                        HIPX(hipGraphLaunch(exec, set->stream));
                        return G;
                    }
                    step(k);
                    return k;
                });
            }
            HIPX(hipGetLastError());
            HIPX(hipEventRecord(e1, set->stream));
            HIPX(hipEventSynchronize(e1));
            float ms = 0;
            HIPX(hipEventElapsedTime(&ms, e0, e1));
            if (gpu_ms) *gpu_ms = ms;
        } catch (...) {
            release();
            throw;
        }
        release();
        for (int r = 0; r < set->n(); ++r) {
            try {
                check_err(set->reps[r].get());
            } catch (const HpdmkError& e) {
                throw HpdmkError(e.code, "replica " + std::to_string(r) + ": " + e.what());
            }
        }
    });
}

int hpdmk_replicas_mc_results(hpdmk_replicas s, double* dE, int* accept) {
    return guard([&] {
        ReplicaSet* set = RS(s);
        const size_t n = set->n_moves;
        for (int r = 0; r < set->n(); ++r) {
            const MoveBuf& mb = set->reps[r]->chain;
            if (dE) HIPX(hipMemcpyAsync(dE + r * n, mb.dE, sizeof(double) * n, hipMemcpyDeviceToHost, set->stream));
            if (accept)
                HIPX(hipMemcpyAsync(accept + r * n, mb.acc, sizeof(int) * n, hipMemcpyDeviceToHost, set->stream));
        }
        HIPX(hipStreamSynchronize(set->stream));
    });
}

}  // extern "C"
