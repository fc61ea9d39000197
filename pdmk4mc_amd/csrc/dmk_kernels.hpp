// Kernel argument blocks and launchers of dmk_build.hip and dmk_mc.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "dmk_common.hpp"
#include "dmk_proposal.hpp"

namespace hpdmk_amd {

constexpr int kMaxBatch = 256;     // trials per batched step
constexpr int kMarkWords = 4096;   // shared-memory bitmap of rewritten arrays: 2 n_pw <= 32 kMarkWords
// candidate arrays of one trial's update (own out, own in, 26 colleagues' in per
// endpoint and expansion level), listed ahead of the decisions for batches of at
// most 64 trials in trees of at most kScanLv expansion levels
constexpr int kScanLv = 4;
constexpr int kScanCand = 2 * kScanLv * 64 * 28;  // 16-bit entries, [endpoint, level][trial][28]
constexpr int kMaxGridD = 53;  // largest 2n+1 of any digits setting
constexpr int kRowStride = 64;     // stride of one 1-D phase row (2n+1 <= 53)
constexpr int kAccPerTrial = 7;    // window, difference, res_t, res_o, scanned_t, scanned_o, overlaps_t
// per trial block: partial sum, particles scanned, hard-sphere overlaps, and
// the learned proposal's feature sums at the target (kFeatSums)
constexpr int kPartStride = 3 + kFeatSums;
constexpr int kStatsPerMove = 8;   // levels t/o, near particles t/o, rewritten arrays t/o,
                                    // difference terms of the pair corrections (as the later trial),
                                    // distinct arrays the move's batch rewrites (first move of a batch)

// Everything later kernels need about one trial move: endpoints, charge and
// the root-to-leaf paths of the origin o and the target t = wrap(o + d).
struct TrialInfo {
    double o[3], t[3], dv[3];
    double q;
    int i;
    int lam_o, lam_t;  // leaf levels
    int path_o[kMaxLevels], path_t[kMaxLevels];
    // fused single-move chain only (prep_body): expansion slot of each path
    // box (-1 outside levels 2..min(leaf level, D-2)) and the near list of
    // each endpoint's leaf; index 0 target, 1 origin
    int pws[2][kMaxLevels];
    int noff[2], ncnt[2];
    int m;        // the move's index in the move list
    int rs;       // tabs slot of its phase rows and McScratch::st slot of its self term and threshold:
                  // m & 1, or 2 + (m & 1) when set up again after the previous move moved its particle
    double u;     // its Metropolis uniform
};

// Fused single-move chain: a move two ahead, prefetched by the update kernel
// of move m for the trial kernel of move m + 1 (which sets up move m + 2):
// its particle and where that particle is after move m.
struct NextMove {
    double o[3], dv[3], q, u;
    int i, m;
};

// Phase rows of one trial (Cplx<T>, kRowStride apart):
//   root:  [endpoint][axis]                     e^{-i (m-n_w) dk_0 x_a}
//   level: [endpoint][level-2][axis][shift+1]    e^{-i (m-n_d) dk_l (x_a + shift L)}
// endpoint 0 = target, 1 = origin; the x-row carries the charge.
__host__ __device__ inline int table_stride(int max_depth) {
    const int nL = max_depth >= 4 ? max_depth - 3 : 0;
    return (6 + 2 * nL * 9) * kRowStride;
}
__host__ __device__ inline int root_row(int ep, int axis) { return (ep * 3 + axis) * kRowStride; }
__host__ __device__ inline int level_row(int nL, int ep, int li, int axis, int shift) {
    return (6 + ((ep * nL + li) * 3 + axis) * 3 + shift + 1) * kRowStride;
}

enum TrialMode : int {
    kTrialOnly = 0,         // dE only
    kTrialMetropolis = 1,   // dE, accept = u < exp(-beta dE), update plan
    kTrialForceAccept = 2,  // dE, accept, update plan (update_shift)
    kTrialBatch = 3,        // partial sums only; the batch scan assembles dE
};

// Launch shape of the trial and update kernels (tile autotuner's search
// space).  k_trial gives each trial `win` window blocks, `diff` blocks per
// (endpoint, difference level) and `near` blocks per endpoint's near field;
// each block strides over its share.  k_apply runs `apply` blocks.
struct Tiles {
    int win = 2, diff = 8, near = 8, apply = 1024;
};
constexpr int kMaxRoleBlocks = 32;  // upper bound of win, diff, near (scratch sizing)
// partial sums per lane of the fused single-move update's decision wave: it
// reads at most 64 kUpd1P trial blocks (the host runs launch shapes with more
// through the four-kernel chain)
constexpr int kUpd1P = 6;
// near-list entries of the next move that the fused update's block 0 writes:
// 4 per thread of its 256, so the fused chain needs 2 max_near <= kUpd1Near
constexpr int kUpd1Near = 4 * 256;
__host__ __device__ inline int trial_blocks(int max_depth, const Tiles& g) {
    const int nL = max_depth >= 4 ? max_depth - 3 : 0;
    return g.win + 2 * nL * g.diff + 2 * g.near;
}
__host__ __device__ inline int max_trial_blocks(int max_depth) {
    const int nL = max_depth >= 4 ? max_depth - 3 : 0;
    return kMaxRoleBlocks * (1 + 2 * nL + 2);
}

// one trial of a batch, assembled from its partials before the decisions
struct TrialRec {
    double dE;       // against the batch's starting state
    double thr;      // decision threshold: -ln(u) / beta (+ H / beta with a learned proposal)
    int room;        // free slots of the target leaf (1 << 30 when the leaf does not change)
    int leaf_t, leaf_o;
    int ov;          // particles within sigma of the target
    int pid;         // particle index
    int pad;
};

// Window part of the pair corrections as a Gram product on the matrix cores
// (FP32 tables): extra blocks of k_trial, each over kGramModes window modes
// in LDS chunks of kGramSW, for batches of 2..kGramMaxK trials.
constexpr int kGramMaxK = 64;
constexpr int kGramSW = 16;
constexpr int kGramChunks = 2;
constexpr int kGramModes = kGramSW * kGramChunks;
__host__ __device__ inline int gram_blocks(int Mw) { return (Mw + kGramModes - 1) / kGramModes; }

// device scratch of the Monte Carlo kernels (sized for kMaxBatch trials)
struct McScratch {
    TrialInfo* info;
    // fused single-move chain: near-field entries of each ring slot's move,
    // [slot][endpoint][1 + max_near]: the count, then {slot start, particle
    // count, image code, leaf} per near-list leaf
    int4* near;
    NextMove* nm;          // fused chain: [2], the move m + 2 prefetched by move m's update (slot m & 1)
    double2* st;           // fused chain: [4], self-pair correction q^2 R_{b_2}(d_min)
                           // (src/energy_shift.cpp:58) and threshold -ln(u) / beta per rows slot
    void* tabs;            // Cplx<T>, table_stride(D) per trial
    double* part;          // per trial and block: kPartStride values
    TrialRec* rec;         // per trial of a batch
    int* accept;           // per trial of the current launch
    double* corr;          // kMaxBatch^2
    double* gram;          // FP32 tables, batches of at most kGramMaxK: window pair terms,
                           // gram_blocks(Mw) shares per pair (null: the pair blocks form them)
    unsigned char* ovl;    // kMaxBatch^2 hard-sphere flags of pair (i, j): 1 t_j near o_i, 2 t_j near t_i
    int* dest;             // rewritten arrays, 2*slot + kind (0 out, 1 in): level-2 in arrays from
                           // the front, the others from the back (2 n_pw entries)
    int* dinfo;            // fused single-move chain: per dest entry the array's level and box cell,
                           // l | cx << 4 | cy << 13 | cz << 22 (dest then in marking order)
    int* plan;             // [0] arrays in dest, [1] accepted moves, [2] level-2 in arrays,
                           // [3] single move: the target leaf has room (marking row)
    unsigned long long* probe;  // optional phase timers (profiling): kProbes sums + counts
    unsigned short* cand = nullptr;  // batches of <= 64: candidate arrays of each trial's update
                                     // ((ep nL + li) 64 + trial) 28 + c, 0xffff none
};

// phase timers of the decision/plan path (wall_clock64 ticks, summed over launches)
enum Probe : int {
    kPrScanAssemble = 0, kPrScanDecide, kPrPlanMark, kPrPlanCompact, kPrPlanRelocate, kPrTrialFinal,
    kPrApplyMatch, kPrApplyTiles, kPrApplyRmw, kPrScanPrep, kPrScanWalk, kPrPlanUnite,
    kProbes
};

struct MoveList {
    const int* idx;     // caller's particle indices
    const double* d;    // displacements, 3 per move
    const double* u;    // Metropolis uniforms
    double* dE;         // final dE per move
    double* dE_stale;   // batched: dE against the batch's starting state
    double* parts;      // optional: window, diff, res_t, res_o, self per move
    int* accept_out;    // optional: per move
    int* stats;         // optional: kStatsPerMove per move
};

// boxes of levels 2..D-2 in pw-slot order
struct PwNodes {
    const int* node;
    const int* level;
    const int* lo;  // slot range of the box's descendant leaves
    const int* hi;
};

struct McArgs {
    McScratch s;
    MoveList ml;
    PwNodes pn;
    int move0;   // first move of this launch in the move list (after *move_ctr when set)
    // when set, the first move is *move_ctr + move0 (steps replayed from a
    // graph; batched learned steps, whose scan advances it) and a launch takes
    // at most n_moves - *move_ctr - move0 trials
    int* move_ctr;
    int n_moves;
    // fused single-move chain: move m's trial state (TrialInfo, phase rows)
    // lives in slot m & 1 = move0 & 1 (a graph's first move, *move_ctr, is
    // even), so kernels find it without reading the counter; the update
    // kernel decides from the trial partials and moves the particle; the
    // trial kernel prepares move m + 1 in the other slot
    int ring;
    // batched chains whose update kernel also sets up the next step
    // (launch_apply_next): the scan moves the accepted particles, so that the
    // next step's set-up reads final positions from the update's first block on
    int scan_reloc;
    int K;       // trials in this launch (at most)
    int mode;    // TrialMode
    double beta;
    // Monte Carlo model (Metropolis and batch modes only): the decision uses
    // lB * dE (Bjerrum length in units of the box length unit), and a target
    // closer than sigma to another particle is rejected (sigma2 = sigma^2, 0 = off)
    double lB;
    double sigma2;
    Tiles tiles;
    Proposal prop;  // learned proposal (Metropolis and batch modes)
};

// single-move steps of launch_trial/launch_decide (plain chain,
// update_shift): the trial kernel's last row of blocks marks the update plan,
// the decision moves the particle
__host__ __device__ inline bool premark(const McArgs& a) {
    return a.K == 1 && (a.mode == kTrialMetropolis || a.mode == kTrialForceAccept);
}

// a batch of at most 64 trials moves its particles in the first block of
// k_apply (one wave ranks the movers), unless its scan does (scan_reloc);
// other steps move them in the plan
__host__ __device__ inline bool apply_relocates(const McArgs& a) {
    return a.mode == kTrialBatch && a.K <= 64 && !a.scan_reloc;
}

// the batch's window pair terms come from the Gram blocks of k_trial
template <typename T>
__host__ __device__ inline bool use_gram(const McArgs& a) {
    return std::is_same<T, float>::value && a.s.gram != nullptr && a.mode == kTrialBatch && a.K >= 2 &&
           a.K <= kGramMaxK;
}
// rows of Gram blocks in front of the trial rows of k_trial's grid
template <typename T>
__host__ __device__ inline int gram_rows(const DevTree<T>& t, const McArgs& a) {
    if (!use_gram<T>(a)) return 0;
    const int tb = trial_blocks(t.max_depth, a.tiles);
    return (gram_blocks(t.Mw) + tb - 1) / tb;
}

// Metropolis steps of every replica of a set (one move each): per-replica
// device state and arguments live in device arrays indexed by blockIdx.z.
template <typename T>
struct ReplicaLaunch {
    const DevTree<T>* trees;  // n
    const McArgs* args;       // n (move_ctr set)
    const McArgs* args_alt;   // n: the same with each replica's other trial-state buffers
    int* move_ctr;            // n step counters
    const int* const* leaves; // n leaf lists
    const int* n_leaves;      // n
    int n, max_depth, max_leaves, apply_blocks;
    int dmax;     // largest grid points per axis over the replicas (shared-memory sizing)
    Tiles tiles;  // trial shape (the same for every replica)
};

// k consecutive Metropolis steps of every replica (one move each): a set-up
// first, then per step the trial, the decision and the update, every update
// but the last setting up the next step in its launch (the steps' trial
// state alternating between args and args_alt)
template <typename T>
void launch_replica_steps(const ReplicaLaunch<T>& r, int k, hipStream_t st);
// move_ctr[0..n) += inc
void launch_advance(int* move_ctr, int n, hipStream_t st, int inc = 1);
template <typename T>
void launch_replica_compact(const ReplicaLaunch<T>& r, hipStream_t st);

// single moves and trial-only batches: dE, decision, update plan
template <typename T>
void launch_decide(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr);
template <typename T>
void launch_setup(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr);
// the update of a batched step (a, scan_reloc set) and the set-up of the next
// (an: its trials' state in the other buffers of the pair; prop: the
// proposal kernel's set-up, else the uniform one) in one launch
template <typename T>
void launch_apply_next(const DevTree<T>& t, const McArgs& a, const McArgs& an, bool prop, hipStream_t st,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// features at the origins and, for the learned proposal, the displacements;
// then the set-up of launch_setup, in one kernel
template <typename T>
void launch_propose_setup(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0 = nullptr,
                          hipEvent_t e1 = nullptr);
template <typename T>
void launch_trial(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr);
template <typename T>
void launch_pair_corr(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr);
template <typename T>
void launch_scan(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr);
template <typename T>
void launch_apply(const DevTree<T>& t, const McArgs& a, hipStream_t st, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr);
template <typename T>
void launch_compact(const DevTree<T>& t, const int* leaves, int n_leaves, hipStream_t st);

void launch_copy16(const void* src, void* dst, size_t n16, hipStream_t st);

template <typename T>
void launch_outgoing(const DevTree<T>& t, const PwNodes& pn, double* acc_root, hipStream_t st);
template <typename T>
void launch_incoming(const DevTree<T>& t, const PwNodes& pn, hipStream_t st);
template <typename T>
void launch_energy(const DevTree<T>& t, const PwNodes& pn, const int* leaves, int n_leaves, double* res,
                   hipStream_t st);

}  // namespace hpdmk_amd
