// Kernel argument blocks and launchers of dmk_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "dmk_common.hpp"

namespace hpdmk_amd {

constexpr int kMaxBatch = 256;  // trials per batched step

// Everything later kernels need about one trial move: endpoints, charge and
// the root-to-leaf paths of the origin o and the target t = wrap(o + d).
struct TrialInfo {
    double o[3], t[3], dv[3];
    double q;
    int i;
    int lam_o, lam_t;  // leaf levels
    int path_o[kMaxLevels], path_t[kMaxLevels];
};

enum TrialMode : int {
    kTrialOnly = 0,         // dE only
    kTrialMetropolis = 1,   // dE and accept = u < exp(-beta dE)
    kTrialForceAccept = 2,  // dE and accept (update_shift after eval_shift_energy)
};

struct TrialArgs {
    const int* idx;     // caller's particle indices (move list)
    const double* d;    // displacements, 3 per move
    const double* u;    // uniforms for the Metropolis test
    double beta;
    int move0;          // first move of this launch in the move list
    int mode;
    double* acc;        // kAccPerTrial FP64 accumulators per trial (self-resetting)
    unsigned* done;     // block counter per trial (self-resetting)
    double* dE;         // per move
    double* parts;      // optional: window, diff, res_t, res_o, self per move
    int* accept;        // per trial of this launch (read by the update)
    int* accept_out;    // optional: per move
    TrialInfo* info;    // per trial of this launch
    int* stats;         // optional, per move: difference levels (t, o), near-field
                        // particles scanned (t, o), arrays an accept rewrites (t, o)
};

constexpr int kAccPerTrial = 6;  // window, difference, res_t, res_o, scanned_t, scanned_o
constexpr int kStatsPerMove = 6;

struct UpdateArgs {
    int n_moves;
    const int* accept;
    const TrialInfo* info;
};

struct CorrArgs {
    int K;
    const TrialInfo* info;
    double* corr;  // K*K, corr[i*K + j] = effect of move i on trial j
};

struct ScanArgs {
    int K, move0;
    const TrialInfo* info;
    const double* corr;
    const double* dE_stale;  // per move
    const double* u;
    double beta;
    double* dE;              // per move (final)
    int* accept;             // per trial of the batch
    int* accept_out;         // optional, per move
};

// boxes of levels 2..D-2 in pw-slot order
struct PwNodes {
    const int* node;
    const int* level;
    const int* lo;  // slot range of the box's descendant leaves
    const int* hi;
};

template <typename T>
void launch_trial(const DevTree<T>& t, const TrialArgs& a, int n_trials, hipStream_t st);
template <typename T>
void launch_update(const DevTree<T>& t, const UpdateArgs& a, hipStream_t st);
template <typename T>
void launch_pair_corr(const DevTree<T>& t, const CorrArgs& a, hipStream_t st);
template <typename T>
void launch_scan(const DevTree<T>& t, const ScanArgs& a, hipStream_t st);
template <typename T>
void launch_outgoing(const DevTree<T>& t, const PwNodes& pn, double* acc_root, hipStream_t st);
template <typename T>
void launch_incoming(const DevTree<T>& t, const PwNodes& pn, hipStream_t st);
template <typename T>
void launch_energy(const DevTree<T>& t, const PwNodes& pn, const int* leaves, int n_leaves, double* res,
                   hipStream_t st);

}  // namespace hpdmk_amd
