// Mixed-precision DIRECT build of the outgoing expansions on FP32 MFMA
// (dmk_nudft_mfma.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "dmk_common.hpp"

namespace hpdmk_amd {

// one work item: a run of a box's descendant leaves (positions in the
// pre-order leaf list); slot -1 is the root.  acc >= 0: the box is split over
// several items and sums into acc_box row `acc`
struct PwWork {
    int slot, node, leaf_lo, leaf_hi, acc;
};

struct PwMfmaPlan {
    PwWork* work = nullptr;
    int n_work = 0;
    double* acc_box = nullptr;  // n_acc x 2 Md
    int* acc_slots = nullptr;   // pw slot of each acc row
    int n_acc = 0;
};

constexpr int kMfmaLeavesPerItem = 8;  // leaves per work item (<= 256)

void launch_outgoing_mfma(const DevTree<float>& t, const int* leaves, const PwMfmaPlan& pl, double* acc_root,
                          hipStream_t st);

}  // namespace hpdmk_amd
