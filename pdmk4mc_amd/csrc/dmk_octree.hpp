// Device build of the octree topology (dmk_octree.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "dmk_host.hpp"

namespace hpdmk_amd {

// The tree build_octree (dmk_host.cpp) builds, node for node, computed on the
// GPU from the host coordinates r (N x 3): Morton keys, radix sort,
// refinement, periodic 2:1 balance, pre-order, colleagues and near lists.
// Returns a tree with n_nodes == 0 when the tree is deeper than the device
// build's 64-bit pre-order keys allow (more than 20 levels); the caller then
// builds it on the host.  Throws HpdmkError like build_octree.
Octree build_octree_device(int n, const double* r, double L, int n_per_leaf, hipStream_t st);

}  // namespace hpdmk_amd
