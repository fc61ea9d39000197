// Arguments of the PROXY (Chebyshev upward pass) kernels, dmk_proxy.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "dmk_common.hpp"
#include "dmk_kernels.hpp"

namespace hpdmk_amd {

struct ProxyArgs {
    int p = 0;                   // Chebyshev order: 18 (digits 3) or 28 (digits 6)
    const double* A = nullptr;   // child-to-parent matrices, lower then upper half, A[h p^2 + l p + a]
    const double* P = nullptr;   // coefficient-to-plane-wave matrix, (2n+1) x p complex (re, im)
    double* coef = nullptr;      // p^3 per pw slot, a fastest
    const int* direct = nullptr; // pw slots whose coefficients come from their particles
    int n_direct = 0;
    const int* c2p = nullptr;    // non-leaf pw slots of levels 2..D-3, grouped by level
    int n_levels = 0;            // levels 2 .. 2 + n_levels - 1 in c2p
    int level_off[kMaxLevels] = {};
    int level_cnt[kMaxLevels] = {};
};

template <typename T>
void launch_outgoing_root(const DevTree<T>& t, double* acc_root, hipStream_t st);
template <typename T>
void launch_outgoing_proxy(const DevTree<T>& t, const PwNodes& pn, double* acc_root, const ProxyArgs& pa,
                           hipStream_t st);

}  // namespace hpdmk_amd
