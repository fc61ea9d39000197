// Arguments of the PROXY (Chebyshev upward pass) kernels, dmk_proxy.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "dmk_common.hpp"
#include "dmk_kernels.hpp"

namespace hpdmk_amd {

// One operand of a batched FP64 matrix-core GEMM (k_gemm64).  Batch entry b
// = b0 inner + b1 starts at p + (tab ? tab[b0] : b0) s0 + b1 s1.  Its rows
// come in segments of seg rows: segment s starts at (stab ? stab[b0 nseg +
// s] : s) sstride further (a negative stab entry: an all-zero segment).
// Element (r, c) of a segment, r < seg: r ld + c (row-major) or r + c ld
// (column-major).
struct GemmOp {
    double* p = nullptr;
    long ld = 0;
    int col = 0;
    const int* tab = nullptr;
    long s0 = 0, s1 = 0;
    int seg = 1 << 30, nseg = 1;
    const int* stab = nullptr;
    long sstride = 0;
};

// C (M x N) = A (M x K) B (K x N) for every batch entry (batch = outer x inner)
struct GemmPass {
    int M = 0, N = 0, K = 0, inner = 1, batch = 0;
    GemmOp a, b, c;
};

// child -> parent transfer of one chunk of parents of one level: the x and y
// passes run per child (any octant), the z pass per parent over its eight
// children at once (absent children are zero segments)
struct ProxyChunk {
    int n_par = 0, n_ch = 0;
    const int* par = nullptr;   // parent pw slots
    const int* ch = nullptr;    // child pw slots
    const int* ox = nullptr;    // per child: x half (0 lower, 1 upper) -> matrix A_ox
    const int* oy = nullptr;    // per child: y half
    const int* seg = nullptr;   // per parent, per octant: the child's index in ch, or -1
};

struct ProxyArgs {
    int p = 0;                   // Chebyshev order: 18 (digits 3), 28 (digits 6), 44 (digits 9)
    int n = 0, d = 0, nz = 0;    // plane waves: k = -n..n per axis (d = 2n + 1), kz <= 0 (nz = n + 1)
    int nw = 0;                  // root (window) plane waves: k = -nw..nw per axis
    const double* A = nullptr;   // child-to-parent matrices, lower then upper half, A[h p^2 + l p + a]
    const double* Az8 = nullptr; // p x 8p: [A_{z(0)} .. A_{z(7)}], z(ci) = ci >> 2 (the z pass of all octants)
    const double* Px = nullptr;  // 2d x p: rows (re, kx) then (im, kx) of the coefficient-to-plane-wave matrix
    const double* Py = nullptr;  // 2d x 2p: [[P_re, -P_im], [P_im, P_re]]
    const double* Pz = nullptr;  // 2nz x 2p: the same for kz = 0..n
    // the root from the level-2 boxes: the coefficient-to-plane-wave matrix
    // at the root's wavenumbers (k - nw) 2 pi / L over a box of size L / 4,
    // stacked like Px, Py, Pz; used when every particle lies in a level-2 box
    const double* Prx = nullptr;
    const double* Pry = nullptr;
    const double* Prz = nullptr;
    const int* root_slots = nullptr;  // pw slots of the level-2 boxes
    int n_root = 0;                   // 0: the root from the particles
    double* coef = nullptr;      // p^3 per pw slot, a fastest
    double* s1 = nullptr;        // scratch: x-pass results (c2p: chunk children p^3; pw: chunk boxes 2d p^2)
    double* s2 = nullptr;        // scratch: y-pass results (c2p: chunk children p^3; pw: chunk boxes 2 d^2 p)
    double* s3 = nullptr;        // scratch: z-pass results of the plane waves (chunk boxes 2 nz d^2)
    int pw_chunk = 0;            // boxes per plane-wave chunk
    const int* direct = nullptr; // pw slots whose coefficients come from their particles
    int n_direct = 0;
    std::vector<ProxyChunk> c2p; // from the deepest parent level up
};

template <typename T>
void launch_outgoing_root(const DevTree<T>& t, double* acc_root, hipStream_t st);
template <typename T>
void launch_outgoing_proxy(const DevTree<T>& t, const PwNodes& pn, double* acc_root, const ProxyArgs& pa,
                           hipStream_t st);
// C = A B over a batch on the matrix cores (exposed for the unit test of the GEMM)
void launch_gemm64(const GemmPass& g, hipStream_t st);

}  // namespace hpdmk_amd
