// Device-resident state of one DMK system (one Monte Carlo replica) on an
// MI355X, shared by the host orchestration (hpdmk_capi.cpp) and the kernels
// (dmk_kernels.hip).  Passed to kernels by value.
//
// HBM layout (all arrays device-resident; T = double in FP64 mode, float in
// mixed mode):
//   nodes      Morton pre-order octree: child[8n], level[n], is_leaf[n], cell[3n]
//              colleague[26n] (same-level neighbours, -1 if absent, offset order
//              (dx+1) + 3(dy+1) + 9(dz+1) with 13 skipped), pw_slot[n]
//   near       per leaf: CSR list of leaves whose particles enter its residual
//              sum, with the periodic image code of each (0..26)
//   particles  slot arrays px,py,pz,pq (double), porig (int); leaf B owns slots
//              [slot_start[B], slot_start[B] + slot_cap[B]), slot_count[B] used
//   expansions root_out[Mw] (complex T), out/in[n_pw_nodes * Md] (complex T)
//              for every node of levels 2..D-2, half grid k_z <= 0, element
//              index i + j*d + k*d*d
//   kernels    Wmat[Mw], Dmat[(D-3) * Md] (T, half-plane weights included)
#pragma once

#include <cstdint>

namespace hpdmk_amd {

constexpr int kMaxLevels = 16;
constexpr int kMaxPoly = 25;
constexpr int kNodeAtLevels = 8;  // dense cell -> node tables up to this many levels (2.4M entries)

template <typename T>
struct Cplx {
    T x, y;
};

template <typename T>
struct DevTree {
    // ---- sizes
    int n_src, n_nodes, max_depth, n_window, n_diff;
    int dw, dd, Mw, Md, n_pw_nodes, total_slots, max_near;
    // ---- scalars (FP64 on device)
    double L;
    double boxsize[kMaxLevels + 1];
    double delta_k[kMaxLevels + 1];
    double cW;                  // dk0^3 / (2 (2 pi)^3), window energy prefactor
    double cD[kMaxLevels];      // dk_l^3 / (2 (2 pi)^3) per level
    double self_window;         // psi0(0) / (2 b_2 C0)
    double self_diff[kMaxLevels];  // psi0(0)(1/(2 b_{l+1} C0) - 1/(2 b_l C0))
    int n_poly;
    double poly[kMaxPoly];      // residual polynomial, ascending powers of u = 2r/b - 1
    // ---- tree
    const int* child;
    const int* level;
    const int8_t* is_leaf;
    const int* cell;
    const int* colleague;
    const int* col_in_slot;  // 26 per node: expansion slot of a non-leaf colleague, else -1
    const int* pw_slot;
    const int* node_at;  // levels 0..D-1 when D <= kNodeAtLevels: node of each cell, -1 below a leaf
                         // (level l at offset (8^l - 1) / 7, cell index cx + 2^l cy + 4^l cz); else null
    const int* near_off;
    const int* near_ids;
    const int8_t* near_img;
    // ---- particles
    const int* slot_start;
    const int* slot_cap;
    int* slot_count;
    double* px;
    double* py;
    double* pz;
    double* pq;
    int* porig;
    int* slot_of;
    int* leaf_of;
    // ---- expansions and kernels
    Cplx<T>* root;
    Cplx<T>* out;
    Cplx<T>* inc;
    const T* Wmat;
    const T* Dmat;
    // the difference tables folded over k_x -> -k_x and k_y -> -k_y (they depend
    // on |k| only): per level (n+1)^3 entries, x' + (n+1) (y' + (n+1) z) holding
    // Dmat at x = n + x', y = n + y' (pair corrections)
    const T* Dfold;
    // ---- sticky error word (0 = ok); set by kernels, read by the host
    int* err;
};

// Error bits written to DevTree::err by device code.
enum DevError : int {
    kErrLeafOverflow = 1,  // an accepted move would overflow the target leaf's slot capacity
};

}  // namespace hpdmk_amd
