// Host-side set-up of a DMK system for the MI355X path: PSWF constants,
// Morton-sorted octree topology, neighbour tables, slot layout and the
// window/difference kernel tables that are uploaded to HBM.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace hpdmk_amd {

// Constants of the prolate spheroidal wave function psi_0^c used by the DMK
// split (bandwidth c, lambda, C0 = int_0^1 psi0, diff0, psi0(0)) and the
// residual polynomial, for digits 3/6/9/12 (include/utils.hpp:20-52).
struct PswfConsts {
    int digits = 0;
    double c = 0, lambda = 0, C0 = 0, diff0 = 0, psi0_0 = 0;
    int n_window = 0, n_diff = 0;
    std::vector<double> poly;  // ascending powers of u = 2x - 1
    std::vector<double> a;     // Legendre coefficients of psi0 (degree k)
    double psi0(double x) const;  // 0 for |x| > 1 (src/pswf.cpp:980-985)
};
PswfConsts pswf_consts(int digits);

// Octree topology with SCTL PtTree semantics: a box is refined while it holds
// more than n_per_leaf points, levels 0 and 1 are always refined, and the tree
// is 2:1 balanced over the 26 periodic neighbours (every non-leaf box has all
// of its 26 same-level neighbours).  Nodes are numbered in Morton pre-order.
struct Octree {
    int n_nodes = 0, max_depth = 0;  // max_depth = number of levels
    std::vector<int> level, cell, child, parent;  // cell: 3 per node; child: 8 per node (-1 for leaves)
    std::vector<int8_t> is_leaf;
    std::vector<int> colleague;  // 26 per node (levels >= 2), -1 when absent
    std::vector<int> near_off, near_ids;  // per node CSR (leaves only non-empty)
    std::vector<int8_t> near_img;
    std::vector<int> order;       // original particle ids in leaf pre-order, Morton order inside a leaf
    std::vector<int> leaf_count;  // per node: particles (leaves), 0 otherwise
    std::vector<int> first_leaf, last_leaf;  // per node: pre-order range of descendant leaves
};
Octree build_octree(int n, const double* r, double L, int n_per_leaf);

// Half-grid tables (weight 2 for k_z < 0, 1 at k_z = 0), include/kernels.hpp:61-113
std::vector<double> window_table(const PswfConsts& p, double L);
std::vector<double> difference_table(const PswfConsts& p, double L, int level);

// Chebyshev tables of the PROXY initialisation (src/upward_pass.cpp,
// include/chebychev.hpp), from the discrete orthogonality of T_a at the
// first-kind nodes x_i: V^-1[a][i] = (2 - delta_a0) / p T_a(x_i).
//   A[h p^2 + l p + a]: T_l(x/2 + (h ? 1/2 : -1/2)) = sum_a A[l][a] T_a(x)
//   P[2 (k p + a) + {0,1}] = sum_i e^{-i (k - n) (pi/3) x_i} V^-1[a][i], k = 0..2n
struct ProxyTables {
    int p = 0;
    std::vector<double> A, P;
};
int proxy_order(int digits);  // 18 for digits <= 3 (src/upward_pass.cpp:9-14), 28 for 6, 44 for 9, else throws
// A: child-to-parent matrices; P[k][a] (complex) = sum_i e^{-i (k - n) hw x_i} (V^-1)[a][i]
// over the Chebyshev nodes x_i: hw = (k step) x (half the box size), pi / 3 on every level
ProxyTables proxy_tables(int p, int n, double hw = M_PI / 3.0);

struct HpdmkError : std::runtime_error {
    int code;
    HpdmkError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

}  // namespace hpdmk_amd
