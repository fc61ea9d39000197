// CPU oracle for the DMK Monte Carlo hot path -- TEST INFRASTRUCTURE ONLY.
//
// This header and oracle/*.cpp restate, on the CPU and in plain C++, what the
// reference (ArrogantGao/PDMK4MC, "hpdmk") computes on its system-init ->
// total-energy -> trial-move dE -> accept/update path, plus its FP64 Ewald
// oracle.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it (as liboracle*.so through ctypes); the product library
// (pdmk4mc_amd) never links it.  Every function cites the reference file:line it
// follows.
//
// Parity status per function (what pins it; see tests/test_oracle_*.py):
//   Prolate0 (psi0, int, lambda)   pinned: residual-polynomial closed form
//                                  (test/pswf_test.cpp:24-33), lambda by two routes
//   Tree build                      pinned by invariants only (leaf cap, 2:1
//                                  balance, 26 colleagues of non-leaf nodes,
//                                  partition).  Identity with SCTL's topology:
//                                  parity unpinned (SCTL source absent).
//   outgoing/incoming/single waves pinned by brute force sums
//   energies                        pinned by the decomposition identity against
//                                  the real-space direct sums and by Ewald
//   shift dE / update               pinned by Ewald dE (test/hpdmk_test.cpp:374-533)
//   Ewald                           pinned by the EwaldSummations.jl known answer
//                                  (test/ewald_test.cpp:144)
#pragma once

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------- PSWF ----
// Prolate spheroidal wave function psi_0^c on [-1, 1] as a Legendre series,
// following the reference's construction (src/pswf.cpp:610-1060): the
// prolate differential operator is symmetric tridiagonal in the normalised
// even Legendre basis; psi_0 is its lowest eigenvector.  Evaluation outside
// [-1,1] returns 0, exactly as the reference's prol0eva does for c < 45
// (src/pswf.cpp:980-985).
struct Prolate0 {
    double c = 0;
    std::vector<double> a;  // a[k]: coefficient of P_k (odd k are zero); int psi0^2 = 2
    Prolate0() = default;
    explicit Prolate0(double c);
    double eval(double x) const;       // src/pswf.cpp:966-995 + :1135
    double int0r(double r) const;      // src/pswf.cpp:997-1027 (200-pt Gauss-Legendre)
    double intx2_0r(double r) const;   // src/pswf.cpp:1029-1060
    double lambda() const;             // src/pswf.cpp:1208-1219
};

void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);

// Constants of include/utils.hpp:20-52 for digits 3/6/9/12 plus derived values.
struct PswfParams {
    int digits = 0;
    double c = 0, lambda = 0, C0 = 0, diff0 = 0, psi0_0 = 0;
    int n_window = 0, n_diff = 0;
    std::vector<double> poly;  // ascending coefficients of P(u), u = 2x-1 in [-1,1]
};
const PswfParams& pswf_params(int digits);  // cached per digits
const Prolate0& prolate0(double c);         // cached per c

// PolyFun::eval of include/pswf.hpp:69-77: P((x-0.5)*2), 0 when |(x-0.5)*2| >= 1.
double residual_poly(const PswfParams& p, double x);

// ---------------------------------------------------------------- tree ----
struct Params {
    int n_per_leaf = 200;
    int digits = 3;
    double L = 1.0;
};

// Restatement of HPDMKPtTree<Real> (include/tree.hpp:28-148, src/tree.cpp).
template <typename Real>
struct DmkTree {
    using cplx = std::complex<Real>;
    Params params;
    PswfParams pw;
    Real L;
    int n_src = 0;

    // Morton pre-order node table (SCTL PtTree semantics, note.md:5-7)
    int n_nodes = 0, max_depth = 0;  // max_depth = number of levels (src/tree.cpp:344-349)
    std::vector<int> level, cell;     // cell: 3 ints per node
    std::vector<int> parent, child;   // child: 8 per node, -1 for leaves
    std::vector<char> is_leaf;
    std::vector<double> center;       // 3 per node
    std::vector<std::vector<int>> level_indices;
    std::vector<double> boxsize;      // boxsize[l] = L/2^l, l in [0, max_depth]
    std::vector<double> sigmas, delta_k;

    // neighbours (src/tree.cpp:136-165): ids and periodic image flags of the
    // neighbour relative to the node (the neighbour's particles are used at x + img*L)
    std::vector<std::vector<int>> colleague, coarsegrain;
    std::vector<std::vector<int8_t>> colleague_img, coarsegrain_img;

    // particles, sorted order (src/tree.cpp:301-323)
    std::vector<Real> r_sorted, q_sorted;
    std::vector<long> invmap;            // original index -> sorted index
    std::vector<std::vector<long>> node_particles;  // exact membership for every node
    Real Q = 0;

    // kernels (src/tree.cpp:84-113, include/kernels.hpp)
    std::vector<Real> window_mat;                 // d_w*d_w*(n_w+1)
    std::vector<std::vector<Real>> diff_mat;      // per level l in [2, max_depth-2]
    std::vector<std::vector<cplx>> outgoing, incoming;  // per node
    std::vector<std::vector<std::vector<cplx>>> shift_mat;  // [level][image code] (include/shiftmat.hpp)
    std::vector<cplx> root_out;                   // half-grid root expansion

    // scratch of the last eval_shift_energy (the reference keeps the same
    // state in path_to_*/outgoing_pw_* members, include/tree.hpp:68-72)
    std::vector<int> path_o, path_t;
    std::vector<std::vector<cplx>> pw_o, pw_t;
    long last_i = -1;
    Real last_dx = 0, last_dy = 0, last_dz = 0;

    DmkTree(const Params& p, int n, const double* r, const double* q);

    int dim_window() const { return (2 * pw.n_window + 1) * (2 * pw.n_window + 1) * (pw.n_window + 1); }
    int dim_diff() const { return (2 * pw.n_diff + 1) * (2 * pw.n_diff + 1) * (pw.n_diff + 1); }

    int locate(Real x, Real y, Real z, std::vector<int>& path) const;  // src/tree.cpp:204-236
    void form_outgoing_pw();   // src/planewave.cpp:24-97 (DIRECT, exact root sum)
    void form_outgoing_pw_root();   // src/planewave.cpp:24-46 (exact sum instead of FINUFFT)
    // PROXY (src/upward_pass.cpp:206-322), oracle/proxy_oracle.cpp
    void form_outgoing_pw_proxy();
    std::vector<double> proxy_from_particles(int node, int order) const;  // upward_pass.cpp:109-149
    std::vector<std::vector<double>> proxy_coeffs;  // per node, order^3 (a fastest), after PROXY
    void form_incoming_pw();   // src/planewave.cpp:99-155
    double eval_energy_window();  // src/energy.cpp:21-39
    double eval_energy_diff();    // src/energy.cpp:42-79
    double eval_energy_res();     // src/energy.cpp:82-173
    double eval_energy_window_direct();  // src/energy_direct.cpp:21-59
    double eval_energy_diff_direct();    // src/energy_direct.cpp:62-93
    double eval_energy_res_direct();     // src/energy_direct.cpp:153-186

    void form_outgoing_pw_single(std::vector<std::vector<cplx>>& pw, const std::vector<int>& path,
                                 Real x, Real y, Real z, Real q);  // src/planewave.cpp:157-181
    // returns dE; parts = {window, diff, res_target, res_origin, self}
    double eval_shift_energy(long i_unsorted, Real dx, Real dy, Real dz, double* parts = nullptr);
    void update_shift(long i_unsorted, Real dx, Real dy, Real dz);  // src/tree.cpp:447-597
    double res_near(long i_sorted, const std::vector<int>& path, Real x, Real y, Real z) const;

    // helpers
    int find_node(int l, int cx, int cy, int cz) const;  // exact node at level l or -1
};

// ---------------------------------------------------------------- PROXY ----
// Chebyshev machinery of include/chebychev.hpp and src/upward_pass.cpp
// (oracle/proxy_oracle.cpp).  Matrices are row-major.
int proxy_order(int digits);                                   // upward_pass.cpp:9-14 (+ digits 6)
std::vector<double> cheb_nodes(int p);                         // chebychev.hpp:35-42
std::vector<double> cheb_vandermonde(const std::vector<double>& x, int p);  // chebychev.hpp:45-60
std::vector<double> lu_inverse(std::vector<double> A, int p);
std::vector<double> cheb_c2p(int p, int half);                 // chebychev.hpp:180-193, 203-296
std::vector<std::complex<double>> cheb_poly2pw(int p, int n);  // upward_pass.cpp:152-180 (conjugated)

// ---------------------------------------------------------------- Ewald ----
// src/ewald.cpp:17-308 with a cell list instead of vesin.
struct Ewald {
    double L, s, alpha, eps, V, r_c, k_c;
    int n;
    int nk;  // k grid half-width
    std::vector<double> q, r;
    std::vector<std::complex<double>> S;  // structure factor on the cube
    std::vector<double> kmat;
    Ewald(double L, double s, double alpha, double eps, int n, const double* q, const double* r);
    void init_planewave_coeffs();
    double energy_short() const;
    double energy_long() const;
    double energy_self() const;
    double compute_energy();
    double potential(double x, double y, double z, long exclude = -1) const;
    // exact Ewald dE of moving particle p from its position to target t
    double shift_energy(long p, double tx, double ty, double tz) const;
    void apply_move(long p, double tx, double ty, double tz);
};

template <typename Real>
inline Real my_mod(Real x, Real L) {  // include/utils.hpp:96-105
    while (x < 0) x += L;
    while (x >= L) x -= L;
    return x;
}

}  // namespace oracle
