// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle/dmk_oracle.hpp).
//
// extern "C" surface of the oracle for ctypes (tests/oracle_lib.py).  Errors
// never cross this boundary as exceptions: each call returns 0 on success and
// a negative code on failure, with the message in oracle_last_error().
#include <chrono>
#include <cstring>
#include <memory>
#include <random>
#include <string>

#include "dmk_oracle.hpp"

using namespace oracle;

namespace {
thread_local std::string g_err;

struct Handle {
    int precision;  // 64 or 32
    std::unique_ptr<DmkTree<double>> t64;
    std::unique_ptr<DmkTree<float>> t32;
};

template <typename F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

template <typename F>
auto with_tree(void* h, F&& f) {
    auto* H = static_cast<Handle*>(h);
    if (H->precision == 64) return f(*H->t64);
    return f(*H->t32);
}
}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// ---- PSWF ---------------------------------------------------------------
// out: c, lambda, C0, diff0, psi0(0), n_window, n_diff, n_poly
int oracle_pswf_params(int digits, double* out) {
    return guard([&] {
        const auto& p = pswf_params(digits);
        out[0] = p.c;
        out[1] = p.lambda;
        out[2] = p.C0;
        out[3] = p.diff0;
        out[4] = p.psi0_0;
        out[5] = p.n_window;
        out[6] = p.n_diff;
        out[7] = (double)p.poly.size();
    });
}
double oracle_prolate0_eval(double c, double x) { return prolate0(c).eval(x); }
double oracle_prolate0_int(double c, double r) { return prolate0(c).int0r(r); }
double oracle_prolate0_intx2(double c, double r) { return prolate0(c).intx2_0r(r); }
double oracle_prolate0_lambda(double c) { return prolate0(c).lambda(); }
int oracle_prolate0_coefs(double c, double* out, int nmax) {
    const auto& a = prolate0(c).a;
    int n = (int)a.size() < nmax ? (int)a.size() : nmax;
    for (int i = 0; i < n; ++i) out[i] = a[i];
    return n;
}
double oracle_residual_poly(int digits, double x) { return residual_poly(pswf_params(digits), x); }

// near-field sum of include/direct_eval.hpp:411-431 for one source against
// n_trg targets (FP64): q_src * sum_t q_t P(2r/b-1)/r over eps < r^2 < b^2
double oracle_direct_eval(const double* r_src, double q_src, int n_trg, const double* r_trg, const double* q_trg,
                          double cutoff, int digits) {
    const auto& p = pswf_params(digits);
    double u = 0;
    for (int t = 0; t < n_trg; ++t) {
        const double dx = r_trg[3 * t] - r_src[0], dy = r_trg[3 * t + 1] - r_src[1], dz = r_trg[3 * t + 2] - r_src[2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > 2.220446049250313e-16 && r2 < cutoff * cutoff) {
            const double rr = std::sqrt(r2);
            u += q_trg[t] * residual_poly(p, rr / cutoff) / rr;
        }
    }
    return u * q_src;
}

// ---- tree ---------------------------------------------------------------
void* oracle_tree_create(int precision, int n_per_leaf, int digits, double L, int n, const double* r,
                         const double* q) {
    Handle* h = nullptr;
    int rc = guard([&] {
        auto H = std::make_unique<Handle>();
        H->precision = precision;
        Params p;
        p.n_per_leaf = n_per_leaf;
        p.digits = digits;
        p.L = L;
        if (precision == 64)
            H->t64 = std::make_unique<DmkTree<double>>(p, n, r, q);
        else if (precision == 32)
            H->t32 = std::make_unique<DmkTree<float>>(p, n, r, q);
        else
            throw std::runtime_error("precision must be 32 or 64");
        h = H.release();
    });
    return rc == 0 ? h : nullptr;
}

void oracle_tree_destroy(void* h) { delete static_cast<Handle*>(h); }

// info: n_nodes, max_depth, n_window, n_diff, dim_window, dim_diff, n_src
int oracle_tree_info(void* h, int* info) {
    return with_tree(h, [&](auto& t) {
        info[0] = t.n_nodes;
        info[1] = t.max_depth;
        info[2] = t.pw.n_window;
        info[3] = t.pw.n_diff;
        info[4] = t.dim_window();
        info[5] = t.dim_diff();
        info[6] = t.n_src;
        return 0;
    });
}

// per node: level, is_leaf, parent, cx, cy, cz, n_particles, n_colleague, n_coarsegrain
int oracle_tree_nodes(void* h, int* out9) {
    return with_tree(h, [&](auto& t) {
        for (int i = 0; i < t.n_nodes; ++i) {
            int* o = out9 + 9 * i;
            o[0] = t.level[i];
            o[1] = t.is_leaf[i];
            o[2] = t.parent[i];
            o[3] = t.cell[3 * i];
            o[4] = t.cell[3 * i + 1];
            o[5] = t.cell[3 * i + 2];
            o[6] = (int)t.node_particles[i].size();
            o[7] = (int)t.colleague[i].size();
            o[8] = (int)t.coarsegrain[i].size();
        }
        return 0;
    });
}

// leaf node id of each ORIGINAL particle index
int oracle_particle_leaf(void* h, int* leaf) {
    return with_tree(h, [&](auto& t) {
        std::vector<int> path;
        for (int i = 0; i < t.n_src; ++i) {
            const long s = t.invmap[i];
            leaf[i] = t.locate(t.r_sorted[3 * s], t.r_sorted[3 * s + 1], t.r_sorted[3 * s + 2], path);
        }
        return 0;
    });
}

// current positions in ORIGINAL index order
int oracle_positions(void* h, double* r) {
    return with_tree(h, [&](auto& t) {
        for (int i = 0; i < t.n_src; ++i)
            for (int d = 0; d < 3; ++d) r[3 * i + d] = (double)t.r_sorted[3 * t.invmap[i] + d];
        return 0;
    });
}

int oracle_form_pw(void* h) {
    return guard([&] {
        with_tree(h, [&](auto& t) {
            t.form_outgoing_pw();
            t.form_incoming_pw();
            return 0;
        });
    });
}

// PROXY outgoing expansions (src/upward_pass.cpp) followed by the incoming ones
int oracle_form_pw_proxy(void* h) {
    return guard([&] {
        with_tree(h, [&](auto& t) {
            t.form_outgoing_pw_proxy();
            t.form_incoming_pw();
            return 0;
        });
    });
}

// Chebyshev coefficients of a node (order^3, a fastest): from_particles = 1
// forms them from the node's own particles, 0 returns those the last PROXY
// build holds (transferred from the children for non-leaf boxes above D-2)
int oracle_proxy_coeffs(void* h, int node, int from_particles, double* out) {
    return with_tree(h, [&](auto& t) {
        const int p = proxy_order(t.params.digits);
        std::vector<double> c;
        if (from_particles)
            c = t.proxy_from_particles(node, p);
        else if (node < (int)t.proxy_coeffs.size())
            c = t.proxy_coeffs[node];
        if (c.empty()) return -1;
        std::copy(c.begin(), c.end(), out);
        return p;
    });
}

// Chebyshev matrices: which 0 V^-1 (p x p), 1 c2p lower half, 2 c2p upper
// half, 3 poly2pw (re/im interleaved, (2n+1) x p); returns the element count
int oracle_cheb_matrix(int which, int p, int n, double* out) {
    try {
        std::vector<double> m;
        if (which == 0) m = lu_inverse(cheb_vandermonde(cheb_nodes(p), p), p);
        if (which == 1 || which == 2) m = cheb_c2p(p, which - 1);
        if (which == 3) {
            const auto P = cheb_poly2pw(p, n);
            for (size_t i = 0; i < P.size(); ++i) {
                out[2 * i] = P[i].real();
                out[2 * i + 1] = P[i].imag();
            }
            return (int)P.size();
        }
        std::copy(m.begin(), m.end(), out);
        return (int)m.size();
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int oracle_proxy_order(int digits) {
    try {
        return proxy_order(digits);
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// which: 0 root outgoing (half grid), 1 node outgoing, 2 node incoming; out: interleaved re/im
int oracle_get_pw(void* h, int which, int node, double* out) {
    return with_tree(h, [&](auto& t) {
        const auto* v = which == 0 ? &t.root_out : (which == 1 ? &t.outgoing[node] : &t.incoming[node]);
        if (v->empty()) return -1;
        for (size_t i = 0; i < v->size(); ++i) {
            out[2 * i] = (double)(*v)[i].real();
            out[2 * i + 1] = (double)(*v)[i].imag();
        }
        return (int)v->size();
    });
}

// single-particle waves of the path of (x,y,z): level l written to out + l*stride (re/im)
int oracle_single_pw(void* h, double x, double y, double z, double q, double* out, int stride, int* path_out) {
    return with_tree(h, [&](auto& t) {
        using R = typename std::decay_t<decltype(t)>::cplx::value_type;
        std::vector<int> path;
        t.locate((R)x, (R)y, (R)z, path);
        std::vector<std::vector<std::complex<R>>> pwv(t.max_depth);
        t.form_outgoing_pw_single(pwv, path, (R)x, (R)y, (R)z, (R)q);
        for (size_t l = 0; l < path.size(); ++l) {
            path_out[l] = path[l];
            for (size_t i = 0; i < pwv[l].size(); ++i) {
                out[l * stride + 2 * i] = (double)pwv[l][i].real();
                out[l * stride + 2 * i + 1] = (double)pwv[l][i].imag();
            }
        }
        return (int)path.size();
    });
}

// out3: window, diff, res
int oracle_eval_energy(void* h, double* out3) {
    return guard([&] {
        with_tree(h, [&](auto& t) {
            out3[0] = t.eval_energy_window();
            out3[1] = t.eval_energy_diff();
            out3[2] = t.eval_energy_res();
            return 0;
        });
    });
}

int oracle_eval_energy_direct(void* h, double* out3) {
    return guard([&] {
        with_tree(h, [&](auto& t) {
            out3[0] = t.eval_energy_window_direct();
            out3[1] = t.eval_energy_diff_direct();
            out3[2] = t.eval_energy_res_direct();
            return 0;
        });
    });
}

int oracle_eval_shift_energy(void* h, long i, double dx, double dy, double dz, double* dE, double* parts5) {
    return guard([&] {
        with_tree(h, [&](auto& t) {
            using R = typename std::decay_t<decltype(t)>::cplx::value_type;
            *dE = t.eval_shift_energy(i, (R)dx, (R)dy, (R)dz, parts5);
            return 0;
        });
    });
}

int oracle_update_shift(void* h, long i, double dx, double dy, double dz) {
    return guard([&] {
        with_tree(h, [&](auto& t) {
            using R = typename std::decay_t<decltype(t)>::cplx::value_type;
            t.update_shift(i, (R)dx, (R)dy, (R)dz);
            return 0;
        });
    });
}

// Metropolis chain: move k proposes idx[k] by d[3k..3k+2]; accepted when
// u[k] < exp(-beta*dE).  Writes dE[k], acc[k]; timing in t2 = {eval_s, update_s}.
int oracle_mc_chain(void* h, int nmoves, const long* idx, const double* d, const double* u, double beta,
                    double* dE, int* acc, double* t2) {
    return guard([&] {
        with_tree(h, [&](auto& t) {
            using R = typename std::decay_t<decltype(t)>::cplx::value_type;
            double te = 0, tu = 0;
            for (int k = 0; k < nmoves; ++k) {
                auto a = std::chrono::steady_clock::now();
                const double e = t.eval_shift_energy(idx[k], (R)d[3 * k], (R)d[3 * k + 1], (R)d[3 * k + 2]);
                auto b = std::chrono::steady_clock::now();
                te += std::chrono::duration<double>(b - a).count();
                dE[k] = e;
                const bool ok = u[k] < std::exp(-beta * e);
                acc[k] = ok;
                if (ok) {
                    auto c = std::chrono::steady_clock::now();
                    t.update_shift(idx[k], (R)d[3 * k], (R)d[3 * k + 1], (R)d[3 * k + 2]);
                    tu += std::chrono::duration<double>(std::chrono::steady_clock::now() - c).count();
                }
            }
            t2[0] = te;
            t2[1] = tu;
            return 0;
        });
    });
}

// Metropolis chain of a charged hard-sphere system: energies scaled by the
// Bjerrum length lB, a target closer than sigma (minimum image, brute force
// over all other particles) to any other particle is an overlap: dE = +inf,
// rejected.  Otherwise accepted when u < exp(-beta lB dE).  This model has no
// reference counterpart; the scaling is the reference Ewald's eps = 1/lB
// (src/ewald.cpp:187).
int oracle_mc_chain_model(void* h, int nmoves, const long* idx, const double* d, const double* u, double beta,
                          double lB, double sigma, double* dE, int* acc) {
    return guard([&] {
        with_tree(h, [&](auto& t) {
            using R = typename std::decay_t<decltype(t)>::cplx::value_type;
            const int n = t.n_src;
            const double L = (double)t.L;
            std::vector<double> r(3 * (size_t)n);
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < 3; ++k) r[3 * i + k] = (double)t.r_sorted[3 * t.invmap[i] + k];
            auto mod = [&](double x) {
                while (x < 0) x += L;
                while (x >= L) x -= L;
                return x;
            };
            for (int m = 0; m < nmoves; ++m) {
                const long p = idx[m];
                double tg[3];
                for (int k = 0; k < 3; ++k) tg[k] = mod(r[3 * p + k] + d[3 * m + k]);
                bool overlap = false;
                for (int j = 0; j < n && sigma > 0 && !overlap; ++j) {
                    if (j == p) continue;
                    double s2 = 0;
                    for (int k = 0; k < 3; ++k) {
                        double x = tg[k] - r[3 * j + k];
                        x -= L * std::rint(x / L);
                        s2 += x * x;
                    }
                    overlap = s2 < sigma * sigma;
                }
                const double e = (double)t.eval_shift_energy(p, (R)d[3 * m], (R)d[3 * m + 1], (R)d[3 * m + 2]);
                dE[m] = overlap ? INFINITY : lB * e;
                const bool ok = !overlap && u[m] < std::exp(-beta * lB * e);
                acc[m] = ok;
                if (ok) {
                    t.update_shift(p, (R)d[3 * m], (R)d[3 * m + 1], (R)d[3 * m + 2]);
                    for (int k = 0; k < 3; ++k) r[3 * p + k] = tg[k];
                }
            }
            return 0;
        });
    });
}

// ---- Ewald ---------------------------------------------------------------
double oracle_ewald_energy(double L, double s, double alpha, double eps, int n, const double* q, const double* r) {
    Ewald e(L, s, alpha, eps, n, q, r);
    return e.compute_energy();
}
double oracle_ewald_potential(double L, double s, double alpha, double eps, int n, const double* q, const double* r,
                              double x, double y, double z) {
    Ewald e(L, s, alpha, eps, n, q, r);
    return e.potential(x, y, z);
}
// Chained Ewald dE for a move list (positions updated after each move when
// apply[k] != 0).  Moves wrap with my_mod (include/utils.hpp:96-105).
int oracle_ewald_shift_chain(double L, double s, double alpha, double eps, int n, const double* q, const double* r,
                             int nmoves, const long* idx, const double* d, const int* apply, double* dE) {
    return guard([&] {
        Ewald e(L, s, alpha, eps, n, q, r);
        for (int k = 0; k < nmoves; ++k) {
            const long p = idx[k];
            const double tx = my_mod(e.r[3 * p] + d[3 * k], L), ty = my_mod(e.r[3 * p + 1] + d[3 * k + 1], L),
                         tz = my_mod(e.r[3 * p + 2] + d[3 * k + 2], L);
            dE[k] = e.shift_energy(p, tx, ty, tz);
            if (apply[k]) e.apply_move(p, tx, ty, tz);
        }
    });
}

// ---- reference test inputs (std::mt19937 default seed, libstdc++ distributions)
// random_init of include/utils.hpp:122-129 (re-seeded on every call)
void oracle_random_init(double* v, long n, double lo, double hi) {
    std::mt19937 g;
    std::uniform_real_distribution<double> dist(lo, hi);
    for (long i = 0; i < n; ++i) v[i] = dist(g);
}
// move sequence of test/hpdmk_test.cpp:409-418 / :486-494: per trial dx,dy,dz ~ U[0,L), i ~ U{0..n-1}
void oracle_test_moves(int n_src, double L, int ntrials, long* idx, double* d) {
    std::mt19937 g;
    std::uniform_real_distribution<double> dist(0, L);
    std::uniform_int_distribution<int> di(0, n_src - 1);
    for (int k = 0; k < ntrials; ++k) {
        d[3 * k] = dist(g);
        d[3 * k + 1] = dist(g);
        d[3 * k + 2] = dist(g);
        idx[k] = di(g);
    }
}
}
