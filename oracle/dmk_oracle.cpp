// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle/dmk_oracle.hpp).
//
// Octree, plane-wave expansions, total energy, trial-move dE and
// accept/update of the reference's HPDMKPtTree, restated in plain C++.
//
// Deliberate, documented differences from the reference binary:
//  * the tree is built here with SCTL PtTree semantics (refine while a box
//    holds more than n_per_leaf points, periodic 2:1 balance over the 26
//    neighbours, Morton pre-order).  SCTL's source is absent, so topology
//    identity is unpinned.  Levels 0 and 1 are always fully refined when the
//    tree has at least three levels: a leaf at level 1 would be skipped by the
//    residual loops of src/energy.cpp:88 and src/energy_direct.cpp:159.
//  * the root expansion is the exact sum over all particles (the reference
//    calls FINUFFT at eps = 10^-(digits+1), src/planewave.cpp:45).
//  * particle membership is kept exact on every level after an accepted move
//    (the reference refreshes only the two lowest path levels,
//    src/tree.cpp:584-596, leaving coarser lists stale).
//  * the near-field 1/r uses a correctly rounded sqrt instead of SCTL's
//    approx_rsqrt<digits> (include/direct_eval.hpp:370).
#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <numeric>

#include "dmk_oracle.hpp"

namespace oracle {

namespace {
constexpr int kMaxLevel = 30;

inline int cell_of(double x, double L, int l) {
    const int n = 1 << l;
    long c = (long)std::floor(x / L * (double)n);
    if (c < 0) c = 0;
    if (c > n - 1) c = n - 1;
    return (int)c;
}

inline uint64_t spread_bits(uint64_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ull) << (3 * b);
    return r;
}

struct BNode {
    int level, cx, cy, cz, parent;
    int child0 = -1;  // children are child0..child0+7
    std::vector<int> pts;
};
}  // namespace

namespace {
// s_p[k] = exp(-i k . p L) on the half grid (include/shiftmat.hpp:12-43)
template <typename Real>
std::vector<std::complex<Real>> shift_vec(int nk, double dk, double L, int imgcode) {
    const int px = imgcode % 3 - 1, py = (imgcode / 3) % 3 - 1, pz = imgcode / 9 - 1;
    const int d = 2 * nk + 1;
    std::vector<std::complex<Real>> s(d * d * (nk + 1));
    for (int k = 0; k <= nk; ++k)
        for (int j = 0; j < d; ++j)
            for (int i = 0; i < d; ++i) {
                const double th = ((i - nk) * px + (j - nk) * py + (k - nk) * pz) * dk * L;
                s[(k * d + j) * d + i] = std::complex<Real>((Real)std::cos(th), (Real)-std::sin(th));
            }
    return s;
}
}  // namespace

template <typename Real>
int DmkTree<Real>::find_node(int l, int cx, int cy, int cz) const {
    int nd = 0;
    while (level[nd] < l) {
        if (is_leaf[nd]) return -1;
        const int sh = l - level[nd] - 1;
        const int bits = ((cx >> sh) & 1) | (((cy >> sh) & 1) << 1) | (((cz >> sh) & 1) << 2);
        nd = child[8 * nd + bits];
    }
    return nd;
}

// src/tree.cpp:268-444 (constructor) + SCTL UpdateRefinement semantics.
template <typename Real>
DmkTree<Real>::DmkTree(const Params& p, int n, const double* r, const double* q)
    : params(p), pw(pswf_params(p.digits)), L((Real)p.L), n_src(n) {
    if (std::is_same<Real, float>::value && p.digits != 3 && p.digits != 6)
        throw std::runtime_error("float only supports 3 and 6 digits accuracy");  // src/tree.cpp:19-23
    const double Ld = p.L;
    // ---- refinement (SCTL: refine while count > n_per_leaf)
    std::vector<BNode> bn;
    bn.push_back(BNode{0, 0, 0, 0, -1, -1, {}});
    bn[0].pts.resize(n);
    std::iota(bn[0].pts.begin(), bn[0].pts.end(), 0);
    auto split = [&](int nd) {
        const int c0 = (int)bn.size();
        const BNode par = bn[nd];
        for (int c = 0; c < 8; ++c)
            bn.push_back(BNode{par.level + 1, 2 * par.cx + (c & 1), 2 * par.cy + ((c >> 1) & 1),
                               2 * par.cz + ((c >> 2) & 1), nd, -1, {}});
        for (int id : bn[nd].pts) {
            const int l1 = par.level + 1;
            const int bx = cell_of(r[3 * id], Ld, l1) & 1;
            const int by = cell_of(r[3 * id + 1], Ld, l1) & 1;
            const int bz = cell_of(r[3 * id + 2], Ld, l1) & 1;
            bn[c0 + (bx | (by << 1) | (bz << 2))].pts.push_back(id);
        }
        bn[nd].pts.clear();
        bn[nd].pts.shrink_to_fit();
        bn[nd].child0 = c0;
    };
    {
        std::vector<int> stack{0};
        while (!stack.empty()) {
            int nd = stack.back();
            stack.pop_back();
            if ((int)bn[nd].pts.size() > p.n_per_leaf && bn[nd].level < kMaxLevel) {
                split(nd);
                for (int c = 0; c < 8; ++c) stack.push_back(bn[nd].child0 + c);
            }
        }
    }
    int maxl = 0;
    for (auto& b : bn) maxl = std::max(maxl, b.level);
    if (maxl + 1 <= 2) throw std::runtime_error("max depth is less than 2, too low for hpdmk");  // src/tree.cpp:352-353
    for (int c = 0; c < 8; ++c)
        if (bn[bn[0].child0 + c].child0 < 0) split(bn[0].child0 + c);
    // ---- periodic 2:1 balance: every non-leaf node has all 26 same-level neighbours
    auto covering = [&](int l, int cx, int cy, int cz) {
        int nd = 0;
        while (bn[nd].level < l && bn[nd].child0 >= 0) {
            const int sh = l - bn[nd].level - 1;
            const int bits = ((cx >> sh) & 1) | (((cy >> sh) & 1) << 1) | (((cz >> sh) & 1) << 2);
            nd = bn[nd].child0 + bits;
        }
        return nd;
    };
    for (bool changed = true; changed;) {
        changed = false;
        for (size_t nd = 0; nd < bn.size(); ++nd) {
            if (bn[nd].child0 < 0 || bn[nd].level < 1) continue;
            const int l = bn[nd].level, nl = 1 << l;
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (!dx && !dy && !dz) continue;
                        const int cx = (bn[nd].cx + dx + nl) % nl, cy = (bn[nd].cy + dy + nl) % nl,
                                  cz = (bn[nd].cz + dz + nl) % nl;
                        const int cov = covering(l, cx, cy, cz);
                        if (bn[cov].level < l) {
                            split(cov);
                            changed = true;
                        }
                    }
        }
    }
    // ---- Morton pre-order numbering
    std::vector<int> order;
    std::function<void(int)> visit = [&](int nd) {
        order.push_back(nd);
        if (bn[nd].child0 >= 0)
            for (int c = 0; c < 8; ++c) visit(bn[nd].child0 + c);
    };
    visit(0);
    n_nodes = (int)order.size();
    std::vector<int> newid(bn.size());
    for (int i = 0; i < n_nodes; ++i) newid[order[i]] = i;
    level.resize(n_nodes);
    cell.resize(3 * n_nodes);
    parent.resize(n_nodes);
    child.assign(8 * n_nodes, -1);
    is_leaf.resize(n_nodes);
    max_depth = 0;
    for (int i = 0; i < n_nodes; ++i) {
        const BNode& b = bn[order[i]];
        level[i] = b.level;
        cell[3 * i] = b.cx;
        cell[3 * i + 1] = b.cy;
        cell[3 * i + 2] = b.cz;
        parent[i] = b.parent < 0 ? -1 : newid[b.parent];
        is_leaf[i] = b.child0 < 0;
        if (b.child0 >= 0)
            for (int c = 0; c < 8; ++c) child[8 * i + c] = newid[b.child0 + c];
        max_depth = std::max(max_depth, b.level + 1);
    }
    level_indices.assign(max_depth, {});
    for (int i = 0; i < n_nodes; ++i) level_indices[level[i]].push_back(i);
    boxsize.resize(max_depth + 1);
    boxsize[0] = Ld;
    for (int i = 1; i <= max_depth; ++i) boxsize[i] = 0.5 * boxsize[i - 1];
    center.resize(3 * n_nodes);
    for (int i = 0; i < n_nodes; ++i)
        for (int d = 0; d < 3; ++d) center[3 * i + d] = Ld * ((cell[3 * i + d] + 0.5) / (double)(1 << level[i]));

    // ---- particles in leaf pre-order, Morton order inside a leaf
    r_sorted.resize(3 * n);
    q_sorted.resize(n);
    invmap.resize(n);
    node_particles.assign(n_nodes, {});
    std::vector<int> first(n_nodes, 0), last(n_nodes, 0);
    long pos = 0;
    for (int i = 0; i < n_nodes; ++i) {
        if (!is_leaf[i]) continue;
        std::vector<int> pts = bn[order[i]].pts;
        std::vector<std::pair<uint64_t, int>> keyed;
        for (int id : pts) {
            uint64_t k = spread_bits(cell_of(r[3 * id], Ld, 21)) | (spread_bits(cell_of(r[3 * id + 1], Ld, 21)) << 1) |
                         (spread_bits(cell_of(r[3 * id + 2], Ld, 21)) << 2);
            keyed.emplace_back(k, id);
        }
        std::sort(keyed.begin(), keyed.end());
        for (auto& kv : keyed) {
            const int id = kv.second;
            for (int d = 0; d < 3; ++d) r_sorted[3 * pos + d] = (Real)r[3 * id + d];
            q_sorted[pos] = (Real)q[id];
            invmap[id] = pos;
            for (int a = i; a >= 0; a = parent[a]) node_particles[a].push_back(pos);
            ++pos;
        }
    }
    double Qd = 0;
    for (int i = 0; i < n; ++i) Qd += q[i] * q[i];  // src/tree.cpp:306-309
    Q = (Real)Qd;

    // ---- wavenumbers (src/tree.cpp:17-81)
    sigmas.resize(max_depth + 1);
    sigmas[0] = Ld / pw.c;
    for (int i = 1; i <= max_depth; ++i) sigmas[i] = 0.5 * sigmas[i - 1];
    delta_k.resize(max_depth + 1);
    delta_k[0] = delta_k[1] = 2 * M_PI / Ld;
    for (int i = 2; i <= max_depth; ++i) delta_k[i] = (2.0 * M_PI / 3.0) / boxsize[i];

    // ---- interaction matrices (src/tree.cpp:84-100, include/kernels.hpp:16-113)
    const Prolate0& f = prolate0(pw.c);
    auto fourier = [&](double x) { return 2 * M_PI * pw.lambda * f.eval(x) / pw.C0; };  // include/pswf.hpp:84-86
    {
        const int nw = pw.n_window, d = 2 * nw + 1;
        window_mat.assign(dim_window(), 0);
        for (int k = 0; k < nw + 1; ++k)
            for (int j = 0; j < d; ++j)
                for (int i = 0; i < d; ++i) {
                    const double kx = (i - nw) * delta_k[1], ky = (j - nw) * delta_k[1], kz = (k - nw) * delta_k[1];
                    const double k2 = kx * kx + ky * ky + kz * kz;
                    double v = (k2 == 0) ? 0.0 : fourier(std::sqrt(k2) * sigmas[2]) / k2;
                    window_mat[k * d * d + j * d + i] = (Real)((k == nw) ? v : 2 * v);
                }
    }
    diff_mat.assign(max_depth, {});
    for (int l = 2; l < max_depth - 1; ++l) {
        const int nd = pw.n_diff, d = 2 * nd + 1;
        auto& D = diff_mat[l];
        D.assign(dim_diff(), 0);
        for (int k = 0; k < nd + 1; ++k)
            for (int j = 0; j < d; ++j)
                for (int i = 0; i < d; ++i) {
                    const double kx = (i - nd) * delta_k[l], ky = (j - nd) * delta_k[l], kz = (k - nd) * delta_k[l];
                    const double k2 = kx * kx + ky * ky + kz * kz;
                    double v;
                    if (k2 == 0)
                        v = pw.diff0 * (sigmas[l + 1] * sigmas[l + 1] - sigmas[l] * sigmas[l]);
                    else
                        v = (fourier(std::sqrt(k2) * sigmas[l + 1]) - fourier(std::sqrt(k2) * sigmas[l])) / k2;
                    D[k * d * d + j * d + i] = (Real)((k == nd) ? v : 2 * v);
                }
    }

    // ---- neighbours (src/tree.cpp:136-165, 405-410)
    colleague.assign(n_nodes, {});
    coarsegrain.assign(n_nodes, {});
    colleague_img.assign(n_nodes, {});
    coarsegrain_img.assign(n_nodes, {});
    for (int l = 2; l < max_depth; ++l)
        for (int nd : level_indices[l]) {
            const int nl = 1 << l;
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (!dx && !dy && !dz) continue;
                        int c[3] = {cell[3 * nd] + dx, cell[3 * nd + 1] + dy, cell[3 * nd + 2] + dz};
                        int img[3];
                        for (int a = 0; a < 3; ++a) {
                            img[a] = c[a] < 0 ? -1 : (c[a] >= nl ? 1 : 0);
                            c[a] = (c[a] + nl) % nl;
                        }
                        const int nb = find_node(l, c[0], c[1], c[2]);
                        if (nb >= 0) {
                            colleague[nd].push_back(nb);
                            colleague_img[nd].push_back((int8_t)((img[0] + 1) + 3 * (img[1] + 1) + 9 * (img[2] + 1)));
                        }
                    }
            if (colleague[nd].size() < 26) {
                const int pa = parent[nd], pl = l - 1, npl = 1 << pl;
                int s[3];
                for (int a = 0; a < 3; ++a) s[a] = (cell[3 * nd + a] & 1) ? 1 : -1;
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            if (!dx && !dy && !dz) continue;
                            if (dx * s[0] < 0 || dy * s[1] < 0 || dz * s[2] < 0) continue;  // src/tree.cpp:159
                            int c[3] = {cell[3 * pa] + dx, cell[3 * pa + 1] + dy, cell[3 * pa + 2] + dz};
                            int img[3];
                            for (int a = 0; a < 3; ++a) {
                                img[a] = c[a] < 0 ? -1 : (c[a] >= npl ? 1 : 0);
                                c[a] = (c[a] + npl) % npl;
                            }
                            const int nb = find_node(pl, c[0], c[1], c[2]);
                            if (nb >= 0 && is_leaf[nb]) {
                                coarsegrain[nd].push_back(nb);
                                coarsegrain_img[nd].push_back(
                                    (int8_t)((img[0] + 1) + 3 * (img[1] + 1) + 9 * (img[2] + 1)));
                            }
                        }
            }
        }

    // ---- expansion storage (src/tree.cpp:413-440)
    outgoing.assign(n_nodes, {});
    incoming.assign(n_nodes, {});
    root_out.assign(dim_window(), cplx(0));
    for (int l = 2; l < max_depth - 1; ++l)
        for (int nd : level_indices[l]) {
            outgoing[nd].assign(dim_diff(), cplx(0));
            incoming[nd].assign(dim_diff(), cplx(0));
        }
    pw_o.assign(max_depth, {});
    pw_t.assign(max_depth, {});
    shift_mat.assign(max_depth, {});
    for (int l = 2; l < max_depth - 1; ++l) {
        shift_mat[l].resize(27);
        for (int c = 0; c < 27; ++c) shift_mat[l][c] = shift_vec<Real>(pw.n_diff, delta_k[l], params.L, c);
    }
}

template <typename Real>
int DmkTree<Real>::locate(Real x, Real y, Real z, std::vector<int>& path) const {
    path.clear();
    int nd = 0;
    path.push_back(0);
    while (!is_leaf[nd]) {
        const int l1 = level[nd] + 1;
        const int bits = (cell_of(x, params.L, l1) & 1) | ((cell_of(y, params.L, l1) & 1) << 1) |
                         ((cell_of(z, params.L, l1) & 1) << 2);
        nd = child[8 * nd + bits];
        path.push_back(nd);
    }
    return nd;
}

namespace {
// phase table e^{-i (m - n) dk x}, m = 0..len-1
template <typename Real>
void phase_row(std::complex<Real>* out, int n, int len, double dk, double x) {
    for (int m = 0; m < len; ++m) {
        const double th = (double)(m - n) * (dk * x);
        out[m] = std::complex<Real>((Real)std::cos(th), (Real)-std::sin(th));
    }
}

// f[kx,ky,kz<=0] (+)= sum_j c_j e^{-i k.x_j} on the half grid (include/nudft.hpp:66-123)
template <typename Real>
void nudft_halfplane(int nk, double dk, const std::vector<long>& idx, const std::vector<Real>& r,
                     const std::vector<Real>& q, std::complex<Real>* f) {
    using cplx = std::complex<Real>;
    const int d = 2 * nk + 1;
    std::vector<cplx> X(d), Y(d), Z(nk + 1);
    for (long s : idx) {
        phase_row(X.data(), nk, d, dk, (double)r[3 * s]);
        phase_row(Y.data(), nk, d, dk, (double)r[3 * s + 1]);
        phase_row(Z.data(), nk, nk + 1, dk, (double)r[3 * s + 2]);
        for (int i = 0; i < d; ++i) X[i] *= q[s];
        for (int k = 0; k <= nk; ++k)
            for (int j = 0; j < d; ++j) {
                const cplx zy = Z[k] * Y[j];
                cplx* fp = f + (k * d + j) * d;
                for (int i = 0; i < d; ++i) fp[i] += zy * X[i];
            }
    }
}
}  // namespace

template <typename Real>
void DmkTree<Real>::form_outgoing_pw_root() {
    std::fill(root_out.begin(), root_out.end(), cplx(0));
    std::vector<long> all(n_src);
    std::iota(all.begin(), all.end(), 0L);
    nudft_halfplane<Real>(pw.n_window, delta_k[0], all, r_sorted, q_sorted, root_out.data());
}

template <typename Real>
void DmkTree<Real>::form_outgoing_pw() {
    form_outgoing_pw_root();
    for (int l = 2; l < max_depth - 1; ++l)
        for (int nd : level_indices[l]) {
            std::fill(outgoing[nd].begin(), outgoing[nd].end(), cplx(0));
            if (!node_particles[nd].empty())
                nudft_halfplane<Real>(pw.n_diff, delta_k[l], node_particles[nd], r_sorted, q_sorted,
                                      outgoing[nd].data());
        }
}


template <typename Real>
void DmkTree<Real>::form_incoming_pw() {
    const int dims = dim_diff();
    for (int l = 2; l < max_depth - 1; ++l) {
        const auto& sv = shift_mat[l];
        for (int nd : level_indices[l]) {
            if (is_leaf[nd]) continue;
            auto& in = incoming[nd];
            if (!node_particles[nd].empty())
                for (int i = 0; i < dims; ++i) in[i] = std::conj(outgoing[nd][i]);
            else
                std::fill(in.begin(), in.end(), cplx(0));
            if (colleague[nd].size() != 26) throw std::runtime_error("non-leaf node without 26 colleagues");
            for (size_t c = 0; c < colleague[nd].size(); ++c) {
                const int nb = colleague[nd][c];
                if (node_particles[nb].empty()) continue;
                const auto& s = sv[colleague_img[nd][c]];
                const auto& o = outgoing[nb];
                for (int i = 0; i < dims; ++i) in[i] += std::conj(o[i] * s[i]);
            }
        }
    }
}

template <typename Real>
double DmkTree<Real>::eval_energy_window() {
    double e = 0;
    for (int i = 0; i < dim_window(); ++i) e += (double)std::norm(root_out[i]) * (double)window_mat[i];
    e *= 1.0 / (2 * std::pow(2 * M_PI, 3)) * std::pow(delta_k[0], 3);
    e -= (double)Q * pw.psi0_0 / (2 * boxsize[2] * pw.C0);
    return e;
}

template <typename Real>
double DmkTree<Real>::eval_energy_diff() {
    double e = 0;
    for (int l = 2; l < max_depth - 1; ++l) {
        const double Cl = 1.0 / (2 * std::pow(2 * M_PI, 3)) * std::pow(delta_k[l], 3);
        const double Sl = pw.psi0_0 * (1.0 / (2 * boxsize[l + 1] * pw.C0) - 1.0 / (2 * boxsize[l] * pw.C0));
        for (int nd : level_indices[l]) {
            if (is_leaf[nd] || node_particles[nd].empty()) continue;
            double acc = 0;
            const auto& o = outgoing[nd];
            const auto& in = incoming[nd];
            for (int i = 0; i < dim_diff(); ++i)
                acc += ((double)o[i].real() * in[i].real() - (double)o[i].imag() * in[i].imag()) * diff_mat[l][i];
            double Qi = 0;
            for (long s : node_particles[nd]) Qi += (double)q_sorted[s] * q_sorted[s];
            e += Cl * acc - Qi * Sl;
        }
    }
    return e;
}

namespace {
inline void img_vec(int code, int* p) {
    p[0] = code % 3 - 1;
    p[1] = (code / 3) % 3 - 1;
    p[2] = code / 9 - 1;
}
}  // namespace

template <typename Real>
double DmkTree<Real>::eval_energy_res() {
    double e = 0;
    for (int l = 2; l < max_depth; ++l) {
        const double b = boxsize[l];
        for (int nd : level_indices[l]) {
            if (!is_leaf[nd] || node_particles[nd].empty()) continue;
            const auto& P = node_particles[nd];
            for (size_t a = 0; a + 1 < P.size(); ++a)  // src/energy.cpp:116-140
                for (size_t c = a + 1; c < P.size(); ++c) {
                    const double dx = (double)r_sorted[3 * P[a]] - r_sorted[3 * P[c]];
                    const double dy = (double)r_sorted[3 * P[a] + 1] - r_sorted[3 * P[c] + 1];
                    const double dz = (double)r_sorted[3 * P[a] + 2] - r_sorted[3 * P[c] + 2];
                    const double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
                    if (rr <= b && rr > 0)
                        e += (double)q_sorted[P[a]] * q_sorted[P[c]] * residual_poly(pw, rr / b) / rr;
                }
            auto pair_sum = [&](int nb, int code) {  // src/energy.cpp:143-173
                int p[3];
                img_vec(code, p);
                for (long i : P)
                    for (long j : node_particles[nb]) {
                        const double dx = (double)r_sorted[3 * i] - (r_sorted[3 * j] + p[0] * params.L);
                        const double dy = (double)r_sorted[3 * i + 1] - (r_sorted[3 * j + 1] + p[1] * params.L);
                        const double dz = (double)r_sorted[3 * i + 2] - (r_sorted[3 * j + 2] + p[2] * params.L);
                        const double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
                        if (rr <= b && rr > 0)
                            e += (double)q_sorted[i] * q_sorted[j] * residual_poly(pw, rr / b) / rr / 2;
                    }
            };
            for (size_t c = 0; c < coarsegrain[nd].size(); ++c)
                if (!node_particles[coarsegrain[nd][c]].empty()) pair_sum(coarsegrain[nd][c], coarsegrain_img[nd][c]);
            for (size_t c = 0; c < colleague[nd].size(); ++c)
                if (!node_particles[colleague[nd][c]].empty()) pair_sum(colleague[nd][c], colleague_img[nd][c]);
        }
    }
    return e;
}

template <typename Real>
double DmkTree<Real>::eval_energy_window_direct() {
    const int nw = pw.n_window, d = 2 * nw + 1;
    const Prolate0& f = prolate0(pw.c);
    double e = 0;
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j)
            for (int k = 0; k < d; ++k) {
                const double kx = (i - nw) * delta_k[0], ky = (j - nw) * delta_k[0], kz = (k - nw) * delta_k[0];
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 == 0) continue;
                std::complex<double> rho = 0;
                for (int s = 0; s < n_src; ++s)
                    rho += std::polar((double)q_sorted[s],
                                      kx * r_sorted[3 * s] + ky * r_sorted[3 * s + 1] + kz * r_sorted[3 * s + 2]);
                const double w = 2 * M_PI * pw.lambda * f.eval(std::sqrt(k2) * sigmas[2]) / pw.C0 / k2;
                e += std::norm(rho) * w;
            }
    e *= 1.0 / (2 * std::pow(2 * M_PI, 3)) * std::pow(delta_k[0], 3);
    e -= (double)Q * pw.psi0_0 / (2 * boxsize[2] * pw.C0);
    return e;
}

template <typename Real>
double DmkTree<Real>::eval_energy_diff_direct() {
    double e = 0;
    for (int l = 2; l < max_depth; ++l)
        for (int nd : level_indices[l]) {
            if (is_leaf[nd] || node_particles[nd].empty()) continue;
            for (long i : node_particles[nd])
                for (long j = 0; j < n_src; ++j) {
                    if (j == i) continue;
                    for (int mx = -1; mx <= 1; ++mx)
                        for (int my = -1; my <= 1; ++my)
                            for (int mz = -1; mz <= 1; ++mz) {
                                const double dx = (double)r_sorted[3 * i] - (r_sorted[3 * j] + mx * params.L);
                                const double dy = (double)r_sorted[3 * i + 1] - (r_sorted[3 * j + 1] + my * params.L);
                                const double dz = (double)r_sorted[3 * i + 2] - (r_sorted[3 * j + 2] + mz * params.L);
                                const double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
                                if (rr == 0) continue;
                                const double dk = -(residual_poly(pw, rr / boxsize[l + 1]) -
                                                    residual_poly(pw, rr / boxsize[l])) / rr;
                                e += (double)q_sorted[i] * q_sorted[j] * dk / 2;
                            }
                }
        }
    return e;
}

template <typename Real>
double DmkTree<Real>::eval_energy_res_direct() {
    double e = 0;
    for (int l = 2; l < max_depth; ++l)
        for (int nd : level_indices[l]) {
            if (!is_leaf[nd] || node_particles[nd].empty()) continue;
            for (long i : node_particles[nd])
                for (long j = 0; j < n_src; ++j) {
                    if (j == i) continue;
                    for (int mx = -1; mx <= 1; ++mx)
                        for (int my = -1; my <= 1; ++my)
                            for (int mz = -1; mz <= 1; ++mz) {
                                const double dx = (double)r_sorted[3 * i] - (r_sorted[3 * j] + mx * params.L);
                                const double dy = (double)r_sorted[3 * i + 1] - (r_sorted[3 * j + 1] + my * params.L);
                                const double dz = (double)r_sorted[3 * i + 2] - (r_sorted[3 * j + 2] + mz * params.L);
                                const double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
                                if (rr == 0) continue;
                                e += (double)q_sorted[i] * q_sorted[j] * residual_poly(pw, rr / boxsize[l]) / rr / 2;
                            }
                }
        }
    return e;
}

template <typename Real>
void DmkTree<Real>::form_outgoing_pw_single(std::vector<std::vector<cplx>>& pwv, const std::vector<int>& path,
                                            Real x, Real y, Real z, Real q) {
    for (int i = 0; i < (int)path.size(); ++i) {
        int nk;
        double dk;
        if (i == 0) {
            nk = pw.n_window;
            dk = delta_k[0];
        } else if (i == 1 || i == max_depth - 1) {
            continue;
        } else {
            nk = pw.n_diff;
            dk = delta_k[i];
        }
        const int d = 2 * nk + 1;
        pwv[i].assign(d * d * (nk + 1), cplx(0));
        std::vector<long> one{0};
        std::vector<Real> rr{x, y, z}, qq{q};
        nudft_halfplane<Real>(nk, dk, one, rr, qq, pwv[i].data());
    }
}

template <typename Real>
double DmkTree<Real>::res_near(long s, const std::vector<int>& path, Real x, Real y, Real z) const {
    // src/energy_shift.cpp:233-302 + include/direct_eval.hpp:309-403
    const int nd = path.back();
    const double b = boxsize[level[nd]];
    const double eps = std::numeric_limits<Real>::epsilon();
    double u = 0;
    auto add = [&](int nb, int code) {
        int p[3];
        img_vec(code, p);
        for (long j : node_particles[nb]) {
            const double dx = (double)x - ((double)r_sorted[3 * j] + p[0] * params.L);
            const double dy = (double)y - ((double)r_sorted[3 * j + 1] + p[1] * params.L);
            const double dz = (double)z - ((double)r_sorted[3 * j + 2] + p[2] * params.L);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > eps && r2 < b * b) {
                const double rr = std::sqrt(r2);
                u += (double)q_sorted[j] * residual_poly(pw, rr / b) / rr;
            }
        }
    };
    add(nd, 13);
    for (size_t c = 0; c < colleague[nd].size(); ++c)
        if (!node_particles[colleague[nd][c]].empty()) add(colleague[nd][c], colleague_img[nd][c]);
    for (size_t c = 0; c < coarsegrain[nd].size(); ++c)
        if (!node_particles[coarsegrain[nd][c]].empty()) add(coarsegrain[nd][c], coarsegrain_img[nd][c]);
    return (double)q_sorted[s] * u;
}

template <typename Real>
double DmkTree<Real>::eval_shift_energy(long i_unsorted, Real dx, Real dy, Real dz, double* parts) {
    // src/energy_shift.cpp:18-70
    const long s = invmap[i_unsorted];
    const Real q = q_sorted[s];
    const Real xo = r_sorted[3 * s], yo = r_sorted[3 * s + 1], zo = r_sorted[3 * s + 2];
    const Real xt = my_mod<Real>(xo + dx, L), yt = my_mod<Real>(yo + dy, L), zt = my_mod<Real>(zo + dz, L);
    double dr = std::sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
    for (int mx = -1; mx <= 1; ++mx)
        for (int my = -1; my <= 1; ++my)
            for (int mz = -1; mz <= 1; ++mz) {
                const double ax = (double)xt - ((double)xo + mx * params.L);
                const double ay = (double)yt - ((double)yo + my * params.L);
                const double az = (double)zt - ((double)zo + mz * params.L);
                dr = std::min(dr, std::sqrt(ax * ax + ay * ay + az * az));
            }
    locate(xt, yt, zt, path_t);
    locate(xo, yo, zo, path_o);
    form_outgoing_pw_single(pw_o, path_o, xo, yo, zo, q);
    form_outgoing_pw_single(pw_t, path_t, xt, yt, zt, q);
    last_i = i_unsorted;
    last_dx = dx;
    last_dy = dy;
    last_dz = dz;

    // window (src/energy_shift.cpp:72-90, include/vecops.hpp:118-145)
    double dEw = 0;
    for (int i = 0; i < dim_window(); ++i) {
        const auto a = pw_t[0][i] - pw_o[0][i];
        const auto g = root_out[i] - pw_o[0][i];
        dEw += ((double)a.real() * g.real() + (double)a.imag() * g.imag()) * window_mat[i];
    }
    dEw = 2.0 * dEw * (1.0 / (2 * std::pow(2 * M_PI, 3)) * std::pow(delta_k[0], 3));

    // difference (src/energy_shift.cpp:92-140, include/vecops.hpp:148-166 and :27-47)
    double dEo = 0, dEt = 0;
    for (int l = 2; l < (int)path_o.size() - 1; ++l) {
        const auto& om = pw_o[l];
        const auto& in = incoming[path_o[l]];
        double acc = 0;
        for (int i = 0; i < dim_diff(); ++i) {
            const double re = (double)om[i].real() * in[i].real() - (double)om[i].imag() * in[i].imag() -
                              ((double)om[i].real() * om[i].real() + (double)om[i].imag() * om[i].imag());
            acc += re * diff_mat[l][i];
        }
        dEo += acc / std::pow(2 * M_PI, 3) * std::pow(delta_k[l], 3);
    }
    for (int l = 2; l < (int)path_t.size() - 1; ++l) {
        const auto& ta = pw_t[l];
        const auto& in = incoming[path_t[l]];
        double acc = 0;
        for (int i = 0; i < dim_diff(); ++i)
            acc += ((double)ta[i].real() * in[i].real() - (double)ta[i].imag() * in[i].imag()) * diff_mat[l][i];
        dEt += acc / std::pow(2 * M_PI, 3) * std::pow(delta_k[l], 3);
    }
    const double dEd = dEt - dEo;
    const double Rt = res_near(s, path_t, xt, yt, zt);
    const double Ro = res_near(s, path_o, xo, yo, zo);
    const double self = (dr == 0 ? 0.0 : residual_poly(pw, dr / boxsize[2]) / dr) * (double)q * q;
    if (parts) {
        parts[0] = dEw;
        parts[1] = dEd;
        parts[2] = Rt;
        parts[3] = Ro;
        parts[4] = self;
    }
    return dEw + dEd + Rt - Ro - self;
}

template <typename Real>
void DmkTree<Real>::update_shift(long i_unsorted, Real dx, Real dy, Real dz) {
    // src/tree.cpp:447-597.  The reference reuses the paths and single-particle
    // waves of the preceding eval_shift_energy call with the same arguments;
    // they are recomputed here, which gives the same values.
    const long s = invmap[i_unsorted];
    const Real q = q_sorted[s];
    const Real xo = r_sorted[3 * s], yo = r_sorted[3 * s + 1], zo = r_sorted[3 * s + 2];
    const Real xt = my_mod<Real>(xo + dx, L), yt = my_mod<Real>(yo + dy, L), zt = my_mod<Real>(zo + dz, L);
    locate(xt, yt, zt, path_t);
    locate(xo, yo, zo, path_o);
    form_outgoing_pw_single(pw_o, path_o, xo, yo, zo, q);
    form_outgoing_pw_single(pw_t, path_t, xt, yt, zt, q);
    r_sorted[3 * s] = xt;
    r_sorted[3 * s + 1] = yt;
    r_sorted[3 * s + 2] = zt;
    for (int i = 0; i < dim_window(); ++i) {
        root_out[i] += pw_t[0][i];
        root_out[i] -= pw_o[0][i];
    }
    const int dims = dim_diff();
    auto apply = [&](const std::vector<int>& path, const std::vector<std::vector<cplx>>& pwv, Real sign) {
        for (int l = 2; l < (int)path.size(); ++l) {
            if (l == max_depth - 1) continue;
            const int nd = path[l];
            const auto& t = pwv[l];
            for (int i = 0; i < dims; ++i) {
                outgoing[nd][i] += sign * t[i];
                incoming[nd][i] += sign * std::conj(t[i]);
            }
            for (size_t c = 0; c < colleague[nd].size(); ++c) {
                const int nb = colleague[nd][c];
                if (is_leaf[nb]) continue;
                // image of nd seen from nb is the opposite of nb's image seen from nd
                const int code = 26 - colleague_img[nd][c];
                if (code == 13) {
                    for (int i = 0; i < dims; ++i) incoming[nb][i] += sign * std::conj(t[i]);
                } else {
                    const auto& sv = shift_mat[l][code];
                    for (int i = 0; i < dims; ++i) incoming[nb][i] += sign * std::conj(t[i] * sv[i]);
                }
            }
        }
    };
    apply(path_t, pw_t, Real(1));
    apply(path_o, pw_o, Real(-1));
    for (int nd : path_o) {
        auto& P = node_particles[nd];
        auto it = std::find(P.begin(), P.end(), s);
        if (it != P.end()) P.erase(it);
    }
    for (int nd : path_t) node_particles[nd].push_back(s);
}

template struct DmkTree<double>;
template struct DmkTree<float>;

}  // namespace oracle
