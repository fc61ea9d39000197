// Host-side set-up of a DMK system (see dmk_host.hpp).
#include "dmk_host.hpp"

#include "../../include/hpdmk.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <unordered_map>

namespace hpdmk_amd {

// ------------------------------------------------------------------ PSWF
// psi_0^c is the lowest eigenfunction of A = -d/dx (1-x^2) d/dx + c^2 x^2.  In
// the normalised even Legendre basis sqrt(k+1/2) P_k the even block of A is a
// symmetric tridiagonal matrix (the system the reference sets up in
// src/pswf.cpp:610-676).  Here it is diagonalised by cyclic Jacobi rotations,
// and every constant the method needs follows in closed form from the Legendre
// coefficients a_k:
//   C0      = int_0^1 psi0           = a_0
//   int_0^1 x^2 psi0                 = a_0/3 + 2 a_2/15
//   lambda  = int_{-1}^1 psi0 / psi0(0) = 2 a_0 / psi0(0)   (integral equation at x = 0)
namespace {
std::vector<double> prolate_legendre_coefs(double c) {
    const int m = 48;
    std::vector<double> A(m * m, 0.0), V(m * m, 0.0);
    const double c2 = c * c;
    for (int i = 0; i < m; ++i) {
        const double k = 2.0 * i;
        A[i * m + i] = k * (k + 1) + c2 * (2 * k * k + 2 * k - 1) / ((2 * k - 1) * (2 * k + 3));
        if (i + 1 < m) {
            const double off = c2 * (k + 1) * (k + 2) / ((2 * k + 3) * std::sqrt((2 * k + 1) * (2 * k + 5)));
            A[i * m + i + 1] = A[(i + 1) * m + i] = off;
        }
        V[i * m + i] = 1.0;
    }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int p = 0; p < m; ++p)
            for (int q = 0; q < m; ++q) {
                tot += A[p * m + q] * A[p * m + q];
                if (p != q) off += A[p * m + q] * A[p * m + q];
            }
        if (off <= 1e-34 * tot) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = A[p * m + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double theta = (A[q * m + q] - A[p * m + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double cs = 1 / std::sqrt(t * t + 1), sn = t * cs;
                for (int k = 0; k < m; ++k) {
                    const double akp = A[k * m + p], akq = A[k * m + q];
                    A[k * m + p] = cs * akp - sn * akq;
                    A[k * m + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < m; ++k) {
                    const double apk = A[p * m + k], aqk = A[q * m + k];
                    A[p * m + k] = cs * apk - sn * aqk;
                    A[q * m + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < m; ++k) {
                    const double vkp = V[k * m + p], vkq = V[k * m + q];
                    V[k * m + p] = cs * vkp - sn * vkq;
                    V[k * m + q] = sn * vkp + cs * vkq;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < m; ++i)
        if (A[i * m + i] < A[best * m + best]) best = i;
    double sgn = V[0 * m + best] < 0 ? -1.0 : 1.0;
    std::vector<double> a(2 * m, 0.0);
    for (int i = 0; i < m; ++i) a[2 * i] = sgn * V[i * m + best] * std::sqrt(2.0 * i + 0.5);
    return a;
}

double legendre_series(const std::vector<double>& a, double x) {
    // Clenshaw recurrence for sum_k a_k P_k(x)
    double b1 = 0, b2 = 0;
    for (int k = (int)a.size() - 1; k >= 0; --k) {
        const double alpha = (2.0 * k + 1) / (k + 1) * x;
        const double beta = -(k + 1.0) / (k + 2.0);
        const double b0 = a[k] + alpha * b1 + beta * b2;
        b2 = b1;
        b1 = b0;
    }
    return b1;
}
}  // namespace

double PswfConsts::psi0(double x) const {
    if (std::fabs(x) > 1) return 0.0;
    return legendre_series(a, x);
}

PswfConsts pswf_consts(int digits) {
    PswfConsts p;
    p.digits = digits;
    switch (digits) {  // include/utils.hpp:21-45 (literal constants of the method)
        case 3:
            p.c = 7.2462000846862793;
            p.n_diff = 6;
            p.poly = {1.627823522210361e-01, -4.553645597616490e-01, 4.171687104204163e-01, -7.073638602709915e-02,
                      -8.957845614474928e-02, 2.617986644718201e-02, 9.633427876507601e-03};
            break;
        case 6:
            p.c = 13.739999771118164;
            p.n_diff = 12;
            p.poly = {5.482525801351582e-02, -2.616592110444692e-01, 4.862652666337138e-01, -3.894296348642919e-01,
                      1.638587821812791e-02, 1.870328434198821e-01, -8.714171086568978e-02, -3.927020727017803e-02,
                      3.728187607052319e-02, 3.153734425831139e-03, -8.651313377285847e-03, 1.725110090795567e-04,
                      1.034762385284044e-03};
            break;
        case 9:
            p.c = 20.736000061035156;
            p.n_diff = 19;
            p.poly = {1.835718730962269e-02, -1.258015846164503e-01, 3.609487248584408e-01, -5.314579651112283e-01,
                      3.447559412892380e-01, 9.664692318551721e-02, -3.124274531849053e-01, 1.322460720579388e-01,
                      9.773007866584822e-02, -1.021958831082768e-01, -3.812847450976566e-03, 3.858117355875043e-02,
                      -8.728545924521301e-03, -9.401196355382909e-03, 4.024549377076924e-03, 1.512806105865091e-03,
                      -9.576734877247042e-04, -1.303457547418901e-04, 1.100385844683190e-04};
            break;
        case 12:
            p.c = 27.870000839233398;
            p.n_diff = 26;
            p.poly = {6.262472576363448e-03, -5.605742936112479e-02, 2.185890864792949e-01, -4.717350304955679e-01,
                      5.669680214206270e-01, -2.511606878849214e-01, -2.744523658778361e-01, 4.582527599363415e-01,
                      -1.397724810121539e-01, -2.131762135835757e-01, 1.995489373508990e-01, 1.793390341864239e-02,
                      -1.035055132403432e-01, 3.035606831075176e-02, 3.153931762550532e-02, -2.033178627450288e-02,
                      -5.406682731236552e-03, 7.543645573618463e-03, 1.437788047407851e-05, -1.928370882351732e-03,
                      2.891658777328665e-04, 3.332996162099811e-04, -8.397699195938912e-05, -3.015837377517983e-05,
                      9.640642701924662e-06};
            break;
        default:
            throw HpdmkError(-1, "double only supports 3, 6, 9 and 12 digits accuracy");  // src/tree.cpp:25-29
    }
    p.a = prolate_legendre_coefs(p.c);
    p.psi0_0 = p.psi0(0.0);
    p.C0 = p.a[0];
    p.lambda = 2 * p.a[0] / p.psi0_0;
    p.diff0 = -2.0 * M_PI * (p.a[0] / 3 + 2 * p.a[2] / 15) / p.C0 * p.c * p.c;
    p.n_window = (int)std::floor(2 * p.c / M_PI);
    return p;
}

std::vector<double> window_table(const PswfConsts& p, double L) {
    const int n = p.n_window, d = 2 * n + 1;
    const double dk = 2 * M_PI / L, sigma = L / p.c / 4;  // sigma_2 = L / (4 c)
    std::vector<double> W((size_t)d * d * (n + 1));
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j < d; ++j)
            for (int i = 0; i < d; ++i) {
                const double kx = (i - n) * dk, ky = (j - n) * dk, kz = (k - n) * dk;
                const double k2 = kx * kx + ky * ky + kz * kz;
                const double v = k2 == 0 ? 0.0 : 2 * M_PI * p.lambda * p.psi0(std::sqrt(k2) * sigma) / p.C0 / k2;
                W[((size_t)k * d + j) * d + i] = (k == n) ? v : 2 * v;
            }
    return W;
}

std::vector<double> difference_table(const PswfConsts& p, double L, int l) {
    const int n = p.n_diff, d = 2 * n + 1;
    const double b = L / std::ldexp(1.0, l);
    const double dk = (2.0 * M_PI / 3.0) / b;
    const double s_l = L / p.c / std::ldexp(1.0, l), s_l1 = 0.5 * s_l;
    auto F = [&](double x) { return 2 * M_PI * p.lambda * p.psi0(x) / p.C0; };
    std::vector<double> D((size_t)d * d * (n + 1));
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j < d; ++j)
            for (int i = 0; i < d; ++i) {
                const double kx = (i - n) * dk, ky = (j - n) * dk, kz = (k - n) * dk;
                const double k2 = kx * kx + ky * ky + kz * kz;
                double v;
                if (k2 == 0)
                    v = p.diff0 * (s_l1 * s_l1 - s_l * s_l);
                else
                    v = (F(std::sqrt(k2) * s_l1) - F(std::sqrt(k2) * s_l)) / k2;
                D[((size_t)k * d + j) * d + i] = (k == n) ? v : 2 * v;
            }
    return D;
}

// ------------------------------------------------------------------ octree
namespace {
inline uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
constexpr int kKeyBits = 21;

struct TNode {
    int level, cx, cy, cz, parent, child0;
    int lo, hi;  // range in the Morton-sorted particle array
};

inline uint64_t node_key(int l, int cx, int cy, int cz) {
    return ((uint64_t)l << 58) | (spread3(cx) | spread3(cy) << 1 | spread3(cz) << 2);
}
}  // namespace

Octree build_octree(int n, const double* r, double L, int n_per_leaf) {
    if (n <= 0) throw HpdmkError(-1, "n_src must be positive");
    if (!(L > 0)) throw HpdmkError(-1, "L must be positive");
    // Morton keys at level 21: cell = floor(x / L * 2^21), clamped into the box
    std::vector<std::pair<uint64_t, int>> keyed(n);
    const double scale = std::ldexp(1.0, kKeyBits);
    for (int i = 0; i < n; ++i) {
        uint64_t c3[3];
        for (int a = 0; a < 3; ++a) {
            const double x = r[3 * i + a];
            if (!(x >= 0 && x < L)) throw HpdmkError(-1, "particle coordinates must lie in [0, L)");
            long v = (long)std::floor(x / L * scale);
            v = std::min<long>(std::max<long>(v, 0), (1l << kKeyBits) - 1);
            c3[a] = (uint64_t)v;
        }
        keyed[i] = {spread3(c3[0]) | spread3(c3[1]) << 1 | spread3(c3[2]) << 2, i};
    }
    std::sort(keyed.begin(), keyed.end());
    std::vector<uint64_t> keys(n);
    for (int i = 0; i < n; ++i) keys[i] = keyed[i].first;

    std::vector<TNode> nodes;
    nodes.push_back({0, 0, 0, 0, -1, -1, 0, n});
    std::unordered_map<uint64_t, int> index;
    index[node_key(0, 0, 0, 0)] = 0;
    auto split = [&](int id) {
        const TNode p = nodes[id];
        const int c0 = (int)nodes.size();
        const int shift = 3 * (kKeyBits - (p.level + 1));
        int lo = p.lo;
        for (int c = 0; c < 8; ++c) {
            // children partition the sorted range by the next Morton digit
            int hi = lo;
            while (hi < p.hi && (int)((keys[hi] >> shift) & 7) == c) ++hi;
            nodes.push_back({p.level + 1, 2 * p.cx + (c & 1), 2 * p.cy + ((c >> 1) & 1), 2 * p.cz + ((c >> 2) & 1),
                             id, -1, lo, hi});
            index[node_key(p.level + 1, nodes.back().cx, nodes.back().cy, nodes.back().cz)] = c0 + c;
            lo = hi;
        }
        nodes[id].child0 = c0;
    };
    for (size_t id = 0; id < nodes.size(); ++id)
        if (nodes[id].hi - nodes[id].lo > n_per_leaf && nodes[id].level < kKeyBits) split((int)id);
    int maxl = 0;
    for (auto& t : nodes) maxl = std::max(maxl, t.level);
    if (maxl + 1 <= 2) throw HpdmkError(-2, "max depth is less than 2, too low for hpdmk");  // src/tree.cpp:352-353
    for (int c = 0; c < 8; ++c)
        if (nodes[1 + c].child0 < 0) split(1 + c);
    // 2:1 balance: for every non-leaf node all 26 same-level neighbours must exist
    auto covering = [&](int l, int cx, int cy, int cz) {
        for (int m = l; m >= 0; --m) {
            auto it = index.find(node_key(m, cx >> (l - m), cy >> (l - m), cz >> (l - m)));
            if (it != index.end()) return it->second;
        }
        return 0;
    };
    for (bool changed = true; changed;) {
        changed = false;
        for (size_t id = 0; id < nodes.size(); ++id) {
            if (nodes[id].child0 < 0 || nodes[id].level < 1) continue;
            const int l = nodes[id].level, nl = 1 << l;
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (!dx && !dy && !dz) continue;
                        const int cov = covering(l, (nodes[id].cx + dx + nl) & (nl - 1),
                                                 (nodes[id].cy + dy + nl) & (nl - 1),
                                                 (nodes[id].cz + dz + nl) & (nl - 1));
                        if (nodes[cov].level < l && nodes[cov].child0 < 0) {
                            split(cov);
                            changed = true;
                        }
                    }
        }
    }
    // Morton pre-order numbering (iterative DFS, children in digit order)
    Octree T;
    std::vector<int> pre;
    pre.reserve(nodes.size());
    std::vector<int> stack{0};
    while (!stack.empty()) {
        const int id = stack.back();
        stack.pop_back();
        pre.push_back(id);
        if (nodes[id].child0 >= 0)
            for (int c = 7; c >= 0; --c) stack.push_back(nodes[id].child0 + c);
    }
    const int nn = (int)pre.size();
    std::vector<int> newid(nodes.size());
    for (int i = 0; i < nn; ++i) newid[pre[i]] = i;
    T.n_nodes = nn;
    T.level.resize(nn);
    T.cell.resize(3 * nn);
    T.parent.resize(nn);
    T.child.assign(8 * nn, -1);
    T.is_leaf.resize(nn);
    T.leaf_count.assign(nn, 0);
    for (int i = 0; i < nn; ++i) {
        const TNode& t = nodes[pre[i]];
        T.level[i] = t.level;
        T.cell[3 * i] = t.cx;
        T.cell[3 * i + 1] = t.cy;
        T.cell[3 * i + 2] = t.cz;
        T.parent[i] = t.parent < 0 ? -1 : newid[t.parent];
        T.is_leaf[i] = t.child0 < 0;
        if (t.child0 >= 0)
            for (int c = 0; c < 8; ++c) T.child[8 * i + c] = newid[t.child0 + c];
        else
            T.leaf_count[i] = t.hi - t.lo;
        T.max_depth = std::max(T.max_depth, t.level + 1);
    }
    T.order.resize(n);
    for (int i = 0; i < n; ++i) T.order[i] = keyed[i].second;  // pre-order of leaves == Morton order
    // descendant-leaf ranges (pre-order makes every subtree contiguous)
    T.first_leaf.assign(nn, -1);
    T.last_leaf.assign(nn, -1);
    for (int i = nn - 1; i >= 0; --i) {
        if (T.is_leaf[i]) {
            T.first_leaf[i] = T.last_leaf[i] = i;
        } else {
            T.first_leaf[i] = T.first_leaf[T.child[8 * i]];
            T.last_leaf[i] = T.last_leaf[T.child[8 * i + 7]];
        }
    }
    // same-level neighbours (src/tree.cpp:136-147)
    auto exact = [&](int l, int cx, int cy, int cz) {
        auto it = index.find(node_key(l, cx, cy, cz));
        return it == index.end() ? -1 : newid[it->second];
    };
    T.colleague.assign(26 * nn, -1);
    for (int i = 0; i < nn; ++i) {
        const int l = T.level[i];
        if (l < 2) continue;
        const int nl = 1 << l;
        int s = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!dx && !dy && !dz) continue;
                    T.colleague[26 * i + s++] = exact(l, (T.cell[3 * i] + dx + nl) & (nl - 1),
                                                      (T.cell[3 * i + 1] + dy + nl) & (nl - 1),
                                                      (T.cell[3 * i + 2] + dz + nl) & (nl - 1));
                }
    }
    // near lists of leaves: the leaf itself, every leaf under an existing
    // colleague, and the coarser leaves that cover missing neighbour cells
    // (src/tree.cpp:147-164, src/energy_shift.cpp:249-297)
    T.near_off.assign(nn + 1, 0);
    std::vector<std::vector<std::pair<int, int>>> nl_lists(nn);
    for (int i = 0; i < nn; ++i) {
        if (!T.is_leaf[i] || T.level[i] < 2) continue;
        auto& lst = nl_lists[i];
        lst.push_back({i, 13});
        const int l = T.level[i], nlv = 1 << l;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!dx && !dy && !dz) continue;
                    int c[3] = {T.cell[3 * i] + dx, T.cell[3 * i + 1] + dy, T.cell[3 * i + 2] + dz};
                    int img[3];
                    for (int a = 0; a < 3; ++a) {
                        img[a] = c[a] < 0 ? -1 : (c[a] >= nlv ? 1 : 0);
                        c[a] = (c[a] + nlv) & (nlv - 1);
                    }
                    const int code = (img[0] + 1) + 3 * (img[1] + 1) + 9 * (img[2] + 1);
                    const int nb = exact(l, c[0], c[1], c[2]);
                    if (nb >= 0) {
                        for (int k = T.first_leaf[nb]; k <= T.last_leaf[nb]; ++k)
                            if (T.is_leaf[k]) lst.push_back({k, code});
                    } else {
                        const int cov = newid[covering(l, c[0], c[1], c[2])];
                        bool seen = false;
                        for (auto& e : lst) seen |= (e.first == cov && e.second == code);
                        if (!seen) lst.push_back({cov, code});
                    }
                }
    }
    for (int i = 0; i < nn; ++i) T.near_off[i + 1] = T.near_off[i] + (int)nl_lists[i].size();
    T.near_ids.resize(T.near_off[nn]);
    T.near_img.resize(T.near_off[nn]);
    for (int i = 0; i < nn; ++i)
        for (size_t k = 0; k < nl_lists[i].size(); ++k) {
            T.near_ids[T.near_off[i] + k] = nl_lists[i][k].first;
            T.near_img[T.near_off[i] + k] = (int8_t)nl_lists[i][k].second;
        }
    return T;
}

// ------------------------------------------------------------------ PROXY tables
int proxy_order(int digits) {
    if (digits <= 3) return 18;  // src/upward_pass.cpp:9-14
    if (digits <= 6) return 28;  // an order of include/chebychev.hpp:86-96; interpolation error < 1e-7
    // digits 9: order 44 interpolates the plane waves (|k| b/2 up to 19 pi/3)
    // to 7e-12; the calculator's 38 only to 3e-8
    if (digits <= 9) return 44;
    throw HpdmkError(HPDMK_ENOTIMPL, "Polynomial order for requested precision not implemented");
}

ProxyTables proxy_tables(int p, int n, double hw) {
    ProxyTables t;
    t.p = p;
    std::vector<double> x(p), Vi((size_t)p * p);
    for (int i = 0; i < p; ++i) x[i] = std::cos(M_PI * ((p - i - 1) + 0.5) / p);  // ascending nodes
    auto cheb = [p](double u, std::vector<double>& T) {
        T.assign(p, 0.0);
        T[0] = 1.0;
        if (p > 1) T[1] = u;
        for (int k = 2; k < p; ++k) T[k] = 2.0 * u * T[k - 1] - T[k - 2];
    };
    std::vector<double> Tv;
    for (int i = 0; i < p; ++i) {
        cheb(x[i], Tv);
        for (int a = 0; a < p; ++a) Vi[(size_t)a * p + i] = (a == 0 ? 1.0 : 2.0) / p * Tv[a];
    }
    t.A.assign(2 * (size_t)p * p, 0.0);
    for (int h = 0; h < 2; ++h)
        for (int i = 0; i < p; ++i) {
            cheb(0.5 * x[i] + (h ? 0.5 : -0.5), Tv);  // T_l at the child's nodes in parent coordinates
            for (int l = 0; l < p; ++l)
                for (int a = 0; a < p; ++a) t.A[(size_t)h * p * p + (size_t)l * p + a] += Vi[(size_t)a * p + i] * Tv[l];
        }
    const int d = 2 * n + 1;
    t.P.assign(2 * (size_t)d * p, 0.0);
    for (int k = 0; k < d; ++k)
        for (int i = 0; i < p; ++i) {
            const double th = (k - n) * hw * x[i];
            const double c = std::cos(th), s = -std::sin(th);
            for (int a = 0; a < p; ++a) {
                t.P[2 * ((size_t)k * p + a)] += c * Vi[(size_t)a * p + i];
                t.P[2 * ((size_t)k * p + a) + 1] += s * Vi[(size_t)a * p + i];
            }
        }
    return t;
}

}  // namespace hpdmk_amd
