// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle/dmk_oracle.hpp).
//
// The PROXY initialisation of the outgoing expansions (src/upward_pass.cpp,
// include/chebychev.hpp) restated in plain C++:
//   * particles -> Chebyshev coefficients C[a,b,c] = sum_j q_j T_a(x^) T_b(y^) T_c(z^)
//     with x^ = 2 (x - centre) / box size (charge2proxycharge, upward_pass.cpp:109-149)
//     at the lowest non-leaf level D-2 (upward_pass.cpp:270-287);
//   * child -> parent transfer by the exact re-expansion
//     T_l(x/2 -+ 1/2) = sum_a A[l,a] T_a(x), applied along each axis
//     (tensorprod_transform, upward_pass.cpp:19-64; matrices of
//     chebychev.hpp:203-296), levels D-3 .. 2 (upward_pass.cpp:289-321);
//   * coefficients -> plane waves rho[k] = e^{-i k.c} sum P[kx,a] P[ky,b] P[kz,c] C[a,b,c],
//     P[k,a] = sum_i e^{-i k (b/2) x_i} (V^-1)[a,i] over the Chebyshev nodes x_i
//     (proxycharge2pw + calc_prox_to_pw + decenter_phase, upward_pass.cpp:67-106,
//     152-203).
// The Vandermonde inverse is taken by LU with partial pivoting, as the
// reference does with Eigen (chebychev.hpp:127-139, upward_pass.cpp:172-174).
//
// Deliberate differences from the reference binary:
//  * a leaf at a level shallower than D-2 gets its coefficients from its own
//    particles.  The reference forms them only at level D-2, so such a leaf's
//    parent reads an empty coefficient vector (upward_pass.cpp:297-305);
//  * orders: digits <= 3 use order 18 (upward_pass.cpp:9-14); digits 6 use
//    order 28, one of the orders the reference's polynomial calculator
//    provides (chebychev.hpp:86-96) and whose interpolation error of the
//    digits-6 plane waves is below 1e-7; digits 9 use order 44: the
//    calculator's next order, 38, interpolates the digits-9 plane waves
//    (|k| b/2 up to 19 pi/3) only to about 1e-7, order 44 to below 1e-10;
//    higher digits throw like the reference.
#include <cmath>
#include <stdexcept>

#include "dmk_oracle.hpp"

namespace oracle {

int proxy_order(int digits) {
    if (digits <= 3) return 18;
    if (digits <= 6) return 28;
    if (digits <= 9) return 44;
    throw std::runtime_error("Polynomial order for requested precision not implemented");
}

std::vector<double> cheb_nodes(int p) {  // include/chebychev.hpp:35-42 on [-1, 1]
    std::vector<double> x(p);
    for (int i = 0; i < p; ++i) x[i] = std::cos(M_PI * ((p - i - 1) + 0.5) / p);
    return x;
}

// V[i][j] = T_j(x_i) (include/chebychev.hpp:45-60), row-major p x p
std::vector<double> cheb_vandermonde(const std::vector<double>& x, int p) {
    const int n = (int)x.size();
    std::vector<double> V((size_t)n * p);
    for (int i = 0; i < n; ++i) {
        double t0 = 1.0, t1 = x[i];
        V[(size_t)i * p] = 1.0;
        if (p > 1) V[(size_t)i * p + 1] = t1;
        for (int j = 2; j < p; ++j) {
            const double t2 = 2.0 * x[i] * t1 - t0;
            V[(size_t)i * p + j] = t2;
            t0 = t1;
            t1 = t2;
        }
    }
    return V;
}

// inverse of a p x p row-major matrix by LU with partial pivoting
std::vector<double> lu_inverse(std::vector<double> A, int p) {
    std::vector<int> piv(p);
    for (int i = 0; i < p; ++i) piv[i] = i;
    for (int k = 0; k < p; ++k) {
        int m = k;
        for (int i = k + 1; i < p; ++i)
            if (std::fabs(A[(size_t)i * p + k]) > std::fabs(A[(size_t)m * p + k])) m = i;
        if (A[(size_t)m * p + k] == 0.0) throw std::runtime_error("singular Vandermonde matrix");
        if (m != k) {
            for (int j = 0; j < p; ++j) std::swap(A[(size_t)k * p + j], A[(size_t)m * p + j]);
            std::swap(piv[k], piv[m]);
        }
        for (int i = k + 1; i < p; ++i) {
            const double f = A[(size_t)i * p + k] / A[(size_t)k * p + k];
            A[(size_t)i * p + k] = f;
            for (int j = k + 1; j < p; ++j) A[(size_t)i * p + j] -= f * A[(size_t)k * p + j];
        }
    }
    std::vector<double> inv((size_t)p * p, 0.0);
    for (int c = 0; c < p; ++c) {  // solve A x = e_{piv^-1}: column c of the inverse
        std::vector<double> y(p);
        for (int i = 0; i < p; ++i) {
            double s = piv[i] == c ? 1.0 : 0.0;
            for (int j = 0; j < i; ++j) s -= A[(size_t)i * p + j] * y[j];
            y[i] = s;
        }
        for (int i = p - 1; i >= 0; --i) {
            double s = y[i];
            for (int j = i + 1; j < p; ++j) s -= A[(size_t)i * p + j] * inv[(size_t)j * p + c];
            inv[(size_t)i * p + c] = s / A[(size_t)i * p + i];
        }
    }
    return inv;
}

// A_h[l][a] with T_l(x/2 + (h ? 1/2 : -1/2)) = sum_a A_h[l][a] T_a(x): the
// transpose of V^-1 V_h (include/chebychev.hpp:180-193, 203-210)
std::vector<double> cheb_c2p(int p, int half) {
    const auto x = cheb_nodes(p);
    const auto Vi = lu_inverse(cheb_vandermonde(x, p), p);
    std::vector<double> xs(p);
    for (int i = 0; i < p; ++i) xs[i] = 0.5 * x[i] + (half ? 0.5 : -0.5);
    const auto Vh = cheb_vandermonde(xs, p);
    std::vector<double> A((size_t)p * p, 0.0);
    for (int a = 0; a < p; ++a)
        for (int l = 0; l < p; ++l) {
            double s = 0;  // (V^-1 V_h)[a][l]
            for (int i = 0; i < p; ++i) s += Vi[(size_t)a * p + i] * Vh[(size_t)i * p + l];
            A[(size_t)l * p + a] = s;
        }
    return A;
}

// P[k][a] = sum_i e^{-i (k - n) (pi/3) x_i} (V^-1)[a][i], k = 0..2n: the
// conjugate of calc_prox_to_pw (upward_pass.cpp:152-180) with hpw (b/2) =
// (2 pi / 3) / b * (b / 2) = pi / 3 on every level l >= 2
std::vector<std::complex<double>> cheb_poly2pw(int p, int n) {
    const auto x = cheb_nodes(p);
    const auto Vi = lu_inverse(cheb_vandermonde(x, p), p);
    const int d = 2 * n + 1;
    std::vector<std::complex<double>> P((size_t)d * p);
    for (int k = 0; k < d; ++k)
        for (int a = 0; a < p; ++a) {
            std::complex<double> s = 0;
            for (int i = 0; i < p; ++i) {
                const double th = (k - n) * (M_PI / 3.0) * x[i];
                s += std::complex<double>(std::cos(th), -std::sin(th)) * Vi[(size_t)a * p + i];
            }
            P[(size_t)k * p + a] = s;
        }
    return P;
}

// C[a + p b + p^2 c] = sum_j q_j T_a(x^) T_b(y^) T_c(z^) (upward_pass.cpp:109-149)
template <typename Real>
std::vector<double> DmkTree<Real>::proxy_from_particles(int nd, int p) const {
    std::vector<double> C((size_t)p * p * p, 0.0);
    const double scale = 2.0 / boxsize[level[nd]];
    std::vector<double> tx(p), ty(p), tz(p);
    auto cheb = [p](double x, std::vector<double>& t) {
        t[0] = 1.0;
        if (p > 1) t[1] = x;
        for (int k = 2; k < p; ++k) t[k] = 2.0 * x * t[k - 1] - t[k - 2];
    };
    for (long s : node_particles[nd]) {
        cheb(scale * ((double)r_sorted[3 * s] - center[3 * nd]), tx);
        cheb(scale * ((double)r_sorted[3 * s + 1] - center[3 * nd + 1]), ty);
        cheb(scale * ((double)r_sorted[3 * s + 2] - center[3 * nd + 2]), tz);
        const double q = (double)q_sorted[s];
        for (int c = 0; c < p; ++c)
            for (int b = 0; b < p; ++b) {
                const double w = q * ty[b] * tz[c];
                double* row = &C[(size_t)p * (b + (size_t)p * c)];
                for (int a = 0; a < p; ++a) row[a] += w * tx[a];
            }
    }
    return C;
}

template <typename Real>
void DmkTree<Real>::form_outgoing_pw_proxy() {
    const int p = proxy_order(params.digits);
    const int n = pw.n_diff, d = 2 * n + 1;
    form_outgoing_pw_root();
    const int lo = 2, hi = max_depth - 2;  // expansion levels
    proxy_coeffs.assign(n_nodes, {});
    // particles -> coefficients: every box of level D-2, and shallower leaves
    for (int l = lo; l <= hi; ++l)
        for (int nd : level_indices[l])
            if ((l == hi || is_leaf[nd]) && !node_particles[nd].empty()) proxy_coeffs[nd] = proxy_from_particles(nd, p);
    // child -> parent (tensorprod_transform, upward_pass.cpp:19-64, 289-321)
    const std::vector<double> A[2] = {cheb_c2p(p, 0), cheb_c2p(p, 1)};
    const size_t p3 = (size_t)p * p * p;
    std::vector<double> t1(p3), t2(p3);
    for (int l = hi - 1; l >= lo; --l)
        for (int nd : level_indices[l]) {
            if (is_leaf[nd]) continue;
            for (int ci = 0; ci < 8; ++ci) {
                const int ch = child[8 * nd + ci];
                if (ch < 0 || proxy_coeffs[ch].empty()) continue;
                if (proxy_coeffs[nd].empty()) proxy_coeffs[nd].assign(p3, 0.0);
                const double* Ax = A[ci & 1].data();
                const double* Ay = A[(ci >> 1) & 1].data();
                const double* Az = A[(ci >> 2) & 1].data();
                const auto& Cc = proxy_coeffs[ch];
                for (int c = 0; c < p; ++c)  // x
                    for (int b = 0; b < p; ++b)
                        for (int l1 = 0; l1 < p; ++l1) {
                            double s = 0;
                            for (int a = 0; a < p; ++a) s += Ax[(size_t)l1 * p + a] * Cc[a + p * (b + (size_t)p * c)];
                            t1[l1 + p * (b + (size_t)p * c)] = s;
                        }
                for (int c = 0; c < p; ++c)  // y
                    for (int l2 = 0; l2 < p; ++l2)
                        for (int l1 = 0; l1 < p; ++l1) {
                            double s = 0;
                            for (int b = 0; b < p; ++b) s += Ay[(size_t)l2 * p + b] * t1[l1 + p * (b + (size_t)p * c)];
                            t2[l1 + p * (l2 + (size_t)p * c)] = s;
                        }
                auto& Cp = proxy_coeffs[nd];
                for (int l3 = 0; l3 < p; ++l3)  // z, accumulated into the parent
                    for (int l2 = 0; l2 < p; ++l2)
                        for (int l1 = 0; l1 < p; ++l1) {
                            double s = 0;
                            for (int c = 0; c < p; ++c) s += Az[(size_t)l3 * p + c] * t2[l1 + p * (l2 + (size_t)p * c)];
                            Cp[l1 + p * (l2 + (size_t)p * l3)] += s;
                        }
            }
        }
    // coefficients -> plane waves, then the de-centring phase e^{-i k.c}
    const auto P = cheb_poly2pw(p, n);
    using cd = std::complex<double>;
    std::vector<cd> g1((size_t)p * p * (n + 1)), g2((size_t)p * d * (n + 1));
    for (int l = lo; l <= hi; ++l)
        for (int nd : level_indices[l]) {
            auto& out = outgoing[nd];
            std::fill(out.begin(), out.end(), cplx(0));
            if (node_particles[nd].empty()) continue;
            const auto& C = proxy_coeffs[nd];
            for (int kz = 0; kz <= n; ++kz)  // z: only k_z <= 0 is stored
                for (int b = 0; b < p; ++b)
                    for (int a = 0; a < p; ++a) {
                        cd s = 0;
                        for (int c = 0; c < p; ++c) s += P[(size_t)kz * p + c] * C[a + p * (b + (size_t)p * c)];
                        g1[a + p * (b + (size_t)p * kz)] = s;
                    }
            for (int kz = 0; kz <= n; ++kz)
                for (int ky = 0; ky < d; ++ky)
                    for (int a = 0; a < p; ++a) {
                        cd s = 0;
                        for (int b = 0; b < p; ++b) s += P[(size_t)ky * p + b] * g1[a + p * (b + (size_t)p * kz)];
                        g2[a + p * (ky + (size_t)d * kz)] = s;
                    }
            const double dk = delta_k[l];
            for (int kz = 0; kz <= n; ++kz)
                for (int ky = 0; ky < d; ++ky)
                    for (int kx = 0; kx < d; ++kx) {
                        cd s = 0;
                        for (int a = 0; a < p; ++a) s += P[(size_t)kx * p + a] * g2[a + p * (ky + (size_t)d * kz)];
                        const double th = dk * ((kx - n) * center[3 * nd] + (ky - n) * center[3 * nd + 1] +
                                                (kz - n) * center[3 * nd + 2]);
                        s *= cd(std::cos(th), -std::sin(th));
                        out[kx + d * (ky + (size_t)d * kz)] = cplx((Real)s.real(), (Real)s.imag());
                    }
        }
}

template std::vector<double> DmkTree<double>::proxy_from_particles(int, int) const;
template std::vector<double> DmkTree<float>::proxy_from_particles(int, int) const;
template void DmkTree<double>::form_outgoing_pw_proxy();
template void DmkTree<float>::form_outgoing_pw_proxy();

}  // namespace oracle
