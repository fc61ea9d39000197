// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle/dmk_oracle.hpp).
//
// Classic FP64 Ewald summation, restating src/ewald.cpp:17-308 of the
// reference (energy = short erfc pairs + long |S(k)|^2 on the cube |n_i| <= n
// + self).  A cell list replaces the vesin neighbour list (ewald.cpp:73-101);
// the trial-move energy change is formed from potential differences, which is
// exact for a pair potential and costs O(N + K) per move instead of two full
// energies (test/hpdmk_test.cpp:428-441).
#include <cmath>
#include <omp.h>

#include "dmk_oracle.hpp"

namespace oracle {

Ewald::Ewald(double L_, double s_, double alpha_, double eps_, int n_, const double* q_, const double* r_)
    : L(L_), s(s_), alpha(alpha_), eps(eps_), V(L_ * L_ * L_), r_c(s_ / alpha_), k_c(2 * s_ * alpha_), n(n_) {
    q.assign(q_, q_ + n);
    r.assign(r_, r_ + 3 * n);
    nk = (int)std::ceil(k_c / (2 * M_PI / L));  // src/ewald.cpp:29
    const int d = 2 * nk + 1;
    const double dk = 2 * M_PI / L;
    kmat.assign((size_t)d * d * d, 0.0);
    for (int ix = 0; ix < d; ++ix)  // src/ewald.cpp:52-71
        for (int iy = 0; iy < d; ++iy)
            for (int iz = 0; iz < d; ++iz) {
                const double kx = (ix - nk) * dk, ky = (iy - nk) * dk, kz = (iz - nk) * dk;
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 == 0.0) continue;
                kmat[((size_t)ix * d + iy) * d + iz] = std::exp(-k2 / (4 * alpha * alpha)) / k2;
            }
    init_planewave_coeffs();
}

static void phase_rows(double x, double y, double z, int nk, double dk, std::complex<double>* X,
                       std::complex<double>* Y, std::complex<double>* Z, double sign) {
    const int d = 2 * nk + 1;
    for (int a = 0; a < d; ++a) {
        const double m = a - nk;
        X[a] = std::polar(1.0, sign * m * dk * x);
        Y[a] = std::polar(1.0, sign * m * dk * y);
        Z[a] = std::polar(1.0, sign * m * dk * z);
    }
}

void Ewald::init_planewave_coeffs() {  // src/ewald.cpp:103-176
    const int d = 2 * nk + 1;
    const size_t K = (size_t)d * d * d;
    const double dk = 2 * M_PI / L;
    S.assign(K, 0.0);
    const int nt = omp_get_max_threads();
    std::vector<std::vector<std::complex<double>>> part(nt);
#pragma omp parallel
    {
        const int t = omp_get_thread_num();
        auto& acc = part[t];
        acc.assign(K, 0.0);
        std::vector<std::complex<double>> X(d), Y(d), Z(d);
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            phase_rows(r[3 * i], r[3 * i + 1], r[3 * i + 2], nk, dk, X.data(), Y.data(), Z.data(), -1.0);
            for (int ix = 0; ix < d; ++ix) {
                const auto t1 = q[i] * X[ix];
                for (int iy = 0; iy < d; ++iy) {
                    const auto t2 = Y[iy] * t1;
                    auto* sp = &acc[((size_t)ix * d + iy) * d];
                    for (int iz = 0; iz < d; ++iz) sp[iz] += Z[iz] * t2;
                }
            }
        }
    }
    for (int t = 0; t < nt; ++t)
        for (size_t k = 0; k < K; ++k) S[k] += part[t][k];
}

namespace {
// all (j, image) within r_c of point x; calls f(j, dist)
template <typename F>
void for_neighbours(const Ewald& e, double x, double y, double z, F&& f) {
    const int M = (int)std::ceil(e.r_c / e.L);
    for (int j = 0; j < e.n; ++j) {
        double bx = x - e.r[3 * j], by = y - e.r[3 * j + 1], bz = z - e.r[3 * j + 2];
        if (e.r_c < 0.5 * e.L) {  // minimum image is the only candidate
            bx -= e.L * std::nearbyint(bx / e.L);
            by -= e.L * std::nearbyint(by / e.L);
            bz -= e.L * std::nearbyint(bz / e.L);
            const double rr = std::sqrt(bx * bx + by * by + bz * bz);
            if (rr < e.r_c) f(j, rr);
        } else {
            for (int mx = -M; mx <= M; ++mx)
                for (int my = -M; my <= M; ++my)
                    for (int mz = -M; mz <= M; ++mz) {
                        const double ax = bx + mx * e.L, ay = by + my * e.L, az = bz + mz * e.L;
                        const double rr = std::sqrt(ax * ax + ay * ay + az * az);
                        if (rr < e.r_c) f(j, rr);
                    }
        }
    }
}
}  // namespace

double Ewald::energy_short() const {  // src/ewald.cpp:192-203 (full list, halved)
    double E = 0;
#pragma omp parallel for reduction(+ : E) schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        double ei = 0;
        for_neighbours(*this, r[3 * i], r[3 * i + 1], r[3 * i + 2], [&](int j, double rr) {
            if (rr > 0) ei += q[i] * q[j] * std::erfc(alpha * rr) / rr;
        });
        E += ei;
    }
    return E / 2.0;
}

double Ewald::energy_long() const {  // src/ewald.cpp:205-214
    double E = 0;
    for (size_t k = 0; k < S.size(); ++k) E += kmat[k] * std::norm(S[k]);
    return E * 2 * M_PI / V;
}

double Ewald::energy_self() const {  // src/ewald.cpp:216-223
    double E = 0;
    for (int i = 0; i < n; ++i) E += -q[i] * q[i] * alpha / std::sqrt(M_PI);
    return E;
}

double Ewald::compute_energy() { return (energy_short() + energy_long() + energy_self()) / eps; }

double Ewald::potential(double x, double y, double z, long exclude) const {  // src/ewald.cpp:281-308
    double ps = 0;
    for_neighbours(*this, x, y, z, [&](int j, double rr) {
        if (j != exclude && rr > 0) ps += q[j] * std::erfc(alpha * rr) / rr;
    });
    const int d = 2 * nk + 1;
    const double dk = 2 * M_PI / L;
    std::vector<std::complex<double>> X(d), Y(d), Z(d), Xe(d), Ye(d), Ze(d);
    phase_rows(x, y, z, nk, dk, X.data(), Y.data(), Z.data(), +1.0);
    const bool ex = exclude >= 0;
    if (ex) phase_rows(r[3 * exclude], r[3 * exclude + 1], r[3 * exclude + 2], nk, dk, Xe.data(), Ye.data(), Ze.data(), -1.0);
    double pl = 0;
#pragma omp parallel for reduction(+ : pl)
    for (int ix = 0; ix < d; ++ix)
        for (int iy = 0; iy < d; ++iy) {
            const auto xy = X[ix] * Y[iy];
            const auto xye = ex ? q[exclude] * Xe[ix] * Ye[iy] : std::complex<double>(0);
            const size_t base = ((size_t)ix * d + iy) * d;
            for (int iz = 0; iz < d; ++iz) {
                std::complex<double> Sk = S[base + iz];
                if (ex) Sk -= xye * Ze[iz];
                pl += kmat[base + iz] * (Sk * (xy * Z[iz])).real();
            }
        }
    return (ps + pl * 4 * M_PI / V) / eps;
}

double Ewald::shift_energy(long p, double tx, double ty, double tz) const {
    return q[p] * (potential(tx, ty, tz, p) - potential(r[3 * p], r[3 * p + 1], r[3 * p + 2], p));
}

void Ewald::apply_move(long p, double tx, double ty, double tz) {
    const int d = 2 * nk + 1;
    const double dk = 2 * M_PI / L;
    std::vector<std::complex<double>> Xt(d), Yt(d), Zt(d), Xo(d), Yo(d), Zo(d);
    phase_rows(tx, ty, tz, nk, dk, Xt.data(), Yt.data(), Zt.data(), -1.0);
    phase_rows(r[3 * p], r[3 * p + 1], r[3 * p + 2], nk, dk, Xo.data(), Yo.data(), Zo.data(), -1.0);
    for (int ix = 0; ix < d; ++ix)
        for (int iy = 0; iy < d; ++iy) {
            const auto at = q[p] * Xt[ix] * Yt[iy], ao = q[p] * Xo[ix] * Yo[iy];
            const size_t base = ((size_t)ix * d + iy) * d;
            for (int iz = 0; iz < d; ++iz) S[base + iz] += at * Zt[iz] - ao * Zo[iz];
        }
    r[3 * p] = tx;
    r[3 * p + 1] = ty;
    r[3 * p + 2] = tz;
}

}  // namespace oracle
