// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle/dmk_oracle.hpp).
//
// Prolate spheroidal wave function psi_0^c and the per-digits DMK constants,
// restating src/pswf.cpp:610-1219 and include/utils.hpp:20-52 of the reference.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>

#include "dmk_oracle.hpp"

namespace oracle {

// Gauss-Legendre rule on [-1,1] (the reference builds it with legerts,
// src/pswf.cpp:82-179); Newton iteration on P_n from the Chebyshev guess.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
        }
        x[n - 1 - i] = z;  // ascending
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

// Lowest eigenpair of the even part of the prolate operator
//   A = -d/dx (1-x^2) d/dx + c^2 x^2
// in the normalised Legendre basis sqrt(k+1/2) P_k, k = 0, 2, 4, ...
// (the same tridiagonal system as prolmatr/prolcoef, src/pswf.cpp:610-676).
// Eigenvalue by Sturm-sequence bisection, eigenvector by inverse iteration
// (the reference uses QL + inverse iteration, src/pswf.cpp:678-845).
Prolate0::Prolate0(double c_) : c(c_) {
    const int m = 64;  // even terms: degree <= 126, far beyond the decay point for c <= 30
    std::vector<double> d(m), e(m - 1);
    const double c2 = c * c;
    for (int i = 0; i < m; ++i) {
        const double k = 2.0 * i;
        const double beta = (2 * k * k + 2 * k - 1) / ((2 * k - 1) * (2 * k + 3));
        d[i] = k * (k + 1) + c2 * beta;
    }
    for (int i = 0; i < m - 1; ++i) {
        const double k = 2.0 * i;
        e[i] = c2 * (k + 1) * (k + 2) / ((2 * k + 3) * std::sqrt((2 * k + 1) * (2 * k + 5)));
    }
    auto count_below = [&](double x) {
        int cnt = 0;
        double qv = d[0] - x;
        if (qv < 0) ++cnt;
        for (int i = 1; i < m; ++i) {
            if (qv == 0) qv = 1e-300;
            qv = d[i] - x - e[i - 1] * e[i - 1] / qv;
            if (qv < 0) ++cnt;
        }
        return cnt;
    };
    double lo = 1e300, hi = d[0];
    for (int i = 0; i < m; ++i) {
        double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < m - 1 ? std::fabs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - r);
    }
    for (int it = 0; it < 400 && hi - lo > 1e-15 * std::max(1.0, std::fabs(hi)); ++it) {
        double mid = 0.5 * (lo + hi);
        if (count_below(mid) >= 1)
            hi = mid;
        else
            lo = mid;
    }
    const double chi0 = 0.5 * (lo + hi);
    const double shift = chi0 - 1e-10 * std::max(1.0, std::fabs(chi0));
    std::vector<double> v(m, 1.0), bb(m), cp(m);
    for (int it = 0; it < 6; ++it) {
        // Thomas algorithm for (T - shift I) y = v
        std::vector<double> dd(m);
        for (int i = 0; i < m; ++i) dd[i] = d[i] - shift;
        cp[0] = e[0] / dd[0];
        bb[0] = v[0] / dd[0];
        for (int i = 1; i < m; ++i) {
            const double den = dd[i] - e[i - 1] * cp[i - 1];
            cp[i] = (i < m - 1) ? e[i] / den : 0.0;
            bb[i] = (v[i] - e[i - 1] * bb[i - 1]) / den;
        }
        for (int i = m - 2; i >= 0; --i) bb[i] -= cp[i] * bb[i + 1];
        double nrm = 0;
        for (double t : bb) nrm += t * t;
        nrm = std::sqrt(nrm);
        for (int i = 0; i < m; ++i) v[i] = bb[i] / nrm;
    }
    if (v[0] < 0)
        for (double& t : v) t = -t;
    // un-normalised Legendre coefficients; the extra sqrt(2) matches the
    // reference's scaling (src/pswf.cpp:991-993), which cancels in every ratio.
    a.assign(2 * m, 0.0);
    for (int i = 0; i < m; ++i) a[2 * i] = std::sqrt(2.0) * v[i] * std::sqrt(2.0 * i + 0.5);
}

double Prolate0::eval(double x) const {
    if (std::fabs(x) > 1) return 0.0;  // src/pswf.cpp:980-985 (c < thresh branch)
    double p0 = 1.0, p1 = x, s = a[0];
    for (size_t k = 1; k < a.size(); ++k) {
        if (k >= 2) {
            double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        s += a[k] * p1;
    }
    return s;
}

struct GL200 {
    std::vector<double> x, w;
    GL200() { gauss_legendre(200, x, w); }
};
static const GL200& gl200() {
    static const GL200 g;
    return g;
}
static const std::vector<double>& gl200_x() { return gl200().x; }
static const std::vector<double>& gl200_w() { return gl200().w; }

double Prolate0::int0r(double r) const {
    const auto& xs = gl200_x();
    const auto& ws = gl200_w();
    double val = 0;
    for (int i = 0; i < 200; ++i) val += ws[i] * r / 2 * eval((xs[i] + 1) * r / 2);
    return val;
}

double Prolate0::intx2_0r(double r) const {
    const auto& xs = gl200_x();
    const auto& ws = gl200_w();
    double val = 0;
    for (int i = 0; i < 200; ++i) {
        const double t = (xs[i] + 1) * r / 2;
        val += ws[i] * r / 2 * eval(t) * t * t;
    }
    return val;
}

double Prolate0::lambda() const {
    const auto& xs = gl200_x();
    const auto& ws = gl200_w();
    double l = 0;
    for (int i = 0; i < 200; ++i) l += ws[i] * eval(xs[i]) * std::cos(c * xs[i] * 0.5);
    return l / eval(0.5);
}

const Prolate0& prolate0(double c) {
    static std::map<double, Prolate0> cache;
    static std::mutex mu;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(c);
    if (it == cache.end()) it = cache.emplace(c, Prolate0(c)).first;
    return it->second;
}

// include/utils.hpp:20-52: bandwidth c, n_diff and the residual polynomial
// (ascending powers of u = 2x - 1; identical literals appear again in
// include/direct_eval.hpp:376-386).
const PswfParams& pswf_params(int digits) {
    static std::map<int, PswfParams> cache;
    static std::mutex mu;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(digits);
    if (it != cache.end()) return it->second;
    PswfParams p;
    p.digits = digits;
    if (digits == 3) {
        p.c = 7.2462000846862793;
        p.n_diff = 6;
        p.poly = {1.627823522210361e-01, -4.553645597616490e-01, 4.171687104204163e-01, -7.073638602709915e-02,
                  -8.957845614474928e-02, 2.617986644718201e-02, 9.633427876507601e-03};
    } else if (digits == 6) {
        p.c = 13.739999771118164;
        p.n_diff = 12;
        p.poly = {5.482525801351582e-02, -2.616592110444692e-01, 4.862652666337138e-01, -3.894296348642919e-01,
                  1.638587821812791e-02, 1.870328434198821e-01, -8.714171086568978e-02, -3.927020727017803e-02,
                  3.728187607052319e-02, 3.153734425831139e-03, -8.651313377285847e-03, 1.725110090795567e-04,
                  1.034762385284044e-03};
    } else if (digits == 9) {
        p.c = 20.736000061035156;
        p.n_diff = 19;
        p.poly = {1.835718730962269e-02, -1.258015846164503e-01, 3.609487248584408e-01, -5.314579651112283e-01,
                  3.447559412892380e-01, 9.664692318551721e-02, -3.124274531849053e-01, 1.322460720579388e-01,
                  9.773007866584822e-02, -1.021958831082768e-01, -3.812847450976566e-03, 3.858117355875043e-02,
                  -8.728545924521301e-03, -9.401196355382909e-03, 4.024549377076924e-03, 1.512806105865091e-03,
                  -9.576734877247042e-04, -1.303457547418901e-04, 1.100385844683190e-04};
    } else if (digits == 12) {
        p.c = 27.870000839233398;
        p.n_diff = 26;
        p.poly = {6.262472576363448e-03, -5.605742936112479e-02, 2.185890864792949e-01, -4.717350304955679e-01,
                  5.669680214206270e-01, -2.511606878849214e-01, -2.744523658778361e-01, 4.582527599363415e-01,
                  -1.397724810121539e-01, -2.131762135835757e-01, 1.995489373508990e-01, 1.793390341864239e-02,
                  -1.035055132403432e-01, 3.035606831075176e-02, 3.153931762550532e-02, -2.033178627450288e-02,
                  -5.406682731236552e-03, 7.543645573618463e-03, 1.437788047407851e-05, -1.928370882351732e-03,
                  2.891658777328665e-04, 3.332996162099811e-04, -8.397699195938912e-05, -3.015837377517983e-05,
                  9.640642701924662e-06};
    } else {
        throw std::runtime_error("digits is not supported");
    }
    const Prolate0& f = prolate0(p.c);
    p.lambda = f.lambda();
    p.C0 = f.int0r(1.0);
    p.diff0 = -2.0 * M_PI * f.intx2_0r(1.0) / p.C0 * p.c * p.c;  // src/tree.cpp:41-42
    p.psi0_0 = f.eval(0.0);
    p.n_window = (int)std::floor(2 * p.c / M_PI);  // src/tree.cpp:58
    return cache.emplace(digits, p).first->second;
}

double residual_poly(const PswfParams& p, double x) {
    const double u = (x - 0.5) * 2.0;
    if (std::fabs(u) >= 1.0) return 0.0;
    double v = 0;
    for (int i = (int)p.poly.size() - 1; i >= 0; --i) v = v * u + p.poly[i];
    return v;
}

}  // namespace oracle
