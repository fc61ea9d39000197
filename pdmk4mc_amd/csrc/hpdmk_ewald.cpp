// C ABI of the FP64 Ewald reporter (include/hpdmk.h): the reference's Ewald
// oracle (include/ewald.hpp:14-57, src/ewald.cpp:17-308) on the GPU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "ewald_kernels.hpp"
#include "hpdmk_handle.hpp"

using namespace hpdmk_amd;
using namespace hpdmk_amd::capi;

namespace {

struct EwaldHandle {
    EwaldDev e{};
    double Q2 = 0;  // sum q^2
    hipStream_t stream = nullptr;
    std::vector<void*> owned;
    double* red = nullptr;     // 2 doubles
    EwaldMove mv{};
    int move_cap = 0;
    long* idx = nullptr;
    double* dmv = nullptr;

    ~EwaldHandle() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : owned) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

EwaldHandle* EH(hpdmk_ewald h) {
    if (!h) throw HpdmkError(HPDMK_EINVAL, "null hpdmk_ewald handle");
    return static_cast<EwaldHandle*>(h);
}

}  // namespace

extern "C" {

int hpdmk_ewald_create(double L, double s, double alpha, double eps, int n, const double* q, const double* r,
                       hpdmk_ewald* out) {
    return guard([&] {
        if (!out || !q || !r || n < 1) throw HpdmkError(HPDMK_EINVAL, "null argument or no particles");
        *out = nullptr;
        if (!(L > 0) || !(s > 0) || !(alpha > 0) || !(eps > 0)) throw HpdmkError(HPDMK_EINVAL, "L, s, alpha, eps must be > 0");
        auto h = std::make_unique<EwaldHandle>();
        EwaldDev& e = h->e;
        e.n = n;
        e.L = L;
        e.alpha = alpha;
        e.eps = eps;
        e.V = L * L * L;
        e.rc = s / alpha;                                          // src/ewald.cpp:18
        e.dk = 2 * M_PI / L;
        e.nk = (int)std::ceil(2 * s * alpha / e.dk);               // src/ewald.cpp:29
        e.d = 2 * e.nk + 1;
        e.K = (size_t)e.d * e.d * (e.nk + 1);
        if (e.d > 256) throw HpdmkError(HPDMK_EINVAL, "more than 256 modes per axis (2 s alpha L / 2 pi > 127)");
        if (!(e.rc < 0.5 * L)) throw HpdmkError(HPDMK_EINVAL, "real-space cutoff s/alpha must be below L/2");
        std::vector<double> x(n), y(n), z(n);
        for (int i = 0; i < n; ++i) {
            x[i] = r[3 * i];
            y[i] = r[3 * i + 1];
            z[i] = r[3 * i + 2];
            h->Q2 += q[i] * q[i];
        }
        std::vector<double> g(e.K, 0.0);
        for (int iz = 0; iz <= e.nk; ++iz)  // src/ewald.cpp:52-71, half cube with weights
            for (int iy = 0; iy < e.d; ++iy)
                for (int ix = 0; ix < e.d; ++ix) {
                    const double kx = (ix - e.nk) * e.dk, ky = (iy - e.nk) * e.dk, kz = iz * e.dk;
                    const double k2 = kx * kx + ky * ky + kz * kz;
                    if (k2 == 0.0) continue;
                    g[ix + (size_t)e.d * (iy + (size_t)e.d * iz)] =
                        (iz == 0 ? 1.0 : 2.0) * std::exp(-k2 / (4 * alpha * alpha)) / k2;
                }
        HIPX(hipStreamCreate(&h->stream));
        auto& o = h->owned;
        e.x = dupload(x, o);
        e.y = dupload(y, o);
        e.z = dupload(z, o);
        std::vector<double> qv(q, q + n);
        e.q = dupload(qv, o);
        e.g = dupload(g, o);
        e.S = dalloc<Cplx<double>>(e.K, o);
        h->red = dalloc<double>(2, o);
        h->mv.rows = dalloc<Cplx<double>>(6 * (size_t)e.d, o);
        h->mv.tgt = dalloc<double>(3, o);
        h->mv.acc = dalloc<double>(1, o);
        HIPX(hipMemset(h->mv.acc, 0, sizeof(double)));
        ewald_structure_factor(e, h->stream);
        HIPX(hipGetLastError());
        HIPX(hipStreamSynchronize(h->stream));
        *out = h.release();
    });
}

int hpdmk_ewald_destroy(hpdmk_ewald h) {
    return guard([&] { delete EH(h); });
}

int hpdmk_ewald_compute_energy(hpdmk_ewald hh, double* E, double* parts) {
    return guard([&] {
        EwaldHandle* h = EH(hh);
        HIPX(hipMemsetAsync(h->red, 0, 2 * sizeof(double), h->stream));
        ewald_energy_parts(h->e, h->red, h->stream);
        HIPX(hipGetLastError());
        double v[2];
        HIPX(hipMemcpyAsync(v, h->red, sizeof(v), hipMemcpyDeviceToHost, h->stream));
        HIPX(hipStreamSynchronize(h->stream));
        const double es = v[0], el = v[1] * 2 * M_PI / h->e.V, eself = -h->Q2 * h->e.alpha / std::sqrt(M_PI);
        if (parts) {
            parts[0] = es / h->e.eps;
            parts[1] = el / h->e.eps;
            parts[2] = eself / h->e.eps;
        }
        if (E) *E = (es + el + eself) / h->e.eps;  // src/ewald.cpp:178-191
    });
}

int hpdmk_ewald_compute_potential(hpdmk_ewald hh, int n_pts, const double* pts, double* phi) {
    return guard([&] {
        EwaldHandle* h = EH(hh);
        if (n_pts < 1 || !pts || !phi) throw HpdmkError(HPDMK_EINVAL, "no points");
        std::vector<void*> tmp;
        std::vector<double> p(pts, pts + 3 * (size_t)n_pts);
        double* dp = nullptr;
        double* dphi = nullptr;
        try {
            dp = dupload(p, tmp);
            dphi = dalloc<double>(n_pts, tmp);
            ewald_potential(h->e, n_pts, dp, dphi, h->stream);
            HIPX(hipGetLastError());
            HIPX(hipMemcpyAsync(phi, dphi, sizeof(double) * n_pts, hipMemcpyDeviceToHost, h->stream));
            HIPX(hipStreamSynchronize(h->stream));
        } catch (...) {
            for (void* q : tmp) (void)hipFree(q);
            throw;
        }
        for (void* q : tmp) (void)hipFree(q);
    });
}

int hpdmk_ewald_shift_chain(hpdmk_ewald hh, int n_moves, const long* idx, const double* d, const int* apply,
                            double* dE) {
    return guard([&] {
        EwaldHandle* h = EH(hh);
        if (n_moves < 1 || !idx || !d || !dE) throw HpdmkError(HPDMK_EINVAL, "empty or null move list");
        for (int k = 0; k < n_moves; ++k)
            if (idx[k] < 0 || idx[k] >= h->e.n) throw HpdmkError(HPDMK_EINVAL, "particle index out of range");
        if (n_moves > h->move_cap) {
            h->idx = dalloc<long>(n_moves, h->owned);
            h->dmv = dalloc<double>(3 * (size_t)n_moves, h->owned);
            h->mv.dE = dalloc<double>(n_moves, h->owned);
            h->move_cap = n_moves;
        }
        HIPX(hipMemcpyAsync(h->idx, idx, sizeof(long) * n_moves, hipMemcpyHostToDevice, h->stream));
        HIPX(hipMemcpyAsync(h->dmv, d, sizeof(double) * 3 * n_moves, hipMemcpyHostToDevice, h->stream));
        h->mv.idx = h->idx;
        h->mv.d = h->dmv;
        for (int k = 0; k < n_moves; ++k) ewald_shift(h->e, h->mv, k, apply && apply[k], h->stream);
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(dE, h->mv.dE, sizeof(double) * n_moves, hipMemcpyDeviceToHost, h->stream));
        HIPX(hipStreamSynchronize(h->stream));
    });
}

}  // extern "C"
