// FP64 Ewald summation on the GPU (src/ewald.cpp:17-308 of the reference):
// the ground truth the DMK energies are reported against at sizes where the
// CPU sum is too slow (N >= 1e5).
//
//   E = [1/2 sum_{pairs, r < r_c} q_i q_j erfc(alpha r)/r
//        + 2 pi/V sum_{k != 0, |n_a| <= n} e^{-k^2/4 alpha^2}/k^2 |S(k)|^2
//        - alpha/sqrt(pi) sum q^2] / eps,       r_c = s/alpha, n = ceil(2 s alpha L / 2 pi)
//
// S(k) = sum_j q_j e^{-i k.x_j} is kept on the half cube k_z >= 0 (the k_z < 0
// half is its conjugate; weights 2 off the k_z = 0 plane are folded into the
// kernel table g), element ix + d (iy + d iz), iz = k_z / dk.  Real-space sums
// use the minimum image (r_c < L/2 is required).  A trial move's energy change
// is exact: the real-space part from the moved particle's pair terms, the
// k-space part as g (2 Re(conj(S) dS) + |dS|^2) with dS the change of S.
#include <hip/hip_runtime.h>

#include "dmk_common.hpp"
#include "ewald_kernels.hpp"

namespace hpdmk_amd {

namespace {

constexpr int ETB = 256;   // threads per block
constexpr int kTileY = 16; // k_y values per structure-factor block
constexpr int kChunk = 128;

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// block-wide sum added to *out with one atomic per block
__device__ __forceinline__ void block_add(double v, double* out) {
    __shared__ double red[ETB / 64];
    v = wave_sum_d(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < ETB / 64; ++w) s += red[w];
        atomicAdd(out, s);
    }
}

__device__ __forceinline__ double min_image(double a, double L) { return a - L * rint(a / L); }

__device__ __forceinline__ double short_kernel(double r2, const EwaldDev& e) {
    if (!(r2 > 0.0) || r2 >= e.rc * e.rc) return 0.0;
    const double r = sqrt(r2);
    return erfc(e.alpha * r) / r;
}

// S(k) for k_z = blockIdx.y dk, k_y in a tile of kTileY, k_x = thread
__global__ void __launch_bounds__(ETB) k_ewald_sk(EwaldDev e) {
    __shared__ double px[kChunk], py[kChunk], pz[kChunk], pq[kChunk];
    __shared__ Cplx<double> yz[kChunk][kTileY];
    const int d = e.d, nk = e.nk;
    const int iy0 = blockIdx.x * kTileY, iz = blockIdx.y, ix = threadIdx.x;
    const double kx = (ix - nk) * e.dk, kz = iz * e.dk;
    double ar[kTileY], ai[kTileY];
#pragma unroll
    for (int t = 0; t < kTileY; ++t) ar[t] = ai[t] = 0.0;
    for (int c0 = 0; c0 < e.n; c0 += kChunk) {
        const int nc = min(kChunk, e.n - c0);
        __syncthreads();
        for (int j = threadIdx.x; j < nc; j += ETB) {
            px[j] = e.x[c0 + j];
            py[j] = e.y[c0 + j];
            pz[j] = e.z[c0 + j];
            pq[j] = e.q[c0 + j];
        }
        __syncthreads();
        for (int w = threadIdx.x; w < nc * kTileY; w += ETB) {
            const int j = w / kTileY, t = w % kTileY;
            const double ky = (iy0 + t - nk) * e.dk;
            double s, c;
            sincos(ky * py[j] + kz * pz[j], &s, &c);
            yz[j][t] = {pq[j] * c, -pq[j] * s};
        }
        __syncthreads();
        if (ix < d) {
            for (int j = 0; j < nc; ++j) {
                double s, c;
                sincos(kx * px[j], &s, &c);
                const double xr = c, xi = -s;
#pragma unroll
                for (int t = 0; t < kTileY; ++t) {
                    const Cplx<double> v = yz[j][t];
                    ar[t] += xr * v.x - xi * v.y;
                    ai[t] += xr * v.y + xi * v.x;
                }
            }
        }
    }
    if (ix < d)
#pragma unroll
        for (int t = 0; t < kTileY; ++t)
            if (iy0 + t < d) e.S[ix + (size_t)d * (iy0 + t + (size_t)d * iz)] = {ar[t], ai[t]};
}

__global__ void __launch_bounds__(ETB) k_ewald_long(EwaldDev e, double* out) {
    double v = 0;
    for (size_t m = blockIdx.x * (size_t)ETB + threadIdx.x; m < e.K; m += (size_t)gridDim.x * ETB) {
        const Cplx<double> s = e.S[m];
        v += e.g[m] * (s.x * s.x + s.y * s.y);
    }
    block_add(v, out);
}

// 1/2 sum_i sum_{j != i} q_i q_j erfc(alpha r)/r, minimum image, r < r_c
__global__ void __launch_bounds__(ETB) k_ewald_short(EwaldDev e, double* out) {
    __shared__ double px[ETB], py[ETB], pz[ETB], pq[ETB];
    const int i = blockIdx.x * ETB + threadIdx.x;
    const bool on = i < e.n;
    const double xi = on ? e.x[i] : 0, yi = on ? e.y[i] : 0, zi = on ? e.z[i] : 0, qi = on ? e.q[i] : 0;
    double v = 0;
    for (int c0 = 0; c0 < e.n; c0 += ETB) {
        __syncthreads();
        const int j = c0 + threadIdx.x;
        if (j < e.n) {
            px[threadIdx.x] = e.x[j];
            py[threadIdx.x] = e.y[j];
            pz[threadIdx.x] = e.z[j];
            pq[threadIdx.x] = e.q[j];
        }
        __syncthreads();
        const int nc = min(ETB, e.n - c0);
        for (int k = 0; k < nc; ++k) {
            const double dx = min_image(xi - px[k], e.L), dy = min_image(yi - py[k], e.L),
                         dz = min_image(zi - pz[k], e.L);
            v += pq[k] * short_kernel(dx * dx + dy * dy + dz * dz, e);
        }
    }
    block_add(0.5 * qi * v, out);
}

// potential at points (src/ewald.cpp:281-308): short part over all particles
// at r > 0 within r_c, long part 4 pi/V sum g Re(S e^{ik.x}); one block per point
__global__ void __launch_bounds__(ETB) k_ewald_potential(EwaldDev e, const double* pts, double* phi) {
    __shared__ double red[ETB / 64];
    const int p = blockIdx.x;
    const double x = pts[3 * p], y = pts[3 * p + 1], z = pts[3 * p + 2];
    double vs = 0, vl = 0;
    for (int j = threadIdx.x; j < e.n; j += ETB) {
        const double dx = min_image(x - e.x[j], e.L), dy = min_image(y - e.y[j], e.L), dz = min_image(z - e.z[j], e.L);
        vs += e.q[j] * short_kernel(dx * dx + dy * dy + dz * dz, e);
    }
    const int d = e.d;
    for (size_t m = threadIdx.x; m < e.K; m += ETB) {
        const int ix = m % d, iy = (m / d) % d, iz = m / ((size_t)d * d);
        double s, c;
        sincos(((ix - e.nk) * x + (iy - e.nk) * y + iz * z) * e.dk, &s, &c);
        const Cplx<double> S = e.S[m];
        vl += e.g[m] * (S.x * c - S.y * s);
    }
    double v = wave_sum_d(vs + vl * 4.0 * M_PI / e.V);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < ETB / 64; ++w) t += red[w];
        phi[p] = t / e.eps;
    }
}

// phase rows of one move: [endpoint t, o][axis][0..d): e^{-i (a - nk) dk x_axis}
// (z rows use a = nk .. 2nk, i.e. k_z >= 0, stored from index 0)
__global__ void k_ewald_rows(EwaldDev e, EwaldMove mv, int k) {
    const int p = (int)mv.idx[k];
    const double o[3] = {e.x[p], e.y[p], e.z[p]};
    double t[3];
    for (int a = 0; a < 3; ++a) {
        double v = o[a] + mv.d[3 * k + a];
        int guard = 0;  // my_mod of include/utils.hpp:96-105
        while (v < 0 && guard++ < 1024) v += e.L;
        while (v >= e.L && guard++ < 2048) v -= e.L;
        t[a] = v;
    }
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        mv.tgt[0] = t[0];
        mv.tgt[1] = t[1];
        mv.tgt[2] = t[2];
    }
    const int d = e.d;
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < 6 * d; w += gridDim.x * blockDim.x) {
        const int ep = w / (3 * d), ax = (w / d) % 3, a = w % d;
        const double m = ax == 2 ? (double)a : (double)(a - e.nk);
        if (ax == 2 && a > e.nk) continue;
        double s, c;
        sincos(m * e.dk * (ep == 0 ? t[ax] : o[ax]), &s, &c);
        mv.rows[w] = {c, -s};
    }
}

// dS(k) = q_p (e^{-ik.t} - e^{-ik.o}) from the rows
__device__ __forceinline__ Cplx<double> delta_s(const EwaldDev& e, const EwaldMove& mv, double q, size_t m) {
    const int d = e.d;
    const int ix = m % d, iy = (m / d) % d, iz = m / ((size_t)d * d);
    const Cplx<double>* R = mv.rows;
    auto prod = [&](int ep) {
        const Cplx<double> X = R[(ep * 3 + 0) * d + ix], Y = R[(ep * 3 + 1) * d + iy], Z = R[(ep * 3 + 2) * d + iz];
        const double xr = X.x * Y.x - X.y * Y.y, xi = X.x * Y.y + X.y * Y.x;
        return Cplx<double>{xr * Z.x - xi * Z.y, xr * Z.y + xi * Z.x};
    };
    const Cplx<double> a = prod(0), b = prod(1);
    return {q * (a.x - b.x), q * (a.y - b.y)};
}

__global__ void __launch_bounds__(ETB) k_ewald_dlong(EwaldDev e, EwaldMove mv, int k) {
    const double q = e.q[mv.idx[k]];
    double v = 0;
    for (size_t m = blockIdx.x * (size_t)ETB + threadIdx.x; m < e.K; m += (size_t)gridDim.x * ETB) {
        const Cplx<double> ds = delta_s(e, mv, q, m), s = e.S[m];
        v += e.g[m] * (2.0 * (s.x * ds.x + s.y * ds.y) + ds.x * ds.x + ds.y * ds.y);
    }
    block_add(v * 2.0 * M_PI / e.V, mv.acc);
}

__global__ void __launch_bounds__(ETB) k_ewald_dshort(EwaldDev e, EwaldMove mv, int k) {
    const int p = (int)mv.idx[k];
    const double ox = e.x[p], oy = e.y[p], oz = e.z[p];
    const double tx = mv.tgt[0], ty = mv.tgt[1], tz = mv.tgt[2];
    double v = 0;
    for (int j = blockIdx.x * ETB + threadIdx.x; j < e.n; j += gridDim.x * ETB) {
        if (j == p) continue;
        const double xj = e.x[j], yj = e.y[j], zj = e.z[j];
        double ax = min_image(tx - xj, e.L), ay = min_image(ty - yj, e.L), az = min_image(tz - zj, e.L);
        double bx = min_image(ox - xj, e.L), by = min_image(oy - yj, e.L), bz = min_image(oz - zj, e.L);
        v += e.q[j] * (short_kernel(ax * ax + ay * ay + az * az, e) - short_kernel(bx * bx + by * by + bz * bz, e));
    }
    block_add(v * e.q[p], mv.acc);
}

__global__ void k_ewald_finish(EwaldDev e, EwaldMove mv, int k) {
    mv.dE[k] = *mv.acc / e.eps;
    *mv.acc = 0.0;
}

__global__ void __launch_bounds__(ETB) k_ewald_apply(EwaldDev e, EwaldMove mv, int k) {
    const int p = (int)mv.idx[k];
    const double q = e.q[p];
    for (size_t m = blockIdx.x * (size_t)ETB + threadIdx.x; m < e.K; m += (size_t)gridDim.x * ETB) {
        const Cplx<double> ds = delta_s(e, mv, q, m);
        e.S[m].x += ds.x;
        e.S[m].y += ds.y;
    }
}

__global__ void k_ewald_move(EwaldDev e, EwaldMove mv, int k) {
    const int p = (int)mv.idx[k];
    e.x[p] = mv.tgt[0];
    e.y[p] = mv.tgt[1];
    e.z[p] = mv.tgt[2];
}

}  // namespace

void ewald_structure_factor(const EwaldDev& e, hipStream_t st) {
    hipLaunchKernelGGL(k_ewald_sk, dim3((e.d + kTileY - 1) / kTileY, e.nk + 1), dim3(ETB), 0, st, e);
}

void ewald_energy_parts(const EwaldDev& e, double* out2, hipStream_t st) {
    hipLaunchKernelGGL(k_ewald_short, dim3((e.n + ETB - 1) / ETB), dim3(ETB), 0, st, e, out2);
    hipLaunchKernelGGL(k_ewald_long, dim3(1024), dim3(ETB), 0, st, e, out2 + 1);
}

void ewald_potential(const EwaldDev& e, int n_pts, const double* pts, double* phi, hipStream_t st) {
    if (n_pts > 0) hipLaunchKernelGGL(k_ewald_potential, dim3(n_pts), dim3(ETB), 0, st, e, pts, phi);
}

void ewald_shift(const EwaldDev& e, const EwaldMove& mv, int k, bool apply, hipStream_t st) {
    hipLaunchKernelGGL(k_ewald_rows, dim3(4), dim3(ETB), 0, st, e, mv, k);
    hipLaunchKernelGGL(k_ewald_dshort, dim3(std::min(1024, (e.n + ETB - 1) / ETB)), dim3(ETB), 0, st, e, mv, k);
    hipLaunchKernelGGL(k_ewald_dlong, dim3(1024), dim3(ETB), 0, st, e, mv, k);
    hipLaunchKernelGGL(k_ewald_finish, dim3(1), dim3(1), 0, st, e, mv, k);
    if (apply) {
        hipLaunchKernelGGL(k_ewald_apply, dim3(1024), dim3(ETB), 0, st, e, mv, k);
        hipLaunchKernelGGL(k_ewald_move, dim3(1), dim3(1), 0, st, e, mv, k);
    }
}

}  // namespace hpdmk_amd
