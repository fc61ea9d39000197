// Outgoing plane-wave expansions of the DIRECT build in mixed precision as
// FP32 matrix-core (MFMA) contractions on gfx950 (src/planewave.cpp:24-84,
// include/nudft.hpp:66-123).
//
// For a box with particles p (charges q_p) the half-grid expansion is
//   rho[i + d (j + d k)] = sum_p X[i, p] YZ[p, j + d k],
//   X[i, p] = q_p e^{-i (i-n) dk x_p},  YZ[p, c] = e^{-i (j-n) dk y_p} e^{-i (k-n) dk z_p},
// a complex GEMM with M = d rows (two 16-row tiles), N = d (n+1) columns
// (16-column tiles) and K = the box's particles.  Each complex tile product
// is four v_mfma_f32_16x16x4f32 (re re, -im im, re im, im re).  The 1-D
// phase rows of 64 particles at a time are formed in FP64 (one sincos per
// particle and axis, powers by complex multiplication) and staged in LDS as
// FP32; the FP32 tile sums are added into FP64 accumulators every 16
// particles, so the sums accumulate in FP64.
//
// Work items are (box, run of its descendant leaves): a box whose leaves are
// split over several items (the root and the large boxes) adds its partial
// sums into an FP64 buffer with atomics, and k_pw_mfma_store converts them.
// Only the live slots of each leaf ([start, start + count)) are read; holes
// left by moves carry q = 0.
#include "dmk_device.hpp"
#include "dmk_kernels.hpp"
#include "dmk_nudft_mfma.hpp"

namespace hpdmk_amd {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kKC = 64;        // particles per staged chunk (sixteen K-steps of 4)
constexpr int kFlushK = 4;     // K-steps between FP64 flushes (16 particles)
constexpr int kRowPad = 32;    // padded length of the x and y rows (d <= 25, two M tiles)
constexpr int kZPad = 16;      // padded length of the z rows (n + 1 <= 13)
constexpr int kMaxLeaves = 256;  // leaves of one work item staged in LDS

}  // namespace

// TPW: column tiles per wave; NWV: waves per block.  d <= 25 (digits <= 6).
template <int TPW, int NWV>
__global__ void __launch_bounds__(64 * NWV) k_pw_mfma(DevTree<float> t, const int* __restrict__ leaves,
                                                      const PwWork* __restrict__ work, double* __restrict__ acc_box,
                                                      double* __restrict__ acc_root) {
    __shared__ float2 X[kKC][kRowPad], Y[kKC][kRowPad], Z[kKC][kZPad];
    __shared__ int lstart[kMaxLeaves], lpre[kMaxLeaves + 1];
    __shared__ int slots[kKC];
    const PwWork w = work[blockIdx.x];
    const bool root = w.slot < 0;
    const int n = root ? t.n_window : t.n_diff;
    const int d = 2 * n + 1, ncol = d * (n + 1), ntile = (ncol + 15) / 16;
    const int l = root ? 0 : t.level[w.node];
    const double dk = t.delta_k[l];
    const int nl = w.leaf_hi - w.leaf_lo;  // <= kMaxLeaves (host splits longer runs)
    // prefix of the live counts of the item's leaves (loads in parallel, sum by one thread from LDS)
    for (int k = threadIdx.x; k < nl; k += blockDim.x) {
        const int lf = leaves[w.leaf_lo + k];
        lstart[k] = t.slot_start[lf];
        lpre[k + 1] = t.slot_count[lf];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        lpre[0] = 0;
        for (int k = 1; k <= nl; ++k) lpre[k] += lpre[k - 1];
    }
    __syncthreads();
    const int total = lpre[nl];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int col_in = lane & 15, kq = lane >> 4;  // B operand column; K index of A and B
    // this lane's columns (j, k) in each of the wave's tiles
    int cj[TPW], ck[TPW];
    bool cv[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        const int c = (wv + NWV * u) * 16 + col_in;
        cv[u] = (wv + NWV * u) < ntile && c < ncol;
        cj[u] = cv[u] ? c % d : 0;
        ck[u] = cv[u] ? c / d : 0;
    }
    f32x4 ar[TPW][2], ai[TPW][2];
    double sr[TPW][2][4], si[TPW][2][4];
#pragma unroll
    for (int u = 0; u < TPW; ++u)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            ar[u][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
            ai[u][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 4; ++r) sr[u][mt][r] = si[u][mt][r] = 0.0;
        }
    for (int base = 0; base < total; base += kKC) {
        __syncthreads();  // the previous chunk's rows are consumed
        if (threadIdx.x < kKC) {  // slot of the chunk's particle (binary search in the prefix)
            const int u = base + threadIdx.x;
            int s = -1;
            if (u < total) {
                int lo = 0, hi = nl - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (lpre[mid] <= u)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                s = lstart[lo] + (u - lpre[lo]);
            }
            slots[threadIdx.x] = s;
        }
        __syncthreads();
        if (threadIdx.x < 3 * kKC) {  // one (particle, axis) row per thread
            const int p = threadIdx.x % kKC, ax = threadIdx.x / kKC;
            const int s = slots[p];
            const double q = s >= 0 ? t.pq[s] : 0.0;
            const double x = s >= 0 ? (ax == 0 ? t.px[s] : (ax == 1 ? t.py[s] : t.pz[s])) : 0.0;
            double sn, cs;
            sincos(dk * x, &sn, &cs);
            const double br = cs, bi = -sn;  // e^{-i dk x}
            const double sc = ax == 0 ? q : 1.0;
            float2* row = ax == 0 ? X[p] : (ax == 1 ? Y[p] : Z[p]);
            double wr = sc, wi = 0.0;  // sc e^{-i m dk x}, m = 0, 1, ...
            if (ax < 2) {
                row[n] = make_float2((float)wr, 0.f);
                for (int m = 1; m <= n; ++m) {
                    const double tr = wr * br - wi * bi;
                    wi = wr * bi + wi * br;
                    wr = tr;
                    row[n + m] = make_float2((float)wr, (float)wi);   // e^{-i m dk x}
                    row[n - m] = make_float2((float)wr, (float)-wi);  // e^{+i m dk x}
                }
                for (int m = d; m < kRowPad; ++m) row[m] = make_float2(0.f, 0.f);
            } else {  // z: k = 0..n is (k - n) <= 0, e^{+i (n - k) dk z}
                row[n] = make_float2(1.f, 0.f);
                for (int m = 1; m <= n; ++m) {
                    const double tr = wr * br - wi * bi;
                    wi = wr * bi + wi * br;
                    wr = tr;
                    row[n - m] = make_float2((float)wr, (float)-wi);
                }
                for (int m = n + 1; m < kZPad; ++m) row[m] = make_float2(0.f, 0.f);
            }
        }
        __syncthreads();
        const int nks = min(kKC, total - base + 3) / 4;  // K-steps holding particles (uniform)
        for (int ks = 0; ks < nks; ++ks) {
            const int p = 4 * ks + kq;
            const float2 a0 = X[p][col_in], a1 = X[p][16 + col_in];  // rows i = col_in (+16) of this K slice
#pragma unroll
            for (int u = 0; u < TPW; ++u) {
                if (!((wv + NWV * u) < ntile)) continue;  // uniform per wave
                float2 yz = make_float2(0.f, 0.f);
                if (cv[u]) {
                    const float2 y = Y[p][cj[u]], z = Z[p][ck[u]];
                    yz = make_float2(y.x * z.x - y.y * z.y, y.x * z.y + y.y * z.x);
                }
                ar[u][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.x, yz.x, ar[u][0], 0, 0, 0);
                ar[u][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a0.y, yz.y, ar[u][0], 0, 0, 0);
                ai[u][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.x, yz.y, ai[u][0], 0, 0, 0);
                ai[u][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.y, yz.x, ai[u][0], 0, 0, 0);
                ar[u][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.x, yz.x, ar[u][1], 0, 0, 0);
                ar[u][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a1.y, yz.y, ar[u][1], 0, 0, 0);
                ai[u][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.x, yz.y, ai[u][1], 0, 0, 0);
                ai[u][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.y, yz.x, ai[u][1], 0, 0, 0);
            }
            if ((ks + 1) % kFlushK == 0 || ks + 1 == nks) {  // FP32 tile sums into FP64 (uniform)
#pragma unroll
                for (int u = 0; u < TPW; ++u)
#pragma unroll
                    for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            sr[u][mt][r] += (double)ar[u][mt][r];
                            si[u][mt][r] += (double)ai[u][mt][r];
                        }
                        ar[u][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
                        ai[u][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
                    }
            }
        }
    }
    // D layout of v_mfma_f32_16x16x4f32: lane holds rows 4 (lane / 16) + r of column lane % 16
    const int M = root ? t.Mw : t.Md;
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        if (!cv[u]) continue;
        const int c = (wv + NWV * u) * 16 + col_in;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * mt + 4 * kq + r;
                if (i >= d) continue;
                const int e = i + d * c;
                if (w.acc >= 0 || root) {
                    double* a = root ? acc_root : acc_box + (size_t)w.acc * 2 * M;
                    atomicAdd(a + 2 * e, sr[u][mt][r]);
                    atomicAdd(a + 2 * e + 1, si[u][mt][r]);
                } else {
                    t.out[(size_t)w.slot * M + e] = {(float)sr[u][mt][r], (float)si[u][mt][r]};
                }
            }
    }
}

// boxes summed over several work items: FP64 sums -> stored expansion
__global__ void k_pw_mfma_store(DevTree<float> t, const int* __restrict__ acc_slots, const double* __restrict__ acc_box) {
    const int b = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= t.Md) return;
    const double* a = acc_box + (size_t)b * 2 * t.Md;
    t.out[(size_t)acc_slots[b] * t.Md + e] = {(float)a[2 * e], (float)a[2 * e + 1]};
}

__global__ void k_zero_n(double* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_store_root_f(DevTree<float> t, const double* acc) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < t.Mw) t.root[e] = {(float)acc[2 * e], (float)acc[2 * e + 1]};
}

void launch_outgoing_mfma(const DevTree<float>& t, const int* leaves, const PwMfmaPlan& pl, double* acc_root,
                          hipStream_t st) {
    const size_t nz = 2 * (size_t)t.Md * pl.n_acc;
    hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, st, acc_root, (size_t)2 * t.Mw);
    if (nz) hipLaunchKernelGGL(k_zero_n, dim3(2048), dim3(256), 0, st, pl.acc_box, nz);
    // 8 waves x 3 column tiles cover the 21 tiles of digits 6 (d = 25) and the 10 of the digits-6 root
    if (pl.n_work > 0)
        hipLaunchKernelGGL((k_pw_mfma<3, 8>), dim3(pl.n_work), dim3(512), 0, st, t, leaves, pl.work, pl.acc_box,
                           acc_root);
    hipLaunchKernelGGL(k_store_root_f, dim3((t.Mw + 255) / 256), dim3(256), 0, st, t, (const double*)acc_root);
    if (pl.n_acc > 0)
        hipLaunchKernelGGL(k_pw_mfma_store, dim3((t.Md + 255) / 256, pl.n_acc), dim3(256), 0, st, t, pl.acc_slots,
                           (const double*)pl.acc_box);
}

}  // namespace hpdmk_amd
