// PROXY initialisation of the outgoing expansions on gfx950 (the reference's
// Chebyshev upward pass, src/upward_pass.cpp:206-322), on the FP64 matrix
// cores (v_mfma_f64_16x16x4f64).
//
//   k_proxy_leaf  particles -> Chebyshev coefficients
//                 C[a + p b + p^2 c] = sum_j q_j T_a(x^) T_b(y^) T_c(z^),
//                 x^ = 2 (x - centre) / box size, for every box of level D-2
//                 and every shallower leaf (upward_pass.cpp:109-149,
//                 270-287): per box the product W^T Tx of the particles'
//                 rows W[j][b + p c] = q_j T_b(y^_j) T_c(z^_j) and Tx[j][a],
//                 formed chunk by chunk in LDS and fed to the matrix cores
//   child -> parent (upward_pass.cpp:19-64, 289-321) and coefficients ->
//                 half-grid plane waves rho[k] = e^{-i k.c} sum P[kx,a]
//                 P[ky,b] P[kz,c] C[a,b,c] (proxycharge2pw + decenter_phase,
//                 upward_pass.cpp:67-106, 183-203): one mode product per
//                 axis, each a batched GEMM (k_gemm64) through scratch
//                 buffers, boxes in chunks that keep the scratch small.
//                 Complex operands are stacked real planes, so a complex
//                 product is one real GEMM: [[P_re, -P_im], [P_im, P_re]]
//                 times [X_re; X_im].  The z pass of the transfer sums the
//                 eight children of a parent in one GEMM of depth 8p.
//   k_pw_out      the de-centring phase and the store of the expansions
//   k_root_sum    the root expansion from the level-2 boxes' coefficients at
//                 the root's wavenumbers (their plane waves, shifted to the
//                 boxes' centres and summed), when every particle lies in a
//                 level-2 box; else the root comes from the particles
//
// Layouts: coefficients of pw slot s at coef + s p^3 (a fastest); the
// matrices of ProxyArgs (dmk_proxy.hpp).
#include <algorithm>

#include "dmk_device.hpp"
#include "dmk_kernels.hpp"
#include "dmk_proxy.hpp"

namespace hpdmk_amd {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4f64 lane layout: A operand A[l % 16][l / 16], B operand
// B[l / 16][l % 16], result D[4 j + l / 16][l % 16] in element j
__device__ __forceinline__ f64x4 mfma64(double a, double b, f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void cheb_row(double x, double* t, int p) {
    double t0 = 1.0, t1 = x;
    t[0] = 1.0;
    t[1] = x;
    for (int k = 2; k < p; ++k) {
        const double t2 = 2.0 * x * t1 - t0;
        t[k] = t2;
        t0 = t1;
        t1 = t2;
    }
}

template <typename T>
__device__ __forceinline__ double box_centre(const DevTree<T>& t, int node, int axis) {
    const int l = t.level[node];
    return (t.cell[3 * node + axis] + 0.5) * t.boxsize[l];
}

// start of batch entry (b0, b1) of an operand
__device__ __forceinline__ long gemm_base(const GemmOp& o, int b0, int b1) {
    return (long)(o.tab ? o.tab[b0] : b0) * o.s0 + (long)b1 * o.s1;
}

// offset of row r of an operand (from its batch start), or -1 in an all-zero segment
__device__ __forceinline__ long gemm_row(const GemmOp& o, int b0, int r) {
    const int s = r / o.seg, rr = r - s * o.seg;
    long so = s;
    if (o.stab) {
        so = o.stab[b0 * o.nseg + s];
        if (so < 0) return -1;
    }
    return so * o.sstride + (o.col ? (long)rr : (long)rr * o.ld);
}

}  // namespace

// Batched GEMM on the FP64 matrix cores: block (batch entry, column tile,
// row tile), four waves of 16 x 32 each (two 16 x 16 tiles), WM of them
// along the rows (block tile 16 WM x 128 / WM); operands read from memory
// (L2) four K-steps at a time, all loads in flight before the products.
template <int WM>
__global__ void __launch_bounds__(256) k_gemm64(GemmPass g) {
    constexpr int BM = 16 * WM, BN = 32 * (4 / WM);
    const int b = blockIdx.x, b0 = b / g.inner, b1 = b - b0 * g.inner;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int m0 = blockIdx.z * BM + (wv % WM) * 16, n0 = blockIdx.y * BN + (wv / WM) * 32;
    const double* A = g.a.p + gemm_base(g.a, b0, b1);
    const double* B = g.b.p + gemm_base(g.b, b0, b1);
    double* C = g.c.p + gemm_base(g.c, b0, b1);
    const int i = m0 + li;
    const long ar = i < g.M ? gemm_row(g.a, b0, i) : -1;
    const long astep = g.a.col ? g.a.ld : 1;
    const long bstep = g.b.col ? g.b.ld : 1;
    f64x4 acc[2] = {f64x4{0.0, 0.0, 0.0, 0.0}, f64x4{0.0, 0.0, 0.0, 0.0}};
    constexpr int U = 4;  // K-steps per round
    for (int k0 = 0; k0 < g.K; k0 += 4 * U) {
        double av[U], bv[U][2];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + 4 * u + lk;
            const bool kin = k < g.K;
            av[u] = (kin && ar >= 0) ? A[ar + (long)k * astep] : 0.0;
            const long br = kin ? gemm_row(g.b, b0, k) : -1;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int n = n0 + 16 * t + li;
                bv[u][t] = (br >= 0 && n < g.N) ? B[br + (long)n * bstep] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int t = 0; t < 2; ++t) acc[t] = mfma64(av[u], bv[u][t], acc[t]);
    }
    const long cstep = g.c.col ? g.c.ld : 1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int r = m0 + 4 * j + lk;
        if (r >= g.M) continue;
        const long cr = gemm_row(g.c, b0, r);
        if (cr < 0) continue;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int n = n0 + 16 * t + li;
            if (n < g.N) C[cr + (long)n * cstep] = acc[t][j];
        }
    }
}

void launch_gemm64(const GemmPass& g, hipStream_t st) {
    if (g.batch <= 0 || g.M <= 0 || g.N <= 0) return;
    if (g.N <= 32)  // narrow products: 64 x 32 block tiles
        hipLaunchKernelGGL(k_gemm64<4>, dim3(g.batch, (g.N + 31) / 32, (g.M + 63) / 64), dim3(256), 0, st, g);
    else
        hipLaunchKernelGGL(k_gemm64<2>, dim3(g.batch, (g.N + 63) / 64, (g.M + 31) / 32), dim3(256), 0, st, g);
}

// particles -> coefficients.  Block (box, 256-column slice of b + p c): wave
// w owns the 16-column tiles 4w .. 4w + 3 of the slice and every tile of a;
// particles are staged 32 at a time (charge-free slots skipped), their
// Chebyshev rows in LDS; per four particles the A operand (rows b + p c) is
// q T_b(y^) T_c(z^), the B operand (columns a) T_a(x^).
template <int P, typename T>
__global__ void __launch_bounds__(256) k_proxy_leaf(DevTree<T> t, PwNodes pn, const int* __restrict__ direct,
                                                    double* __restrict__ coef) {
    constexpr int PA = (P + 15) / 16 * 16, AT = PA / 16;
    constexpr int CTW = 4, CB = 4 * CTW * 16;
    constexpr int CHN = 32;
    __shared__ double tx[CHN][PA], ty[CHN][P], tz[CHN][P], qs[CHN], xs[CHN][3];
    __shared__ double ctr[3], scale;
    __shared__ int cnt;
    const int slot = direct[blockIdx.x];
    const int node = pn.node[slot];
    if (threadIdx.x < 3) ctr[threadIdx.x] = box_centre(t, node, threadIdx.x);
    if (threadIdx.x == 0) scale = 2.0 / t.boxsize[t.level[node]];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int bc0 = blockIdx.y * CB + wv * CTW * 16;
    int yb[CTW], zc[CTW];
    bool on[CTW];
#pragma unroll
    for (int ct = 0; ct < CTW; ++ct) {
        const int bc = bc0 + 16 * ct + li;
        on[ct] = bc < P * P;
        yb[ct] = on[ct] ? bc % P : 0;
        zc[ct] = on[ct] ? bc / P : 0;
    }
    f64x4 acc[CTW][AT];
#pragma unroll
    for (int ct = 0; ct < CTW; ++ct)
#pragma unroll
        for (int ta = 0; ta < AT; ++ta) acc[ct][ta] = f64x4{0.0, 0.0, 0.0, 0.0};
    const int s0 = pn.lo[slot], s1 = pn.hi[slot];
    for (int base = s0; base < s1; base += CHN) {
        __syncthreads();  // the previous chunk is consumed
        if (threadIdx.x < 64) {  // wave 0 compacts the chunk's charged slots
            const int s = base + threadIdx.x;
            const double q = (threadIdx.x < CHN && s < s1) ? t.pq[s] : 0.0;
            const bool live = q != 0.0;
            const unsigned long long bal = __ballot(live);
            const unsigned long long lt = threadIdx.x == 0 ? 0ull : (~0ull >> (64 - threadIdx.x));
            if (live) {
                const int m = __popcll(bal & lt);
                xs[m][0] = scale * (t.px[s] - ctr[0]);
                xs[m][1] = scale * (t.py[s] - ctr[1]);
                xs[m][2] = scale * (t.pz[s] - ctr[2]);
                qs[m] = q;
            }
            if (threadIdx.x == 0) cnt = __popcll(bal);
        }
        __syncthreads();
        const int n = cnt;
        for (int job = threadIdx.x; job < 3 * CHN; job += 256) {
            const int m = job / 3, ax = job % 3;
            double* row = ax == 0 ? tx[m] : (ax == 1 ? ty[m] : tz[m]);
            const int len = ax == 0 ? PA : P;
            if (m < n) {
                cheb_row(xs[m][ax], row, P);
                for (int a = P; a < len; ++a) row[a] = 0.0;
            } else {
                for (int a = 0; a < len; ++a) row[a] = 0.0;  // rows past the chunk's particles add nothing
                if (ax == 0) qs[m] = 0.0;
            }
        }
        __syncthreads();
        for (int s4 = 0; s4 < n; s4 += 4) {
            const int m = s4 + lk;
            double bop[AT];
#pragma unroll
            for (int ta = 0; ta < AT; ++ta) bop[ta] = tx[m][16 * ta + li];
            const double qm = qs[m];
#pragma unroll
            for (int ct = 0; ct < CTW; ++ct) {
                const double aop = on[ct] ? qm * ty[m][yb[ct]] * tz[m][zc[ct]] : 0.0;
#pragma unroll
                for (int ta = 0; ta < AT; ++ta) acc[ct][ta] = mfma64(aop, bop[ta], acc[ct][ta]);
            }
        }
    }
    double* C = coef + (size_t)slot * P * P * P;
#pragma unroll
    for (int ct = 0; ct < CTW; ++ct)
#pragma unroll
        for (int ta = 0; ta < AT; ++ta)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int bc = bc0 + 16 * ct + 4 * j + lk, a = 16 * ta + li;
                if (bc < P * P && a < P) C[a + (size_t)P * bc] = acc[ct][ta][j];
            }
}

// the de-centring phase e^{-i k.c} and the store of the expansions of the
// boxes b0 .. b0 + gridDim.x - 1 from the z pass (planes re, im of nz d^2)
template <typename T>
__global__ void __launch_bounds__(256) k_pw_out(DevTree<T> t, PwNodes pn, const double* __restrict__ s3, int n,
                                                int b0) {
    __shared__ double2 ph[3][kMaxGridD];  // e^{-i dk (k - n) c} per axis
    const int slot = b0 + blockIdx.x;
    const int node = pn.node[slot];
    const int l = pn.level[slot];
    const int d = 2 * n + 1, nz = n + 1, d2 = d * d;
    const double dk = t.delta_k[l];
    for (int k = threadIdx.x; k < 3 * d; k += blockDim.x) {
        const int ax = k / d, kk = k - ax * d;
        double s, c;
        sincos(dk * (kk - n) * box_centre(t, node, ax), &s, &c);
        ph[ax][kk] = make_double2(c, -s);
    }
    __syncthreads();
    const double* Y = s3 + (size_t)blockIdx.x * 2 * nz * d2;
    Cplx<T>* out = t.out + (size_t)slot * t.Md;
    for (int e = threadIdx.x; e < nz * d2; e += blockDim.x) {
        const int kz = e / d2, kxy = e - kz * d2, kx = kxy % d, ky = kxy / d;
        const double2 px = ph[0][kx], py = ph[1][ky], pz = ph[2][kz];
        const double xyr = px.x * py.x - px.y * py.y, xyi = px.x * py.y + px.y * py.x;
        const double pr = xyr * pz.x - xyi * pz.y, pi = xyr * pz.y + xyi * pz.x;  // e^{-i k.c}
        const double ar = Y[e], ai = Y[nz * d2 + e];
        out[e] = {(T)(ar * pr - ai * pi), (T)(ar * pi + ai * pr)};
    }
}

// the root expansion from the level-2 boxes' plane waves (z pass results in
// s3, planes re, im of nzw dw^2 per box): root[k] = sum_b e^{-i k.c_b} Y_b[k]
// over the boxes in slot order, in FP64, then stored in the tree's precision
template <typename T>
__global__ void __launch_bounds__(256) k_root_sum(DevTree<T> t, PwNodes pn, const int* __restrict__ slots, int nb,
                                                  const double* __restrict__ s3, double* __restrict__ acc) {
    // e^{-i dk (k - nw) c} for the four level-2 cell centres c = (i + 1/2) L / 4 of an axis
    __shared__ double2 ph[4][kMaxGridD];
    const int nw = t.n_window, dw = t.dw, nzw = nw + 1, d2 = dw * dw;
    const double dk = t.delta_k[0], b2 = t.boxsize[2];
    for (int k = threadIdx.x; k < 4 * dw; k += blockDim.x) {
        const int ci = k / dw, kk = k - ci * dw;
        double s, c;
        sincos(dk * (kk - nw) * ((ci + 0.5) * b2), &s, &c);
        ph[ci][kk] = make_double2(c, -s);
    }
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nzw * d2) return;
    const int kz = e / d2, kxy = e - kz * d2, kx = kxy % dw, ky = kxy / dw;
    double sr = 0.0, si = 0.0;
    for (int b = 0; b < nb; ++b) {
        const int node = pn.node[slots[b]];
        const double2 px = ph[t.cell[3 * node]][kx], py = ph[t.cell[3 * node + 1]][ky], pz = ph[t.cell[3 * node + 2]][kz];
        const double xyr = px.x * py.x - px.y * py.y, xyi = px.x * py.y + px.y * py.x;
        const double pr = xyr * pz.x - xyi * pz.y, pi = xyr * pz.y + xyi * pz.x;  // e^{-i k.c_b}
        const double* Y = s3 + (size_t)b * 2 * nzw * d2;
        const double ar = Y[e], ai = Y[nzw * d2 + e];
        sr += ar * pr - ai * pi;
        si += ar * pi + ai * pr;
    }
    acc[2 * e] = sr;
    acc[2 * e + 1] = si;
    t.root[e] = {(T)sr, (T)si};
}

template <int P, typename T>
static void launch_proxy_p(const DevTree<T>& t, const PwNodes& pn, double* acc_root, const ProxyArgs& pa,
                           hipStream_t st) {
    constexpr int CB = 256;
    if (pa.n_direct > 0)
        hipLaunchKernelGGL((k_proxy_leaf<P, T>), dim3(pa.n_direct, (P * P + CB - 1) / CB), dim3(256), 0, st, t, pn,
                           pa.direct, pa.coef);
    const int p = P;
    const long p2 = (long)p * p, p3 = p2 * p;
    // child -> parent, deepest parents first
    for (const ProxyChunk& c : pa.c2p) {
        GemmPass x;  // T1[l1 + p (b + p c)] = sum_a C[a + p (b + p c)] A_ox[l1][a]
        x.M = (int)p2, x.K = p, x.N = p, x.batch = c.n_ch;
        x.a.p = pa.coef, x.a.ld = p, x.a.tab = c.ch, x.a.s0 = p3;
        x.b.p = const_cast<double*>(pa.A), x.b.ld = p, x.b.col = 1, x.b.tab = c.ox, x.b.s0 = p2;
        x.c.p = pa.s1, x.c.ld = p, x.c.s0 = p3;
        launch_gemm64(x, st);
        GemmPass y;  // T2[l1 + p l2 + p^2 c] = sum_b A_oy[l2][b] T1[l1 + p (b + p c)]
        y.M = p, y.K = p, y.N = p, y.inner = p, y.batch = c.n_ch * p;
        y.a.p = const_cast<double*>(pa.A), y.a.ld = p, y.a.tab = c.oy, y.a.s0 = p2;
        y.b.p = pa.s1, y.b.ld = p, y.b.s0 = p3, y.b.s1 = p2;
        y.c.p = pa.s2, y.c.ld = p, y.c.s0 = p3, y.c.s1 = p2;
        launch_gemm64(y, st);
        GemmPass z;  // C_parent[(l1 + p l2) + p^2 l3] = sum_{ci, c} A_{z(ci)}[l3][c] T2_ci[(l1 + p l2) + p^2 c]
        z.M = p, z.K = 8 * p, z.N = (int)p2, z.batch = c.n_par;
        z.a.p = const_cast<double*>(pa.Az8), z.a.ld = 8 * p, z.a.s0 = 0;
        z.b.p = pa.s2, z.b.ld = p2, z.b.s0 = 0, z.b.seg = p, z.b.nseg = 8, z.b.stab = c.seg, z.b.sstride = p3;
        z.c.p = pa.coef, z.c.ld = p2, z.c.tab = c.par, z.c.s0 = p3;
        launch_gemm64(z, st);
    }
    // coefficients -> plane waves: the x, y and z passes of nb boxes (slots
    // tab[i], or b0 + i) with the stacked matrices of grid d, nz into s3
    auto pw3 = [&](const int* tab, int b0, int nb, const double* Mx, const double* My, const double* Mz, int d,
                   int nz) {
        const long d2 = (long)d * d;
        GemmPass x;  // Y1[(plane d + kx) + 2d (b + p c)] = sum_a C[a + p (b + p c)] Mx[plane d + kx][a]
        x.M = (int)p2, x.K = p, x.N = 2 * d, x.batch = nb;
        x.a.p = pa.coef + (size_t)b0 * p3, x.a.ld = p, x.a.tab = tab, x.a.s0 = p3;
        x.b.p = const_cast<double*>(Mx), x.b.ld = p, x.b.col = 1;
        x.c.p = pa.s1, x.c.ld = 2 * d, x.c.s0 = 2 * d * p2;
        launch_gemm64(x, st);
        GemmPass y;  // Y2[kx + d ky + d^2 c + d^2 p plane] = sum_{plane', b} My[plane d + ky][plane' p + b] Y1[plane' d + kx + 2d (b + p c)]
        y.M = 2 * d, y.K = 2 * p, y.N = d, y.inner = p, y.batch = nb * p;
        y.a.p = const_cast<double*>(My), y.a.ld = 2 * p;
        y.b.p = pa.s1, y.b.ld = 2 * d, y.b.s0 = 2 * d * p2, y.b.s1 = 2 * d * p, y.b.seg = p, y.b.nseg = 2,
        y.b.sstride = d;
        y.c.p = pa.s2, y.c.ld = d, y.c.s0 = 2 * d2 * p, y.c.s1 = d2, y.c.seg = d, y.c.nseg = 2, y.c.sstride = d2 * p;
        launch_gemm64(y, st);
        GemmPass z;  // Y3[kxy + d^2 kz + d^2 nz plane] = sum_{plane', c} Mz[plane nz + kz][plane' p + c] Y2[kxy + d^2 c + d^2 p plane']
        z.M = 2 * nz, z.K = 2 * p, z.N = (int)d2, z.batch = nb;
        z.a.p = const_cast<double*>(Mz), z.a.ld = 2 * p;
        z.b.p = pa.s2, z.b.ld = d2, z.b.s0 = 2 * d2 * p, z.b.seg = p, z.b.nseg = 2, z.b.sstride = d2 * p;
        z.c.p = pa.s3, z.c.ld = d2, z.c.s0 = 2 * d2 * nz, z.c.seg = nz, z.c.nseg = 2, z.c.sstride = d2 * nz;
        launch_gemm64(z, st);
    };
    // the root from the level-2 boxes, at the root's wavenumbers
    if (pa.n_root > 0) {
        const int dw = 2 * pa.nw + 1, nzw = pa.nw + 1;
        pw3(pa.root_slots, 0, pa.n_root, pa.Prx, pa.Pry, pa.Prz, dw, nzw);
        hipLaunchKernelGGL(k_root_sum<T>, dim3((nzw * dw * dw + 255) / 256), dim3(256), 0, st, t, pn, pa.root_slots,
                           pa.n_root, pa.s3, acc_root);
    }
    // every expansion box, in chunks of pa.pw_chunk
    for (int b0 = 0; b0 < t.n_pw_nodes; b0 += pa.pw_chunk) {
        const int nb = std::min(pa.pw_chunk, t.n_pw_nodes - b0);
        pw3(nullptr, b0, nb, pa.Px, pa.Py, pa.Pz, pa.d, pa.nz);
        hipLaunchKernelGGL(k_pw_out<T>, dim3(nb), dim3(256), 0, st, t, pn, pa.s3, pa.n, b0);
    }
}

template <typename T>
void launch_outgoing_proxy(const DevTree<T>& t, const PwNodes& pn, double* acc_root, const ProxyArgs& pa,
                           hipStream_t st) {
    if (pa.n_root == 0) launch_outgoing_root(t, acc_root, st);  // from the particles
    if (pa.p == 18)
        launch_proxy_p<18>(t, pn, acc_root, pa, st);  // digits 3
    else if (pa.p == 28)
        launch_proxy_p<28>(t, pn, acc_root, pa, st);  // digits 6
    else
        launch_proxy_p<44>(t, pn, acc_root, pa, st);  // digits 9
}

template void launch_outgoing_proxy<double>(const DevTree<double>&, const PwNodes&, double*, const ProxyArgs&,
                                            hipStream_t);
template void launch_outgoing_proxy<float>(const DevTree<float>&, const PwNodes&, double*, const ProxyArgs&,
                                           hipStream_t);

}  // namespace hpdmk_amd
