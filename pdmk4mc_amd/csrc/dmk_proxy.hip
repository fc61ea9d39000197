// PROXY initialisation of the outgoing expansions on gfx950 (the reference's
// Chebyshev upward pass, src/upward_pass.cpp:206-322): three FP64 kernels.
//
//   k_proxy_leaf  particles -> Chebyshev coefficients
//                 C[a + p b + p^2 c] = sum_j q_j T_a(x^) T_b(y^) T_c(z^),
//                 x^ = 2 (x - centre) / box size, for every box of level D-2
//                 and every shallower leaf (upward_pass.cpp:109-149, 270-287)
//   k_proxy_c2p   child -> parent by the exact re-expansion along each axis,
//                 one level at a time from D-3 up to 2 (upward_pass.cpp:19-64,
//                 289-321)
//   k_proxy_pw    coefficients -> half-grid plane waves
//                 rho[k] = e^{-i k.c} sum P[kx,a] P[ky,b] P[kz,c] C[a,b,c]
//                 (proxycharge2pw + decenter_phase, upward_pass.cpp:67-106, 183-203)
//
// Layouts: coefficients of pw slot s at coef + s p^3 (a fastest); A holds the
// two child-to-parent matrices (lower and upper half, A[h p^2 + l p + a]);
// Pm the coefficient-to-plane-wave matrix, Pm[k p + a] complex, k = 0..2n.
// A thread block owns one box; each thread keeps the sums of a fixed set of
// index pairs in registers and the per-slab operands are staged in LDS.
#include "dmk_device.hpp"
#include "dmk_kernels.hpp"
#include "dmk_proxy.hpp"

namespace hpdmk_amd {

namespace {

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

}  // namespace

// particles -> coefficients.  Thread t owns the pairs (b, c) = t + NT r and
// all a; particles are staged 64 at a time (charge-free slots skipped).
template <int P, int NT, typename T>
__global__ void __launch_bounds__(NT) k_proxy_leaf(DevTree<T> t, PwNodes pn, ProxyArgs pa) {
    constexpr int R = (P * P + NT - 1) / NT;
    constexpr int CH = 64;
    __shared__ double tx[CH][P], ty[CH][P], tz[CH][P], qs[CH];
    __shared__ double ctr[3], scale;
    __shared__ int cnt;
    const int slot = pa.direct[blockIdx.x];
    const int node = pn.node[slot];
    if (threadIdx.x < 3) ctr[threadIdx.x] = box_centre(t, node, threadIdx.x);
    if (threadIdx.x == 0) scale = 2.0 / t.boxsize[t.level[node]];
    double acc[R][P];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int a = 0; a < P; ++a) acc[r][a] = 0.0;
    const int s0 = pn.lo[slot], s1 = pn.hi[slot];
    __shared__ double xs[CH][3];
    for (int base = s0; base < s1; base += CH) {
        __syncthreads();  // the previous chunk is consumed
        if (threadIdx.x < 64) {  // wave 0 compacts the chunk's charged slots
            const int s = base + threadIdx.x;
            const double q = s < s1 ? t.pq[s] : 0.0;
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
        for (int job = threadIdx.x; job < 3 * n; job += NT) {
            const int m = job / 3, ax = job % 3;
            cheb_row(xs[m][ax], ax == 0 ? tx[m] : (ax == 1 ? ty[m] : tz[m]), P);
        }
        __syncthreads();
        for (int m = 0; m < n; ++m) {
            double w[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int bc = threadIdx.x + NT * r;
                w[r] = bc < P * P ? qs[m] * ty[m][bc % P] * tz[m][bc / P] : 0.0;
            }
#pragma unroll
            for (int a = 0; a < P; ++a) {
                const double xa = tx[m][a];
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r][a] += w[r] * xa;
            }
        }
    }
    double* C = pa.coef + (size_t)slot * P * P * P;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int bc = threadIdx.x + NT * r;
        if (bc < P * P)
#pragma unroll
            for (int a = 0; a < P; ++a) C[a + (size_t)P * bc] = acc[r][a];
    }
}

// child -> parent for the non-leaf boxes of one level.  Thread t owns the
// pairs (l1, l2) = t + NT r and all l3.  Per child and per z slab c:
// Tm[l1, b] = sum_a Ax[l1, a] S[a, b] (LDS), then
// acc[l1, l2, l3] += Az[l3, c] sum_b Ay[l2, b] Tm[l1, b].
template <int P, int NT, typename T>
__global__ void __launch_bounds__(NT) k_proxy_c2p(DevTree<T> t, PwNodes pn, ProxyArgs pa, const int* parents) {
    constexpr int R = (P * P + NT - 1) / NT;
    __shared__ double A[2][P * P], S[P * P], Tm[P * P];
    const int slot = parents[blockIdx.x];
    const int node = pn.node[slot];
    for (int e = threadIdx.x; e < 2 * P * P; e += NT) A[e / (P * P)][e % (P * P)] = pa.A[e];
    double acc[R][P];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int l3 = 0; l3 < P; ++l3) acc[r][l3] = 0.0;
    for (int ci = 0; ci < 8; ++ci) {
        const int ch = t.child[8 * node + ci];
        if (ch < 0) continue;  // uniform
        const double* Cc = pa.coef + (size_t)t.pw_slot[ch] * P * P * P;
        const double* Ax = A[ci & 1];
        const double* Ay = A[(ci >> 1) & 1];
        const double* Az = A[(ci >> 2) & 1];
        for (int c = 0; c < P; ++c) {
            __syncthreads();  // S and Tm of the previous slab are consumed
            for (int e = threadIdx.x; e < P * P; e += NT) S[e] = Cc[(size_t)P * P * c + e];
            __syncthreads();
            for (int e = threadIdx.x; e < P * P; e += NT) {
                const int l1 = e % P, b = e / P;
                double s = 0.0;
#pragma unroll
                for (int a = 0; a < P; ++a) s += Ax[l1 * P + a] * S[a + P * b];
                Tm[l1 + P * b] = s;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int pr = threadIdx.x + NT * r;
                if (pr >= P * P) continue;
                const int l1 = pr % P, l2 = pr / P;
                double w = 0.0;
#pragma unroll
                for (int b = 0; b < P; ++b) w += Ay[l2 * P + b] * Tm[l1 + P * b];
#pragma unroll
                for (int l3 = 0; l3 < P; ++l3) acc[r][l3] += Az[l3 * P + c] * w;
            }
        }
    }
    double* C = pa.coef + (size_t)slot * P * P * P;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int pr = threadIdx.x + NT * r;
        if (pr < P * P)
#pragma unroll
            for (int l3 = 0; l3 < P; ++l3) C[pr + (size_t)P * P * l3] = acc[r][l3];
    }
}

// coefficients -> plane waves of every expansion box.  Thread t owns the
// mode pair (kx, ky) = t (d^2 <= NT) and all kz <= 0.  Per z slab c of the
// coefficients: H[kx, b] = sum_a P[kx, a] S[a, b] (LDS), then
// acc[kz] += P[kz, c] sum_b P[ky, b] H[kx, b]; finally the de-centring phase.
template <int P, int NT, typename T>
__global__ void __launch_bounds__(NT) k_proxy_pw(DevTree<T> t, PwNodes pn, ProxyArgs pa) {
    constexpr int NZ_MAX = 14;  // n + 1 for n <= 13
    __shared__ double S[P * P];
    __shared__ double2 Pm[27 * P], Hs[27 * P];
    const int slot = blockIdx.x;
    const int node = pn.node[slot];
    const int l = pn.level[slot];
    const int n = t.n_diff, d = t.dd;
    for (int e = threadIdx.x; e < d * P; e += NT) Pm[e] = make_double2(pa.P[2 * e], pa.P[2 * e + 1]);
    const double* C = pa.coef + (size_t)slot * P * P * P;
    const int kx = threadIdx.x % d, ky = threadIdx.x / d;
    const bool own = threadIdx.x < d * d;
    double ar[NZ_MAX], ai[NZ_MAX];
#pragma unroll
    for (int k = 0; k < NZ_MAX; ++k) ar[k] = ai[k] = 0.0;
    for (int c = 0; c < P; ++c) {
        __syncthreads();  // S and Hs of the previous slab are consumed
        for (int e = threadIdx.x; e < P * P; e += NT) S[e] = C[(size_t)P * P * c + e];
        __syncthreads();
        for (int e = threadIdx.x; e < d * P; e += NT) {
            const int k = e % d, b = e / d;
            double hr = 0.0, hi = 0.0;
#pragma unroll
            for (int a = 0; a < P; ++a) {
                const double2 pk = Pm[k * P + a];
                const double sv = S[a + P * b];
                hr += pk.x * sv;
                hi += pk.y * sv;
            }
            Hs[k + d * b] = make_double2(hr, hi);
        }
        __syncthreads();
        if (own) {
            double gr = 0.0, gi = 0.0;
#pragma unroll
            for (int b = 0; b < P; ++b) {
                const double2 py = Pm[ky * P + b], h = Hs[kx + d * b];
                gr += py.x * h.x - py.y * h.y;
                gi += py.x * h.y + py.y * h.x;
            }
#pragma unroll
            for (int kz = 0; kz < NZ_MAX; ++kz) {
                if (kz > n) break;
                const double2 pz = Pm[kz * P + c];
                ar[kz] += pz.x * gr - pz.y * gi;
                ai[kz] += pz.x * gi + pz.y * gr;
            }
        }
    }
    if (!own) return;
    const double dk = t.delta_k[l];
    const double cx = box_centre(t, node, 0), cy = box_centre(t, node, 1), cz = box_centre(t, node, 2);
    double sxy, cxy;
    sincos(dk * ((kx - n) * cx + (ky - n) * cy), &sxy, &cxy);
    Cplx<T>* out = t.out + (size_t)slot * t.Md;
#pragma unroll
    for (int kz = 0; kz < NZ_MAX; ++kz) {
        if (kz > n) break;
        double sz, cz_;
        sincos(dk * (kz - n) * cz, &sz, &cz_);
        const double pr = cxy * cz_ - sxy * sz, pi = -(sxy * cz_ + cxy * sz);  // e^{-i k.c}
        out[kx + d * ky + (size_t)d * d * kz] = {(T)(ar[kz] * pr - ai[kz] * pi), (T)(ar[kz] * pi + ai[kz] * pr)};
    }
}

template <int P, int NL, int NC, int NW, typename T>
static void launch_proxy_p(const DevTree<T>& t, const PwNodes& pn, const ProxyArgs& pa, hipStream_t st) {
    if (pa.n_direct > 0) hipLaunchKernelGGL((k_proxy_leaf<P, NL, T>), dim3(pa.n_direct), dim3(NL), 0, st, t, pn, pa);
    for (int li = pa.n_levels - 1; li >= 0; --li) {
        const int cnt = pa.level_cnt[li];
        if (cnt > 0)
            hipLaunchKernelGGL((k_proxy_c2p<P, NC, T>), dim3(cnt), dim3(NC), 0, st, t, pn, pa,
                               pa.c2p + pa.level_off[li]);
    }
    if (t.n_pw_nodes > 0) hipLaunchKernelGGL((k_proxy_pw<P, NW, T>), dim3(t.n_pw_nodes), dim3(NW), 0, st, t, pn, pa);
}

template <typename T>
void launch_outgoing_proxy(const DevTree<T>& t, const PwNodes& pn, double* acc_root, const ProxyArgs& pa,
                           hipStream_t st) {
    launch_outgoing_root(t, acc_root, st);
    if (pa.p == 18)
        launch_proxy_p<18, 256, 256, 192>(t, pn, pa, st);  // digits 3: d = 13, d^2 = 169 mode pairs
    else
        launch_proxy_p<28, 512, 512, 640>(t, pn, pa, st);  // digits 6: d = 25, d^2 = 625 mode pairs
}

template void launch_outgoing_proxy<double>(const DevTree<double>&, const PwNodes&, double*, const ProxyArgs&,
                                            hipStream_t);
template void launch_outgoing_proxy<float>(const DevTree<float>&, const PwNodes&, double*, const ProxyArgs&,
                                           hipStream_t);

}  // namespace hpdmk_amd
