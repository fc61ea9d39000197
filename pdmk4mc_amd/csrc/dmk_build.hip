// HIP kernels (gfx950) that (re)build the expansions and evaluate the total
// energy: root and per-box outgoing plane waves, incoming assembly, window,
// difference and residual energies.  Phase arguments are formed in FP64;
// sums accumulate in FP64; T is the storage type (double or float).
#include "dmk_device.hpp"
#include "dmk_kernels.hpp"

namespace hpdmk_amd {

// ------------------------------------------------------------------ expansions
// outgoing root expansion, split over particle chunks; FP64 atomics into acc[2*Mw]
template <typename T>
__global__ void __launch_bounds__(TB) k_pw_root(DevTree<T> t, double* acc, int chunk) {
    constexpr int TILE = 16;
    __shared__ Cplx<T> tab[TILE][kRow];
    const int nw = t.n_window, dw = t.dw;
    const int e = blockIdx.x * TB + threadIdx.x;
    const int s0 = blockIdx.y * chunk, s1 = min(s0 + chunk, t.total_slots);
    double sr = 0, si = 0;
    for (int base = s0; base < s1; base += TILE) {
        __syncthreads();
        for (int p = threadIdx.x / 64; p < TILE; p += blockDim.x / 64) {
            const int s = base + p;
            const bool ok = s < s1;
            phase_rows(tab[p], tab[p] + dw, tab[p] + 2 * dw, nw, t.delta_k[0], ok ? t.px[s] : 0.0,
                       ok ? t.py[s] : 0.0, ok ? t.pz[s] : 0.0, ok ? t.pq[s] : 0.0, threadIdx.x & 63, 64);
        }
        __syncthreads();
        if (e < t.Mw) {
            const int i = e % dw, r = e / dw, j = r % dw, k = r / dw;
            for (int p = 0; p < TILE; ++p) {
                const Cplx<T> v = cmul(cmul(tab[p][i], tab[p][dw + j]), tab[p][2 * dw + k]);
                sr += (double)v.x;
                si += (double)v.y;
            }
        }
    }
    if (e < t.Mw) {
        atomicAdd(acc + 2 * e, sr);
        atomicAdd(acc + 2 * e + 1, si);
    }
}

template <typename T>
__global__ void k_store_root(DevTree<T> t, const double* acc) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < t.Mw) t.root[e] = {(T)acc[2 * e], (T)acc[2 * e + 1]};
}

// outgoing expansion of every box of levels 2..D-2 (src/planewave.cpp:48-84):
// grid (ceil(Md/TB), n_pw_nodes); the box's particles are the slot range of
// its descendant leaves (unused slots carry q = 0)
template <typename T>
__global__ void __launch_bounds__(TB) k_pw_level(DevTree<T> t, PwNodes pn) {
    constexpr int TILE = 16;
    __shared__ Cplx<T> tab[TILE][kRow];
    const int slot = blockIdx.y;
    const int l = pn.level[slot];
    const int nd = t.n_diff, dd = t.dd;
    const int e = blockIdx.x * TB + threadIdx.x;
    const int s0 = pn.lo[slot], s1 = pn.hi[slot];
    double sr = 0, si = 0;
    for (int base = s0; base < s1; base += TILE) {
        __syncthreads();
        for (int p = threadIdx.x / 64; p < TILE; p += blockDim.x / 64) {
            const int s = base + p;
            const bool ok = s < s1;
            phase_rows(tab[p], tab[p] + dd, tab[p] + 2 * dd, nd, t.delta_k[l], ok ? t.px[s] : 0.0,
                       ok ? t.py[s] : 0.0, ok ? t.pz[s] : 0.0, ok ? t.pq[s] : 0.0, threadIdx.x & 63, 64);
        }
        __syncthreads();
        if (e < t.Md) {
            const int i = e % dd, r = e / dd, j = r % dd, k = r / dd;
            for (int p = 0; p < TILE; ++p) {
                const Cplx<T> v = cmul(cmul(tab[p][i], tab[p][dd + j]), tab[p][2 * dd + k]);
                sr += (double)v.x;
                si += (double)v.y;
            }
        }
    }
    if (e < t.Md) t.out[(size_t)slot * t.Md + e] = {(T)sr, (T)si};
}

// incoming expansion of non-leaf boxes (src/planewave.cpp:99-155):
// in_B = conj(sum over B and its 26 colleagues of out_C s_p), p = image of C adjacent to B
template <typename T>
__global__ void __launch_bounds__(TB) k_incoming(DevTree<T> t, PwNodes pn) {
    __shared__ Cplx<T> sh[9 * 64];
    const int slot = blockIdx.y;
    const int node = pn.node[slot];
    if (t.is_leaf[node]) return;
    const int l = pn.level[slot];
    const int nd = t.n_diff, dd = t.dd;
    // sh[(3*axis + s) * dd + m] = e^{-i (m - n) dk_l (s-1) L}
    for (int q = threadIdx.x; q < 9 * dd; q += blockDim.x) {
        const int ax = q / (3 * dd), s = (q / dd) % 3, m = q % dd;
        (void)ax;
        double sn, cs;
        sincos((double)(m - nd) * (t.delta_k[l] * (s - 1) * t.L), &sn, &cs);
        sh[q] = {(T)cs, (T)(-sn)};
    }
    __syncthreads();
    const int e = blockIdx.x * TB + threadIdx.x;
    if (e >= t.Md) return;
    const int i = e % dd, r = e / dd, j = r % dd, k = r / dd;
    const Cplx<T> self = t.out[(size_t)slot * t.Md + e];
    double ar = (double)self.x, ai = (double)self.y;
    const int nl = 1 << l;
    int sidx = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!dx && !dy && !dz) continue;
                const int C = t.colleague[26 * node + sidx++];
                if (C < 0) continue;
                const int cx = t.cell[3 * node] + dx, cy = t.cell[3 * node + 1] + dy, cz = t.cell[3 * node + 2] + dz;
                const int px = cx < 0 ? -1 : (cx >= nl ? 1 : 0);
                const int py = cy < 0 ? -1 : (cy >= nl ? 1 : 0);
                const int pz = cz < 0 ? -1 : (cz >= nl ? 1 : 0);
                const Cplx<T> s = cmul(cmul(sh[(0 + px + 1) * dd + i], sh[(3 + py + 1) * dd + j]), sh[(6 + pz + 1) * dd + k]);
                const Cplx<T> v = cmul(t.out[(size_t)t.pw_slot[C] * t.Md + e], s);
                ar += (double)v.x;
                ai += (double)v.y;
            }
    t.inc[(size_t)slot * t.Md + e] = {(T)ar, (T)(-ai)};
}

// ------------------------------------------------------------------ energies
template <typename T>
__global__ void __launch_bounds__(TB) k_energy_window(DevTree<T> t, double* res) {
    __shared__ double red[TB / 64];
    const int e = blockIdx.x * TB + threadIdx.x;
    double v = 0;
    if (e < t.Mw) {
        const Cplx<T> g = t.root[e];
        v = ((double)g.x * g.x + (double)g.y * g.y) * (double)t.Wmat[e];
    }
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) atomicAdd(res + 0, s * t.cW);
}

template <typename T>
__global__ void __launch_bounds__(TB) k_energy_diff(DevTree<T> t, PwNodes pn, double* res) {
    __shared__ double red[TB / 64];
    const int slot = blockIdx.y;
    const int node = pn.node[slot];
    if (t.is_leaf[node]) return;
    const int l = pn.level[slot];
    const int e = blockIdx.x * TB + threadIdx.x;
    double v = 0;
    if (e < t.Md) {
        const Cplx<T> o = t.out[(size_t)slot * t.Md + e], in = t.inc[(size_t)slot * t.Md + e];
        v = ((double)o.x * in.x - (double)o.y * in.y) * (double)t.Dmat[(size_t)(l - 2) * t.Md + e] * t.cD[l];
    }
    if (blockIdx.x == 0)  // self-interaction term -Q_B S_l (src/energy.cpp:67-71)
        for (int s = pn.lo[slot] + threadIdx.x; s < pn.hi[slot]; s += blockDim.x) v -= t.pq[s] * t.pq[s] * t.self_diff[l];
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) atomicAdd(res + 1, s);
}

// residual energy (src/energy.cpp:82-173): grid (max_near, n_leaf_list);
// block = (leaf B, near entry); pairs counted from B's side with weight 1/2
template <typename T>
__global__ void __launch_bounds__(TB) k_energy_res(DevTree<T> t, const int* leaves, double* res) {
    __shared__ double red[TB / 64];
    const int leaf = leaves[blockIdx.y];
    const int k = blockIdx.x;
    const int off = t.near_off[leaf], cnt = t.near_off[leaf + 1] - off;
    if (k >= cnt) return;
    const int nb = t.near_ids[off + k];
    const int code = t.near_img[off + k];
    const double sx = (code % 3 - 1) * t.L, sy = ((code / 3) % 3 - 1) * t.L, sz = (code / 9 - 1) * t.L;
    const double bb = t.boxsize[t.level[leaf]];
    const T b2 = (T)(bb * bb), inv_b = (T)(1.0 / bb);
    const int a0 = t.slot_start[leaf], na = t.slot_count[leaf];
    const int c0 = t.slot_start[nb], nc = t.slot_count[nb];
    double v = 0;
    for (int p = threadIdx.x; p < na; p += blockDim.x) {
        const int s = a0 + p;
        const double x = t.px[s], y = t.py[s], z = t.pz[s];
        double acc = 0;
        for (int q = 0; q < nc; ++q) {
            const int s2 = c0 + q;
            if (s2 == s) continue;
            const T dx = (T)(x - (t.px[s2] + sx)), dy = (T)(y - (t.py[s2] + sy)), dz = (T)(z - (t.pz[s2] + sz));
            const T r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > T(0) && r2 < b2) acc += (double)((T)t.pq[s2] * residual_poly_over_r<T>(sqrt(r2), inv_b, t.poly, t.n_poly));
        }
        v += 0.5 * acc * t.pq[s];
    }
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) atomicAdd(res + 2, s);
}

// zero fill used instead of hipMemsetAsync, so that any of these sequences can
// be captured into a HIP graph without memset nodes
__global__ void k_zero(double* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// streaming copy of n 16-byte words (HBM bandwidth probe and byte-counter
// calibration): each word is read once and written once
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void launch_copy16(const void* src, void* dst, size_t n16, hipStream_t st) {
    hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(TB), 0, st, (const uint4*)src, (uint4*)dst, n16);
}

template <typename T>
void launch_outgoing_root(const DevTree<T>& t, double* acc_root, hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3((2 * t.Mw + TB - 1) / TB), dim3(TB), 0, st, acc_root, 2 * t.Mw);
    const int chunk = 2048;
    const int nch = (t.total_slots + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_pw_root<T>, dim3((t.Mw + TB - 1) / TB, nch), dim3(TB), 0, st, t, acc_root, chunk);
    hipLaunchKernelGGL(k_store_root<T>, dim3((t.Mw + TB - 1) / TB), dim3(TB), 0, st, t, (const double*)acc_root);
}

template <typename T>
void launch_outgoing(const DevTree<T>& t, const PwNodes& pn, double* acc_root, hipStream_t st) {
    launch_outgoing_root(t, acc_root, st);
    if (t.n_pw_nodes > 0)
        hipLaunchKernelGGL(k_pw_level<T>, dim3((t.Md + TB - 1) / TB, t.n_pw_nodes), dim3(TB), 0, st, t, pn);
}

template <typename T>
void launch_incoming(const DevTree<T>& t, const PwNodes& pn, hipStream_t st) {
    if (t.n_pw_nodes > 0)
        hipLaunchKernelGGL(k_incoming<T>, dim3((t.Md + TB - 1) / TB, t.n_pw_nodes), dim3(TB), 0, st, t, pn);
}

template <typename T>
void launch_energy(const DevTree<T>& t, const PwNodes& pn, const int* leaves, int n_leaves, double* res,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(TB), 0, st, res, 3);
    hipLaunchKernelGGL(k_energy_window<T>, dim3((t.Mw + TB - 1) / TB), dim3(TB), 0, st, t, res);
    if (t.n_pw_nodes > 0)
        hipLaunchKernelGGL(k_energy_diff<T>, dim3((t.Md + TB - 1) / TB, t.n_pw_nodes), dim3(TB), 0, st, t, pn, res);
    if (n_leaves > 0)
        hipLaunchKernelGGL(k_energy_res<T>, dim3(t.max_near, n_leaves), dim3(TB), 0, st, t, leaves, res);
}

#define INSTANTIATE(T)                                                                                     \
    template void launch_outgoing_root<T>(const DevTree<T>&, double*, hipStream_t);                      \
    template void launch_outgoing<T>(const DevTree<T>&, const PwNodes&, double*, hipStream_t);           \
    template void launch_incoming<T>(const DevTree<T>&, const PwNodes&, hipStream_t);                    \
    template void launch_energy<T>(const DevTree<T>&, const PwNodes&, const int*, int, double*, hipStream_t);
INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace hpdmk_amd
