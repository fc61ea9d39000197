// Device build of the octree topology (src/tree.cpp:268-444 with SCTL
// UpdateRefinement semantics, src/tree.cpp:136-165 neighbours), the same tree
// as build_octree (dmk_host.cpp) node for node:
//   1. Morton keys of the particles (21 bits per axis) and a stable radix
//      sort of (key, particle index);
//   2. refinement level by level: a box is internal while it holds more than
//      n_per_leaf particles (below level 21); a child's particle range is two
//      binary searches in the sorted keys; levels 0 and 1 are always refined;
//   3. periodic 2:1 balance: every internal box of level l >= 1 needs its 26
//      same-level neighbour cells, i.e. their parents internal at level l - 1.
//      Processing levels from the deepest up adds exactly the forced splits, so
//      the result is the unique smallest balanced refinement, which is what the
//      host's repeated sweeps reach;
//   4. nodes = the root and the 8 children of every internal box; Morton
//      pre-order by a radix sort of (full-depth key, level);
//   5. per node: level, cell, parent, children, leaf ranges, colleagues; per
//      leaf the near list (itself, the leaves under each existing neighbour,
//      the coarser leaf covering each missing neighbour cell).
// Boxes of one level are kept as sorted Morton prefixes, so every lookup is a
// binary search; the host loop only carries the per-level counts.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "dmk_octree.hpp"

namespace hpdmk_amd {

namespace {

constexpr int kKeyBits = 21;
constexpr int OB = 256;

__host__ __device__ inline uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
__host__ __device__ inline uint32_t compact3(uint64_t v) {
    v &= 0x1249249249249249ull;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ull;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00full;
    v = (v ^ (v >> 8)) & 0x1f0000ff0000ffull;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffull;
    v = (v ^ (v >> 32)) & 0x1fffff;
    return (uint32_t)v;
}
__device__ inline uint64_t encode(int cx, int cy, int cz) {
    return spread3((uint64_t)cx) | spread3((uint64_t)cy) << 1 | spread3((uint64_t)cz) << 2;
}

// first index in [0, n) whose value is >= v (n when none)
__device__ inline int lower_bound64(const uint64_t* a, int n, uint64_t v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}
// index of v in the sorted a[0, n), -1 when absent
__device__ inline int find64(const uint64_t* a, int n, uint64_t v) {
    const int i = lower_bound64(a, n, v);
    return i < n && a[i] == v ? i : -1;
}

__global__ void k_morton(int n, const double* __restrict__ r, double L, uint64_t* keys, int* vals, int* err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double scale = (double)(1 << kKeyBits);
    uint64_t c3[3];
    for (int a = 0; a < 3; ++a) {
        const double x = r[3 * (size_t)i + a];
        if (!(x >= 0 && x < L)) atomicOr(err, 1);
        long v = (long)floor(x / L * scale);
        v = v < 0 ? 0 : (v > (1l << kKeyBits) - 1 ? (1l << kKeyBits) - 1 : v);
        c3[a] = (uint64_t)v;
    }
    keys[i] = spread3(c3[0]) | spread3(c3[1]) << 1 | spread3(c3[2]) << 2;
    vals[i] = i;
}

// children of the internal boxes P (level l, sorted prefixes): child prefix,
// its particle range and whether it is internal at level l + 1
__global__ void k_children(int np, const uint64_t* __restrict__ P, int l, const uint64_t* __restrict__ keys, int n,
                           int cap, uint64_t* cq, int* flag) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= 8 * np) return;
    const uint64_t q = P[w >> 3] * 8 + (w & 7);
    const int s = 3 * (kKeyBits - (l + 1));
    const int lo = lower_bound64(keys, n, q << s), hi = lower_bound64(keys, n, (q + 1) << s);
    cq[w] = q;
    flag[w] = (hi - lo > cap) && (l + 1 < kKeyBits);
}

// parents (level l - 1) of the 26 periodic neighbour cells of every box in P (level l)
__global__ void k_balance(int np, const uint64_t* __restrict__ P, int l, uint64_t* out) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= 26 * np) return;
    const uint64_t p = P[w / 26];
    int j = w % 26;
    j = j < 13 ? j : j + 1;  // offset (dx+1) + 3(dy+1) + 9(dz+1), 13 (self) skipped
    const int nl = 1 << l;
    const int cx = ((int)compact3(p) + j % 3 - 1 + nl) & (nl - 1);
    const int cy = ((int)compact3(p >> 1) + (j / 3) % 3 - 1 + nl) & (nl - 1);
    const int cz = ((int)compact3(p >> 2) + j / 9 - 1 + nl) & (nl - 1);
    out[w] = encode(cx, cy, cz) >> 3;
}

// sort keys of the nodes: (prefix << 3 (Dm - l)) << 5 | l, flat node index
__global__ void k_node_keys(int nn, const uint64_t* __restrict__ prefix, const int* __restrict__ level, int Dm,
                            uint64_t* key, int* val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int l = level[i];
    key[i] = (prefix[i] << (3 * (Dm - l))) << 5 | (uint64_t)l;
    val[i] = i;
}

struct Levels {
    // per level: sorted prefixes of all nodes (lv) and of the internal ones (iv)
    const uint64_t* node[kKeyBits + 1];
    int nnode[kKeyBits + 1];
    int base[kKeyBits + 1];  // flat index of the level's first node
    const uint64_t* inl[kKeyBits + 1];
    int ninl[kKeyBits + 1];
    int nlev;                // levels holding nodes
};

__device__ inline int flat_of(const Levels& V, int l, uint64_t q) {
    if (l < 0 || l >= V.nlev) return -1;
    const int k = find64(V.node[l], V.nnode[l], q);
    return k < 0 ? -1 : V.base[l] + k;
}

// per node (pre-order i): level, cell, parent, children, leaf flag and count
__global__ void k_node_info(int nn, Levels V, const int* __restrict__ pre2flat, const int* __restrict__ flat2pre,
                            const uint64_t* __restrict__ fprefix, const int* __restrict__ flevel,
                            const uint64_t* __restrict__ keys, int n, int* level, int* cell, int* parent, int* child,
                            int8_t* is_leaf, int* leaf_count, int* err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int f = pre2flat[i], l = flevel[f];
    const uint64_t q = fprefix[f];
    level[i] = l;
    cell[3 * i] = (int)compact3(q);
    cell[3 * i + 1] = (int)compact3(q >> 1);
    cell[3 * i + 2] = (int)compact3(q >> 2);
    const int pf = l == 0 ? -1 : flat_of(V, l - 1, q >> 3);
    if (l > 0 && pf < 0) atomicOr(err, 2);  // every box below the root has a parent box
    parent[i] = pf < 0 ? -1 : flat2pre[pf];
    const int r = l < V.nlev ? find64(V.inl[l], V.ninl[l], q) : -1;
    is_leaf[i] = r < 0;
    for (int c = 0; c < 8; ++c) child[8 * i + c] = r < 0 ? -1 : flat2pre[V.base[l + 1] + 8 * r + c];
    if (r < 0) {
        const int s = 3 * (kKeyBits - l);
        leaf_count[i] = lower_bound64(keys, n, (q + 1) << s) - lower_bound64(keys, n, q << s);
    } else {
        leaf_count[i] = 0;
    }
}

// descendant-leaf range of every node: its pre-order subtree is [i, end)
// with end the first node whose full-depth key passes the box
__global__ void k_leaf_range(int nn, const uint64_t* __restrict__ skeys, const int* __restrict__ level, int Dm,
                             const int* __restrict__ leafpos, const int* __restrict__ leaves, int n_leaves,
                             int* first_leaf, int* last_leaf, int* err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int l = level[i];
    const uint64_t full = skeys[i] >> 5;
    const uint64_t span = 1ull << (3 * (Dm - l));
    const int end = lower_bound64(skeys, nn, (full + span) << 5);
    const int a = leafpos[i], b = leafpos[end] - 1;  // every subtree holds a leaf: a <= b
    if (!(a <= b && a < n_leaves && b >= 0)) atomicOr(err, 4);
    first_leaf[i] = leaves[min(max(a, 0), n_leaves - 1)];
    last_leaf[i] = leaves[min(max(b, 0), n_leaves - 1)];
}

__global__ void k_colleagues(int nn, Levels V, const int* __restrict__ flat2pre, const int* __restrict__ level,
                             const int* __restrict__ cell, int* colleague) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= 26 * nn) return;
    const int i = w / 26, s = w % 26, l = level[i];
    if (l < 2) {
        colleague[w] = -1;
        return;
    }
    const int j = s < 13 ? s : s + 1;
    const int nl = 1 << l;
    const int cx = (cell[3 * i] + j % 3 - 1 + nl) & (nl - 1);
    const int cy = (cell[3 * i + 1] + (j / 3) % 3 - 1 + nl) & (nl - 1);
    const int cz = (cell[3 * i + 2] + j / 9 - 1 + nl) & (nl - 1);
    const int f = flat_of(V, l, encode(cx, cy, cz));
    colleague[w] = f < 0 ? -1 : flat2pre[f];
}

// Near list of leaf i (count only when ids is null): itself, then for each
// neighbour offset in (dz, dy, dx) order the leaves under the existing
// neighbour box, or the coarser leaf covering the missing cell (once per
// image code).  Leaves below level 2 have none.
__global__ void k_near(int nn, Levels V, const int* __restrict__ flat2pre, const int* __restrict__ level,
                       const int* __restrict__ cell, const int8_t* __restrict__ is_leaf,
                       const int* __restrict__ first_leaf, const int* __restrict__ last_leaf,
                       const int* __restrict__ leafpos, const int* __restrict__ leaves, const int* __restrict__ off,
                       int* cnt, int* ids, int8_t* img) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    if (!is_leaf[i] || level[i] < 2) {
        if (!ids) cnt[i] = 0;
        return;
    }
    const int l = level[i], nl = 1 << l;
    int pos = ids ? off[i] : 0;
    auto put = [&](int id, int code) {
        if (ids) {
            ids[pos] = id;
            img[pos] = (int8_t)code;
        }
        ++pos;
    };
    put(i, 13);
    int cov_id[26], cov_code[26], ncov = 0;
    for (int j = 0; j < 27; ++j) {
        if (j == 13) continue;
        int c[3] = {cell[3 * i] + j % 3 - 1, cell[3 * i + 1] + (j / 3) % 3 - 1, cell[3 * i + 2] + j / 9 - 1};
        int im[3];
        for (int a = 0; a < 3; ++a) {
            im[a] = c[a] < 0 ? -1 : (c[a] >= nl ? 1 : 0);
            c[a] = (c[a] + nl) & (nl - 1);
        }
        const int code = (im[0] + 1) + 3 * (im[1] + 1) + 9 * (im[2] + 1);
        const int f = flat_of(V, l, encode(c[0], c[1], c[2]));
        if (f >= 0) {
            const int nb = flat2pre[f];
            for (int k = leafpos[first_leaf[nb]]; k <= leafpos[last_leaf[nb]]; ++k) put(leaves[k], code);
        } else {
            int cov = -1;
            for (int m = l - 1; m >= 0 && cov < 0; --m) {
                const int g = flat_of(V, m, encode(c[0] >> (l - m), c[1] >> (l - m), c[2] >> (l - m)));
                if (g >= 0) cov = flat2pre[g];
            }
            bool seen = false;
            for (int k = 0; k < ncov; ++k) seen |= cov_id[k] == cov && cov_code[k] == code;
            if (!seen) {
                cov_id[ncov] = cov;
                cov_code[ncov] = code;
                ++ncov;
                put(cov, code);
            }
        }
    }
    if (!ids) cnt[i] = pos;
}

__global__ void k_leaf_flags(int nn, const int8_t* __restrict__ is_leaf, int* f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) f[i] = is_leaf[i];
    if (i == nn) f[i] = 0;
}
__global__ void k_leaf_list(int nn, const int8_t* __restrict__ is_leaf, const int* __restrict__ leafpos, int* leaves) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn && is_leaf[i]) leaves[leafpos[i]] = i;
}
__global__ void k_scatter(int nn, const int* __restrict__ pre2flat, int* flat2pre) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) flat2pre[pre2flat[i]] = i;
}
__global__ void k_fill_level(int cnt, uint64_t* prefix_out, int* level_out, const uint64_t* __restrict__ src, int l) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    prefix_out[i] = src[i];
    level_out[i] = l;
}

inline int grid(long n) { return (int)((n + OB - 1) / OB); }

// device buffers of one build, freed on exit
struct Pool {
    std::vector<void*> p;
    ~Pool() {
        for (void* q : p) (void)hipFree(q);
    }
    template <typename V>
    V* get(size_t n) {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(V));
        if (e != hipSuccess) throw HpdmkError(-3, std::string("hipMalloc (octree build): ") + hipGetErrorString(e));
        p.push_back(q);
        return static_cast<V*>(q);
    }
};

#define OCK(call)                                                                                       \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) throw HpdmkError(-3, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename F>
void cub_call(Pool& pool, F&& f) {
    size_t bytes = 0;
    OCK(f(nullptr, bytes));
    void* tmp = pool.get<char>(bytes);
    OCK(f(tmp, bytes));
}

}  // namespace

Octree build_octree_device(int n, const double* r, double L, int n_per_leaf, hipStream_t st) {
    if (n <= 0) throw HpdmkError(-1, "n_src must be positive");
    if (!(L > 0)) throw HpdmkError(-1, "L must be positive");
    Pool pool;
    // ---- 1. Morton keys, stable radix sort of (key, index)
    double* dr = pool.get<double>(3 * (size_t)n);
    OCK(hipMemcpyAsync(dr, r, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, st));
    uint64_t* k0 = pool.get<uint64_t>(n);
    uint64_t* keys = pool.get<uint64_t>(n);
    int* v0 = pool.get<int>(n);
    int* order = pool.get<int>(n);
    int* err = pool.get<int>(1);
    OCK(hipMemsetAsync(err, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_morton, dim3(grid(n)), dim3(OB), 0, st, n, dr, L, k0, v0, err);
    cub_call(pool, [&](void* t, size_t& b) {
        return hipcub::DeviceRadixSort::SortPairs(t, b, k0, keys, v0, order, n, 0, 3 * kKeyBits, st);
    });
    int herr = 0;
    OCK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
    OCK(hipStreamSynchronize(st));
    if (herr) throw HpdmkError(-1, "particle coordinates must lie in [0, L)");

    // ---- 2. refinement, level by level (internal prefixes per level)
    std::vector<uint64_t*> P(kKeyBits + 1, nullptr);
    std::vector<int> np(kKeyBits + 1, 0);
    int* dcount = pool.get<int>(1);
    if (n > n_per_leaf) {
        P[0] = pool.get<uint64_t>(1);
        OCK(hipMemsetAsync(P[0], 0, sizeof(uint64_t), st));
        np[0] = 1;
    }
    auto children_of = [&](int l, bool internal_only, uint64_t*& out, int& cnt) {
        const int m = 8 * np[l];
        uint64_t* cq = pool.get<uint64_t>(m);
        int* flag = pool.get<int>(m);
        hipLaunchKernelGGL(k_children, dim3(grid(m)), dim3(OB), 0, st, np[l], P[l], l, keys, n, n_per_leaf, cq, flag);
        if (!internal_only) {
            out = cq;
            cnt = m;
            return;
        }
        out = pool.get<uint64_t>(m);
        cub_call(pool, [&](void* t, size_t& b) { return hipcub::DeviceSelect::Flagged(t, b, cq, flag, out, dcount, m, st); });
        OCK(hipMemcpyAsync(&cnt, dcount, sizeof(int), hipMemcpyDeviceToHost, st));
        OCK(hipStreamSynchronize(st));
    };
    for (int l = 0; l < kKeyBits && np[l] > 0; ++l) children_of(l, true, P[l + 1], np[l + 1]);
    // levels with nodes: the root and the children of internal boxes
    int maxl = 0;
    for (int l = 0; l <= kKeyBits; ++l)
        if (np[l] > 0) maxl = l + 1;
    if (maxl + 1 <= 2) throw HpdmkError(-2, "max depth is less than 2, too low for hpdmk");  // src/tree.cpp:352-353
    {  // levels 0 and 1 are always refined
        std::vector<uint64_t> all8(8);
        for (int c = 0; c < 8; ++c) all8[c] = (uint64_t)c;
        P[1] = pool.get<uint64_t>(8);
        OCK(hipMemcpyAsync(P[1], all8.data(), sizeof(uint64_t) * 8, hipMemcpyHostToDevice, st));
        np[1] = 8;
    }
    // ---- 3. periodic 2:1 balance, deepest level first
    int top = 0;
    for (int l = 0; l <= kKeyBits; ++l)
        if (np[l] > 0) top = l;
    for (int l = top; l >= 2; --l) {
        if (np[l] == 0) continue;
        const int m = 26 * np[l] + np[l - 1];
        uint64_t* cand = pool.get<uint64_t>(m);
        hipLaunchKernelGGL(k_balance, dim3(grid(26 * np[l])), dim3(OB), 0, st, np[l], P[l], l, cand);
        if (np[l - 1] > 0)
            OCK(hipMemcpyAsync(cand + 26 * np[l], P[l - 1], sizeof(uint64_t) * np[l - 1], hipMemcpyDeviceToDevice, st));
        uint64_t* sorted = pool.get<uint64_t>(m);
        cub_call(pool, [&](void* t, size_t& b) {
            return hipcub::DeviceRadixSort::SortKeys(t, b, cand, sorted, m, 0, 3 * (l - 1) + 1, st);
        });
        uint64_t* uniq = pool.get<uint64_t>(m);
        cub_call(pool, [&](void* t, size_t& b) { return hipcub::DeviceSelect::Unique(t, b, sorted, uniq, dcount, m, st); });
        OCK(hipMemcpyAsync(&np[l - 1], dcount, sizeof(int), hipMemcpyDeviceToHost, st));
        OCK(hipStreamSynchronize(st));
        P[l - 1] = uniq;
    }
    // ---- 4. nodes: root, then the children of every internal box, per level
    Levels V{};
    int nlev = 0;
    for (int l = 0; l <= kKeyBits; ++l)
        if (np[l] > 0) nlev = l + 2;
    if (nlev > kKeyBits + 1) nlev = kKeyBits + 1;
    const int Dm = nlev - 1;  // deepest level
    if (3 * Dm + 5 > 64) return Octree{};  // deeper than the pre-order keys hold: the host builds it
    std::vector<int> nnode(nlev, 0);
    nnode[0] = 1;
    for (int l = 1; l < nlev; ++l) nnode[l] = 8 * np[l - 1];
    int nn = 0;
    for (int l = 0; l < nlev; ++l) {
        V.base[l] = nn;
        nn += nnode[l];
    }
    uint64_t* fprefix = pool.get<uint64_t>(nn);
    int* flevel = pool.get<int>(nn);
    {
        const uint64_t zero = 0;
        const int l0 = 0;
        OCK(hipMemcpyAsync(fprefix, &zero, sizeof(uint64_t), hipMemcpyHostToDevice, st));
        OCK(hipMemcpyAsync(flevel, &l0, sizeof(int), hipMemcpyHostToDevice, st));
    }
    for (int l = 1; l < nlev; ++l) {
        uint64_t* cq = nullptr;
        int cnt = 0;
        children_of(l - 1, false, cq, cnt);  // sorted: parents sorted, digits ascending
        hipLaunchKernelGGL(k_fill_level, dim3(grid(cnt)), dim3(OB), 0, st, cnt, fprefix + V.base[l],
                           flevel + V.base[l], cq, l);
    }
    V.nlev = nlev;
    for (int l = 0; l < nlev; ++l) {
        V.node[l] = fprefix + V.base[l];
        V.nnode[l] = nnode[l];
        V.inl[l] = P[l];
        V.ninl[l] = np[l];
    }
    // Morton pre-order: sort (full-depth key, level)
    uint64_t* nk0 = pool.get<uint64_t>(nn);
    uint64_t* skeys = pool.get<uint64_t>(nn);
    int* nv0 = pool.get<int>(nn);
    int* pre2flat = pool.get<int>(nn);
    int* flat2pre = pool.get<int>(nn);
    hipLaunchKernelGGL(k_node_keys, dim3(grid(nn)), dim3(OB), 0, st, nn, fprefix, flevel, Dm, nk0, nv0);
    cub_call(pool, [&](void* t, size_t& b) {
        return hipcub::DeviceRadixSort::SortPairs(t, b, nk0, skeys, nv0, pre2flat, nn, 0, 3 * Dm + 5, st);
    });
    hipLaunchKernelGGL(k_scatter, dim3(grid(nn)), dim3(OB), 0, st, nn, pre2flat, flat2pre);
    // ---- 5. per-node tables
    int* level = pool.get<int>(nn);
    int* cell = pool.get<int>(3 * (size_t)nn);
    int* parent = pool.get<int>(nn);
    int* child = pool.get<int>(8 * (size_t)nn);
    int8_t* is_leaf = pool.get<int8_t>(nn);
    int* leaf_count = pool.get<int>(nn);
    hipLaunchKernelGGL(k_node_info, dim3(grid(nn)), dim3(OB), 0, st, nn, V, pre2flat, flat2pre, fprefix, flevel, keys,
                       n, level, cell, parent, child, is_leaf, leaf_count, err);
    int* lflag = pool.get<int>(nn + 1);
    int* leafpos = pool.get<int>(nn + 1);
    hipLaunchKernelGGL(k_leaf_flags, dim3(grid(nn + 1)), dim3(OB), 0, st, nn, is_leaf, lflag);
    cub_call(pool, [&](void* t, size_t& b) { return hipcub::DeviceScan::ExclusiveSum(t, b, lflag, leafpos, nn + 1, st); });
    int n_leaves = 0;
    OCK(hipMemcpyAsync(&n_leaves, leafpos + nn, sizeof(int), hipMemcpyDeviceToHost, st));
    OCK(hipStreamSynchronize(st));
    int* leaves = pool.get<int>(n_leaves);
    hipLaunchKernelGGL(k_leaf_list, dim3(grid(nn)), dim3(OB), 0, st, nn, is_leaf, leafpos, leaves);
    int* first_leaf = pool.get<int>(nn);
    int* last_leaf = pool.get<int>(nn);
    hipLaunchKernelGGL(k_leaf_range, dim3(grid(nn)), dim3(OB), 0, st, nn, skeys, level, Dm, leafpos, leaves, n_leaves,
                       first_leaf, last_leaf, err);
    OCK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
    OCK(hipStreamSynchronize(st));
    if (herr) throw HpdmkError(-2, "device octree build: inconsistent tree (code " + std::to_string(herr) + ")");
    int* colleague = pool.get<int>(26 * (size_t)nn);
    hipLaunchKernelGGL(k_colleagues, dim3(grid(26l * nn)), dim3(OB), 0, st, nn, V, flat2pre, level, cell, colleague);
    int* ncnt = pool.get<int>(nn + 1);
    int* noff = pool.get<int>(nn + 1);
    hipLaunchKernelGGL(k_near, dim3(grid(nn)), dim3(OB), 0, st, nn, V, flat2pre, level, cell, is_leaf, first_leaf,
                       last_leaf, leafpos, leaves, nullptr, ncnt, nullptr, nullptr);
    OCK(hipMemsetAsync(ncnt + nn, 0, sizeof(int), st));
    cub_call(pool, [&](void* t, size_t& b) { return hipcub::DeviceScan::ExclusiveSum(t, b, ncnt, noff, nn + 1, st); });
    int n_near = 0;
    OCK(hipMemcpyAsync(&n_near, noff + nn, sizeof(int), hipMemcpyDeviceToHost, st));
    OCK(hipStreamSynchronize(st));
    int* nids = pool.get<int>(n_near);
    int8_t* nimg = pool.get<int8_t>(n_near);
    hipLaunchKernelGGL(k_near, dim3(grid(nn)), dim3(OB), 0, st, nn, V, flat2pre, level, cell, is_leaf, first_leaf,
                       last_leaf, leafpos, leaves, noff, ncnt, nids, nimg);
    OCK(hipGetLastError());

    // ---- download into the host structure
    Octree T;
    T.n_nodes = nn;
    T.max_depth = Dm + 1;
    auto down = [&](auto& vec, const auto* src, size_t cnt) {
        vec.resize(cnt);
        if (cnt) OCK(hipMemcpyAsync(vec.data(), src, sizeof(vec[0]) * cnt, hipMemcpyDeviceToHost, st));
    };
    down(T.level, level, nn);
    down(T.cell, cell, 3 * (size_t)nn);
    down(T.parent, parent, nn);
    down(T.child, child, 8 * (size_t)nn);
    down(T.is_leaf, is_leaf, nn);
    down(T.leaf_count, leaf_count, nn);
    down(T.first_leaf, first_leaf, nn);
    down(T.last_leaf, last_leaf, nn);
    down(T.colleague, colleague, 26 * (size_t)nn);
    down(T.near_off, noff, nn + 1);
    down(T.near_ids, nids, n_near);
    down(T.near_img, nimg, n_near);
    down(T.order, order, n);
    OCK(hipStreamSynchronize(st));
    return T;
}

}  // namespace hpdmk_amd
