// Allocation-free variant of hpdmk_eval_shift_energy_batch (include/hpdmk.h):
// the move list and the energies are staged in a file-scope device array
// instead of buffers the library allocates, so the call makes no hipMalloc or
// hipFree (for callers that cannot allocate: code under HIP graph capture,
// tight serving loops, harnesses that forbid it).  The kernels are the ones
// hpdmk_eval_shift_energy_batch launches.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "hpdmk_handle.hpp"

using namespace hpdmk_amd;
using namespace hpdmk_amd::capi;

namespace {

// Staging of one call: caller indices (int), displacements (3 doubles per
// move) and energies (1 double per move), at most kNoAllocMaxMoves moves.
constexpr int kNoAllocMaxMoves = 4096;
constexpr size_t kIdxOff = 0;                                              // bytes
constexpr size_t kDOff = kIdxOff + sizeof(int) * kNoAllocMaxMoves;         // 16 KiB
constexpr size_t kEOff = kDOff + sizeof(double) * 3 * kNoAllocMaxMoves;    // + 96 KiB
constexpr size_t kScratchBytes = kEOff + sizeof(double) * kNoAllocMaxMoves;  // + 32 KiB = 144 KiB

// Every GPU the library runs on holds its own copy until the process ends.
// In host code `scratch` names a host-side stand-in, not the device copy: host
// code uses only the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char* g_ptr[kMaxDevices] = {};

// device address of `scratch` on `dev`, looked up once per device with that
// device current
char* scratch_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw HpdmkError(HPDMK_EINVAL, "device id out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        int cur = 0;
        HIPX(hipGetDevice(&cur));
        if (cur != dev) HIPX(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) HIPX(hipSetDevice(cur));
        if (e != hipSuccess)
            throw HpdmkError(HPDMK_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[dev] = static_cast<char*>(p);
    }
    return g_ptr[dev];
}

}  // namespace

extern "C" int hpdmk_eval_shift_energy_batch_noalloc(hpdmk_tree t, int n, const long* idx, const double* d,
                                                     double* dE) {
    return guard([&] {
        Handle* h = H(t);
        if (n < 0) throw HpdmkError(HPDMK_EINVAL, "n must be non-negative");
        if (n > kNoAllocMaxMoves)
            throw HpdmkError(HPDMK_EINVAL, "hpdmk_eval_shift_energy_batch_noalloc takes at most " +
                                               std::to_string(kNoAllocMaxMoves) + " moves per call");
        if (n == 0) return;
        if (!idx || !d || !dE) throw HpdmkError(HPDMK_EINVAL, "null argument");
        for (int k = 0; k < n; ++k)
            if (idx[k] < 0 || idx[k] >= h->n_src) throw HpdmkError(HPDMK_EINVAL, "particle index out of range");
        char* base = scratch_on(h->device);
        MoveBuf b;  // views into the scratch array; nothing is allocated
        b.n = b.cap = n;
        b.idx = reinterpret_cast<int*>(base + kIdxOff);
        b.d = reinterpret_cast<double*>(base + kDOff);
        b.dE = reinterpret_cast<double*>(base + kEOff);
        std::vector<int> i32(idx, idx + n);
        HIPX(hipMemcpyAsync(b.idx, i32.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->stream));
        HIPX(hipMemcpyAsync(b.d, d, sizeof(double) * 3 * n, hipMemcpyHostToDevice, h->stream));
        for (int m0 = 0; m0 < n; m0 += kMaxBatch) {
            const int k = std::min(kMaxBatch, n - m0);
            dispatch(h, [&](auto& dv) {
                McArgs a = mc_args(h, b, m0, k, kTrialOnly, 0.0);
                a.ml.stats = nullptr;
                launch_setup(dv, a, h->stream);
                launch_trial(dv, a, h->stream);
                launch_decide(dv, a, h->stream);
            });
        }
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(dE, b.dE, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
        check_err(h);  // synchronises the stream
    });
}
