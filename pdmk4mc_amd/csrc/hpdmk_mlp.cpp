// C ABI of the proposal-MLP training (include/hpdmk.h): maximum likelihood of
// weighted displacements, Adam, on the GPU.  No reference counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "dmk_proposal.hpp"
#include "hpdmk_handle.hpp"
#include "mlp_kernels.hpp"

using namespace hpdmk_amd;
using namespace hpdmk_amd::capi;

namespace {

struct MlpDevice {
    std::vector<void*> owned;
    hipStream_t stream = nullptr;
    MlpData d{};
    float* w = nullptr;
    ~MlpDevice() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : owned) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

void upload_data(MlpDevice& m, int n, const double* feat, const double* delta, const double* weight,
                 const double* cfg3) {
    if (n < 1 || !feat || !delta || !weight || !cfg3) throw HpdmkError(HPDMK_EINVAL, "empty or null training data");
    if (!(cfg3[1] > 0) || !(cfg3[2] > cfg3[1]) || !(cfg3[0] >= 0))
        throw HpdmkError(HPDMK_EINVAL, "need 0 < sigma_min < sigma_max and max_mean >= 0");
    std::vector<float> X(feat, feat + (size_t)kFeat * n), D(delta, delta + 3 * (size_t)n), W(weight, weight + n);
    HIPX(hipStreamCreate(&m.stream));
    m.d.X = dupload(X, m.owned);
    m.d.D = dupload(D, m.owned);
    m.d.wt = dupload(W, m.owned);
    m.d.max_mean = (float)cfg3[0];
    m.d.smin = (float)cfg3[1];
    m.d.smax = (float)cfg3[2];
}

}  // namespace

extern "C" {

int hpdmk_mlp_eval(int n, const float* weights, const double* feat, const double* delta, const double* weight,
                   const double* cfg3, double* mu_sigma, float* grad, double* loss) {
    return guard([&] {
        if (!weights) throw HpdmkError(HPDMK_EINVAL, "null weights");
        MlpDevice m;
        upload_data(m, n, feat, delta, weight, cfg3);
        std::vector<float> w(weights, weights + kMlpParams);
        m.w = dupload(w, m.owned);
        std::vector<unsigned short> packed(kPackSize);
        mlp_pack_weights(weights, packed.data());
        const unsigned short* wp = dupload(packed, m.owned);
        const int nb = mlp_blocks(n);
        float* part = dalloc<float>((size_t)nb * kMlpParams, m.owned);
        float* lp = dalloc<float>(nb, m.owned);
        float* g = dalloc<float>(kMlpParams, m.owned);
        float* out = dalloc<float>(6 * (size_t)n, m.owned);
        launch_mlp_grad(m.d, nullptr, 0, n, m.w, part, lp, m.stream);
        launch_mlp_adam(m.w, nullptr, nullptr, part, nb, g, 1.0f, AdamStep{0, 0, 0, 0, 1, 1}, m.stream);
        Proposal pr{};
        pr.learned = 1;
        pr.w = m.w;
        pr.wp = wp;
        pr.max_mean = cfg3[0];
        pr.smin = cfg3[1];
        pr.smax = cfg3[2];
        launch_mlp_infer(pr, m.d.X, n, out, m.stream);
        HIPX(hipGetLastError());
        std::vector<float> o(6 * (size_t)n), l(nb);
        HIPX(hipMemcpyAsync(o.data(), out, sizeof(float) * o.size(), hipMemcpyDeviceToHost, m.stream));
        HIPX(hipMemcpyAsync(l.data(), lp, sizeof(float) * nb, hipMemcpyDeviceToHost, m.stream));
        if (grad) HIPX(hipMemcpyAsync(grad, g, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m.stream));
        HIPX(hipStreamSynchronize(m.stream));
        if (mu_sigma) std::copy(o.begin(), o.end(), mu_sigma);
        if (loss) *loss = std::accumulate(l.begin(), l.end(), 0.0);
    });
}

int hpdmk_mlp_train(int n, const double* feat, const double* delta, const double* weight, const double* cfg,
                    float* weights, double* epoch_loss) {
    return guard([&] {
        if (!weights || !cfg) throw HpdmkError(HPDMK_EINVAL, "null weights or configuration");
        const double lr = cfg[3];
        const int epochs = (int)cfg[4], batch = (int)cfg[5];
        const unsigned seed = (unsigned)cfg[6];
        const bool by_count = cfg[7] != 0.0;  // signed weights (policy gradient): mean over the minibatch
        if (!(lr > 0) || epochs < 1 || batch < 1) throw HpdmkError(HPDMK_EINVAL, "need lr > 0, epochs >= 1, batch >= 1");
        MlpDevice m;
        upload_data(m, n, feat, delta, weight, cfg);
        std::vector<float> w(weights, weights + kMlpParams);
        m.w = dupload(w, m.owned);
        std::vector<float> zero(kMlpParams, 0.0f);
        float* mom = dupload(zero, m.owned);
        float* var = dupload(zero, m.owned);
        const int bs = std::min(batch, n);
        const int nbmax = mlp_blocks(bs);
        float* part = dalloc<float>((size_t)nbmax * kMlpParams, m.owned);
        float* lp = dalloc<float>((size_t)nbmax * ((n + bs - 1) / bs), m.owned);
        int* sel = dalloc<int>(n, m.owned);
        std::vector<int> perm(n);
        std::iota(perm.begin(), perm.end(), 0);
        std::mt19937 rng(seed);
        double wsum_all = 0;
        for (int i = 0; i < n; ++i) wsum_all += by_count ? 1.0 : weight[i];
        if (!(wsum_all > 0)) throw HpdmkError(HPDMK_EINVAL, "all sample weights are zero");
        const float b1 = 0.9f, b2 = 0.999f;
        long step = 0;
        for (int ep = 0; ep < epochs; ++ep) {
            std::shuffle(perm.begin(), perm.end(), rng);
            HIPX(hipMemcpyAsync(sel, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, m.stream));
            int nbat = 0;
            for (int b0 = 0; b0 < n; b0 += bs, ++nbat) {
                const int cnt = std::min(bs, n - b0);
                double wsum = 0;
                for (int k = 0; k < cnt; ++k) wsum += by_count ? 1.0 : weight[perm[b0 + k]];
                ++step;
                const AdamStep st{(float)lr, b1, b2, 1e-8f, 1.0f - std::pow(b1, (float)step),
                                  1.0f - std::pow(b2, (float)step)};
                launch_mlp_grad(m.d, sel, b0, cnt, m.w, part, lp + (size_t)nbat * nbmax, m.stream);
                launch_mlp_adam(m.w, mom, var, part, mlp_blocks(cnt), nullptr, wsum > 0 ? (float)(1.0 / wsum) : 0.0f, st,
                                m.stream);
            }
            HIPX(hipGetLastError());
            if (epoch_loss) {
                std::vector<float> l((size_t)nbat * nbmax, 0.0f);
                // blocks past a short last minibatch did not write: count only written ones
                HIPX(hipMemcpyAsync(l.data(), lp, sizeof(float) * l.size(), hipMemcpyDeviceToHost, m.stream));
                HIPX(hipStreamSynchronize(m.stream));
                double tot = 0;
                for (int b = 0; b < nbat; ++b) {
                    const int cnt = std::min(bs, n - b * bs);
                    for (int k = 0; k < mlp_blocks(cnt); ++k) tot += l[(size_t)b * nbmax + k];
                }
                epoch_loss[ep] = tot / wsum_all;
            }
        }
        HIPX(hipMemcpyAsync(weights, m.w, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m.stream));
        HIPX(hipStreamSynchronize(m.stream));
    });
}

}  // extern "C"
