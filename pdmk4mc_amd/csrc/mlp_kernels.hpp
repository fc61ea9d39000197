// Launchers of the proposal-MLP training kernels (mlp_train.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "dmk_proposal.hpp"

namespace hpdmk_amd {

// training samples on the device: features X[n][8], displacements D[n][3],
// weights wt[n]; output scaling of the proposal
struct MlpData {
    const float* X;
    const float* D;
    const float* wt;
    float max_mean, smin, smax;
};

struct AdamStep {
    float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
};

int mlp_blocks(int n);
// per-block partial gradients (mlp_blocks(n) x kMlpParams) and losses of samples sel[b0 .. b0 + n) (sel may be null)
void launch_mlp_grad(const MlpData& d, const int* sel, int b0, int n, const float* w, float* part, float* loss_part,
                     hipStream_t st);
// grad = scale * sum of partials; then an Adam step unless s.lr <= 0
void launch_mlp_adam(float* w, float* m, float* v, const float* part, int nblk, float* grad, float scale,
                     const AdamStep& s, hipStream_t st);
// the deployed BF16 matrix-core inference of n samples X[n][8]: out[n][6] = mu, sigma
void launch_mlp_infer(const Proposal& p, const float* X, int n, float* out, hipStream_t st);

}  // namespace hpdmk_amd
