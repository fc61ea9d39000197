// Device state and launchers of the FP64 Ewald reporter (ewald.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "dmk_common.hpp"

namespace hpdmk_amd {

struct EwaldDev {
    int n;        // particles
    int nk;       // modes per axis: k_a = (i - nk) dk, |i - nk| <= nk
    int d;        // 2 nk + 1
    size_t K;     // half-cube modes d * d * (nk + 1)
    double L, alpha, eps, V, rc, dk;
    double *x, *y, *z, *q;  // particle coordinates and charges (FP64)
    Cplx<double>* S;        // structure factor on the half cube
    const double* g;        // e^{-k^2/4 alpha^2}/k^2 times the half-cube weight (0 at k = 0)
};

// one trial move of a chain: idx/d are the move list, rows/tgt/acc/dE scratch
struct EwaldMove {
    const long* idx;
    const double* d;
    Cplx<double>* rows;  // 6 d: [endpoint t, o][axis][mode]
    double* tgt;         // 3
    double* acc;         // 1 (self-resetting)
    double* dE;          // per move
};

void ewald_structure_factor(const EwaldDev& e, hipStream_t st);
// out2[0] += short-range sum, out2[1] += sum g |S|^2 (callers zero out2 first)
void ewald_energy_parts(const EwaldDev& e, double* out2, hipStream_t st);
void ewald_potential(const EwaldDev& e, int n_pts, const double* pts, double* phi, hipStream_t st);
// dE of move k into mv.dE[k]; with apply, S and the particle follow the move
void ewald_shift(const EwaldDev& e, const EwaldMove& mv, int k, bool apply, hipStream_t st);

}  // namespace hpdmk_amd
