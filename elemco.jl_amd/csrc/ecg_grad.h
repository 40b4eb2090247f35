// Gradient of the closed-shell CC correlation energy with respect to h and the DF factors.
#pragma once
#include "ecg_cc.h"

namespace ecg {
// dEdh: n x n, dEdBL / dEdBR: n x n x naux (column-major, any may be null); T, U host
// arrays (converged amplitudes and Lambda multipliers).  Hermitian integrals: dEdBL
// receives dE/dB and dEdBR is not written; two-sided factors: dE/dBL and dE/dBR.
void energy_grad(Context& c, int method, const double* T1h, const double* T2h, const double* U1h, const double* U2h,
                 double* dEdh, double* dEdBL, double* dEdBR);
}  // namespace ecg
