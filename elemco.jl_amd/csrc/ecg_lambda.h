// Lambda (CC Lagrange multiplier) residual, iterations and the one-particle density.
#pragma once
#include "ecg_cc.h"

namespace ecg {
void lambda_residual_host(Context& c, int method, Prec pr, const double* T1h, const double* T2h, const double* U1h,
                          const double* U2h, double* R1h, double* R2h);
void lambda_solve(Context& c, int method, Prec pr, const SolverOpts& o, const double* T1h, const double* T2h, double* U1h,
                  double* U2h, int* iterations, int* converged);
void one_rdm_host(Context& c, const double* T1h, const double* T2h, const double* U1h, const double* U2h,
                  double* D1h, double* dD1h);
}  // namespace ecg
