// DF preparation in front of the closed-shell path: MO density-fitting factors from AO
// three-index integrals (generate_AO_DF_integrals / generate_3idx_integrals,
// src/integrals/dftools.jl:28-125).
#pragma once
#include <vector>

namespace ecg {

// sqrtinvchol (src/tools/tensortools.jl:377-391), host side: pivoted Cholesky
// A[p,p] = U^T U (largest remaining diagonal first, stopped once it is <= tol, the absolute
// criterion of LAPACK dpstrf), then X = U^{-1} at full rank or the minimum-norm solution of
// U[1:r,:] X = I_r (X = U^T (U U^T)^{-1}) below it, rows put back in the original order.
// Returns M (n x r, column-major) with M M^T = A^{-1} (the pseudo-inverse at reduced rank).
std::vector<double> sqrtinvchol(const double* A, int n, double tol, int& rank);

// B[p,q,L] = sum_{mu,nu,P} C[mu,p] AAP[mu,nu,P] M[P,L] C[nu,q] on HIP device `device` with
// FP64 MFMA contractions; C nao x nmo, AAP nao x nao x nfit, M nfit x r, B nmo x nmo x r
// (all host, column-major).
void df_ao_to_mo_device(int device, const double* C, int nao, int nmo, const double* AAP, int nfit,
                        const double* M, int r, double* B);

}  // namespace ecg
