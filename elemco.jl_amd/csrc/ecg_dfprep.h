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

// Frozen orbitals on density-fitted integrals, the DF counterpart of freeze_orbs_in_dump
// (src/integrals/dumptools.jl:20-62) and of freeze_core! / freeze_nvirt! after the Fock
// matrix of all occupied orbitals (src/cc/ccdriver.jl:93-98, src/infos/ecinfos.jl:331-392):
// (pq|rs) = sum_Q BL[p,q,Q] BR[r,s,Q] (BR = BL when br is null).  Orbitals of `orbs`
// (1-based) below nocc form the core c: core_fock[p,q] = sum_c 2 (pq|cc) - (pc|cq),
// e0 += 2 sum_c h[c,c] + core_fock[c,c], h <- (h + core_fock)[kept, kept], factors <- their
// kept (p,q) block; frozen virtuals are dropped.  Numbers outside 1..n and repeats are
// ignored.  Outputs: kept x kept (x naux) arrays, column-major.
struct DfFreeze {
  int ncore = 0, nvirt = 0, nkept = 0;
  double e0 = 0.0;
  std::vector<double> h, BL, BR;   // BR empty when br is null
};
DfFreeze df_freeze(int n, int nocc, int naux, const double* h, const double* bl, const double* br, double e0,
                   const std::vector<int>& orbs);

}  // namespace ecg
