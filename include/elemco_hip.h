/* elemco_hip.h -- C ABI of the MI355X (gfx950) closed-shell coupled-cluster library.
 *
 * Drop-in replacement for the closed-shell hot path of ElemCo.jl v0.14.0
 * (file:line citations are relative to that source tree).  All arrays are
 * column-major (Julia layout, first index fastest) and FP64 unless stated.
 * Host pointers are owned by the caller; the context owns all device memory.
 * Every call is synchronous (the library's HIP stream is drained on return).
 * Status: 0 = success, negative = error; ecg_last_error() returns the message
 * (the Julia binding raises it with error(), the reference's convention).
 * Orbital spaces: occupied = first nocc orbitals, virtual = the next nvirt
 * (setup_space_fd!, src/infos/ecinfos.jl:171-192); n = nocc + nvirt.
 * Integrals are density-fitted: (pq|rs) = sum_Q BL[p,q,Q] * BR[r,s,Q]
 * (BR = NULL means BR = BL, the Hermitian case; layout of mmL[p,q,L] in
 * src/integrals/dftools.jl:102).  A four-index FCIDUMP is brought to this form
 * by an exact factorisation of its (pq|rs) supermatrix (host side).
 */
#ifndef ELEMCO_HIP_H
#define ELEMCO_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ecg_ctx ecg_ctx;

enum { ECG_CCSD = 0, ECG_DCSD = 1 };
enum { ECG_FP64 = 0, ECG_FP32 = 1, ECG_BF16X3 = 2 };

/* Iteration options; defaults of CcOptions/DiisOptions (src/infos/options.jl:128-221,272-282). */
typedef struct {
  double thr;        /* squared residual norm threshold (1e-10) */
  double conven;     /* energy threshold factor: |dE| < sqrt(thr)*conven (0.1) */
  double shifts;     /* singles level shift (0.15) */
  double shiftp;     /* doubles level shift (0.2) */
  double resthr;     /* DIIS start threshold (10.0) */
  int maxit;         /* maximum iterations (50) */
  int maxdiis;       /* DIIS subspace (6) */
  int refine_every;  /* mixed precision: FP64 residual every k-th iteration (0: never) */
  int graphs;        /* 1: the residual's launch sequence is captured once in a HIP graph and
                        replayed every iteration (for launch-bound small systems); 0: eager */
} ecg_solver_opts;

void ecg_default_opts(ecg_solver_opts* opts);

/* Create a context on HIP device `device`.  h: n x n one-electron integrals,
 * BL (and BR or NULL): n x n x naux DF factors, e0: core energy.
 * workspace_bytes = 0 picks a size from (nocc, nvirt, naux).
 * Replaces the FCIDUMP set-up of ccdriver (src/cc/ccdriver.jl:37-76, read_fcidump
 * src/integrals/dump.jl:326) for the closed-shell path. */
int ecg_create(ecg_ctx** ctx, int nocc, int nvirt, int naux, const double* h, const double* BL,
               const double* BR, double e0, int device, size_t workspace_bytes);
int ecg_destroy(ecg_ctx* ctx);
/* Message of the last failed call on ctx; with ctx = NULL, of the last failed
 * ecg_create* or ecg_fcidump_* call of the calling thread. */
const char* ecg_last_error(const ecg_ctx* ctx);

/* FCIDUMP input, host side (no GPU): read_fcidump (src/integrals/dump.jl:326-346)
 * for closed-shell (IUHF = 0) files, plain or gzip-compressed.  The header
 * (read_header, dump.jl:360-405) is kept with the reference's typing; integrals
 * come from the NPY files named by NPY2/NPY1 when present (Fortran-ordered
 * float64, int2 = <pq|rs> at [p,q,r+s(s+1)/2], r <= s; core energy ENUC;
 * dump.jl:442-461, 731-745; a missing file falls back to the integral lines) or
 * from the integral lines (read_integrals!, dump.jl:573-604; set_int1!/set_int2!,
 * dump.jl:472-544: particle symmetry always, 8-fold symmetry unless ST > 0).
 * Errors: message in ecg_last_error(NULL). */
typedef struct ecg_fcidump ecg_fcidump;
int ecg_fcidump_read(const char* path, ecg_fcidump** fd);
/* An in-memory dump from arrays: the FDump constructor of the reference
 * (src/integrals/dump.jl:107-128) for four-index integrals already in memory.  h: n x n,
 * eri: (pq|rs) at [p,q,r,s] (n^4, chemists' notation); header NORB = norb, NELEC, MS2,
 * ISYM = 1, ORBSYM all 1, and ST when st > 0 (similarity-transformed integrals: particle
 * symmetry only).  The arrays are copied; the dump then serves ecg_fcidump_df,
 * ecg_fcidump_freeze, ecg_fcidump_write and ecg_create_from_fcidump_handle. */
int ecg_fcidump_from_integrals(int norb, int nelec, int ms2, int st, const double* h, const double* eri, double e0,
                               ecg_fcidump** fd);
int ecg_fcidump_free(ecg_fcidump* fd);
/* NORB, NELEC (-1 if absent), MS2 (0 if absent), ST and the core energy; any
 * output may be NULL. */
int ecg_fcidump_info(const ecg_fcidump* fd, int* norb, int* nelec, int* ms2, int* st, double* e0);
/* headvar (dump.jl:664-696): the values of header key `key` (upper case) joined by
 * ',' into buf (NUL-terminated), and its type ('i' Int, 'f' Float64, 's' String) in
 * *type.  Returns 1 if the key is absent, -2 if buflen is too small. */
int ecg_fcidump_header(const ecg_fcidump* fd, const char* key, char* buf, int buflen, char* type);
/* Name of header key number `index` (0-based, sorted); 1 past the last key. */
int ecg_fcidump_header_key(const ecg_fcidump* fd, int index, char* buf, int buflen);
/* h[p,q] (n x n) and (pq|rs) (n^4, chemists' notation, column-major); either may be NULL. */
int ecg_fcidump_integrals(const ecg_fcidump* fd, double* h, double* eri);
/* Exact DF factors of the integrals: (pq|rs) = sum_Q BL[p,q,Q] BR[r,s,Q] by
 * diagonally pivoted LDL^T of the (pq),(rs) supermatrix (Cholesky for ST = 0),
 * stopped at pivots <= rtol * the largest diagonal (rtol <= 0: 1e-14).  Computed on
 * the first call and kept with fd; a call with BL = NULL only returns *naux and
 * *has_br (BR differs from BL: ST > 0 with negative pivots), so BL and BR are
 * n x n x naux arrays the caller allocates.  No reference counterpart: the
 * reference keeps four-index integrals; the GPU path takes factors. */
int ecg_fcidump_df(ecg_fcidump* fd, double rtol, int* naux, double* BL, double* BR, int* has_br);
/* ecg_fcidump_read + ecg_fcidump_df + ecg_create with nocc = NELEC / 2: the
 * ccdriver(EC, method; fcidump) set-up (src/cc/ccdriver.jl:37-76) in one call. */
int ecg_create_from_fcidump(ecg_ctx** ctx, const char* path, double rtol, int device, size_t workspace_bytes);
/* The same from a dump already in memory (read, possibly frozen); fd stays the caller's. */
int ecg_create_from_fcidump_handle(ecg_ctx** ctx, ecg_fcidump* fd, double rtol, int device, size_t workspace_bytes);
/* freeze_orbs_in_dump (src/integrals/dumptools.jl:20-62; the @freeze_orbs macro,
 * src/ElemCo.jl:703-725), closed-shell branch: orbitals orbs[0..norbs-1] (1-based) of
 * the occupied space (the lowest NELEC/2) become a frozen core folded into h and the
 * core energy (core Fock sum_c 2 (pq|cc) - (pc|cq); int0 += 2 sum_c h_cc + core Fock_cc),
 * those of the virtual space are dropped; the integrals keep the other orbitals in order,
 * NORB and NELEC shrink, ORBSYM becomes all 1.  Numbers outside 1..NORB and repeats are
 * ignored, as in the reference.  *ncore / *nfrozvirt (may be NULL): orbitals frozen in
 * each space.  Drops DF factors computed before. */
int ecg_fcidump_freeze(ecg_fcidump* fd, const int* orbs, int norbs, int* ncore, int* nfrozvirt);
/* write_fcidump (src/integrals/dump.jl:752-945) for IUHF = 0: header (NORB, NELEC, MS2,
 * ISYM, ORBSYM first; IUHF/ST/III omitted when 0), then "%23.15e %3i %3i %3i %3i" lines
 * of (pq|rs) in the reference's loop order (8-fold unique set, or the particle-symmetric
 * set for ST > 0), h, and the core energy; values with |v| <= tol are skipped
 * (tol < 0: the reference's default 1e-12). */
int ecg_fcidump_write(const ecg_fcidump* fd, const char* path, double tol);
/* Frozen orbitals on density-fitted integrals (host side): freeze_core! / freeze_nvirt!
 * after the Fock matrix of all occupied orbitals (src/cc/ccdriver.jl:93-98,
 * src/infos/ecinfos.jl:331-392), i.e. freeze_orbs_in_dump (src/integrals/dumptools.jl:20-62)
 * on (pq|rs) = sum_Q BL[p,q,Q] BR[r,s,Q] (BR NULL: = BL).  orbs[0..norbs-1] (1-based) below
 * nocc are the core, folded into h and the core energy (core Fock sum_c 2 (pq|cc) - (pc|cq)),
 * the others are dropped; out of range numbers and repeats are ignored.  Outputs: *nkept,
 * *ncore, h_out (nkept^2), BL_out / BR_out (nkept^2 * naux; BR_out only with BR), *e0_out;
 * a call with NULL arrays returns the counts only.  The frozen system is then
 * ecg_create(nocc - ncore, nkept - nocc + ncore, naux, h_out, BL_out, BR_out, e0_out, ...). */
int ecg_df_freeze(int n, int nocc, int naux, const double* h, const double* BL, const double* BR, double e0,
                  const int* orbs, int norbs, int* nkept, int* ncore, double* h_out, double* BL_out, double* BR_out,
                  double* e0_out);
/* DF preparation in front of the path: MO factors from AO three-index integrals, as
 * generate_AO_DF_integrals / generate_3idx_integrals (src/integrals/dftools.jl:28-125)
 * after the integral library has produced them.  C: nao x nmo MO coefficients,
 * AAP: (mu nu|P) nao x nao x nfit, PQ: metric (P|Q) nfit x nfit.  M = sqrtinvchol(PQ, tol)
 * (src/tools/tensortools.jl:377-391: pivoted Cholesky stopped at pivots <= tol, then
 * U^{-1}, or the minimum-norm solution when the metric is rank-deficient; host side),
 * B[p,q,L] = sum C[mu,p] (mu nu|P) M[P,L] C[nu,q] by FP64 MFMA contractions on `device`.
 * B must hold nmo*nmo*nfit doubles; the first *naux = rank slices are written. */
int ecg_df_ao_to_mo(const double* C, int nao, int nmo, const double* AAP, const double* PQ, int nfit, double tol,
                    int device, int* naux, double* B);
/* ecg_df_ao_to_mo + h_MO = C^T h_AO C + ecg_create: the DF set-up of a CC calculation
 * (`@dfhf; @cc`, examples/h2o.jl; dftools.jl:92-125) on the GPU path. */
int ecg_create_from_ao(ecg_ctx** ctx, int nocc, const double* C, int nao, int nmo, const double* h_ao,
                       const double* AAP, const double* PQ, int nfit, double tol, double e0, int device,
                       size_t workspace_bytes);
/* Orbital-space sizes of a context (the reference's EC.space 'o', 'v' and the DF
 * basis, src/infos/ecinfos.jl:231-244); any output may be NULL. */
int ecg_dims(const ecg_ctx* ctx, int* nocc, int* nvirt, int* naux);

/* Closed-shell Fock matrix (n x n), orbital energies (n) and HF energy:
 * gen_fock (src/scf/fockfactory.jl:27-30), calc_fock_matrix / calc_HF_energy
 * (src/cc/cctools.jl:35-76).  Any output pointer may be NULL. */
int ecg_fock(ecg_ctx* ctx, double* fock, double* eps, double* ehf);

/* MP2 amplitudes and energy [E, ESS, EOS]: calc_MP2 (src/cc/cc.jl:1011-1024). */
int ecg_mp2(ecg_ctx* ctx, double* T1, double* T2, double* emp2);

/* Closed-shell CCSD (method=ECG_CCSD) or DCSD (ECG_DCSD) residual for given
 * amplitudes T1 (v x o), T2 (v x v x o x o): calc_cc_resid(EC, T1, T2; dc)
 * (src/cc/cc.jl:1343-1489). */
int ecg_cc_residual(ecg_ctx* ctx, int method, int precision, const double* T1, const double* T2,
                    double* R1, double* R2);
/* Same with device pointers (already resident in HBM). */
int ecg_cc_residual_dev(ecg_ctx* ctx, int method, int precision, const double* dT1, const double* dT2,
                        double* dR1, double* dR2);

/* Allocation-free variant of ecg_cc_residual_dev.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  FP64 only
 * (precision must be ECG_FP64; products run on the unpacked FP64 MFMA kernel).
 * Every contraction temporary, ladder integral block and split-K partial sum lives
 * in one static device array of ECG_NOALLOC_SCRATCH_BYTES per GPU, held from the
 * library's first use of that GPU until the process ends; the dressed factors were
 * allocated by ecg_create.  A system whose residual does not fit the array is
 * rejected with status -2 and an ecg_last_error message (at 256 MiB: o = 12,
 * v = 108, naux = 360 fits; o = 20, v = 180 does not).  All calls on one GPU share
 * the array: calls that overlap in time on one GPU (contexts driven from several
 * threads) race; calls one after another are safe, since nothing in the array
 * has to start at zero.  Results equal ecg_cc_residual_dev's in FP64 up to
 * summation order. */
#define ECG_NOALLOC_SCRATCH_BYTES (256u * 1024u * 1024u)
int ecg_cc_residual_dev_noalloc(ecg_ctx* ctx, int method, int precision, const double* dT1, const double* dT2,
                                double* dR1, double* dR2);

/* Device-resident iterations of cc_iterations! (src/cc/cc.jl:2518-2648).
 * state_init: amplitudes from T1/T2 (host), or the MP2 guess when both are NULL;
 * resets DIIS.  state_step: one iteration (residual, Hylleraas energy, update,
 * DIIS) in the given precision, without FP64 correction of a low-precision
 * residual; info = [E_H, ESS, EOS, E_proj, dE, NormR, NormT, converged(0/1)].
 * state_get: copy the current amplitudes to the host. */
int ecg_cc_state_init(ecg_ctx* ctx, const double* T1, const double* T2, const ecg_solver_opts* opts);
int ecg_cc_state_step(ecg_ctx* ctx, int method, int precision, const ecg_solver_opts* opts, double* info);
int ecg_cc_state_get(ecg_ctx* ctx, double* T1, double* T2);

/* calc_cc for the closed-shell path (src/cc/cc.jl:2451-2516): iterate from the
 * MP2 guess until converged.  energies = [E_H, ESS, EOS, E_proj, dE, NormR] of
 * the last iteration; T1/T2 receive the converged amplitudes when not NULL.
 * precision = ECG_FP32 or ECG_BF16X3 (mixed, defect correction): iterations
 * use the residual R_low(T) + delta with low-precision products; delta =
 * R_64 - R_low is refreshed by an FP64 residual at the current T every
 * opts->refine_every-th iteration (when > 0), whenever the corrected NormR
 * drops below 100*thr, and whenever it stalls (less than 20% below its
 * smallest value since the last refresh).  Convergence (the reference's test)
 * is only declared on such an exact iteration, so the result is the FP64
 * solution; ecg_stats reports how many FP64 residuals were used. */
int ecg_cc_solve(ecg_ctx* ctx, int method, int precision, const ecg_solver_opts* opts, double* T1,
                 double* T2, double* energies, int* iterations, int* converged);

/* Many independent systems: ecg_cc_solve for each of nsys contexts (created
 * on the same or different devices), run concurrently -- one host thread and
 * the context's own HIP stream per system, at most max_threads at a time
 * (0: min(nsys, 16)).  The reference drives one system per process
 * (ccdriver, src/cc/ccdriver.jl:37-76); this is its batched form.  Outputs
 * are per system: energies 6*nsys (layout of ecg_cc_solve), iterations and
 * converged nsys each.  status[k] (nsys, may be NULL) receives each system's
 * status; the return value is 0 if all succeeded, else the first failure's
 * status (its message is in ecg_last_error of that context). */
int ecg_cc_solve_batch(int nsys, ecg_ctx* const* ctxs, int method, int precision, const ecg_solver_opts* opts,
                       int max_threads, double* energies, int* iterations, int* converged, int* status);

/* Perturbative triples for the closed-shell path.  lambda = 0:
 * calc_pertT_closed_shell (src/cc/cc_triples.jl:39-148) with amplitudes T1, T2;
 * lambda = 1: calc_LambdapertT_closed_shell (cc_triples.jl:159-288), which also
 * needs the contravariant multipliers U1, U2.  Returns E(T) and E[T]. */
int ecg_pert_triples(ecg_ctx* ctx, int lambda, int precision, const double* T1, const double* T2,
                     const double* U1, const double* U2, double* ET, double* ETb);
/* The same restricted to triplets first .. first+count-1 of the reference's loop
 * order (cc_triples.jl:61-71: k, j<=k, i<=j, i=j=k skipped).  E(T) and E[T] are
 * sums over triplets, so the results of disjoint ranges add up: this is the unit
 * in which (T) is distributed over devices or processes. */
int ecg_pert_triples_range(ecg_ctx* ctx, int lambda, int precision, const double* T1, const double* T2,
                           const double* U1, const double* U2, long first, long count, double* ET,
                           double* ETb);

/* Lambda residual R = U x Jacobian: calc_ccsd_vector_times_Jacobian(EC, U1, U2; dc)
 * (src/cc/cc_lagrange.jl:9-212) at amplitudes T1, T2. */
int ecg_lambda_residual(ecg_ctx* ctx, int method, int precision, const double* T1, const double* T2,
                        const double* U1, const double* U2, double* R1, double* R2);

/* Lambda iterations lm_cc_iterations! (src/cc/cc_lagrange.jl:1155-1223) at
 * converged amplitudes T1, T2; U1/U2 receive the multipliers.  precision =
 * ECG_FP32 or ECG_BF16X3: the defect-corrected iterations of ecg_cc_solve
 * applied to the Lambda residual (low-precision products plus an FP64 correction
 * refreshed every refine_every-th iteration, near convergence and on stalls;
 * convergence only on an exact iteration), so the result is the FP64 solution. */
int ecg_lambda_solve(ecg_ctx* ctx, int method, int precision, const ecg_solver_opts* opts, const double* T1,
                     const double* T2, double* U1, double* U2, int* iterations, int* converged);

/* Gradient of the correlation energy (the backward of the CC energy functional;
 * the reference has no counterpart -- its Lambda equations, lm_cc_iterations!
 * src/cc/cc_lagrange.jl:1155-1223, and calc_1RDM :793-837 are the pieces it
 * builds on).  At converged amplitudes T1, T2 (ecg_cc_solve) and multipliers U1,
 * U2 (ecg_lambda_solve) of the same method: dE_dh[p,q] = dE/dh[p,q] (n x n) and
 * dE_dB[p,q,Q] = dE/dB[p,q,Q] (n x n x naux), every entry of h and B treated as
 * independent; either output may be NULL.  dE/dF at fixed B equals dE/dh.
 * Hermitian integrals (BR = NULL at ecg_create); two-sided factors: ecg_energy_grad_lr. */
int ecg_energy_grad(ecg_ctx* ctx, int method, const double* T1, const double* T2, const double* U1,
                    const double* U2, double* dE_dh, double* dE_dB);
/* The same for two-sided DF factors (a context made with BR, e.g. from a similarity-
 * transformed "ST" FCIDUMP): dE_dBL and dE_dBR (n x n x naux each, either may be NULL)
 * are the derivatives with respect to BL and BR as independent arrays.  Along directions
 * that keep the integrals' particle symmetry (pq|rs) = (rs|pq) -- e.g. BL -> BL + X S^(1/2),
 * BR -> BR + X S^(1/2) sign(D) for the factors BL = L |D|^(1/2), BR = L sign(D) |D|^(1/2) of
 * ecg_fcidump_df -- their sum <dE_dBL, dBL> + <dE_dBR, dBR> is the derivative of the CC
 * energy.  dE/dh = dE/dF at fixed factors, as for ecg_energy_grad.  Contexts with one
 * factor (BR = NULL) are refused (use ecg_energy_grad). */
int ecg_energy_grad_lr(ecg_ctx* ctx, int method, const double* T1, const double* T2, const double* U1,
                       const double* U2, double* dE_dh, double* dE_dBL, double* dE_dBR);

/* One-particle density: calc_1RDM (src/cc/cc_lagrange.jl:793-837); D1, dD1 n x n. */
int ecg_one_rdm(ecg_ctx* ctx, const double* T1, const double* T2, const double* U1, const double* U2,
                double* D1, double* dD1);

/* Instrumentation (no counterpart in the reference, whose timings come from
 * print_time, src/tools/utils.jl:51-57): enable = 1 records HIP events around
 * every MFMA GEMM launch on the library stream and zeroes the counters.
 * stats = [last residual ms, last ladder ms, timed GEMM ms, timed GEMM launches,
 *          issued GEMM flops, FP64 iterations of the last ecg_cc_solve,
 *          algorithmic GEMM bytes (operands once, C written once or read and written),
 *          triplets per block of the last (T) call]
 * (8 doubles).  The BF16x3 (T) evaluates blocks of consecutive triplets (same j, k)
 * per launch; the block size (1, 2, 4 or 8) is measured on the context's first (T)
 * call of at least 64 triplets and kept; results do not depend on it. */
int ecg_profile(ecg_ctx* ctx, int enable);
int ecg_stats(ecg_ctx* ctx, double* stats);

/* ---- Open-shell (unrestricted) coupled cluster: UCCSD (ECG_UCCSD), UCCD (ECG_UCCD, no
 * singles) and UDCSD (ECG_UDCSD).  Replaces the unrestricted branch of calc_cc
 * (src/cc/cc.jl:2451-2516) with its residual calc_cc_resid(EC, T1a, T1b, T2a, T2b, T2ab; dc)
 * (src/cc/cc.jl:1496-1934),
 * calc_UMP2 (cc.jl:1034-1060) and the open-shell Fock / HF energy (calc_fock_matrix,
 * calc_HF_energy, src/cc/cctools.jl:35-76).  The alpha spin orbitals are the lowest nocca
 * orbitals, the beta ones the lowest noccb (the default spaces o / O of setup_space_fd!,
 * src/infos/ecinfos.jl:171-192).  Amplitudes and residuals in the reference's spin blocks,
 * column-major: T1a [va, na], T1b [vb, nb], T2a [va, va, na, na], T2b [vb, vb, nb, nb] (both
 * antisymmetric), T2ab [va, vb, na, nb], with na = nocca, nb = noccb, va = norb - na,
 * vb = norb - nb.  Hermitian integrals, (pq|rs) = sum_Q B[p,q,Q] B[r,s,Q].  Status codes and
 * errors as for the closed-shell context (message: ecg_ulast_error). */
enum { ECG_UCCSD = 0, ECG_UCCD = 1, ECG_UDCSD = 2 };
typedef struct ecg_uctx ecg_uctx;
int ecg_ucreate(ecg_uctx** ctx, int nocca, int noccb, int norb, int naux, const double* h, const double* B,
                double e0, int device, size_t workspace_bytes);
/* From an FCIDUMP file (ST = 0): nocca = (NELEC + ms2) / 2, noccb = (NELEC - ms2) / 2, with ms2
 * the header's MS2 when ms2 < 0 (the reference's `@opt wf ms2=...` overrides it). */
int ecg_ucreate_from_fcidump(ecg_uctx** ctx, const char* path, int ms2, double rtol, int device,
                             size_t workspace_bytes);
int ecg_udestroy(ecg_uctx* ctx);
/* Message of the last failed call on ctx; with ctx = NULL, of the last failed create call. */
const char* ecg_ulast_error(const ecg_uctx* ctx);
int ecg_udims(const ecg_uctx* ctx, int* norb, int* nocca, int* noccb, int* naux);
/* alpha and beta Fock matrices (gen_fock(EC, :alpha / :beta), n x n each) and the HF energy
 * 0.5 (sum eps_o^a + sum eps_o^b + sum h_oo^a + sum h_oo^b) + E0; any output may be NULL. */
int ecg_ufock(ecg_uctx* ctx, double* focka, double* fockb, double* ehf);
/* UMP2 amplitudes and energy (doubles plus the Fock part of the singles energy). */
int ecg_ump2(ecg_uctx* ctx, double* T1a, double* T1b, double* T2a, double* T2b, double* T2ab, double* emp2);
/* The residual of the given amplitudes (UCCD: the singles are taken as zero and R1 is zero). */
int ecg_ucc_residual(ecg_uctx* ctx, int method, const double* T1a, const double* T1b, const double* T2a,
                     const double* T2b, const double* T2ab, double* R1a, double* R1b, double* R2a, double* R2b,
                     double* R2ab);
/* cc_iterations! (src/cc/cc.jl:2518-2648) from the UMP2 guess with the reference's
 * unrestricted DIIS dots; energies = [E_H, E_proj, dE, NormR] of the last iteration, the
 * converged amplitudes in T1a .. T2ab when given (any may be NULL).  Uses opts->thr, conven,
 * maxit, shifts, shiftp, maxdiis, resthr (FP64 throughout). */
int ecg_ucc_solve(ecg_uctx* ctx, int method, const ecg_solver_opts* opts, double* T1a, double* T1b, double* T2a,
                  double* T2b, double* T2ab, double* energies, int* iterations, int* converged);
/* (T) of UCCSD: calc_pertT_unrestricted (src/cc/cc_triples.jl:297-320, same-spin and mixed-spin
 * parts :357-640) at the given amplitudes, or at those of the last ecg_ucc_solve when all five
 * are NULL; E(T) and E[T].  Diagonal-Fock denominators and the f_ov term of the reference.  Holds
 * five (nvirta + nvirtb)^3 (nocca + noccb)^3 arrays in the workspace. */
int ecg_upert_triples(ecg_uctx* ctx, const double* T1a, const double* T1b, const double* T2a, const double* T2b,
                      const double* T2ab, double* ET, double* ETb);

#ifdef __cplusplus
}
#endif
#endif
