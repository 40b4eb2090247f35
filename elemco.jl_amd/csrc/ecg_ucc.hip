// Open-shell (unrestricted) CCSD / CCD / DCSD on the device.
//
// The reference's UCCSD / UDCSD residual (calc_cc_resid(EC, T1a, T1b, T2a, T2b, T2ab; dc),
// src/cc/cc.jl:1496-1934) works on alpha/beta spin blocks with dressed four-index integral
// files.  Here the
// same equations run on spin orbitals ([occupied alpha, occupied beta, virtual alpha, virtual
// beta]) in the intermediate form of Stanton and Gauss (J. Chem. Phys. 94, 4334 (1991)) with
// the full Fock matrix (non-canonical ROHF orbitals), as a sequence of contractions on the
// MFMA GEMM engine.  Integrals come from spin-orbital DF factors B[p,q,Q] (zero between
// different spins); five antisymmetrised blocks <oo||vv>, <ov||vv>, <oo||ov>, <ov||vo>,
// <oo||oo> are stored, the other blocks the equations use are their permutations, and the
// particle-particle ladder sum_ef <ab|ef> tau_ij^ef is assembled from the factors block by
// block (df_ladder_half), so no v^4 array exists.  The alpha-alpha, beta-beta and alpha-beta
// blocks of the residual are the reference's R2a, R2b, R2ab.
//
// UDCSD: the reference's dressed equations differ from UCCSD only in terms quadratic in T2
// with the undressed <kl|cd> (cc.jl:1724-1808): the hole-hole term is dropped, the quadratic
// parts of x_ij / x_ab are halved and the ring terms keep only <kl|cd> with (k,c) in one
// amplitude and (l,d) in the other.  Here those differences are taken out of the UCCSD terms
// (with T2 alone, the T1 parts of tau / tau~ untouched): W_mbej gets 1/2 sum_nf <mn|fe>
// T2[b,f,j,n] back, the doubles' copies of F_be and F_mj a quarter of their T2 parts, and
// 1/4 T2[a,b,m,n] sum_ef <mn||ef> T2[e,f,i,j] is subtracted.
#include "ecg_ucc.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ecg {

namespace {

Tensor t2(double* p, long a, long b) { return Tensor(p, {a, b}); }
Tensor t3(double* p, long a, long b, long c) { return Tensor(p, {a, b, c}); }
Tensor t4(double* p, long a, long b, long c, long d) { return Tensor(p, {a, b, c, d}); }
double* cp(const double* p) { return const_cast<double*>(p); }

// w[Q] = sum_m X[m,m,Q] over an O x O x Q block
__global__ void utrace_kernel(int O, int nq, const double* __restrict__ X, double* w) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  double s = 0.0;
  const double* x = X + (long)q * O * O;
  for (int m = 0; m < O; ++m) s += x[m + (long)m * O];
  w[q] = s;
}

// T2[a,b,i,j] (+)= -R2/(e_a + e_b - e_i - e_j + shift), T1[a,i] (+)= -R1/(e_a - e_i + shift)
__global__ void uupdate_doubles_kernel(int V, int O, const double* __restrict__ eo, const double* __restrict__ ev,
                                       double shift, const double* __restrict__ R, double* T, int assign) {
  const long n = (long)V * V * O * O;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x) {
    long r = idx;
    const int a = (int)(r % V); r /= V;
    const int b = (int)(r % V); r /= V;
    const int i = (int)(r % O);
    const int j = (int)(r / O);
    const double d = -R[idx] / (ev[a] + ev[b] - eo[i] - eo[j] + shift);
    T[idx] = assign ? d : T[idx] + d;
  }
}

__global__ void uupdate_singles_kernel(int V, int O, const double* __restrict__ eo, const double* __restrict__ ev,
                                       double shift, const double* __restrict__ R, double* T, int assign) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= V * O) return;
  const int a = idx % V, i = idx / V;
  const double d = -R[idx] / (ev[a] - eo[i] + shift);
  T[idx] = assign ? d : T[idx] + d;
}

unsigned ugrid(long n) { return (unsigned)std::min<long>((n + 255) / 256, 16384); }

}  // namespace

void UContext::init(int nocca, int noccb, int norb, int naux, const double* h_host, const double* B_host,
                    double e0_in, size_t workspace_bytes) {
  ECG_CHECK(norb > 0 && naux > 0 && nocca >= 0 && noccb >= 0 && nocca <= norb && noccb <= norb,
            "open-shell sizes out of range");
  ECG_CHECK(nocca + noccb > 0 && nocca + noccb < 2 * norb, "no occupied or no virtual spin orbitals");
  n = norb; na = nocca; nb = noccb; nq = naux; e0 = e0_in;
  O = na + nb;
  V = (n - na) + (n - nb);
  N = O + V;
  sp.clear(); sg.clear();
  for (int i = 0; i < na; ++i) { sp.push_back(i); sg.push_back(0); }
  for (int i = 0; i < nb; ++i) { sp.push_back(i); sg.push_back(1); }
  for (int a = na; a < n; ++a) { sp.push_back(a); sg.push_back(0); }
  for (int a = nb; a < n; ++a) { sp.push_back(a); sg.push_back(1); }
  ECG_HIP(hipSetDevice(device));
  ECG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  ops.stream = stream;
  ops.gemm.stream = stream;
  // spin-orbital factor blocks: B[p,q,Q] for equal spins, zero otherwise
  auto block = [&](int r0, int rows, int c0, int cols, DevBuf& d) {
    std::vector<double> b((size_t)rows * cols * nq, 0.0);
    for (int Q = 0; Q < nq; ++Q)
      for (int c = 0; c < cols; ++c)
        for (int r = 0; r < rows; ++r) {
          const int p = r0 + r, q = c0 + c;
          if (sg[p] != sg[q]) continue;
          b[(size_t)r + (size_t)rows * (c + (size_t)cols * Q)] = B_host[sp[p] + (size_t)n * (sp[q] + (size_t)n * Q)];
        }
    d.alloc(sizeof(double) * b.size());
    ECG_HIP(hipMemcpyAsync(d.p, b.data(), sizeof(double) * b.size(), hipMemcpyHostToDevice, stream));
    ECG_HIP(hipStreamSynchronize(stream));
  };
  block(0, O, 0, O, Boo);
  block(0, O, O, V, Bov);
  block(O, V, 0, O, Bvo);
  block(O, V, O, V, Bvv);
  const long o = O, v = V;
  size_t want = sizeof(double) * (size_t)(12 * n2() + 4 * v * o * o * o + o * o * o * o + 16 * v * v * v) +
                (size_t)256 * 1024 * 1024;
  // room for the (T) arrays (five V^3 O^3) when they are at most 32 GB
  const size_t t_need = sizeof(double) * (size_t)(5 * v * v * v * o * o * o) + (size_t)1024 * 1024;
  if (t_need <= ((size_t)32 << 30)) want = std::max(want, t_need);
  if (workspace_bytes) want = workspace_bytes;
  ops.arena.buf.alloc(want);
  ops.arena.reset();
  // Fock matrix f = h + sum_m <pm||qm> per block: h + B tr - sum_m B[p,m] B[m,q]
  std::vector<double> hso((size_t)N * N, 0.0);
  for (int q = 0; q < N; ++q)
    for (int p = 0; p < N; ++p)
      if (sg[p] == sg[q]) hso[(size_t)p + (size_t)N * q] = h_host[sp[p] + (size_t)n * sp[q]];
  double* tr = ops.arena.take(nq);
  hipLaunchKernelGGL(utrace_kernel, dim3((nq + 255) / 256), dim3(256), 0, stream, O, nq, Boo.d(), tr);
  const int rows[4] = {O, O, V, V}, cols[4] = {O, V, O, V}, r0[4] = {0, 0, O, O}, c0[4] = {0, O, 0, O};
  DevBuf* fb[4] = {&foo, &fov, &fvo, &fvv};
  const DevBuf* Bq[4] = {&Boo, &Bov, &Bvo, &Bvv};
  const DevBuf* Bxm[4] = {&Boo, &Boo, &Bvo, &Bvo};
  const DevBuf* Bmy[4] = {&Boo, &Bov, &Boo, &Bov};
  std::vector<double> fso((size_t)N * N, 0.0), tmp;
  for (int b = 0; b < 4; ++b) {
    tmp.assign((size_t)rows[b] * cols[b], 0.0);
    for (int c = 0; c < cols[b]; ++c)
      for (int r = 0; r < rows[b]; ++r) tmp[(size_t)r + (size_t)rows[b] * c] = hso[(size_t)(r0[b] + r) + (size_t)N * (c0[b] + c)];
    fb[b]->alloc(sizeof(double) * tmp.size());
    ECG_HIP(hipMemcpyAsync(fb[b]->p, tmp.data(), sizeof(double) * tmp.size(), hipMemcpyHostToDevice, stream));
    Tensor f = t2(fb[b]->d(), rows[b], cols[b]);
    ops.contract_p(Prec::F64, 1.0, t3(Bq[b]->d(), rows[b], cols[b], nq), "pqQ", Tensor(tr, {(long)nq}), "Q", 1.0, f,
                   "pq");
    ops.contract_p(Prec::F64, -1.0, t3(Bxm[b]->d(), rows[b], O, nq), "pmQ", t3(Bmy[b]->d(), O, cols[b], nq), "mqQ",
                   1.0, f, "pq");
    ECG_HIP(hipMemcpyAsync(tmp.data(), fb[b]->p, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, stream));
    ECG_HIP(hipStreamSynchronize(stream));
    for (int c = 0; c < cols[b]; ++c)
      for (int r = 0; r < rows[b]; ++r) fso[(size_t)(r0[b] + r) + (size_t)N * (c0[b] + c)] = tmp[(size_t)r + (size_t)rows[b] * c];
  }
  fock.alloc(sizeof(double) * fso.size());
  ECG_HIP(hipMemcpyAsync(fock.p, fso.data(), sizeof(double) * fso.size(), hipMemcpyHostToDevice, stream));
  std::vector<double> e(N);
  for (int p = 0; p < N; ++p) e[p] = fso[(size_t)p * (N + 1)];
  eps.alloc(sizeof(double) * N);
  ECG_HIP(hipMemcpyAsync(eps.p, e.data(), sizeof(double) * N, hipMemcpyHostToDevice, stream));
  // calc_HF_energy, open-shell branch (src/cc/cctools.jl:71-74)
  double se = 0.0, sh = 0.0;
  for (int i = 0; i < O; ++i) {
    se += e[i];
    sh += hso[(size_t)i * (N + 1)];
  }
  ehf = 0.5 * (se + sh) + e0;
  // antisymmetrised blocks  <pq||rs> = (pr|qs) - (ps|qr)
  auto anti = [&](DevBuf& dst, long d0, long d1, long d2, long d3, const DevBuf& Bpr, long p0, long p1, const DevBuf& Bqs,
                  long q0, long q1, const DevBuf& Bps, long s0, long s1, const DevBuf& Bqr, long t0, long t1) {
    dst.alloc(sizeof(double) * d0 * d1 * d2 * d3);
    Tensor A = t4(dst.d(), d0, d1, d2, d3);
    ops.contract_p(Prec::F64, 1.0, t3(Bpr.d(), p0, p1, nq), "prQ", t3(Bqs.d(), q0, q1, nq), "qsQ", 0.0, A, "pqrs");
    ops.contract_p(Prec::F64, -1.0, t3(Bps.d(), s0, s1, nq), "psQ", t3(Bqr.d(), t0, t1, nq), "qrQ", 1.0, A, "pqrs");
  };
  anti(oovv, o, o, v, v, Bov, o, v, Bov, o, v, Bov, o, v, Bov, o, v);
  anti(ovvv, o, v, v, v, Bov, o, v, Bvv, v, v, Bov, o, v, Bvv, v, v);
  anti(ooov, o, o, o, v, Boo, o, o, Bov, o, v, Bov, o, v, Boo, o, o);
  anti(ovvo, o, v, v, o, Bov, o, v, Bvo, v, o, Boo, o, o, Bvv, v, v);
  anti(oooo, o, o, o, o, Boo, o, o, Boo, o, o, Boo, o, o, Boo, o, o);
  ECG_HIP(hipStreamSynchronize(stream));
  ops.arena.reset();
}

UContext::~UContext() {
  if (stream) (void)hipStreamDestroy(stream);
}

// R(T) on spin orbitals: Stanton-Gauss intermediates with the full Fock matrix.
void UContext::residual(int method, const double* T1d, const double* T2d, double* R1d, double* R2d) {
  const bool singles = has_singles(method), dc = method == 2;
  const long o = O, v = V;
  const Prec F = Prec::F64;
  ops.arena.reset();
  Tensor T1 = t2(cp(T1d), v, o), T2 = t4(cp(T2d), v, v, o, o);
  Tensor R1 = t2(R1d, v, o), R2 = t4(R2d, v, v, o, o);
  Tensor Fov = t2(fov.d(), o, v);
  Tensor Aoovv = t4(oovv.d(), o, o, v, v), Aovvv = t4(ovvv.d(), o, v, v, v), Aooov = t4(ooov.d(), o, o, o, v),
         Aovvo = t4(ovvo.d(), o, v, v, o);
  const int p_ba[4] = {1, 0, 2, 3}, p_ij[4] = {0, 1, 3, 2}, p_sym[4] = {1, 0, 3, 2};
  auto take = [&](long count) { return ops.arena.take(count); };
  // tt = T1 T1, tau~ = T2 + (tt - tt_ba)/2, tau = T2 + tt - tt_ba
  double* tt = take(n2());
  double* taut = take(n2());
  double* tau = take(n2());
  Tensor Ttt = t4(tt, v, v, o, o), Ttaut = t4(taut, v, v, o, o), Ttau = t4(tau, v, v, o, o);
  ops.contract_p(F, 1.0, T1, "ai", T1, "bj", 0.0, Ttt, "abij");
  ops.copy(n2(), T2d, taut);
  ops.axpby(n2(), 0.5, tt, 1.0, taut);
  ops.permute(Ttt, p_ba, Ttaut, -0.5, 1.0);
  ops.copy(n2(), T2d, tau);
  ops.axpby(n2(), 1.0, tt, 1.0, tau);
  ops.permute(Ttt, p_ba, Ttau, -1.0, 1.0);
  // F_ae, F_mi, F_me
  double* fae = take(v * v);
  double* fmi = take(o * o);
  double* fme = take(o * v);
  Tensor Fae = t2(fae, v, v), Fmi = t2(fmi, o, o), Fme = t2(fme, o, v);
  ops.copy(v * v, fvv.d(), fae);
  ops.contract_p(F, -0.5, Fov, "me", T1, "am", 1.0, Fae, "ae");
  ops.contract_p(F, 1.0, T1, "fm", Aovvv, "mafe", 1.0, Fae, "ae");
  ops.contract_p(F, -0.5, Ttaut, "afmn", Aoovv, "mnef", 1.0, Fae, "ae");
  ops.copy(o * o, foo.d(), fmi);
  ops.contract_p(F, 0.5, T1, "ei", Fov, "me", 1.0, Fmi, "mi");
  ops.contract_p(F, 1.0, T1, "en", Aooov, "mnie", 1.0, Fmi, "mi");
  ops.contract_p(F, 0.5, Ttaut, "efin", Aoovv, "mnef", 1.0, Fmi, "mi");
  ops.copy(o * v, fov.d(), fme);
  ops.contract_p(F, 1.0, T1, "fn", Aoovv, "mnef", 1.0, Fme, "me");
  // W_mnij = <mn||ij> + P(ij) T1[e,j] <mn||ie> + 1/4 tau[e,f,i,j] <mn||ef>
  double* wmn = take(o * o * o * o);
  Tensor W = t4(wmn, o, o, o, o);
  ops.copy(o * o * o * o, oooo.d(), wmn);
  ops.contract_p(F, 1.0, T1, "ej", Aooov, "mnie", 1.0, W, "mnij");
  ops.contract_p(F, -1.0, T1, "ei", Aooov, "mnje", 1.0, W, "mnij");
  ops.contract_p(F, 0.25, Ttau, "efij", Aoovv, "mnef", 1.0, W, "mnij");
  // W_mbej = <mb||ej> + T1[f,j] <mb||ef> + T1[b,n] <mn||je> - (T2/2 + tt)[f,b,j,n] <mn||ef>
  double* wmb = take(o * v * v * o);
  Tensor W2 = t4(wmb, o, v, v, o);
  double* g = take(n2());
  Tensor G = t4(g, v, v, o, o);
  ops.copy(o * v * v * o, ovvo.d(), wmb);
  ops.contract_p(F, 1.0, T1, "fj", Aovvv, "mbef", 1.0, W2, "mbej");
  ops.contract_p(F, 1.0, T1, "bn", Aooov, "mnje", 1.0, W2, "mbej");
  ops.axpby(n2(), 0.5, T2d, 0.0, g);
  ops.axpby(n2(), 1.0, tt, 1.0, g);
  ops.contract_p(F, -1.0, G, "fbjn", Aoovv, "mnef", 1.0, W2, "mbej");
  if (dc) {
    // exchange part of the T2 ring term out: + 1/2 <mn|fe> T2[b,f,j,n],  <mn|fe> = (mf|ne)
    size_t mk = ops.arena.mark();
    double* kx = take(n2());
    Tensor Kx = t4(kx, o, o, v, v);
    Tensor Bo = t3(Bov.d(), o, v, nq);
    ops.contract_p(F, 1.0, Bo, "mfQ", Bo, "neQ", 0.0, Kx, "mnef");
    ops.contract_p(F, 0.5, Kx, "mnef", T2, "bfjn", 1.0, W2, "mbej");
    ops.arena.release_to(mk);
  }
  // singles
  if (singles) {
    ops.copy(v * o, fvo.d(), R1d);
    ops.contract_p(F, 1.0, T1, "ei", Fae, "ae", 1.0, R1, "ai");
    ops.contract_p(F, -1.0, T1, "am", Fmi, "mi", 1.0, R1, "ai");
    ops.contract_p(F, 1.0, T2, "aeim", Fme, "me", 1.0, R1, "ai");
    ops.contract_p(F, 1.0, T1, "fn", Aovvo, "nafi", 1.0, R1, "ai");     // - T1[f,n] <an||fi>
    ops.contract_p(F, -0.5, T2, "efim", Aovvv, "maef", 1.0, R1, "ai");
    ops.contract_p(F, 0.5, T2, "aemn", Aooov, "nmie", 1.0, R1, "ai");   // - 1/2 T2[a,e,m,n] <nm||ei>
  } else {
    ops.fill(v * o, 0.0, R1d);
  }
  // doubles: <ab||ij> = <ij||ab>
  const int p_vvoo[4] = {2, 3, 0, 1};
  ops.permute(Aoovv, p_vvoo, R2, 1.0, 0.0);
  {
    // P(ab) T2[a,e,i,j] (F_be - T1[b,m] F_me / 2)
    double* hb = take(v * v);
    Tensor H = t2(hb, v, v);
    ops.copy(v * v, fae, hb);
    ops.contract_p(F, -0.5, T1, "bm", Fme, "me", 1.0, H, "be");
    if (dc) ops.contract_p(F, 0.25, T2, "bfmn", Aoovv, "mnef", 1.0, H, "be");   // x_ab halved
    ops.contract_p(F, 1.0, T2, "aeij", H, "be", 1.0, R2, "abij");
    ops.contract_p(F, -1.0, T2, "beij", H, "ae", 1.0, R2, "abij");
    // - P(ij) T2[a,b,i,m] (F_mj + F_me T1[e,j] / 2)
    double* g2 = take(o * o);
    Tensor G2 = t2(g2, o, o);
    ops.copy(o * o, fmi, g2);
    ops.contract_p(F, 0.5, Fme, "me", T1, "ej", 1.0, G2, "mj");
    if (dc) ops.contract_p(F, -0.25, T2, "efjn", Aoovv, "mnef", 1.0, G2, "mj");  // x_ij halved
    ops.contract_p(F, -1.0, T2, "abim", G2, "mj", 1.0, R2, "abij");
    ops.contract_p(F, 1.0, T2, "abjm", G2, "mi", 1.0, R2, "abij");
  }
  // 1/2 tau[a,b,m,n] W_mnij
  ops.contract_p(F, 0.5, Ttau, "abmn", W, "mnij", 1.0, R2, "abij");
  {
    // 1/2 tau[e,f,i,j] W_abef without forming W_abef:
    //   sum_ef <ab|ef> tau[e,f,i,j]                 (DF ladder, half + symmetrisation)
    //   - P(ab) T1[b,m] Z[a,m,i,j],  Z = -1/2 <ma||ef> tau[e,f,i,j]
    //   + 1/4 tau[a,b,m,n] X[m,n,i,j],  X = 1/2 <mn||ef> tau[e,f,i,j]
    size_t mk = ops.arena.mark();
    double* r2r = take(n2());
    double* taux = take(n2());
    Tensor R2r = t4(r2r, v, v, o, o), Taux = t4(taux, o, o, v, v);
    ops.fill(n2(), 0.0, r2r);
    const int p_x[4] = {2, 3, 0, 1};
    ops.permute(Ttau, p_x, Taux, 1.0, 0.0);
    df_ladder_half(ops, v, o, nq, F, Bvv.d(), Bvv.d(), taux, r2r);
    ops.axpby(n2(), 1.0, r2r, 1.0, R2d);
    ops.permute(R2r, p_sym, R2, 1.0, 1.0);
    ops.arena.release_to(mk);
    double* z = take(v * o * o * o);
    Tensor Z = t4(z, v, o, o, o);
    ops.contract_p(F, -0.5, Aovvv, "maef", Ttau, "efij", 0.0, Z, "amij");
    ops.contract_p(F, -1.0, T1, "bm", Z, "amij", 1.0, R2, "abij");
    ops.contract_p(F, 1.0, T1, "am", Z, "bmij", 1.0, R2, "abij");
    double* x = take(o * o * o * o);
    Tensor X = t4(x, o, o, o, o);
    ops.contract_p(F, 0.5, Aoovv, "mnef", Ttau, "efij", 0.0, X, "mnij");
    ops.contract_p(F, 0.25, Ttau, "abmn", X, "mnij", 1.0, R2, "abij");
    if (dc) {
      // hole-hole term out: - 1/4 T2[a,b,m,n] sum_ef <mn||ef> T2[e,f,i,j]
      ops.contract_p(F, 1.0, Aoovv, "mnef", T2, "efij", 0.0, X, "mnij");
      ops.contract_p(F, -0.25, T2, "abmn", X, "mnij", 1.0, R2, "abij");
    }
  }
  {
    // P(ij) P(ab) [T2[a,e,i,m] W_mbej - T1[a,m] T1[e,i] <mb||ej>]
    double* y2 = take(o * v * o * o);
    Tensor Y2 = t4(y2, o, v, o, o);
    ops.contract_p(F, 1.0, T1, "ei", Aovvo, "mbej", 0.0, Y2, "mbij");
    double* x4 = take(n2());
    Tensor X4 = t4(x4, v, v, o, o);
    ops.contract_p(F, 1.0, T2, "aeim", W2, "mbej", 0.0, X4, "abij");
    ops.contract_p(F, -1.0, T1, "am", Y2, "mbij", 1.0, X4, "abij");
    ops.axpby(n2(), 1.0, x4, 1.0, R2d);
    ops.permute(X4, p_ba, R2, -1.0, 1.0);
    ops.permute(X4, p_ij, R2, -1.0, 1.0);
    ops.permute(X4, p_sym, R2, 1.0, 1.0);
  }
  // P(ij) T1[e,i] <ab||ej>  with <ab||ej> = -<je||ab>;  - P(ab) T1[a,m] <mb||ij>  with <mb||ij> = <ij||mb>
  ops.contract_p(F, -1.0, T1, "ei", Aovvv, "jeab", 1.0, R2, "abij");
  ops.contract_p(F, 1.0, T1, "ej", Aovvv, "ieab", 1.0, R2, "abij");
  ops.contract_p(F, -1.0, T1, "am", Aooov, "ijmb", 1.0, R2, "abij");
  ops.contract_p(F, 1.0, T1, "bm", Aooov, "ijma", 1.0, R2, "abij");
  ops.arena.reset();
}

// E = 1/4 T2.<ij||ab> + T1.f_ov + 1/2 T1 T1 <ij||ab>  (calc_doubles_energy + calc_singles_energy,
// cc.jl:123-154,176-193, summed over spin blocks)
double UContext::proj_energy(const double* T1d, const double* T2d) {
  const long o = O, v = V;
  ops.arena.reset();
  double* vv = ops.arena.take(n2());
  double* tt = ops.arena.take(n2());
  double* fv = ops.arena.take(v * o);
  Tensor A = t4(oovv.d(), o, o, v, v), B = t4(vv, v, v, o, o), Ttt = t4(tt, v, v, o, o);
  const int p_vvoo[4] = {2, 3, 0, 1};
  ops.permute(A, p_vvoo, B, 1.0, 0.0);
  Tensor T1 = t2(cp(T1d), v, o);
  ops.contract_p(Prec::F64, 1.0, T1, "ai", T1, "bj", 0.0, Ttt, "abij");
  const int p_t[2] = {1, 0};
  Tensor fo = t2(fov.d(), o, v), fvt = t2(fv, v, o);
  ops.permute(fo, p_t, fvt, 1.0, 0.0);
  const double e = 0.25 * dot(T2d, vv, n2()) + dot(T1d, fv, n1()) + 0.5 * dot(tt, vv, n2());
  ops.arena.reset();
  return e;
}

// calc_UMP2 (cc.jl:1034-1060): T2 = -<ab||ij>/D, T1 = -f_vo/D (no shifts); emp2 = [E] with E the
// doubles energy plus the Fock part of the singles energy
void UContext::mp2(double* T1d, double* T2d, double* emp2) {
  const long o = O, v = V;
  ops.arena.reset();
  double* vv = ops.arena.take(n2());
  double* fv = ops.arena.take(v * o);
  Tensor A = t4(oovv.d(), o, o, v, v), B = t4(vv, v, v, o, o);
  const int p_vvoo[4] = {2, 3, 0, 1}, p_t[2] = {1, 0};
  ops.permute(A, p_vvoo, B, 1.0, 0.0);
  Tensor fo = t2(fov.d(), o, v), fvt = t2(fv, v, o);
  ops.permute(fo, p_t, fvt, 1.0, 0.0);
  const double* eo = eps.d();
  const double* ev = eps.d() + O;
  hipLaunchKernelGGL(uupdate_doubles_kernel, dim3(ugrid(n2())), dim3(256), 0, stream, V, O, eo, ev, 0.0, vv, T2d, 1);
  hipLaunchKernelGGL(uupdate_singles_kernel, dim3((unsigned)((n1() + 255) / 256)), dim3(256), 0, stream, V, O, eo, ev,
                     0.0, fvo.d(), T1d, 1);
  ECG_HIP(hipGetLastError());
  emp2[0] = 0.25 * dot(T2d, vv, n2()) + dot(T1d, fv, n1());
  ops.arena.reset();
}

// cc_iterations! (cc.jl:2518-2648) for UCCSD / UCCD / UDCSD from the UMP2 guess: residual, Hylleraas
// energy, update with level shifts, DIIS with the reference's unrestricted dots (singles 1,
// doubles 1/4 per spin-orbital element: calc_u_singles_dot, calc_samespin_doubles_dot,
// calc_ab_doubles_dot, cctools.jl:453-540), convergence on the squared residual norm and
// the energy change.  energies = [E_H, E_proj, dE, NormR].
int UContext::solve(int method, const SolverOpts& o, double* energies, int* converged) {
  const bool singles = has_singles(method);
  const long m1 = n1(), m2 = n2();
  T1.ensure(sizeof(double) * m1);
  T2.ensure(sizeof(double) * m2);
  R1.ensure(sizeof(double) * m1);
  R2.ensure(sizeof(double) * m2);
  double emp2[1];
  mp2(T1.d(), T2.d(), emp2);
  if (!singles) ops.fill(m1, 0.0, T1.d());
  const int md = o.maxdiis;
  ECG_CHECK(md >= 1 && md <= 64, "maxdiis out of range");
  diis.maxdiis = md; diis.resthr = o.resthr; diis.next = 0; diis.ndim = 0;
  diis.bmat.assign((size_t)(md + 1) * (md + 1), 0.0);
  if ((int)diis.amps.size() != md) {
    diis.amps.clear();
    diis.res.clear();
    for (int k = 0; k < md; ++k) {
      diis.amps.emplace_back(sizeof(double) * (m1 + m2));
      diis.res.emplace_back(sizeof(double) * (m1 + m2));
    }
  }
  const double thren = std::sqrt(o.thr) * o.conven;
  const double* eo = eps.d();
  const double* ev = eps.d() + O;
  double Eh = 0.0, En = 0.0, NormR = 0.0;
  int conv = 0, it = 0;
  fp64_iterations = 0;
  for (it = 1; it <= o.maxit; ++it) {
    residual(method, T1.d(), T2.d(), R1.d(), R2.d());
    fp64_iterations++;
    // calc_hylleraas (cc.jl:286-308): projected energy + 1/4 T2.R2 + R1.T1
    Eh = proj_energy(T1.d(), T2.d()) + 0.25 * dot(T2.d(), R2.d(), m2) + dot(R1.d(), T1.d(), m1);
    NormR = dot(R1.d(), R1.d(), m1) + 0.25 * dot(R2.d(), R2.d(), m2);
    hipLaunchKernelGGL(uupdate_doubles_kernel, dim3(ugrid(m2)), dim3(256), 0, stream, V, O, eo, ev, o.shiftp, R2.d(),
                       T2.d(), 0);
    if (singles)
      hipLaunchKernelGGL(uupdate_singles_kernel, dim3((unsigned)((m1 + 255) / 256)), dim3(256), 0, stream, V, O, eo,
                         ev, o.shifts, R1.d(), T1.d(), 0);
    ECG_HIP(hipGetLastError());
    // perform! (diis.jl:246-305)
    Diis& D = diis;
    if (D.ndim < md) D.ndim++;
    const int ithis = D.next, nd = D.ndim;
    ops.copy(m1, T1.d(), D.amps[ithis].d());
    ops.copy(m2, T2.d(), D.amps[ithis].d() + m1);
    ops.copy(m1, R1.d(), D.res[ithis].d());
    ops.copy(m2, R2.d(), D.res[ithis].d() + m1);
    auto dotv = [&](const double* x, const double* y) { return dot(x, y, m1) + 0.25 * dot(x + m1, y + m1, m2); };
    auto B = [&](int i, int j) -> double& { return D.bmat[(size_t)i * (md + 1) + j]; };
    const double* rt = D.res[ithis].d();
    const double thisdot = dotv(rt, rt);
    for (int i = 0; i < nd; ++i) {
      if (i != ithis) {
        const double d = dotv(rt, D.res[i].d());
        B(i, ithis) = d;
        B(ithis, i) = d;
      } else {
        B(ithis, ithis) = thisdot;
      }
      B(i, nd) = -1.0;
      B(nd, i) = -1.0;
    }
    B(nd, nd) = 0.0;
    const int m = nd + 1;
    std::vector<double> Bm((size_t)m * m), Vv, w;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) Bm[(size_t)i * m + j] = B(i, j);
    jacobi_eig(m, Bm, Vv, w);
    double smax = 0.0;
    for (double x : w) smax = std::max(smax, std::fabs(x));
    std::vector<double> coeff(m, 0.0);
    for (int k = 0; k < m; ++k) {
      if (std::fabs(w[k]) < 2.220446049250313e-16 * smax || w[k] == 0.0) continue;
      const double proj = -Vv[(size_t)nd * m + k];
      for (int i = 0; i < m; ++i) coeff[i] += Vv[(size_t)i * m + k] * proj / w[k];
    }
    bool extrap = true;
    if (nd == 1 && D.next == 0 && thisdot > D.resthr) {
      D.ndim = 0;
      extrap = false;
    } else if (nd < md) {
      D.next = nd;
    } else {
      double cmin = 1e10;
      for (int i = 0; i < nd; ++i)
        if (std::fabs(coeff[i]) < cmin && (i != ithis || nd == 1)) {
          cmin = std::fabs(coeff[i]);
          D.next = i;
        }
    }
    if (extrap) {
      ops.axpby(m1, coeff[0], D.amps[0].d(), 0.0, T1.d());
      ops.axpby(m2, coeff[0], D.amps[0].d() + m1, 0.0, T2.d());
      for (int i = 1; i < nd; ++i) {
        ops.axpby(m1, coeff[i], D.amps[i].d(), 1.0, T1.d());
        ops.axpby(m2, coeff[i], D.amps[i].d() + m1, 1.0, T2.d());
      }
    }
    En = proj_energy(T1.d(), T2.d());
    if (NormR < o.thr && std::fabs(En - Eh) < thren) {
      conv = 1;
      break;
    }
  }
  if (it > o.maxit) it = o.maxit;
  energies[0] = Eh; energies[1] = En; energies[2] = En - Eh; energies[3] = NormR;
  *converged = conv;
  return it;
}

// (T) of UCCSD: calc_pertT_unrestricted (src/cc/cc_triples.jl:297-320) with its same-spin and
// mixed-spin parts (:357-640) in spin-orbital form, on the full [a,b,c,i,j,k] tensors:
//   Z = sum_e T2[a,e,j,k] <bc||ei> - sum_m T2[b,c,i,m] <ma||jk>,  W = P(a/bc) P(i/jk) Z,
//   X = W / (e_i + e_j + e_k - e_a - e_b - e_c),  E[T] = W.X / 36,
//   E(T) = E[T] + X.P(a/bc) P(i/jk) (T1[a,i] <jk||bc> + f[i,a] T2[b,c,j,k]) / 36
// (the singles and Fock terms of the reference's IntX / IntY; diagonal-Fock denominators).
// Memory: five V^3 O^3 arrays in the contraction arena.
__global__ void divide_by_denominator_kernel(int V, int O, const double* __restrict__ eo, const double* __restrict__ ev,
                                             const double* __restrict__ W, double* X) {
  const long v3 = (long)V * V * V, n = v3 * O * O * O;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x) {
    long r = idx;
    const int a = (int)(r % V); r /= V;
    const int b = (int)(r % V); r /= V;
    const int c = (int)(r % V); r /= V;
    const int i = (int)(r % O); r /= O;
    const int j = (int)(r % O);
    const int k = (int)(r / O);
    X[idx] = W[idx] / (eo[i] + eo[j] + eo[k] - ev[a] - ev[b] - ev[c]);
  }
}

void UContext::pert_triples(const double* T1d, const double* T2d, double* e) {
  const long o = O, v = V, n6 = v * v * v * o * o * o;
  ops.arena.reset();
  ECG_CHECK(ops.arena.buf.bytes >= sizeof(double) * (size_t)(5 * n6) + 4096,
            "open-shell (T): the V^3 O^3 arrays exceed the workspace (raise workspace_bytes)");
  double* z = ops.arena.take(n6);
  double* w = ops.arena.take(n6);
  double* x = ops.arena.take(n6);
  double* y = ops.arena.take(n6);
  double* pw = ops.arena.take(n6);
  auto t6 = [&](double* p) { return Tensor(p, {v, v, v, o, o, o}); };
  Tensor Z = t6(z), W = t6(w), Y = t6(y), PW = t6(pw);
  Tensor T1 = t2(cp(T1d), v, o), T2 = t4(cp(T2d), v, v, o, o);
  Tensor Aoovv = t4(oovv.d(), o, o, v, v), Aovvv = t4(ovvv.d(), o, v, v, v), Aooov = t4(ooov.d(), o, o, o, v);
  const Prec F = Prec::F64;
  // <bc||ei> = -<ie||bc>;  <ma||jk> = <jk||ma>
  ops.contract_p(F, -1.0, T2, "aejk", Aovvv, "iebc", 0.0, Z, "abcijk");
  ops.contract_p(F, -1.0, T2, "bcim", Aooov, "jkma", 1.0, Z, "abcijk");
  // P(a/bc) P(i/jk): X - X(i<->j) - X(i<->k), then the same on a, b, c
  const int s_ij[6] = {0, 1, 2, 4, 3, 5}, s_ik[6] = {0, 1, 2, 5, 4, 3}, s_ab[6] = {1, 0, 2, 3, 4, 5},
            s_ac[6] = {2, 1, 0, 3, 4, 5};
  auto pp = [&](Tensor& src, double* tmp_p, Tensor& tmp, Tensor& dst) {
    ops.copy(n6, src.p, tmp_p);
    ops.permute(src, s_ij, tmp, -1.0, 1.0);
    ops.permute(src, s_ik, tmp, -1.0, 1.0);
    ops.copy(n6, tmp_p, dst.p);
    ops.permute(tmp, s_ab, dst, -1.0, 1.0);
    ops.permute(tmp, s_ac, dst, -1.0, 1.0);
  };
  pp(Z, pw, PW, W);
  hipLaunchKernelGGL(divide_by_denominator_kernel, dim3(ugrid(n6)), dim3(256), 0, stream, V, O, eps.d(), eps.d() + O,
                     w, x);
  ECG_HIP(hipGetLastError());
  const double etb = dot(w, x, n6) / 36.0;
  // disconnected: T1[a,i] <jk||bc> + f[i,a] T2[b,c,j,k]
  Tensor Fov = t2(fov.d(), o, v);
  ops.contract_p(F, 1.0, T1, "ai", Aoovv, "jkbc", 0.0, Z, "abcijk");
  ops.contract_p(F, 1.0, Fov, "ia", T2, "bcjk", 1.0, Z, "abcijk");
  pp(Z, pw, PW, Y);
  const double ed = dot(x, y, n6) / 36.0;
  e[0] = etb + ed;
  e[1] = etb;
  ops.arena.reset();
}

void UContext::to_blocks(const double* T1so, const double* T2so, double* T1a, double* T1b, double* T2a, double* T2b,
                         double* T2ab) const {
  const long va = n - na, vb = n - nb, o = O, v = V;
  auto so2 = [&](long a, long b, long i, long j) { return T2so[a + v * (b + v * (i + o * j))]; };
  if (T1a)
    for (long i = 0; i < na; ++i)
      for (long a = 0; a < va; ++a) T1a[a + va * i] = T1so[a + v * i];
  if (T1b)
    for (long i = 0; i < nb; ++i)
      for (long a = 0; a < vb; ++a) T1b[a + vb * i] = T1so[(va + a) + v * (na + i)];
  if (T2a)
    for (long j = 0; j < na; ++j)
      for (long i = 0; i < na; ++i)
        for (long b = 0; b < va; ++b)
          for (long a = 0; a < va; ++a) T2a[a + va * (b + va * (i + (long)na * j))] = so2(a, b, i, j);
  if (T2b)
    for (long j = 0; j < nb; ++j)
      for (long i = 0; i < nb; ++i)
        for (long b = 0; b < vb; ++b)
          for (long a = 0; a < vb; ++a) T2b[a + vb * (b + vb * (i + (long)nb * j))] = so2(va + a, va + b, na + i, na + j);
  if (T2ab)
    for (long j = 0; j < nb; ++j)
      for (long i = 0; i < na; ++i)
        for (long b = 0; b < vb; ++b)
          for (long a = 0; a < va; ++a) T2ab[a + va * (b + vb * (i + (long)na * j))] = so2(a, va + b, i, na + j);
}

void UContext::from_blocks(const double* T1a, const double* T1b, const double* T2a, const double* T2b,
                           const double* T2ab, double* T1so, double* T2so) const {
  const long va = n - na, vb = n - nb, o = O, v = V;
  std::fill(T1so, T1so + n1(), 0.0);
  std::fill(T2so, T2so + n2(), 0.0);
  auto at = [&](long a, long b, long i, long j) -> double& { return T2so[a + v * (b + v * (i + o * j))]; };
  for (long i = 0; i < na; ++i)
    for (long a = 0; a < va; ++a) T1so[a + v * i] = T1a[a + va * i];
  for (long i = 0; i < nb; ++i)
    for (long a = 0; a < vb; ++a) T1so[(va + a) + v * (na + i)] = T1b[a + vb * i];
  for (long j = 0; j < na; ++j)
    for (long i = 0; i < na; ++i)
      for (long b = 0; b < va; ++b)
        for (long a = 0; a < va; ++a) at(a, b, i, j) = T2a[a + va * (b + va * (i + (long)na * j))];
  for (long j = 0; j < nb; ++j)
    for (long i = 0; i < nb; ++i)
      for (long b = 0; b < vb; ++b)
        for (long a = 0; a < vb; ++a) at(va + a, va + b, na + i, na + j) = T2b[a + vb * (b + vb * (i + (long)nb * j))];
  for (long j = 0; j < nb; ++j)
    for (long i = 0; i < na; ++i)
      for (long b = 0; b < vb; ++b)
        for (long a = 0; a < va; ++a) {
          const double x = T2ab[a + va * (b + vb * (i + (long)na * j))];
          at(a, va + b, i, na + j) = x;
          at(va + b, a, i, na + j) = -x;
          at(a, va + b, na + j, i) = -x;
          at(va + b, a, na + j, i) = x;
        }
}

}  // namespace ecg
