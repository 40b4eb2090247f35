// Gradient of the closed-shell CCSD/DCSD correlation energy with respect to the
// one-electron integrals h and the DF factors: B for Hermitian integrals (BR = BL = B),
// the left and right factors BL, BR for two-sided ones, (pq|rs) = sum_Q BL[p,q,Q] BR[r,s,Q].
//
// At converged amplitudes T and Lambda multipliers U the energy is stationary in
// both, so dE/dx = dL/dx for the Lagrangian
//     L(T, U; h, B) = E(T) + sum U1 R1(T) + sum U2 R2(T)
// (R the residual of calc_cc_resid, src/cc/cc.jl:1343-1489; E the doubles plus
// singles energy, cc.jl:97-167; U as solved by lm_cc_iterations!,
// src/cc/cc_lagrange.jl:1155-1223).  This file evaluates dL/dh and dL/dB by
// reverse mode through every DF contraction of the residual, in the order of
// Context::residual: each product  C = A (x) B  with an integral factor
// contributes  Gbar_A += Gbar_C (x) B  and  Gbar_B += A (x) Gbar_C, the
// dressed Fock matrix f^ = h^ + 2 J^ - K^ and the energy are differentiated the
// same way, and the T1 similarity transform of the factors and of h is reversed
// last.  The particle-particle ladder term, sum_cd sum_Q L^_ac^Q L^_bd^Q T_cd^ij,
// gives  Gbar_vv[a,c,Q] += 2 sum_bd M[a,b,c,d] L^_vv[b,d,Q],  M = sum_ij U2 T,
// assembled blockwise like the ladder itself, so no v^4 array is formed.  With two-sided
// factors every product of a left and a right factor sends its adjoints to the left and
// the right blocks separately (GL, GR; for Hermitian integrals the two are one buffer),
// and the ladder gives GL_vv += M R^_vv and GR_vv += M L^_vv.  The two-sided gradient is
// the derivative of the residual's function of (BL, BR); along directions that keep the
// (pq),(rs) supermatrix symmetric (the integrals' particle symmetry) it is the derivative
// of the CC energy.
// dE/dh equals the transposed dressed one-particle density dD1 of calc_1RDM.
#include "ecg_grad.h"

#include <algorithm>
#include <vector>

namespace ecg {

namespace {

Tensor V1(double* p, long a) { return Tensor(p, {a}); }
Tensor V2(double* p, long a, long b) { return Tensor(p, {a, b}); }
Tensor V3(double* p, long a, long b, long c) { return Tensor(p, {a, b, c}); }
Tensor V4(double* p, long a, long b, long c, long d) { return Tensor(p, {a, b, c, d}); }
double* mu(const double* p) { return const_cast<double*>(p); }

struct Buf {
  DevBuf b;
  long n = 0;
  Buf(TensorOps& ops, long count) : b(sizeof(double) * count), n(count) { ops.fill(count, 0.0, b.d()); }
  double* d() const { return b.d(); }
};

}  // namespace

void energy_grad(Context& c, int method, const double* T1h, const double* T2h, const double* U1h, const double* U2h,
                 double* dEdh, double* dEdBL, double* dEdBR) {
  const bool herm = c.hermitian;
  const bool dc = (method == 1);
  const double fac = dc ? 0.5 : 1.0;
  const long o = c.no, v = c.nv, Q = c.nq, n = c.n;
  const long n1 = o * v, n2 = v * v * o * o;
  TensorOps& ops = c.ops;
  hipStream_t s = c.stream;
  ops.gemm.stream = s;
  ops.arena.reset();

  // ---- inputs and amplitude combinations
  DevBuf dT1(sizeof(double) * n1), dT2(sizeof(double) * n2), dU1(sizeof(double) * n1), dU2(sizeof(double) * n2);
  ECG_HIP(hipMemcpyAsync(dT1.p, T1h, dT1.bytes, hipMemcpyHostToDevice, s));
  ECG_HIP(hipMemcpyAsync(dT2.p, T2h, dT2.bytes, hipMemcpyHostToDevice, s));
  ECG_HIP(hipMemcpyAsync(dU1.p, U1h, dU1.bytes, hipMemcpyHostToDevice, s));
  ECG_HIP(hipMemcpyAsync(dU2.p, U2h, dU2.bytes, hipMemcpyHostToDevice, s));
  c.dress(dT1.d());   // dressed factors Ld (Rd), dressed h (dh), dressed Fock (df), trace of the right oo block
  const Blocks& D = c.Ld;                      // dressed left factor
  const Blocks& DR = herm ? c.Ld : c.Rd;       // dressed right factor
  const Blocks& U = c.L;                       // undressed left factor
  const Blocks& UR = herm ? c.L : c.R;         // undressed right factor

  Tensor t1 = V2(dT1.d(), v, o), tT = V4(dT2.d(), v, v, o, o), tU1 = V2(dU1.d(), v, o);
  const int p_ba[4] = {1, 0, 2, 3}, p_ij[4] = {0, 1, 3, 2}, p_sym[4] = {1, 0, 3, 2};
  DevBuf Tt(sizeof(double) * n2), Tp(sizeof(double) * n2), Tm(sizeof(double) * n2), U2s(sizeof(double) * n2),
      L2r(sizeof(double) * n2);
  Tensor tTt = V4(Tt.d(), v, v, o, o), tTp = V4(Tp.d(), v, v, o, o), tTm = V4(Tm.d(), v, v, o, o),
         tU2 = V4(U2s.d(), v, v, o, o), tL2r = V4(L2r.d(), v, v, o, o);
  ops.axpby(n2, 2.0, dT2.d(), 0.0, Tt.d());
  ops.permute(tT, p_ba, tTt, -1.0, 1.0);          // T~ = 2T - T[b,a,i,j]
  ops.axpby(n2, 2.0, dT2.d(), 0.0, Tp.d());
  ops.permute(tT, p_ij, tTp, -1.0, 1.0);          // 2T[c,d,i,k] - T[c,d,k,i]
  ops.copy(n2, dT2.d(), Tm.d());
  ops.permute(tT, p_ba, tTm, -1.0, 1.0);          // T - T[b,a,i,j]
  {
    // U2 enters only through U2 . R2 with R2[a,b,i,j] = R2[b,a,j,i]: use its symmetric part
    Tensor raw = V4(dU2.d(), v, v, o, o);
    ops.axpby(n2, 0.5, dU2.d(), 0.0, U2s.d());
    ops.permute(raw, p_sym, tU2, 0.5, 1.0);
  }
  ops.axpby(n2, 2.0, U2s.d(), 0.0, L2r.d());      // adjoint of R2r: U2 + U2[b,a,j,i]

  // ---- adjoints: dressed factor blocks (left g*, right gr*; one buffer for Hermitian
  // integrals), dressed Fock blocks, <kl|cd>, <ij|ab>
  Buf Goo(ops, o * o * Q), Gov(ops, o * v * Q), Gvo(ops, v * o * Q), Gvv(ops, v * v * Q);
  Buf GRoo(ops, herm ? 0 : o * o * Q), GRov(ops, herm ? 0 : o * v * Q), GRvo(ops, herm ? 0 : v * o * Q),
      GRvv(ops, herm ? 0 : v * v * Q);
  Buf Foo(ops, o * o), Fov(ops, o * v), Fvo(ops, v * o), Fvv(ops, v * v);   // d L / d f^
  Buf Gkl(ops, n2);                                                       // d L / d <kl|cd> [k,l,c,d]
  Tensor goo = V3(Goo.d(), o, o, Q), gov = V3(Gov.d(), o, v, Q), gvo = V3(Gvo.d(), v, o, Q),
         gvv = V3(Gvv.d(), v, v, Q);
  Tensor groo = herm ? goo : V3(GRoo.d(), o, o, Q), grov = herm ? gov : V3(GRov.d(), o, v, Q),
         grvo = herm ? gvo : V3(GRvo.d(), v, o, Q);   // the right vv adjoint is written by the ladder only
  Tensor loo = V3(D.oo.d(), o, o, Q), lov = V3(D.ov.d(), o, v, Q), lvo = V3(D.vo.d(), v, o, Q),
         lvv = V3(D.vv.d(), v, v, Q);
  Tensor roo = V3(DR.oo.d(), o, o, Q), rov = V3(DR.ov.d(), o, v, Q), rvo = V3(DR.vo.d(), v, o, Q),
         rvv = V3(DR.vv.d(), v, v, Q);
  Tensor foo = V2(Foo.d(), o, o), fov = V2(Fov.d(), o, v), fvo = V2(Fvo.d(), v, o), fvv = V2(Fvv.d(), v, v);
  Tensor gkl = V4(Gkl.d(), o, o, v, v);
  const Prec F = Prec::F64;

  // R1 = f^_vo + T~[a,b,i,j] f^[j,b] - (kc|ji)^ T~[c,a,k,j] + sum_Q L^[a,c,Q] Z[c,i,Q]
  ops.axpby(n1, 1.0, dU1.d(), 1.0, Fvo.d());
  ops.contract_p(F, 1.0, tU1, "ai", tTt, "abij", 1.0, fov, "jb");
  {
    DevBuf g(sizeof(double) * o * o * v * o);
    Tensor tg = V4(g.d(), o, o, v, o);
    ops.contract_p(F, -1.0, tU1, "ai", tTt, "cakj", 0.0, tg, "kjci");
    ops.contract_p(F, 1.0, tg, "kjci", roo, "jiQ", 1.0, gov, "kcQ");
    ops.contract_p(F, 1.0, tg, "kjci", lov, "kcQ", 1.0, groo, "jiQ");
  }
  {
    DevBuf z(sizeof(double) * v * o * Q), gz(sizeof(double) * v * o * Q);
    Tensor tz = V3(z.d(), v, o, Q), tgz = V3(gz.d(), v, o, Q);
    ops.contract_p(F, 1.0, tTp, "cdik", rov, "kdQ", 0.0, tz, "ciQ");
    ops.contract_p(F, 1.0, tU1, "ai", tz, "ciQ", 1.0, gvv, "acQ");
    ops.contract_p(F, 1.0, tU1, "ai", lvv, "acQ", 0.0, tgz, "ciQ");
    ops.contract_p(F, 1.0, tgz, "ciQ", tTp, "cdik", 1.0, grov, "kdQ");
  }
  // R2: hole ladder W[k,l,i,j] = (ki|lj)^ (+ <kl|cd> T[c,d,i,j]),  R2 = W T[a,b,k,l]
  {
    DevBuf g(sizeof(double) * o * o * o * o);
    Tensor tg = V4(g.d(), o, o, o, o);
    ops.contract_p(F, 1.0, tU2, "abij", tT, "abkl", 0.0, tg, "klij");
    ops.contract_p(F, 1.0, tg, "klij", roo, "ljQ", 1.0, goo, "kiQ");
    ops.contract_p(F, 1.0, tg, "klij", loo, "kiQ", 1.0, groo, "ljQ");
    if (!dc) ops.contract_p(F, 1.0, tg, "klij", tT, "cdij", 1.0, gkl, "klcd");
  }
  // R2 += (ai|bj)^
  ops.contract_p(F, 1.0, tU2, "abij", rvo, "bjQ", 1.0, gvo, "aiQ");
  ops.contract_p(F, 1.0, tU2, "abij", lvo, "aiQ", 1.0, grvo, "bjQ");
  DevBuf X(sizeof(double) * n2);
  if (!dc) {
    // R2 += <kl|cd> T[a,d,k,j] T[c,b,i,l]
    Tensor tx = V4(X.d(), o, v, v, o);
    ops.contract_p(F, 1.0, tU2, "abij", tT, "cbil", 0.0, tx, "lcaj");
    ops.contract_p(F, 1.0, tx, "lcaj", tT, "adkj", 1.0, gkl, "klcd");
  }
  // R2r = x_ad T[d,b,i,j] - x_ki T[a,b,k,j],  x_ad = f^_vv - fac <kl|cd> T~[c,a,k,l],
  //       x_ki = f^_oo + fac <kl|cd> T~[c,d,i,l]
  {
    DevBuf gad(sizeof(double) * v * v), gki(sizeof(double) * o * o);
    Tensor tgad = V2(gad.d(), v, v), tgki = V2(gki.d(), o, o);
    ops.contract_p(F, 1.0, tL2r, "abij", tT, "dbij", 0.0, tgad, "ad");
    ops.contract_p(F, -1.0, tL2r, "abij", tT, "abkj", 0.0, tgki, "ki");
    ops.axpby(v * v, 1.0, gad.d(), 1.0, Fvv.d());
    ops.axpby(o * o, 1.0, gki.d(), 1.0, Foo.d());
    ops.contract_p(F, -fac, tgad, "ad", tTt, "cakl", 1.0, gkl, "klcd");
    ops.contract_p(F, fac, tgki, "ki", tTt, "cdil", 1.0, gkl, "klcd");
  }
  // particle-particle ladder  sum_Q sum_cd L^[a,c,Q] R^[b,d,Q] T[c,d,i,j]  (in R2 once):
  //   GL_vv[a,c,Q] += sum_{b,d} M[a,b,c,d] R^[b,d,Q],  GR_vv[a,c,Q] += sum_{b,d} M[a,b,c,d] L^[b,d,Q]
  //   (M[a,b,c,d] = sum_ij U2[a,b,i,j] T[c,d,i,j] = M[b,a,d,c]); Hermitian: Gvv += 2 M L^
  {
    DevBuf LvvT(sizeof(double) * v * v * Q);   // L^vv as [Q, b, d]
    DevBuf RvvT(sizeof(double) * (herm ? 1 : v * v * Q));   // R^vv as [Q, b, d]
    Tensor tLT = V3(LvvT.d(), Q, v, v), tRT = V3(RvvT.d(), Q, v, v);
    const int p_q[3] = {2, 0, 1};
    ops.permute(lvv, p_q, tLT, 1.0, 0.0);
    if (!herm) ops.permute(rvv, p_q, tRT, 1.0, 0.0);
    size_t free_b = ops.arena.free_bytes();
    const long per_a = v * v * v;
    long ab = std::max<long>(1, std::min<long>(v, (long)(free_b * 0.8 / sizeof(double)) / per_a));
    size_t mark = ops.arena.mark();
    double* M = ops.arena.take(ab * per_a);
    for (long a0 = 0; a0 < v; a0 += ab) {
      const long na = std::min<long>(ab, v - a0);
      // M[a',c,b,d] = sum_ij U2[a0+a',b,i,j] T[c,d,i,j]   (m = (a',b), n = (c,d), k = (i,j))
      GemmOp g1;
      g1.M = (int)(na * v); g1.N = (int)(v * v); g1.K = (int)(o * o);
      g1.A = U2s.d() + a0; g1.a_s0 = 1; g1.a_d0 = (int)na; g1.a_s1 = v; g1.a_k = v * v;
      g1.B = dT2.d(); g1.b_s0 = 1; g1.b_k = v * v;
      g1.C = M; g1.c_ms0 = 1; g1.c_md0 = (int)na; g1.c_ms1 = na * v;      // a' -> 1, b -> na*v
      g1.c_ns0 = na; g1.c_nd0 = (int)v; g1.c_ns1 = na * v * v;              // c -> na, d -> na*v*v
      g1.alpha = 1.0; g1.beta = 0.0;
      ops.gemm.run_prec(g1, F);
      // G_vv[a0+a',c,Q] += w sum_{(b,d)} M[(a',c),(b,d)] X^T[Q,(b,d)]   (m = (a',c), k = (b,d), n = Q)
      auto adjoint = [&](const double* XT, double* G, double w) {
        GemmOp g2;
        g2.M = (int)(na * v); g2.N = (int)Q; g2.K = (int)(v * v);
        g2.A = M; g2.a_s0 = 1; g2.a_k = na * v;
        g2.B = XT; g2.b_s0 = 1; g2.b_k = Q;
        g2.C = G + a0; g2.c_ms0 = 1; g2.c_md0 = (int)na; g2.c_ms1 = v; g2.c_ns0 = v * v;
        g2.alpha = w; g2.beta = 1.0;
        ops.gemm.run_prec(g2, F);
      };
      if (herm) {
        adjoint(LvvT.d(), Gvv.d(), 2.0);
      } else {
        adjoint(RvvT.d(), Gvv.d(), 1.0);    // left factor's adjoint takes the right factor
        adjoint(LvvT.d(), GRvv.d(), 1.0);
      }
    }
    ops.arena.release_to(mark);
  }
  // Y[a,k,i,c] = (ai|kc)^ + 1/2 <kl|cd> T~[a,d,i,l];  R2r += Y T~[c,b,k,j]
  {
    Tensor ty = V4(X.d(), v, o, o, v);
    ops.contract_p(F, 1.0, tL2r, "abij", tTt, "cbkj", 0.0, ty, "akic");
    ops.contract_p(F, 1.0, ty, "akic", rov, "kcQ", 1.0, gvo, "aiQ");
    ops.contract_p(F, 1.0, ty, "akic", lvo, "aiQ", 1.0, grov, "kcQ");
    ops.contract_p(F, 0.5, ty, "akic", tTt, "adil", 1.0, gkl, "klcd");
  }
  if (!dc) {
    // R2r -= <kl|cd> T[d,a,k,i] (T - T[b,a])[c,b,l,j]
    Tensor tz = V4(X.d(), o, v, v, o);
    ops.contract_p(F, -1.0, tL2r, "abij", tTm, "cblj", 0.0, tz, "lcai");
    ops.contract_p(F, 1.0, tz, "lcai", tT, "daki", 1.0, gkl, "klcd");
  }
  {
    // R2r -= (ac|ki)^ T[c,b,k,j] + (bc|ki)^ T[a,c,k,j]
    Tensor tg = V4(X.d(), v, o, v, o);
    ops.contract_p(F, -1.0, tL2r, "abij", tT, "cbkj", 0.0, tg, "akci");
    ops.contract_p(F, -1.0, tL2r, "axij", tT, "ackj", 1.0, tg, "xkci");
    ops.contract_p(F, 1.0, tg, "akci", roo, "kiQ", 1.0, gvv, "acQ");
    ops.contract_p(F, 1.0, tg, "akci", lvv, "acQ", 1.0, groo, "kiQ");
  }
  // <kl|cd> = sum_Q L[k,c,Q] R[l,d,Q] of the undressed ov blocks (identical to the dressed ones)
  ops.contract_p(F, 1.0, gkl, "klcd", rov, "ldQ", 1.0, gov, "kcQ");
  ops.contract_p(F, 1.0, gkl, "klcd", lov, "kcQ", 1.0, grov, "ldQ");
  // energy: E = sum (2T - T[b,a]) <ij|ab> + sum (2 t t - (t t)[b,a]) <ij|ab> + 2 sum T1 f_ov
  DevBuf Fock_ov(sizeof(double) * o * v);   // d L / d f_ov[i,a] of the undressed Fock matrix
  Tensor fk = V2(Fock_ov.d(), o, v);
  {
    Tensor gab = V4(X.d(), v, v, o, o);
    DevBuf tt(sizeof(double) * n2);
    Tensor ttt = V4(tt.d(), v, v, o, o);
    ops.contract_p(F, 1.0, t1, "ai", t1, "bj", 0.0, ttt, "abij");
    ops.copy(n2, Tt.d(), X.d());
    ops.axpby(n2, 2.0, tt.d(), 1.0, X.d());
    ops.permute(ttt, p_ba, gab, -1.0, 1.0);
    ops.contract_p(F, 1.0, gab, "abij", rov, "jbQ", 1.0, gov, "iaQ");
    ops.contract_p(F, 1.0, gab, "abij", lov, "iaQ", 1.0, grov, "jbQ");
    const int p_t[2] = {1, 0};
    ops.permute(t1, p_t, fk, 2.0, 0.0);
  }
  // dressed Fock f^ = h^ + 2 L^ tr - sum_k L^[p,k] L^[k,q],  tr[Q] = sum_k L^oo[k,k,Q]
  std::vector<double> eye((size_t)o * o, 0.0);
  for (long k = 0; k < o; ++k) eye[(size_t)k * o + k] = 1.0;
  DevBuf I(sizeof(double) * o * o);
  ECG_HIP(hipMemcpyAsync(I.p, eye.data(), I.bytes, hipMemcpyHostToDevice, s));
  Tensor ti = V2(I.d(), o, o);
  {
    Buf Gtr(ops, Q);
    Tensor gtr = V1(Gtr.d(), Q), tr = V1(c.trace.d(), Q);
    // f^[p,q] = h^ + 2 sum_Q L^[p,q,Q] tr[Q] - sum_kQ L^[p,k,Q] R^[k,q,Q],  tr[Q] = sum_k R^oo[k,k,Q]
    Tensor* Fb[4] = {&foo, &fov, &fvo, &fvv};
    Tensor* Lq[4] = {&loo, &lov, &lvo, &lvv};
    Tensor* Gq[4] = {&goo, &gov, &gvo, &gvv};
    Tensor* Lk[4] = {&loo, &loo, &lvo, &lvo};
    Tensor* Gk[4] = {&goo, &goo, &gvo, &gvo};
    Tensor* Rk[4] = {&roo, &rov, &roo, &rov};
    Tensor* Gr[4] = {&groo, &grov, &groo, &grov};
    for (int b = 0; b < 4; ++b) {
      ops.contract_p(F, 2.0, *Fb[b], "pq", tr, "Q", 1.0, *Gq[b], "pqQ");
      ops.contract_p(F, 2.0, *Fb[b], "pq", *Lq[b], "pqQ", 1.0, gtr, "Q");
      ops.contract_p(F, -1.0, *Fb[b], "pq", *Rk[b], "kqQ", 1.0, *Gk[b], "pkQ");
      ops.contract_p(F, -1.0, *Fb[b], "pq", *Lk[b], "pkQ", 1.0, *Gr[b], "kqQ");
    }
    ops.contract_p(F, 1.0, ti, "kl", gtr, "Q", 1.0, groo, "klQ");
  }
  // undressed Fock (energy's f_ov): f = h + 2 L tr_u - sum_k L[p,k] R[k,q], tr_u[Q] = sum_k R[k,k,Q]
  Buf UGoo(ops, o * o * Q), UGov(ops, o * v * Q), UGvo(ops, v * o * Q), UGvv(ops, v * v * Q);
  Buf URGoo(ops, herm ? 0 : o * o * Q), URGov(ops, herm ? 0 : o * v * Q), URGvo(ops, herm ? 0 : v * o * Q),
      URGvv(ops, herm ? 0 : v * v * Q);
  Tensor ugoo = V3(UGoo.d(), o, o, Q), ugov = V3(UGov.d(), o, v, Q);
  Tensor urgoo = herm ? ugoo : V3(URGoo.d(), o, o, Q), urgov = herm ? ugov : V3(URGov.d(), o, v, Q);
  Tensor uoo = V3(U.oo.d(), o, o, Q), uov = V3(U.ov.d(), o, v, Q);
  Tensor ruoo = V3(UR.oo.d(), o, o, Q), ruov = V3(UR.ov.d(), o, v, Q);
  {
    Buf utr(ops, Q), gutr(ops, Q);
    Tensor tu = V1(utr.d(), Q), tgu = V1(gutr.d(), Q);
    ops.contract_p(F, 1.0, ti, "kl", ruoo, "klQ", 0.0, tu, "Q");
    ops.contract_p(F, 2.0, fk, "ia", tu, "Q", 1.0, ugov, "iaQ");
    ops.contract_p(F, 2.0, fk, "ia", uov, "iaQ", 0.0, tgu, "Q");
    ops.contract_p(F, 1.0, ti, "kl", tgu, "Q", 1.0, urgoo, "klQ");
    ops.contract_p(F, -1.0, fk, "ia", ruov, "kaQ", 1.0, ugoo, "ikQ");
    ops.contract_p(F, -1.0, fk, "ia", uoo, "ikQ", 1.0, urgov, "kaQ");
  }
  // reverse of the T1 similarity transform of each factor:
  //   oo^ = oo + ov t,  vv^ = vv - t ov,  vo^ = vo + vv^ t - t oo,  ov^ = ov
  auto undress = [&](Buf& Go, Buf& Gv, Buf& Gvo_, Buf& Gvv_, Buf& UGo, Buf& UGv, Buf& UGvo_, Buf& UGvv_) {
    Tensor g_oo = V3(Go.d(), o, o, Q), g_vo = V3(Gvo_.d(), v, o, Q), g_vv = V3(Gvv_.d(), v, v, Q);
    Tensor ug_oo = V3(UGo.d(), o, o, Q), ug_ov = V3(UGv.d(), o, v, Q);
    ops.contract_p(F, 1.0, g_vo, "aiQ", t1, "bi", 1.0, g_vv, "abQ");
    ops.axpby(v * o * Q, 1.0, Gvo_.d(), 1.0, UGvo_.d());
    ops.contract_p(F, -1.0, t1, "aj", g_vo, "aiQ", 1.0, ug_oo, "jiQ");
    ops.axpby(v * v * Q, 1.0, Gvv_.d(), 1.0, UGvv_.d());
    ops.contract_p(F, -1.0, t1, "ai", g_vv, "abQ", 1.0, ug_ov, "ibQ");
    ops.axpby(o * o * Q, 1.0, Go.d(), 1.0, UGo.d());
    ops.contract_p(F, 1.0, g_oo, "ijQ", t1, "bj", 1.0, ug_ov, "ibQ");
    ops.axpby(o * v * Q, 1.0, Gv.d(), 1.0, UGv.d());
  };
  undress(Goo, Gov, Gvo, Gvv, UGoo, UGov, UGvo, UGvv);
  if (!herm) undress(GRoo, GRov, GRvo, GRvv, URGoo, URGov, URGvo, URGvv);
  // dh = (1 - t) h (1 + t) reversed; d L / d h^ = d L / d f^
  Buf Hoo(ops, o * o), Hov(ops, o * v), Hvo(ops, v * o), Hvv(ops, v * v);
  Tensor hoo = V2(Hoo.d(), o, o), hov = V2(Hov.d(), o, v);
  ops.contract_p(F, 1.0, fvo, "ai", t1, "bi", 1.0, fvv, "ab");
  ops.axpby(v * o, 1.0, Fvo.d(), 1.0, Hvo.d());
  ops.contract_p(F, -1.0, t1, "aj", fvo, "ai", 1.0, hoo, "ji");
  ops.axpby(v * v, 1.0, Fvv.d(), 1.0, Hvv.d());
  ops.contract_p(F, -1.0, t1, "ai", fvv, "ab", 1.0, hov, "ib");
  ops.axpby(o * o, 1.0, Foo.d(), 1.0, Hoo.d());
  ops.contract_p(F, 1.0, foo, "ij", t1, "bj", 1.0, hov, "ib");
  ops.axpby(o * v, 1.0, Fov.d(), 1.0, Hov.d());
  ops.axpby(o * v, 1.0, Fock_ov.d(), 1.0, Hov.d());

  // ---- assemble dE/dh (n x n) and dE/dB (n x n x Q) on the host
  auto get = [&](const DevBuf& b, long count) {
    std::vector<double> h((size_t)count);
    ECG_HIP(hipMemcpyAsync(h.data(), b.p, sizeof(double) * count, hipMemcpyDeviceToHost, s));
    return h;
  };
  std::vector<double> hoo_h = get(Hoo.b, o * o), hov_h = get(Hov.b, o * v), hvo_h = get(Hvo.b, v * o),
                      hvv_h = get(Hvv.b, v * v);
  std::vector<double> boo = get(UGoo.b, o * o * Q), bov = get(UGov.b, o * v * Q), bvo = get(UGvo.b, v * o * Q),
                      bvv = get(UGvv.b, v * v * Q);
  std::vector<double> roo_h, rov_h, rvo_h, rvv_h;
  if (!herm) {
    roo_h = get(URGoo.b, o * o * Q); rov_h = get(URGov.b, o * v * Q);
    rvo_h = get(URGvo.b, v * o * Q); rvv_h = get(URGvv.b, v * v * Q);
  }
  ECG_HIP(hipStreamSynchronize(s));
  auto put = [&](double* dst, const std::vector<double>& src, long rows, long cols, long r0, long c0, long nq) {
    for (long q = 0; q < nq; ++q)
      for (long j = 0; j < cols; ++j)
        for (long i = 0; i < rows; ++i)
          dst[(size_t)q * n * n + (size_t)(c0 + j) * n + (r0 + i)] = src[(size_t)q * rows * cols + (size_t)j * rows + i];
  };
  if (dEdh) {
    put(dEdh, hoo_h, o, o, 0, 0, 1);
    put(dEdh, hov_h, o, v, 0, o, 1);
    put(dEdh, hvo_h, v, o, o, 0, 1);
    put(dEdh, hvv_h, v, v, o, o, 1);
  }
  if (dEdBL) {
    put(dEdBL, boo, o, o, 0, 0, Q);
    put(dEdBL, bov, o, v, 0, o, Q);
    put(dEdBL, bvo, v, o, o, 0, Q);
    put(dEdBL, bvv, v, v, o, o, Q);
  }
  if (dEdBR && !herm) {
    put(dEdBR, roo_h, o, o, 0, 0, Q);
    put(dEdBR, rov_h, o, v, 0, o, Q);
    put(dEdBR, rvo_h, v, o, o, 0, Q);
    put(dEdBR, rvv_h, v, v, o, o, Q);
  }
  ops.arena.reset();
}

}  // namespace ecg
