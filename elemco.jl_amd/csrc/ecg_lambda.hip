// Closed-shell Lambda equations on the device: the vector x Jacobian residual,
// the one-particle density and the Lambda iterations.
//
// Restates calc_ccsd_vector_times_Jacobian (src/cc/cc_lagrange.jl:9-212),
// calc_dU2 (:222-236), calc_1RDM (:793-837), calc_3ext_times_T2 (:948-952),
// calc_correlation_norm (:1065-1077) and lm_cc_iterations! (:1155-1223).
// The dressed four-index blocks d_xxxx of the reference are assembled from the
// T1-dressed DF factors, <pq|rs>^ = sum_Q L^[p,r,Q] R^[q,s,Q].  The kext term
// K[r,s,m,n] = sum_pq <pq|rs> dU2[p,q,m,n] (dU2: U2 dressed with T1 on the bra
// side) enters as  R2 += sum_ab U2[a,b,m,n] <ab|ef>^, the transpose of the CC
// ladder (assembled on MFMA from the factor blocks, half of it plus symmetry), and
// R1 += 2 sum_j K^[e,j,m,j] = 2 sum_Q sum_a L^[a,e,Q] (sum_bj U2[a,b,m,j] R^[b,j,Q]).
#include "ecg_lambda.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace ecg {

namespace {

Tensor T1v(double* p, long a) { return Tensor(p, {a}); }
Tensor T2d(double* p, long a, long b) { return Tensor(p, {a, b}); }
Tensor T3d(double* p, long a, long b, long c) { return Tensor(p, {a, b, c}); }
Tensor T4d(double* p, long a, long b, long c, long d) { return Tensor(p, {a, b, c, d}); }
double* cp(const double* p) { return const_cast<double*>(p); }

// Dressed blocks and intermediates that stay fixed during the Lambda iterations.
struct LambdaFixed {
  DevBuf T1, T2;                                   // converged amplitudes
  DevBuf doooo, dvoov, dvovv, doovo, dvovo, dvooo;  // dressed integrals (reference index order)
  DevBuf vT;                                       // vT_ovoo[m,b,i,j]
  DevBuf Lt, Rt;                                   // dressed vv factors transposed, [e,a,Q]
  DevBuf dfoo, dfov, dfvv, fov;                    // dressed Fock blocks, undressed f_ov
};

// copy block [r0:r0+m, c0:c0+k] of an n x n device matrix into a dense m x k one
void block2d(Context& c, const double* src, int r0, int m, int c0, int k, DevBuf& dst) {
  dst.ensure(sizeof(double) * (size_t)m * k);
  ECG_HIP(hipMemcpy2DAsync(dst.p, sizeof(double) * m, src + (size_t)c0 * c.n + r0, sizeof(double) * c.n,
                           sizeof(double) * m, k, hipMemcpyDeviceToDevice, c.stream));
}

// calc_dressed_ints(EC, T1; calc_d_vovv=true) + calc_vT2_intermediates (cc_lagrange.jl:1165-1166)
void prepare(Context& c, const double* T1h, const double* T2h, LambdaFixed& F) {
  const long o = c.no, v = c.nv, Q = c.nq;
  TensorOps& ops = c.ops;
  F.T1.ensure(sizeof(double) * c.nvo());
  F.T2.ensure(sizeof(double) * c.nvvoo());
  ECG_HIP(hipMemcpyAsync(F.T1.p, T1h, F.T1.bytes, hipMemcpyHostToDevice, c.stream));
  ECG_HIP(hipMemcpyAsync(F.T2.p, T2h, F.T2.bytes, hipMemcpyHostToDevice, c.stream));
  ops.arena.reset();
  c.dress(F.T1.d());
  const Blocks& Rh = c.hermitian ? c.Ld : c.Rd;
  const Blocks& Lh = c.Ld;
  auto t3 = [&](const DevBuf& b, long x, long y) { return T3d(b.d(), x, y, Q); };
  F.doooo.ensure(sizeof(double) * o * o * o * o);
  F.dvoov.ensure(sizeof(double) * c.nvvoo());
  F.dvovv.ensure(sizeof(double) * v * o * v * v);
  F.doovo.ensure(sizeof(double) * o * o * v * o);
  F.dvovo.ensure(sizeof(double) * c.nvvoo());
  F.dvooo.ensure(sizeof(double) * v * o * o * o);
  Tensor x;
  // d_oooo[m,n,i,j] = (mi|nj)^
  x = T4d(F.doooo.d(), o, o, o, o);
  ops.contract_p(Prec::F64, 1.0, t3(Lh.oo, o, o), "miQ", t3(Rh.oo, o, o), "njQ", 0.0, x, "mnij");
  // d_voov[a,m,i,e] = (ai|me)^
  x = T4d(F.dvoov.d(), v, o, o, v);
  ops.contract_p(Prec::F64, 1.0, t3(Lh.vo, v, o), "aiQ", t3(Rh.ov, o, v), "meQ", 0.0, x, "amie");
  // d_vovv[a,m,f,e] = (af|me)^
  x = T4d(F.dvovv.d(), v, o, v, v);
  ops.contract_p(Prec::F64, 1.0, t3(Lh.vv, v, v), "afQ", t3(Rh.ov, o, v), "meQ", 0.0, x, "amfe");
  // d_oovo[m,n,e,i] = (me|ni)^
  x = T4d(F.doovo.d(), o, o, v, o);
  ops.contract_p(Prec::F64, 1.0, t3(Lh.ov, o, v), "meQ", t3(Rh.oo, o, o), "niQ", 0.0, x, "mnei");
  // d_vovo[a,m,e,i] = (ae|mi)^
  x = T4d(F.dvovo.d(), v, o, v, o);
  ops.contract_p(Prec::F64, 1.0, t3(Lh.vv, v, v), "aeQ", t3(Rh.oo, o, o), "miQ", 0.0, x, "amei");
  // d_vooo[b,m,j,i] = (bj|mi)^
  x = T4d(F.dvooo.d(), v, o, o, o);
  ops.contract_p(Prec::F64, 1.0, t3(Lh.vo, v, o), "bjQ", t3(Rh.oo, o, o), "miQ", 0.0, x, "bmji");
  // vT_ovoo[m,b,i,j] = d_vovv[b,m,d,c] T2[d,c,j,i]  (cc_lagrange.jl:948-952)
  F.vT.ensure(sizeof(double) * o * v * o * o);
  x = T4d(F.vT.d(), o, v, o, o);
  ops.contract_p(Prec::F64, 1.0, T4d(F.dvovv.d(), v, o, v, v), "bmdc", T4d(F.T2.d(), v, v, o, o), "dcji", 0.0, x,
                 "mbij");
  // transposed dressed vv factors for the Lambda kext: Lt[e,a,Q] = L^[a,e,Q]
  F.Lt.ensure(sizeof(double) * v * v * Q);
  Tensor lt = T3d(F.Lt.d(), v, v, Q);
  const int p_t[3] = {1, 0, 2};
  ops.permute(t3(Lh.vv, v, v), p_t, lt, 1.0, 0.0);
  if (!c.hermitian) {
    F.Rt.ensure(sizeof(double) * v * v * Q);
    Tensor rt = T3d(F.Rt.d(), v, v, Q);
    ops.permute(t3(Rh.vv, v, v), p_t, rt, 1.0, 0.0);
  }
  // Fock blocks
  block2d(c, c.df.d(), 0, (int)o, 0, (int)o, F.dfoo);
  block2d(c, c.df.d(), 0, (int)o, (int)o, (int)v, F.dfov);
  block2d(c, c.df.d(), (int)o, (int)v, (int)o, (int)v, F.dfvv);
  block2d(c, c.fock.d(), 0, (int)o, (int)o, (int)v, F.fov);
  ECG_HIP(hipStreamSynchronize(c.stream));
  ops.arena.reset();
}

// calc_1RDM (cc_lagrange.jl:793-837) on device blocks: D (without T1 terms) and dD.
struct Rdm {
  DevBuf oo, ov, vo, vv, doo, dov, dvo, dvv;
};

void one_rdm(Context& c, const double* T1, const double* T2, const double* U1, const double* U2, Rdm& D) {
  const long o = c.no, v = c.nv;
  TensorOps& ops = c.ops;
  for (DevBuf* b : {&D.oo, &D.doo}) b->ensure(sizeof(double) * o * o);
  for (DevBuf* b : {&D.vv, &D.dvv}) b->ensure(sizeof(double) * v * v);
  for (DevBuf* b : {&D.ov, &D.vo, &D.dov, &D.dvo}) b->ensure(sizeof(double) * v * o);
  Tensor t1 = T2d(cp(T1), v, o), u1 = T2d(cp(U1), v, o);
  Tensor t2 = T4d(cp(T2), v, v, o, o), u2 = T4d(cp(U2), v, v, o, o);
  Tensor Doo = T2d(D.oo.d(), o, o), Dvv = T2d(D.vv.d(), v, v), Dov = T2d(D.ov.d(), o, v), Dvo = T2d(D.vo.d(), v, o);
  ops.contract_p(Prec::F64, -2.0, u2, "cdik", t2, "cdjk", 0.0, Doo, "ij");
  ops.contract_p(Prec::F64, 2.0, u2, "bckl", t2, "ackl", 0.0, Dvv, "ab");
  const int p_t[2] = {1, 0};
  ops.permute(u1, p_t, Dov, 1.0, 0.0);
  ops.contract_p(Prec::F64, 2.0, u1, "ck", t2, "acik", 0.0, Dvo, "ai");
  ops.contract_p(Prec::F64, -1.0, u1, "ck", t2, "acki", 1.0, Dvo, "ai");
  Tensor dDoo = T2d(D.doo.d(), o, o), dDvv = T2d(D.dvv.d(), v, v), dDvo = T2d(D.dvo.d(), v, o);
  ops.copy(o * o, D.oo.d(), D.doo.d());
  ops.copy(v * v, D.vv.d(), D.dvv.d());
  ops.copy(v * o, D.ov.d(), D.dov.d());
  ops.copy(v * o, D.vo.d(), D.dvo.d());
  ops.contract_p(Prec::F64, -1.0, Dov, "ic", t1, "cj", 1.0, dDoo, "ij");
  ops.contract_p(Prec::F64, 1.0, t1, "ak", Dov, "kb", 1.0, dDvv, "ab");
  ops.axpby(v * o, 2.0, T1, 1.0, D.dvo.d());
  ops.contract_p(Prec::F64, -1.0, Dvv, "ac", t1, "ci", 1.0, dDvo, "ai");
  ops.contract_p(Prec::F64, 1.0, t1, "ak", dDoo, "ki", 1.0, dDvo, "ai");
}

// calc_ccsd_vector_times_Jacobian (cc_lagrange.jl:9-212); all pointers on device
void residual(Context& c, LambdaFixed& F, int method, Prec pr, const double* U1d, const double* U2d, double* R1d,
              double* R2d) {
  const bool dc = (method == 1);
  const long o = c.no, v = c.nv, Q = c.nq, n2 = c.nvvoo();
  TensorOps& ops = c.ops;
  ops.arena.reset();
  auto take4 = [&]() { return ops.arena.take(n2); };
  const Blocks& Rh = c.hermitian ? c.Ld : c.Rd;
  const Blocks& Rb = c.hermitian ? c.L : c.R;
  Tensor t2 = T4d(F.T2.d(), v, v, o, o);
  Tensor u1 = T2d(cp(U1d), v, o), u2 = T4d(cp(U2d), v, v, o, o);
  Tensor r1 = T2d(R1d, v, o), r2 = T4d(R2d, v, v, o, o);
  Tensor oovv = T4d(c.oovv.d(), o, o, v, v);
  Tensor doovo = T4d(F.doovo.d(), o, o, v, o), dvovv = T4d(F.dvovv.d(), v, o, v, v);
  Tensor dfov = T2d(F.dfov.d(), o, v);
  Rdm D;
  one_rdm(c, F.T1.d(), F.T2.d(), U1d, U2d, D);
  Tensor D1oo = T2d(D.oo.d(), o, o), D1vv = T2d(D.vv.d(), v, v);
  // R1[e,m] = 2 f[m,e]
  const int p_t[2] = {1, 0};
  ops.permute(T2d(F.fov.d(), o, v), p_t, r1, 2.0, 0.0);
  // R2[e,f,m,n] = 2 <mn|ef> - <nm|ef>
  const int p_a[4] = {2, 3, 0, 1}, p_b[4] = {2, 3, 1, 0};
  ops.permute(oovv, p_a, r2, 2.0, 0.0);
  ops.permute(oovv, p_b, r2, -1.0, 1.0);
  {
    // R2 += (d_oooo[m,n,i,j] (+ <mn|cd> T2[c,d,i,j])) U2[e,f,i,j]
    double* W = ops.arena.take(o * o * o * o);
    Tensor tw = T4d(W, o, o, o, o);
    ops.copy(o * o * o * o, F.doooo.d(), W);
    if (!dc) ops.contract_p(pr, 1.0, oovv, "mncd", t2, "cdij", 1.0, tw, "mnij");
    ops.contract_p(pr, 1.0, tw, "mnij", u2, "efij", 1.0, r2, "efmn");
  }
  {
    // kext: R2 += sum_ab U2[a,b,m,n] <ab|ef>^  (transposed DF ladder, symmetrised)
    size_t mk = ops.arena.mark();
    double* Ux = take4();
    double* P = take4();
    Tensor tux = T4d(Ux, o, o, v, v), tp = T4d(P, v, v, o, o);
    const int p_x[4] = {2, 3, 0, 1}, p_s[4] = {1, 0, 3, 2};
    ops.permute(u2, p_x, tux, 1.0, 0.0);
    ops.fill(n2, 0.0, P);
    c.ladder_df(pr, F.Lt.d(), c.hermitian ? F.Lt.d() : F.Rt.d(), Ux, P);
    ops.axpby(n2, 1.0, P, 1.0, R2d);
    ops.permute(tp, p_s, r2, 1.0, 1.0);
    // R1 += 2 sum_Q sum_a L^[a,e,Q] Z[a,m,Q],  Z[a,m,Q] = sum_bj U2[a,b,m,j] R^[b,j,Q]
    double* Z = ops.arena.take(v * o * Q);
    Tensor tz = T3d(Z, v, o, Q);
    ops.contract_p(pr, 1.0, u2, "abmj", T3d(Rh.vo.d(), v, o, Q), "bjQ", 0.0, tz, "amQ");
    ops.contract_p(Prec::F64, 2.0, T3d(c.Ld.vv.d(), v, v, Q), "aeQ", tz, "amQ", 1.0, r1, "em");
    ops.arena.release_to(mk);
  }
  double fac = dc ? 0.25 : 0.5;
  double* tR2 = take4();
  Tensor ttr = T4d(tR2, v, v, o, o);
  ops.contract_p(pr, fac, D1oo, "mk", oovv, "knef", 0.0, ttr, "efmn");
  ops.contract_p(pr, -fac, D1vv, "ce", oovv, "mncf", 1.0, ttr, "efmn");
  double* tT = take4();
  Tensor ttT = T4d(tT, v, v, o, o);
  const int p_ij[4] = {0, 1, 3, 2};
  ops.axpby(n2, 2.0, F.T2.d(), 0.0, tT);
  ops.permute(t2, p_ij, ttT, -1.0, 1.0);
  double* xoo = ops.arena.take(o * o);
  double* xvv = ops.arena.take(v * v);
  Tensor txoo = T2d(xoo, o, o), txvv = T2d(xvv, v, v);
  ops.contract_p(pr, 1.0, ttT, "cdil", oovv, "mlcd", 0.0, txoo, "mi");
  ops.contract_p(pr, 1.0, ttT, "ackl", oovv, "klec", 0.0, txvv, "ae");
  {
    // tR2 += U2[a,f,i,n] (d_voov[a,m,i,e] + <km|ce> T~[a,c,i,k])
    double* Y = take4();
    Tensor ty = T4d(Y, v, o, o, v);
    ops.copy(n2, F.dvoov.d(), Y);
    ops.contract_p(pr, 1.0, oovv, "kmce", ttT, "acik", 1.0, ty, "amie");
    ops.contract_p(pr, 1.0, u2, "afin", ty, "amie", 1.0, ttr, "efmn");
  }
  ops.contract_p(pr, 0.5, u1, "em", dfov, "nf", 1.0, ttr, "efmn");
  ops.contract_p(pr, 0.5, u1, "an", dvovv, "amfe", 1.0, ttr, "efmn");
  ops.contract_p(pr, -0.5, u1, "fi", doovo, "mnei", 1.0, ttr, "efmn");
  // pR2 = 2 tR2 - tR2[e,f,n,m]
  double* pR2 = take4();
  Tensor tpr = T4d(pR2, v, v, o, o);
  ops.axpby(n2, 2.0, tR2, 0.0, pR2);
  ops.permute(ttr, p_ij, tpr, -1.0, 1.0);
  // D2[a,b,i,j] = U2[a,c,i,k] T~[c,b,k,j]
  double* D2 = tR2;   // tR2 no longer needed
  Tensor td2 = T4d(D2, v, v, o, o);
  ops.contract_p(pr, 1.0, u2, "acik", ttT, "cbkj", 0.0, td2, "abij");
  ops.contract_p(pr, 2.0, td2, "adml", dvovv, "aled", 1.0, r1, "em");
  ops.contract_p(pr, -2.0, td2, "edil", doovo, "lmdi", 1.0, r1, "em");
  if (!dc) ops.contract_p(pr, -1.0, td2, "fcnl", oovv, "mlce", 1.0, tpr, "efmn");
  // Dbar[a,b,i,j] = U2[a,c,i,k] T2[c,b,k,j] + U2[c,a,i,k] T2[b,c,k,j]
  ops.contract_p(pr, 1.0, u2, "acik", t2, "cbkj", 0.0, td2, "abij");
  ops.contract_p(pr, 1.0, u2, "caik", t2, "bckj", 1.0, td2, "abij");
  if (!dc) ops.contract_p(pr, 1.0, td2, "ednk", oovv, "kmfd", 1.0, tpr, "efmn");
  ops.contract_p(pr, -2.0, td2, "acmk", dvovv, "akce", 1.0, r1, "em");
  ops.contract_p(pr, 2.0, td2, "ecik", doovo, "mkci", 1.0, r1, "em");
  {
    // D[i,j,k,l] = U2[c,d,i,j] T2[c,d,k,l]
    double* Do = ops.arena.take(o * o * o * o);
    Tensor tdo = T4d(Do, o, o, o, o);
    ops.contract_p(pr, 1.0, u2, "cdij", t2, "cdkl", 0.0, tdo, "ijkl");
    ops.contract_p(pr, 2.0, tdo, "mjkl", doovo, "klej", 1.0, r1, "em");
    if (!dc) ops.contract_p(pr, 1.0, tdo, "mnkl", oovv, "klef", 1.0, r2, "efmn");
  }
  ops.contract_p(pr, -2.0, u2, "ebij", T4d(F.vT.d(), o, v, o, o), "mbij", 1.0, r1, "em");
  double fac2 = dc ? 0.5 : 1.0;
  {
    // pR2 += U2[a,f,m,n] (f^_ae - fac x_ae) - U2[e,f,i,n] (f^_mi + fac x_mi)
    double* gvv = ops.arena.take(v * v);
    double* goo = ops.arena.take(o * o);
    ops.copy(v * v, F.dfvv.d(), gvv);
    ops.axpby(v * v, -fac2, xvv, 1.0, gvv);
    ops.copy(o * o, F.dfoo.d(), goo);
    ops.axpby(o * o, fac2, xoo, 1.0, goo);
    ops.contract_p(pr, 1.0, u2, "afmn", T2d(gvv, v, v), "ae", 1.0, tpr, "efmn");
    ops.contract_p(pr, -1.0, u2, "efin", T2d(goo, o, o), "mi", 1.0, tpr, "efmn");
  }
  Tensor dvovo = T4d(F.dvovo.d(), v, o, v, o);
  ops.contract_p(pr, -1.0, u2, "afin", dvovo, "amei", 1.0, tpr, "efmn");
  ops.contract_p(pr, -1.0, u2, "ebin", dvovo, "bmfi", 1.0, tpr, "efmn");
  // R2 += pR2 + pR2[f,e,n,m]
  const int p_s[4] = {1, 0, 3, 2};
  ops.axpby(n2, 1.0, pR2, 1.0, R2d);
  ops.permute(tpr, p_s, r2, 1.0, 1.0);
  // R1 += dD[p,q] (2 <qm|pe> - <qm|ep>),  <qm|pe> = (qp|me), undressed, over all orbitals:
  //   2 sum_Q c[Q] R[m,e,Q] with c[Q] = sum_pq dD[p,q] L[q,p,Q]
  //   - sum_Q sum_q L[q,e,Q] Z[m,q,Q] with Z[m,q,Q] = sum_p R[m,p,Q] dD[p,q]
  {
    double* cQ = ops.arena.take(Q);
    Tensor tc = T1v(cQ, Q);
    Tensor dd_oo = T2d(D.doo.d(), o, o), dd_ov = T2d(D.dov.d(), o, v), dd_vo = T2d(D.dvo.d(), v, o),
           dd_vv = T2d(D.dvv.d(), v, v);
    ops.contract_p(Prec::F64, 1.0, dd_oo, "pq", T3d(c.L.oo.d(), o, o, Q), "qpQ", 0.0, tc, "Q");
    ops.contract_p(Prec::F64, 1.0, dd_ov, "pq", T3d(c.L.vo.d(), v, o, Q), "qpQ", 1.0, tc, "Q");
    ops.contract_p(Prec::F64, 1.0, dd_vo, "pq", T3d(c.L.ov.d(), o, v, Q), "qpQ", 1.0, tc, "Q");
    ops.contract_p(Prec::F64, 1.0, dd_vv, "pq", T3d(c.L.vv.d(), v, v, Q), "qpQ", 1.0, tc, "Q");
    ops.contract_p(Prec::F64, 2.0, T3d(Rb.ov.d(), o, v, Q), "meQ", tc, "Q", 1.0, r1, "em");
    double* Zo = ops.arena.take(o * o * Q);
    double* Zv = ops.arena.take(o * v * Q);
    Tensor tzo = T3d(Zo, o, o, Q), tzv = T3d(Zv, o, v, Q);
    // Z[m,q,Q] for q occupied (Zo) and virtual (Zv)
    ops.contract_p(Prec::F64, 1.0, T3d(Rb.oo.d(), o, o, Q), "mpQ", dd_oo, "pq", 0.0, tzo, "mqQ");
    ops.contract_p(Prec::F64, 1.0, T3d(Rb.ov.d(), o, v, Q), "mpQ", dd_vo, "pq", 1.0, tzo, "mqQ");
    ops.contract_p(Prec::F64, 1.0, T3d(Rb.oo.d(), o, o, Q), "mpQ", dd_ov, "pq", 0.0, tzv, "mqQ");
    ops.contract_p(Prec::F64, 1.0, T3d(Rb.ov.d(), o, v, Q), "mpQ", dd_vv, "pq", 1.0, tzv, "mqQ");
    ops.contract_p(Prec::F64, -1.0, T3d(c.L.ov.d(), o, v, Q), "qeQ", tzo, "mqQ", 1.0, r1, "em");
    ops.contract_p(Prec::F64, -1.0, T3d(c.L.vv.d(), v, v, Q), "qeQ", tzv, "mqQ", 1.0, r1, "em");
  }
  ops.contract_p(pr, -2.0, u2, "ebij", T4d(F.dvooo.d(), v, o, o, o), "bmji", 1.0, r1, "em");
  ops.contract_p(Prec::F64, 1.0, D1oo, "mk", dfov, "ke", 1.0, r1, "em");
  ops.contract_p(Prec::F64, -1.0, D1vv, "de", dfov, "md", 1.0, r1, "em");
  {
    double* gvv = ops.arena.take(v * v);
    double* goo = ops.arena.take(o * o);
    ops.copy(v * v, F.dfvv.d(), gvv);
    ops.axpby(v * v, -1.0, xvv, 1.0, gvv);
    ops.copy(o * o, F.dfoo.d(), goo);
    ops.axpby(o * o, 1.0, xoo, 1.0, goo);
    ops.contract_p(Prec::F64, 1.0, u1, "am", T2d(gvv, v, v), "ae", 1.0, r1, "em");
    ops.contract_p(Prec::F64, -1.0, u1, "ei", T2d(goo, o, o), "mi", 1.0, r1, "em");
  }
  ops.arena.reset();
}

// assemble an n x n host matrix from four device blocks
void gather_nn(Context& c, const DevBuf& oo, const DevBuf& ov, const DevBuf& vo, const DevBuf& vv, double* out) {
  const int o = c.no, v = c.nv, n = c.n;
  std::vector<double> b;
  auto put = [&](const DevBuf& src, int r0, int m, int c0, int k) {
    b.resize((size_t)m * k);
    ECG_HIP(hipMemcpyAsync(b.data(), src.p, sizeof(double) * m * k, hipMemcpyDeviceToHost, c.stream));
    ECG_HIP(hipStreamSynchronize(c.stream));
    for (int col = 0; col < k; ++col)
      for (int r = 0; r < m; ++r) out[(size_t)(c0 + col) * n + r0 + r] = b[(size_t)col * m + r];
  };
  put(oo, 0, o, 0, o);
  put(ov, 0, o, o, v);
  put(vo, o, v, 0, o);
  put(vv, o, v, o, v);
}

}  // namespace

void lambda_residual_host(Context& c, int method, Prec pr, const double* T1h, const double* T2h, const double* U1h,
                          const double* U2h, double* R1h, double* R2h) {
  LambdaFixed F;
  prepare(c, T1h, T2h, F);
  const long n1 = c.nvo(), n2 = c.nvvoo();
  DevBuf U1(sizeof(double) * n1), U2(sizeof(double) * n2), R1(sizeof(double) * n1), R2(sizeof(double) * n2);
  ECG_HIP(hipMemcpyAsync(U1.p, U1h, U1.bytes, hipMemcpyHostToDevice, c.stream));
  ECG_HIP(hipMemcpyAsync(U2.p, U2h, U2.bytes, hipMemcpyHostToDevice, c.stream));
  residual(c, F, method, pr, U1.d(), U2.d(), R1.d(), R2.d());
  ECG_HIP(hipMemcpyAsync(R1h, R1.p, R1.bytes, hipMemcpyDeviceToHost, c.stream));
  ECG_HIP(hipMemcpyAsync(R2h, R2.p, R2.bytes, hipMemcpyDeviceToHost, c.stream));
  ECG_HIP(hipStreamSynchronize(c.stream));
}

void one_rdm_host(Context& c, const double* T1h, const double* T2h, const double* U1h, const double* U2h, double* D1h,
                  double* dD1h) {
  const long n1 = c.nvo(), n2 = c.nvvoo();
  DevBuf T1(sizeof(double) * n1), T2(sizeof(double) * n2), U1(sizeof(double) * n1), U2(sizeof(double) * n2);
  ECG_HIP(hipMemcpyAsync(T1.p, T1h, T1.bytes, hipMemcpyHostToDevice, c.stream));
  ECG_HIP(hipMemcpyAsync(T2.p, T2h, T2.bytes, hipMemcpyHostToDevice, c.stream));
  ECG_HIP(hipMemcpyAsync(U1.p, U1h, U1.bytes, hipMemcpyHostToDevice, c.stream));
  ECG_HIP(hipMemcpyAsync(U2.p, U2h, U2.bytes, hipMemcpyHostToDevice, c.stream));
  c.ops.arena.reset();
  Rdm D;
  one_rdm(c, T1.d(), T2.d(), U1.d(), U2.d(), D);
  if (D1h) gather_nn(c, D.oo, D.ov, D.vo, D.vv, D1h);
  if (dD1h) gather_nn(c, D.doo, D.dov, D.dvo, D.dvv, dD1h);
}

// lm_cc_iterations! closed-shell (cc_lagrange.jl:1155-1223).  pr != F64: defect
// correction as in Context::solve -- iterations use R_low(U) + delta with delta =
// R_64 - R_low refreshed at the current U every refine_every-th iteration, when
// the corrected norm drops below 100*thr and when it stalls; convergence is
// declared only on such an exact iteration; after two refreshes in a row whose
// next low-precision iteration already stalls (the rounding floor of the
// corrected residual), the remaining iterations are plain FP64 ones with a fresh
// DIIS history.
void lambda_solve(Context& c, int method, Prec pr, const SolverOpts& o, const double* T1h, const double* T2h,
                  double* U1h, double* U2h, int* iterations, int* converged) {
  LambdaFixed F;
  prepare(c, T1h, T2h, F);
  const long n1 = c.nvo(), n2 = c.nvvoo(), v = c.nv, no = c.no;
  TensorOps& ops = c.ops;
  DevBuf U1(sizeof(double) * n1), U2(sizeof(double) * n2), R1(sizeof(double) * n1), R2(sizeof(double) * n2);
  // transform_amplitudes2lagrange_multipliers! / add_singles2doubles! (cctools.jl:887-928):
  // tau = T2 + T1 T1; U2 = 2 tau - tau[a,b,j,i]; U1 = 2 T1
  {
    ops.arena.reset();
    double* tau = ops.arena.take(n2);
    Tensor ttau = T4d(tau, v, v, no, no), tu2 = T4d(U2.d(), v, v, no, no);
    Tensor t1 = T2d(F.T1.d(), v, no);
    ops.copy(n2, F.T2.d(), tau);
    ops.contract_p(Prec::F64, 1.0, t1, "ai", t1, "bj", 1.0, ttau, "abij");
    const int p_ij[4] = {0, 1, 3, 2};
    ops.axpby(n2, 2.0, tau, 0.0, U2.d());
    ops.permute(ttau, p_ij, tu2, -1.0, 1.0);
    ops.axpby(n1, 2.0, F.T1.d(), 0.0, U1.d());
    ops.arena.reset();
  }
  c.diis_reset(o.maxdiis, o.resthr);
  const double thren = std::sqrt(o.thr) * o.conven;
  double ltnorm = 0.0;
  int conv = 0, it = 0;
  // <Lambda|Psi> = U2.T2 + U2.(T1 T1) + U1.T1 (calc_correlation_norm)
  auto corr_norm = [&]() {
    double s = ops.dot_host(n2, U2.d(), F.T2.d()) + ops.dot_host(n1, U1.d(), F.T1.d());
    ops.arena.reset();
    double* tt = ops.arena.take(n2);
    Tensor ttt = T4d(tt, v, v, no, no), t1 = T2d(F.T1.d(), v, no);
    ops.contract_p(Prec::F64, 1.0, t1, "ai", t1, "bj", 0.0, ttt, "abij");
    s += ops.dot_host(n2, U2.d(), tt);
    ops.arena.reset();
    return s;
  };
  const bool mixed = (pr != Prec::F64);
  FILE* log = std::getenv("ECG_ITERATION_LOG") ? stderr : nullptr;   // per-iteration lines, opt-in
  DevBuf D1, D2;   // FP64 correction of the low-precision residual
  bool have_delta = false, refresh_next = false, fp64_only = false, after_refresh = false;
  int quick_stalls = 0;
  double best = HUGE_VAL;
  int exact_count = 0;
  for (it = 1; it <= o.maxit; ++it) {
    const bool refresh =
        mixed && !fp64_only && (refresh_next || (o.refine_every > 0 && (it % o.refine_every) == 0));
    const bool exact = !mixed || refresh || fp64_only;
    if (mixed && !fp64_only && refresh) {
      // exact residual first; R_low (for delta = R_64 - R_low) only when R_64 does not yet
      // meet thr, as in Context::state_step
      residual(c, F, method, Prec::F64, U1.d(), U2.d(), R1.d(), R2.d());
      const double nr = (2.0 / 3.0) * c.doubles_dot(R2.d(), R2.d(), -0.5) + 0.5 * ops.dot_host(n1, R1.d(), R1.d());
      if (nr >= o.thr) {
        D1.ensure(sizeof(double) * n1);
        D2.ensure(sizeof(double) * n2);
        residual(c, F, method, pr, U1.d(), U2.d(), D1.d(), D2.d());
        ops.axpby(n1, 1.0, R1.d(), -1.0, D1.d());
        ops.axpby(n2, 1.0, R2.d(), -1.0, D2.d());
        have_delta = true;
      }
    } else {
      residual(c, F, method, fp64_only ? Prec::F64 : pr, U1.d(), U2.d(), R1.d(), R2.d());
      if (mixed && !fp64_only && have_delta) {
        ops.axpby(n1, 1.0, D1.d(), 1.0, R1.d());
        ops.axpby(n2, 1.0, D2.d(), 1.0, R2.d());
      }
    }
    if (exact) exact_count++;
    // update_doubles!/update_singles! with the CC denominators and shifts
    c.update(U1.d(), U2.d(), R1.d(), R2.d(), o.shifts, o.shiftp);
    // calc_contra_doubles_norm (cctools.jl:487-490): sum (2R + R_ba) R / 3; singles: R.R / 2
    double NormR2 = (2.0 / 3.0) * c.doubles_dot(R2.d(), R2.d(), -0.5);
    double NormR1 = 0.5 * ops.dot_host(n1, R1.d(), R1.d());
    c.diis_perform(U1.d(), U2.d(), R1.d(), R2.d(), 1);
    double prev = ltnorm;
    ltnorm = corr_norm();
    double NormR = NormR1 + NormR2;
    if (log) {
      std::fprintf(log, "%4d %16.10f %12.3e %12.3e %s\n", it, ltnorm, ltnorm - prev, NormR, exact ? "fp64" : "low");
      std::fflush(log);
    }
    if (exact && NormR < o.thr && std::fabs(ltnorm - prev) < thren) {
      conv = 1;
      break;
    }
    if (mixed && !fp64_only) {
      if (refresh) best = HUGE_VAL;
      const bool stalled = NormR > 0.8 * best;
      if (!refresh && after_refresh) {
        quick_stalls = stalled ? quick_stalls + 1 : 0;
        fp64_only = quick_stalls >= 2;
        // the DIIS history holds corrected residuals that do not match the exact ones
        // (DIIS keeps its largest-coefficient vectors, so they would persist): restart it
        if (fp64_only) c.diis_reset(o.maxdiis, o.resthr);
      }
      after_refresh = refresh;
      best = std::min(best, NormR);
      refresh_next = NormR < 100.0 * o.thr || stalled;
    }
  }
  c.fp64_iterations = exact_count;
  if (it > o.maxit) it = o.maxit;
  ECG_HIP(hipMemcpyAsync(U1h, U1.p, U1.bytes, hipMemcpyDeviceToHost, c.stream));
  ECG_HIP(hipMemcpyAsync(U2h, U2.p, U2.bytes, hipMemcpyDeviceToHost, c.stream));
  ECG_HIP(hipStreamSynchronize(c.stream));
  *iterations = it;
  *converged = conv;
}

}  // namespace ecg
