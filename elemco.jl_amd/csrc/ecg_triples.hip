// Closed-shell (T) and Lambda-(T) on the device.
//
// Restates calc_pertT_closed_shell (src/cc/cc_triples.jl:39-148) and
// calc_LambdapertT_closed_shell (:159-288).  For each triplet i<=j<=k (i=j=k
// skipped; weight 2 if i<j<k, else 1):
//   K[a,b,c]  = 6 particle terms  sum_d T2 . <..|dk>   (v x v^2 x v products)
//             - 6 hole terms      sum_l T2 . <lc|jk>   (v^2 x v x o products)
//   X = (4K - 2K_acb - 2K_cba - 2K_bac + K_cab + K_bca) / (e_i+e_j+e_k-e_a-e_b-e_c)
//   E[T] += w sum K.X   (Lambda: sum KU.X with KU built from U2 and the transposed integrals)
//   IntX[:,i] += w X.<jk|bc>,  IntY[:,i] += w X.T2[b,c,j,k]  (and the j, k analogues)
// E(T) = E[T] + sum T1.IntX + sum f_ov.IntY   (Lambda: 1/2 U1.IntX + f_ov.IntY).
// The twelve products fall into three classes by the index of K that the
// amplitude factor carries (a, b or c).  Each class is one segmented MFMA GEMM
// launch (four K segments: two particle terms of length v, two hole terms of
// length o) whose output orientation lets every operand stream contiguously;
// a transposed copy of the particle integrals and a negated transposed copy of
// the hole integrals make that possible.  K is therefore written three times
// per triplet instead of twelve.  The combine step stages, for each sorted
// triple of 8-wide virtual tiles, the six 8x8x8 boxes of K at the permutations
// of that triple in LDS (K is read once), evaluates every output block that is
// a permutation of the triple, and writes per-block partial sums that a second
// kernel reduces in a fixed order (deterministic, no atomics).
#include "ecg_triples.h"

#include <algorithm>
#include <cmath>

namespace ecg {

namespace {

constexpr int TT = 8;                       // combine tile edge
constexpr int NPART = 6 * TT + 1;           // per-block partial record: 6 vectors x 8 + energy

struct CombArgs {
  int nv, no, i, j, k, ntile;
  const double* K;      // v^3, [a,b,c]
  const double* KU;     // Lambda: second intermediate, else null
  const double* eo;     // occupied orbital energies
  const double* ev;     // virtual orbital energies
  const double* W;      // <ij|ab> as [a,b,i,j]
  const double* Y;      // T2 (or covariant U2) [a,b,i,j] for IntY
  const int* tri;       // sorted tile triples (A <= B <= C), 3 ints each, one per workgroup
  double* part;         // [ntile^3][NPART], record of output block at + nt*(bt + nt*ct)
};

// One workgroup per sorted tile triple (A,B,C).  The six boxes K[perm(A,B,C)]
// (8x8x8 each) are staged in LDS once and serve every distinct output block
// (at,bt,ct) that is a permutation of (A,B,C): all six index permutations of
// K that the combine of such a block reads lie in those boxes, so K is read
// once per triplet instead of six times.
__global__ __launch_bounds__(256) void triples_combine_kernel(CombArgs p) {
  __shared__ double box[6][TT * TT * TT];
  __shared__ double wsh[6][TT * TT];    // <jk|bc>-type slices (0-2) and Y slices (3-5) of the block
  __shared__ double esh[3][TT];         // virtual orbital energies of the block's a, b, c tiles
  __shared__ double pab[6][4][TT];      // per-wave partials of the a- and b-indexed vectors
  __shared__ double pcc[6][TT];         // c-indexed vectors
  __shared__ double ered[4];
  const int nt = p.ntile;
  const long v = p.nv;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int la = lane & 7, lb = lane >> 3;   // element e = la + 8 lb + 64 lc, lc = w and w + 4
  const int T[3] = {p.tri[3 * blockIdx.x], p.tri[3 * blockIdx.x + 1], p.tri[3 * blockIdx.x + 2]};
  // permutation table: which of the three coordinates sits at K index 1,2,3
  const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}, {2, 0, 1}, {1, 2, 0}};
  // slot s holds K at tile coordinates (T[perm[s][0]], T[perm[s][1]], T[perm[s][2]])
  for (int s = 0; s < 6; ++s) {
    const int b1 = T[perm[s][0]] * TT, b2 = T[perm[s][1]] * TT, b3 = T[perm[s][2]] * TT;
    for (int e = t; e < TT * TT * TT; e += blockDim.x) {
      int x1 = b1 + e % TT, x2 = b2 + (e / TT) % TT, x3 = b3 + e / (TT * TT);
      bool ok = x1 < v && x2 < v && x3 < v;
      box[s][e] = ok ? p.K[(long)x1 + v * x2 + v * v * x3] : 0.0;
    }
  }
  const double eijk = p.eo[p.i] + p.eo[p.j] + p.eo[p.k];
  const long ooff[3] = {v * v * (p.j + (long)p.no * p.k), v * v * (p.i + (long)p.no * p.k),
                        v * v * (p.i + (long)p.no * p.j)};
  for (int o = 0; o < 6; ++o) {
    // output block tile coordinates; skip permutations equal to an earlier one (ties)
    const int tg[3] = {T[perm[o][0]], T[perm[o][1]], T[perm[o][2]]};
    bool dup = false;
    for (int o2 = 0; o2 < o; ++o2)
      dup = dup || (T[perm[o2][0]] == tg[0] && T[perm[o2][1]] == tg[1] && T[perm[o2][2]] == tg[2]);
    if (dup) continue;
    const int base[3] = {tg[0] * TT, tg[1] * TT, tg[2] * TT};
    // stage the 8x8 slices of W and Y this block multiplies X with, and the orbital energies:
    // slice q%3 = 0 over (b,c) at (j,k), 1 over (a,c) at (i,k), 2 over (a,b) at (i,j)
    for (int u = t; u < 6 * TT * TT; u += blockDim.x) {
      const int q = u / (TT * TT), x = u % TT, y = (u / TT) % TT;
      const int role = q % 3;
      const int d1 = role == 0 ? 1 : 0, d2 = role == 2 ? 1 : 2;   // the two coordinates of the slice
      const int c1 = base[d1] + x, c2 = base[d2] + y;
      const double* src = q < 3 ? p.W : p.Y;
      const bool ok = c1 < v && c2 < v;
      wsh[q][x + TT * y] = ok ? src[c1 + v * c2 + ooff[role]] : 0.0;
    }
    if (t < 3 * TT) {
      const int d = t / TT, x = base[d] + t % TT;
      esh[d][t % TT] = x < v ? p.ev[x] : 0.0;
    }
    // the box holding permutation q of this block's coordinates
    int slot[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int w0 = tg[perm[q][0]], w1 = tg[perm[q][1]], w2 = tg[perm[q][2]];
      int sl = 0;
      for (int s = 5; s >= 0; --s)
        if (T[perm[s][0]] == w0 && T[perm[s][1]] == w1 && T[perm[s][2]] == w2) sl = s;
      slot[q] = sl;
    }
    __syncthreads();   // boxes (first block) and slices staged
    double esum = 0.0;
    double rs[2][6];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int lc = w + 4 * h;
      const int a = base[0] + la, b = base[1] + lb, c = base[2] + lc;
#pragma unroll
      for (int q = 0; q < 6; ++q) rs[h][q] = 0.0;
      if (a < v && b < v && c < v) {
        // local coordinates (la,lb,lc) read in the order of each permutation
        const int lo[3] = {la, lb, lc};
        double kv[6];
#pragma unroll
        for (int q = 0; q < 6; ++q)
          kv[q] = box[slot[q]][lo[perm[q][0]] + TT * lo[perm[q][1]] + TT * TT * lo[perm[q][2]]];
        const double x = (4.0 * kv[0] - 2.0 * kv[1] - 2.0 * kv[2] - 2.0 * kv[3] + kv[4] + kv[5]) /
                         (eijk - esh[0][la] - esh[1][lb] - esh[2][lc]);
        const double kk = p.KU ? p.KU[(long)a + v * b + v * v * c] : kv[0];
        esum += kk * x;
        rs[h][0] = x * wsh[0][lb + TT * lc];   // IntX[a,i]
        rs[h][1] = x * wsh[1][la + TT * lc];   // IntX[b,j]
        rs[h][2] = x * wsh[2][la + TT * lb];   // IntX[c,k]
        rs[h][3] = x * wsh[3][lb + TT * lc];   // IntY[a,i]
        rs[h][4] = x * wsh[4][la + TT * lc];   // IntY[b,j]
        rs[h][5] = x * wsh[5][la + TT * lb];   // IntY[c,k]
      }
    }
    // fixed-order reductions: a-indexed vectors over (lb, lc), b-indexed over (la, lc),
    // c-indexed over (la, lb); lanes by xor trees, waves through LDS
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int role = q % 3;
      if (role == 0) {
        double s = rs[0][q] + rs[1][q];
        s += __shfl_xor(s, 8, 64);
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        if (lb == 0) pab[q][w][la] = s;
      } else if (role == 1) {
        double s = rs[0][q] + rs[1][q];
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        s += __shfl_xor(s, 4, 64);
        if (la == 0) pab[q][w][lb] = s;
      } else {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          double s = rs[h][q];
          for (int m = 1; m < 64; m <<= 1) s += __shfl_xor(s, m, 64);
          if (lane == 0) pcc[q][w + 4 * h] = s;
        }
      }
    }
    for (int m = 1; m < 64; m <<= 1) esum += __shfl_xor(esum, m, 64);
    if (lane == 0) ered[w] = esum;
    __syncthreads();
    double* out = p.part + ((long)tg[0] + (long)nt * (tg[1] + (long)nt * tg[2])) * NPART;
    if (t < 6 * TT) {
      const int q = t / TT, l = t % TT;
      out[t] = (q % 3 == 2) ? pcc[q][l] : ((pab[q][0][l] + pab[q][1][l]) + (pab[q][2][l] + pab[q][3][l]));
    }
    if (t == 0) out[48] = (ered[0] + ered[1]) + (ered[2] + ered[3]);
    __syncthreads();   // slices and partial buffers are reused by the next output block
  }
}

// Reduce the per-block partials of one triplet in a fixed order.
// Workgroups g < 2*ntile: vector type = g / ntile (0: IntX, 1: IntY) and virtual
// tile tx = g % ntile; the workgroup adds the i, j and k contributions (roles
// 0, 1, 2) in turn, so with i = j or j = k the two updates of one element are
// ordered.  Workgroups g >= 2*ntile: the energy partials of the c-tile slab
// ct = g - 2*ntile, added to eacc[ct].  Every element has one writer per
// launch and launches are stream-ordered, so the result is deterministic.
constexpr int RED_T = 256;
__global__ __launch_bounds__(RED_T) void triples_reduce_kernel(int nv, int ntile, int i, int j, int k, double w,
                                                               const double* __restrict__ part, double* IntX,
                                                               double* IntY, double* eacc) {
  __shared__ double red[RED_T];
  const int g = blockIdx.x, t = threadIdx.x;
  const long nt2 = (long)ntile * ntile;
  if (g < 2 * ntile) {
    const int type = g / ntile, tx = g % ntile;
    const int l = t & (TT - 1), s8 = t / TT;      // 32 strands per vector element
    constexpr int NS = RED_T / TT;
    for (int role = 0; role < 3; ++role) {
      const int q = 3 * type + role;
      double s = 0.0;
      for (long u = s8; u < nt2; u += NS) {
        const int u1 = (int)(u % ntile), u2 = (int)(u / ntile);
        int at, bt, ct;
        if (role == 0) { at = tx; bt = u1; ct = u2; }
        else if (role == 1) { bt = tx; at = u1; ct = u2; }
        else { ct = tx; at = u1; bt = u2; }
        const long blk = at + (long)ntile * (bt + (long)ntile * ct);
        s += part[blk * NPART + q * TT + l];
      }
      red[t] = s;
      __syncthreads();
      for (int h = NS / 2; h > 0; h >>= 1) {
        if (s8 < h) red[t] += red[t + h * TT];
        __syncthreads();
      }
      const int x = tx * TT + l;
      if (s8 == 0 && x < nv) {
        const int occ = role == 0 ? i : (role == 1 ? j : k);
        double* dst = (type == 0 ? IntX : IntY) + x + (long)nv * occ;
        *dst += w * red[t];
      }
      __syncthreads();
    }
  } else {
    const int ct = g - 2 * ntile;
    const double* src = part + (long)ct * nt2 * NPART + 48;
    double s = 0.0;
    for (long u = t; u < nt2; u += RED_T) s += src[u * NPART];
    red[t] = s;
    __syncthreads();
    for (int h = RED_T / 2; h > 0; h >>= 1) {
      if (t < h) red[t] += red[t + h];
      __syncthreads();
    }
    if (t == 0) eacc[ct] += w * red[0];
  }
}

Tensor T3d(double* p, long a, long b, long c) { return Tensor(p, {a, b, c}); }
Tensor T4d(double* p, long a, long b, long c, long d) { return Tensor(p, {a, b, c, d}); }

// Integrals of one (T) evaluation: W[x,y,d,m] (v,v,v,o), its transpose
// Wt[x,y,d,m] = W[y,x,d,m], and Hn[c,l,m,n] = -H[l,c,m,n] (v,o,o,o) for the
// hole integrals H[l,c,m,n].
struct TripleInts {
  const double* W;
  const double* Wt;
  const double* Hn;
};

// K[a,b,c] for triplet (i,j,k) from amplitudes T (v,v,o,o): three segmented
// GEMM launches, one per output orientation (cc_triples.jl:83-111 lists the
// twelve products).  Term numbers below follow the order of the particle terms
// (1-6) and hole terms (7-12) of that loop.
void build_K(GemmEngine& g, Prec pr, int v, int o, int i, int j, int k, const double* T, const TripleInts& I,
             double* K) {
  const long vv = (long)v * v, vvv = vv * v, vo = (long)v * o, vvo = vv * o;
  auto Tp = [&](int p, int q) { return T + vv * (p + (long)o * q); };          // T[:,:,p,q]
  auto Wm = [&](int m) { return I.W + vvv * m; };                               // W[:,:,:,m]
  auto Wtm = [&](int m) { return I.Wt + vvv * m; };                             // Wt[:,:,:,m]
  auto Hn = [&](int m, int n) { return I.Hn + vo * (m + (long)o * n); };      // Hn[:,:,m,n]
  auto seg = [](GemmSeg& s, const double* A, long a_k, const double* B, long b_k, int K) {
    s.A = A; s.a_k = a_k; s.B = B; s.b_k = b_k; s.K = K;
  };
  // ---- class a: K[a,(b,c)], m = a, n = (b,c) contiguous
  {
    GemmOp op; op.M = v; op.N = (int)vv; op.nseg = 4;
    op.a_s0 = 1; op.b_s0 = 1;
    seg(op.seg[0], Tp(i, j), v, Wm(k), vv, v);        // 1: T^{ij}[a,d] W_k[b,c,d]
    seg(op.seg[1], Tp(i, k), v, Wtm(j), vv, v);       // 3: T^{ik}[a,d] W_j[c,b,d]
    seg(op.seg[2], Hn(k, i), v, T + vv * j, vvo, o);  // 11: -H[l,a,k,i] T[b,c,j,l]
    seg(op.seg[3], Hn(j, i), v, T + vvo * k, vv, o);  // 12: -H[l,a,j,i] T[b,c,l,k]
    op.C = K; op.c_ms0 = 1; op.c_ns0 = v;
    op.alpha = 1.0; op.beta = 0.0;
    g.run_prec(op, pr);
  }
  // ---- class b: m = (a,c) contiguous in the operands, written at a + vv*c; n = b
  {
    GemmOp op; op.M = (int)vv; op.N = v; op.nseg = 4;
    op.a_s0 = 1; op.b_s0 = 1;
    seg(op.seg[0], Wm(k), vv, Tp(j, i), v, v);        // 2: W_k[a,c,d] T^{ji}[b,d]
    seg(op.seg[1], Wtm(i), vv, Tp(j, k), v, v);       // 5: W_i[c,a,d] T^{jk}[b,d]
    seg(op.seg[2], T + vv * i, vvo, Hn(k, j), v, o);  // 9: T[a,c,i,l] (-H[l,b,k,j])
    seg(op.seg[3], T + vvo * k, vv, Hn(i, j), v, o);  // 10: T[a,c,l,k] (-H[l,b,i,j])
    op.C = K; op.c_ms0 = 1; op.c_md0 = v; op.c_ms1 = vv; op.c_ns0 = v;
    op.alpha = 1.0; op.beta = 1.0;
    g.run_prec(op, pr);
  }
  // ---- class c: m = (a,b) contiguous, n = c
  {
    GemmOp op; op.M = (int)vv; op.N = v; op.nseg = 4;
    op.a_s0 = 1; op.b_s0 = 1;
    seg(op.seg[0], Wm(j), vv, Tp(k, i), v, v);        // 4: W_j[a,b,d] T^{ki}[c,d]
    seg(op.seg[1], Wtm(i), vv, Tp(k, j), v, v);       // 6: W_i[b,a,d] T^{kj}[c,d]
    seg(op.seg[2], T + vv * i, vvo, Hn(j, k), v, o);  // 7: T[a,b,i,l] (-H[l,c,j,k])
    seg(op.seg[3], T + vvo * j, vv, Hn(i, k), v, o);  // 8: T[a,b,l,j] (-H[l,c,i,k])
    op.C = K; op.c_ms0 = 1; op.c_ns0 = vv;
    op.alpha = 1.0; op.beta = 1.0;
    g.run_prec(op, pr);
  }
}

// Transposed and negated-transposed integral copies for build_K.
void make_triple_ints(TensorOps& ops, int v, int o, const double* W, const double* H, DevBuf& Wt, DevBuf& Hn) {
  const int p_yx[4] = {1, 0, 2, 3};
  Wt.alloc(sizeof(double) * (long)v * v * v * o);
  Hn.alloc(sizeof(double) * (long)v * o * o * o);
  Tensor tw = T4d(const_cast<double*>(W), v, v, v, o), twt = T4d(Wt.d(), v, v, v, o);
  ops.permute(tw, p_yx, twt, 1.0, 0.0);
  Tensor th = T4d(const_cast<double*>(H), o, v, o, o), thn = T4d(Hn.d(), v, o, o, o);
  ops.permute(th, p_yx, thn, -1.0, 0.0);
}

}  // namespace

void pert_triples(Context& c, bool lambda, Prec pr, const double* T1h, const double* T2h, const double* U1h,
                  const double* U2h, long first, long count, double* e) {
  const int v = c.nv, o = c.no;
  const long n1 = c.nvo(), n2 = c.nvvoo();
  const long vvv = (long)v * v * v;
  hipStream_t s = c.stream;
  TensorOps& ops = c.ops;
  ops.gemm.stream = s;
  const Blocks& Rb = c.hermitian ? c.L : c.R;
  DevBuf dT1(sizeof(double) * n1), dT2(sizeof(double) * n2);
  ECG_HIP(hipMemcpyAsync(dT1.p, T1h, dT1.bytes, hipMemcpyHostToDevice, s));
  ECG_HIP(hipMemcpyAsync(dT2.p, T2h, dT2.bytes, hipMemcpyHostToDevice, s));
  // undressed integrals: vvvo[a,b,c,k] = (ac|bk),  ovoo[l,c,j,k] = (lj|ck)
  DevBuf vvvo(sizeof(double) * vvv * o), ovoo(sizeof(double) * (long)o * v * o * o);
  {
    Tensor tw = T4d(vvvo.d(), v, v, v, o), th = T4d(ovoo.d(), o, v, o, o);
    ops.arena.reset();
    ops.contract_p(Prec::F64, 1.0, T3d(c.L.vv.d(), v, v, c.nq), "acQ", T3d(Rb.vo.d(), v, o, c.nq), "bkQ", 0.0, tw,
                   "abck");
    ops.contract_p(Prec::F64, 1.0, T3d(c.L.oo.d(), o, o, c.nq), "ljQ", T3d(Rb.vo.d(), v, o, c.nq), "ckQ", 0.0, th,
                   "lcjk");
  }
  DevBuf vvvoT, ovooN;
  make_triple_ints(ops, v, o, vvvo.d(), ovoo.d(), vvvoT, ovooN);
  const TripleInts IT{vvvo.d(), vvvoT.d(), ovooN.d()};
  // Lambda: covariant U2 = (2U + U[b,a,i,j])/3 (cctools.jl:935-938) and the
  // integrals <ck|ab> as [a,b,c,k] = (ca|kb), <jk|lc> as [l,c,j,k] = (jl|kc)
  DevBuf dU1, dU2, vvvoL, ovooL, vvvoLT, ovooLN;
  TripleInts IL = IT;
  if (lambda) {
    dU1.alloc(sizeof(double) * n1);
    dU2.alloc(sizeof(double) * n2);
    ECG_HIP(hipMemcpyAsync(dU1.p, U1h, dU1.bytes, hipMemcpyHostToDevice, s));
    DevBuf raw(sizeof(double) * n2);
    ECG_HIP(hipMemcpyAsync(raw.p, U2h, raw.bytes, hipMemcpyHostToDevice, s));
    Tensor tr = T4d(raw.d(), v, v, o, o), tu = T4d(dU2.d(), v, v, o, o);
    const int p_ba[4] = {1, 0, 2, 3};
    ops.axpby(n2, 2.0 / 3.0, raw.d(), 0.0, dU2.d());
    ops.permute(tr, p_ba, tu, 1.0 / 3.0, 1.0);
    if (!c.hermitian) {
      // (ca|kb) = (ac|bk) and (jl|kc) = (lj|ck) hold only for Hermitian real integrals
      vvvoL.alloc(vvvo.bytes);
      ovooL.alloc(ovoo.bytes);
      Tensor tw = T4d(vvvoL.d(), v, v, v, o), th = T4d(ovooL.d(), o, v, o, o);
      ops.arena.reset();
      ops.contract_p(Prec::F64, 1.0, T3d(c.L.vv.d(), v, v, c.nq), "caQ", T3d(Rb.ov.d(), o, v, c.nq), "kbQ", 0.0, tw,
                     "abck");
      ops.contract_p(Prec::F64, 1.0, T3d(c.L.oo.d(), o, o, c.nq), "jlQ", T3d(Rb.ov.d(), o, v, c.nq), "kcQ", 0.0, th,
                     "lcjk");
      make_triple_ints(ops, v, o, vvvoL.d(), ovooL.d(), vvvoLT, ovooLN);
      IL = TripleInts{vvvoL.d(), vvvoLT.d(), ovooLN.d()};
    }
  }
  const int ntile = (v + TT - 1) / TT;
  const long nblocks = (long)ntile * ntile * ntile;
  DevBuf K(sizeof(double) * vvv), KU, part(sizeof(double) * nblocks * NPART);
  if (lambda) KU.alloc(sizeof(double) * vvv);
  DevBuf IntX(sizeof(double) * n1), IntY(sizeof(double) * n1), eacc(sizeof(double) * ntile);
  ops.fill(n1, 0.0, IntX.d());
  ops.fill(n1, 0.0, IntY.d());
  ops.fill(ntile, 0.0, eacc.d());
  // sorted tile triples A <= B <= C: one combine workgroup each
  std::vector<int> trih;
  for (int C = 0; C < ntile; ++C)
    for (int B = 0; B <= C; ++B)
      for (int A = 0; A <= B; ++A) { trih.push_back(A); trih.push_back(B); trih.push_back(C); }
  const unsigned ntri = (unsigned)(trih.size() / 3);
  DevBuf tri(sizeof(int) * trih.size());
  ECG_HIP(hipMemcpyAsync(tri.p, trih.data(), tri.bytes, hipMemcpyHostToDevice, s));
  CombArgs ca{};
  ca.nv = v; ca.no = o; ca.ntile = ntile;
  ca.tri = static_cast<const int*>(tri.p);
  ca.K = K.d(); ca.KU = lambda ? KU.d() : nullptr;
  ca.eo = c.eps.d(); ca.ev = c.eps.d() + o;
  ca.W = c.vvoo.d();
  ca.Y = lambda ? dU2.d() : dT2.d();
  ca.part = part.d();
  // loop order of cc_triples.jl:61-71
  long tidx = -1;
  const long last = count < 0 ? (1L << 62) : first + count;
  for (int k = 0; k < o; ++k)
    for (int j = 0; j <= k; ++j)
      for (int i = 0; i <= j; ++i) {
        if (i == j && j == k) continue;
        ++tidx;
        if (tidx < first || tidx >= last) continue;
        double w = (i < j && j < k) ? 2.0 : 1.0;
        build_K(ops.gemm, pr, v, o, i, j, k, dT2.d(), IT, K.d());
        if (lambda) build_K(ops.gemm, pr, v, o, i, j, k, dU2.d(), IL, KU.d());
        ca.i = i; ca.j = j; ca.k = k;
        hipLaunchKernelGGL(triples_combine_kernel, dim3(ntri), dim3(256), 0, s, ca);
        hipLaunchKernelGGL(triples_reduce_kernel, dim3(3 * ntile), dim3(RED_T), 0, s, v, ntile, i, j, k, w,
                           part.d(), IntX.d(), IntY.d(), eacc.d());
        ECG_HIP(hipGetLastError());
      }
  // E(T) = E[T] + T1.IntX + f_ov.IntY  (Lambda: 1/2 U1.IntX)
  std::vector<double> hx(n1), hy(n1), hf((size_t)c.n * c.n), he(ntile);
  ECG_HIP(hipMemcpyAsync(hx.data(), IntX.p, IntX.bytes, hipMemcpyDeviceToHost, s));
  ECG_HIP(hipMemcpyAsync(hy.data(), IntY.p, IntY.bytes, hipMemcpyDeviceToHost, s));
  ECG_HIP(hipMemcpyAsync(he.data(), eacc.p, eacc.bytes, hipMemcpyDeviceToHost, s));
  ECG_HIP(hipMemcpyAsync(hf.data(), c.fock.p, hf.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  ECG_HIP(hipStreamSynchronize(s));
  double eb = 0.0;
  for (int t = 0; t < ntile; ++t) eb += he[t];
  const double* t1 = lambda ? U1h : T1h;
  double fac1 = lambda ? 0.5 : 1.0;
  double en = eb;
  for (int i = 0; i < o; ++i)
    for (int a = 0; a < v; ++a) {
      long ai = a + (long)v * i;
      en += fac1 * t1[ai] * hx[ai];
      en += hf[(size_t)(o + a) * c.n + i] * hy[ai];   // f[i, o+a] (occupied row, virtual column)
    }
  e[0] = en;
  e[1] = eb;
  ops.arena.reset();
}

}  // namespace ecg
