// Closed-shell (T) and Lambda-(T) on the device.
//
// Restates calc_pertT_closed_shell (src/cc/cc_triples.jl:39-148) and
// calc_LambdapertT_closed_shell (:159-288).  For each triplet i<=j<=k (i=j=k
// skipped; weight 2 if i<j<k, else 1):
//   K[a,b,c]  = 6 particle terms  sum_d T2 . <..|dk>   (v x v^2 x v MFMA GEMMs)
//             - 6 hole terms      sum_l T2 . <lc|jk>   (v^2 x v x o GEMMs)
//   X = (4K - 2K_acb - 2K_cba - 2K_bac + K_cab + K_bca) / (e_i+e_j+e_k-e_a-e_b-e_c)
//   E[T] += w sum K.X   (Lambda: sum KU.X with KU built from U2 and the transposed integrals)
//   IntX[:,i] += w X.<jk|bc>,  IntY[:,i] += w X.T2[b,c,j,k]  (and the j, k analogues)
// E(T) = E[T] + sum T1.IntX + sum f_ov.IntY   (Lambda: 1/2 U1.IntX + f_ov.IntY).
// Each GEMM writes K through split-stride addressing, oriented so that both
// operands stream contiguously; the combine step reads K through LDS-staged
// 8x8x8 boxes of all six index permutations and writes per-block partial sums
// that a second kernel reduces in a fixed order (deterministic, no atomics).
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
  double* part;         // [nblocks][NPART]
};

__global__ __launch_bounds__(256) void triples_combine_kernel(CombArgs p) {
  __shared__ double box[6][TT * TT * TT];
  __shared__ double red[6][TT * TT * TT];
  __shared__ double ered[4];
  const int nt = p.ntile;
  int blk = blockIdx.x;
  const int at = blk % nt, bt = (blk / nt) % nt, ct = blk / (nt * nt);
  const int a0 = at * TT, b0 = bt * TT, c0 = ct * TT;
  const long v = p.nv;
  const int base[3] = {a0, b0, c0};
  // permutation table: tile coordinate (0=a,1=b,2=c) placed at K index 1,2,3
  const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}, {2, 0, 1}, {1, 2, 0}};
  for (int q = 0; q < 6; ++q) {
    for (int e = threadIdx.x; e < TT * TT * TT; e += blockDim.x) {
      int u = e % TT, w = (e / TT) % TT, z = e / (TT * TT);
      int x1 = base[perm[q][0]] + u, x2 = base[perm[q][1]] + w, x3 = base[perm[q][2]] + z;
      bool ok = x1 < v && x2 < v && x3 < v;
      box[q][e] = ok ? p.K[(long)x1 + v * x2 + v * v * x3] : 0.0;
    }
  }
  __syncthreads();
  const double eijk = p.eo[p.i] + p.eo[p.j] + p.eo[p.k];
  const long ooff_jk = v * v * (p.j + (long)p.no * p.k);
  const long ooff_ik = v * v * (p.i + (long)p.no * p.k);
  const long ooff_ij = v * v * (p.i + (long)p.no * p.j);
  double esum = 0.0;
  for (int e = threadIdx.x; e < TT * TT * TT; e += blockDim.x) {
    int la = e % TT, lb = (e / TT) % TT, lc = e / (TT * TT);
    int a = a0 + la, b = b0 + lb, c = c0 + lc;
    double r[6] = {0, 0, 0, 0, 0, 0};
    if (a < v && b < v && c < v) {
      // local coordinates (la,lb,lc) read in the order of each permutation
      const int lo[3] = {la, lb, lc};
      double kv[6];
#pragma unroll
      for (int q = 0; q < 6; ++q)
        kv[q] = box[q][lo[perm[q][0]] + TT * lo[perm[q][1]] + TT * TT * lo[perm[q][2]]];
      double x = (4.0 * kv[0] - 2.0 * kv[1] - 2.0 * kv[2] - 2.0 * kv[3] + kv[4] + kv[5]) /
                 (eijk - p.ev[a] - p.ev[b] - p.ev[c]);
      double kk = p.KU ? p.KU[(long)a + v * b + v * v * c] : kv[0];
      esum += kk * x;
      r[0] = x * p.W[b + v * c + ooff_jk];   // IntX[a,i]
      r[1] = x * p.W[a + v * c + ooff_ik];   // IntX[b,j]
      r[2] = x * p.W[a + v * b + ooff_ij];   // IntX[c,k]
      r[3] = x * p.Y[b + v * c + ooff_jk];   // IntY[a,i]
      r[4] = x * p.Y[a + v * c + ooff_ik];   // IntY[b,j]
      r[5] = x * p.Y[a + v * b + ooff_ij];   // IntY[c,k]
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) red[q][e] = r[q];
  }
  // energy: wavefront reduction, then across the four waves
  for (int s = 32; s > 0; s >>= 1) esum += __shfl_down(esum, s, 64);
  if ((threadIdx.x & 63) == 0) ered[threadIdx.x >> 6] = esum;
  __syncthreads();
  double* out = p.part + (long)blk * NPART;
  if (threadIdx.x < 48) {
    int q = threadIdx.x / TT, l = threadIdx.x % TT;
    // vectors 0,3 are indexed by a (sum over lb,lc); 1,4 by b; 2,5 by c
    int role = q % 3;
    double s = 0.0;
    for (int x = 0; x < TT; ++x)
      for (int y = 0; y < TT; ++y) {
        int la, lb, lc;
        if (role == 0) { la = l; lb = x; lc = y; }
        else if (role == 1) { lb = l; la = x; lc = y; }
        else { lc = l; la = x; lb = y; }
        s += red[q][la + TT * lb + TT * TT * lc];
      }
    out[threadIdx.x] = s;
  }
  if (threadIdx.x == 0) out[48] = ered[0] + ered[1] + ered[2] + ered[3];
}

// Reduce the per-block partials of one triplet in a fixed order and add
// w * (them) to IntX/IntY (v x o) and to the energy accumulator.
__global__ void triples_reduce_kernel(int nv, int ntile, int i, int j, int k, double w, const double* __restrict__ part,
                                      double* IntX, double* IntY, double* eacc) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int nvec = 2 * nv;
  if (t < nvec) {
    // one thread per (IntX or IntY, x) adds the i, j and k contributions in turn:
    // with i = j or j = k two of them target the same element
    int type = t / nv, x = t % nv;
    int tx = x / TT, l = x % TT;
    for (int role = 0; role < 3; ++role) {
      int q = 3 * type + role;
      double s = 0.0;
      for (int u = 0; u < ntile; ++u)
        for (int z = 0; z < ntile; ++z) {
          int at, bt, ct;
          if (role == 0) { at = tx; bt = u; ct = z; }
          else if (role == 1) { bt = tx; at = u; ct = z; }
          else { ct = tx; at = u; bt = z; }
          long blk = at + (long)ntile * (bt + (long)ntile * ct);
          s += part[blk * NPART + q * TT + l];
        }
      int occ = role == 0 ? i : (role == 1 ? j : k);
      double* dst = (type == 0 ? IntX : IntY) + x + (long)nv * occ;
      *dst += w * s;
    }
  } else if (t == nvec) {
    long nb = (long)ntile * ntile * ntile;
    double s = 0.0;
    for (long b = 0; b < nb; ++b) s += part[b * NPART + 48];
    *eacc += w * s;
  }
}

Tensor T3d(double* p, long a, long b, long c) { return Tensor(p, {a, b, c}); }
Tensor T4d(double* p, long a, long b, long c, long d) { return Tensor(p, {a, b, c, d}); }

// Accumulate K[a,b,c] for triplet (i,j,k) from amplitudes T (v,v,o,o), particle
// integrals W4[x,y,d,m] (v,v,v,o) and hole integrals H4[l,c,m,n] (o,v,o,o).
void build_K(GemmEngine& g, Prec pr, int v, int o, int i, int j, int k, const double* T, const double* W4,
             const double* H4, double* K) {
  const long vv = (long)v * v, vvv = vv * v;
  auto Tp = [&](int p, int q) { return T + vv * (p + (long)o * q); };   // T[:,:,p,q]
  auto Wm = [&](int m) { return W4 + vvv * m; };                        // W[:,:,:,m]
  auto Hm = [&](int m, int n) { return H4 + (long)o * v * (m + (long)o * n); };  // H[:,:,m,n]
  bool first = true;
  // C(m, n) = K[...]: helpers set the split strides of the output
  auto run = [&](GemmOp& op) {
    op.alpha = 1.0;
    op.beta = first ? 0.0 : 1.0;
    first = false;
    g.run_prec(op, pr);
  };
  // ---- particle terms: K += sum_d A(m,d) B(d,n)
  {  // 1: T^{ij}[a,d] W_k[b,c,d]      m=a, n=(b,c)
    GemmOp op; op.M = v; op.N = (int)vv; op.K = v;
    op.A = Tp(i, j); op.a_s0 = 1; op.a_k = v;
    op.B = Wm(k); op.b_s0 = 1; op.b_k = vv;
    op.C = K; op.c_ms0 = 1; op.c_ns0 = v;
    run(op);
  }
  {  // 2: T^{ji}[b,d] W_k[a,c,d]      m=(a,c), n=b
    GemmOp op; op.M = (int)vv; op.N = v; op.K = v;
    op.A = Wm(k); op.a_s0 = 1; op.a_k = vv;
    op.B = Tp(j, i); op.b_s0 = 1; op.b_k = v;
    op.C = K; op.c_ms0 = 1; op.c_md0 = v; op.c_ms1 = vv; op.c_ns0 = v;
    run(op);
  }
  {  // 3: T^{ik}[a,d] W_j[c,b,d]      m=a, n=(c,b)
    GemmOp op; op.M = v; op.N = (int)vv; op.K = v;
    op.A = Tp(i, k); op.a_s0 = 1; op.a_k = v;
    op.B = Wm(j); op.b_s0 = 1; op.b_k = vv;
    op.C = K; op.c_ms0 = 1; op.c_ns0 = vv; op.c_nd0 = v; op.c_ns1 = v;
    run(op);
  }
  {  // 4: T^{ki}[c,d] W_j[a,b,d]      m=(a,b), n=c
    GemmOp op; op.M = (int)vv; op.N = v; op.K = v;
    op.A = Wm(j); op.a_s0 = 1; op.a_k = vv;
    op.B = Tp(k, i); op.b_s0 = 1; op.b_k = v;
    op.C = K; op.c_ms0 = 1; op.c_ns0 = vv;
    run(op);
  }
  {  // 5: T^{jk}[b,d] W_i[c,a,d]      m=(c,a), n=b
    GemmOp op; op.M = (int)vv; op.N = v; op.K = v;
    op.A = Wm(i); op.a_s0 = 1; op.a_k = vv;
    op.B = Tp(j, k); op.b_s0 = 1; op.b_k = v;
    op.C = K; op.c_ms0 = vv; op.c_md0 = v; op.c_ms1 = 1; op.c_ns0 = v;
    run(op);
  }
  {  // 6: T^{kj}[c,d] W_i[b,a,d]      m=(b,a), n=c
    GemmOp op; op.M = (int)vv; op.N = v; op.K = v;
    op.A = Wm(i); op.a_s0 = 1; op.a_k = vv;
    op.B = Tp(k, j); op.b_s0 = 1; op.b_k = v;
    op.C = K; op.c_ms0 = v; op.c_md0 = v; op.c_ms1 = 1; op.c_ns0 = vv;
    run(op);
  }
  // ---- hole terms: K -= sum_l A(m,l) B(l,n),  A = T2 pairs (v^2 x o), B = H slices (o x v)
  auto hole = [&](const double* A, long a_k, const double* Hs, long c_ms0, int c_md0, long c_ms1, long c_ns0) {
    GemmOp op; op.M = (int)vv; op.N = v; op.K = o;
    op.A = A; op.a_s0 = 1; op.a_k = a_k;
    op.B = Hs; op.b_s0 = o; op.b_k = 1;
    op.C = K; op.c_ms0 = c_ms0; op.c_md0 = c_md0; op.c_ms1 = c_ms1; op.c_ns0 = c_ns0;
    op.alpha = -1.0; op.beta = 1.0;
    g.run_prec(op, pr);
  };
  const long vvo = vv * o;
  // 7: T[b,a,l,i] = T[a,b,i,l] ; H[l,c,j,k]      m=(a,b), n=c
  hole(T + vv * i, vvo, Hm(j, k), 1, 1 << 30, 0, vv);
  // 8: T[a,b,l,j] ; H[l,c,i,k]                    m=(a,b), n=c
  hole(T + vvo * j, vv, Hm(i, k), 1, 1 << 30, 0, vv);
  // 9: T[c,a,l,i] = T[a,c,i,l] ; H[l,b,k,j]      m=(a,c), n=b
  hole(T + vv * i, vvo, Hm(k, j), 1, v, vv, v);
  // 10: T[a,c,l,k] ; H[l,b,i,j]                  m=(a,c), n=b
  hole(T + vvo * k, vv, Hm(i, j), 1, v, vv, v);
  // 11: T[c,b,l,j] = T[b,c,j,l] ; H[l,a,k,i]     m=(b,c), n=a
  hole(T + vv * j, vvo, Hm(k, i), v, 1 << 30, 0, 1);
  // 12: T[b,c,l,k] ; H[l,a,j,i]                  m=(b,c), n=a
  hole(T + vvo * k, vv, Hm(j, i), v, 1 << 30, 0, 1);
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
  // Lambda: covariant U2 = (2U + U[b,a,i,j])/3 (cctools.jl:935-938) and the
  // integrals <ck|ab> as [a,b,c,k] = (ca|kb), <jk|lc> as [l,c,j,k] = (jl|kc)
  DevBuf dU1, dU2, vvvoL, ovooL;
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
    if (c.hermitian) {
      // (ca|kb) = (ac|bk) and (jl|kc) = (lj|ck) for Hermitian real integrals
    } else {
      vvvoL.alloc(vvvo.bytes);
      ovooL.alloc(ovoo.bytes);
      Tensor tw = T4d(vvvoL.d(), v, v, v, o), th = T4d(ovooL.d(), o, v, o, o);
      ops.arena.reset();
      ops.contract_p(Prec::F64, 1.0, T3d(c.L.vv.d(), v, v, c.nq), "caQ", T3d(Rb.ov.d(), o, v, c.nq), "kbQ", 0.0, tw,
                     "abck");
      ops.contract_p(Prec::F64, 1.0, T3d(c.L.oo.d(), o, o, c.nq), "jlQ", T3d(Rb.ov.d(), o, v, c.nq), "kcQ", 0.0, th,
                     "lcjk");
    }
  }
  const double* WL = lambda && !c.hermitian ? vvvoL.d() : vvvo.d();
  const double* HL = lambda && !c.hermitian ? ovooL.d() : ovoo.d();
  const int ntile = (v + TT - 1) / TT;
  const long nblocks = (long)ntile * ntile * ntile;
  DevBuf K(sizeof(double) * vvv), KU, part(sizeof(double) * nblocks * NPART);
  if (lambda) KU.alloc(sizeof(double) * vvv);
  DevBuf IntX(sizeof(double) * n1), IntY(sizeof(double) * n1), eacc(sizeof(double));
  ops.fill(n1, 0.0, IntX.d());
  ops.fill(n1, 0.0, IntY.d());
  ops.fill(1, 0.0, eacc.d());
  CombArgs ca{};
  ca.nv = v; ca.no = o; ca.ntile = ntile;
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
        build_K(ops.gemm, pr, v, o, i, j, k, dT2.d(), vvvo.d(), ovoo.d(), K.d());
        if (lambda) build_K(ops.gemm, pr, v, o, i, j, k, dU2.d(), WL, HL, KU.d());
        ca.i = i; ca.j = j; ca.k = k;
        hipLaunchKernelGGL(triples_combine_kernel, dim3((unsigned)nblocks), dim3(256), 0, s, ca);
        hipLaunchKernelGGL(triples_reduce_kernel, dim3((2 * v + 1 + 255) / 256), dim3(256), 0, s, v, ntile, i, j, k,
                           w, part.d(), IntX.d(), IntY.d(), eacc.d());
        ECG_HIP(hipGetLastError());
      }
  // E(T) = E[T] + T1.IntX + f_ov.IntY  (Lambda: 1/2 U1.IntX)
  std::vector<double> hx(n1), hy(n1), hf((size_t)c.n * c.n);
  double eb = 0.0;
  ECG_HIP(hipMemcpyAsync(hx.data(), IntX.p, IntX.bytes, hipMemcpyDeviceToHost, s));
  ECG_HIP(hipMemcpyAsync(hy.data(), IntY.p, IntY.bytes, hipMemcpyDeviceToHost, s));
  ECG_HIP(hipMemcpyAsync(&eb, eacc.p, sizeof(double), hipMemcpyDeviceToHost, s));
  ECG_HIP(hipMemcpyAsync(hf.data(), c.fock.p, hf.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  ECG_HIP(hipStreamSynchronize(s));
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
