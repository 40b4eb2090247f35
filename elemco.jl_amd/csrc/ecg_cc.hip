// Closed-shell CCSD / DCSD on the device: T1 dressing of DF factors, residual,
// energies, amplitude update, DIIS and the iteration loop.
//
// Equations follow src/cc/cc.jl of ElemCo.jl v0.14.0 (calc_cc_resid, :1343-1489)
// with the integrals represented by DF factors, (pq|rs) = sum_Q L[p,q,Q] R[r,s,Q]
// (R = L for Hermitian integrals).  The T1 dressing of cc.jl:453-813 is the
// similarity transform  L^ = (1 - t) L (1 + t)  of each n x n factor (t holds
// T1 in its virtual-occupied block), so every dressed block <pq|rs>^ used by the
// reference is  sum_Q L^[p,r,Q] R^[q,s,Q].  The particle-particle ladder of
// cc.jl:1400-1425 (kext on undressed integrals, back-transformed with T1) equals
//   K^ij_ab = sum_Q L^[a,i,Q] R^[b,j,Q] + sum_Q sum_cd L^[a,c,Q] R^[b,d,Q] T[c,d,i,j]
// and its singles part  sum_k 2K^ik_ak - K^ki_ak  is assembled from the same factors.
#include "ecg_cc.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ecg {

namespace {

Tensor T2d(double* p, long a, long b) { return Tensor(p, {a, b}); }
Tensor T3d(double* p, long a, long b, long c) { return Tensor(p, {a, b, c}); }
Tensor T4d(double* p, long a, long b, long c, long d) { return Tensor(p, {a, b, c, d}); }
Tensor T1v(double* p, long a) { return Tensor(p, {a}); }
double* cp(const double* p) { return const_cast<double*>(p); }

// ---------------------------------------------------------------- kernels
// T2[a,b,i,j] += -R2/(e_a + e_b - e_i - e_j + shift)   (cctools.jl:245-258)
__global__ void update_doubles_kernel(int nv, int no, const double* __restrict__ eo, const double* __restrict__ ev,
                                      double shift, const double* __restrict__ R, double* T, int assign) {
  long n = (long)nv * nv * no * no;
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long st = (long)gridDim.x * blockDim.x;
  for (; idx < n; idx += st) {
    long r = idx;
    int a = (int)(r % nv); r /= nv;
    int b = (int)(r % nv); r /= nv;
    int i = (int)(r % no); r /= no;
    int j = (int)r;
    double d = -R[idx] / (ev[a] + ev[b] - eo[i] - eo[j] + shift);
    T[idx] = assign ? d : T[idx] + d;
  }
}

// T1[a,i] += -R1/(e_a - e_i + shift)   (cctools.jl:198-210)
__global__ void update_singles_kernel(int nv, int no, const double* __restrict__ eo, const double* __restrict__ ev,
                                      double shift, const double* __restrict__ R, long ldr, double* T, int assign) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nv * no) return;
  int a = idx % nv, i = idx / nv;
  double d = -R[a + (long)i * ldr] / (ev[a] - eo[i] + shift);
  T[idx] = assign ? d : T[idx] + d;
}

// Per-(i,j) block sums  d = sum_ab X[a,b,i,j] Y[a,b,i,j],  e = sum_ab X[b,a,i,j] Y[a,b,i,j].
// The v x v block is walked in 32 x 32 tiles: X and Y of tile (A, B) and X of the mirrored
// tile (B, A) are read along their unit-stride index (coalesced), the mirrored tile is
// transposed through LDS.  Each thread keeps 4 elements of a tile; four groups of 256
// threads walk alternate tiles (one workgroup per pair: at o = 16 a single 256-thread group
// per pair left one workgroup per CU waiting on its loads).
__global__ __launch_bounds__(1024) void cs_dots_kernel(int nv, const double* __restrict__ X,
                                                       const double* __restrict__ Y, double* part) {
  __shared__ double xt[4][32][33];   // per group: mirrored tile, [b][a] as read
  __shared__ double sd[16], se[16];
  const long off = (long)blockIdx.x * nv * nv;
  const double* x = X + off;
  const double* y = Y + off;
  // four groups of 256 threads take tiles t0 + g of the (tb, ta) walk; every thread runs the
  // same number of steps (a group past the last tile idles), so all reach every barrier
  const int g = threadIdx.x >> 8, tl = threadIdx.x & 255;
  const int tx = tl & 31, ty = tl >> 5;   // 32 x 8 threads per group
  const int nt = (nv + 31) / 32, ntt = nt * nt;
  double d = 0.0, e = 0.0;
  for (int t0 = 0; t0 < ntt; t0 += 4) {
    const int t = t0 + g;
    const bool on = t < ntt;
    const int tb = on ? t / nt : 0, ta = on ? t - tb * nt : 0;
    // mirrored tile: rows b in tile tb (unit stride), columns a in tile ta
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int b = tb * 32 + tx, a = ta * 32 + ty + 8 * k;
      xt[g][ty + 8 * k][tx] = (on && a < nv && b < nv) ? x[b + (long)a * nv] : 0.0;
    }
    __syncthreads();
    if (on) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int a = ta * 32 + tx, b = tb * 32 + ty + 8 * k;
        if (a < nv && b < nv) {
          const long ti = a + (long)b * nv;
          const double yv = y[ti];
          d += x[ti] * yv;
          e += xt[g][tx][ty + 8 * k] * yv;   // X[b, a]
        }
      }
    }
  }
  for (int s = 32; s > 0; s >>= 1) {
    d += __shfl_down(d, s, 64);
    e += __shfl_down(e, s, 64);
  }
  if ((threadIdx.x & 63) == 0) { sd[threadIdx.x >> 6] = d; se[threadIdx.x >> 6] = e; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ds = 0.0, es = 0.0;
    for (int w = 0; w < 16; ++w) { ds += sd[w]; es += se[w]; }
    part[2 * blockIdx.x] = ds;
    part[2 * blockIdx.x + 1] = es;
  }
}

// w[Q] = sum_k X[k,k,Q] over an o x o x Q block
__global__ void trace_kernel(int no, int nq, const double* __restrict__ X, double* w) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  double s = 0.0;
  const double* x = X + (long)q * no * no;
  for (int k = 0; k < no; ++k) s += x[k + (long)k * no];
  w[q] = s;
}

// diagonal of an m x m column-major block
__global__ void diag_kernel(int m, const double* __restrict__ X, double* d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) d[i] = X[i + (long)i * m];
}

unsigned nblk(long n) { return (unsigned)std::min<long>((n + 255) / 256, 65535); }

// Extract block [r0:r0+m, c0:c0+k, :] of a column-major n x n x nq host array.
void extract_block(const double* src, int n, int nq, int r0, int m, int c0, int k, std::vector<double>& out) {
  out.resize((size_t)m * k * nq);
  for (int q = 0; q < nq; ++q)
    for (int c = 0; c < k; ++c)
      std::memcpy(&out[((size_t)q * k + c) * m], src + ((size_t)q * n + (c0 + c)) * n + r0, sizeof(double) * m);
}

void upload(DevBuf& d, const std::vector<double>& h, hipStream_t s) {
  d.alloc(sizeof(double) * std::max<size_t>(h.size(), 1));
  if (!h.empty()) ECG_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s));
}

}  // namespace

// ---------------------------------------------------------------- set-up
void Context::init(int nocc, int nvirt, int naux, const double* h_host, const double* BL_host,
                   const double* BR_host, double e0_in, size_t workspace_bytes) {
  ECG_CHECK(nocc > 0 && nvirt > 0 && naux > 0, "sizes must be positive");
  no = nocc; nv = nvirt; n = no + nv; nq = naux; e0 = e0_in;
  hermitian = (BR_host == nullptr || BR_host == BL_host);
  ECG_HIP(hipSetDevice(device));
  ECG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  for (auto& e : ev) ECG_HIP(hipEventCreate(&e));
  ops.stream = stream;
  ops.gemm.stream = stream;
  std::vector<double> tmp;
  // one-electron integrals, stored n x n
  tmp.assign(h_host, h_host + (size_t)n * n);
  upload(h, tmp, stream);
  auto load_factor = [&](const double* src, Blocks& B) {
    extract_block(src, n, nq, 0, no, 0, no, tmp); upload(B.oo, tmp, stream);
    extract_block(src, n, nq, 0, no, no, nv, tmp); upload(B.ov, tmp, stream);
    extract_block(src, n, nq, no, nv, 0, no, tmp); upload(B.vo, tmp, stream);
    extract_block(src, n, nq, no, nv, no, nv, tmp); upload(B.vv, tmp, stream);
    ECG_HIP(hipStreamSynchronize(stream));
  };
  load_factor(BL_host, L);
  if (!hermitian) load_factor(BR_host, R);
  // dressed factors and integrals: allocated with the context, rewritten by every residual
  auto alloc_dressed = [&](Blocks& D) {
    D.oo.alloc(sizeof(double) * (size_t)no * no * nq);
    D.ov.alloc(sizeof(double) * (size_t)no * nv * nq);
    D.vo.alloc(sizeof(double) * (size_t)nv * no * nq);
    D.vv.alloc(sizeof(double) * (size_t)nv * nv * nq);
  };
  alloc_dressed(Ld);
  if (!hermitian) alloc_dressed(Rd);
  dh.alloc(sizeof(double) * (size_t)n * n);
  df.alloc(sizeof(double) * (size_t)n * n);
  trace.alloc(sizeof(double) * (size_t)nq);
  // workspace: contraction temporaries and ladder blocks
  long ovov = nvvoo();
  size_t want = sizeof(double) * (size_t)(18 * ovov + 4L * nv * nv * nq + 64L * 1024 * 1024);
  if (workspace_bytes) want = workspace_bytes;
  ops.arena.buf.alloc(want);
  build_fock();
}

// undressed closed-shell Fock matrix (fockfactory.jl:27-30) and the fixed
// blocks <kl|cd> (klcd of cc.jl:1390) and <ij|ab> as [a,b,i,j]
void Context::build_fock() {
  const Blocks& Rb = hermitian ? L : R;
  ops.arena.reset();
  double* w = ops.arena.take(nq);
  hipLaunchKernelGGL(trace_kernel, dim3((nq + 255) / 256), dim3(256), 0, stream, no, nq, Rb.oo.d(), w);
  fock.alloc(sizeof(double) * n * n);
  // f = h + 2 J - K, assembled block by block into an n x n column-major matrix
  double* foo = ops.arena.take((long)no * no);
  double* fov = ops.arena.take((long)no * nv);
  double* fvo = ops.arena.take((long)nv * no);
  double* fvv = ops.arena.take((long)nv * nv);
  auto jk = [&](const DevBuf& Lpq, const DevBuf& Lpo, const DevBuf& Roq, int m, int k, double* out) {
    Tensor o2 = T2d(out, m, k);
    ops.contract_p(Prec::F64, 2.0, T3d(Lpq.d(), m, k, nq), "pqQ", T1v(w, nq), "Q", 0.0, o2, "pq");
    ops.contract_p(Prec::F64, -1.0, T3d(Lpo.d(), m, no, nq), "pkQ", T3d(Roq.d(), no, k, nq), "kqQ", 1.0, o2, "pq");
  };
  jk(L.oo, L.oo, Rb.oo, no, no, foo);
  jk(L.ov, L.oo, Rb.ov, no, nv, fov);
  jk(L.vo, L.vo, Rb.oo, nv, no, fvo);
  jk(L.vv, L.vo, Rb.ov, nv, nv, fvv);
  // scatter blocks into the full matrix and add h
  std::vector<double> hb((size_t)n * n), fo((size_t)n * n), tmpb;
  ECG_HIP(hipMemcpyAsync(hb.data(), h.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, stream));
  auto get = [&](double* d, int m, int k) {
    tmpb.resize((size_t)m * k);
    ECG_HIP(hipMemcpyAsync(tmpb.data(), d, sizeof(double) * m * k, hipMemcpyDeviceToHost, stream));
    ECG_HIP(hipStreamSynchronize(stream));
  };
  auto put = [&](int r0, int c0, int m, int k) {
    for (int c = 0; c < k; ++c)
      for (int r = 0; r < m; ++r) fo[(size_t)(c0 + c) * n + r0 + r] = hb[(size_t)(c0 + c) * n + r0 + r] + tmpb[(size_t)c * m + r];
  };
  get(foo, no, no); put(0, 0, no, no);
  get(fov, no, nv); put(0, no, no, nv);
  get(fvo, nv, no); put(no, 0, nv, no);
  get(fvv, nv, nv); put(no, no, nv, nv);
  ECG_HIP(hipMemcpyAsync(fock.p, fo.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, stream));
  std::vector<double> e(n);
  for (int p = 0; p < n; ++p) e[p] = fo[(size_t)p * n + p];
  upload(eps, e, stream);
  eref_hf = e0;
  for (int i = 0; i < no; ++i) eref_hf += e[i] + hb[(size_t)i * n + i];
  // fixed integral blocks
  oovv.alloc(sizeof(double) * nvvoo());
  vvoo.alloc(sizeof(double) * nvvoo());
  Tensor t_oovv = T4d(oovv.d(), no, no, nv, nv), t_vvoo = T4d(vvoo.d(), nv, nv, no, no);
  ops.contract_p(Prec::F64, 1.0, T3d(L.ov.d(), no, nv, nq), "kcQ", T3d(Rb.ov.d(), no, nv, nq), "ldQ", 0.0,
                 t_oovv, "klcd");
  ops.contract_p(Prec::F64, 1.0, T3d(L.ov.d(), no, nv, nq), "iaQ", T3d(Rb.ov.d(), no, nv, nq), "jbQ", 0.0,
                 t_vvoo, "abij");
  if (!hermitian) {
    // <ab|ij> = (ai|bj) differs from <ij|ab> for non-Hermitian integrals
    abij_buf.alloc(sizeof(double) * nvvoo());
    Tensor t_ab = T4d(abij_buf.d(), nv, nv, no, no);
    ops.contract_p(Prec::F64, 1.0, T3d(L.vo.d(), nv, no, nq), "aiQ", T3d(Rb.vo.d(), nv, no, nq), "bjQ", 0.0,
                   t_ab, "abij");
  }
  ECG_HIP(hipStreamSynchronize(stream));
  ops.arena.reset();
}

double Context::hf_energy() { return eref_hf; }

// ---------------------------------------------------------------- dressing
// L^ = (1 - t) L (1 + t) per aux index, block form:
//   oo += ov T1 ;  vv -= T1 ov ;  vo += vv^ T1 - T1 oo ;  ov unchanged.
void Context::dress(const double* T1d) {
  auto dress_one = [&](const Blocks& S, Blocks& D) {
    long ooq = (long)no * no * nq, ovq = (long)no * nv * nq, vvq = (long)nv * nv * nq;
    D.oo.ensure(sizeof(double) * ooq);
    D.ov.ensure(sizeof(double) * ovq);
    D.vo.ensure(sizeof(double) * ovq);
    D.vv.ensure(sizeof(double) * vvq);
    ops.copy(ovq, S.ov.d(), D.ov.d());
    ops.copy(ooq, S.oo.d(), D.oo.d());
    ops.copy(vvq, S.vv.d(), D.vv.d());
    ops.copy(ovq, S.vo.d(), D.vo.d());
    Tensor t1 = T2d(cp(T1d), nv, no);
    Tensor doo = T3d(D.oo.d(), no, no, nq), dvv = T3d(D.vv.d(), nv, nv, nq), dvo = T3d(D.vo.d(), nv, no, nq);
    ops.contract_p(Prec::F64, 1.0, T3d(S.ov.d(), no, nv, nq), "ibQ", t1, "bj", 1.0, doo, "ijQ");
    ops.contract_p(Prec::F64, -1.0, t1, "ai", T3d(S.ov.d(), no, nv, nq), "ibQ", 1.0, dvv, "abQ");
    ops.contract_p(Prec::F64, 1.0, dvv, "abQ", t1, "bi", 1.0, dvo, "aiQ");
    ops.contract_p(Prec::F64, -1.0, t1, "aj", T3d(S.oo.d(), no, no, nq), "jiQ", 1.0, dvo, "aiQ");
  };
  dress_one(L, Ld);
  if (!hermitian) dress_one(R, Rd);
  const Blocks& Rh = hermitian ? Ld : Rd;
  // dressed one-electron integrals (cc.jl:783-790) and Fock matrix (cc.jl:795-812),
  // both kept as full n x n column-major matrices
  dh.ensure(sizeof(double) * n * n);
  df.ensure(sizeof(double) * n * n);
  trace.ensure(sizeof(double) * nq);
  hipLaunchKernelGGL(trace_kernel, dim3((nq + 255) / 256), dim3(256), 0, stream, no, nq, Rh.oo.d(), trace.d());
  size_t mark = ops.arena.mark();
  double* hb[4];
  int rows[4] = {no, no, nv, nv}, cols[4] = {no, nv, no, nv}, r0[4] = {0, 0, no, no}, c0[4] = {0, no, 0, no};
  for (int b = 0; b < 4; ++b) {
    hb[b] = ops.arena.take((long)rows[b] * cols[b]);
    ECG_HIP(hipMemcpy2DAsync(hb[b], sizeof(double) * rows[b], h.d() + (size_t)c0[b] * n + r0[b], sizeof(double) * n,
                             sizeof(double) * rows[b], cols[b], hipMemcpyDeviceToDevice, stream));
  }
  Tensor t1 = T2d(cp(T1d), nv, no);
  Tensor hoo = T2d(hb[0], no, no), hov = T2d(hb[1], no, nv), hvo = T2d(hb[2], nv, no), hvv = T2d(hb[3], nv, nv);
  // h^ = (1 - t) h (1 + t):  vv -= T1 ov ;  vo += vv^ T1 - T1 oo ;  oo += ov T1
  ops.contract_p(Prec::F64, -1.0, t1, "ai", hov, "ib", 1.0, hvv, "ab");
  ops.contract_p(Prec::F64, 1.0, hvv, "ab", t1, "bi", 1.0, hvo, "ai");
  ops.contract_p(Prec::F64, -1.0, t1, "aj", hoo, "ji", 1.0, hvo, "ai");
  ops.contract_p(Prec::F64, 1.0, hov, "ib", t1, "bj", 1.0, hoo, "ij");
  for (int b = 0; b < 4; ++b)
    ECG_HIP(hipMemcpy2DAsync(dh.d() + (size_t)c0[b] * n + r0[b], sizeof(double) * n, hb[b], sizeof(double) * rows[b],
                             sizeof(double) * rows[b], cols[b], hipMemcpyDeviceToDevice, stream));
  // dressed Fock: df = dh + 2 J^ - K^
  const DevBuf* Lq[4] = {&Ld.oo, &Ld.ov, &Ld.vo, &Ld.vv};
  const DevBuf* Lk[4] = {&Ld.oo, &Ld.oo, &Ld.vo, &Ld.vo};
  const DevBuf* Rk[4] = {&Rh.oo, &Rh.ov, &Rh.oo, &Rh.ov};
  for (int b = 0; b < 4; ++b) {
    Tensor o2 = T2d(hb[b], rows[b], cols[b]);
    ops.contract_p(Prec::F64, 2.0, T3d(Lq[b]->d(), rows[b], cols[b], nq), "pqQ", T1v(trace.d(), nq), "Q", 1.0, o2, "pq");
    ops.contract_p(Prec::F64, -1.0, T3d(Lk[b]->d(), rows[b], no, nq), "pkQ", T3d(Rk[b]->d(), no, cols[b], nq), "kqQ",
                   1.0, o2, "pq");
    ECG_HIP(hipMemcpy2DAsync(df.d() + (size_t)c0[b] * n + r0[b], sizeof(double) * n, hb[b], sizeof(double) * rows[b],
                             sizeof(double) * rows[b], cols[b], hipMemcpyDeviceToDevice, stream));
  }
  ops.arena.release_to(mark);
}

// ---------------------------------------------------------------- particle-particle ladder
// P[a,b,i,j] = sum_Q sum_cd L^[a,c,Q] R^[b,d,Q] T2[c,d,i,j]  is added to R2r, whose
// final symmetrisation R2 += R2r + R2r[b,a,j,i] restores the full term, because
// P[a,b,i,j] = P[b,a,j,i] ((ac|bd) = (bd|ac)).  For a block of a the integrals
// V[a',b',c,d] (b = a0 + b' >= a0) are assembled on MFMA and contracted at once;
// the v^4 tensor never exists and only b >= a0 is formed (half the work).  Inside the
// diagonal block the fused path (df_ladder_half_planes) forms the pairs a <= b once
// (weight 1/2 for a = b); the FP64-V path forms them twice, each weighted by 1/2.
void Context::ladder(Prec pr, const double* T2x, double* R2r) {
  const Blocks& Rh = hermitian ? Ld : Rd;
  if (!capturing) ECG_HIP(hipEventRecord(ev[2], stream));
  ladder_df(pr, Ld.vv.d(), Rh.vv.d(), T2x, R2r);
  if (!capturing) ECG_HIP(hipEventRecord(ev[3], stream));
}

// out[a,b,i,j] += sum_Q sum_cd Lf[a,c,Q] Rf[b,d,Q] X[c,d,i,j] for b >= (start of a's block),
// diagonal blocks weighted 1/2 (see ladder); X is supplied as Xx[i,j,c,d].
void Context::ladder_df(Prec pr, const double* Lf, const double* Rf, const double* T2x, double* R2r) {
  df_ladder_half(ops, nv, no, nq, pr, Lf, Rf, T2x, R2r);
}

// a-rows per block of the fused ladder (fewer when the arena is short).  With the diagonal
// blocks formed as triangles every pair a <= b is computed once whatever the block, so the
// size only trades launch count against product size: the c2 iteration (o=40 v=360) takes
// FP64 1708 / 1702 / 1701 ms and BF16 hi/lo 398 / 401 / 406 ms with at most 32 / 64 / 128
// rows (the arena holds about 55 rows of FP64 planes there)
constexpr long kLadderMaxRows = 64;

// R2r[a0+a',a0+b',ij] += w Y[t,ij] for the triangle rows t = b'(b'+1)/2 + a' (a' <= b' < na),
// w = 1 off the diagonal and 1/2 on it (the later symmetrisation adds the transposed half)
__global__ void ladder_tri_scatter_kernel(int na, long v, long a0, long ntri, long nij, const double* __restrict__ Y,
                                          double* R2r) {
  const long n = ntri * nij;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long t = e % ntri, ij = e / ntri;
    int b = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((long)b * (b + 1) / 2 > t) --b;
    while ((long)(b + 1) * (b + 2) / 2 <= t) ++b;
    const int a = (int)(t - (long)b * (b + 1) / 2);
    R2r[(a0 + a) + v * (a0 + b) + v * v * ij] += (a == b ? 0.5 : 1.0) * Y[e];
  }
}

// Zero the last 32-element slab of every row of a plane (rows of ld elements of esz bytes):
// the K padding past v^2, which no V product writes
__global__ void zero_plane_tail_kernel(char* plane, long rows, long ld, int esz) {
  const long nw = (long)esz * 32 / 8, n = rows * nw;   // 8-byte words per row tail
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long r = e / nw, w = e - r * nw;
    reinterpret_cast<unsigned long long*>(plane + (r * ld + ld - 32) * esz)[w] = 0ull;
  }
}

// The ladder with the V blocks written straight into the packed planes of the contraction's
// A operand (GemmEngine::run_planes with a PlaneOut): V never exists in FP64 and is not read
// back for packing.  The V products run on the factors transposed to [c, a, Q] (one pass
// each per residual), so their rows are m = c + v a' and their columns n = d + v b', and
// their stores walk c: consecutive positions c + v d of a plane row, the layout the
// contraction's own packing of V[a',b',c,d] would write.  Per block of na a-rows:
//  - b' >= na: plane rows a' + na (b' - na), contracted straight into R2r;
//  - b' < na (the diagonal block): only a' <= b', at rows b'(b'+1)/2 + a' (tiles with none
//    are skipped), contracted into a small temporary and added to R2r with weight 1, or
//    1/2 for a' = b' (the symmetrisation supplies the (b',a') half),
// so every pair a <= b is formed once.  K (= v^2) is padded to a multiple of 32 with zeros.
// Needs the packed-plane kernels, o^2 >= 128 and v >= 64; returns false otherwise.
static bool df_ladder_half_planes(TensorOps& ops, long nv, long no, int nq, Prec pr, const double* Lf,
                                  const double* Rf, const double* T2x, double* R2r) {
  GemmEngine& G = ops.gemm;
  const long v = nv, vv = v * v, oo = no * no;
  if (G.transient || oo < 128 || v < 64) return false;
  if ((pr == Prec::F64 && G.f64_kernel != 1) || (pr == Prec::F32 && G.f32_kernel != 1)) return false;
  const int kind = pr == Prec::BF16X3 ? 0 : (pr == Prec::F32 ? 1 : 2);
  const long esz = kind == 2 ? 8 : 4;   // bytes per plane element (BF16: hi + lo)
  const long kpad = ((vv + 31) / 32) * 32;
  size_t mark = ops.arena.mark();
  const int p_ca[3] = {1, 0, 2};
  double* Lt = ops.arena.take(vv * nq);
  {
    Tensor in(const_cast<double*>(Lf), {v, v, (long)nq}), out(Lt, {v, v, (long)nq});
    ops.permute(in, p_ca, out, 1.0, 0.0);
  }
  double* Rt = Lt;
  if (Rf != Lf) {
    Rt = ops.arena.take(vv * nq);
    Tensor in(const_cast<double*>(Rf), {v, v, (long)nq}), out(Rt, {v, v, (long)nq});
    ops.permute(in, p_ca, out, 1.0, 0.0);
  }
  // planes of a block of ab a-rows: the triangle (ab (ab+1)/2 rows), then the ab (v - a0 - ab)
  // off-diagonal rows, plus 256 rows that a contraction's last tile reads (results
  // discarded); the temporary of the triangle's contraction
  const long free_b = (long)ops.arena.free_bytes() - 256 * kpad * esz;
  long ab = std::min<long>(v, kLadderMaxRows);
  auto need = [&](long r) { return (r * v + r * (r + 1) / 2) * kpad * esz + 8 * (r * (r + 1) / 2) * oo; };
  while (ab > 1 && need(ab) > (long)(0.8 * (double)free_b)) --ab;
  if (need(ab) > (long)(0.8 * (double)free_b)) {   // not even one a-row of planes: FP64 V blocks
    ops.arena.release_to(mark);
    return false;
  }
  const long nrv = ab * v + ab * (ab + 1) / 2 + 256;
  char* vp = reinterpret_cast<char*>(ops.arena.take((nrv * kpad * esz + 7) / 8));
  double* ytri = ops.arena.take(ab * (ab + 1) / 2 * oo);
  if (kpad != vv) {
    // BF16: the hi and lo planes are one run of 2 nrv rows of 2-byte elements
    const long rows = kind == 0 ? 2 * nrv : nrv;
    const int pe = kind == 0 ? 2 : (int)esz;
    hipLaunchKernelGGL(zero_plane_tail_kernel, dim3((unsigned)std::min<long>((rows * pe * 4 + 255) / 256, 65535)),
                       dim3(256), 0, ops.stream, vp, rows, kpad, pe);
    ECG_HIP(hipGetLastError());
  }
  auto planes_at = [&](GemmOp& h, long r0) {
    h.a_planes_ld = kpad;
    if (kind == 2) {
      h.a_planes = reinterpret_cast<const double*>(vp) + r0 * kpad;
    } else if (kind == 1) {
      h.a_planes = reinterpret_cast<const float*>(vp) + r0 * kpad;
    } else {
      h.a_planes = reinterpret_cast<const uint16_t*>(vp) + r0 * kpad;
      h.a_planes_lo = reinterpret_cast<const uint16_t*>(vp) + nrv * kpad + r0 * kpad;
    }
  };
  G.drop_kept_planes();
  for (long a0 = 0; a0 < v; a0 += ab) {
    const long na = std::min<long>(ab, v - a0);
    const long nb = v - a0;
    const long ntri = na * (na + 1) / 2;
    // V[a',b',c,d] = sum_Q L^[a0+a',c,Q] R^[a0+b',d,Q]   (m = c + v a', n = d + v b', k = Q)
    PlaneOut po;
    po.hi = vp;
    po.lo = kind == 0 ? static_cast<void*>(reinterpret_cast<uint16_t*>(vp) + nrv * kpad) : nullptr;
    po.ld = kpad; po.dm = (int)v; po.dn = (int)v; po.kn = v;
    GemmOp g;
    g.M = (int)(na * v); g.N = (int)(na * v); g.K = nq;
    g.A = Lt + v * a0; g.a_s0 = 1; g.a_k = vv;
    g.B = Rt + v * a0; g.b_s0 = 1; g.b_k = vv;
    g.alpha = 1.0; g.beta = 0.0;
    po.tri = 1;
    G.run_planes(g, kind, &po);                                   // the triangle a' <= b' < na
    if (nb > na) {
      GemmOp go = g;
      go.N = (int)((nb - na) * v);
      go.B = Rt + v * (a0 + na);
      PlaneOut pq = po;
      pq.tri = 0;
      pq.ra = 1; pq.rb = na;
      pq.hi = kind == 0 ? static_cast<void*>(reinterpret_cast<uint16_t*>(vp) + ntri * kpad)
                        : static_cast<void*>(vp + ntri * kpad * esz);
      pq.lo = kind == 0 ? static_cast<void*>(reinterpret_cast<uint16_t*>(vp) + nrv * kpad + ntri * kpad) : nullptr;
      G.run_planes(go, kind, &pq);                                // b' >= na
    }
    // R2r[a0+a',a0+b',i,j] += w sum_{cd} V[a',b',c,d] T[c,d,i,j]
    GemmOp h;
    h.N = (int)oo; h.K = (int)vv;
    h.B = T2x; h.b_s0 = 1; h.b_d0 = 1 << 30; h.b_s1 = 0; h.b_k = oo;
    h.b_keep = true;
    if (nb > na) {
      GemmOp ho = h;
      ho.M = (int)(na * (nb - na));
      planes_at(ho, ntri);
      ho.C = R2r + a0 + (a0 + na) * v; ho.c_ms0 = 1; ho.c_md0 = (int)na; ho.c_ms1 = v;
      ho.c_ns0 = vv; ho.c_nd0 = 1 << 30; ho.c_ns1 = 0;
      ho.alpha = 1.0; ho.beta = 1.0;
      G.run_planes(ho, kind);
    }
    h.M = (int)ntri;
    planes_at(h, 0);
    h.C = ytri; h.c_ms0 = 1; h.c_md0 = 1 << 30; h.c_ms1 = 0; h.c_ns0 = ntri; h.c_nd0 = 1 << 30; h.c_ns1 = 0;
    h.alpha = 1.0; h.beta = 0.0;
    G.run_planes(h, kind);
    const long n = ntri * oo;
    hipLaunchKernelGGL(ladder_tri_scatter_kernel, dim3((unsigned)std::min<long>((n + 255) / 256, 65535)), dim3(256),
                       0, ops.stream, (int)na, v, a0, ntri, oo, ytri, R2r);
    ECG_HIP(hipGetLastError());
  }
  G.drop_kept_planes();
  ops.arena.release_to(mark);
  return true;
}

void df_ladder_half(TensorOps& ops, long nv, long no, int nq, Prec pr, const double* Lf, const double* Rf,
                    const double* T2x, double* R2r) {
  if (df_ladder_half_planes(ops, nv, no, nq, pr, Lf, Rf, T2x, R2r)) return;
  size_t mark = ops.arena.mark();
  size_t free_b = ops.arena.free_bytes();
  const long v = nv;
  const long per_a = v * v * v;
  long ab = std::max<long>(1, std::min<long>(v, (long)(free_b * 0.8 / sizeof(double)) / per_a));
  double* V = ops.arena.take(ab * per_a);
  // T2x is read-only here: its packed BF16 planes serve every block's second product
  ops.gemm.drop_kept_planes();
  for (long a0 = 0; a0 < v; a0 += ab) {
    const long na = std::min<long>(ab, v - a0);
    const long nb = v - a0;
    // V[a',b',c,d] = sum_Q L^[a0+a',c,Q] R^[a0+b',d,Q]   (m = (a',c), n = (b',d), k = Q)
    GemmOp g;
    g.M = (int)(na * v); g.N = (int)(nb * v); g.K = nq;
    g.A = Lf + a0; g.a_s0 = 1; g.a_d0 = (int)na; g.a_s1 = v; g.a_k = v * v;
    g.B = Rf + a0; g.b_s0 = 1; g.b_d0 = (int)nb; g.b_s1 = v; g.b_k = v * v;
    g.C = V; g.c_ms0 = 1; g.c_md0 = (int)na; g.c_ms1 = na * nb;
    g.c_ns0 = na; g.c_nd0 = (int)nb; g.c_ns1 = na * nb * v;
    g.alpha = 1.0; g.beta = 0.0;
    ops.gemm.run_prec(g, pr);
    // R2r[a0+a',a0+b',i,j] += w sum_{cd} V[a',b',c,d] T[c,d,i,j]  (m = (a',b'), n = (i,j), k = (c,d));
    // T is supplied as Tx[i,j,c,d] so both operands stream along m / n.
    for (int part = 0; part < 2; ++part) {
      const long b0 = part == 0 ? 0 : na;           // diagonal block, then b' >= na
      const long nbp = part == 0 ? na : nb - na;
      if (nbp <= 0) continue;
      GemmOp h;
      h.M = (int)(na * nbp); h.N = (int)(no * no); h.K = (int)(v * v);
      h.A = V + na * b0; h.a_s0 = 1; h.a_d0 = (int)na; h.a_s1 = na; h.a_k = na * nb;
      h.B = T2x; h.b_s0 = 1; h.b_d0 = 1 << 30; h.b_s1 = 0; h.b_k = no * no;
      h.C = R2r + a0 + (a0 + b0) * v; h.c_ms0 = 1; h.c_md0 = (int)na; h.c_ms1 = v;
      h.c_ns0 = v * v; h.c_nd0 = 1 << 30; h.c_ns1 = 0;
      h.alpha = part == 0 ? 0.5 : 1.0; h.beta = 1.0;
      h.b_keep = true;
      ops.gemm.run_prec(h, pr);
    }
  }
  ops.gemm.drop_kept_planes();
  ops.arena.release_to(mark);
}

// ---------------------------------------------------------------- residual
// method: 0 = CCSD, 1 = DCSD.  calc_cc_resid (cc.jl:1343-1489) with use_kext.
void Context::residual(int method, Prec pr, const double* T1d, const double* T2d_, double* R1d, double* R2d) {
  const bool dc = (method == 1);
  const long v = nv, o = no, Q = nq;
  residual_calls++;
  ops.arena.reset();
  if (!capturing) ECG_HIP(hipEventRecord(ev[0], stream));
  dress(T1d);
  const Blocks& Rh = hermitian ? Ld : Rd;
  const Blocks& Ru = hermitian ? L : R;
  double* T2 = cp(T2d_);
  auto take4 = [&]() { return ops.arena.take(nvvoo()); };
  Tensor tT2 = T4d(T2, v, v, o, o);
  // T~ = 2 T - T[b,a,i,j]
  double* Tt = take4();
  Tensor tTt = T4d(Tt, v, v, o, o);
  const int p_ba[4] = {1, 0, 2, 3};
  ops.axpby(nvvoo(), 2.0, T2, 0.0, Tt);
  ops.permute(tT2, p_ba, tTt, -1.0, 1.0);
  Tensor klcd = T4d(oovv.d(), o, o, v, v);
  // ---- singles (cc.jl:1357-1378 and the kext part :1413-1421)
  Tensor tR1 = T2d(R1d, v, o);
  {
    // R1 = f^_vo  (dh_vo plus the 2J - K part of the kext singles term)
    ECG_HIP(hipMemcpy2DAsync(R1d, sizeof(double) * v, df.d() + no, sizeof(double) * n, sizeof(double) * v, o,
                             hipMemcpyDeviceToDevice, stream));
    // + T~[a,b,i,j] f^[j,b]
    double* fov = ops.arena.take(o * v);
    ECG_HIP(hipMemcpy2DAsync(fov, sizeof(double) * o, df.d() + (size_t)no * n, sizeof(double) * n, sizeof(double) * o,
                             v, hipMemcpyDeviceToDevice, stream));
    ops.contract_p(Prec::F64, 1.0, tTt, "abij", T2d(fov, o, v), "jb", 1.0, tR1, "ai");
    // - d_oovo[k,j,c,i] T~[c,a,k,j],  d_oovo[k,j,c,i] = (kc|ji)^
    double* doovo = ops.arena.take(o * o * v * o);
    Tensor tdoovo = T4d(doovo, o, o, v, o);
    ops.contract_p(Prec::F64, 1.0, T3d(Ld.ov.d(), o, v, Q), "kcQ", T3d(Rh.oo.d(), o, o, Q), "jiQ", 0.0, tdoovo, "kjci");
    ops.contract_p(Prec::F64, -1.0, tdoovo, "kjci", tTt, "cakj", 1.0, tR1, "ai");
    // + sum_Q sum_c L^[a,c,Q] Z[c,i,Q],  Z[c,i,Q] = sum_kd R[k,d,Q] (2T[c,d,i,k] - T[c,d,k,i])
    double* Tp = take4();
    Tensor tTp = T4d(Tp, v, v, o, o);
    const int p_ij[4] = {0, 1, 3, 2};
    ops.axpby(nvvoo(), 2.0, T2, 0.0, Tp);
    ops.permute(tT2, p_ij, tTp, -1.0, 1.0);
    double* Z = ops.arena.take(v * o * Q);
    Tensor tZ = T3d(Z, v, o, Q);
    ops.contract_p(pr, 1.0, tTp, "cdik", T3d(Ru.ov.d(), o, v, Q), "kdQ", 0.0, tZ, "ciQ");
    ops.contract_p(Prec::F64, 1.0, T3d(Ld.vv.d(), v, v, Q), "acQ", tZ, "ciQ", 1.0, tR1, "ai");
  }
  // ---- doubles
  Tensor tR2 = T4d(R2d, v, v, o, o);
  {
    // hole-hole ladder: W[k,l,i,j] = (ki|lj)^ (+ <kl|cd> T[c,d,i,j] for CCSD); R2 = W T[a,b,k,l]
    double* W = ops.arena.take(o * o * o * o);
    Tensor tW = T4d(W, o, o, o, o);
    ops.contract_p(Prec::F64, 1.0, T3d(Ld.oo.d(), o, o, Q), "kiQ", T3d(Rh.oo.d(), o, o, Q), "ljQ", 0.0, tW, "klij");
    if (!dc) ops.contract_p(pr, 1.0, klcd, "klcd", tT2, "cdij", 1.0, tW, "klij");
    ops.contract_p(pr, 1.0, tW, "klij", tT2, "abkl", 0.0, tR2, "abij");
  }
  // particle-particle ladder (kext): sum_Q L^[a,i,Q] R^[b,j,Q] + DF ladder.  The three
  // o^2 v^2 Q assemblies of the doubles (this one, d_voov and d_vovo below) run in FP32
  // products when the residual's products are FP32 or BF16 hi/lo: the solver's FP64
  // defect correction removes their rounding with that of the T products.  (In BF16 hi/lo
  // they raised the corrected residual's floor: one more exact iteration per solve.)
  const Prec pa = pr == Prec::F64 ? Prec::F64 : Prec::F32;
  ops.contract_p(pa, 1.0, T3d(Ld.vo.d(), v, o, Q), "aiQ", T3d(Rh.vo.d(), v, o, Q), "bjQ", 1.0, tR2, "abij");
  double* X = take4();
  if (!dc) {
    // <kl|cd> T[a,d,k,j] T[c,b,i,l]   (cc.jl:1435-1439)
    Tensor tX = T4d(X, o, v, v, o);
    ops.contract_p(pr, 1.0, klcd, "klcd", tT2, "adkj", 0.0, tX, "lcaj");
    ops.contract_p(pr, 1.0, tX, "lcaj", tT2, "cbil", 1.0, tR2, "abij");
  }
  // x_ad, x_ki (cc.jl:1441-1452)
  const double fac = dc ? 0.5 : 1.0;
  double* xad = ops.arena.take(v * v);
  double* xki = ops.arena.take(o * o);
  Tensor txad = T2d(xad, v, v), txki = T2d(xki, o, o);
  ECG_HIP(hipMemcpy2DAsync(xad, sizeof(double) * v, df.d() + (size_t)no * n + no, sizeof(double) * n,
                           sizeof(double) * v, v, hipMemcpyDeviceToDevice, stream));
  ECG_HIP(hipMemcpy2DAsync(xki, sizeof(double) * o, df.d(), sizeof(double) * n, sizeof(double) * o, o,
                           hipMemcpyDeviceToDevice, stream));
  ops.contract_p(pr, -fac, klcd, "klcd", tTt, "cakl", 1.0, txad, "ad");
  ops.contract_p(pr, fac, klcd, "klcd", tTt, "cdil", 1.0, txki, "ki");
  // R2r = x_ad T[d,b,i,j] - x_ki T[a,b,k,j]   (cc.jl:1455-1460)
  double* R2r = take4();
  Tensor tR2r = T4d(R2r, v, v, o, o);
  ops.contract_p(pr, 1.0, txad, "ad", tT2, "dbij", 0.0, tR2r, "abij");
  ops.contract_p(pr, -1.0, txki, "ki", tT2, "abkj", 1.0, tR2r, "abij");
  {
    // DF particle-particle ladder (half of it; completed by the symmetrisation below)
    size_t mk = ops.arena.mark();
    double* T2x = take4();
    Tensor tT2x = T4d(T2x, o, o, v, v);
    const int p_x[4] = {2, 3, 0, 1};
    ops.permute(tT2, p_x, tT2x, 1.0, 0.0);
    ladder(pr, T2x, R2r);
    ops.arena.release_to(mk);
  }
  {
    // Y[a,k,i,c] = d_voov[a,k,i,c] + 1/2 <kl|cd> T~[a,d,i,l];  R2r += Y T~[c,b,k,j]   (cc.jl:1462-1468)
    Tensor tY = T4d(X, v, o, o, v);
    ops.contract_p(pa, 1.0, T3d(Ld.vo.d(), v, o, Q), "aiQ", T3d(Ru.ov.d(), o, v, Q), "kcQ", 0.0, tY, "akic");
    ops.contract_p(pr, 0.5, klcd, "klcd", tTt, "adil", 1.0, tY, "akic");
    ops.contract_p(pr, 1.0, tY, "akic", tTt, "cbkj", 1.0, tR2r, "abij");
  }
  if (!dc) {
    // - <kl|cd> T[d,a,k,i] (T[c,b,l,j] - T[b,c,l,j])   (cc.jl:1470-1475)
    Tensor tZ2 = T4d(X, o, v, v, o);
    ops.contract_p(pr, 1.0, klcd, "klcd", tT2, "daki", 0.0, tZ2, "lcai");
    double* Tm = take4();
    Tensor tTm = T4d(Tm, v, v, o, o);
    ops.copy(nvvoo(), T2, Tm);
    ops.permute(tT2, p_ba, tTm, -1.0, 1.0);
    ops.contract_p(pr, -1.0, tZ2, "lcai", tTm, "cblj", 1.0, tR2r, "abij");
  }
  {
    // - d_vovo[a,k,c,i] T[c,b,k,j] - d_vovo[b,k,c,i] T[a,c,k,j],  d_vovo[a,k,c,i] = (ac|ki)^   (cc.jl:1476-1481)
    Tensor tD = T4d(X, v, o, v, o);
    ops.contract_p(pa, 1.0, T3d(Ld.vv.d(), v, v, Q), "acQ", T3d(Rh.oo.d(), o, o, Q), "kiQ", 0.0, tD, "akci");
    ops.contract_p(pr, -1.0, tD, "akci", tT2, "cbkj", 1.0, tR2r, "abij");
    ops.contract_p(pr, -1.0, tD, "bkci", tT2, "ackj", 1.0, tR2r, "abij");
  }
  // R2 += R2r + R2r[b,a,j,i]
  const int p_sym[4] = {1, 0, 3, 2};
  ops.axpby(nvvoo(), 1.0, R2r, 1.0, R2d);
  ops.permute(tR2r, p_sym, tR2, 1.0, 1.0);
  if (!capturing) {
    ECG_HIP(hipEventRecord(ev[1], stream));
    ECG_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    ECG_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    last_residual_ms = ms;
    ECG_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
    last_ladder_ms = ms;
  }
  ops.arena.reset();
}

// Lower bound of the caller region residual_noalloc needs: the residual's own
// temporaries at their peak (five o^2 v^2 tensors, d_oovo, Z[c,i,Q], W[k,l,i,j],
// x_ad, x_ki, f_ov) and one a-row (v^3) of the ladder's integral block.  The
// contraction planner's transposed operands come on top; if they do not fit, the
// arena stops the call with an error before any launch that would overrun it.
size_t Context::residual_noalloc_min_bytes() const {
  const long v = nv, o = no, Q = nq;
  const long d = 5 * nvvoo() + o * o * o * v + v * o * Q + o * o * o * o + v * v + o * o + o * v + v * v * v;
  return sizeof(double) * (size_t)d + 16 * 256;
}

void Context::residual_noalloc(int method, const double* T1d, const double* T2d, double* R1d, double* R2d,
                               void* scratch, size_t bytes) {
  const size_t need = residual_noalloc_min_bytes();
  if (need > bytes)
    throw Error(-2, "system too large for the allocation-free residual: needs at least " +
                        std::to_string(need >> 20) + " MiB of scratch, the region holds " +
                        std::to_string(bytes >> 20) + " MiB");
  // the arena and the engine's split-K partials run from the caller's region for this call
  struct Restore {
    TensorOps& ops;
    ~Restore() {
      ops.arena.ext = nullptr;
      ops.arena.ext_bytes = 0;
      ops.arena.reset();
      ops.gemm.transient = nullptr;
    }
  } restore{ops};
  ops.arena.ext = static_cast<char*>(scratch);
  ops.arena.ext_bytes = bytes;
  ops.arena.reset();
  ops.gemm.transient = &ops.arena;
  residual(method, Prec::F64, T1d, T2d, R1d, R2d);
}

// The residual from a captured HIP graph: the first call for a (method, precision,
// buffers) key runs eagerly (sizing every scratch buffer and producing this call's
// result) and then records the same launch sequence; later calls replay it.  The
// sequence depends only on the sizes, the key's pointers and the scratch buffers,
// so a replay recomputes the residual of the current T1/T2.
void Context::residual_graph(int method, Prec pr, const double* T1d, const double* T2d, double* R1d, double* R2d) {
  auto scratch = [&](ResidualGraph& g) {
    g.ws = ops.gemm.ws.p; g.wsb = ops.gemm.ws.bytes;
    g.pk = ops.gemm.pk.p; g.pkb = ops.gemm.pk.bytes;
    g.keep = ops.gemm.pk_keep.p; g.keepb = ops.gemm.pk_keep.bytes;
    g.arena = ops.arena.base();
  };
  for (size_t k = 0; k < rgraphs.size(); ++k) {
    ResidualGraph& g = rgraphs[k];
    if (g.method != method || g.pr != pr || g.t1 != T1d || g.t2 != T2d || g.r1 != R1d || g.r2 != R2d) continue;
    ResidualGraph now;
    scratch(now);
    if (now.ws == g.ws && now.wsb == g.wsb && now.pk == g.pk && now.pkb == g.pkb && now.keep == g.keep &&
        now.keepb == g.keepb && now.arena == g.arena) {
      residual_calls++;
      ECG_HIP(hipGraphLaunch(g.exec, stream));
      return;
    }
    (void)hipGraphExecDestroy(g.exec);   // a recorded scratch buffer moved: record again
    rgraphs.erase(rgraphs.begin() + (long)k);
    break;
  }
  residual(method, pr, T1d, T2d, R1d, R2d);
  ResidualGraph g;
  g.method = method; g.pr = pr; g.t1 = T1d; g.t2 = T2d; g.r1 = R1d; g.r2 = R2d;
  scratch(g);
  hipGraph_t graph = nullptr;
  capturing = true;
  ECG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  try {
    residual(method, pr, T1d, T2d, R1d, R2d);
  } catch (...) {
    (void)hipStreamEndCapture(stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    capturing = false;
    throw;
  }
  capturing = false;
  ECG_HIP(hipStreamEndCapture(stream, &graph));
  ResidualGraph after;
  scratch(after);
  ECG_CHECK(after.ws == g.ws && after.pk == g.pk && after.keep == g.keep && after.arena == g.arena,
            "residual graph: a scratch buffer was reallocated during capture");
  hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  ECG_HIP(e);
  rgraphs.push_back(g);
}

// ---------------------------------------------------------------- energies, norms
// d = sum X[a,b,i,j] Y[a,b,i,j],  e = sum X[b,a,i,j] Y[a,b,i,j]  (FP64 accumulation)
void Context::doubles_de(const double* X, const double* Y, double& d, double& e) {
  long np = (long)no * no;
  dots_part.ensure(sizeof(double) * 2 * np);
  launch_pair_dots(stream, nv, np, X, Y, dots_part.d());
  dots_host.resize(2 * np);
  ECG_HIP(hipMemcpyAsync(dots_host.data(), dots_part.p, sizeof(double) * 2 * np, hipMemcpyDeviceToHost, stream));
  ECG_HIP(hipStreamSynchronize(stream));
  d = 0.0;
  e = 0.0;
  for (long k = 0; k < np; ++k) { d += dots_host[2 * k]; e += dots_host[2 * k + 1]; }
}

double Context::doubles_dot(const double* X, const double* Y, double cex) {
  double d, e;
  doubles_de(X, Y, d, e);
  return d - cex * e;
}

// calc_doubles_energy (cc.jl:157-167): ESS = d - e, EOS = d with d = sum T V, e = sum T_ba V
void Context::doubles_energy(const double* T2d_, double* e) {
  double d, ex;
  doubles_de(T2d_, vvoo.d(), d, ex);
  e[1] = d - ex;
  e[2] = d;
  e[0] = e[1] + e[2];
}

// calc_singles_energy (cc.jl:97-114)
void Context::singles_energy(const double* T1d, bool fock_only, double* e) {
  double ss = 0.0, os = 0.0;
  size_t mark = ops.arena.mark();
  if (!fock_only) {
    double* tt = ops.arena.take(nvvoo());
    Tensor t1 = T2d(cp(T1d), nv, no), ttt = T4d(tt, nv, nv, no, no);
    ops.contract_p(Prec::F64, 1.0, t1, "ai", t1, "bj", 0.0, ttt, "abij");
    double d, ex;
    doubles_de(tt, vvoo.d(), d, ex);
    ss = d - ex;
    os = d;
  }
  double* fvo = ops.arena.take(nvo());
  // f_ov[i,a] pairs with T1[a,i]: copy f[o,v] transposed via the permute kernel
  double* fovb = ops.arena.take(nvo());
  ECG_HIP(hipMemcpy2DAsync(fovb, sizeof(double) * no, fock.d() + (size_t)no * n, sizeof(double) * n,
                           sizeof(double) * no, nv, hipMemcpyDeviceToDevice, stream));
  Tensor src = T2d(fovb, no, nv), dst = T2d(fvo, nv, no);
  const int pt[2] = {1, 0};
  ops.permute(src, pt, dst, 1.0, 0.0);
  double f = 2.0 * ops.dot_host(nvo(), T1d, fvo);
  ops.arena.release_to(mark);
  e[1] = ss; e[2] = os; e[0] = ss + os + f;
}

// calc_hylleraas (cc.jl:195-241): e = [E, ESS, EOS, pE, pESS, pEOS]
void Context::hylleraas(const double* T1d, const double* T2d_, const double* R1d, const double* R2d, double* e) {
  double pd, pex, rd, rex;
  doubles_de(T2d_, vvoo.d(), pd, pex);
  doubles_de(T2d_, R2d, rd, rex);
  double s1[3];
  singles_energy(T1d, false, s1);
  double ET1SS = s1[1], ET1OS = s1[2];
  double pET1 = s1[0] - ET1SS - ET1OS;   // 2 f.T1
  double rET1 = 2.0 * ops.dot_host(nvo(), R1d, T1d);
  double ET2d = pd + rd, ET2ex = pex + rex;
  double ET2SS = ET2d - ET2ex + ET1SS, ET2OS = ET2d + ET1OS;
  double ET2 = (ET2d - ET2ex) + ET2d + (ET1SS + ET1OS) + pET1 + rET1;
  double pSS = pd - pex + ET1SS, pOS = pd + ET1OS;
  double pE = (pd - pex) + pd + pET1 + ET1SS + ET1OS;
  e[0] = ET2; e[1] = ET2SS; e[2] = ET2OS; e[3] = pE; e[4] = pSS; e[5] = pOS;
}

void Context::update(double* T1d, double* T2d_, const double* R1d, const double* R2d, double shifts, double shiftp) {
  const double* eo = eps.d();
  const double* ev = eps.d() + no;
  hipLaunchKernelGGL(update_doubles_kernel, dim3(nblk(nvvoo())), dim3(256), 0, stream, nv, no, eo, ev, shiftp, R2d,
                     T2d_, 0);
  hipLaunchKernelGGL(update_singles_kernel, dim3((unsigned)((nvo() + 255) / 256)), dim3(256), 0, stream, nv, no, eo,
                     ev, shifts, R1d, (long)nv, T1d, 0);
  ECG_HIP(hipGetLastError());
}

// MP2 amplitudes (cc.jl:1011-1024): T2 = -<ab|ij>/D, T1 = -f_vo/(e_a - e_i); kernels only
void Context::mp2_amplitudes(double* T1d, double* T2d_) {
  const double* eo = eps.d();
  const double* ev = eps.d() + no;
  hipLaunchKernelGGL(update_doubles_kernel, dim3(nblk(nvvoo())), dim3(256), 0, stream, nv, no, eo, ev, 0.0,
                     abij(), T2d_, 1);
  hipLaunchKernelGGL(update_singles_kernel, dim3((unsigned)((nvo() + 255) / 256)), dim3(256), 0, stream, nv, no, eo,
                     ev, 0.0, fock.d() + no, (long)n, T1d, 1);
  ECG_HIP(hipGetLastError());
}

// Per-(i,j) partial sums [d, e] of two v x v x o x o tensors (pair_dots in ecg_cc.h)
void launch_pair_dots(hipStream_t s, int nv, long npairs, const double* X, const double* Y, double* part) {
  hipLaunchKernelGGL(cs_dots_kernel, dim3((unsigned)npairs), dim3(1024), 0, s, nv, X, Y, part);
  ECG_HIP(hipGetLastError());
}

// MP2 guess (cc.jl:1011-1024) and its energy
void Context::mp2_guess(double* T1d, double* T2d_, double* emp2) {
  mp2_amplitudes(T1d, T2d_);
  double e2[3], e1[3];
  doubles_energy(T2d_, e2);
  singles_energy(T1d, true, e1);
  for (int k = 0; k < 3; ++k) emp2[k] = e2[k] + e1[k];
}

// ---------------------------------------------------------------- DIIS (diis.jl:34-305)
void Context::diis_reset(int maxd, double resthr) {
  diis.maxdiis = maxd;
  diis.resthr = resthr;
  diis.next = 0;
  diis.ndim = 0;
  diis.bmat.assign((size_t)(maxd + 1) * (maxd + 1), 0.0);
  if ((int)diis.amps.size() != maxd) {
    diis.amps.clear();
    diis.res.clear();
    for (int k = 0; k < maxd; ++k) {
      diis.amps.emplace_back(sizeof(double) * (nvo() + nvvoo()));
      diis.res.emplace_back(sizeof(double) * (nvo() + nvvoo()));
    }
  }
}

// symmetric eigen-decomposition by cyclic Jacobi (small matrices)
void jacobi_eig(int m, std::vector<double>& A, std::vector<double>& V, std::vector<double>& w) {
  V.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) V[(size_t)i * m + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) off += A[(size_t)p * m + q] * A[(size_t)p * m + q];
    if (off < 1e-300) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        double apq = A[(size_t)p * m + q];
        if (std::fabs(apq) < 1e-300) continue;
        double app = A[(size_t)p * m + p], aqq = A[(size_t)q * m + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; ++k) {
          double akp = A[(size_t)k * m + p], akq = A[(size_t)k * m + q];
          A[(size_t)k * m + p] = c * akp - s * akq;
          A[(size_t)k * m + q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          double apk = A[(size_t)p * m + k], aqk = A[(size_t)q * m + k];
          A[(size_t)p * m + k] = c * apk - s * aqk;
          A[(size_t)q * m + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          double vkp = V[(size_t)k * m + p], vkq = V[(size_t)k * m + q];
          V[(size_t)k * m + p] = c * vkp - s * vkq;
          V[(size_t)k * m + q] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(m);
  for (int i = 0; i < m; ++i) w[i] = A[(size_t)i * m + i];
}

void Context::diis_perform(double* A1, double* A2, const double* R1d, const double* R2d, int dots) {
  DiisState& D = diis;
  const int md = D.maxdiis;
  if (D.ndim < md) D.ndim++;
  const int ithis = D.next, nd = D.ndim;
  const long n1 = nvo(), n2 = nvvoo();
  ops.copy(n1, A1, D.amps[ithis].d());
  ops.copy(n2, A2, D.amps[ithis].d() + n1);
  ops.copy(n1, R1d, D.res[ithis].d());
  ops.copy(n2, R2d, D.res[ithis].d() + n1);
  // closed-shell custom dots (cctools.jl:453-456,511-514) or the plain dot
  auto dotv = [&](const double* x, const double* y) {
    if (dots == 0) return 2.0 * ops.dot_host(n1, x, y) + doubles_dot(x + n1, y + n1, 0.5) * 2.0;
    return ops.dot_host(n1 + n2, x, y);
  };
  auto B = [&](int i, int j) -> double& { return D.bmat[(size_t)i * (md + 1) + j]; };
  const double* rt = D.res[ithis].d();
  double thisdot = dotv(rt, rt);
  for (int i = 0; i < nd; ++i) {
    if (i != ithis) {
      double d = dotv(rt, D.res[i].d());
      B(i, ithis) = d;
      B(ithis, i) = d;
    } else {
      B(ithis, ithis) = thisdot;
    }
    B(i, nd) = -1.0;
    B(nd, i) = -1.0;
  }
  B(nd, nd) = 0.0;
  // coeffs = pinv(bmat) * rhs, through the SVD (= |eigenvalues| of the symmetric B)
  int m = nd + 1;
  std::vector<double> Bm((size_t)m * m), V, w;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) Bm[(size_t)i * m + j] = B(i, j);
  jacobi_eig(m, Bm, V, w);
  double smax = 0.0;
  for (double x : w) smax = std::max(smax, std::fabs(x));
  std::vector<double> coeff(m, 0.0);
  for (int k = 0; k < m; ++k) {
    if (std::fabs(w[k]) < 2.220446049250313e-16 * smax || w[k] == 0.0) continue;
    double proj = -V[(size_t)nd * m + k];  // (V^T rhs)_k with rhs = -e_nd
    for (int i = 0; i < m; ++i) coeff[i] += V[(size_t)i * m + k] * proj / w[k];
  }
  if (nd == 1 && D.next == 0 && thisdot > D.resthr) {
    D.ndim = 0;
    return;
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
  // A = sum_i c_i amps_i
  ops.axpby(n1, coeff[0], D.amps[0].d(), 0.0, A1);
  ops.axpby(n2, coeff[0], D.amps[0].d() + n1, 0.0, A2);
  for (int i = 1; i < nd; ++i) {
    ops.axpby(n1, coeff[i], D.amps[i].d(), 1.0, A1);
    ops.axpby(n2, coeff[i], D.amps[i].d() + n1, 1.0, A2);
  }
}

// ---------------------------------------------------------------- iterations
Context::~Context() {
  for (auto& g : rgraphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

void Context::state_init(const double* T1h, const double* T2h, const SolverOpts& o) {
  const long n1 = nvo(), n2 = nvvoo();
  have_delta = false;
  T1.ensure(sizeof(double) * n1);
  T2.ensure(sizeof(double) * n2);
  R1.ensure(sizeof(double) * n1);
  R2.ensure(sizeof(double) * n2);
  if (T1h && T2h) {
    ECG_HIP(hipMemcpyAsync(T1.p, T1h, sizeof(double) * n1, hipMemcpyHostToDevice, stream));
    ECG_HIP(hipMemcpyAsync(T2.p, T2h, sizeof(double) * n2, hipMemcpyHostToDevice, stream));
  } else {
    double e[3];
    mp2_guess(T1.d(), T2.d(), e);
  }
  diis_reset(o.maxdiis, o.resthr);
  iter = 0;
  ECG_HIP(hipStreamSynchronize(stream));
}

// One pass of the loop body of cc_iterations! (cc.jl:2546-2629), closed-shell branch.
void Context::state_step(int method, Prec pr, const SolverOpts& o, double* info, bool refresh) {
  const long n1 = nvo();
  iter++;
  auto resid = [&](Prec p, double* r1, double* r2) {
    if (o.graphs && !ops.gemm.timed) residual_graph(method, p, T1.d(), T2.d(), r1, r2);
    else residual(method, p, T1.d(), T2.d(), r1, r2);
  };
  if (pr != Prec::F64 && refresh) {
    // the exact residual R_64 at this T drives this iteration; delta = R_64 - R_low corrects
    // the following low-precision ones.  When R_64 already meets thr (the iteration that
    // converges) R_low is not formed: a further iteration would refresh again anyway
    // (the refresh follows every corrected norm below 100 thr)
    resid(Prec::F64, R1.d(), R2.d());
    const double nr = 2.0 * doubles_dot(R2.d(), R2.d(), 0.5) + 2.0 * ops.dot_host(n1, R1.d(), R1.d());
    if (nr >= o.thr) {
      D1.ensure(sizeof(double) * n1);
      D2.ensure(sizeof(double) * nvvoo());
      resid(pr, D1.d(), D2.d());
      ops.axpby(n1, 1.0, R1.d(), -1.0, D1.d());
      ops.axpby(nvvoo(), 1.0, R2.d(), -1.0, D2.d());
      have_delta = true;
    }
  } else {
    resid(pr, R1.d(), R2.d());
    if (pr != Prec::F64 && have_delta) {
      ops.axpby(n1, 1.0, D1.d(), 1.0, R1.d());
      ops.axpby(nvvoo(), 1.0, D2.d(), 1.0, R2.d());
    }
  }
  // calc_doubles_norm (cctools.jl:477-480): sum (2T - T_ba) T
  double NormT2 = 2.0 * doubles_dot(T2.d(), T2.d(), 0.5);
  double NormR2 = 2.0 * doubles_dot(R2.d(), R2.d(), 0.5);
  double Eh[6];
  hylleraas(T1.d(), T2.d(), R1.d(), R2.d(), Eh);
  double NormT1 = 2.0 * ops.dot_host(n1, T1.d(), T1.d());
  double NormR1 = 2.0 * ops.dot_host(n1, R1.d(), R1.d());
  update(T1.d(), T2.d(), R1.d(), R2.d(), o.shifts, o.shiftp);
  diis_perform(T1.d(), T2.d(), R1.d(), R2.d(), 0);
  double e2[3], e1[3];
  doubles_energy(T2.d(), e2);
  singles_energy(T1.d(), false, e1);
  double En = e2[0] + e1[0];
  double dE = En - Eh[0];
  double NormR = NormR1 + NormR2;
  const double thren = std::sqrt(o.thr) * o.conven;
  info[0] = Eh[0]; info[1] = Eh[1]; info[2] = Eh[2]; info[3] = Eh[3];
  info[4] = dE; info[5] = NormR; info[6] = 1.0 + NormT1 + NormT2;
  info[7] = (NormR < o.thr && std::fabs(dE) < thren) ? 1.0 : 0.0;
}

// cc_iterations! closed-shell (cc.jl:2518-2648) from the amplitudes already in T1/T2.
// energies: [E_H, ESS, EOS, E_proj, dE, NormR]; returns the number of iterations.
// Mixed precision (pr != F64), defect correction: iterations with the residual
// R_low + delta; delta is refreshed with an FP64 residual (an exact iteration)
// every refine_every-th iteration (when > 0), whenever the corrected residual
// norm drops below 100*thr, and whenever it stalls (less than 20% below the
// smallest norm since the last refresh: the low-precision rounding floor).
// Convergence is declared only on an exact (FP64) iteration, so the fixed
// point is the FP64 solution.  The corrected residual has a floor set by the
// rounding of the low-precision products (their error does not shrink with
// T - T_ref); when the first low-precision iteration after each of two refreshes
// in a row already stalls, the floor is reached and the remaining iterations are
// plain FP64 ones with a fresh DIIS history (needed only for thresholds below
// that floor).
int Context::solve(int method, Prec pr, const SolverOpts& o, double* energies, int* converged, FILE* log) {
  // FP64 with graphs: whole iterations on the device, replayed from one captured graph
  if (pr == Prec::F64 && o.graphs) return dev_solve(method, o, energies, converged, log);
  diis_reset(o.maxdiis, o.resthr);
  iter = 0;
  have_delta = false;
  double info[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int conv = 0, it = 0;
  const bool mixed = (pr != Prec::F64);
  bool refresh_next = false, fp64_only = false, after_refresh = false;
  int quick_stalls = 0;
  double best = HUGE_VAL;
  fp64_iterations = 0;
  for (it = 1; it <= o.maxit; ++it) {
    const bool refresh =
        mixed && !fp64_only && (refresh_next || (o.refine_every > 0 && (it % o.refine_every) == 0));
    const bool exact = !mixed || refresh || fp64_only;
    if (fp64_only) state_step(method, Prec::F64, o, info, false);
    else state_step(method, pr, o, info, refresh);
    if (exact) fp64_iterations++;
    if (log) {
      std::fprintf(log, "%4d %12.8f %16.10f %12.3e %12.3e %s\n", it, info[6], info[0], info[4], info[5],
                   exact ? "fp64" : "low");
      std::fflush(log);
    }
    if (info[7] != 0.0 && exact) {
      conv = 1;
      break;
    }
    if (mixed && !fp64_only) {
      if (refresh) best = HUGE_VAL;
      const bool stalled = info[5] > 0.8 * best;
      if (!refresh && after_refresh) {
        quick_stalls = stalled ? quick_stalls + 1 : 0;
        fp64_only = quick_stalls >= 2;
        // the DIIS history holds corrected residuals that do not match the exact ones
        // (DIIS keeps its largest-coefficient vectors, so they would persist): restart it
        if (fp64_only) diis_reset(o.maxdiis, o.resthr);
      }
      after_refresh = refresh;
      best = std::min(best, info[5]);
      refresh_next = info[5] < 100.0 * o.thr || stalled;
    }
  }
  if (it > o.maxit) it = o.maxit;
  for (int k = 0; k < 6; ++k) energies[k] = info[k];
  *converged = conv;
  return it;
}

}  // namespace ecg
