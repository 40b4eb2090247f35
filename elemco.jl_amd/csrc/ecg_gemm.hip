// MFMA GEMM kernels for gfx950 (see ecg_gemm.h for the operand addressing).
//
// FP64 path: v_mfma_f64_16x16x4_f64.  A 256-thread workgroup (4 waves, 2x2)
// computes a 128x128 tile of C; each wave owns 64x64 = 4x4 MFMA tiles held in
// registers.  K advances in slabs of 16 staged through double-buffered LDS
// (register prefetch of the next slab overlaps the MFMAs of the current one).
// MFMA operand layout of the 16x16x4 f64 builtin on gfx950 (checked with a lane probe):
//   A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16*(i%4), slot i/4.
#include "ecg_gemm.h"
#include "ecg_gemm_dev.h"
#include "ecg_tensor.h"

#include <algorithm>

namespace ecg {

typedef double d4_t __attribute__((ext_vector_type(4)));

// K runs over the concatenation of nseg segments, each padded to a multiple of
// the K slab (BKS) so that a slab never straddles two segments; slab k0 of the
// padded range belongs to segment s with kp[s] <= k0 < kp[s+1].
struct GemmParams {
  int M, N, K, kchunk, ksplit, tiles_m, tiles_n;   // K: padded total
  int nseg;
  const double* sA[MAXSEG]; long sak[MAXSEG];
  const double* sB[MAXSEG]; long sbk[MAXSEG];
  int sK[MAXSEG], kp[MAXSEG + 1];
  long a_s0, a_s1, a_z; int a_d0;
  long b_s0, b_s1, b_z; int b_d0;
  double* C; long c_ms0, c_ms1, c_ns0, c_ns1, c_z; int c_md0, c_nd0;
  double alpha, beta;
  int cplain;   // 1: C is a plain column-major partial-sum slab (split-K workspace)
  long ldc_ws;  // leading dimension of the workspace slab
};

constexpr int BM = 128, BN = 128, BKS = 16, LDP = BM + 16;

typedef float f4_t __attribute__((ext_vector_type(4)));

// Per-precision pieces: operand type in LDS, accumulator, the 16x16x4 MFMA and
// the accumulator row layout (lane probe on gfx950: f64 slot r of lane l holds
// row l/16 + 4r, f32 slot r holds row 4(l/16) + r; the column is l%16 for both).
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4_t acc_t;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct Mfma<float> {
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// T = double: FP64 MFMA.  T = float: operands rounded to FP32 when staged in
// LDS, FP32 MFMA accumulation, FP64 epilogue (alpha, beta and C stay FP64).
// SEG: several K segments (the slab's segment is looked up per slab); without
// it the single product's base pointers and strides stay in registers.
template <typename T, bool AK, bool BK, bool SEG>
__global__ __launch_bounds__(256, 2) void gemm_kernel(GemmParams p) {
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T As[2][BKS][LDP];
  __shared__ T Bs[2][BKS][LDP];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, p.tiles_m, tm, tn);
  const int bm = tm * BM, bn = tn * BN;
  const int z = blockIdx.z;
  const int bz = z / p.ksplit, kz = z - bz * p.ksplit;
  const int kbeg = kz * p.kchunk;
  const int kend = min(p.K, kbeg + p.kchunk);

  // per-thread row / column offsets (fixed along K)
  long aoff[8], boff[8];
  bool aok[8], bok[8];
  if (!AK) {
    int m = bm + (t & 127);
    bool ok = m < p.M;
    long o = ok ? split_off(m, p.a_d0, p.a_s0, p.a_s1) : 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) { aoff[j] = o; aok[j] = ok; }
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      int m = bm + (t >> 4) + 16 * j;
      aok[j] = m < p.M;
      aoff[j] = aok[j] ? split_off(m, p.a_d0, p.a_s0, p.a_s1) : 0;
    }
  }
  if (!BK) {
    int n = bn + (t & 127);
    bool ok = n < p.N;
    long o = ok ? split_off(n, p.b_d0, p.b_s0, p.b_s1) : 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) { boff[j] = o; bok[j] = ok; }
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      int n = bn + (t >> 4) + 16 * j;
      bok[j] = n < p.N;
      boff[j] = bok[j] ? split_off(n, p.b_d0, p.b_s0, p.b_s1) : 0;
    }
  }

  double ra[8], rb[8];
  unsigned amask = 0, bmask = 0;   // bit j: element j of the prefetched slab is in range
  // branch-free: out-of-range elements read element 0 of the segment (always
  // valid) and are zeroed when stored to LDS.  The segment of slab k0 is uniform
  // across the workgroup; selects with constant indices keep the kernel
  // arguments in SGPRs.
  auto gload = [&](int k0) {
    const double* A = p.sA[0];
    const double* B = p.sB[0];
    long a_k = p.sak[0], b_k = p.sbk[0];
    int kl = k0, klen = p.sK[0];
    if constexpr (SEG) {
#pragma unroll
      for (int q = 1; q < MAXSEG; ++q) {
        if (q < p.nseg && k0 >= p.kp[q]) {
          A = p.sA[q]; B = p.sB[q]; a_k = p.sak[q]; b_k = p.sbk[q];
          kl = k0 - p.kp[q]; klen = p.sK[q];
        }
      }
    }
    A += (long)bz * p.a_z;
    B += (long)bz * p.b_z;
    // one 64-bit product per operand and slab: rows (AK) or k (otherwise) advance by a fixed stride
    const int kta = AK ? (kl + (t & 15)) : (kl + (t >> 7));
    const int ktb = BK ? (kl + (t & 15)) : (kl + (t >> 7));
    const long ka = (long)kta * a_k, kb = (long)ktb * b_k;
    amask = bmask = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = AK ? kta : kta + 2 * j;
      const bool ok = aok[j] && k < klen;
      const long off = AK ? aoff[j] + ka : aoff[0] + ka + (long)(2 * j) * a_k;
      ra[j] = A[ok ? off : 0];
      amask |= (unsigned)ok << j;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = BK ? ktb : ktb + 2 * j;
      const bool ok = bok[j] && k < klen;
      const long off = BK ? boff[j] + kb : boff[0] + kb + (long)(2 * j) * b_k;
      rb[j] = B[ok ? off : 0];
      bmask |= (unsigned)ok << j;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const T x = ((amask >> j) & 1) ? (T)ra[j] : (T)0;
      if (AK) As[buf][t & 15][(t >> 4) + 16 * j] = x;
      else    As[buf][(t >> 7) + 2 * j][t & 127] = x;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const T x = ((bmask >> j) & 1) ? (T)rb[j] : (T)0;
      if (BK) Bs[buf][t & 15][(t >> 4) + 16 * j] = x;
      else    Bs[buf][(t >> 7) + 2 * j][t & 127] = x;
    }
  };

  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int li = lane & 15, lk = lane >> 4;
  int cur = 0;
  if (kbeg < kend) {
    gload(kbeg);
    sstore(0);
  }
  __syncthreads();
  // operand fragments of k-step kk of the current LDS buffer
  auto frag = [&](T (&a)[4], T (&b)[4], int kk) {
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = As[cur][kk + lk][wm * 64 + i * 16 + li];
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = Bs[cur][kk + lk][wn * 64 + j * 16 + li];
  };
  for (int k0 = kbeg; k0 < kend; k0 += BKS) {
    const bool more = k0 + BKS < kend;
    // the first fragments are read before the prefetch, whose address arithmetic then
    // covers their LDS latency
    T a[4], b[4];
    frag(a, b, 0);
    // unconditional prefetch (the last slab reloads itself): a conditional load would make
    // the compiler wait for it before the MFMAs below instead of before the LDS store
    gload(more ? k0 + BKS : k0);
    asm volatile("" ::: "memory");   // keep the prefetch loads issued ahead of the MFMAs
#pragma unroll
    for (int kk = 0; kk < BKS; kk += 4) {
      T an[4], bn[4];
      if (kk + 4 < BKS) frag(an, bn, kk + 4);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(a[i], b[j], acc[i][j]);
      if (kk + 4 < BKS) {
#pragma unroll
        for (int i = 0; i < 4; ++i) { a[i] = an[i]; b[i] = bn[i]; }
      }
    }
    if (more) sstore(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }

  // epilogue
  double* C = p.C + (long)z * p.c_z;  // z-slab for split-K workspaces, batch for direct output
  if (!p.cplain) C = p.C + (long)bz * p.c_z;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int n = bn + wn * 64 + j * 16 + li;
    if (n >= p.N) continue;
    long cn = p.cplain ? (long)n * p.ldc_ws : split_off(n, p.c_nd0, p.c_ns0, p.c_ns1);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int m = bm + wm * 64 + i * 16 + Mfma<T>::row(lk, r);
        if (m >= p.M) continue;
        long cm = p.cplain ? (long)m : split_off(m, p.c_md0, p.c_ms0, p.c_ms1);
        double v = p.alpha * (double)acc[i][j][r];
        if (p.beta != 0.0) v += p.beta * C[cm + cn];
        C[cm + cn] = v;
      }
    }
  }
}

// C = alpha * sum_s W[s] + beta * C over split-K partial slabs W[s] (M x N plain).
__global__ void splitk_reduce_kernel(const double* __restrict__ W, int ksplit, long slab, int M, int N,
                                     double* C, long c_ms0, long c_ms1, int c_md0, long c_ns0,
                                     long c_ns1, int c_nd0, long c_z, double alpha, double beta) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int bz = blockIdx.y;
  long total = (long)M * N;
  if (idx >= total) return;
  int m = (int)(idx % M), n = (int)(idx / M);
  const double* w = W + (long)bz * ksplit * slab + idx;
  double s = 0.0;
  for (int q = 0; q < ksplit; ++q) s += w[(long)q * slab];
  double* c = C + (long)bz * c_z + split_off(m, c_md0, c_ms0, c_ms1) + split_off(n, c_nd0, c_ns0, c_ns1);
  double v = alpha * s;
  if (beta != 0.0) v += beta * (*c);
  *c = v;
}

template <typename T>
void launch_gemm(bool ak, bool bk, bool seg, dim3 grid, hipStream_t stream, const GemmParams& p) {
  if (seg) {
    if (!ak && !bk) gemm_kernel<T, false, false, true><<<grid, dim3(256), 0, stream>>>(p);
    else if (ak && !bk) gemm_kernel<T, true, false, true><<<grid, dim3(256), 0, stream>>>(p);
    else if (!ak && bk) gemm_kernel<T, false, true, true><<<grid, dim3(256), 0, stream>>>(p);
    else gemm_kernel<T, true, true, true><<<grid, dim3(256), 0, stream>>>(p);
  } else {
    if (!ak && !bk) gemm_kernel<T, false, false, false><<<grid, dim3(256), 0, stream>>>(p);
    else if (ak && !bk) gemm_kernel<T, true, false, false><<<grid, dim3(256), 0, stream>>>(p);
    else if (!ak && bk) gemm_kernel<T, false, true, false><<<grid, dim3(256), 0, stream>>>(p);
    else gemm_kernel<T, true, true, false><<<grid, dim3(256), 0, stream>>>(p);
  }
}

void GemmEngine::run(const GemmOp& op) { run_prec(op, prec); }

void GemmEngine::collect_timing() {
  for (size_t k = 0; k + 1 < evused; k += 2) {
    ECG_HIP(hipEventSynchronize(evpool[k + 1]));
    float ms = 0.f;
    ECG_HIP(hipEventElapsedTime(&ms, evpool[k], evpool[k + 1]));
    timed_ms += ms;
    timed_launches++;
  }
  evused = 0;
}

void GemmEngine::run_prec(const GemmOp& op, Prec pr) {
  if (op.M <= 0 || op.N <= 0 || op.batch <= 0) return;
  ECG_CHECK(op.nseg > 0 || op.K > 0, "gemm: K must be positive");
  if (pr == Prec::BF16X3) {
    // BF16 hi/lo products pay for packing the operands (8 bytes read and 4
    // written per element) only when the output is large in both directions;
    // skinny products (M N / (M + N) < 256) and batched ones use FP32 products.
    const double mn = (double)op.M * op.N / ((double)op.M + op.N);
    if (op.batch == 1 && mn >= 256.0) {
      run_bf16x3(op);
      return;
    }
    pr = Prec::F32;
  }
  ECG_CHECK(!transient || pr == Prec::F64, "gemm: products into a caller's scratch region are FP64 only");
  if (pr == Prec::F64 && f64_kernel == 1 && op.batch == 1 && !transient &&
      (double)op.M * op.N / ((double)op.M + op.N) >= 1024.0) {
    // large FP64 products: operands packed into FP64 planes (16 bytes moved per element)
    // for the direct-to-LDS FP64 kernel.  The packing costs about 100 / (M N / (M + N)) of
    // the product's time and the kernel gains about 14 %, hence the threshold.
    run_f64p(op);
    return;
  }
  if (pr == Prec::F32 && f32_kernel == 1 && op.batch == 1 &&
      (double)op.M * op.N / ((double)op.M + op.N) >= 256.0) {
    // large FP32 products: operands packed into FP32 planes (12 bytes moved per element)
    // for the direct-to-LDS FP32 kernel
    run_f32p(op);
    return;
  }
  ECG_CHECK(pr == Prec::F64 || pr == Prec::F32, "gemm: unknown precision");
  GemmParams p{};
  GemmSeg one;
  one.A = op.A; one.a_k = op.a_k; one.B = op.B; one.b_k = op.b_k; one.K = op.K;
  const GemmSeg* sg = op.nseg > 0 ? op.seg : &one;
  p.nseg = op.nseg > 0 ? op.nseg : 1;
  ECG_CHECK(p.nseg <= MAXSEG, "gemm: too many K segments");
  long ktot = 0;
  p.kp[0] = 0;
  for (int s = 0; s < p.nseg; ++s) {
    ECG_CHECK(sg[s].K > 0 && sg[s].A && sg[s].B, "gemm: empty K segment");
    p.sA[s] = sg[s].A; p.sak[s] = sg[s].a_k; p.sB[s] = sg[s].B; p.sbk[s] = sg[s].b_k; p.sK[s] = sg[s].K;
    p.kp[s + 1] = p.kp[s] + ((sg[s].K + BKS - 1) / BKS) * BKS;
    ktot += sg[s].K;
  }
  for (int s = p.nseg; s < MAXSEG; ++s) { p.sA[s] = sg[0].A; p.sB[s] = sg[0].B; p.kp[s + 1] = p.kp[s]; }
  flops += 2LL * op.M * op.N * ktot * op.batch;
  bytes += 8.0 * op.batch * ((double)(op.M + op.N) * ktot + (op.beta != 0.0 ? 2.0 : 1.0) * op.M * op.N);
  p.M = op.M; p.N = op.N; p.K = p.kp[p.nseg];
  p.tiles_m = (op.M + BM - 1) / BM;
  p.tiles_n = (op.N + BN - 1) / BN;
  p.a_s0 = op.a_s0; p.a_s1 = op.a_s1; p.a_z = op.a_z; p.a_d0 = op.a_d0;
  p.b_s0 = op.b_s0; p.b_s1 = op.b_s1; p.b_z = op.b_z; p.b_d0 = op.b_d0;
  p.alpha = op.alpha; p.beta = op.beta;
  // split K when the output grid alone cannot fill the CUs (about 4 workgroups per CU,
  // each with at least 8 slabs of K): a one-tile product such as x_ki (M = N = o, K = o v^2)
  // otherwise runs on a quarter of the CUs
  long tiles = (long)p.tiles_m * p.tiles_n * op.batch;
  int ksplit = 1;
  if (tiles < 2 * cus && p.K >= 4 * BKS * 8) {
    ksplit = (int)std::min<long>((4 * cus + tiles - 1) / tiles, p.K / (BKS * 8));
    ksplit = std::max(1, std::min(ksplit, 1024));
  }
  long slab = (long)op.M * op.N;
  size_t tmark = 0;
  if (transient) {
    // partial slabs come from the caller's region: split only as far as they fit there
    tmark = transient->mark();
    const size_t per = sizeof(double) * (size_t)slab * op.batch + 256;
    ksplit = (int)std::max<long>(1, std::min<long>(ksplit, (long)(transient->free_bytes() / per)));
  }
  int kchunk = (p.K + ksplit - 1) / ksplit;
  kchunk = ((kchunk + BKS - 1) / BKS) * BKS;
  ksplit = (p.K + kchunk - 1) / kchunk;
  p.kchunk = kchunk; p.ksplit = ksplit;
  double* wsp = nullptr;
  if (ksplit > 1) {
    if (transient) {
      wsp = transient->take(slab * ksplit * op.batch);
    } else {
      ws.ensure(sizeof(double) * slab * ksplit * op.batch);
      wsp = ws.d();
    }
    p.C = wsp; p.cplain = 1; p.ldc_ws = op.M; p.c_z = slab;
  } else {
    p.C = op.C; p.cplain = 0;
    p.c_ms0 = op.c_ms0; p.c_ms1 = op.c_ms1; p.c_md0 = op.c_md0;
    p.c_ns0 = op.c_ns0; p.c_ns1 = op.c_ns1; p.c_nd0 = op.c_nd0; p.c_z = op.c_z;
  }
  if (ksplit > 1) { p.alpha = 1.0; p.beta = 0.0; }
  dim3 grid(p.tiles_m * p.tiles_n, 1, op.batch * ksplit);
  ECG_CHECK(grid.z <= 65535, "gemm: batch*ksplit exceeds grid limit");
  // K-contiguous operand loads; every segment must share the contiguity of the first
  const bool ak = (sg[0].a_k == 1 && op.a_s0 != 1);
  const bool bk = (sg[0].b_k == 1 && op.b_s0 != 1);
  for (int s = 1; s < p.nseg; ++s)
    ECG_CHECK((sg[s].a_k == 1 && op.a_s0 != 1) == ak && (sg[s].b_k == 1 && op.b_s0 != 1) == bk,
              "gemm: K segments differ in operand contiguity");
  hipEvent_t e0 = timing_begin();
  if (pr == Prec::F64) launch_gemm<double>(ak, bk, p.nseg > 1, grid, stream, p);
  else launch_gemm<float>(ak, bk, p.nseg > 1, grid, stream, p);
  ECG_HIP(hipGetLastError());
  timing_end(e0);
  if (ksplit > 1) reduce_slabs(op, ksplit, wsp);
  if (transient) transient->release_to(tmark);
}

// Event pair around one GEMM kernel launch when timing is on (pairs are read
// back by collect_timing); returns the start event or null.
hipEvent_t GemmEngine::timing_begin() {
  if (!timed) return nullptr;
  while (evpool.size() < evused + 2) {
    hipEvent_t e;
    ECG_HIP(hipEventCreate(&e));
    evpool.push_back(e);
  }
  hipEvent_t e0 = evpool[evused];
  evused += 2;
  ECG_HIP(hipEventRecord(e0, stream));
  return e0;
}

void GemmEngine::timing_end(hipEvent_t e0) {
  if (e0) ECG_HIP(hipEventRecord(evpool[evused - 1], stream));
}

void GemmEngine::reduce_slabs(const GemmOp& op, int ksplit, const double* wsp) {
  const long slab = (long)op.M * op.N;
  dim3 g2((unsigned)((slab + 255) / 256), op.batch);
  hipLaunchKernelGGL(splitk_reduce_kernel, g2, dim3(256), 0, stream, wsp, ksplit, slab, op.M, op.N, op.C,
                     op.c_ms0, op.c_ms1, op.c_md0, op.c_ns0, op.c_ns1, op.c_nd0, op.c_z, op.alpha, op.beta);
  ECG_HIP(hipGetLastError());
}

}  // namespace ecg
