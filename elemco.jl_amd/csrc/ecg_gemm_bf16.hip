// BF16 hi/lo ("BF16x3") products on gfx950 MFMA.
//
// Every FP64 operand value x is split as x ~ h + l with h = bf16(x) and
// l = bf16(x - h) (|x - h - l| <= 2^-17 |x| up to rounding), and a product is
// formed from three BF16 MFMA terms h_a h_b + h_a l_b + l_a h_b accumulated in
// FP32 (the l_a l_b term, ~2^-16 relative, is dropped).  C is updated in FP64.
//
// 1. pack_bf16x3_kernel turns one operand, given by the GEMM engine's
//    split-stride (and segmented-K) addressing, into two planes P[r][k] of
//    BF16 (k fastest), zero-padded to the tile multiples in r and to 32 in each
//    K segment, through a 64x32 LDS transpose tile.
// 2. gemm_bf16x3_kernel: 256x128 output tile per 512-thread workgroup (8 waves
//    as 4x2, 64x64 per wave = 4x4 tiles of v_mfma_f32_16x16x32_bf16), K slabs of
//    32 double-buffered in LDS (16-byte chunks, XOR-swizzled so a 16-lane row
//    read touches every bank once), register prefetch two slabs ahead, FP64
//    split-stride epilogue or split-K partial slabs.
// MFMA operand layout (16x16x32 BF16, gfx950): A[i][k] in lane i + 16*(k/8),
// element k%8; B[k][j] in lane j + 16*(k/8), element k%8; D[i][j] in lane
// j + 16*(i/4), slot i%4.
//
// The same machinery serves FP32 products on packed planes (one FP32 plane per
// operand, v_mfma_f32_16x16x4f32: A[i][k] in lane i + 16*k, k < 4; D as above):
// pack_f32p_kernel stores each 32-wide K slab permuted, position 8*(k%4) + k/4
// holding k, so the eight k values a lane needs for the slab's eight MFMA steps
// (k = 4s + lane/16, s < 8) are two consecutive 16-byte chunks.
#include "ecg_gemm.h"
#include "ecg_gemm_dev.h"

#include <algorithm>

namespace ecg {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d2_t __attribute__((ext_vector_type(2)));
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

constexpr int XM = 256, XN = 128, XK = 32;   // output tile and K slab
constexpr int PR = 64;                       // pack tile rows

struct PackArgs {
  int rows, kpad, nseg;
  const double* src[MAXSEG];
  long sk[MAXSEG];
  int sK[MAXSEG], kp[MAXSEG + 1];
  long s0, s1;
  int d0;
  u16* hi;   // [rows_pad][kpad]
  u16* lo;
  float* f32;   // FP32 plane [rows_pad][kpad] (pack_f32p_kernel)
};

__device__ __forceinline__ u16 bf16_bits(float f) {
  __bf16 b = (__bf16)f;
  return __builtin_bit_cast(u16, b);
}
__device__ __forceinline__ float bf16_val(u16 u) { return (float)__builtin_bit_cast(__bf16, u); }

// 1-D grid of (rows_pad / 64) * (kpad / 32) blocks (row tile fastest), 256 threads
__global__ __launch_bounds__(256) void pack_bf16x3_kernel(PackArgs p, int nrt) {
  __shared__ u16 hs[XK][PR + 2];
  __shared__ u16 ls[XK][PR + 2];
  const int t = threadIdx.x;
  const long bid = blockIdx.x;
  const int r0 = (int)(bid % nrt) * PR, k0 = (int)(bid / nrt) * XK;
  const double* src = p.src[0];
  long sk = p.sk[0];
  int kl0 = k0, klen = p.sK[0];
#pragma unroll
  for (int q = 1; q < MAXSEG; ++q)
    if (q < p.nseg && k0 >= p.kp[q]) { src = p.src[q]; sk = p.sk[q]; kl0 = k0 - p.kp[q]; klen = p.sK[q]; }
  // read: row r0 + (t % 64) (rows fastest: coalesced for row-contiguous sources), 8 k values each
  const int row = t & (PR - 1);
  const int r = r0 + row;
  const bool rok = r < p.rows;
  const long roff = rok ? split_off(r, p.d0, p.s0, p.s1) : 0;
#pragma unroll
  for (int j = 0; j < XK / 4; ++j) {
    const int kk = (t >> 6) + 4 * j;
    const int k = kl0 + kk;
    const bool ok = rok && k < klen;
    const double x = src[ok ? roff + (long)k * sk : 0];
    const double v = ok ? x : 0.0;
    const u16 h = bf16_bits((float)v);
    const u16 l = bf16_bits((float)(v - (double)bf16_val(h)));
    hs[kk][row] = h;
    ls[kk][row] = l;
  }
  __syncthreads();
  // write: row r0 + t/4, 8 consecutive k (one 16-byte chunk) per plane
  const int wr = t >> 2, ch = t & 3;
  u32x4_t vh, vl;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    vh[e] = (unsigned)hs[8 * ch + 2 * e][wr] | ((unsigned)hs[8 * ch + 2 * e + 1][wr] << 16);
    vl[e] = (unsigned)ls[8 * ch + 2 * e][wr] | ((unsigned)ls[8 * ch + 2 * e + 1][wr] << 16);
  }
  const size_t off = (size_t)(r0 + wr) * p.kpad + k0 + 8 * ch;
  *reinterpret_cast<u32x4_t*>(p.hi + off) = vh;
  *reinterpret_cast<u32x4_t*>(p.lo + off) = vl;
}

// FP32 plane with the slab permutation (position 8*(k%4) + k/4 of each 32-wide slab holds
// k); grid, reads and K segments as pack_bf16x3_kernel.
__global__ __launch_bounds__(256) void pack_f32p_kernel(PackArgs p, int nrt) {
  __shared__ float fs[XK][PR + 1];
  const int t = threadIdx.x;
  const long bid = blockIdx.x;
  const int r0 = (int)(bid % nrt) * PR, k0 = (int)(bid / nrt) * XK;
  const double* src = p.src[0];
  long sk = p.sk[0];
  int kl0 = k0, klen = p.sK[0];
#pragma unroll
  for (int q = 1; q < MAXSEG; ++q)
    if (q < p.nseg && k0 >= p.kp[q]) { src = p.src[q]; sk = p.sk[q]; kl0 = k0 - p.kp[q]; klen = p.sK[q]; }
  const int row = t & (PR - 1);
  const int r = r0 + row;
  const bool rok = r < p.rows;
  const long roff = rok ? split_off(r, p.d0, p.s0, p.s1) : 0;
#pragma unroll
  for (int j = 0; j < XK / 4; ++j) {
    const int kk = (t >> 6) + 4 * j;
    const int k = kl0 + kk;
    const bool ok = rok && k < klen;
    const double x = src[ok ? roff + (long)k * sk : 0];
    fs[kk][row] = ok ? (float)x : 0.f;
  }
  __syncthreads();
  // write: row r0 + t/4, positions 8q .. 8q+7 (q = t%4) = k = q, q+4, ..., q+28
  const int wr = t >> 2, q = t & 3;
  f4_t v0, v1;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    v0[e] = fs[q + 4 * e][wr];
    v1[e] = fs[q + 4 * (e + 4)][wr];
  }
  float* dst = p.f32 + (size_t)(r0 + wr) * p.kpad + k0 + 8 * q;
  *reinterpret_cast<f4_t*>(dst) = v0;
  *reinterpret_cast<f4_t*>(dst + 4) = v1;
}

// FP64 plane for gemm_f64p_dma_kernel, K slabs of 16: position 4*(k%4) + k/4 of each 16-wide
// slab holds k (so a lane's four k values of a slab, k = 4 s + lane/16, are two consecutive
// 16-byte chunks); grid, reads and K segments as pack_bf16x3_kernel (a 32-wide tile = two
// slabs).  PackArgs::f32 points at the plane's bytes.
__global__ __launch_bounds__(256) void pack_f64p_kernel(PackArgs p, int nrt) {
  __shared__ double fs[XK][PR + 1];
  const int t = threadIdx.x;
  const long bid = blockIdx.x;
  const int r0 = (int)(bid % nrt) * PR, k0 = (int)(bid / nrt) * XK;
  const double* src = p.src[0];
  long sk = p.sk[0];
  int kl0 = k0, klen = p.sK[0];
#pragma unroll
  for (int q = 1; q < MAXSEG; ++q)
    if (q < p.nseg && k0 >= p.kp[q]) { src = p.src[q]; sk = p.sk[q]; kl0 = k0 - p.kp[q]; klen = p.sK[q]; }
  const int row = t & (PR - 1);
  const int r = r0 + row;
  const bool rok = r < p.rows;
  const long roff = rok ? split_off(r, p.d0, p.s0, p.s1) : 0;
#pragma unroll
  for (int j = 0; j < XK / 4; ++j) {
    const int kk = (t >> 6) + 4 * j;
    const int k = kl0 + kk;
    const bool ok = rok && k < klen;
    const double x = src[ok ? roff + (long)k * sk : 0];
    fs[kk][row] = ok ? x : 0.0;
  }
  __syncthreads();
  // write: row r0 + t/4, positions 8q .. 8q+7 of the tile: slab q/2, in-slab positions
  // 8 (q%2) + e = 4 a + b with k%16 = a + 4 b, a = 2 (q%2) + e/4, b = e%4
  const int wr = t >> 2, q = t & 3;
  double v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = fs[16 * (q >> 1) + 2 * (q & 1) + (e >> 2) + 4 * (e & 3)][wr];
  double* dst = reinterpret_cast<double*>(p.f32) + (size_t)(r0 + wr) * p.kpad + k0 + 8 * q;
#pragma unroll
  for (int e = 0; e < 8; e += 2) *reinterpret_cast<d2_t*>(dst + e) = d2_t{v[e], v[e + 1]};
}

struct X3Params {
  int M, N, kpad, kchunk, ksplit, tiles_m, tiles_n;
  // K segments: slab k0 in [kb[s], kb[s+1]) reads row m of A at Ah[s] + m*lda[s] + (k0 - kb[s]);
  // kb[s] are multiples of the slab (32; 16 in FP64 products)
  int nseg, kb[MAXSEG + 1];
  const u16* Ah[MAXSEG]; const u16* Al[MAXSEG]; long lda[MAXSEG];
  const u16* Bh[MAXSEG]; const u16* Bl[MAXSEG]; long ldb[MAXSEG];
  long az[MAXSEG], bz[MAXSEG], cbz;   // per batch index (blockIdx.y) offsets of A, B, C
  double* C; float* Cf; long c_ms0, c_ms1, c_ns0, c_ns1, c_z; int c_md0, c_nd0;
  double alpha, beta;
  int cplain; long ldc_ws;
  int box_class, box_v;   // boxed output (see PackedOp)
  // plane output (see PlaneOut): 0 none, 1 BF16 hi/lo planes, 2 FP32 plane, 3 FP64 plane
  int pout_kind, pout_dm, pout_dn, pout_tri;
  long pout_ra, pout_rb, pout_kn, pout_ld;
  void* pout_hi;
  void* pout_lo;
};

// Store x as element (m, n) of a plane output (PlaneOut): m = c + dm a', n = d + dn b' goes to
// row a' ra + b' rb, position k = c + kn d, in the layout pack_*_kernel writes for the kind:
// BF16 hi and lo planes unpermuted, with the same split of x as the packing; FP32 32-wide
// slabs with position 8 (k % 4) + (k % 32) / 4 holding k; FP64 16-wide slabs with
// 4 (k % 4) + (k % 16) / 4.
__device__ __forceinline__ void plane_put(const X3Params& p, int a, int c, int b, int d, double x) {
  if (p.pout_tri && a > b) return;
  const long row = p.pout_tri ? (long)b * (b + 1) / 2 + a : (long)a * p.pout_ra + (long)b * p.pout_rb;
  const long k = (long)c + p.pout_kn * (long)d;
  if (p.pout_kind == 3) {
    const long pos = (k & ~15L) | ((k & 3) << 2) | ((k >> 2) & 3);
    reinterpret_cast<double*>(p.pout_hi)[row * p.pout_ld + pos] = x;
  } else if (p.pout_kind == 2) {
    const long pos = (k & ~31L) | ((k & 3) << 3) | ((k >> 2) & 7);
    reinterpret_cast<float*>(p.pout_hi)[row * p.pout_ld + pos] = (float)x;
  } else {
    const u16 h = bf16_bits((float)x);
    const u16 l = bf16_bits((float)(x - (double)bf16_val(h)));
    reinterpret_cast<u16*>(p.pout_hi)[row * p.pout_ld + k] = h;
    reinterpret_cast<u16*>(p.pout_lo)[row * p.pout_ld + k] = l;
  }
}

// Offset of element (m, n) of a boxed output (PackedOp::box_class).
__device__ __forceinline__ long box_off(int m, int n, int cls, int bv) {
  const int y2 = m / bv, y1 = m - y2 * bv;
  const int x1 = cls == 1 ? n : y1, x2 = cls == 1 ? y1 : (cls == 2 ? n : y2), x3 = cls == 3 ? n : y2;
  const long nt = (bv + 7) >> 3;
  return (((long)(x1 >> 3) + nt * ((x2 >> 3) + nt * (long)(x3 >> 3))) << 9) + (x1 & 7) + ((x2 & 7) << 3) +
         ((x3 & 7) << 6);
}

// A triangular plane output (PlaneOut::tri) stores nothing of a tile whose every row has a'
// greater than every column's b': such a workgroup returns before its K loop.
__device__ __forceinline__ bool x3_tile_skipped(const X3Params& p, int bm, int bn) {
  if (!p.pout_tri) return false;
  const int nlast = min(bn + XN, p.N) - 1;
  return bm / p.pout_dm > nlast / p.pout_dn;
}

__device__ __forceinline__ int swz(int row, int ch) { return ch ^ ((row >> 1) & 3); }

// Epilogue of a 256x128 tile: element (row, col) of MFMA tile (i,j) sits in lane
// col + 16*(row/4), slot row%4 (TO: lane row + 16*(col/4), slot col%4).  Per
// column group: all 16 offsets, then all loads of C, then all stores (a load
// issued after a store to C could alias it, which would serialise every element).
// L64: the FP64 16x16x4 MFMA's layout (slot r of lane (li, lk) holds row lk + 4 r of
// column li) instead of the FP32/BF16 one (row 4 lk + r).
template <bool TO, typename A4 = f4_t, bool L64 = false, bool PO = false>
__device__ __forceinline__ void x3_epilogue(const X3Params& p, const A4 (&acc)[4][4], int bm, int bn, int wm,
                                            int wn, int li, int lk, int kz, long zb) {
  if constexpr (PO) {
    // plane output (the host launches it with TO, lanes along m): (a', c) once per row
    // block and (b', d) once per column, 32-bit divisions
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int q = L64 ? lk + 4 * r : 4 * lk + r;
          const int n = bn + wn * 64 + j * 16 + (TO ? q : li);
          const int m = bm + wm * 64 + i * 16 + (TO ? li : q);
          const int a = m / p.pout_dm, b = n / p.pout_dn;
          if (n < p.N && m < p.M)
            plane_put(p, a, m - a * p.pout_dm, b, n - b * p.pout_dn, p.alpha * (double)acc[i][j][r]);
        }
    return;
  }
  double* C = p.C + (p.cplain ? (long)kz * p.c_z : 0) + zb * p.cbz;
  float* Cf = p.Cf ? p.Cf + zb * p.cbz : nullptr;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    long off[16];
    bool ok[16];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = L64 ? lk + 4 * r : 4 * lk + r;
        const int n = bn + wn * 64 + j * 16 + (TO ? q : li);
        const int m = bm + wm * 64 + i * 16 + (TO ? li : q);
        const int e = 4 * i + r;
        ok[e] = n < p.N && m < p.M;
        long oe;
        if (p.box_class) {
          oe = box_off(m, n, p.box_class, p.box_v);
        } else {
          const long cn = p.cplain ? (long)n * p.ldc_ws : split_off(n, p.c_nd0, p.c_ns0, p.c_ns1);
          const long cm = p.cplain ? (long)m : split_off(m, p.c_md0, p.c_ms0, p.c_ms1);
          oe = cm + cn;
        }
        off[e] = ok[e] ? oe : 0;
      }
    const bool rd = p.beta != 0.0;
    if (Cf) {
      float cv[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) cv[e] = (rd && ok[e]) ? Cf[off[e]] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int e = 4 * i + r;
          if (ok[e]) Cf[off[e]] = (float)(p.alpha * (double)acc[i][j][r] + (rd ? p.beta * (double)cv[e] : 0.0));
        }
    } else {
      double cv[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) cv[e] = (rd && ok[e]) ? C[off[e]] : 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int e = 4 * i + r;
          if (ok[e]) C[off[e]] = p.alpha * (double)acc[i][j][r] + (rd ? p.beta * cv[e] : 0.0);
        }
    }
  }
}

// TO (transposed output): the MFMAs take B as their first operand, so each
// accumulator holds D^T -- lanes run along m and the four slots along n -- and
// the epilogue stores stay contiguous when m is the unit-stride output index.
// MS (multi-segment): slab k0 selects its segment's planes and row strides;
// without it all slabs read one pair of planes with the row stride kpad and the
// row offsets stay loop-invariant.
template <bool TO, bool MS>
__global__ __launch_bounds__(512, 1) void gemm_bf16x3_kernel(X3Params p) {
  __shared__ u32x4_t As[2][2][XM][4];   // [buffer][plane][row][16-byte chunk]
  __shared__ u32x4_t Bs[2][2][XN][4];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave & 3, wn = wave >> 2;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, p.tiles_m, tm, tn);
  const int bm = tm * XM, bn = tn * XN;
  const int kz = blockIdx.z;
  const long zb = blockIdx.y;   // batch index (packed products)
  const int kbeg = kz * p.kchunk;
  const int kend = min(p.kpad, kbeg + p.kchunk);

  // global chunks: A 2 planes x 256 rows x 4 chunks (4 per thread), B 2 x 128 x 4 (2 per thread).
  // Two register sets: the loads of slab s+2 are issued while slab s is multiplied, so every
  // load has two slabs of MFMA work to arrive (one slab, ~0.6 us, is shorter than the latency).
  struct Regs { u32x4_t a[4], b[2]; };
  Regs r0, r1;
  auto gload = [&](Regs& r, int k0) {
    // segment of this slab (uniform; static indices only, so the arguments stay in SGPRs)
    const u16* ah = p.Ah[0]; const u16* al = p.Al[0]; const u16* bh = p.Bh[0]; const u16* bl = p.Bl[0];
    long lda = p.kpad, ldb = p.kpad;
    int kl = k0;
    if constexpr (MS) {
      lda = p.lda[0]; ldb = p.ldb[0];
      long oa = p.az[0], ob = p.bz[0];
#pragma unroll
      for (int q = 1; q < MAXSEG; ++q)
        if (q < p.nseg && k0 >= p.kb[q]) {
          ah = p.Ah[q]; al = p.Al[q]; bh = p.Bh[q]; bl = p.Bl[q];
          lda = p.lda[q]; ldb = p.ldb[q]; kl = k0 - p.kb[q];
          oa = p.az[q]; ob = p.bz[q];
        }
      ah += zb * oa; al += zb * oa; bh += zb * ob; bl += zb * ob;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int idx = t + 512 * j;
      const int pl = idx >> 10, row = (idx >> 2) & (XM - 1), ch = idx & 3;
      const u16* base = pl ? al : ah;
      r.a[j] = *reinterpret_cast<const u32x4_t*>(base + (size_t)(bm + row) * lda + kl + 8 * ch);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int idx = t + 512 * j;
      const int pl = idx >> 9, row = (idx >> 2) & (XN - 1), ch = idx & 3;
      const u16* base = pl ? bl : bh;
      r.b[j] = *reinterpret_cast<const u32x4_t*>(base + (size_t)(bn + row) * ldb + kl + 8 * ch);
    }
  };
  auto sstore = [&](const Regs& r, int buf) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int idx = t + 512 * j;
      const int pl = idx >> 10, row = (idx >> 2) & (XM - 1), ch = idx & 3;
      As[buf][pl][row][swz(row, ch)] = r.a[j];
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int idx = t + 512 * j;
      const int pl = idx >> 9, row = (idx >> 2) & (XN - 1), ch = idx & 3;
      Bs[buf][pl][row][swz(row, ch)] = r.b[j];
    }
  };

  f4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  const int li = lane & 15, lk = lane >> 4;
  auto mfma_slab = [&](int buf) {
    bf16x8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = wm * 64 + i * 16 + li;
      ah[i] = __builtin_bit_cast(bf16x8_t, As[buf][0][r][swz(r, lk)]);
      al[i] = __builtin_bit_cast(bf16x8_t, As[buf][1][r][swz(r, lk)]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = wn * 64 + j * 16 + li;
      bh[j] = __builtin_bit_cast(bf16x8_t, Bs[buf][0][c][swz(c, lk)]);
      bl[j] = __builtin_bit_cast(bf16x8_t, Bs[buf][1][c][swz(c, lk)]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if constexpr (TO) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
        } else {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
      }
  };
  // slab k0 of this split, clamped to the first one past the end (such loads are never stored)
  auto slab = [&](int k0) { return k0 < kend ? k0 : kbeg; };
  if (kbeg < kend) {
    gload(r0, kbeg);
    sstore(r0, 0);
    gload(r1, slab(kbeg + XK));
  }
  __syncthreads();
  // slab s sits in LDS buffer s%2; register set (s+1)%2 holds slab s+1 in flight
  for (int k0 = kbeg; k0 < kend; k0 += 2 * XK) {
    gload(r0, slab(k0 + 2 * XK));
    asm volatile("" ::: "memory");   // keep the prefetch loads issued ahead of the MFMAs
    mfma_slab(0);
    const bool more1 = k0 + XK < kend;
    if (more1) sstore(r1, 1);
    __syncthreads();
    if (!more1) break;
    gload(r1, slab(k0 + 3 * XK));
    asm volatile("" ::: "memory");
    mfma_slab(1);
    if (k0 + 2 * XK < kend) sstore(r0, 0);
    __syncthreads();
  }

  x3_epilogue<TO>(p, acc, bm, bn, wm, wn, li, lk, kz, zb);
}

// Operand planes and row strides of the segment holding slab k0 (uniform; static
// argument indices only, so the arguments stay in SGPRs).
struct SlabSrc {
  const u16 *ah, *al, *bh, *bl;
  long lda, ldb;
  int kl;
};
template <bool MS>
__device__ __forceinline__ SlabSrc slab_src(const X3Params& p, int k0, long zb) {
  SlabSrc r{p.Ah[0], p.Al[0], p.Bh[0], p.Bl[0], p.kpad, p.kpad, k0};
  if constexpr (MS) {
    r.lda = p.lda[0]; r.ldb = p.ldb[0];
    long oa = p.az[0], ob = p.bz[0];
#pragma unroll
    for (int q = 1; q < MAXSEG; ++q)
      if (q < p.nseg && k0 >= p.kb[q]) {
        r.ah = p.Ah[q]; r.al = p.Al[q]; r.bh = p.Bh[q]; r.bl = p.Bl[q];
        r.lda = p.lda[q]; r.ldb = p.ldb[q]; r.kl = k0 - p.kb[q];
        oa = p.az[q]; ob = p.bz[q];
      }
    r.ah += zb * oa; r.al += zb * oa; r.bh += zb * ob; r.bl += zb * ob;
  }
  return r;
}

typedef __attribute__((address_space(3))) void* lds_vptr;

// 16 bytes per lane from global memory into LDS at M0 + 16 * lane (lane = position in
// the wave).  Issued as inline assembly: the compiler then neither tracks these loads
// (the kernel waits for them with explicit vmcnt waits) nor assumes they may alias
// every LDS read that follows (which made it wait for all loads in flight).  One
// wait state separates the M0 write from the load that reads it.  M0 is named in the
// clobber list; the AMDGPU backend treats it as a reserved register (it never keeps a
// value of its own there and warns about the clobber, silenced below), and in the built
// kernels every M0 write is one of these.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void dma16(const void* src, lds_vptr dst) {
  const unsigned m0 = (unsigned)(__UINTPTR_TYPE__)dst;
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(m0), "v"(src) : "memory", "m0");
}

// 16 bytes per lane from global memory at base + off (base uniform, in SGPRs; off a
// per-lane byte offset) into LDS at M0 + 16 * lane; otherwise as dma16.
__device__ __forceinline__ void dma16s(const void* base, unsigned off, lds_vptr dst) {
  const unsigned m0 = (unsigned)(__UINTPTR_TYPE__)dst;
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, %2" ::"s"(m0), "v"(off), "s"(base)
               : "memory", "m0");
}
#pragma clang diagnostic pop

// One slab (32 k) of both operands' hi and lo planes, HBM/L2 -> LDS: A is 32 blocks of
// 16 rows x 64 bytes (2 planes x 16), B 16 blocks; wave w moves blocks w, w+8, ... (4 of
// A, 2 of B: six 1 KB loads per wave).  Lane l of a block loads row 16 c + l/4 into slot
// l%4, i.e. chunk (l%4) ^ ((l/8)%4) of that row: a uniform block base plus a per-lane
// byte offset that is the same for every block of an operand.  With several K segments
// (MS) the segment's planes and strides change along K; the per-lane 64-bit addresses
// then cost fewer scalar registers than uniform bases of every block (which made the
// compiler spill scalar registers inside the loop).
template <bool MS>
__device__ __forceinline__ void issue_bf16x3_slab(const X3Params& p, u32x4_t (&As)[2][XM][4],
                                                  u32x4_t (&Bs)[2][XN][4], int k0, long zb, int bm, int bn,
                                                  int wave, int lane) {
  const int lrow = lane >> 2, lch = (lane & 3) ^ ((lrow >> 1) & 3);
  const SlabSrc g = slab_src<MS>(p, k0, zb);
  const unsigned offa = (unsigned)(lrow * g.lda + 8 * lch) * 2u;
  const unsigned offb = (unsigned)(lrow * g.ldb + 8 * lch) * 2u;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int c = wave + 8 * j, pl = c >> 4, r0 = 16 * (c & 15);
    const u16* pa = pl ? g.al : g.ah;
    if constexpr (MS) dma16(pa + (size_t)(bm + r0 + lrow) * g.lda + g.kl + 8 * lch, (lds_vptr)&As[pl][r0][0]);
    else dma16s(pa + (size_t)(bm + r0) * g.lda + g.kl, offa, (lds_vptr)&As[pl][r0][0]);
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int c = wave + 8 * j, pl = c >> 3, r0 = 16 * (c & 7);
    const u16* pb = pl ? g.bl : g.bh;
    if constexpr (MS) dma16(pb + (size_t)(bn + r0 + lrow) * g.ldb + g.kl + 8 * lch, (lds_vptr)&Bs[pl][r0][0]);
    else dma16s(pb + (size_t)(bn + r0) * g.ldb + g.kl, offb, (lds_vptr)&Bs[pl][r0][0]);
  }
}

// The same product with the operand slabs moved HBM/L2 -> LDS by the load unit
// (global_load_lds_dwordx4: lane l of a wave writes 16 bytes at M0 + 16 l), three
// LDS buffers in rotation: slab s+2 is in flight while slab s is multiplied, with
// no register staging.  Each lane fetches the chunk that its LDS slot holds under
// the XOR swizzle, so the LDS layout (and the MFMA reads) are those of
// gemm_bf16x3_kernel.  The three buffers are separate LDS objects and the loop is
// unrolled by three, so the compiler sees which buffer a DMA writes and which one
// the MFMAs read (no conservative wait on the loads in flight).
template <bool TO, bool MS, bool PO = false>
__global__ __launch_bounds__(512, 1) void gemm_bf16x3_dma_kernel(X3Params p) {
  __shared__ u32x4_t A0[2][XM][4], A1[2][XM][4], A2[2][XM][4];   // [plane][row][16-byte chunk]
  __shared__ u32x4_t B0[2][XN][4], B1[2][XN][4], B2[2][XN][4];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);   // wave index in an SGPR
  const int wm = wave & 3, wn = wave >> 2;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, p.tiles_m, tm, tn);
  const int bm = tm * XM, bn = tn * XN;
  if (PO && x3_tile_skipped(p, bm, bn)) return;
  const int kz = blockIdx.z;
  const long zb = blockIdx.y;
  const int kbeg = kz * p.kchunk;
  const int kend = min(p.kpad, kbeg + p.kchunk);

  // slab k0 -> one buffer (issue_bf16x3_slab)
  auto issue = [&](u32x4_t (&As)[2][XM][4], u32x4_t (&Bs)[2][XN][4], int k0) {
    issue_bf16x3_slab<MS>(p, As, Bs, k0, zb, bm, bn, wave, lane);
  };

  f4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
  const int li = lane & 15, lk = lane >> 4;
  auto compute = [&](const u32x4_t (&As)[2][XM][4], const u32x4_t (&Bs)[2][XN][4]) {
    bf16x8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = wm * 64 + i * 16 + li;
      ah[i] = __builtin_bit_cast(bf16x8_t, As[0][r][swz(r, lk)]);
      al[i] = __builtin_bit_cast(bf16x8_t, As[1][r][swz(r, lk)]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = wn * 64 + j * 16 + li;
      bh[j] = __builtin_bit_cast(bf16x8_t, Bs[0][c][swz(c, lk)]);
      bl[j] = __builtin_bit_cast(bf16x8_t, Bs[1][c][swz(c, lk)]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if constexpr (TO) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
        } else {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
      }
  };
  // slab k0 of this split, clamped to the first one past the end (such loads fill a
  // buffer that is never multiplied)
  auto slab = [&](int k0) { return k0 < kend ? k0 : kbeg; };
  // this wave's loads of the slab issued before the most recent one have landed
  auto wait_prev = []() { asm volatile("s_waitcnt vmcnt(6)" ::: "memory"); };
  // workgroup barrier without the memory fence of the HIP barrier (which waits for all
  // loads, including the slab still in flight): the slab made visible by it is the one
  // wait_prev covers, and every LDS read of the buffer reloaded next has completed (the
  // MFMAs used it); the empty asm statements keep the compiler from moving LDS accesses
  // across it
  auto barrier = []() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };
  if (kbeg < kend) {
    issue(A0, B0, kbeg);
    issue(A1, B1, slab(kbeg + XK));
    wait_prev();
  }
  barrier();
  for (int k0 = kbeg; k0 < kend; k0 += 3 * XK) {
    issue(A2, B2, slab(k0 + 2 * XK));
    compute(A0, B0);
    wait_prev();
    barrier();
    if (k0 + XK >= kend) break;
    issue(A0, B0, slab(k0 + 3 * XK));
    compute(A1, B1);
    wait_prev();
    barrier();
    if (k0 + 2 * XK >= kend) break;
    issue(A1, B1, slab(k0 + 4 * XK));
    compute(A2, B2);
    wait_prev();
    barrier();
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // no load into LDS outlives the loop
  x3_epilogue<TO, f4_t, false, PO>(p, acc, bm, bn, wm, wn, li, lk, kz, zb);
}

// FP32 planes and row strides (in floats) of the segment holding slab k0: the packed
// FP32 product keeps its planes in X3Params::Ah / Bh (as FP32 pointers).
template <typename E>
struct SlabSrcT {
  const E *a, *b;
  long lda, ldb;
  int kl;
};
using SlabSrcF = SlabSrcT<float>;
// slab k0 of a packed FP32 / FP64 product: plane pointers, row strides and in-segment k of
// the K segment holding k0 (one segment unless MS), advanced to batch member zb
template <bool MS, typename E = float>
__device__ __forceinline__ SlabSrcT<E> slab_src_planes(const X3Params& p, int k0, long zb) {
  SlabSrcT<E> r{reinterpret_cast<const E*>(p.Ah[0]), reinterpret_cast<const E*>(p.Bh[0]), p.kpad, p.kpad, k0};
  if constexpr (MS) {
    r.lda = p.lda[0]; r.ldb = p.ldb[0];
    long oa = p.az[0], ob = p.bz[0];
#pragma unroll
    for (int q = 1; q < MAXSEG; ++q)
      if (q < p.nseg && k0 >= p.kb[q]) {
        r.a = reinterpret_cast<const E*>(p.Ah[q]); r.b = reinterpret_cast<const E*>(p.Bh[q]);
        r.lda = p.lda[q]; r.ldb = p.ldb[q]; r.kl = k0 - p.kb[q];
        oa = p.az[q]; ob = p.bz[q];
      }
    r.a += zb * oa; r.b += zb * ob;
  }
  return r;
}
template <bool MS>
__device__ __forceinline__ SlabSrcF slab_src_f32(const X3Params& p, int k0, long zb) {
  return slab_src_planes<MS, float>(p, k0, zb);
}

// Slot of chunk c of row r in a 128-byte LDS row of the FP32 / FP64 plane kernels:
// c ^ plane_swz(r), plane_swz a permutation of r % 8 that swaps its two low bits.  A
// fragment read (ds_read_b128, lane l: row r0 + l % 16, chunk 2 (l / 16) + h) is served in
// passes of 16 lanes, {0..7, 16..23} and its shifts by 8 and 32, i.e. rows r..r+7 at chunks c
// and c ^ 2, over 256 bytes of banks (two rows).  The four even rows of a pass get swizzles
// {0, 1, 4, 5} and the odd rows {2, 3, 6, 7}: no two of one parity have XOR 0 or 2, so the
// 16 lanes hit 16 distinct 16-byte bank groups (the plain r % 8 swizzle leaves rows r, r+2
// of a pass in one group: a two-way conflict on every read).  Passes of eight consecutive
// lanes, or of lanes {0..3, 16..19, 32..35, 48..51}, are conflict-free under it as well.
__device__ __forceinline__ int plane_swz(int r) { return (r & 4) | ((r & 1) << 1) | ((r >> 1) & 1); }

// FP32 product on packed FP32 planes (pack_f32p_kernel): the tile, waves, LDS buffers,
// direct HBM -> LDS slab loads and epilogue of gemm_bf16x3_dma_kernel; a slab row is 32
// floats = 8 chunks of 16 bytes, chunk c of row r stored at slot c ^ plane_swz(r).  Per slab
// and wave: 16 reads of 16 bytes and 128 v_mfma_f32_16x16x4f32 (eight k steps of the 4x4
// tiles).
template <bool TO, bool MS, bool PO = false>
__global__ __launch_bounds__(512, 1) void gemm_f32p_dma_kernel(X3Params p) {
  __shared__ u32x4_t A0[XM][8], A1[XM][8], A2[XM][8];   // [row][16-byte chunk]
  __shared__ u32x4_t B0[XN][8], B1[XN][8], B2[XN][8];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);   // wave index in an SGPR
  const int wm = wave & 3, wn = wave >> 2;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, p.tiles_m, tm, tn);
  const int bm = tm * XM, bn = tn * XN;
  if (PO && x3_tile_skipped(p, bm, bn)) return;
  const int kz = blockIdx.z;
  const long zb = blockIdx.y;
  const int kbeg = kz * p.kchunk;
  const int kend = min(p.kpad, kbeg + p.kchunk);

  // slab k0 -> one buffer: A is 32 blocks of 8 rows x 128 bytes, B 16 blocks; wave w moves
  // blocks w, w+8, ... (4 of A, 2 of B).  Lane l of a block loads row 8 q + l/8 into slot
  // l%8, i.e. the chunk that the swizzle stores there: a uniform block base plus a per-lane
  // byte offset that is the same for every block of a wave.
  // row r of a block (r = 8 q + l/8, q = wave + 8 j) has r % 8 = l/8
  const int lrow = lane >> 3, lch = (lane & 7) ^ plane_swz(lrow);
  auto issue = [&](u32x4_t (&As)[XM][8], u32x4_t (&Bs)[XN][8], int k0) {
    const SlabSrcF g = slab_src_f32<MS>(p, k0, zb);
    const unsigned offa = (unsigned)(lrow * g.lda + 4 * lch) * 4u;
    const unsigned offb = (unsigned)(lrow * g.ldb + 4 * lch) * 4u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r0 = 8 * (wave + 8 * j);
      dma16s(g.a + (size_t)(bm + r0) * g.lda + g.kl, offa, (lds_vptr)&As[r0][0]);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int r0 = 8 * (wave + 8 * j);
      dma16s(g.b + (size_t)(bn + r0) * g.ldb + g.kl, offb, (lds_vptr)&Bs[r0][0]);
    }
  };

  f4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
  const int li = lane & 15, lk = lane >> 4;
  const int sw = plane_swz(li);   // = plane_swz(r) of every fragment row r (r = li mod 16)
  // a slab in two halves of four k steps (chunks 2 lk + h): 32 fragment registers live
  auto compute = [&](const u32x4_t (&As)[XM][8], const u32x4_t (&Bs)[XN][8]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      f4_t a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = wm * 64 + i * 16 + li;
        a[i] = __builtin_bit_cast(f4_t, As[r][(2 * lk + h) ^ sw]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = wn * 64 + j * 16 + li;
        b[j] = __builtin_bit_cast(f4_t, Bs[c][(2 * lk + h) ^ sw]);
      }
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            if constexpr (TO) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[j][s], a[i][s], acc[i][j], 0, 0, 0);
            else acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][s], b[j][s], acc[i][j], 0, 0, 0);
          }
    }
  };
  auto slab = [&](int k0) { return k0 < kend ? k0 : kbeg; };
  auto wait_prev = []() { asm volatile("s_waitcnt vmcnt(6)" ::: "memory"); };
  auto barrier = []() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };
  if (kbeg < kend) {
    issue(A0, B0, kbeg);
    issue(A1, B1, slab(kbeg + XK));
    wait_prev();
  }
  barrier();
  for (int k0 = kbeg; k0 < kend; k0 += 3 * XK) {
    issue(A2, B2, slab(k0 + 2 * XK));
    compute(A0, B0);
    wait_prev();
    barrier();
    if (k0 + XK >= kend) break;
    issue(A0, B0, slab(k0 + 3 * XK));
    compute(A1, B1);
    wait_prev();
    barrier();
    if (k0 + 2 * XK >= kend) break;
    issue(A1, B1, slab(k0 + 4 * XK));
    compute(A2, B2);
    wait_prev();
    barrier();
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // no load into LDS outlives the loop
  x3_epilogue<TO, f4_t, false, PO>(p, acc, bm, bn, wm, wn, li, lk, kz, zb);
}

constexpr int XK64 = 16;   // K slab of the FP64 packed product (128 bytes per row, as FP32's 32)

// FP64 product on packed FP64 planes (pack_f64p_kernel; one plane per operand, single K
// segment, plane pointers in X3Params::Ah / Bh, row strides kpad doubles): the tile, waves,
// three LDS buffers, SGPR-base direct-to-LDS slab loads and swizzle of gemm_f32p_dma_kernel
// with 16-wide K slabs of doubles and v_mfma_f64_16x16x4f64 (64 per slab and wave, as
// many MFMA cycles as the FP32 kernel's 128).
template <bool TO, bool MS = false, bool PO = false>
__global__ __launch_bounds__(512, 1) void gemm_f64p_dma_kernel(X3Params p) {
  __shared__ u32x4_t As[3][XM][8];   // [buffer][row][16-byte chunk]
  __shared__ u32x4_t Bs[3][XN][8];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);   // wave index in an SGPR
  const int wm = wave & 3, wn = wave >> 2;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, p.tiles_m, tm, tn);
  const int bm = tm * XM, bn = tn * XN;
  if (PO && x3_tile_skipped(p, bm, bn)) return;
  const int kz = blockIdx.z;
  const long zb = blockIdx.y;
  const int kbeg = kz * p.kchunk;
  const int kend = min(p.kpad, kbeg + p.kchunk);
  const double* pa = reinterpret_cast<const double*>(p.Ah[0]);
  const double* pb = reinterpret_cast<const double*>(p.Bh[0]);
  const long ld = p.kpad;

  // slab k0 -> one buffer: blocks of 8 rows x 128 bytes as in gemm_f32p_dma_kernel
  const int lrow = lane >> 3, lch = (lane & 7) ^ plane_swz(lrow);
  const unsigned off = (unsigned)(lrow * ld + 2 * lch) * 8u;
  auto issue = [&](int buf, int k0) {
    if constexpr (MS) {
      // K segments and batch members (PackedOp with f64): per-slab plane, stride and offset
      const SlabSrcT<double> g = slab_src_planes<true, double>(p, k0, zb);
      const unsigned offa = (unsigned)(lrow * g.lda + 2 * lch) * 8u;
      const unsigned offb = (unsigned)(lrow * g.ldb + 2 * lch) * 8u;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r0 = 8 * (wave + 8 * j);
        dma16s(g.a + (size_t)(bm + r0) * g.lda + g.kl, offa, (lds_vptr)&As[buf][r0][0]);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int r0 = 8 * (wave + 8 * j);
        dma16s(g.b + (size_t)(bn + r0) * g.ldb + g.kl, offb, (lds_vptr)&Bs[buf][r0][0]);
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r0 = 8 * (wave + 8 * j);
        dma16s(pa + (size_t)(bm + r0) * ld + k0, off, (lds_vptr)&As[buf][r0][0]);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int r0 = 8 * (wave + 8 * j);
        dma16s(pb + (size_t)(bn + r0) * ld + k0, off, (lds_vptr)&Bs[buf][r0][0]);
      }
    }
  };

  d4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
  const int li = lane & 15, lk = lane >> 4;
  const int sw = plane_swz(li);   // = plane_swz(r) of every fragment row r (r = li mod 16)
  // a slab in two halves of two k steps (chunk 2 lk + h holds k steps 2 h and 2 h + 1)
  auto compute = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      d2_t a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = wm * 64 + i * 16 + li;
        a[i] = __builtin_bit_cast(d2_t, As[buf][r][(2 * lk + h) ^ sw]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = wn * 64 + j * 16 + li;
        b[j] = __builtin_bit_cast(d2_t, Bs[buf][c][(2 * lk + h) ^ sw]);
      }
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            if constexpr (TO) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j][s], a[i][s], acc[i][j], 0, 0, 0);
            else acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][s], b[j][s], acc[i][j], 0, 0, 0);
          }
    }
  };
  auto slab = [&](int k0) { return k0 < kend ? k0 : kbeg; };
  auto wait_prev = []() { asm volatile("s_waitcnt vmcnt(6)" ::: "memory"); };
  auto barrier = []() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };
  // one slab per iteration with the buffer index a plain variable (the loads are inline
  // assembly the compiler does not track): the accumulators keep their registers across
  // iterations, which the three-way unrolled loop of the FP32 kernel does not afford here
  if (kbeg < kend) {
    issue(0, kbeg);
    issue(1, slab(kbeg + XK64));
    wait_prev();
  }
  barrier();
  int buf = 0;
#pragma nounroll
  for (int k0 = kbeg; k0 < kend; k0 += XK64) {
    issue(buf == 0 ? 2 : buf - 1, slab(k0 + 2 * XK64));   // the buffer of slab k0 - XK64
    compute(buf);
    wait_prev();
    barrier();
    buf = buf == 2 ? 0 : buf + 1;
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // no load into LDS outlives the loop
  x3_epilogue<TO, d4_t, true, PO>(p, acc, bm, bn, wm, wn, li, lk, kz, zb);
}

}  // namespace

void GemmEngine::run_bf16x3(const GemmOp& op) { run_planes(op, 0); }

void GemmEngine::run_f32p(const GemmOp& op) { run_planes(op, 1); }

void GemmEngine::run_f64p(const GemmOp& op) { run_planes(op, 2); }

// Operands packed on the fly (kind 0: BF16 hi/lo planes; 1: one FP32 plane; 2: one FP64
// plane), then the packed-plane kernel; split K when the tiles alone cannot fill the GPU.
void GemmEngine::run_planes(const GemmOp& op, int kind, const PlaneOut* po) {
  ECG_CHECK(op.batch == 1, "gemm: batched packed products are not supported");
  ECG_CHECK(!transient, "gemm: no packed products into a caller's scratch region");
  GemmSeg one;
  one.A = op.A; one.a_k = op.a_k; one.B = op.B; one.b_k = op.b_k; one.K = op.K;
  const GemmSeg* sg = op.nseg > 0 ? op.seg : &one;
  const int nseg = op.nseg > 0 ? op.nseg : 1;
  ECG_CHECK(nseg <= MAXSEG, "gemm: too many K segments");
  PackArgs pa{}, pb{};
  pa.nseg = pb.nseg = nseg;
  int kpad = 0;
  long ktot = 0;
  pa.kp[0] = pb.kp[0] = 0;
  for (int s = 0; s < nseg; ++s) {
    ECG_CHECK(sg[s].K > 0 && (sg[s].A || op.a_planes) && sg[s].B, "gemm: empty K segment");
    pa.src[s] = sg[s].A; pa.sk[s] = sg[s].a_k; pa.sK[s] = sg[s].K;
    pb.src[s] = sg[s].B; pb.sk[s] = sg[s].b_k; pb.sK[s] = sg[s].K;
    kpad += ((sg[s].K + XK - 1) / XK) * XK;
    pa.kp[s + 1] = pb.kp[s + 1] = kpad;
    ktot += sg[s].K;
  }
  for (int s = nseg; s < MAXSEG; ++s) {
    pa.src[s] = pa.src[0]; pb.src[s] = pb.src[0];
    pa.kp[s + 1] = pb.kp[s + 1] = kpad;
  }
  const int tiles_m = (op.M + XM - 1) / XM, tiles_n = (op.N + XN - 1) / XN;
  long long mn = (long long)op.M * op.N;
  if (po && po->tri) {
    // the tiles a triangular plane output computes (x3_tile_skipped)
    mn = 0;
    for (int tm = 0; tm < tiles_m; ++tm)
      for (int tn = 0; tn < tiles_n; ++tn) {
        const int nlast = std::min((tn + 1) * XN, op.N) - 1;
        if ((tm * XM) / po->dm > nlast / po->dn) continue;
        mn += (long long)(std::min((tm + 1) * XM, op.M) - tm * XM) * (nlast + 1 - tn * XN);
      }
  }
  flops += 2LL * mn * ktot;
  bytes += 8.0 * ((double)(op.M + op.N) * ktot + (op.beta != 0.0 ? 2.0 : 1.0) * op.M * op.N);
  const long Mp = (long)tiles_m * XM, Np = (long)tiles_n * XN;
  // bytes per element of a packed operand (two BF16 planes or one FP32 plane: 4; FP64: 8)
  const size_t esz = kind == 2 ? 8 : 4;
  // A given as planes: no A packing, no room for it in pk
  const bool pack_a = op.a_planes == nullptr;
  ECG_CHECK(pack_a || (nseg == 1 && op.a_planes_ld == kpad && (kind != 0 || op.a_planes_lo)),
            "gemm: packed A planes need one K segment, rows of K padded to 32 and (BF16) both planes");
  ECG_CHECK(!po || (op.beta == 0.0 && po->hi && (kind != 0 || po->lo) && po->dm > 0 && po->dn > 0),
            "gemm: plane output needs beta = 0 and its planes");
  const long Ma = pack_a ? Mp : 0;
  char *a_base, *b_base;
  bool pack_b = true;
  if (op.b_keep) {
    // B planes kept between products: reuse them when B's addressing is the kept one
    std::vector<long> key{kind, nseg, op.N, kpad, op.b_s0, op.b_s1, op.b_d0};
    for (int s = 0; s < nseg; ++s) {
      key.push_back((long)(__UINTPTR_TYPE__)sg[s].B);
      key.push_back(sg[s].b_k);
      key.push_back(sg[s].K);
    }
    if (key == keep_key) {
      pack_b = false;
    } else {
      keep_key.clear();   // invalid while the planes are being replaced
      pk_keep.ensure(esz * (size_t)Np * kpad);
      keep_key = key;
    }
    if (pack_a) pk.ensure(esz * (size_t)Mp * kpad);
    a_base = static_cast<char*>(pk.p);
    b_base = static_cast<char*>(pk_keep.p);
  } else {
    pk.ensure(esz * (size_t)(Ma + Np) * kpad);
    a_base = static_cast<char*>(pk.p);
    b_base = a_base + esz * (size_t)Ma * kpad;
  }
  u16* Ah = pack_a ? reinterpret_cast<u16*>(a_base) : const_cast<u16*>(static_cast<const u16*>(op.a_planes));
  u16* Al = pack_a ? Ah + (size_t)Mp * kpad   // BF16 lo plane (kind 0)
                   : const_cast<u16*>(static_cast<const u16*>(op.a_planes_lo));
  u16* Bh = reinterpret_cast<u16*>(b_base);
  u16* Bl = Bh + (size_t)Np * kpad;
  // an FP32 or FP64 plane starts at the hi plane's address, rows of kpad elements
  pa.rows = op.M; pa.kpad = kpad; pa.s0 = op.a_s0; pa.s1 = op.a_s1; pa.d0 = op.a_d0; pa.hi = Ah; pa.lo = Al;
  pb.rows = op.N; pb.kpad = kpad; pb.s0 = op.b_s0; pb.s1 = op.b_s1; pb.d0 = op.b_d0; pb.hi = Bh; pb.lo = Bl;
  pa.f32 = reinterpret_cast<float*>(Ah);
  pb.f32 = reinterpret_cast<float*>(Bh);
  const long nba = (Mp / PR) * (long)(kpad / XK), nbb = (Np / PR) * (long)(kpad / XK);
  ECG_CHECK(nba < (1L << 31) && nbb < (1L << 31), "gemm: operand too large for the packing grid");
  if (kind == 2) {
    if (pack_a) hipLaunchKernelGGL(pack_f64p_kernel, dim3((unsigned)nba), dim3(256), 0, stream, pa, (int)(Mp / PR));
    if (pack_b) hipLaunchKernelGGL(pack_f64p_kernel, dim3((unsigned)nbb), dim3(256), 0, stream, pb, (int)(Np / PR));
  } else if (kind == 1) {
    if (pack_a) hipLaunchKernelGGL(pack_f32p_kernel, dim3((unsigned)nba), dim3(256), 0, stream, pa, (int)(Mp / PR));
    if (pack_b) hipLaunchKernelGGL(pack_f32p_kernel, dim3((unsigned)nbb), dim3(256), 0, stream, pb, (int)(Np / PR));
  } else {
    if (pack_a) hipLaunchKernelGGL(pack_bf16x3_kernel, dim3((unsigned)nba), dim3(256), 0, stream, pa, (int)(Mp / PR));
    if (pack_b) hipLaunchKernelGGL(pack_bf16x3_kernel, dim3((unsigned)nbb), dim3(256), 0, stream, pb, (int)(Np / PR));
  }
  ECG_HIP(hipGetLastError());

  X3Params x{};
  x.M = op.M; x.N = op.N; x.kpad = kpad; x.tiles_m = tiles_m; x.tiles_n = tiles_n;
  x.nseg = 1; x.kb[0] = 0;
  for (int s = 1; s <= MAXSEG; ++s) x.kb[s] = kpad;
  for (int s = 0; s < MAXSEG; ++s) {
    x.Ah[s] = Ah; x.Al[s] = Al; x.Bh[s] = Bh; x.Bl[s] = Bl; x.lda[s] = kpad; x.ldb[s] = kpad;
  }
  x.alpha = op.alpha; x.beta = op.beta;
  // split K when the tiles alone cannot occupy the CUs (one workgroup per CU; `cus` of them,
  // 256 unless several contexts share the device); for deep products with a few rounds of
  // tiles, split as far as it cuts the partly filled last round (rounds per unit of work
  // ceil(tiles ks / cus) / ks, at least 4096 of K per split, at least 5 % gained): the
  // ladder's N = o^2 contractions have K = v^2
  const long tiles = (long)tiles_m * tiles_n;
  int ksplit = 1;
  if (po) {
    // plane output: no partial sums
  } else if (tiles < cus && kpad >= 8 * XK) {
    ksplit = (int)std::min<long>((2 * cus + tiles - 1) / tiles, kpad / (4 * XK));
    ksplit = std::max(1, std::min(ksplit, 64));
  } else if (tiles < 16 * cus && kpad >= 2 * 4096) {
    auto rounds = [&](int ks) { return (double)((tiles * ks + cus - 1) / cus) / ks; };
    double best = rounds(1);
    for (int ks = 2; ks <= 8 && kpad / ks >= 4096; ++ks)
      if (rounds(ks) < 0.95 * best) { best = rounds(ks); ksplit = ks; }
  }
  int kchunk = (kpad + ksplit - 1) / ksplit;
  kchunk = ((kchunk + XK - 1) / XK) * XK;
  ksplit = (kpad + kchunk - 1) / kchunk;
  x.kchunk = kchunk; x.ksplit = ksplit;
  const long slab = (long)op.M * op.N;
  if (ksplit > 1) {
    ws.ensure(sizeof(double) * slab * ksplit);
    x.C = ws.d(); x.cplain = 1; x.ldc_ws = op.M; x.c_z = slab;
    x.alpha = 1.0; x.beta = 0.0;
  } else {
    x.C = op.C; x.cplain = 0;
    x.c_ms0 = op.c_ms0; x.c_ms1 = op.c_ms1; x.c_md0 = op.c_md0;
    x.c_ns0 = op.c_ns0; x.c_ns1 = op.c_ns1; x.c_nd0 = op.c_nd0;
  }
  ECG_CHECK(ksplit <= 65535, "gemm: split-K grid limit");
  hipEvent_t e0 = timing_begin();
  const dim3 grid((unsigned)tiles, 1, (unsigned)ksplit);
  // transposed accumulators (lanes along m) when m is the output's unit stride, as in the
  // split-K partial slabs: a store instruction then writes 4 whole 128-byte rows of C
  // instead of touching 16
  const bool to = po || x.cplain || (x.c_ms0 == 1 && x.c_ns0 != 1);
  if (po) {
    x.pout_kind = kind == 2 ? 3 : (kind == 1 ? 2 : 1);
    x.pout_dm = po->dm; x.pout_dn = po->dn;
    x.pout_ra = po->ra; x.pout_rb = po->rb; x.pout_kn = po->kn; x.pout_ld = po->ld;
    x.pout_hi = po->hi; x.pout_lo = po->lo; x.pout_tri = po->tri;
    // plane output: lanes along m, the direct-to-LDS kernels
    if (kind == 2) hipLaunchKernelGGL((gemm_f64p_dma_kernel<true, false, true>), grid, dim3(512), 0, stream, x);
    else if (kind == 1) hipLaunchKernelGGL((gemm_f32p_dma_kernel<true, false, true>), grid, dim3(512), 0, stream, x);
    else hipLaunchKernelGGL((gemm_bf16x3_dma_kernel<true, false, true>), grid, dim3(512), 0, stream, x);
  } else if (kind == 2) {
    if (to) hipLaunchKernelGGL((gemm_f64p_dma_kernel<true>), grid, dim3(512), 0, stream, x);
    else hipLaunchKernelGGL((gemm_f64p_dma_kernel<false>), grid, dim3(512), 0, stream, x);
  } else if (kind == 1) {
    if (to) hipLaunchKernelGGL((gemm_f32p_dma_kernel<true, false>), grid, dim3(512), 0, stream, x);
    else hipLaunchKernelGGL((gemm_f32p_dma_kernel<false, false>), grid, dim3(512), 0, stream, x);
  } else if (bf16_kernel == 1) {
    if (to) hipLaunchKernelGGL((gemm_bf16x3_dma_kernel<true, false>), grid, dim3(512), 0, stream, x);
    else hipLaunchKernelGGL((gemm_bf16x3_dma_kernel<false, false>), grid, dim3(512), 0, stream, x);
  } else {
    if (to) hipLaunchKernelGGL((gemm_bf16x3_kernel<true, false>), grid, dim3(512), 0, stream, x);
    else hipLaunchKernelGGL((gemm_bf16x3_kernel<false, false>), grid, dim3(512), 0, stream, x);
  }
  ECG_HIP(hipGetLastError());
  timing_end(e0);
  if (ksplit > 1) reduce_slabs(op, ksplit, ws.d());
}

void GemmEngine::pack_rows_bf16x3(const double* src, long rows, int d0, long s0, long s1, int K, long sk, int kpad,
                                  uint16_t* hi, uint16_t* lo) {
  ECG_CHECK(K > 0 && kpad % XK == 0 && kpad >= K, "pack: kpad must be a multiple of 32 covering K");
  const long rt = (rows + PR - 1) / PR;
  ECG_CHECK(rt * PR < (1L << 31) && rt * (kpad / XK) < (1L << 31), "pack: operand too large for the packing grid");
  PackArgs pa{};
  pa.rows = (int)rows; pa.kpad = kpad; pa.nseg = 1;
  for (int s = 0; s < MAXSEG; ++s) { pa.src[s] = src; pa.sk[s] = sk; pa.sK[s] = K; }
  pa.kp[0] = 0;
  for (int s = 1; s <= MAXSEG; ++s) pa.kp[s] = kpad;
  pa.s0 = s0; pa.s1 = s1; pa.d0 = d0; pa.hi = hi; pa.lo = lo;
  hipLaunchKernelGGL(pack_bf16x3_kernel, dim3((unsigned)(rt * (kpad / XK))), dim3(256), 0, stream, pa, (int)rt);
  ECG_HIP(hipGetLastError());
}

void GemmEngine::pack_rows_f32(const double* src, long rows, int d0, long s0, long s1, int K, long sk, int kpad,
                               float* out) {
  ECG_CHECK(K > 0 && kpad % XK == 0 && kpad >= K, "pack: kpad must be a multiple of 32 covering K");
  const long rt = (rows + PR - 1) / PR;
  ECG_CHECK(rt * PR < (1L << 31) && rt * (kpad / XK) < (1L << 31), "pack: operand too large for the packing grid");
  PackArgs pa{};
  pa.rows = (int)rows; pa.kpad = kpad; pa.nseg = 1;
  for (int s = 0; s < MAXSEG; ++s) { pa.src[s] = src; pa.sk[s] = sk; pa.sK[s] = K; }
  pa.kp[0] = 0;
  for (int s = 1; s <= MAXSEG; ++s) pa.kp[s] = kpad;
  pa.s0 = s0; pa.s1 = s1; pa.d0 = d0; pa.f32 = out;
  hipLaunchKernelGGL(pack_f32p_kernel, dim3((unsigned)(rt * (kpad / XK))), dim3(256), 0, stream, pa, (int)rt);
  ECG_HIP(hipGetLastError());
}

void GemmEngine::pack_rows_f64(const double* src, long rows, int d0, long s0, long s1, int K, long sk, int kpad,
                               double* out) {
  ECG_CHECK(K > 0 && kpad % XK == 0 && kpad >= K, "pack: kpad must be a multiple of 32 covering K");
  const long rt = (rows + PR - 1) / PR;
  ECG_CHECK(rt * PR < (1L << 31) && rt * (kpad / XK) < (1L << 31), "pack: operand too large for the packing grid");
  PackArgs pa{};
  pa.rows = (int)rows; pa.kpad = kpad; pa.nseg = 1;
  for (int s = 0; s < MAXSEG; ++s) { pa.src[s] = src; pa.sk[s] = sk; pa.sK[s] = K; }
  pa.kp[0] = 0;
  for (int s = 1; s <= MAXSEG; ++s) pa.kp[s] = kpad;
  pa.s0 = s0; pa.s1 = s1; pa.d0 = d0; pa.f32 = reinterpret_cast<float*>(out);
  hipLaunchKernelGGL(pack_f64p_kernel, dim3((unsigned)(rt * (kpad / XK))), dim3(256), 0, stream, pa, (int)rt);
  ECG_HIP(hipGetLastError());
}

void GemmEngine::run_bf16x3_packed(const PackedOp& op) {
  ECG_CHECK(op.nseg >= 1 && op.nseg <= MAXSEG, "gemm: packed product needs 1..4 K segments");
  ECG_CHECK((op.C != nullptr) != (op.Cf != nullptr), "gemm: packed product needs exactly one output");
  X3Params x{};
  x.M = op.M; x.N = op.N;
  x.tiles_m = (op.M + XM - 1) / XM; x.tiles_n = (op.N + XN - 1) / XN;
  x.nseg = op.nseg;
  // segment lengths in the concatenated K: FP64 products run slabs of 16, so an FP64
  // segment ends at its data length rounded up to 16; BF16 hi/lo and FP32 products keep
  // whole 32-wide slabs per segment (a BF16 slab that straddles two segments, addressed
  // per lane, measured 11 % slower per product than the 10 % of K it saves); the last
  // slab's tail past the last segment must lie in that segment's zero padding, else every
  // segment keeps its full row length
  const int gran = op.f64 ? 16 : XK, slabk = gran;
  int klen[MAXSEG];
  int ktot = 0;
  for (int s = 0; s < op.nseg; ++s) {
    const PackedSeg& g = op.seg[s];
    ECG_CHECK(g.K > 0 && g.K % XK == 0 && g.Kt >= 0 && g.Kt <= g.K && g.Ah && g.Bh &&
                  (op.f32 || op.f64 || (g.Al && g.Bl)) && g.lda >= g.K && g.ldb >= g.K,
              "gemm: packed segment needs K a positive multiple of 32, Kt <= K and row strides >= K");
    const int kt = g.Kt > 0 ? g.Kt : g.K;
    klen[s] = ((kt + gran - 1) / gran) * gran;
    ktot += klen[s];
  }
  int kpad = ((ktot + slabk - 1) / slabk) * slabk;
  if (klen[op.nseg - 1] + (kpad - ktot) > op.seg[op.nseg - 1].K) {
    kpad = 0;
    for (int s = 0; s < op.nseg; ++s) kpad += (klen[s] = op.seg[s].K);
  }
  int kacc = 0;
  for (int s = 0; s < op.nseg; ++s) {
    const PackedSeg& g = op.seg[s];
    x.kb[s] = kacc;
    x.Ah[s] = g.Ah; x.Al[s] = g.Al; x.lda[s] = g.lda;
    x.Bh[s] = g.Bh; x.Bl[s] = g.Bl; x.ldb[s] = g.ldb;
    x.az[s] = g.a_z; x.bz[s] = g.b_z;
    kacc += klen[s];
  }
  for (int s = op.nseg; s < MAXSEG; ++s) {
    x.kb[s] = kpad;
    x.Ah[s] = x.Ah[0]; x.Al[s] = x.Al[0]; x.lda[s] = x.lda[0];
    x.Bh[s] = x.Bh[0]; x.Bl[s] = x.Bl[0]; x.ldb[s] = x.ldb[0];
    x.az[s] = x.az[0]; x.bz[s] = x.bz[0];
  }
  x.kb[MAXSEG] = kpad;
  ECG_CHECK(op.batch >= 1 && op.batch <= 65535, "gemm: packed batch out of range");
  x.cbz = op.c_z;
  x.kpad = kpad; x.kchunk = kpad; x.ksplit = 1;
  x.C = op.C; x.Cf = op.Cf; x.cplain = 0;
  x.c_ms0 = op.c_ms0; x.c_ms1 = op.c_ms1; x.c_md0 = op.c_md0;
  x.c_ns0 = op.c_ns0; x.c_ns1 = op.c_ns1; x.c_nd0 = op.c_nd0;
  x.alpha = op.alpha; x.beta = op.beta;
  x.box_class = op.box_class; x.box_v = op.box_v;
  ECG_CHECK(op.box_class == 0 || ((long)op.box_v * op.box_v == op.M && op.box_v == op.N),
            "gemm: a boxed output needs M = box_v^2 and N = box_v");
  flops += 2LL * op.M * op.N * op.algorithmic_k * op.batch;
  bytes += op.batch * ((op.Cf ? 4.0 : 8.0) * (op.beta != 0.0 ? 2.0 : 1.0) * op.M * op.N +
                       (op.f64 ? 8.0 : 4.0) * (double)(op.M + op.N) * op.algorithmic_k);
  const long tiles = (long)x.tiles_m * x.tiles_n;
  ECG_CHECK(tiles < (1L << 31), "gemm: too many tiles");
  hipEvent_t e0 = timing_begin();
  // transposed accumulators when the stores run along m: unit m stride, or the box's fastest index
  const bool to = op.box_class ? op.box_class != 1 : (x.c_ms0 == 1 && x.c_ns0 != 1);
  const dim3 grid((unsigned)tiles, (unsigned)op.batch, 1);
  if (op.f64) {
    ECG_CHECK(op.C != nullptr, "gemm: a packed FP64 product writes an FP64 output");
    if (to) hipLaunchKernelGGL((gemm_f64p_dma_kernel<true, true>), grid, dim3(512), 0, stream, x);
    else hipLaunchKernelGGL((gemm_f64p_dma_kernel<false, true>), grid, dim3(512), 0, stream, x);
  } else if (op.f32) {
    if (to) hipLaunchKernelGGL((gemm_f32p_dma_kernel<true, true>), grid, dim3(512), 0, stream, x);
    else hipLaunchKernelGGL((gemm_f32p_dma_kernel<false, true>), grid, dim3(512), 0, stream, x);
  } else if (bf16_kernel == 1) {
    if (to) hipLaunchKernelGGL((gemm_bf16x3_dma_kernel<true, true>), grid, dim3(512), 0, stream, x);
    else hipLaunchKernelGGL((gemm_bf16x3_dma_kernel<false, true>), grid, dim3(512), 0, stream, x);
  } else {
    if (to) hipLaunchKernelGGL((gemm_bf16x3_kernel<true, true>), grid, dim3(512), 0, stream, x);
    else hipLaunchKernelGGL((gemm_bf16x3_kernel<false, true>), grid, dim3(512), 0, stream, x);
  }
  ECG_HIP(hipGetLastError());
  timing_end(e0);
}

}  // namespace ecg
