// Tensor-GEMM engine for gfx950: C(m,n) = alpha * sum_k A(m,k) B(k,n) + beta * C(m,n).
//
// Rows (m) of A and C and columns (n) of B and C may each be a pair of tensor
// indices with independent strides ("split strides"):
//     A(m,k) = A[(m % a_d0)*a_s0 + (m / a_d0)*a_s1 + k*a_k]
//     B(k,n) = B[(n % b_d0)*b_s0 + (n / b_d0)*b_s1 + k*b_k]
//     C(m,n) = C[(m % c_md0)*c_ms0 + (m / c_md0)*c_ms1 + (n % c_nd0)*c_ns0 + (n / c_nd0)*c_ns1]
// so permuted outputs (the (T) intermediates, the DF ladder blocks) are written
// in place without a separate transpose pass.  An optional batch index z adds
// z*a_z, z*b_z, z*c_z to the three base pointers.
//
// Segmented K: with nseg > 0 the product is a sum over up to MAXSEG segments,
//     C = alpha * sum_s sum_{k < seg[s].K} A_s(m,k) B_s(k,n) + beta * C,
// where segment s has its own base pointers and K strides (seg[s].A, a_k, B,
// b_k) and shares the m / n addressing of the op.  One launch then accumulates
// several products into the same output tile (the (T) intermediates).
#pragma once
#include "ecg_common.h"

#include <vector>

namespace ecg {

constexpr int MAXSEG = 4;
struct Arena;

struct GemmSeg {
  const double* A = nullptr; long a_k = 0;
  const double* B = nullptr; long b_k = 0;
  int K = 0;
};

struct GemmOp {
  int M = 0, N = 0, K = 0, batch = 1;
  const double* A = nullptr; long a_s0 = 1, a_s1 = 0, a_k = 0, a_z = 0; int a_d0 = 1 << 30;
  const double* B = nullptr; long b_s0 = 1, b_s1 = 0, b_k = 0, b_z = 0; int b_d0 = 1 << 30;
  double* C = nullptr; long c_ms0 = 1, c_ms1 = 0, c_ns0 = 0, c_ns1 = 0, c_z = 0;
  int c_md0 = 1 << 30, c_nd0 = 1 << 30;
  double alpha = 1.0, beta = 0.0;
  int nseg = 0;            // 0: single product (A, a_k, B, b_k, K); else seg[0..nseg)
  GemmSeg seg[MAXSEG];
  // BF16 hi/lo products: B's data does not change until GemmEngine::drop_kept_planes(), so
  // its packed planes are kept and reused by the next product with the same B operand
  bool b_keep = false;
  // Packed-plane products: A already packed in the layout the product's own packing
  // writes (BF16 hi plane a_planes and lo plane a_planes_lo, or the FP32 / FP64 plane),
  // rows of a_planes_ld = K padded to 32 elements, readable up to M rounded up to 256
  // rows; A, a_s0/a_s1/a_d0 and a_k are then unused and A is not packed
  const void* a_planes = nullptr;
  const void* a_planes_lo = nullptr;
  long a_planes_ld = 0;

  // Plain column-major helpers: A is M x K (lda), or K x M when transposed.
  GemmOp& a_colmajor(const double* p, long lda, bool trans) {
    A = p; a_d0 = 1 << 30; a_s1 = 0;
    if (!trans) { a_s0 = 1; a_k = lda; } else { a_s0 = lda; a_k = 1; }
    return *this;
  }
  GemmOp& b_colmajor(const double* p, long ldb, bool trans) {
    B = p; b_d0 = 1 << 30; b_s1 = 0;
    if (!trans) { b_s0 = ldb; b_k = 1; } else { b_s0 = 1; b_k = ldb; }
    return *this;
  }
  GemmOp& c_colmajor(double* p, long ldc) {
    C = p; c_md0 = 1 << 30; c_ms0 = 1; c_ms1 = 0; c_nd0 = 1 << 30; c_ns0 = ldc; c_ns1 = 0;
    return *this;
  }
};

// A product on operands already packed as BF16 hi/lo planes (pack_rows_bf16x3):
// segment s reads A rows m at Ah/Al + m*lda and B rows n at Bh/Bl + n*ldb, each
// row K-contiguous and zero-padded to a multiple of 32 (K is that padded
// length).  Rows up to the next multiple of 256 (A) and 128 (B) past M and N are
// read, so the planes need that much readable memory after the last row (the
// results of those rows are discarded).  C is FP64 or FP32, split-stride
// addressed as in GemmOp.
struct PackedSeg {
  const uint16_t* Ah = nullptr; const uint16_t* Al = nullptr; long lda = 0;
  const uint16_t* Bh = nullptr; const uint16_t* Bl = nullptr; long ldb = 0;
  int K = 0;
  // rows hold data in their first Kt elements and zeros up to K (0: Kt = K).  FP64 products
  // then run the segment over Kt rounded up to 16 (their slab); BF16 hi/lo and FP32 ones
  // over K
  int Kt = 0;
  long a_z = 0, b_z = 0;   // element offsets of A and B (both planes) per batch index
};
struct PackedOp {
  int M = 0, N = 0, nseg = 0;
  PackedSeg seg[MAXSEG];
  double* C = nullptr; float* Cf = nullptr;   // exactly one is set
  long c_ms0 = 1, c_ms1 = 0, c_ns0 = 0, c_ns1 = 0;
  int c_md0 = 1 << 30, c_nd0 = 1 << 30;
  double alpha = 1.0, beta = 0.0;
  long long algorithmic_k = 0;                 // unpadded K summed over segments (flop count)
  // box_class = 1, 2, 3: instead of the split strides, C is a tensor K[x1,x2,x3] (each
  // index < box_v) stored in 8x8x8 boxes: element at ((x1/8) + nt (x2/8 + nt x3/8)) * 512
  // + x1%8 + 8 (x2%8) + 64 (x3%8), nt = ceil(box_v / 8).  With m = y1 + box_v y2 the
  // coordinates are (n, y1, y2) for class 1, (y1, n, y2) for 2 and (y1, y2, n) for 3.
  int box_class = 0, box_v = 0;
  // batch: products z = 0 .. batch-1 read the segments' planes at + z * a_z / b_z and
  // write C at + z * c_z (elements)
  int batch = 1;
  long c_z = 0;
  // f32: the planes are FP32 (pack_rows_f32; Ah / Bh hold the FP32 plane pointers, Al / Bl
  // are unused, row strides and batch offsets count floats) and the products are FP32 MFMA
  bool f32 = false;
  // f64: FP64 planes (pack_rows_f64; plane pointers in Ah / Bh) multiplied by the FP64 MFMA
  // kernel into the FP64 output C
  bool f64 = false;
};

// Output of a packed-plane product written into packed planes of a later product's A
// operand instead of C: element (m, n), m = c + dm a' (c < dm), n = d + dn b' (d < dn),
// goes to plane row a' ra + b' rb, K position c + kn d, in the layout of the product
// kind's packing (BF16 hi and lo planes, the FP32 or the FP64 plane; rows of ld elements).
// The product stores with lanes along m, i.e. along consecutive positions of a row.
// With tri, only elements with a' <= b' are stored, at row b' (b' + 1) / 2 + a' (ra, rb
// unused), and tiles that hold none are not computed.
struct PlaneOut {
  void* hi = nullptr;   // BF16 hi plane, or the FP32 / FP64 plane
  void* lo = nullptr;   // BF16 lo plane
  long ld = 0;
  int dm = 1, dn = 1;
  long ra = 0, rb = 0, kn = 0;
  int tri = 0;
};

// Precision of the products: FP64 MFMA, FP32 MFMA (inputs rounded to FP32,
// FP32 accumulation), or BF16 hi/lo split (three BF16 MFMA products per pair,
// FP32 accumulation).  Results are always returned in FP64.
struct GemmEngine {
  hipStream_t stream = nullptr;
  DevBuf ws;          // split-K partial sums
  Prec prec = Prec::F64;
  long long flops = 0;   // running count of 2*M*N*K issued (for reporting)
  double bytes = 0.0;    // running count of algorithmic bytes: 8 (M K + K N) + 8 M N (16 M N when beta != 0)
  // optional per-launch timing with HIP events on `stream` (off by default)
  bool timed = false;
  double timed_ms = 0.0;
  long long timed_launches = 0;
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  void collect_timing();   // synchronises and accumulates recorded launches
  void run(const GemmOp& op);
  void run_prec(const GemmOp& op, Prec p);
  // BF16 hi/lo products (ecg_gemm_bf16.hip): operands packed into zero-padded,
  // K-contiguous BF16 hi and lo planes in `pk`, 256x128 MFMA tiles
  DevBuf pk;
  // BF16 hi/lo GEMM kernel: 0 register-staged slabs (gemm_bf16x3_kernel), 1 slabs loaded
  // straight into LDS (gemm_bf16x3_dma_kernel)
  int bf16_kernel = 1;
  void run_bf16x3(const GemmOp& op);
  // FP32 products on FP32 planes packed on the fly (ecg_gemm_bf16.hip, the (T)'s FP32 kernel)
  void run_f32p(const GemmOp& op);
  // FP64 products on FP64 planes packed on the fly (K slabs of 16)
  void run_f64p(const GemmOp& op);
  // kind 0 BF16 hi/lo, 1 FP32, 2 FP64 planes; with po the result goes to planes of the
  // same kind (beta 0, K not split)
  void run_planes(const GemmOp& op, int kind, const PlaneOut* po = nullptr);
  int f32_kernel = 1;   // large FP32 products: 0 generic FP32 MFMA kernel, 1 packed FP32 planes
  int f64_kernel = 1;   // large FP64 products: 0 generic FP64 MFMA kernel, 1 packed FP64 planes
  // packed B planes kept for products with GemmOp::b_keep (one operand, identified by its
  // addressing); drop_kept_planes() forgets them (the caller's B may change afterwards)
  DevBuf pk_keep;
  std::vector<long> keep_key;
  void drop_kept_planes() { keep_key.clear(); }
  // Pack rows r < rows of src (row offset split as in GemmOp: (r % d0)*s0 + (r / d0)*s1,
  // element k at + k*sk, k < K) into hi/lo planes [r][kpad], kpad a multiple of 32 >= K.
  // Rows are written up to the next multiple of 64.
  void pack_rows_bf16x3(const double* src, long rows, int d0, long s0, long s1, int K, long sk, int kpad,
                        uint16_t* hi, uint16_t* lo);
  void run_bf16x3_packed(const PackedOp& op);   // also the FP32 packed product (op.f32)
  // Pack rows as pack_rows_bf16x3 into one FP32 plane [r][kpad]; within each 32-wide K slab
  // position 8*(k%4) + k/4 holds k (the order the FP32 packed product reads).
  void pack_rows_f32(const double* src, long rows, int d0, long s0, long s1, int K, long sk, int kpad, float* out);
  void pack_rows_f64(const double* src, long rows, int d0, long s0, long s1, int K, long sk, int kpad, double* out);
  void reduce_slabs(const GemmOp& op, int ksplit, const double* wsp);   // split-K partial slabs -> C
  // While set, split-K partial sums are taken from this region (and returned after each
  // product) instead of `ws`, no operand is packed into planes, and only FP64 products
  // run: the engine then allocates nothing (the allocation-free entry points).
  Arena* transient = nullptr;
  // CUs a product may expect to have to itself when the split-K rules decide how far to
  // split (256: the whole device; fewer while several contexts run products concurrently)
  int cus = 256;
  hipEvent_t timing_begin();
  void timing_end(hipEvent_t e0);
};

}  // namespace ecg
