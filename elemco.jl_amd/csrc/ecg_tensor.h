// Dense column-major device tensors and the contraction planner.
//
// A Tensor is a device pointer plus extents (first index fastest, as in the
// Julia arrays of the reference).  contract() evaluates
//     C[c-labels] = alpha * sum A[a-labels] * B[b-labels] + beta * C
// by transposing operands only when the GEMM engine cannot address them
// directly (transpose-transpose-GEMM-transpose, as TensorOperations.jl does
// for the reference's @tensor lines).
#pragma once
#include "ecg_common.h"
#include "ecg_gemm.h"

#include <array>
#include <string>

namespace ecg {

constexpr int kMaxDim = 6;

struct Tensor {
  double* p = nullptr;
  int nd = 0;
  std::array<long, kMaxDim> dim{};
  Tensor() = default;
  Tensor(double* ptr, std::initializer_list<long> d) : p(ptr), nd((int)d.size()) {
    int i = 0;
    for (long x : d) dim[i++] = x;
  }
  long size() const {
    long s = 1;
    for (int i = 0; i < nd; ++i) s *= dim[i];
    return s;
  }
  long stride(int i) const {
    long s = 1;
    for (int k = 0; k < i; ++k) s *= dim[k];
    return s;
  }
};

// Scratch arena: bump allocation from one device block, reset per phase.  The
// block is the library's own buffer, or, while `ext` is set, a caller-supplied
// region (the allocation-free entry points), which is never grown or freed.
struct Arena {
  DevBuf buf;
  char* ext = nullptr;
  size_t ext_bytes = 0;
  size_t top = 0;
  size_t high = 0;
  double* take(long n);
  void reset() { top = 0; }
  size_t mark() const { return top; }
  void release_to(size_t m) { top = m; }
  char* base() const { return ext ? ext : static_cast<char*>(buf.p); }
  size_t capacity() const { return ext ? ext_bytes : buf.bytes; }
  size_t free_bytes() const { return capacity() > top ? capacity() - top : 0; }
};

struct TensorOps {
  hipStream_t stream = nullptr;
  GemmEngine gemm;
  Arena arena;
  Prec prec = Prec::F64;

  // out[perm] : out index d takes input index perm[d];  out = alpha*permute(in) + beta*out
  void permute(const Tensor& in, const int* perm, Tensor& out, double alpha, double beta);
  // C = alpha*A*B + beta*C with index labels (each label a single character)
  void contract(double alpha, const Tensor& A, const char* la, const Tensor& B, const char* lb,
                double beta, Tensor& C, const char* lc);
  void contract_p(Prec p, double alpha, const Tensor& A, const char* la, const Tensor& B,
                  const char* lb, double beta, Tensor& C, const char* lc);
  // y = alpha*x + beta*y over n elements
  void axpby(long n, double alpha, const double* x, double beta, double* y);
  void fill(long n, double v, double* y);
  void copy(long n, const double* x, double* y) { axpby(n, 1.0, x, 0.0, y); }
  // sum_i x[i]*y[i] accumulated in FP64 (device result written to *res_dev)
  void dot(long n, const double* x, const double* y, double* res_dev);
  double dot_host(long n, const double* x, const double* y);
  DevBuf red;  // reduction scratch
};

}  // namespace ecg
