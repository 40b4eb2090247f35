// Test-only entry points (libecg_testhooks.so): expose the tensor engine so the
// GPU tests can check single contractions and permutations against numpy.
// Not part of the product C ABI.
#include <algorithm>
#include <cstring>
#include <string>

#include "../ecg_dfprep.h"
#include "../ecg_tensor.h"

namespace {

ecg::Tensor make(double* p, int nd, const long* dims) {
  ecg::Tensor t;
  t.p = p;
  t.nd = nd;
  for (int i = 0; i < nd; ++i) t.dim[i] = dims[i];
  return t;
}

struct Env {
  ecg::TensorOps ops;
  hipStream_t s = nullptr;
  Env() {
    (void)hipStreamCreate(&s);
    ops.stream = s;
    ops.gemm.stream = s;
  }
};

Env& env() {
  static Env e;
  return e;
}

void put_err(char* err, int len, const std::string& m) {
  if (err && len > 0) {
    std::strncpy(err, m.c_str(), len - 1);
    err[len - 1] = 0;
  }
}

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// One wave: lane l supplies a[l] and b[l] as the A and B operand registers of
// the 16x16x4 MFMA and writes its four accumulator slots to out[4l + r].
__global__ void mfma_probe_kernel(const double* a, const double* b, double* out, const float* af, const float* bf,
                                  float* outf) {
  int l = threadIdx.x;
  d4_t acc = {0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[l], b[l], acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[4 * l + r] = acc[r];
  f4_t accf = {0.f, 0.f, 0.f, 0.f};
  accf = __builtin_amdgcn_mfma_f32_16x16x4f32(af[l], bf[l], accf, 0, 0, 0);
  for (int r = 0; r < 4; ++r) outf[4 * l + r] = accf[r];
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// One wave: lane l supplies a[8l..8l+7] and b[8l..8l+7] (small integers, exact
// in BF16) as the operand registers of the 16x16x32 BF16 MFMA and writes its
// four FP32 accumulator slots to out[4l + r].
__global__ void mfma_probe_bf16_kernel(const float* a, const float* b, float* out) {
  int l = threadIdx.x;
  bf16x8_t va, vb;
  for (int t = 0; t < 8; ++t) { va[t] = (__bf16)a[8 * l + t]; vb[t] = (__bf16)b[8 * l + t]; }
  f4_t acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[4 * l + r] = acc[r];
}

}  // namespace

extern "C" {

__global__ void fill_pattern_kernel(double* p, long n, double s) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (long)gridDim.x * blockDim.x) p[i] = s * (double)((i * 2654435761L) % 1000 - 500) * 1e-3;
}

__global__ void copy_kernel(const double* __restrict__ x, double* __restrict__ y, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

// Time `reps` evaluations of one contraction on device-resident operands; returns ms per call.
double ecgt_bench_contract(int prec, int na, const long* adims, const char* la, int nb, const long* bdims,
                           const char* lb, int nc, const long* cdims, const char* lc, int reps) {
  try {
    Env& e = env();
    ecg::Tensor tA = make(nullptr, na, adims), tB = make(nullptr, nb, bdims), tC = make(nullptr, nc, cdims);
    ecg::DevBuf dA(sizeof(double) * tA.size()), dB(sizeof(double) * tB.size()), dC(sizeof(double) * tC.size());
    tA.p = dA.d(); tB.p = dB.d(); tC.p = dC.d();
    hipLaunchKernelGGL(fill_pattern_kernel, dim3(4096), dim3(256), 0, e.s, dA.d(), tA.size(), 1.0);
    hipLaunchKernelGGL(fill_pattern_kernel, dim3(4096), dim3(256), 0, e.s, dB.d(), tB.size(), 0.5);
    size_t need = sizeof(double) * (size_t)(2 * (tA.size() + tB.size() + tC.size()) + 4096);
    if (e.ops.arena.buf.bytes < need) e.ops.arena.buf.alloc(need);
    e.ops.arena.reset();
    e.ops.contract_p(static_cast<ecg::Prec>(prec), 1.0, tA, la, tB, lb, 0.0, tC, lc);  // warm-up
    hipEvent_t a, b;
    ECG_HIP(hipEventCreate(&a));
    ECG_HIP(hipEventCreate(&b));
    ECG_HIP(hipEventRecord(a, e.s));
    for (int r = 0; r < reps; ++r) e.ops.contract_p(static_cast<ecg::Prec>(prec), 1.0, tA, la, tB, lb, 0.0, tC, lc);
    ECG_HIP(hipEventRecord(b, e.s));
    ECG_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    ECG_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / reps;
  } catch (const std::exception&) {
    return -1.0;
  }
}

// sqrtinvchol of the DF preparation (host only): M (n x n, first *rank columns used).
int ecgt_sqrtinvchol(const double* A, int n, double tol, int* rank, double* M) {
  try {
    int r = 0;
    const std::vector<double> m = ecg::sqrtinvchol(A, n, tol, r);
    std::copy(m.begin(), m.end(), M);
    *rank = r;
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// Selects the FP64 GEMM kernel for large products of the hooks' engine
// (GemmEngine::f64_kernel: 0 generic, 1 packed FP64 planes); returns the previous choice.
int ecgt_set_f64_kernel(int k) {
  Env& e = env();
  const int prev = e.ops.gemm.f64_kernel;
  e.ops.gemm.f64_kernel = k;
  return prev;
}

// Selects the FP32 GEMM kernel for large products of the hooks' engine
// (GemmEngine::f32_kernel: 0 generic, 1 packed FP32 planes); returns the previous choice.
int ecgt_set_f32_kernel(int k) {
  Env& e = env();
  const int prev = e.ops.gemm.f32_kernel;
  e.ops.gemm.f32_kernel = k;
  return prev;
}

// Selects the BF16 hi/lo GEMM kernel of the hooks' engine (GemmEngine::bf16_kernel);
// returns the previous choice.
int ecgt_set_bf16_kernel(int k) {
  Env& e = env();
  const int prev = e.ops.gemm.bf16_kernel;
  e.ops.gemm.bf16_kernel = k;
  return prev;
}

// Lane-layout probe of the FP64 and FP32 16x16x4 MFMA builtins (64 lanes, in bounds).
int ecgt_mfma_probe(const double* a, const double* b, double* out, const float* af, const float* bf, float* outf) {
  try {
    ecg::DevBuf da(64 * 8), db(64 * 8), dout(256 * 8), daf(64 * 4), dbf(64 * 4), doutf(256 * 4);
    ECG_HIP(hipMemcpy(da.p, a, 512, hipMemcpyHostToDevice));
    ECG_HIP(hipMemcpy(db.p, b, 512, hipMemcpyHostToDevice));
    ECG_HIP(hipMemcpy(daf.p, af, 256, hipMemcpyHostToDevice));
    ECG_HIP(hipMemcpy(dbf.p, bf, 256, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, 0, da.d(), db.d(), dout.d(), daf.f(), dbf.f(),
                       doutf.f());
    ECG_HIP(hipDeviceSynchronize());
    ECG_HIP(hipMemcpy(out, dout.p, 2048, hipMemcpyDeviceToHost));
    ECG_HIP(hipMemcpy(outf, doutf.p, 1024, hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// Plain device copy of n doubles (reads 8n bytes, writes 8n bytes), the
// calibration kernel for the FETCH_SIZE / WRITE_SIZE counters; returns ms per copy.
double ecgt_copy_bench(long n, int reps) {
  try {
    Env& e = env();
    ecg::DevBuf x(sizeof(double) * n), y(sizeof(double) * n);
    hipLaunchKernelGGL(fill_pattern_kernel, dim3(4096), dim3(256), 0, e.s, x.d(), n, 1.0);
    hipEvent_t a, b;
    ECG_HIP(hipEventCreate(&a));
    ECG_HIP(hipEventCreate(&b));
    ECG_HIP(hipEventRecord(a, e.s));
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e.s, x.d(), y.d(), n);
    ECG_HIP(hipEventRecord(b, e.s));
    ECG_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    ECG_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / reps;
  } catch (const std::exception&) {
    return -1.0;
  }
}

// Lane-layout probe of the 16x16x32 BF16 MFMA (64 lanes, 8 operands each, in bounds).
int ecgt_mfma_probe_bf16(const float* a, const float* b, float* out) {
  try {
    ecg::DevBuf da(512 * 4), db(512 * 4), dout(256 * 4);
    ECG_HIP(hipMemcpy(da.p, a, 2048, hipMemcpyHostToDevice));
    ECG_HIP(hipMemcpy(db.p, b, 2048, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mfma_probe_bf16_kernel, dim3(1), dim3(64), 0, 0, da.f(), db.f(), dout.f());
    ECG_HIP(hipDeviceSynchronize());
    ECG_HIP(hipMemcpy(out, dout.p, 1024, hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// C[lc] = alpha * A[la] B[lb] + beta * C[lc] on the device; host arrays in and out.
int ecgt_contract(int prec, double alpha, const double* A, int na, const long* adims, const char* la,
                  const double* B, int nb, const long* bdims, const char* lb, double beta, double* C, int nc,
                  const long* cdims, const char* lc, char* err, int errlen) {
  try {
    Env& e = env();
    ecg::Tensor tA = make(nullptr, na, adims), tB = make(nullptr, nb, bdims), tC = make(nullptr, nc, cdims);
    ecg::DevBuf dA(sizeof(double) * tA.size()), dB(sizeof(double) * tB.size()), dC(sizeof(double) * tC.size());
    tA.p = dA.d(); tB.p = dB.d(); tC.p = dC.d();
    ECG_HIP(hipMemcpy(dA.p, A, dA.bytes, hipMemcpyHostToDevice));
    ECG_HIP(hipMemcpy(dB.p, B, dB.bytes, hipMemcpyHostToDevice));
    ECG_HIP(hipMemcpy(dC.p, C, dC.bytes, hipMemcpyHostToDevice));
    size_t need = sizeof(double) * (size_t)(2 * (tA.size() + tB.size() + tC.size()) + 4096);
    if (e.ops.arena.buf.bytes < need) e.ops.arena.buf.alloc(need);
    e.ops.arena.reset();
    e.ops.contract_p(static_cast<ecg::Prec>(prec), alpha, tA, la, tB, lb, beta, tC, lc);
    ECG_HIP(hipStreamSynchronize(e.s));
    ECG_HIP(hipMemcpy(C, dC.p, dC.bytes, hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception& ex) {
    put_err(err, errlen, ex.what());
    return -1;
  }
}

// out = alpha * permute(in) + beta * out, out index d = in index perm[d].
int ecgt_permute(const double* in, int nd, const long* idims, const int* perm, double alpha, double beta,
                 double* out, char* err, int errlen) {
  try {
    Env& e = env();
    ecg::Tensor ti = make(nullptr, nd, idims);
    long od[ecg::kMaxDim];
    for (int d = 0; d < nd; ++d) od[d] = idims[perm[d]];
    ecg::Tensor to = make(nullptr, nd, od);
    ecg::DevBuf di(sizeof(double) * ti.size()), dout(sizeof(double) * to.size());
    ti.p = di.d();
    to.p = dout.d();
    ECG_HIP(hipMemcpy(di.p, in, di.bytes, hipMemcpyHostToDevice));
    ECG_HIP(hipMemcpy(dout.p, out, dout.bytes, hipMemcpyHostToDevice));
    e.ops.permute(ti, perm, to, alpha, beta);
    ECG_HIP(hipStreamSynchronize(e.s));
    ECG_HIP(hipMemcpy(out, dout.p, dout.bytes, hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception& ex) {
    put_err(err, errlen, ex.what());
    return -1;
  }
}

}  // extern "C"
