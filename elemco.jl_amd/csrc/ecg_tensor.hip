// Permutation / elementwise / reduction kernels and the contraction planner.
#include "ecg_tensor.h"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace ecg {

// ------------------------------------------------------------------ arena
double* Arena::take(long n) {
  size_t bytes = ((size_t)std::max<long>(n, 1) * sizeof(double) + 255) & ~(size_t)255;
  ECG_CHECK(top + bytes <= capacity(), ext ? "no-alloc scratch exhausted (the system is too large for it)"
                                            : "scratch arena exhausted (raise the workspace size)");
  double* p = reinterpret_cast<double*>(base() + top);
  top += bytes;
  high = std::max(high, top);
  return p;
}

// ------------------------------------------------------------------ permute
struct PermArgs {
  int nd;
  long n;                      // total elements
  long odim[kMaxDim];          // out extents
  long ostr[kMaxDim];          // out strides
  long istr_o[kMaxDim];        // input stride of the index at out position d
  double alpha, beta;
};

// Same fastest index on both sides: one thread per element, out order.
__global__ void permute_linear_kernel(const double* __restrict__ in, double* __restrict__ out, PermArgs a) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.n) return;
  long r = idx, ioff = 0;
#pragma unroll
  for (int d = 0; d < kMaxDim; ++d) {
    if (d >= a.nd) break;
    long x = r % a.odim[d];
    r /= a.odim[d];
    ioff += x * a.istr_o[d];
  }
  double v = a.alpha * in[ioff];
  if (a.beta != 0.0) v += a.beta * out[idx];
  out[idx] = v;
}

// Different fastest indices: 32x32 tiles over (X = input-fastest, Y = output-fastest)
// staged through LDS so both the read and the write are coalesced.
struct PermTileArgs {
  long dimX, dimY;
  long ixs, iys, oxs, oys;     // strides of X and Y in input and output
  int nrest;
  long rdim[kMaxDim], ristr[kMaxDim], rostr[kMaxDim];
  long tilesX, tilesY;
  double alpha, beta;
};

__global__ __launch_bounds__(256) void permute_tile_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                           PermTileArgs a) {
  __shared__ double tile[32][33];
  long b = blockIdx.x;
  long tx = b % a.tilesX; b /= a.tilesX;
  long ty = b % a.tilesY; b /= a.tilesY;
  long ioff = 0, ooff = 0;
  for (int d = 0; d < a.nrest; ++d) {
    long x = b % a.rdim[d];
    b /= a.rdim[d];
    ioff += x * a.ristr[d];
    ooff += x * a.rostr[d];
  }
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;  // 32 x 8
  long x0 = tx * 32, y0 = ty * 32;
  for (int yy = ly; yy < 32; yy += 8) {
    long x = x0 + lx, y = y0 + yy;
    if (x < a.dimX && y < a.dimY) tile[yy][lx] = in[ioff + x * a.ixs + y * a.iys];
  }
  __syncthreads();
  for (int xx = ly; xx < 32; xx += 8) {
    long y = y0 + lx, x = x0 + xx;
    if (x < a.dimX && y < a.dimY) {
      long o = ooff + x * a.oxs + y * a.oys;
      double v = a.alpha * tile[lx][xx];
      if (a.beta != 0.0) v += a.beta * out[o];
      out[o] = v;
    }
  }
}

void TensorOps::permute(const Tensor& in, const int* perm, Tensor& out, double alpha, double beta) {
  const int nd = in.nd;
  ECG_CHECK(out.nd == nd, "permute: rank mismatch");
  for (int d = 0; d < nd; ++d) ECG_CHECK(out.dim[d] == in.dim[perm[d]], "permute: extent mismatch");
  long n = in.size();
  if (n == 0) return;
  if (nd <= 1 || perm[0] == 0) {
    PermArgs a{};
    a.nd = std::max(nd, 1);
    a.n = n;
    a.alpha = alpha;
    a.beta = beta;
    for (int d = 0; d < nd; ++d) {
      a.odim[d] = out.dim[d];
      a.ostr[d] = out.stride(d);
      a.istr_o[d] = in.stride(perm[d]);
    }
    if (nd == 0) { a.odim[0] = 1; a.istr_o[0] = 0; }
    hipLaunchKernelGGL(permute_linear_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       in.p, out.p, a);
    ECG_HIP(hipGetLastError());
    return;
  }
  PermTileArgs a{};
  int X = 0;              // input dim 0
  int Y = perm[0];        // input dim that is fastest in out
  int oX = -1;
  for (int d = 0; d < nd; ++d) if (perm[d] == X) oX = d;
  a.dimX = in.dim[X];
  a.dimY = in.dim[Y];
  a.ixs = in.stride(X);
  a.iys = in.stride(Y);
  a.oxs = out.stride(oX);
  a.oys = out.stride(0);
  a.nrest = 0;
  for (int d = 0; d < nd; ++d) {
    int id = perm[d];
    if (id == X || id == Y) continue;
    a.rdim[a.nrest] = in.dim[id];
    a.ristr[a.nrest] = in.stride(id);
    a.rostr[a.nrest] = out.stride(d);
    a.nrest++;
  }
  a.tilesX = (a.dimX + 31) / 32;
  a.tilesY = (a.dimY + 31) / 32;
  long rest = 1;
  for (int d = 0; d < a.nrest; ++d) rest *= a.rdim[d];
  a.alpha = alpha;
  a.beta = beta;
  long blocks = a.tilesX * a.tilesY * rest;
  ECG_CHECK(blocks < (1L << 31), "permute: grid too large");
  hipLaunchKernelGGL(permute_tile_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, in.p, out.p, a);
  ECG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ elementwise, reductions
__global__ void axpby_kernel(long n, double alpha, const double* __restrict__ x, double beta, double* y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long st = (long)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    double v = alpha * x[i];
    if (beta != 0.0) v += beta * y[i];
    y[i] = v;
  }
}

__global__ void fill_kernel(long n, double v, double* y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long st = (long)gridDim.x * blockDim.x;
  for (; i < n; i += st) y[i] = v;
}

static unsigned grid_for(long n) { return (unsigned)std::min<long>((n + 255) / 256, 16384); }

void TensorOps::axpby(long n, double alpha, const double* x, double beta, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, stream, n, alpha, x, beta, y);
  ECG_HIP(hipGetLastError());
}

void TensorOps::fill(long n, double v, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, stream, n, v, y);
  ECG_HIP(hipGetLastError());
}

// Wavefront-level reduction of a double over 64 lanes.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__global__ __launch_bounds__(256) void dot_partial_kernel(long n, const double* __restrict__ x,
                                                          const double* __restrict__ y, double* part) {
  __shared__ double ws[4];
  double s = 0.0;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long st = (long)gridDim.x * blockDim.x;
  for (; i < n; i += st) s += x[i] * y[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ void sum_partial_kernel(int np, const double* __restrict__ part, double* res) {
  __shared__ double ws[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *res = ws[0] + ws[1] + ws[2] + ws[3];
}

void TensorOps::dot(long n, const double* x, const double* y, double* res_dev) {
  unsigned g = (unsigned)std::min<long>(std::max<long>((n + 255) / 256, 1), 2048);
  red.ensure(sizeof(double) * 4096);
  hipLaunchKernelGGL(dot_partial_kernel, dim3(g), dim3(256), 0, stream, n, x, y, red.d());
  hipLaunchKernelGGL(sum_partial_kernel, dim3(1), dim3(256), 0, stream, (int)g, red.d(), res_dev);
  ECG_HIP(hipGetLastError());
}

double TensorOps::dot_host(long n, const double* x, const double* y) {
  red.ensure(sizeof(double) * 4096);
  dot(n, x, y, red.d() + 4095);
  double r = 0.0;
  ECG_HIP(hipMemcpyAsync(&r, red.d() + 4095, sizeof(double), hipMemcpyDeviceToHost, stream));
  ECG_HIP(hipStreamSynchronize(stream));
  return r;
}

// ------------------------------------------------------------------ contraction planner
namespace {

struct Lab {
  const Tensor* t;
  std::string l;
  int pos(char c) const { auto k = l.find(c); return k == std::string::npos ? -1 : (int)k; }
  long dim(char c) const { return t->dim[pos(c)]; }
  long str(char c) const { return t->stride(pos(c)); }
  // labels g (in this order) are consecutive in memory: stride(next) == stride(prev) * dim(prev);
  // extent-1 labels never move the address and are skipped
  bool contig(const std::string& g) const {
    int prev = -1;
    for (size_t i = 0; i < g.size(); ++i) {
      if (dim(g[i]) == 1) continue;
      if (prev >= 0 && str(g[i]) != str(g[prev]) * dim(g[prev])) return false;
      prev = (int)i;
    }
    return true;
  }
  // stride of the first label with extent > 1 (0 if none)
  long first_str(const std::string& g) const {
    for (char x : g)
      if (dim(x) != 1) return str(x);
    return 0;
  }
};

long prod_dims(const Lab& t, const std::string& g) {
  long s = 1;
  for (char c : g) s *= t.dim(c);
  return s;
}

// All (order, split) pairs of a label set: each yields groups g1 | g2.
void splits(const std::string& labels, std::vector<std::pair<std::string, std::string>>& out) {
  std::string s = labels;
  std::sort(s.begin(), s.end());
  do {
    for (size_t k = (s.empty() ? 0 : 1); k <= s.size(); ++k) out.push_back({s.substr(0, k), s.substr(k)});
    if (s.empty()) out.push_back({"", ""});
  } while (std::next_permutation(s.begin(), s.end()));
}

void perms(const std::string& labels, std::vector<std::string>& out) {
  std::string s = labels;
  std::sort(s.begin(), s.end());
  do out.push_back(s); while (std::next_permutation(s.begin(), s.end()));
}

}  // namespace

void TensorOps::contract(double alpha, const Tensor& A, const char* la, const Tensor& B, const char* lb,
                         double beta, Tensor& C, const char* lc) {
  contract_p(prec, alpha, A, la, B, lb, beta, C, lc);
}

void TensorOps::contract_p(Prec pr, double alpha, const Tensor& A, const char* la, const Tensor& B,
                           const char* lb, double beta, Tensor& C, const char* lc) {
  Lab a{&A, la}, b{&B, lb}, c{&C, lc};
  ECG_CHECK((int)a.l.size() == A.nd && (int)b.l.size() == B.nd && (int)c.l.size() == C.nd,
            "contract: label count does not match tensor rank");
  std::string M, N, K;
  for (char x : c.l) {
    bool ia = a.pos(x) >= 0, ib = b.pos(x) >= 0;
    ECG_CHECK(ia != ib, std::string("contract: output label '") + x + "' must be in exactly one operand");
    (ia ? M : N) += x;
  }
  for (char x : a.l) {
    if (c.pos(x) >= 0) continue;
    ECG_CHECK(b.pos(x) >= 0, std::string("contract: label '") + x + "' not contracted");
    K += x;
  }
  for (char x : b.l)
    ECG_CHECK(c.pos(x) >= 0 || a.pos(x) >= 0, std::string("contract: label '") + x + "' not contracted");
  for (char x : M) ECG_CHECK(a.dim(x) == c.dim(x), "contract: extent mismatch (A,C)");
  for (char x : N) ECG_CHECK(b.dim(x) == c.dim(x), "contract: extent mismatch (B,C)");
  for (char x : K) ECG_CHECK(a.dim(x) == b.dim(x), "contract: extent mismatch (A,B)");

  std::vector<std::pair<std::string, std::string>> ms, ns;
  std::vector<std::string> ks;
  splits(M, ms);
  splits(N, ns);
  perms(K, ks);

  // score candidates: permuting an operand costs its size, an uncoalesced
  // operand counts as if it had to be permuted twice, C stores that are not
  // stride-1 along m cost a fraction of C's size.
  struct Best {
    double cost = 1e300;
    std::pair<std::string, std::string> m, n;
    std::string k;
    bool permA = false, permB = false, tmpC = false;
  } best;
  const double sA = (double)A.size(), sB = (double)B.size(), sC = (double)C.size();
  for (auto& mg : ms) {
    bool cm = c.contig(mg.first) && c.contig(mg.second);
    for (auto& ng : ns) {
      bool cn = c.contig(ng.first) && c.contig(ng.second);
      bool tmpC = !(cm && cn);
      for (auto& kg : ks) {
        bool aok = a.contig(mg.first) && a.contig(mg.second) && a.contig(kg);
        bool bok = b.contig(ng.first) && b.contig(ng.second) && b.contig(kg);
        double cost = 0;
        bool pA = !aok, pB = !bok;
        // the MFMA GEMM streams operands fastest along m (A) and n (B); a
        // k-contiguous operand would need a transposing LDS store, so it is
        // transposed in HBM first (a few tenths of a millisecond at these sizes)
        if (aok && !(a.first_str(mg.first) == 1 || prod_dims(a, M) == 1)) pA = true;
        if (bok && !(b.first_str(ng.first) == 1 || prod_dims(b, N) == 1)) pB = true;
        if (pA) cost += 2 * sA;
        if (pB) cost += 2 * sB;
        if (tmpC) cost += 3 * sC;
        else if (c.first_str(mg.first) != 1) cost += 0.25 * sC;
        if (cost < best.cost) {
          best.cost = cost; best.m = mg; best.n = ng; best.k = kg;
          best.permA = pA; best.permB = pB; best.tmpC = tmpC;
        }
      }
    }
  }
  const std::string Mo = best.m.first + best.m.second;
  const std::string No = best.n.first + best.n.second;
  const std::string& Ko = best.k;
  size_t mark = arena.mark();

  // materialise permuted operands
  Tensor Ause = A, Buse = B;
  std::string lau = a.l, lbu = b.l;
  auto make_perm = [&](const Lab& src, const std::string& order, Tensor& dst, std::string& ldst) {
    dst = Tensor();
    dst.nd = (int)order.size();
    for (int d = 0; d < dst.nd; ++d) dst.dim[d] = src.dim(order[d]);
    dst.p = arena.take(dst.size());
    int perm[kMaxDim];
    for (int d = 0; d < dst.nd; ++d) perm[d] = src.pos(order[d]);
    permute(*src.t, perm, dst, 1.0, 0.0);
    ldst = order;
  };
  if (best.permA) make_perm(a, Mo + Ko, Ause, lau);
  if (best.permB) make_perm(b, No + Ko, Buse, lbu);
  Lab au{&Ause, lau}, bu{&Buse, lbu};
  auto m1 = best.m.first, m2 = best.m.second, n1 = best.n.first, n2 = best.n.second;
  // the split used for C must be compatible with the one used for A / B
  Tensor Cuse = C;
  std::string lcu = c.l;
  bool tmpC = best.tmpC;
  if (!tmpC) {
    Lab cl{&C, lc};
    if (!(cl.contig(m1) && cl.contig(m2) && cl.contig(n1) && cl.contig(n2))) tmpC = true;
  }
  if (tmpC) {
    Cuse = Tensor();
    std::string order = m1 + m2 + n1 + n2;
    Cuse.nd = (int)order.size();
    for (int d = 0; d < Cuse.nd; ++d) Cuse.dim[d] = c.dim(order[d]);
    Cuse.p = arena.take(Cuse.size());
    lcu = order;
  }
  Lab cu{&Cuse, lcu};

  GemmOp op;
  op.M = (int)prod_dims(c, Mo);
  op.N = (int)prod_dims(c, No);
  op.K = (int)(Ko.empty() ? 1 : prod_dims(a, Ko));
  auto grp = [](const Lab& t, const std::string& g1, const std::string& g2, long& s0, long& s1, int& d0) {
    s0 = t.first_str(g1);
    d0 = g2.empty() ? (1 << 30) : (int)prod_dims(t, g1);
    s1 = t.first_str(g2);
  };
  op.A = Ause.p;
  grp(au, m1, m2, op.a_s0, op.a_s1, op.a_d0);
  op.a_k = au.first_str(Ko);
  op.B = Buse.p;
  grp(bu, n1, n2, op.b_s0, op.b_s1, op.b_d0);
  op.b_k = bu.first_str(Ko);
  op.C = Cuse.p;
  grp(cu, m1, m2, op.c_ms0, op.c_ms1, op.c_md0);
  grp(cu, n1, n2, op.c_ns0, op.c_ns1, op.c_nd0);
  op.alpha = tmpC ? 1.0 : alpha;
  op.beta = tmpC ? 0.0 : beta;
  gemm.stream = stream;
  gemm.run_prec(op, pr);
  if (tmpC) {
    int perm[kMaxDim];
    for (int d = 0; d < C.nd; ++d) perm[d] = cu.pos(c.l[d]);
    permute(Cuse, perm, C, alpha, beta);
  }
  arena.release_to(mark);
}

}  // namespace ecg
