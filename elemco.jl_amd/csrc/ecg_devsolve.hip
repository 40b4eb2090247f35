// Device-resident CC iterations: one pass of the loop body of cc_iterations!
// (src/cc/cc.jl:2546-2629, closed-shell branch) with every reduction, the DIIS linear
// solve (src/solvers/diis.jl:246-305) and the convergence test on the device, so an
// iteration needs no host round trip until its info vector is read.  The launch
// sequence of an iteration depends only on the sizes and the context's buffers, so it is
// captured once as a HIP graph and replayed (one graph launch per iteration); the DIIS
// ring state (next slot, subspace size, B matrix) lives in device memory and advances
// inside the replays.
//
// The arithmetic follows the host-driven Context::state_step term by term: pair sums of
// the closed-shell dots are added in the same order (one thread, pair by pair), the
// Hylleraas and projected energies use the same expressions, and the DIIS solve is the
// same Jacobi pseudo-inverse of the bordered B matrix, so the iterations agree with the
// host-driven ones to rounding.
#include "ecg_cc.h"

#include <algorithm>
#include <cmath>
#include <cstring>

// The scalar arithmetic below (energies, norms, the DIIS solve) mirrors host code that is
// compiled without fused multiply-adds; contracting it here would change results in the
// last bit, so it is kept off.
#pragma clang fp contract(off)

namespace ecg {

namespace {

// slots of the device scalar array
enum : int {
  S_NT2 = 0,      // pair sums (d, e) of (T2, T2)
  S_NR2 = 2,      // (R2, R2)
  S_PD = 4,       // (T2, <ij|ab>)
  S_RD = 6,       // (T2, R2)
  S_TT = 8,       // (T1 T1, <ij|ab>)
  S_FT1 = 10,     // T1 . f_vo
  S_RT1 = 11,     // R1 . T1
  S_NT1 = 12,     // T1 . T1
  S_NR1 = 13,     // R1 . R1
  S_DS = 14,      // DIIS singles dots R1 . res_i[singles], i < 6
  S_DP = 20,      // DIIS pair sums (d, e) of (R2, res_i[doubles]), i < 6
  S_NPD = 32,     // (T2_new, <ij|ab>)
  S_NTT = 34,     // (T1_new T1_new, <ij|ab>)
  S_NFT1 = 36,    // T1_new . f_vo
  S_INFO = 40,    // info[8] of the iteration
  S_NEXT = 48,    // DIIS: next slot
  S_NDIM = 49,    //       subspace size
  S_ND = 50,      //       vectors used by this iteration's extrapolation
  S_EXTRAP = 51,  //       1: extrapolate
  S_COEF = 52,    //       coefficients (7)
  S_B = 60,       //       B matrix (7 x 7, row-major, stride 7)
  S_ITER = 110,   // iterations done
  S_COUNT = 112
};
constexpr int MAXD = 6;

struct Slots {
  double* amps[MAXD];
  double* res[MAXD];
};

// sum of the per-pair partials (d, e) in pair order (the order of Context::doubles_de's host sum)
__global__ void pair_sum_kernel(long np, const double* __restrict__ part, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double d = 0.0, e = 0.0;
  for (long k = 0; k < np; ++k) {
    d += part[2 * k];
    e += part[2 * k + 1];
  }
  out[0] = d;
  out[1] = e;
}

// amps[next] = (T1, T2), res[next] = (R1, R2)
__global__ void diis_store_kernel(Slots sl, long n1, long n2, const double* __restrict__ T1,
                                  const double* __restrict__ T2, const double* __restrict__ R1,
                                  const double* __restrict__ R2, const double* __restrict__ sc) {
  const int k = (int)sc[S_NEXT];
  double* a = sl.amps[k];
  double* r = sl.res[k];
  const long n = n1 + n2;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    if (i < n1) {
      a[i] = T1[i];
      r[i] = R1[i];
    } else {
      a[i] = T2[i - n1];
      r[i] = R2[i - n1];
    }
  }
}

// cyclic Jacobi eigen-decomposition of a symmetric m x m matrix (row-major A, m <= 7)
__device__ void jacobi_eig_dev(int m, double* A, double* V, double* w) {
  for (int i = 0; i < m * m; ++i) V[i] = 0.0;
  for (int i = 0; i < m; ++i) V[i * m + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) off += A[p * m + q] * A[p * m + q];
    if (off < 1e-300) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = A[p * m + q];
        if (fabs(apq) < 1e-300) continue;
        const double app = A[p * m + p], aqq = A[q * m + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; ++k) {
          const double akp = A[k * m + p], akq = A[k * m + q];
          A[k * m + p] = c * akp - s * akq;
          A[k * m + q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          const double apk = A[p * m + k], aqk = A[q * m + k];
          A[p * m + k] = c * apk - s * aqk;
          A[q * m + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          const double vkp = V[k * m + p], vkq = V[k * m + q];
          V[k * m + p] = c * vkp - s * vkq;
          V[k * m + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < m; ++i) w[i] = A[i * m + i];
}

// perform! (diis.jl:246-305) with the closed-shell custom dots: B-matrix row of this
// iteration, pseudo-inverse solve, skip / replacement rules; writes the coefficients, the
// extrapolation flag and the ring state
__global__ void diis_solve_kernel(double* sc, int md, double resthr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int ndim = (int)sc[S_NDIM];
  const int next = (int)sc[S_NEXT];
  if (ndim < md) ndim++;
  const int ithis = next, nd = ndim;
  auto B = [&](int i, int j) -> double& { return sc[S_B + i * (MAXD + 1) + j]; };
  auto dotv = [&](int i) { return 2.0 * sc[S_DS + i] + (sc[S_DP + 2 * i] - 0.5 * sc[S_DP + 2 * i + 1]) * 2.0; };
  const double thisdot = dotv(ithis);
  for (int i = 0; i < nd; ++i) {
    if (i != ithis) {
      const double d = dotv(i);
      B(i, ithis) = d;
      B(ithis, i) = d;
    } else {
      B(ithis, ithis) = thisdot;
    }
    B(i, nd) = -1.0;
    B(nd, i) = -1.0;
  }
  B(nd, nd) = 0.0;
  const int m = nd + 1;
  double A[(MAXD + 1) * (MAXD + 1)], V[(MAXD + 1) * (MAXD + 1)], w[MAXD + 1], coeff[MAXD + 1];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) A[i * m + j] = B(i, j);
  jacobi_eig_dev(m, A, V, w);
  double smax = 0.0;
  for (int i = 0; i < m; ++i) smax = fmax(smax, fabs(w[i]));
  for (int i = 0; i < m; ++i) coeff[i] = 0.0;
  for (int k = 0; k < m; ++k) {
    if (fabs(w[k]) < 2.220446049250313e-16 * smax || w[k] == 0.0) continue;
    const double proj = -V[nd * m + k];
    for (int i = 0; i < m; ++i) coeff[i] += V[i * m + k] * proj / w[k];
  }
  int nnext = next;
  double extrap = 1.0;
  if (nd == 1 && next == 0 && thisdot > resthr) {
    ndim = 0;
    extrap = 0.0;
  } else if (nd < md) {
    nnext = nd;
  } else {
    double cmin = 1e10;
    for (int i = 0; i < nd; ++i)
      if (fabs(coeff[i]) < cmin && (i != ithis || nd == 1)) {
        cmin = fabs(coeff[i]);
        nnext = i;
      }
  }
  sc[S_NDIM] = ndim;
  sc[S_NEXT] = nnext;
  sc[S_ND] = nd;
  sc[S_EXTRAP] = extrap;
  for (int i = 0; i < MAXD + 1; ++i) sc[S_COEF + i] = i < m ? coeff[i] : 0.0;
}

// (T1, T2) = sum_{i < nd} c_i amps_i when the DIIS step extrapolates
__global__ void diis_extrapolate_kernel(Slots sl, long n1, long n2, double* T1, double* T2,
                                        const double* __restrict__ sc) {
  if (sc[S_EXTRAP] == 0.0) return;
  const int nd = (int)sc[S_ND];
  const long n = n1 + n2;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double v = sc[S_COEF] * sl.amps[0][i];
    for (int k = 1; k < nd; ++k) {
      const double t = sc[S_COEF + k] * sl.amps[k][i];
      v = t + 1.0 * v;
    }
    if (i < n1) T1[i] = v;
    else T2[i - n1] = v;
  }
}

// energies, norms, info and the convergence test of the iteration (state_step)
__global__ void iter_finish_kernel(double* sc, double thr, double conven) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  // calc_hylleraas (cc.jl:195-241) at the amplitudes before the update
  const double pd = sc[S_PD], pex = sc[S_PD + 1], rd = sc[S_RD], rex = sc[S_RD + 1];
  const double ss = sc[S_TT] - sc[S_TT + 1], os = sc[S_TT];
  const double s0 = ss + os + 2.0 * sc[S_FT1];
  const double ET1SS = ss, ET1OS = os;
  const double pET1 = s0 - ET1SS - ET1OS;
  const double rET1 = 2.0 * sc[S_RT1];
  const double ET2d = pd + rd, ET2ex = pex + rex;
  const double ET2SS = ET2d - ET2ex + ET1SS, ET2OS = ET2d + ET1OS;
  const double ET2 = (ET2d - ET2ex) + ET2d + (ET1SS + ET1OS) + pET1 + rET1;
  const double pE = (pd - pex) + pd + pET1 + ET1SS + ET1OS;
  // norms (cctools.jl:453-480)
  const double NormT2 = 2.0 * (sc[S_NT2] - 0.5 * sc[S_NT2 + 1]);
  const double NormR2 = 2.0 * (sc[S_NR2] - 0.5 * sc[S_NR2 + 1]);
  const double NormT1 = 2.0 * sc[S_NT1], NormR1 = 2.0 * sc[S_NR1];
  // projected energy of the new amplitudes (calc_doubles_energy + calc_singles_energy)
  const double e2 = (sc[S_NPD] - sc[S_NPD + 1]) + sc[S_NPD];
  const double nss = sc[S_NTT] - sc[S_NTT + 1], nos = sc[S_NTT];
  const double e1 = nss + nos + 2.0 * sc[S_NFT1];
  const double En = e2 + e1;
  const double dE = En - ET2;
  const double NormR = NormR1 + NormR2;
  const double thren = sqrt(thr) * conven;
  sc[S_INFO + 0] = ET2;
  sc[S_INFO + 1] = ET2SS;
  sc[S_INFO + 2] = ET2OS;
  sc[S_INFO + 3] = pE;
  sc[S_INFO + 4] = dE;
  sc[S_INFO + 5] = NormR;
  sc[S_INFO + 6] = 1.0 + NormT1 + NormT2;
  sc[S_INFO + 7] = (NormR < thr && fabs(dE) < thren) ? 1.0 : 0.0;
  sc[S_ITER] += 1.0;
}

unsigned grid_n(long n) { return (unsigned)std::min<long>((n + 255) / 256, 4096); }

}  // namespace

// Buffers and state of the device-resident iterations; the DIIS ring state is zeroed.
void Context::dev_reset(const SolverOpts& o) {
  ECG_CHECK(o.maxdiis >= 1 && o.maxdiis <= MAXD, "device iterations: maxdiis must lie in 1..6");
  const long n1 = nvo(), n2 = nvvoo(), np = (long)no * no;
  DevIter& d = dv;
  d.scal.ensure(sizeof(double) * S_COUNT);
  d.part.ensure(sizeof(double) * 2 * np);
  d.tt.ensure(sizeof(double) * n2);
  d.fvo.ensure(sizeof(double) * n1);
  if (!d.info_host) ECG_HIP(hipHostMalloc(reinterpret_cast<void**>(&d.info_host), sizeof(double) * 8));
  if (!d.done) ECG_HIP(hipEventCreateWithFlags(&d.done, hipEventDisableTiming));
  ops.fill(S_COUNT, 0.0, d.scal.d());
  // f_vo[a,i] = f[i, o+a] (the Fock matrix is fixed)
  ECG_HIP(hipMemcpy2DAsync(d.tt.p, sizeof(double) * no, fock.d() + (size_t)no * n, sizeof(double) * n,
                           sizeof(double) * no, nv, hipMemcpyDeviceToDevice, stream));
  Tensor src(d.tt.d(), {(long)no, (long)nv}), dst(d.fvo.d(), {(long)nv, (long)no});
  const int pt[2] = {1, 0};
  ops.permute(src, pt, dst, 1.0, 0.0);
  // DIIS vectors of the host-driven solver serve as the ring; zeroed so that dots with
  // slots not yet written are finite
  diis_reset(o.maxdiis, o.resthr);
  for (int k = 0; k < o.maxdiis; ++k) {
    ops.fill(n1 + n2, 0.0, diis.amps[k].d());
    ops.fill(n1 + n2, 0.0, diis.res[k].d());
  }
  d.maxdiis = o.maxdiis;
}

// Enqueue one iteration on the context's stream (eagerly or under stream capture).
void Context::dev_step(int method, const SolverOpts& o) {
  const long n1 = nvo(), n2 = nvvoo(), np = (long)no * no;
  DevIter& d = dv;
  double* sc = d.scal.d();
  auto pair = [&](const double* X, const double* Y, int slot) {
    launch_pair_dots(stream, nv, np, X, Y, d.part.d());
    hipLaunchKernelGGL(pair_sum_kernel, dim3(1), dim3(64), 0, stream, np, d.part.d(), sc + slot);
  };
  auto tt_of = [&](const double* t1) {
    Tensor a(const_cast<double*>(t1), {(long)nv, (long)no}), tt(d.tt.d(), {(long)nv, (long)nv, (long)no, (long)no});
    ops.contract_p(Prec::F64, 1.0, a, "ai", a, "bj", 0.0, tt, "abij");
  };
  residual(method, Prec::F64, T1.d(), T2.d(), R1.d(), R2.d());
  pair(T2.d(), T2.d(), S_NT2);
  pair(R2.d(), R2.d(), S_NR2);
  pair(T2.d(), vvoo.d(), S_PD);
  pair(T2.d(), R2.d(), S_RD);
  tt_of(T1.d());
  pair(d.tt.d(), vvoo.d(), S_TT);
  ops.dot(n1, T1.d(), d.fvo.d(), sc + S_FT1);
  ops.dot(n1, R1.d(), T1.d(), sc + S_RT1);
  ops.dot(n1, T1.d(), T1.d(), sc + S_NT1);
  ops.dot(n1, R1.d(), R1.d(), sc + S_NR1);
  update(T1.d(), T2.d(), R1.d(), R2.d(), o.shifts, o.shiftp);
  Slots sl{};
  for (int k = 0; k < MAXD; ++k) {
    const int kk = std::min(k, d.maxdiis - 1);
    sl.amps[k] = diis.amps[kk].d();
    sl.res[k] = diis.res[kk].d();
  }
  hipLaunchKernelGGL(diis_store_kernel, dim3(grid_n(n1 + n2)), dim3(256), 0, stream, sl, n1, n2, T1.d(), T2.d(),
                     R1.d(), R2.d(), sc);
  for (int k = 0; k < d.maxdiis; ++k) {
    ops.dot(n1, R1.d(), diis.res[k].d(), sc + S_DS + k);
    pair(R2.d(), diis.res[k].d() + n1, S_DP + 2 * k);
  }
  hipLaunchKernelGGL(diis_solve_kernel, dim3(1), dim3(64), 0, stream, sc, d.maxdiis, o.resthr);
  hipLaunchKernelGGL(diis_extrapolate_kernel, dim3(grid_n(n1 + n2)), dim3(256), 0, stream, sl, n1, n2, T1.d(),
                     T2.d(), (const double*)sc);
  pair(T2.d(), vvoo.d(), S_NPD);
  tt_of(T1.d());
  pair(d.tt.d(), vvoo.d(), S_NTT);
  ops.dot(n1, T1.d(), d.fvo.d(), sc + S_NFT1);
  hipLaunchKernelGGL(iter_finish_kernel, dim3(1), dim3(64), 0, stream, sc, o.thr, o.conven);
  ECG_HIP(hipGetLastError());
}

// One iteration: the first after dev_reset runs eagerly (it sizes every scratch buffer),
// later ones replay the captured graph of the same (method, options, buffers); the info
// vector is copied to pinned host memory (d.info_host) on the stream.
void Context::dev_launch(int method, const SolverOpts& o) {
  DevIter& d = dv;
  auto key_now = [&]() {
    std::vector<const void*> k{T1.p, T2.p, R1.p, R2.p, d.scal.p, d.part.p, d.tt.p, d.fvo.p, ops.gemm.ws.p,
                               ops.gemm.pk.p, ops.gemm.pk_keep.p, ops.arena.base(), ops.red.p};
    for (int i = 0; i < d.maxdiis; ++i) {
      k.push_back(diis.amps[i].p);
      k.push_back(diis.res[i].p);
    }
    return k;
  };
  if (ops.gemm.timed) {   // per-launch timing records events: run eagerly
    dev_step(method, o);
    d.eager_done = true;
    ECG_HIP(hipMemcpyAsync(d.info_host, d.scal.d() + S_INFO, sizeof(double) * 8, hipMemcpyDeviceToHost, stream));
    return;
  }
  // the split-K share is part of the key: it fixes the captured launches' grids and splits
  const std::vector<double> okey{(double)method, o.thr,       o.conven, o.shifts, o.shiftp, o.resthr,
                                 (double)o.maxdiis, (double)ops.gemm.cus};
  if (d.exec && (d.key != key_now() || d.okey != okey)) {
    (void)hipGraphExecDestroy(d.exec);
    d.exec = nullptr;
  }
  if (d.exec) {
    residual_calls++;
    ECG_HIP(hipGraphLaunch(d.exec, stream));
  } else if (!d.eager_done) {
    dev_step(method, o);
    d.eager_done = true;
  } else {
    hipGraph_t graph = nullptr;
    const std::vector<const void*> before = key_now();
    capturing = true;
    ECG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      dev_step(method, o);
    } catch (...) {
      (void)hipStreamEndCapture(stream, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      capturing = false;
      throw;
    }
    capturing = false;
    ECG_HIP(hipStreamEndCapture(stream, &graph));
    ECG_CHECK(key_now() == before, "device iteration graph: a buffer was reallocated during capture");
    hipError_t e = hipGraphInstantiate(&d.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    ECG_HIP(e);
    d.key = before;
    d.okey = okey;
    residual_calls++;
    ECG_HIP(hipGraphLaunch(d.exec, stream));
  }
  ECG_HIP(hipMemcpyAsync(d.info_host, d.scal.d() + S_INFO, sizeof(double) * 8, hipMemcpyDeviceToHost, stream));
}

// cc_iterations! with device-resident iterations (FP64): one graph launch and one host
// synchronisation per iteration.
int Context::dev_solve(int method, const SolverOpts& o, double* energies, int* converged, FILE* log) {
  dev_reset(o);
  dv.eager_done = false;
  iter = 0;
  int it = 0, conv = 0;
  double info[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  fp64_iterations = 0;
  for (it = 1; it <= o.maxit; ++it) {
    dev_launch(method, o);
    ECG_HIP(hipStreamSynchronize(stream));
    std::memcpy(info, dv.info_host, sizeof(info));
    iter++;
    fp64_iterations++;
    if (log) {
      std::fprintf(log, "%4d %12.8f %16.10f %12.3e %12.3e dev\n", it, info[6], info[0], info[4], info[5]);
      std::fflush(log);
    }
    if (info[7] != 0.0) {
      conv = 1;
      break;
    }
  }
  if (it > o.maxit) it = o.maxit;
  for (int k = 0; k < 6; ++k) energies[k] = info[k];
  *converged = conv;
  return it;
}

Context::DevIter::~DevIter() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (info_host) (void)hipHostFree(info_host);
  if (done) (void)hipEventDestroy(done);
}

}  // namespace ecg
