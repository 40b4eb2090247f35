// extern "C" boundary of the library (declarations and reference citations in include/elemco_hip.h).
#include "../../include/elemco_hip.h"

#include <algorithm>
#include <cstdlib>
#include <atomic>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "ecg_cc.h"
#include "ecg_grad.h"
#include "ecg_lambda.h"
#include "ecg_triples.h"

struct ecg_ctx {
  ecg::Context c;
};

namespace {

thread_local std::string g_create_error;

ecg::SolverOpts to_opts(const ecg_solver_opts* o) {
  ecg::SolverOpts s;
  if (!o) return s;
  s.thr = o->thr; s.conven = o->conven; s.shifts = o->shifts; s.shiftp = o->shiftp;
  s.resthr = o->resthr; s.maxit = o->maxit; s.maxdiis = o->maxdiis; s.refine_every = o->refine_every; s.graphs = o->graphs;
  return s;
}

ecg::Prec to_prec(int p) {
  ECG_CHECK(p == ECG_FP64 || p == ECG_FP32 || p == ECG_BF16X3, "unknown precision");
  return static_cast<ecg::Prec>(p);
}

void check_method(int m) { ECG_CHECK(m == ECG_CCSD || m == ECG_DCSD, "method must be ECG_CCSD or ECG_DCSD"); }

template <class F>
int guarded(ecg_ctx* ctx, F&& f) {
  if (!ctx) return -1;
  try {
    ECG_HIP(hipSetDevice(ctx->c.device));
    f(ctx->c);
    ctx->c.last_error.clear();
    return 0;
  } catch (const ecg::Error& e) {
    ctx->c.last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->c.last_error = e.what();
    return -9;
  }
}

void h2d(void* d, const void* h, size_t bytes, hipStream_t s) {
  ECG_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
}
void d2h(void* h, const void* d, size_t bytes, hipStream_t s) {
  ECG_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
}

}  // namespace

extern "C" {

void ecg_default_opts(ecg_solver_opts* o) {
  if (!o) return;
  o->thr = 1e-10; o->conven = 0.1; o->shifts = 0.15; o->shiftp = 0.2; o->resthr = 10.0;
  o->maxit = 50; o->maxdiis = 6; o->refine_every = 0; o->graphs = 0;
}

int ecg_create(ecg_ctx** out, int nocc, int nvirt, int naux, const double* h, const double* BL, const double* BR,
               double e0, int device, size_t workspace_bytes) {
  if (!out) return -1;
  *out = nullptr;
  auto ctx = std::make_unique<ecg_ctx>();
  try {
    ECG_CHECK(h && BL, "h and BL must not be NULL");
    int ndev = 0;
    ECG_HIP(hipGetDeviceCount(&ndev));
    ECG_CHECK(device >= 0 && device < ndev, "no such HIP device");
    ctx->c.device = device;
    ctx->c.init(nocc, nvirt, naux, h, BL, BR, e0, workspace_bytes);
  } catch (const ecg::Error& e) {
    g_create_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return -9;
  }
  *out = ctx.release();
  return 0;
}

int ecg_destroy(ecg_ctx* ctx) {
  if (!ctx) return -1;
  (void)hipSetDevice(ctx->c.device);
  delete ctx;
  return 0;
}

const char* ecg_last_error(const ecg_ctx* ctx) {
  if (!ctx) return g_create_error.c_str();
  return ctx->c.last_error.c_str();
}

int ecg_fock(ecg_ctx* ctx, double* fock, double* eps, double* ehf) {
  return guarded(ctx, [&](ecg::Context& c) {
    if (fock) d2h(fock, c.fock.p, sizeof(double) * c.n * c.n, c.stream);
    if (eps) d2h(eps, c.eps.p, sizeof(double) * c.n, c.stream);
    ECG_HIP(hipStreamSynchronize(c.stream));
    if (ehf) *ehf = c.hf_energy();
  });
}

int ecg_mp2(ecg_ctx* ctx, double* T1, double* T2, double* emp2) {
  return guarded(ctx, [&](ecg::Context& c) {
    ecg::DevBuf d1(sizeof(double) * c.nvo()), d2(sizeof(double) * c.nvvoo());
    double e[3];
    c.mp2_guess(d1.d(), d2.d(), e);
    if (T1) d2h(T1, d1.p, d1.bytes, c.stream);
    if (T2) d2h(T2, d2.p, d2.bytes, c.stream);
    ECG_HIP(hipStreamSynchronize(c.stream));
    if (emp2) std::memcpy(emp2, e, sizeof(e));
  });
}

int ecg_cc_residual(ecg_ctx* ctx, int method, int precision, const double* T1, const double* T2, double* R1,
                    double* R2) {
  return guarded(ctx, [&](ecg::Context& c) {
    check_method(method);
    ECG_CHECK(T1 && T2 && R1 && R2, "NULL amplitude or residual pointer");
    ecg::DevBuf d1(sizeof(double) * c.nvo()), d2(sizeof(double) * c.nvvoo());
    ecg::DevBuf r1(d1.bytes), r2(d2.bytes);
    h2d(d1.p, T1, d1.bytes, c.stream);
    h2d(d2.p, T2, d2.bytes, c.stream);
    c.residual(method, to_prec(precision), d1.d(), d2.d(), r1.d(), r2.d());
    d2h(R1, r1.p, r1.bytes, c.stream);
    d2h(R2, r2.p, r2.bytes, c.stream);
    ECG_HIP(hipStreamSynchronize(c.stream));
  });
}

int ecg_cc_residual_dev(ecg_ctx* ctx, int method, int precision, const double* dT1, const double* dT2, double* dR1,
                        double* dR2) {
  return guarded(ctx, [&](ecg::Context& c) {
    check_method(method);
    ECG_CHECK(dT1 && dT2 && dR1 && dR2, "NULL device pointer");
    c.residual(method, to_prec(precision), dT1, dT2, dR1, dR2);
    ECG_HIP(hipStreamSynchronize(c.stream));
  });
}

int ecg_cc_state_init(ecg_ctx* ctx, const double* T1, const double* T2, const ecg_solver_opts* opts) {
  return guarded(ctx, [&](ecg::Context& c) {
    ECG_CHECK((T1 == nullptr) == (T2 == nullptr), "give both T1 and T2 or neither");
    c.state_init(T1, T2, to_opts(opts));
  });
}

int ecg_cc_state_step(ecg_ctx* ctx, int method, int precision, const ecg_solver_opts* opts, double* info) {
  return guarded(ctx, [&](ecg::Context& c) {
    check_method(method);
    ECG_CHECK(c.T1.p != nullptr, "ecg_cc_state_init has not been called");
    double tmp[8];
    c.state_step(method, to_prec(precision), to_opts(opts), info ? info : tmp);
    ECG_HIP(hipStreamSynchronize(c.stream));
  });
}

int ecg_cc_state_get(ecg_ctx* ctx, double* T1, double* T2) {
  return guarded(ctx, [&](ecg::Context& c) {
    ECG_CHECK(c.T1.p != nullptr, "ecg_cc_state_init has not been called");
    if (T1) d2h(T1, c.T1.p, sizeof(double) * c.nvo(), c.stream);
    if (T2) d2h(T2, c.T2.p, sizeof(double) * c.nvvoo(), c.stream);
    ECG_HIP(hipStreamSynchronize(c.stream));
  });
}

int ecg_cc_solve(ecg_ctx* ctx, int method, int precision, const ecg_solver_opts* opts, double* T1, double* T2,
                 double* energies, int* iterations, int* converged) {
  return guarded(ctx, [&](ecg::Context& c) {
    check_method(method);
    ecg::SolverOpts o = to_opts(opts);
    c.state_init(nullptr, nullptr, o);
    double e[6];
    int conv = 0;
    int it = c.solve(method, to_prec(precision), o, e, &conv, std::getenv("ECG_ITERATION_LOG") ? stderr : nullptr);
    if (T1) d2h(T1, c.T1.p, sizeof(double) * c.nvo(), c.stream);
    if (T2) d2h(T2, c.T2.p, sizeof(double) * c.nvvoo(), c.stream);
    ECG_HIP(hipStreamSynchronize(c.stream));
    if (energies) std::memcpy(energies, e, sizeof(e));
    if (iterations) *iterations = it;
    if (converged) *converged = conv;
  });
}

int ecg_cc_solve_batch(int nsys, ecg_ctx* const* ctxs, int method, int precision, const ecg_solver_opts* opts,
                       int max_threads, double* energies, int* iterations, int* converged, int* status) {
  if (nsys < 0 || (nsys > 0 && (!ctxs || !energies || !iterations || !converged))) return -1;
  for (int k = 0; k < nsys; ++k)
    if (!ctxs[k]) return -1;
  std::vector<int> st(nsys, 0);
  // longest first (cost of an iteration ~ v^4 (o^2 + naux)), so the large systems do not
  // start last and leave the device to one stream at the end
  std::vector<int> order(nsys);
  std::vector<double> cost(nsys);
  for (int k = 0; k < nsys; ++k) {
    const ecg::Context& c = ctxs[k]->c;
    const double v = c.nv;
    cost[k] = v * v * v * v * ((double)c.no * c.no + c.nq);
    order[k] = k;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (int q = next.fetch_add(1); q < nsys; q = next.fetch_add(1)) {
      const int k = order[q];
      st[k] = guarded(ctxs[k], [&](ecg::Context& c) {
        check_method(method);
        ecg::SolverOpts o = to_opts(opts);
        c.state_init(nullptr, nullptr, o);
        int conv = 0;
        iterations[k] = c.solve(method, to_prec(precision), o, energies + 6 * k, &conv, nullptr);
        converged[k] = conv;
        ECG_HIP(hipStreamSynchronize(c.stream));
      });
    }
  };
  int nt = max_threads > 0 ? max_threads : 16;
  nt = std::max(1, std::min(nt, nsys));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  int first = 0;
  for (int k = 0; k < nsys; ++k) {
    if (status) status[k] = st[k];
    if (st[k] != 0 && first == 0) first = st[k];
  }
  return first;
}

int ecg_pert_triples(ecg_ctx* ctx, int lambda, int precision, const double* T1, const double* T2, const double* U1,
                     const double* U2, double* ET, double* ETb) {
  return guarded(ctx, [&](ecg::Context& c) {
    ECG_CHECK(T1 && T2, "NULL amplitudes");
    ECG_CHECK(!lambda || (U1 && U2), "Lambda-(T) needs U1 and U2");
    double e[2];
    ecg::pert_triples(c, lambda != 0, to_prec(precision), T1, T2, U1, U2, 0, -1, e);
    if (ET) *ET = e[0];
    if (ETb) *ETb = e[1];
  });
}

int ecg_pert_triples_range(ecg_ctx* ctx, int lambda, int precision, const double* T1, const double* T2,
                           const double* U1, const double* U2, long first, long count, double* ET, double* ETb) {
  return guarded(ctx, [&](ecg::Context& c) {
    ECG_CHECK(T1 && T2, "NULL amplitudes");
    ECG_CHECK(!lambda || (U1 && U2), "Lambda-(T) needs U1 and U2");
    ECG_CHECK(first >= 0 && count >= 0, "triplet range must be non-negative");
    double e[2];
    ecg::pert_triples(c, lambda != 0, to_prec(precision), T1, T2, U1, U2, first, count, e);
    if (ET) *ET = e[0];
    if (ETb) *ETb = e[1];
  });
}

int ecg_lambda_residual(ecg_ctx* ctx, int method, int precision, const double* T1, const double* T2,
                        const double* U1, const double* U2, double* R1, double* R2) {
  return guarded(ctx, [&](ecg::Context& c) {
    check_method(method);
    ECG_CHECK(T1 && T2 && U1 && U2 && R1 && R2, "NULL pointer");
    ecg::lambda_residual_host(c, method, to_prec(precision), T1, T2, U1, U2, R1, R2);
  });
}

int ecg_lambda_solve(ecg_ctx* ctx, int method, int precision, const ecg_solver_opts* opts, const double* T1,
                     const double* T2, double* U1, double* U2, int* iterations, int* converged) {
  return guarded(ctx, [&](ecg::Context& c) {
    check_method(method);
    ECG_CHECK(T1 && T2 && U1 && U2, "NULL pointer");
    int it = 0, conv = 0;
    ecg::lambda_solve(c, method, to_prec(precision), to_opts(opts), T1, T2, U1, U2, &it, &conv);
    if (iterations) *iterations = it;
    if (converged) *converged = conv;
  });
}

int ecg_energy_grad(ecg_ctx* ctx, int method, const double* T1, const double* T2, const double* U1, const double* U2,
                    double* dE_dh, double* dE_dB) {
  return guarded(ctx, [&](ecg::Context& c) {
    check_method(method);
    ECG_CHECK(T1 && T2 && U1 && U2, "NULL pointer");
    ecg::energy_grad(c, method, T1, T2, U1, U2, dE_dh, dE_dB);
  });
}

int ecg_one_rdm(ecg_ctx* ctx, const double* T1, const double* T2, const double* U1, const double* U2, double* D1,
                double* dD1) {
  return guarded(ctx, [&](ecg::Context& c) {
    ECG_CHECK(T1 && T2 && U1 && U2, "NULL pointer");
    ecg::one_rdm_host(c, T1, T2, U1, U2, D1, dD1);
  });
}

int ecg_profile(ecg_ctx* ctx, int enable) {
  return guarded(ctx, [&](ecg::Context& c) {
    ECG_HIP(hipStreamSynchronize(c.stream));
    c.ops.gemm.collect_timing();
    c.ops.gemm.timed = enable != 0;
    c.ops.gemm.timed_ms = 0.0;
    c.ops.gemm.timed_launches = 0;
    c.ops.gemm.flops = 0;
    c.ops.gemm.bytes = 0.0;
  });
}

int ecg_stats(ecg_ctx* ctx, double* stats) {
  return guarded(ctx, [&](ecg::Context& c) {
    ECG_CHECK(stats, "NULL stats");
    ECG_HIP(hipStreamSynchronize(c.stream));
    c.ops.gemm.collect_timing();
    stats[0] = c.last_residual_ms;
    stats[1] = c.last_ladder_ms;
    stats[2] = c.ops.gemm.timed_ms;
    stats[3] = (double)c.ops.gemm.timed_launches;
    stats[4] = (double)c.ops.gemm.flops;
    stats[5] = (double)c.fp64_iterations;
    stats[6] = c.ops.gemm.bytes;
    stats[7] = (double)c.triples_last_batch;
  });
}

}  // extern "C"
