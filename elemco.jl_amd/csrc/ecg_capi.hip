// extern "C" boundary of the library (declarations and reference citations in include/elemco_hip.h).
#include "../../include/elemco_hip.h"

#include <algorithm>
#include <cstdlib>
#include <atomic>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ecg_cc.h"
#include "ecg_dfprep.h"
#include "ecg_fcidump.h"
#include "ecg_grad.h"
#include "ecg_lambda.h"
#include "ecg_triples.h"
#include "ecg_ucc.h"

struct ecg_ctx {
  ecg::Context c;
};

struct ecg_uctx {
  ecg::UContext c;
};

struct ecg_fcidump {
  ecg::Fcidump fd;
  bool df_done = false;
  double df_rtol = 0.0;
  ecg::DfFactors df;
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

template <class F>
int uguarded(ecg_uctx* ctx, F&& f) {
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

void check_umethod(int m) {
  ECG_CHECK(m == ECG_UCCSD || m == ECG_UCCD || m == ECG_UDCSD, "method must be ECG_UCCSD, ECG_UCCD or ECG_UDCSD");
}

// Region of the allocation-free entry point: one copy per GPU, held from the
// library's first use of that GPU until the process ends.  Host code never passes
// `scratch` itself to a kernel or a copy (on the host the name only stands for the
// symbol): it uses the device address looked up for the current device.
static __device__ __align__(256) unsigned char scratch[ECG_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be the current device; looked
// up on the first call for that device and kept.
void* scratch_on(int device) {
  ECG_CHECK(device >= 0 && device < kMaxDevices, "device index out of range for the allocation-free scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    ECG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ECG_CHECK(p != nullptr, "hipGetSymbolAddress returned no address for the allocation-free scratch");
    g_scratch_dev[device] = p;
  }
  return g_scratch_dev[device];
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

}  // extern "C"

namespace {

// host-side calls without a context: errors go to the thread's create-error slot
template <class F>
int host_guarded(F&& f) {
  try {
    return f();
  } catch (const ecg::Error& e) {
    g_create_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return -9;
  }
}

int copy_text(const std::string& s, char* buf, int buflen) {
  if (!buf || buflen < (int)s.size() + 1) {
    g_create_error = "buffer too small";
    return -2;
  }
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return 0;
}

const ecg::DfFactors& fcidump_df(ecg_fcidump* f, double rtol) {
  if (rtol <= 0.0) rtol = 1e-14;
  if (!f->df_done || f->df_rtol != rtol) {
    f->df = ecg::df_factorise(f->fd.eri, f->fd.norb, !f->fd.simtra(), rtol);
    f->df_done = true;
    f->df_rtol = rtol;
  }
  return f->df;
}

}  // namespace

extern "C" {

int ecg_fcidump_read(const char* path, ecg_fcidump** out) {
  if (!out) return -1;
  *out = nullptr;
  return host_guarded([&] {
    ECG_CHECK(path, "path must not be NULL");
    auto f = std::make_unique<ecg_fcidump>();
    f->fd = ecg::read_fcidump(path);
    *out = f.release();
    return 0;
  });
}

int ecg_fcidump_from_integrals(int norb, int nelec, int ms2, int st, const double* h, const double* eri, double e0,
                               ecg_fcidump** out) {
  if (!out) return -1;
  *out = nullptr;
  return host_guarded([&] {
    ECG_CHECK(norb > 0 && h && eri, "norb must be positive and h, eri given");
    ECG_CHECK(nelec >= 0 && ms2 >= 0 && st >= 0, "NELEC, MS2 and ST must be non-negative");
    auto f = std::make_unique<ecg_fcidump>();
    ecg::Fcidump& fd = f->fd;
    const size_t n = (size_t)norb;
    fd.norb = norb;
    fd.int0 = e0;
    fd.h.assign(h, h + n * n);
    fd.eri.assign(eri, eri + n * n * n * n);
    auto put = [&](const std::string& key, std::vector<std::string> vals) {
      fd.head[key] = std::move(vals);
      fd.head_type[key] = 'i';
    };
    put("NORB", {std::to_string(norb)});
    put("NELEC", {std::to_string(nelec)});
    put("MS2", {std::to_string(ms2)});
    put("ISYM", {"1"});
    put("ORBSYM", std::vector<std::string>(n, "1"));
    if (st > 0) put("ST", {std::to_string(st)});
    *out = f.release();
    return 0;
  });
}

int ecg_fcidump_free(ecg_fcidump* fd) {
  if (!fd) return -1;
  delete fd;
  return 0;
}

int ecg_fcidump_info(const ecg_fcidump* f, int* norb, int* nelec, int* ms2, int* st, double* e0) {
  if (!f) return -1;
  return host_guarded([&] {
    if (norb) *norb = f->fd.norb;
    if (nelec) *nelec = (int)f->fd.head_int("NELEC", -1);
    if (ms2) *ms2 = (int)f->fd.head_int("MS2", 0);
    if (st) *st = (int)f->fd.head_int("ST", 0);
    if (e0) *e0 = f->fd.int0;
    return 0;
  });
}

int ecg_fcidump_header(const ecg_fcidump* f, const char* key, char* buf, int buflen, char* type) {
  if (!f || !key) return -1;
  return host_guarded([&] {
    auto it = f->fd.head.find(key);
    if (it == f->fd.head.end()) return 1;
    std::string s;
    for (size_t k = 0; k < it->second.size(); ++k) s += (k ? "," : "") + it->second[k];
    if (type) *type = f->fd.head_type.at(key);
    return copy_text(s, buf, buflen);
  });
}

int ecg_fcidump_header_key(const ecg_fcidump* f, int index, char* buf, int buflen) {
  if (!f || index < 0) return -1;
  return host_guarded([&] {
    if (index >= (int)f->fd.head.size()) return 1;
    auto it = f->fd.head.begin();
    std::advance(it, index);
    return copy_text(it->first, buf, buflen);
  });
}

int ecg_fcidump_integrals(const ecg_fcidump* f, double* h, double* eri) {
  if (!f) return -1;
  if (h) std::memcpy(h, f->fd.h.data(), sizeof(double) * f->fd.h.size());
  if (eri) std::memcpy(eri, f->fd.eri.data(), sizeof(double) * f->fd.eri.size());
  return 0;
}

int ecg_fcidump_df(ecg_fcidump* f, double rtol, int* naux, double* BL, double* BR, int* has_br) {
  if (!f) return -1;
  return host_guarded([&] {
    const ecg::DfFactors& d = fcidump_df(f, rtol);
    if (naux) *naux = d.naux;
    if (has_br) *has_br = d.BR.empty() ? 0 : 1;
    if (BL) std::memcpy(BL, d.BL.data(), sizeof(double) * d.BL.size());
    if (BR) std::memcpy(BR, (d.BR.empty() ? d.BL : d.BR).data(), sizeof(double) * d.BL.size());
    return 0;
  });
}

int ecg_fcidump_freeze(ecg_fcidump* f, const int* orbs, int norbs, int* ncore, int* nfrozvirt) {
  if (!f || norbs < 0 || (norbs > 0 && !orbs)) return -1;
  return host_guarded([&] {
    const ecg::FreezeCount c = ecg::freeze_orbs_in_dump(f->fd, std::vector<int>(orbs, orbs + norbs));
    if (c.ncore + c.nvirt > 0) {
      f->df_done = false;
      f->df = ecg::DfFactors();
    }
    if (ncore) *ncore = c.ncore;
    if (nfrozvirt) *nfrozvirt = c.nvirt;
    return 0;
  });
}

int ecg_fcidump_write(const ecg_fcidump* f, const char* path, double tol) {
  if (!f || !path) return -1;
  return host_guarded([&] {
    ecg::write_fcidump(f->fd, path, tol < 0.0 ? 1e-12 : tol);
    return 0;
  });
}

int ecg_df_freeze(int n, int nocc, int naux, const double* h, const double* BL, const double* BR, double e0,
                  const int* orbs, int norbs, int* nkept, int* ncore, double* h_out, double* BL_out, double* BR_out,
                  double* e0_out) {
  return host_guarded([&] {
    ECG_CHECK(norbs >= 0 && (norbs == 0 || orbs), "ecg_df_freeze: invalid orbital list");
    const ecg::DfFreeze f = ecg::df_freeze(n, nocc, naux, h, BL, BR, e0, std::vector<int>(orbs, orbs + norbs));
    if (nkept) *nkept = f.nkept;
    if (ncore) *ncore = f.ncore;
    if (e0_out) *e0_out = f.e0;
    if (h_out) std::memcpy(h_out, f.h.data(), sizeof(double) * f.h.size());
    if (BL_out) std::memcpy(BL_out, f.BL.data(), sizeof(double) * f.BL.size());
    if (BR_out && BR) std::memcpy(BR_out, f.BR.data(), sizeof(double) * f.BR.size());
    return 0;
  });
}

int ecg_df_ao_to_mo(const double* C, int nao, int nmo, const double* AAP, const double* PQ, int nfit, double tol,
                    int device, int* naux, double* B) {
  return host_guarded([&] {
    ECG_CHECK(C && AAP && PQ && B && naux, "ecg_df_ao_to_mo: NULL argument");
    ECG_CHECK(nao > 0 && nmo > 0 && nfit > 0, "ecg_df_ao_to_mo: sizes must be positive");
    int ndev = 0;
    ECG_HIP(hipGetDeviceCount(&ndev));
    ECG_CHECK(device >= 0 && device < ndev, "no such HIP device");
    int r = 0;
    const std::vector<double> M = ecg::sqrtinvchol(PQ, nfit, tol, r);
    ecg::df_ao_to_mo_device(device, C, nao, nmo, AAP, nfit, M.data(), r, B);
    *naux = r;
    return 0;
  });
}

int ecg_create_from_ao(ecg_ctx** out, int nocc, const double* C, int nao, int nmo, const double* h_ao,
                       const double* AAP, const double* PQ, int nfit, double tol, double e0, int device,
                       size_t workspace_bytes) {
  if (!out) return -1;
  *out = nullptr;
  std::vector<double> B, h;
  int naux = 0;
  int st = host_guarded([&] {
    ECG_CHECK(h_ao, "ecg_create_from_ao: NULL argument");
    ECG_CHECK(nocc > 0 && nocc < nmo, "ecg_create_from_ao: nocc must lie between 1 and nmo - 1");
    B.resize((size_t)nmo * nmo * nfit);
    return 0;
  });
  if (st != 0) return st;
  st = ecg_df_ao_to_mo(C, nao, nmo, AAP, PQ, nfit, tol, device, &naux, B.data());
  if (st != 0) return st;
  // h[p,q] = sum C[mu,p] h_ao[mu,nu] C[nu,q]
  std::vector<double> t((size_t)nao * nmo, 0.0);
  for (int q = 0; q < nmo; ++q)
    for (int nu = 0; nu < nao; ++nu) {
      const double c = C[nu + (size_t)nao * q];
      for (int mu = 0; mu < nao; ++mu) t[mu + (size_t)nao * q] += h_ao[mu + (size_t)nao * nu] * c;
    }
  h.assign((size_t)nmo * nmo, 0.0);
  for (int q = 0; q < nmo; ++q)
    for (int p = 0; p < nmo; ++p) {
      double s = 0.0;
      for (int mu = 0; mu < nao; ++mu) s += C[mu + (size_t)nao * p] * t[mu + (size_t)nao * q];
      h[p + (size_t)nmo * q] = s;
    }
  return ecg_create(out, nocc, nmo - nocc, naux, h.data(), B.data(), nullptr, e0, device, workspace_bytes);
}

int ecg_dims(const ecg_ctx* ctx, int* nocc, int* nvirt, int* naux) {
  if (!ctx) return -1;
  if (nocc) *nocc = ctx->c.no;
  if (nvirt) *nvirt = ctx->c.nv;
  if (naux) *naux = ctx->c.nq;
  return 0;
}

int ecg_create_from_fcidump(ecg_ctx** out, const char* path, double rtol, int device, size_t workspace_bytes) {
  if (!out) return -1;
  *out = nullptr;
  ecg_fcidump* f = nullptr;
  int st = ecg_fcidump_read(path, &f);
  if (st != 0) return st;
  std::unique_ptr<ecg_fcidump> hold(f);
  return ecg_create_from_fcidump_handle(out, f, rtol, device, workspace_bytes);
}

int ecg_create_from_fcidump_handle(ecg_ctx** out, ecg_fcidump* f, double rtol, int device, size_t workspace_bytes) {
  if (!out || !f) return -1;
  *out = nullptr;
  int st = host_guarded([&] {
    const long nelec = f->fd.head_int("NELEC", -1);
    ECG_CHECK(nelec >= 0, "NELEC option not found in fcidump");
    ECG_CHECK(nelec % 2 == 0 && f->fd.head_int("MS2", 0) == 0,
              "the closed-shell path needs an even NELEC and MS2 = 0");
    const int nocc = (int)(nelec / 2);
    ECG_CHECK(nocc > 0 && nocc < f->fd.norb, "NELEC / 2 must lie between 1 and NORB - 1");
    fcidump_df(f, rtol);
    return 0;
  });
  if (st != 0) return st;
  const ecg::DfFactors& d = f->df;
  const int n = f->fd.norb, nocc = (int)(f->fd.head_int("NELEC", 0) / 2);
  return ecg_create(out, nocc, n - nocc, d.naux, f->fd.h.data(), d.BL.data(), d.BR.empty() ? nullptr : d.BR.data(),
                    f->fd.int0, device, workspace_bytes);
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

int ecg_cc_residual_dev_noalloc(ecg_ctx* ctx, int method, int precision, const double* dT1, const double* dT2,
                                double* dR1, double* dR2) {
  return guarded(ctx, [&](ecg::Context& c) {
    check_method(method);
    ECG_CHECK(precision == ECG_FP64, "the allocation-free residual computes in FP64 only (precision ECG_FP64)");
    ECG_CHECK(dT1 && dT2 && dR1 && dR2, "NULL device pointer");
    void* region = scratch_on(c.device);   // guarded() made c.device current
    c.residual_noalloc(method, dT1, dT2, dR1, dR2, region, ECG_NOALLOC_SCRATCH_BYTES);
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
  // Concurrent systems share the CUs, so each product splits K only as far as a quarter of
  // the device needs (64 CUs): on the 64-system batch (nocc 4..16) 1.92 s against 2.00 s with
  // splits sized for all 256 CUs (32: 2.03 s, 16: 2.34 s).  Restored on return.
  std::vector<int> cus_before(nsys);
  for (int k = 0; k < nsys; ++k) {
    cus_before[k] = ctxs[k]->c.ops.gemm.cus;
    if (nsys > 1) ctxs[k]->c.ops.gemm.cus = 64;
  }
  struct Restore {
    int n; ecg_ctx* const* c; const std::vector<int>& v;
    ~Restore() { for (int k = 0; k < n; ++k) c[k]->c.ops.gemm.cus = v[k]; }
  } restore{nsys, ctxs, cus_before};
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
  if (opts && opts->graphs && precision == ECG_FP64) {
    // device-resident iterations: one host thread drives every system, one graph launch
    // per system and iteration; an iteration's info is read when its event has completed,
    // and the system's next iteration is launched at once (no barrier across systems)
    const ecg::SolverOpts o = to_opts(opts);
    std::vector<int> it(nsys, 0), active(nsys, 0);
    for (int q = 0; q < nsys; ++q) {
      const int k = order[q];
      st[k] = guarded(ctxs[k], [&](ecg::Context& c) {
        check_method(method);
        c.state_init(nullptr, nullptr, o);
        c.dev_reset(o);
        c.dv.eager_done = false;
        c.fp64_iterations = 0;
        c.dev_launch(method, o);
        ECG_HIP(hipEventRecord(c.dv.done, c.stream));
      });
      active[k] = st[k] == 0;
      converged[k] = 0;
      iterations[k] = 0;
    }
    int left = 0;
    for (int k = 0; k < nsys; ++k) left += active[k];
    while (left > 0) {
      bool progressed = false;
      for (int q = 0; q < nsys; ++q) {
        const int k = order[q];
        if (!active[k]) continue;
        bool finished = false;
        st[k] = guarded(ctxs[k], [&](ecg::Context& c) {
          const hipError_t e = hipEventQuery(c.dv.done);
          if (e == hipErrorNotReady) return;
          ECG_HIP(e);
          progressed = true;
          const double* info = c.dv.info_host;
          it[k]++;
          c.fp64_iterations++;
          if (info[7] != 0.0 || it[k] >= o.maxit) {
            for (int j = 0; j < 6; ++j) energies[6 * k + j] = info[j];
            converged[k] = info[7] != 0.0;
            iterations[k] = it[k];
            finished = true;
            return;
          }
          c.dev_launch(method, o);
          ECG_HIP(hipEventRecord(c.dv.done, c.stream));
        });
        if (finished || st[k] != 0) {
          active[k] = 0;
          left--;
        }
      }
      if (!progressed) std::this_thread::yield();
    }
    int first = 0;
    for (int k = 0; k < nsys; ++k) {
      if (status) status[k] = st[k];
      if (st[k] != 0 && first == 0) first = st[k];
    }
    return first;
  }
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
    ECG_CHECK(c.hermitian, "ecg_energy_grad: two-sided DF factors (BR given): use ecg_energy_grad_lr");
    ecg::energy_grad(c, method, T1, T2, U1, U2, dE_dh, dE_dB, nullptr);
  });
}

int ecg_energy_grad_lr(ecg_ctx* ctx, int method, const double* T1, const double* T2, const double* U1,
                       const double* U2, double* dE_dh, double* dE_dBL, double* dE_dBR) {
  return guarded(ctx, [&](ecg::Context& c) {
    check_method(method);
    ECG_CHECK(T1 && T2 && U1 && U2, "NULL pointer");
    ECG_CHECK(!c.hermitian, "ecg_energy_grad_lr: the context has one DF factor (BR = NULL): use ecg_energy_grad");
    ecg::energy_grad(c, method, T1, T2, U1, U2, dE_dh, dE_dBL, dE_dBR);
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

// ---------------------------------------------------------------- open-shell (UCCSD / UCCD / UDCSD)
int ecg_ucreate(ecg_uctx** out, int nocca, int noccb, int norb, int naux, const double* h, const double* B, double e0,
                int device, size_t workspace_bytes) {
  if (!out) return -1;
  *out = nullptr;
  auto ctx = std::make_unique<ecg_uctx>();
  ctx->c.device = device;
  const int st = host_guarded([&] {
    ECG_CHECK(h && B, "NULL integrals");
    ctx->c.init(nocca, noccb, norb, naux, h, B, e0, workspace_bytes);
    return 0;
  });
  if (st != 0) return st;
  *out = ctx.release();
  return 0;
}

int ecg_ucreate_from_fcidump(ecg_uctx** out, const char* path, int ms2, double rtol, int device,
                             size_t workspace_bytes) {
  if (!out || !path) return -1;
  *out = nullptr;
  ecg_fcidump* f = nullptr;
  int st = ecg_fcidump_read(path, &f);
  if (st != 0) return st;
  std::unique_ptr<ecg_fcidump> hold(f);
  return host_guarded([&] {
    const long nelec = f->fd.head_int("NELEC", -1);
    ECG_CHECK(nelec >= 0, "NELEC option not found in fcidump");
    const long m = ms2 >= 0 ? ms2 : f->fd.head_int("MS2", 0);
    ECG_CHECK(m >= 0 && m <= nelec && (nelec + m) % 2 == 0, "NELEC and MS2 do not give whole alpha/beta counts");
    ECG_CHECK(!f->fd.simtra(), "the open-shell path takes Hermitian integrals (ST = 0)");
    const ecg::DfFactors& d = fcidump_df(f, rtol);
    ECG_CHECK(d.BR.empty(), "the open-shell path takes Hermitian integrals");
    auto ctx = std::make_unique<ecg_uctx>();
    ctx->c.device = device;
    ctx->c.init((int)((nelec + m) / 2), (int)((nelec - m) / 2), f->fd.norb, d.naux, f->fd.h.data(), d.BL.data(),
                f->fd.int0, workspace_bytes);
    *out = ctx.release();
    return 0;
  });
}

int ecg_udestroy(ecg_uctx* ctx) {
  if (!ctx) return -1;
  (void)hipSetDevice(ctx->c.device);
  delete ctx;
  return 0;
}

const char* ecg_ulast_error(const ecg_uctx* ctx) {
  if (!ctx) return g_create_error.c_str();
  return ctx->c.last_error.c_str();
}

int ecg_udims(const ecg_uctx* ctx, int* norb, int* nocca, int* noccb, int* naux) {
  if (!ctx) return -1;
  if (norb) *norb = ctx->c.n;
  if (nocca) *nocca = ctx->c.na;
  if (noccb) *noccb = ctx->c.nb;
  if (naux) *naux = ctx->c.nq;
  return 0;
}

int ecg_ufock(ecg_uctx* ctx, double* focka, double* fockb, double* ehf) {
  return uguarded(ctx, [&](ecg::UContext& c) {
    const int n = c.n, N = c.N;
    std::vector<double> f((size_t)N * N);
    ECG_HIP(hipMemcpyAsync(f.data(), c.fock.p, sizeof(double) * f.size(), hipMemcpyDeviceToHost, c.stream));
    ECG_HIP(hipStreamSynchronize(c.stream));
    // spin orbital of spatial orbital p with spin s
    auto so = [&](int p, int s) {
      for (int k = 0; k < N; ++k)
        if (c.sp[k] == p && c.sg[k] == s) return k;
      return -1;
    };
    for (int s = 0; s < 2; ++s) {
      double* dst = s == 0 ? focka : fockb;
      if (!dst) continue;
      for (int q = 0; q < n; ++q)
        for (int p = 0; p < n; ++p) dst[p + (size_t)n * q] = f[(size_t)so(p, s) + (size_t)N * so(q, s)];
    }
    if (ehf) *ehf = c.ehf;
  });
}

int ecg_ump2(ecg_uctx* ctx, double* T1a, double* T1b, double* T2a, double* T2b, double* T2ab, double* emp2) {
  return uguarded(ctx, [&](ecg::UContext& c) {
    ecg::DevBuf d1(sizeof(double) * c.n1()), d2(sizeof(double) * c.n2());
    double e[1];
    c.mp2(d1.d(), d2.d(), e);
    std::vector<double> h1(c.n1()), h2(c.n2());
    d2h(h1.data(), d1.p, d1.bytes, c.stream);
    d2h(h2.data(), d2.p, d2.bytes, c.stream);
    ECG_HIP(hipStreamSynchronize(c.stream));
    c.to_blocks(h1.data(), h2.data(), T1a, T1b, T2a, T2b, T2ab);
    if (emp2) *emp2 = e[0];
  });
}

int ecg_ucc_residual(ecg_uctx* ctx, int method, const double* T1a, const double* T1b, const double* T2a,
                     const double* T2b, const double* T2ab, double* R1a, double* R1b, double* R2a, double* R2b,
                     double* R2ab) {
  return uguarded(ctx, [&](ecg::UContext& c) {
    check_umethod(method);
    ECG_CHECK(T1a && T1b && T2a && T2b && T2ab, "NULL amplitude pointer");
    std::vector<double> h1(c.n1()), h2(c.n2());
    c.from_blocks(T1a, T1b, T2a, T2b, T2ab, h1.data(), h2.data());
    if (method == ECG_UCCD) std::fill(h1.begin(), h1.end(), 0.0);
    ecg::DevBuf d1(sizeof(double) * c.n1()), d2(sizeof(double) * c.n2()), r1(d1.bytes), r2(d2.bytes);
    h2d(d1.p, h1.data(), d1.bytes, c.stream);
    h2d(d2.p, h2.data(), d2.bytes, c.stream);
    c.residual(method, d1.d(), d2.d(), r1.d(), r2.d());
    d2h(h1.data(), r1.p, r1.bytes, c.stream);
    d2h(h2.data(), r2.p, r2.bytes, c.stream);
    ECG_HIP(hipStreamSynchronize(c.stream));
    c.to_blocks(h1.data(), h2.data(), R1a, R1b, R2a, R2b, R2ab);
  });
}

int ecg_ucc_solve(ecg_uctx* ctx, int method, const ecg_solver_opts* opts, double* T1a, double* T1b, double* T2a,
                  double* T2b, double* T2ab, double* energies, int* iterations, int* converged) {
  return uguarded(ctx, [&](ecg::UContext& c) {
    check_umethod(method);
    double e[4];
    int conv = 0;
    const int it = c.solve(method, to_opts(opts), e, &conv);
    if (T1a || T1b || T2a || T2b || T2ab) {
      std::vector<double> h1(c.n1()), h2(c.n2());
      d2h(h1.data(), c.T1.p, sizeof(double) * c.n1(), c.stream);
      d2h(h2.data(), c.T2.p, sizeof(double) * c.n2(), c.stream);
      ECG_HIP(hipStreamSynchronize(c.stream));
      c.to_blocks(h1.data(), h2.data(), T1a, T1b, T2a, T2b, T2ab);
    }
    if (energies) std::memcpy(energies, e, sizeof(e));
    if (iterations) *iterations = it;
    if (converged) *converged = conv;
  });
}

int ecg_upert_triples(ecg_uctx* ctx, const double* T1a, const double* T1b, const double* T2a, const double* T2b,
                      const double* T2ab, double* ET, double* ETb) {
  return uguarded(ctx, [&](ecg::UContext& c) {
    double e[2];
    if (T1a || T1b || T2a || T2b || T2ab) {
      ECG_CHECK(T1a && T1b && T2a && T2b && T2ab, "give all five amplitude blocks or none");
      std::vector<double> h1(c.n1()), h2(c.n2());
      c.from_blocks(T1a, T1b, T2a, T2b, T2ab, h1.data(), h2.data());
      ecg::DevBuf d1(sizeof(double) * c.n1()), d2(sizeof(double) * c.n2());
      h2d(d1.p, h1.data(), d1.bytes, c.stream);
      h2d(d2.p, h2.data(), d2.bytes, c.stream);
      c.pert_triples(d1.d(), d2.d(), e);
    } else {
      ECG_CHECK(c.T1.p != nullptr, "no amplitudes: run ecg_ucc_solve first or pass them");
      c.pert_triples(c.T1.d(), c.T2.d(), e);
    }
    if (ET) *ET = e[0];
    if (ETb) *ETb = e[1];
  });
}

}  // extern "C"
