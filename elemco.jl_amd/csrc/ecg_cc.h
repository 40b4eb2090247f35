// Closed-shell coupled-cluster context: integrals, dressing, residuals, solver.
#pragma once
#include "ecg_tensor.h"

namespace ecg {

// Orbital blocks of an n x n (x naux) quantity, each stored densely.
struct Blocks {
  DevBuf oo, ov, vo, vv;
};

struct SolverOpts {
  double thr = 1e-10, conven = 0.1, shifts = 0.15, shiftp = 0.2, resthr = 10.0;
  int maxit = 50, maxdiis = 6;
  int refine_every = 0;  // mixed precision: FP64 residual every k-th iteration (0 = off)
  int graphs = 0;        // replay the residual from a captured HIP graph
};

struct DiisState {
  int maxdiis = 6, next = 0, ndim = 0;
  double resthr = 10.0;
  std::vector<DevBuf> amps, res;
  std::vector<double> bmat;  // (maxdiis+1)^2, row-major
};

struct Context {
  int no = 0, nv = 0, n = 0, nq = 0;
  int device = 0;
  bool hermitian = true;     // right DF factor equals the left one
  double e0 = 0.0;
  hipStream_t stream = nullptr;
  TensorOps ops;
  std::string last_error;

  // undressed inputs
  DevBuf h;                  // n x n one-electron integrals
  Blocks L, R;               // DF factors per block, [p,q,Q]
  DevBuf fock, eps;          // closed-shell Fock matrix (n x n), orbital energies (n)
  DevBuf oovv;               // <kl|cd> = (kc|ld), [k,l,c,d]
  DevBuf vvoo;               // <ij|ab> reordered to [a,b,i,j]
  // dressed (per residual call)
  Blocks Ld, Rd;
  DevBuf dh, df;             // dressed one-electron integrals and Fock matrix (n x n)
  DevBuf trace;              // sum_k Rd_oo[k,k,Q]
  // amplitude state of the device-resident solver
  DevBuf T1, T2, R1, R2;
  DiisState diis;
  double eref_hf = 0.0;
  long long residual_calls = 0;

  long nvo() const { return (long)nv * no; }
  long nvvoo() const { return (long)nv * nv * no * no; }

  void init(int nocc, int nvirt, int naux, const double* h_host, const double* BL_host,
            const double* BR_host, double e0_in, size_t workspace_bytes);
  void build_fock();
  void dress(const double* T1d);                       // T1 on device
  void residual(int method, Prec pr, const double* T1d, const double* T2d, double* R1d, double* R2d);
  // The FP64 residual with every contraction temporary and split-K partial sum taken
  // from the caller's device region [scratch, scratch + bytes): nothing is allocated
  // (the dressed factors are allocated with the context).  Throws if the system needs
  // more than the region holds.
  void residual_noalloc(int method, const double* T1d, const double* T2d, double* R1d, double* R2d, void* scratch,
                        size_t bytes);
  size_t residual_noalloc_min_bytes() const;   // lower bound of the region that call needs
  // adds the (a <= b half of the) DF particle-particle ladder to R2r; T2x: amplitudes as [i,j,c,d]
  void ladder(Prec pr, const double* T2x, double* R2r);
  void ladder_df(Prec pr, const double* Lf, const double* Rf, const double* T2x, double* R2r);
  // energies
  double hf_energy();
  void mp2_guess(double* T1d, double* T2d, double* emp2 /*[E,ESS,EOS]*/);
  void mp2_amplitudes(double* T1d, double* T2d);       // the guess' amplitudes only (kernels, no allocation)
  void doubles_energy(const double* T2d, double* e /*[E,ESS,EOS]*/);
  void singles_energy(const double* T1d, bool fock_only, double* e);
  void hylleraas(const double* T1d, const double* T2d, const double* R1d, const double* R2d,
                 double* e /*[E,ESS,EOS,pE,pESS,pEOS]*/);
  // amplitude update T += -R/(D+shift)
  void update(double* T1d, double* T2d, const double* R1d, const double* R2d, double shifts, double shiftp);
  // closed-shell norms and dots of the reference (cctools.jl)
  double doubles_dot(const double* X, const double* Y, double cex);  // sum (X - cex X_ba) Y
  void doubles_de(const double* X, const double* Y, double& d, double& e);
  DevBuf dots_part;
  std::vector<double> dots_host;
  double singles_dot(const double* X, const double* Y) { return ops.dot_host(nvo(), X, Y); }
  // DIIS (diis.jl) on device vectors; dots: 0 = closed-shell custom dots, 1 = plain
  void diis_reset(int maxdiis, double resthr);
  void diis_perform(double* A1, double* A2, const double* R1d, const double* R2d, int dots);
  // device-resident CC iterations (cc.jl:2518-2648)
  int iter = 0;
  void state_init(const double* T1h, const double* T2h, const SolverOpts& o);
  // info = [E_H, ESS, EOS, E_proj, dE, NormR, NormT, converged]
  // Low precision: the residual is R_low(T) + delta, delta = R_64(T_ref) - R_low(T_ref)
  // from the last refresh (none after state_init); refresh = true evaluates both at
  // the current T, stores the new delta and uses the exact FP64 residual.
  void state_step(int method, Prec pr, const SolverOpts& o, double* info, bool refresh = false);
  DevBuf D1, D2;             // FP64 residual correction (defect) of the mixed-precision solver
  bool have_delta = false;
  int solve(int method, Prec pr, const SolverOpts& o, double* energies, int* converged, FILE* log);
  // <ab|ij> as [a,b,i,j] (equals vvoo for Hermitian integrals)
  DevBuf abij_buf;
  const double* abij() const { return hermitian ? vvoo.d() : abij_buf.d(); }
  // timing of the last residual call (HIP events on the library stream)
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double last_residual_ms = 0.0, last_ladder_ms = 0.0;
  int fp64_iterations = 0;   // FP64 iterations of the last solve (mixed precision)
  DevBuf tws[25];            // (T) work buffers: kept between calls, grown on demand
  int triples_batch[3][2] = {{0, 0}, {0, 0}, {0, 0}};   // measured (T) block size of the packed paths [BF16x3, FP32, FP64][Lambda?]; 0: not yet
  int triples_last_batch = 1;      // block size used by the last (T) call
  // Residual launch sequences captured as HIP graphs (SolverOpts::graphs), one per
  // (method, precision, buffers); a graph is recaptured when a scratch buffer it
  // recorded (GEMM workspaces, contraction arena) has been reallocated since.
  struct ResidualGraph {
    int method = 0;
    Prec pr = Prec::F64;
    const double *t1 = nullptr, *t2 = nullptr;
    double *r1 = nullptr, *r2 = nullptr;
    const void *ws = nullptr, *pk = nullptr, *keep = nullptr, *arena = nullptr;
    size_t wsb = 0, pkb = 0, keepb = 0;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<ResidualGraph> rgraphs;
  bool capturing = false;    // a residual is being recorded (no event records, no host waits)
  void residual_graph(int method, Prec pr, const double* T1d, const double* T2d, double* R1d, double* R2d);
  // Device-resident FP64 iterations (ecg_devsolve.hip): every reduction, the DIIS solve and
  // the convergence test on the device; the iteration is captured as one HIP graph and
  // replayed, with one host read of its info vector per iteration.
  struct DevIter {
    DevBuf scal, part, tt, fvo;          // device scalars, pair-dot partials, T1 T1, f_vo
    int maxdiis = 0;
    bool eager_done = false;             // the first iteration after dev_reset ran eagerly
    hipGraphExec_t exec = nullptr;
    std::vector<const void*> key;        // buffers the graph recorded
    std::vector<double> okey;            // method and options the graph recorded
    double* info_host = nullptr;         // pinned: info[8] of the last iteration
    hipEvent_t done = nullptr;           // recorded after an iteration (batch driver)
    ~DevIter();
  } dv;
  void dev_reset(const SolverOpts& o);
  void dev_step(int method, const SolverOpts& o);
  void dev_launch(int method, const SolverOpts& o);
  int dev_solve(int method, const SolverOpts& o, double* energies, int* converged, FILE* log);
  ~Context();
};

// Half of the DF particle-particle ladder: R2r[a,b,i,j] += sum_Q sum_cd Lf[a,c,Q] Rf[b,d,Q]
// X[c,d,i,j] for b >= the start of a's block of a-rows (diagonal blocks weighted 1/2), X given
// as Xx[i,j,c,d]; when the full term P satisfies P[a,b,i,j] = P[b,a,j,i] it is completed by
// R2 += R2r + R2r[b,a,j,i].  The integrals V[a',b',c,d] of a block of a-rows are assembled on
// MFMA in the free part of the contraction arena; no v^4 array is formed.
void df_ladder_half(TensorOps& ops, long nv, long no, int nq, Prec pr, const double* Lf, const double* Rf,
                    const double* T2x, double* R2r);

// Symmetric eigen-decomposition of a small m x m row-major matrix A (destroyed) by cyclic
// Jacobi sweeps: eigenvectors in the columns of V, eigenvalues w (the DIIS solves).
void jacobi_eig(int m, std::vector<double>& A, std::vector<double>& V, std::vector<double>& w);

// Per-(i,j) block sums of two [a,b,i,j] tensors (v = nv, npairs = o^2 blocks):
// part[2 p] = sum_ab X[a,b,p] Y[a,b,p],  part[2 p + 1] = sum_ab X[b,a,p] Y[a,b,p].
void launch_pair_dots(hipStream_t s, int nv, long npairs, const double* X, const double* Y, double* part);

}  // namespace ecg
