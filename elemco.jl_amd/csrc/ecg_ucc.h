// Open-shell (unrestricted) coupled cluster on the device: UCCSD / UCCD residual, UMP2 guess,
// energies and the iteration loop, on spin orbitals ordered [occupied alpha, occupied beta,
// virtual alpha, virtual beta].
#pragma once
#include "ecg_cc.h"

namespace ecg {

struct UContext {
  int n = 0, na = 0, nb = 0, nq = 0;   // spatial orbitals, alpha / beta electrons, DF functions
  int O = 0, V = 0, N = 0;             // occupied / virtual / all spin orbitals
  int device = 0;
  double e0 = 0.0, ehf = 0.0;
  hipStream_t stream = nullptr;
  TensorOps ops;
  std::string last_error;
  std::vector<int> sp, sg;             // spatial index and spin (0 alpha, 1 beta) of each spin orbital
  DevBuf Boo, Bov, Bvo, Bvv;           // spin-orbital DF factors [p,q,Q] per block
  DevBuf fock, eps;                    // N x N spin-orbital Fock matrix, its diagonal
  DevBuf foo, fov, fvo, fvv;           // its blocks
  // antisymmetrised integrals <pq||rs>: the other blocks the equations use are permutations
  // of these (vvvo, vovv from ovvv; ovoo, oovo from ooov; vovo from ovvo; vvoo from oovv)
  DevBuf oovv, ovvv, ooov, ovvo, oooo;
  DevBuf T1, T2, R1, R2;               // amplitude state of the iterations
  struct Diis {
    int maxdiis = 6, next = 0, ndim = 0;
    double resthr = 10.0;
    std::vector<DevBuf> amps, res;
    std::vector<double> bmat;
  } diis;
  int fp64_iterations = 0;

  long n1() const { return (long)V * O; }
  long n2() const { return (long)V * V * O * O; }
  void init(int nocca, int noccb, int norb, int naux, const double* h, const double* B, double e0_in,
            size_t workspace_bytes);
  // method: 0 UCCSD, 1 UCCD (T1 = 0, R1 set to zero), 2 UDCSD (distinguishable-cluster doubles)
  static bool has_singles(int method) { return method != 1; }
  // residual R(T) of the method at device amplitudes T1d, T2d
  void residual(int method, const double* T1d, const double* T2d, double* R1d, double* R2d);
  void mp2(double* T1d, double* T2d, double* emp2);
  double proj_energy(const double* T1d, const double* T2d);
  double dot(const double* a, const double* b, long n) { return ops.dot_host(n, a, b); }
  int solve(int method, const SolverOpts& o, double* energies, int* converged);
  // (T) of UCCSD on spin orbitals at amplitudes T1d, T2d (device): e = [E(T), E[T]]
  void pert_triples(const double* T1d, const double* T2d, double* e);
  // spin blocks of the reference (T1a, T1b, T2a, T2b, T2ab) <-> spin-orbital arrays (host)
  void to_blocks(const double* T1so, const double* T2so, double* T1a, double* T1b, double* T2a, double* T2b,
                 double* T2ab) const;
  void from_blocks(const double* T1a, const double* T1b, const double* T2a, const double* T2b, const double* T2ab,
                   double* T1so, double* T2so) const;
  ~UContext();
};

}  // namespace ecg
