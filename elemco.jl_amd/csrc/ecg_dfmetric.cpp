// sqrtinvchol of the DF metric and frozen orbitals on DF factors (host C++; see ecg_dfprep.h).
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "ecg_dfprep.h"

namespace ecg {

std::vector<double> sqrtinvchol(const double* A, int n, double tol, int& rank) {
  if (n <= 0) throw std::runtime_error("sqrtinvchol: empty matrix");
  // right-looking pivoted Cholesky on a full copy; R holds the Schur complement, U the rows
  std::vector<double> R(A, A + (size_t)n * n);
  auto r_ = [&](long i, long j) -> double& { return R[i + (size_t)n * j]; };
  std::vector<double> U((size_t)n * n, 0.0);   // U[k + n j], row k, column j (pivot order)
  auto u_ = [&](long k, long j) -> double& { return U[k + (size_t)n * j]; };
  std::vector<int> p(n);
  for (int i = 0; i < n; ++i) p[i] = i;
  int r = 0;
  for (int k = 0; k < n; ++k) {
    int j = k;
    for (int i = k + 1; i < n; ++i)
      if (r_(i, i) > r_(j, j)) j = i;
    if (!(r_(j, j) > tol)) break;
    if (j != k) {
      for (int c = 0; c < n; ++c) std::swap(r_(k, c), r_(j, c));
      for (int c = 0; c < n; ++c) std::swap(r_(c, k), r_(c, j));
      std::swap(p[k], p[j]);
      for (int q = 0; q < k; ++q) std::swap(u_(q, k), u_(q, j));
    }
    const double piv = std::sqrt(r_(k, k));
    u_(k, k) = piv;
    for (int c = k + 1; c < n; ++c) u_(k, c) = r_(k, c) / piv;
    for (int c = k + 1; c < n; ++c) {
      const double uc = u_(k, c);
      if (uc == 0.0) continue;
      for (int i = k + 1; i < n; ++i) r_(i, c) -= u_(k, i) * uc;
    }
    r = k + 1;
  }
  if (r == 0) throw std::runtime_error("sqrtinvchol: no pivot above the tolerance");
  rank = r;
  // X (n x r, pivot order): U[0:r,:] X = I_r
  std::vector<double> X((size_t)n * r, 0.0);
  auto x_ = [&](long i, long j) -> double& { return X[i + (size_t)n * j]; };
  if (r == n) {
    // upper-triangular inverse, column by column (back substitution)
    for (int c = 0; c < n; ++c) {
      x_(c, c) = 1.0 / u_(c, c);
      for (int i = c - 1; i >= 0; --i) {
        double s = 0.0;
        for (int q = i + 1; q <= c; ++q) s += u_(i, q) * x_(q, c);
        x_(i, c) = -s / u_(i, i);
      }
    }
  } else {
    // minimum norm: X = U^T G^{-1}, G = U U^T (r x r, SPD) = L L^T
    std::vector<double> G((size_t)r * r, 0.0);
    for (int a = 0; a < r; ++a)
      for (int b = 0; b <= a; ++b) {
        double s = 0.0;
        for (int c = std::max(a, b); c < n; ++c) s += u_(a, c) * u_(b, c);
        G[a + (size_t)r * b] = G[b + (size_t)r * a] = s;
      }
    for (int c = 0; c < r; ++c) {   // Cholesky in place, lower triangle
      double d = G[c + (size_t)r * c];
      for (int q = 0; q < c; ++q) d -= G[c + (size_t)r * q] * G[c + (size_t)r * q];
      if (!(d > 0.0)) throw std::runtime_error("sqrtinvchol: U U^T not positive definite");
      d = std::sqrt(d);
      G[c + (size_t)r * c] = d;
      for (int i = c + 1; i < r; ++i) {
        double s = G[i + (size_t)r * c];
        for (int q = 0; q < c; ++q) s -= G[i + (size_t)r * q] * G[c + (size_t)r * q];
        G[i + (size_t)r * c] = s / d;
      }
    }
    // row i of X: solve G y = U[:, i] (a length-r vector), x_i = y^T
    std::vector<double> y(r);
    for (int i = 0; i < n; ++i) {
      for (int a = 0; a < r; ++a) y[a] = u_(a, i);
      for (int a = 0; a < r; ++a) {   // L z = u
        double s = y[a];
        for (int q = 0; q < a; ++q) s -= G[a + (size_t)r * q] * y[q];
        y[a] = s / G[a + (size_t)r * a];
      }
      for (int a = r - 1; a >= 0; --a) {   // L^T y = z
        double s = y[a];
        for (int q = a + 1; q < r; ++q) s -= G[q + (size_t)r * a] * y[q];
        y[a] = s / G[a + (size_t)r * a];
      }
      for (int a = 0; a < r; ++a) x_(i, a) = y[a];
    }
  }
  // undo the pivot order: M[p[i], :] = X[i, :]
  std::vector<double> M((size_t)n * r);
  for (int a = 0; a < r; ++a)
    for (int i = 0; i < n; ++i) M[p[i] + (size_t)n * a] = x_(i, a);
  return M;
}

DfFreeze df_freeze(int n, int nocc, int naux, const double* h, const double* bl, const double* br, double e0,
                   const std::vector<int>& orbs) {
  if (n <= 0 || naux <= 0 || nocc < 0 || nocc > n) throw std::runtime_error("df_freeze: invalid sizes");
  if (!h || !bl) throw std::runtime_error("df_freeze: NULL integrals");
  const double* R = br ? br : bl;
  std::vector<char> fr(n, 0);
  for (int o : orbs)
    if (o >= 1 && o <= n) fr[o - 1] = 1;
  DfFreeze out;
  std::vector<int> core, kept;
  for (int p = 0; p < n; ++p) {
    if (!fr[p]) kept.push_back(p);
    else if (p < nocc) core.push_back(p);
    else ++out.nvirt;
  }
  out.ncore = (int)core.size();
  const size_t nn = (size_t)n * n;
  auto L = [&](int p, int q, int Q) { return bl[p + (size_t)n * q + nn * Q]; };
  auto Rr = [&](int p, int q, int Q) { return R[p + (size_t)n * q + nn * Q]; };
  // d[Q] = sum_c R[c,c,Q]: the Coulomb part is sum_Q L[p,q,Q] d[Q]
  std::vector<double> d(naux, 0.0);
  for (int Q = 0; Q < naux; ++Q)
    for (int c : core) d[Q] += Rr(c, c, Q);
  std::vector<double> cf(nn, 0.0);
  for (int Q = 0; Q < naux; ++Q) {
    const double dq = 2.0 * d[Q];
    const double* lq = bl + nn * Q;
    for (size_t i = 0; i < nn; ++i) cf[i] += lq[i] * dq;
  }
  // exchange: cf[p,q] -= sum_{c,Q} L[p,c,Q] R[c,q,Q]
  for (int Q = 0; Q < naux; ++Q)
    for (int c : core)
      for (int q = 0; q < n; ++q) {
        const double r = Rr(c, q, Q);
        if (r == 0.0) continue;
        const double* lc = bl + (size_t)n * c + nn * Q;   // L[:, c, Q]
        double* col = cf.data() + (size_t)n * q;
        for (int p = 0; p < n; ++p) col[p] -= lc[p] * r;
      }
  out.e0 = e0;
  for (int c : core) out.e0 += 2.0 * h[c + (size_t)n * c] + cf[c + (size_t)n * c];
  const int m = (int)kept.size();
  out.nkept = m;
  out.h.resize((size_t)m * m);
  for (int b = 0; b < m; ++b)
    for (int a = 0; a < m; ++a)
      out.h[a + (size_t)m * b] = h[kept[a] + (size_t)n * kept[b]] + cf[kept[a] + (size_t)n * kept[b]];
  auto slice = [&](auto&& src, std::vector<double>& dst) {
    dst.resize((size_t)m * m * naux);
    for (int Q = 0; Q < naux; ++Q)
      for (int b = 0; b < m; ++b)
        for (int a = 0; a < m; ++a) dst[a + (size_t)m * b + (size_t)m * m * Q] = src(kept[a], kept[b], Q);
  };
  slice(L, out.BL);
  if (br) slice(Rr, out.BR);
  return out;
}

}  // namespace ecg
