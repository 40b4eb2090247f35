"""DF preparation (CPU restatement; TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

MO density-fitting factors from AO three-index integrals, as
``generate_AO_DF_integrals`` / ``generate_3idx_integrals``
(src/integrals/dftools.jl:28-125) do after the integral library has produced
(mu nu|P) and the metric (P|Q):

    M = sqrtinvchol((P|Q), tol)                     (src/tools/tensortools.jl:377-391)
    AAL[mu,nu,L] = sum_P (mu nu|P) M[P,L]           (dftools.jl:64)
    mmL[p,q,L]   = sum_{mu,nu} C[mu,p] AAL[mu,nu,L] C[nu,q]   (dftools.jl:111-112)

``sqrtinvchol``: pivoted Cholesky A[p,p] = U^T U (largest remaining diagonal first,
stopped once it is <= tol, LAPACK dpstrf's absolute criterion), then the
minimum-norm solution X of U[1:r,:] X = I_r (X = U^{-1} at full rank) with the pivot
order undone, so M M^T is the (pseudo)inverse of A.

Pin (tests/test_oracle.py): sum_L mmL mmL equals the closed form
C^T C^T (mu nu|P) pinv((P|Q)) (P|lambda sigma) C C, and M M^T = pinv((P|Q)), for full-rank
and rank-deficient metrics.
"""
from __future__ import annotations

import numpy as np


def pivoted_cholesky(A: np.ndarray, tol: float):
    """A[p, p] = U^T U over the kept rank r (U r x n, upper trapezoidal in pivot order)."""
    A = np.array(A, dtype=np.float64)
    n = A.shape[0]
    p = np.arange(n)
    U = np.zeros((n, n))
    R = A.copy()
    r = 0
    for k in range(n):
        d = np.diag(R)[k:]
        j = k + int(np.argmax(d))
        if d[j - k] <= tol:
            break
        # swap k and j in the remaining matrix, the permutation and the computed rows
        R[[k, j], :] = R[[j, k], :]
        R[:, [k, j]] = R[:, [j, k]]
        p[[k, j]] = p[[j, k]]
        U[:k, [k, j]] = U[:k, [j, k]]
        piv = np.sqrt(R[k, k])
        U[k, k] = piv
        U[k, k + 1:] = R[k, k + 1:] / piv
        R[k + 1:, k + 1:] -= np.outer(U[k, k + 1:], U[k, k + 1:])
        r = k + 1
    return U[:r, :], p, r


def sqrtinvchol(A: np.ndarray, tol: float = 1e-8):
    """``sqrtinvchol`` (tensortools.jl:377-391): M (n x r) with M M^T = A^{-1} (pseudo)."""
    U, p, r = pivoted_cholesky(A, tol)
    if r == A.shape[0]:
        X = np.linalg.solve(U, np.eye(r))
    else:
        X = np.linalg.lstsq(U, np.eye(r), rcond=None)[0]   # minimum-norm solution
    M = np.empty_like(X)
    M[p, :] = X
    return M


def df_ao_to_mo(C: np.ndarray, AAP: np.ndarray, PQ: np.ndarray, tol: float = 1e-8):
    """mmL[p,q,L] of ``generate_3idx_integrals`` (dftools.jl:92-125) from AO integrals."""
    M = sqrtinvchol(PQ, tol)
    AAL = np.einsum("mnP,PL->mnL", AAP, M)
    return np.einsum("mp,mnL,nq->pqL", C, AAL, C)
