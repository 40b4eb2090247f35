"""DF preparation: sqrtinvchol of the metric (host side of ecg_df_ao_to_mo, reference
src/tools/tensortools.jl:377-391), checked on the CPU through the test-hook library
against the oracle (oracle/dfprep.py) and the closed form M M^T = pinv((P|Q)).  The
device half (the AO -> MO contractions) is checked in tests/test_gpu_dfprep.py."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOOKS = os.path.join(ROOT, "elemco.jl_amd", "libecg_testhooks.so")


def _metric(nfit, rank, seed):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((nfit, rank))
    return G @ G.T + (1e-3 * np.eye(nfit) if rank >= nfit else 0.0)


def _sqrtinvchol(A, tol):
    h = ctypes.CDLL(HOOKS)
    dp = ctypes.POINTER(ctypes.c_double)
    h.ecgt_sqrtinvchol.argtypes = [dp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_int), dp]
    n = A.shape[0]
    A = np.asfortranarray(A)
    M = np.zeros((n, n), order="F")
    r = ctypes.c_int()
    assert h.ecgt_sqrtinvchol(A.ctypes.data_as(dp), n, tol, ctypes.byref(r), M.ctypes.data_as(dp)) == 0
    return np.asfortranarray(M.reshape(-1, order="F")[: n * r.value].reshape(n, r.value, order="F"))


@pytest.mark.parametrize("nfit,rank,tol", [(30, 60, 1e-10), (40, 17, 1e-8), (25, 25, 1e-10)])
def test_sqrtinvchol_matches_oracle_and_pseudo_inverse(nfit, rank, tol):
    from oracle.dfprep import sqrtinvchol
    A = _metric(nfit, rank, nfit)
    M = _sqrtinvchol(A, tol)
    Mo = sqrtinvchol(A, tol)
    assert M.shape == Mo.shape == (nfit, min(nfit, rank))
    scale = np.max(np.abs(Mo))
    assert np.max(np.abs(M - Mo)) < 1e-9 * scale      # same pivots, same triangular solve
    P = np.linalg.pinv(A, rcond=1e-12) if rank < nfit else np.linalg.inv(A)
    assert np.max(np.abs(M @ M.T - P)) < 1e-9 * np.max(np.abs(P))
