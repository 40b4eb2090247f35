"""DF preparation on the GPU path (ecg_df_ao_to_mo, ecg_create_from_ao; reference
src/integrals/dftools.jl:92-125 with sqrtinvchol, src/tools/tensortools.jl:377-391).

The factors are compared with the oracle (same pivoted Cholesky, numpy contractions) to
1e-10 relative.  End to end: AO integrals built so that their four-index integrals equal
those of a synthetic MO density-fitted system (random orthogonal C, random metric
(P|Q) = L L^T, (mu nu|P) = sum_Q B_ao[mu,nu,Q] L[P,Q]) give the same CCSD energy as that
system (1e-10 Eh).
"""
import numpy as np
import pytest

import elemco_amd as E
from elemco_amd import lib
from gpu_helpers import requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]


@pytest.mark.parametrize("nao,nmo,nfit,rank", [(14, 11, 40, 60), (20, 20, 37, 21)])
def test_df_ao_to_mo_matches_oracle(nao, nmo, nfit, rank):
    from oracle.dfprep import df_ao_to_mo
    rng = np.random.default_rng(nao + rank)
    G = rng.standard_normal((nfit, rank))
    PQ = G @ G.T
    AAP = rng.standard_normal((nao, nao, nfit))
    AAP = AAP + AAP.transpose(1, 0, 2)
    C = rng.standard_normal((nao, nmo))
    B = lib.df_ao_to_mo(C, AAP, PQ, tol=1e-9)
    Bo = df_ao_to_mo(C, AAP, PQ, 1e-9)
    assert B.shape == Bo.shape == (nmo, nmo, min(nfit, rank))
    assert np.max(np.abs(B - Bo)) < 1e-10 * np.max(np.abs(Bo))


def test_create_from_ao_reproduces_the_mo_system():
    from oracle.synthetic import make_df_system
    from elemco_amd.synthetic import conditioned_scale
    o, v, q, seed = 4, 20, 60, 11
    scale = conditioned_scale(o, v, q)
    s = make_df_system(o, v, q, seed, scale=scale)
    n = o + v
    rng = np.random.default_rng(3)
    C, _ = np.linalg.qr(rng.standard_normal((n, n)))
    L = np.linalg.cholesky(np.eye(q) + 0.1 * (lambda X: X @ X.T)(rng.standard_normal((q, q))))
    B_ao = np.einsum("mp,pqQ,nq->mnQ", C, s.B, C)
    AAP = np.einsum("mnQ,PQ->mnP", B_ao, L)
    PQ = L @ L.T
    h_ao = C @ s.h @ C.T
    ctx = lib.Context.from_ao(o, C, h_ao, AAP, PQ, tol=1e-12, e0=s.e0)
    assert (ctx.nocc, ctx.nvirt, ctx.naux) == (o, v, q)
    EC = E.ECInfo.synthetic(o, v, q, seed, scale=scale)
    opts = EC.opts()
    e1 = ctx.cc_solve(E.ECG_CCSD, E.ECG_FP64, opts)[0]
    e2 = EC.ctx.cc_solve(E.ECG_CCSD, E.ECG_FP64, opts)[0]
    assert abs(float(e1[0]) - float(e2[0])) < 1e-10
    ctx.close()
