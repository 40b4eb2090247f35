"""The DF particle ladder in several a-row blocks, against the oracle's residual.

The ladder assembles V[a',b',c,d] for a block of a-rows at a time, as many rows
as the contraction arena leaves room for (80 % of the free arena, v^3 doubles
per row), and contracts every block with the same amplitudes T[i,j,c,d].  For
products large enough to run on packed planes (BF16 hi/lo or FP32:
M N / (M + N) >= 256; FP64: >= 1024) the amplitudes' planes are packed once per
residual and kept for every block (GemmOp::b_keep).  The context here is made
with a workspace small enough that the ladder runs in several blocks, each with
a diagonal and an off-diagonal part, so the kept planes are reused.

With o^2 >= 128 and v >= 64 (the o=12 and o=40 cases here) the V blocks are written
straight into the contraction's packed planes (the factors transposed once per residual
to [c, a, Q]), each block's diagonal part as its a' <= b' triangle contracted into a
temporary and added with weights 1 and 1/2; v = 66 has v^2 not a multiple of 32 (zeroed K
padding in the planes); o=20 v=50 runs the V blocks in FP64 and packs them.

The workspace is the residual's other arena tensors at the time of the ladder
(five o^2 v^2 tensors, W o^4, d_oovo o^3 v, Z v o Q and the Fock-like blocks)
plus room for `rows` ladder rows.

Bars (max-abs against the oracle's calc_cc_resid, cc.jl:1343-1489): FP64 1e-10;
FP32 and BF16 hi/lo products 1e-5 and 1e-4 of max|R2| (their per-product rounding,
about 2^-24 sqrt(K) and 2^-17 relative, times the cancellation inside the residual).
"""
import numpy as np
import pytest

from gpu_helpers import requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]

import elemco_amd as E  # noqa: E402
from elemco_amd.synthetic import conditioned_scale, make_system  # noqa: E402
from oracle import ccsd as O  # noqa: E402
from oracle.synthetic import make_df_system  # noqa: E402


def _workspace(o, v, q, rows):
    fixed = 5 * o * o * v * v + o**4 + o**3 * v + v * o * q + o * v + v * v + o * o
    return int(8 * (fixed + rows * v**3 / 0.8)) + (64 << 10)


@pytest.mark.parametrize("cfg,rows,prec,tol", [
    ((20, 50, 80, 31), 16, E.ECG_BF16X3, 1e-4),
    ((20, 50, 80, 31), 16, E.ECG_FP32, 1e-5),
    ((12, 64, 100, 33), 16, E.ECG_BF16X3, 1e-4),
    ((12, 64, 100, 33), 16, E.ECG_FP32, 1e-5),
    ((12, 66, 100, 34), 16, E.ECG_BF16X3, 1e-4),
    ((12, 66, 100, 34), 16, E.ECG_FP32, 1e-5),
    ((12, 66, 100, 34), 20, E.ECG_FP64, 0.0),
    ((40, 120, 100, 32), 68, E.ECG_FP64, 0.0),
])
def test_multi_block_ladder_residual_matches_oracle(cfg, rows, prec, tol):
    o, v, q, seed = cfg
    sc = 0.7 * conditioned_scale(o, v, q)
    _, h, B, e0 = make_system(o, v, q, seed, sc)
    ctx = E.Context(o, v, q, h, B, None, e0, workspace_bytes=_workspace(o, v, q, rows))
    ECo = O.ECInfo.from_df(make_df_system(o, v, q, seed, sc))
    O.calc_fock_matrix(ECo)
    O.calc_MP2(ECo)
    T1, T2 = ECo.files["T_vo"], ECo.files["T_vvoo"]
    try:
        for dc in (False, True):
            R1o, R2o = O.calc_cc_resid(ECo, T1, T2, dc=dc)
            meth = E.ECG_DCSD if dc else E.ECG_CCSD
            R1, R2 = ctx.cc_residual(T1, T2, meth, prec)
            R1b, R2b = ctx.cc_residual(T1, T2, meth, prec)     # second call: planes packed again
            assert np.array_equal(R2, R2b) and np.array_equal(R1, R1b)
            err = max(float(np.max(np.abs(R1 - R1o))), float(np.max(np.abs(R2 - R2o))))
            scale = float(np.max(np.abs(R2o)))
            print(f"  o={o} v={v} prec={prec} dc={dc}: max|R - R_oracle| = {err:.3e} (max|R2| {scale:.3e})",
                  flush=True)
            if prec == E.ECG_FP64:
                assert err < 1e-10
            else:
                assert err < tol * scale
    finally:
        ctx.close()


@pytest.mark.parametrize("prec,tol", [(E.ECG_FP64, 0.0), (E.ECG_BF16X3, 1e-4)])
def test_two_sided_factors_fused_ladder_matches_oracle(prec, tol):
    """Non-Hermitian integrals (pq|rs) = sum_Q BL[p,q,Q] BR[r,s,Q] with BR != BL, at o=12 v=66
    naux=100: the fused ladder then transposes both factors (the Hermitian residual passes
    one factor twice).  The integrals keep the particle symmetry (pq|rs) = (rs|pq) of the
    reference's similarity-transformed dumps (which the half ladder and the final
    symmetrisation use) but not (pq|rs) = (qp|rs): B[p,q,P] is perturbed asymmetrically,
    BR = B and BL = B D with D symmetric.  The oracle takes the same integrals as the
    four-index array <pq|rs> = sum_Q BL[p,r,Q] BR[q,s,Q]; MP2 amplitudes of the oracle.
    Bars as above."""
    o, v, q, seed = 12, 66, 100, 35
    sc = 0.7 * conditioned_scale(o, v, q)
    _, h, B0, e0 = make_system(o, v, q, seed, sc)
    rng = np.random.default_rng(seed)
    Bns = B0 + 0.1 * float(np.std(B0)) * rng.standard_normal(B0.shape)
    X = 0.05 * rng.standard_normal((q, q))
    D = np.eye(q) + 0.5 * (X + X.T)
    BR = np.asfortranarray(Bns)
    BL = np.asfortranarray(np.einsum("pqP,PQ->pqQ", Bns, D))
    ctx = E.Context(o, v, q, h, BL, BR, e0, workspace_bytes=_workspace(o, v, q, 16))
    int2 = np.einsum("prQ,qsQ->pqrs", BL, BR, optimize=True)
    assert np.allclose(int2, int2.transpose(1, 0, 3, 2))     # particle symmetry
    ECo = O.ECInfo(O.FourIndexInts(int2), h, e0, o)
    O.calc_fock_matrix(ECo)
    O.calc_MP2(ECo)
    T1, T2 = ECo.files["T_vo"], ECo.files["T_vvoo"]
    try:
        for dc in (False, True):
            R1o, R2o = O.calc_cc_resid(ECo, T1, T2, dc=dc)
            R1, R2 = ctx.cc_residual(T1, T2, E.ECG_DCSD if dc else E.ECG_CCSD, prec)
            err = max(float(np.max(np.abs(R1 - R1o))), float(np.max(np.abs(R2 - R2o))))
            scale = float(np.max(np.abs(R2o)))
            print(f"  two-sided o={o} v={v} prec={prec} dc={dc}: max|R - R_oracle| = {err:.3e} "
                  f"(max|R2| {scale:.3e})", flush=True)
            if prec == E.ECG_FP64:
                assert err < 1e-10
            else:
                assert err < tol * scale
    finally:
        ctx.close()
