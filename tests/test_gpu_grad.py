"""Gradient of the CC correlation energy (ecg_energy_grad) on the GPU.

Bars:
* dE/dh equals the transposed dressed one-particle density dD1 of
  ecg_one_rdm (calc_1RDM, src/cc/cc_lagrange.jl:793-837; an independent code
  path, itself pinned to finite differences in tests/test_oracle.py) to 1e-10;
* dE/dB and dE/dh along seeded random directions agree with Richardson-
  extrapolated central finite differences of fully re-solved energies
  (thr 1e-20) to 1e-7 relative -- the invariant that defines the gradient.
"""
import numpy as np
import pytest

from gpu_helpers import requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]

import elemco_amd as E  # noqa: E402
from elemco_amd.synthetic import conditioned_scale, make_system  # noqa: E402

CFG = (4, 16, 40, 5)
OPTS = dict(thr=1e-20, maxit=400)


def solve(no, h, B, method):
    ctx = E.Context(no, h.shape[0] - no, B.shape[2], h, B)
    e, T1, T2, it, conv = ctx.cc_solve(method, E.ECG_FP64, E.default_opts(**OPTS))
    assert conv, it
    return float(e[0]), ctx, T1, T2


def richardson(f, step):
    fd = lambda s: (f(s) - f(-s)) / (2.0 * s)
    return (4.0 * fd(step / 2.0) - fd(step)) / 3.0


@pytest.mark.parametrize("method", [E.ECG_CCSD, E.ECG_DCSD])
def test_energy_gradient_matches_density_and_finite_differences(method):
    o, v, q, seed = CFG
    _, h, B, _ = make_system(o, v, q, seed, 0.7 * conditioned_scale(o, v, q))
    _, ctx, T1, T2 = solve(o, h, B, method)
    U1, U2, it, conv = ctx.lambda_solve(T1, T2, method, E.default_opts(**OPTS))
    assert conv, it
    gh, gb = ctx.energy_grad(T1, T2, U1, U2, method)
    assert gh.shape == h.shape and gb.shape == B.shape
    # dE/dh against the Lambda one-particle density
    _, dD1 = ctx.one_rdm(T1, T2, U1, U2)
    assert float(np.max(np.abs(gh - dD1.T))) < 1e-10 * max(1.0, float(np.max(np.abs(gh))))
    rng = np.random.default_rng(7)
    # dE/dB along a random direction symmetric in (p, q)
    dB = float(np.std(B)) * rng.standard_normal(B.shape)
    dB = 0.5 * (dB + dB.transpose(1, 0, 2))
    fd = richardson(lambda s: solve(o, h, B + s * dB, method)[0], 1e-3)
    an = float(np.sum(gb * dB))
    assert abs(fd - an) < 1e-7 * abs(an), (fd, an)
    # dE/dh along a random symmetric direction
    dh = rng.standard_normal(h.shape)
    dh = 0.5 * (dh + dh.T)
    fd = richardson(lambda s: solve(o, h + s * dh, B, method)[0], 1e-3)
    an = float(np.sum(gh * dh))
    assert abs(fd - an) < 1e-7 * abs(an), (fd, an)


@pytest.mark.parametrize("method", [E.ECG_CCSD, E.ECG_DCSD])
def test_two_sided_gradient_on_st_fixture(method):
    """ecg_energy_grad_lr on the reference's non-Hermitian H2O_ST1 fixture (two-sided exact DF
    factors BL = L|D|^1/2, BR = BL sign(D)): dE/dh equals dD1^T (1e-10), and along directions
    that keep the particle symmetry of the integrals (dBL = X, dBR = X sign(D)) and along a
    random dh, the analytic derivative equals Richardson-extrapolated central differences of
    re-solved energies to 1e-7 relative."""
    import os
    from conftest import GOLDEN
    from elemco_amd.lib import Fcidump
    fd = Fcidump(os.path.join(GOLDEN, "H2O_ST1.FCIDUMP.gz"))
    BL, BR = fd.df_factors()
    h, e0, o = fd.int1.copy(order="F"), fd.int0, fd.nelec // 2
    fd.close()
    assert BR is not None
    sgn = np.sign(np.einsum("pqQ,pqQ->Q", BL, BR))
    assert np.max(np.abs(BR - BL * sgn)) < 1e-14 and (sgn < 0).any()

    def solve2(hh, bl, br):
        ctx = E.Context(o, hh.shape[0] - o, bl.shape[2], hh, bl, br, e0)
        e, T1, T2, it, conv = ctx.cc_solve(method, E.ECG_FP64, E.default_opts(**OPTS))
        assert conv, it
        return float(e[0]), ctx, T1, T2

    _, ctx, T1, T2 = solve2(h, BL, BR)
    U1, U2, it, conv = ctx.lambda_solve(T1, T2, method, E.default_opts(**OPTS))
    assert conv, it
    with pytest.raises(E.ElemcoError, match="ecg_energy_grad_lr"):
        ctx.energy_grad(T1, T2, U1, U2, method)
    gh, gl, gr = ctx.energy_grad_lr(T1, T2, U1, U2, method)
    _, dD1 = ctx.one_rdm(T1, T2, U1, U2)
    ctx.close()
    assert float(np.max(np.abs(gh - dD1.T))) < 1e-10 * max(1.0, float(np.max(np.abs(gh))))
    rng = np.random.default_rng(23)
    X = float(np.std(BL)) * rng.standard_normal(BL.shape)
    fd_b = richardson(lambda s: solve2(h, BL + s * X, BR + s * X * sgn)[0], 1e-3)
    an_b = float(np.sum(gl * X) + np.sum(gr * X * sgn))
    assert abs(fd_b - an_b) < 1e-7 * abs(an_b), (fd_b, an_b)
    dh = rng.standard_normal(h.shape)
    fd_h = richardson(lambda s: solve2(h + s * dh, BL, BR)[0], 1e-3)
    an_h = float(np.sum(gh * dh))
    assert abs(fd_h - an_h) < 1e-7 * abs(an_h), (fd_h, an_h)
