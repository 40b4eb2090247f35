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
