"""Mixed-precision CCSD at the full o=40, v=360, naux=1200 size (seed 2).

The literal N(0, 0.05^2) factors of this size give a divergent CCSD (E_MP2 of
order -1e5 Eh), so converged energies are compared on the conditioned system of
the same sizes (factor width scaled by conditioned_scale; the shapes and every
product are those of the benchmark workload).

Bars:
* the FP32 and BF16 hi/lo defect-corrected solves end within 1e-6 Eh of the
  FP64 solve (the target of the mixed path);
* every solve ends on an exact FP64 iteration whose squared residual norm is
  below thr, and the mixed solves used fewer FP64 residuals than iterations;
* the residual at the converged amplitudes, evaluated afresh in FP64 through
  ecg_cc_residual, has a squared norm below thr (so the fixed point is the FP64
  one, independently of the solver's bookkeeping), and the FP32 / BF16 hi/lo
  residuals there differ from it by less than 1e-4 (FP32) and 1e-3 (BF16 hi/lo)
  of max|R2| at the MP2 guess, and by more than zero (the low-precision
  products ran).
"""
import numpy as np
import pytest

from gpu_helpers import requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]

import elemco_amd as E  # noqa: E402
from elemco_amd.synthetic import conditioned_scale  # noqa: E402

CFG = (40, 360, 1200, 2)
THR = 1e-10


@pytest.fixture(scope="module")
def c2():
    o, v, q, seed = CFG
    print(f"\nc2 system o={o} v={v} naux={q} seed={seed} (conditioned)", flush=True)
    EC = E.ECInfo.synthetic(o, v, q, seed, scale=conditioned_scale(o, v, q))
    out = {"EC": EC}
    yield out
    EC.ctx.close()


def _solve(EC, prec):
    opts = E.default_opts(maxit=60, thr=THR)
    e, T1, T2, it, conv = EC.ctx.cc_solve(E.ECG_CCSD, prec, opts)
    n64 = EC.ctx.stats()["fp64_iterations"]
    print(f"  precision {prec}: E={float(e[0]):.12f} it={it} fp64 it={n64} NormR={float(e[5]):.2e}", flush=True)
    return float(e[0]), float(e[5]), T1, T2, it, conv, n64


def test_c2_fp64_solve(c2):
    e, normr, T1, T2, it, conv, _ = _solve(c2["EC"], E.ECG_FP64)
    assert conv and normr < THR
    assert np.isfinite(e) and -10.0 < e < 0.0
    c2["fp64"] = (e, T1, T2)


@pytest.mark.parametrize("prec", [E.ECG_BF16X3, E.ECG_FP32])
def test_c2_mixed_solve_matches_fp64(c2, prec):
    if "fp64" not in c2:
        pytest.skip("needs the FP64 solve")
    e64 = c2["fp64"][0]
    e, normr, T1, T2, it, conv, n64 = _solve(c2["EC"], prec)
    assert conv and normr < THR
    assert 1 <= n64 < it
    assert abs(e - e64) < 1e-6, (e, e64)
    c2[prec] = (e, T1, T2)


def test_c2_fixed_point_is_the_fp64_one(c2):
    """Fresh FP64 residual at the BF16x3 solution: squared norm (the reference's
    calc_doubles_norm + singles, cctools.jl:453-480) below thr; low-precision residuals
    there within their per-product error of it."""
    if E.ECG_BF16X3 not in c2:
        pytest.skip("needs the BF16x3 solve")
    ctx = c2["EC"].ctx
    _, T1, T2 = c2[E.ECG_BF16X3]
    R1, R2 = ctx.cc_residual(T1, T2, E.ECG_CCSD, E.ECG_FP64)
    norm = 2.0 * float(np.sum(R1 * R1)) + float(np.sum((2.0 * R2 - R2.transpose(1, 0, 2, 3)) * R2))
    assert norm < THR, norm
    T1m, T2m, _ = ctx.mp2()
    _, R2m = ctx.cc_residual(T1m, T2m, E.ECG_CCSD, E.ECG_FP64)
    scale = float(np.max(np.abs(R2m)))
    for prec, tol in ((E.ECG_FP32, 1e-4), (E.ECG_BF16X3, 1e-3)):
        S1, S2 = ctx.cc_residual(T1, T2, E.ECG_CCSD, prec)
        d = max(float(np.max(np.abs(S1 - R1))), float(np.max(np.abs(S2 - R2))))
        print(f"  precision {prec}: max|R_low - R_64| = {d:.3e} (scale {scale:.3e})", flush=True)
        assert d < tol * scale and d > 0.0
