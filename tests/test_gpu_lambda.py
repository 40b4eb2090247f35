"""GPU parity of the Lambda residual, 1-RDM, Lambda iterations and Lambda-(T) (FP64).

Bars: residual and densities within 1e-10 of the oracle; the reference's
Lambda-CCSD(T) energy (test/h2o.jl:9) within 1e-6 Eh and the oracle within 1e-9.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from gpu_helpers import requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]

import elemco_amd as E  # noqa: E402
from oracle import ccsd as O  # noqa: E402
from oracle import lagrange as OL  # noqa: E402
from oracle.synthetic import make_df_system  # noqa: E402


def _converged_oracle(cfg, method="ccsd"):
    s = make_df_system(*cfg)
    ECo = O.ECInfo.from_df(s)
    ECo.options["thr"] = 1e-14
    ECo.options["maxit"] = 200
    O.ccdriver(ECo, method)
    return s, ECo


@pytest.mark.parametrize("dc", [False, True])
def test_lambda_residual(dc):
    cfg = (4, 15, 40, 3)
    s, ECo = _converged_oracle(cfg, "dcsd" if dc else "ccsd")
    T1, T2 = ECo.files["T_vo"], ECo.files["T_vvoo"]
    rng = np.random.default_rng(21)
    U1 = 0.05 * rng.standard_normal(T1.shape)
    U2 = 0.05 * rng.standard_normal(T2.shape)
    U2 = U2 + U2.transpose(1, 0, 3, 2)
    O.calc_dressed_ints(ECo, T1, calc_d_vovv=True)
    OL.calc_3ext_times_T2(ECo, T2)
    R1o, R2o = OL.calc_ccsd_vector_times_Jacobian(ECo, U1, U2, dc=dc)
    EC = E.ECInfo.synthetic(*cfg)
    E.calc_fock_matrix(EC)
    R1, R2 = EC.ctx.lambda_residual(T1, T2, U1, U2, E.ECG_DCSD if dc else E.ECG_CCSD)
    assert np.max(np.abs(R1 - R1o)) < 1e-10
    assert np.max(np.abs(R2 - R2o)) < 1e-10


def test_one_rdm():
    cfg = (4, 15, 40, 3)
    s, ECo = _converged_oracle(cfg)
    T1, T2 = ECo.files["T_vo"], ECo.files["T_vvoo"]
    rng = np.random.default_rng(8)
    U1 = 0.05 * rng.standard_normal(T1.shape)
    U2 = 0.05 * rng.standard_normal(T2.shape)
    D1o, dD1o = OL.calc_1RDM(ECo, U1, U2, T1, T2)
    EC = E.ECInfo.synthetic(*cfg)
    D1, dD1 = EC.ctx.one_rdm(T1, T2, U1, U2)
    assert np.max(np.abs(D1 - D1o)) < 1e-12
    assert np.max(np.abs(dD1 - dD1o)) < 1e-12


def test_lambda_solve_and_lambda_triples():
    cfg = (5, 19, 72, 0)
    s, ECo = _converged_oracle(cfg)
    OL.calc_lm_cc(ECo)
    from oracle import triples as OT
    et_o, etb_o = OT.calc_LpertT_closed_shell(ECo)
    EC = E.ECInfo.synthetic(*cfg)
    E.calc_fock_matrix(EC)
    T1, T2 = ECo.files["T_vo"], ECo.files["T_vvoo"]
    U1, U2, it, conv = EC.ctx.lambda_solve(T1, T2, opts=E.default_opts(thr=1e-14, maxit=200))
    assert conv
    assert float(np.max(np.abs(U2 - ECo.files["U_vvoo"]))) < 1e-8
    assert float(np.max(np.abs(U1 - ECo.files["U_vo"]))) < 1e-8
    et, etb = EC.ctx.pert_triples(T1, T2, ECo.files["U_vo"], ECo.files["U_vvoo"])
    assert abs(et - et_o) < 1e-10 and abs(etb - etb_o) < 1e-10


def test_h2o_lambda_ccsd_t_reference_energy():
    """@cc λccsd(t) on test/files/H2O.FCIDUMP (test/h2o.jl:9,24-25)."""
    EC = E.ECInfo.from_fcidump(os.path.join(GOLDEN, "H2O.FCIDUMP.gz"))
    e = E.ccdriver(EC, "λccsd(t)")
    assert abs(e["ΛCCSD(T)c"] - (-0.326915143863)) < 1e-6
    from oracle.fcidump import read_fcidump
    eo = O.ccdriver(O.ECInfo.from_fcidump(read_fcidump(os.path.join(GOLDEN, "H2O.FCIDUMP.gz"))), "λccsd(t)")
    assert abs(e["ΛCCSD(T)c"] - eo["ΛCCSD(T)c"]) < 1e-9


@pytest.mark.parametrize("method", [E.ECG_CCSD, E.ECG_DCSD])
@pytest.mark.parametrize("prec", [E.ECG_FP32, E.ECG_BF16X3])
def test_mixed_lambda_solve_converges_to_fp64(method, prec):
    """Defect-corrected low-precision Lambda iterations end at the FP64 multipliers:
    the FP64 Lambda residual at the mixed solution meets thr = 1e-16 (products
    rounded to FP32 without correction leave ~1e-7 relative residual errors, i.e.
    a squared norm far above it), the solution agrees with the FP64 solve to 1e-8
    (both within ~1e-8 of the fixed point at this threshold), convergence comes
    on an FP64 iteration, and fewer FP64 residuals than iterations are used.
    Sizes as in the mixed CC solve test (BF16 hi/lo products are used only for
    products with M N / (M + N) >= 256)."""
    from elemco_amd.synthetic import conditioned_scale
    no, nv, nq, seed = (6, 40, 90, 21) if prec == E.ECG_FP32 else (10, 90, 300, 5)
    EC = E.ECInfo.synthetic(no, nv, nq, seed, scale=0.7 * conditioned_scale(no, nv, nq))
    opts = E.default_opts(maxit=300, thr=1e-16)
    _, T1, T2, _, conv = EC.ctx.cc_solve(method, E.ECG_FP64, E.default_opts(maxit=300, thr=1e-20))
    assert conv
    U1a, U2a, it64, c64 = EC.ctx.lambda_solve(T1, T2, method, opts)
    U1b, U2b, itlo, clo = EC.ctx.lambda_solve(T1, T2, method, opts, precision=prec)
    n64 = EC.ctx.stats()["fp64_iterations"]
    assert c64 and clo
    assert 1 <= n64 < itlo
    # the contravariant residual norm of lm_cc_iterations!: (2 R + R_ba) . R / 3 + R1 . R1 / 2
    R1, R2 = EC.ctx.lambda_residual(T1, T2, U1b, U2b, method)
    norm = float(np.sum((2.0 * R2 + R2.transpose(1, 0, 2, 3)) * R2) / 3.0 + 0.5 * np.sum(R1 * R1))
    assert norm < 1e-16, norm
    assert float(np.max(np.abs(U1b - U1a))) < 1e-8 and float(np.max(np.abs(U2b - U2a))) < 1e-8
