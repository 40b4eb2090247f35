"""Open-shell UCCSD / UCCD / UDCSD on the GPU (ecg_ucreate*, ecg_ucc_residual, ecg_ucc_solve) against
the CPU oracle (oracle/uccsd.py) and the reference's own open-shell test energies.

Bars: residual spin blocks within 1e-10 (max-abs) of the oracle for random amplitudes;
UMP2 amplitudes 1e-12; HF / MP2 energies 1e-10 of the oracle; converged UCCSD / UCCD
energies within 1e-9 Eh of the oracle solved to the same threshold and 1e-6 Eh of
test/h2o_cation.jl and test/h2o_triplet.jl; the closed-shell limit equals the closed-shell
GPU path's CCSD (DCSD) energy (1e-9 Eh).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from gpu_helpers import requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]

import elemco_amd as E  # noqa: E402
from oracle import uccsd as U  # noqa: E402
from oracle.fcidump import read_fcidump  # noqa: E402

CASES = [("H2O_CATION.FCIDUMP.gz", 1), ("H2O.FCIDUMP.gz", 2)]


def _oracle(name, ms2):
    fd = read_fcidump(os.path.join(GOLDEN, name))
    na, nb = (fd.nelec + ms2) // 2, (fd.nelec - ms2) // 2
    return U.USystem(fd.int1, fd.int2.transpose(0, 2, 1, 3), fd.int0, na, nb), fd


@pytest.mark.parametrize("name,ms2", CASES)
def test_ucc_residual_fock_mp2_match_oracle(name, ms2):
    S, fd = _oracle(name, ms2)
    ctx = E.UContext.from_fcidump(os.path.join(GOLDEN, name), ms2=ms2)
    try:
        assert (ctx.nocca, ctx.noccb) == (S.na, S.nb)
        fa, fb, ehf = ctx.fock()
        assert abs(ehf - U.hf_energy(S)) < 1e-10
        T1o, T2o, emp2o = U.calc_UMP2(S)
        Tg, emp2 = ctx.mp2()
        assert abs(emp2 - emp2o) < 1e-10
        for g, ref in zip(Tg, S.to_blocks(T1o, T2o)):
            assert np.max(np.abs(g - ref)) < 1e-12
        rng = np.random.default_rng(7)
        blocks = [0.05 * rng.standard_normal(z.shape) for z in ctx.zeros()]
        blocks[2] = blocks[2] - blocks[2].transpose(1, 0, 2, 3)
        blocks[2] = blocks[2] - blocks[2].transpose(0, 1, 3, 2)
        blocks[3] = blocks[3] - blocks[3].transpose(1, 0, 2, 3)
        blocks[3] = blocks[3] - blocks[3].transpose(0, 1, 3, 2)
        T1s, T2s = S.from_blocks(*blocks)
        Ro = S.to_blocks(*U.calc_ucc_resid(S, T1s, T2s))
        Rg = ctx.residual(blocks, E.ECG_UCCSD)
        for k, (g, ref) in enumerate(zip(Rg, Ro)):
            assert np.max(np.abs(g - ref)) < 1e-10, (k, np.max(np.abs(g - ref)))
        # UCCD: the singles are zero
        T1z = np.zeros_like(T1s)
        Rod = S.to_blocks(*U.calc_ucc_resid(S, T1z, T2s))
        Rgd = ctx.residual(blocks, E.ECG_UCCD)
        assert np.max(np.abs(Rgd[0])) == 0.0 and np.max(np.abs(Rgd[1])) == 0.0
        for g, ref in zip(Rgd[2:], Rod[2:]):
            assert np.max(np.abs(g - ref)) < 1e-10
        # UDCSD: every distinguishable-cluster change contributes for random amplitudes
        Rodc = S.to_blocks(*U.calc_ucc_resid(S, T1s, T2s, dc=True))
        Rgdc = ctx.residual(blocks, E.ECG_UDCSD)
        for k, (g, ref) in enumerate(zip(Rgdc, Rodc)):
            assert np.max(np.abs(g - ref)) < 1e-10, (k, np.max(np.abs(g - ref)))
        assert np.max(np.abs(Rgdc[4] - Rg[4])) > 1e-4
    finally:
        ctx.close()


@pytest.mark.parametrize("name,ms2,method,eref", [
    ("H2O_CATION.FCIDUMP.gz", 1, "uccsd", -0.232320803220),
    ("H2O_CATION.FCIDUMP.gz", 1, "uccd", -0.213495204514),
    ("H2O_CATION.FCIDUMP.gz", 1, "udcsd", -0.243223819179),
    ("H2O.FCIDUMP.gz", 2, "uccsd", -0.276560985613),
])
def test_ucc_solve_matches_oracle_and_reference(name, ms2, method, eref):
    S, fd = _oracle(name, ms2)
    eo, _, _, conv_o, _ = U.calc_ucc(S, thr=1e-14, singles=(method != "uccd"), dc=(method == "udcsd"))
    assert conv_o
    ctx = E.UContext.from_fcidump(os.path.join(GOLDEN, name), ms2=ms2)
    try:
        m = {"uccsd": E.ECG_UCCSD, "uccd": E.ECG_UCCD, "udcsd": E.ECG_UDCSD}[method]
        e, T, it, conv = ctx.solve(m, E.default_opts(thr=1e-14, maxit=100))
        assert conv, it
        print(f"\n{name} {method}: E = {e[0]:.12f} (oracle {eo:.12f}, reference {eref:.12f}) it = {it}", flush=True)
        assert abs(e[0] - eo) < 1e-9
        assert abs(e[0] - eref) < 1e-6
    finally:
        ctx.close()


@pytest.mark.parametrize("um,cm", [("ECG_UCCSD", "ECG_CCSD"), ("ECG_UDCSD", "ECG_DCSD")])
def test_ucc_closed_shell_limit_equals_ccsd(um, cm):
    """nocca = noccb on the closed-shell H2O fixture: the UCCSD (UDCSD) energy is the closed-shell
    CCSD (DCSD) one."""
    path = os.path.join(GOLDEN, "H2O.FCIDUMP.gz")
    ctx = E.UContext.from_fcidump(path, ms2=0)
    cs = E.Context.from_fcidump(path)
    try:
        opts = E.default_opts(thr=1e-14, maxit=100)
        eu = ctx.solve(getattr(E, um), opts)[0][0]
        ec = cs.cc_solve(getattr(E, cm), E.ECG_FP64, opts)[0][0]
        assert abs(eu - ec) < 1e-9, (eu, ec)
    finally:
        ctx.close()
        cs.close()


def test_ucc_triples_match_oracle_and_reference():
    """(T) of UCCSD (ecg_upert_triples) on the H2O triplet: equals the oracle's spin-orbital (T)
    at the same converged amplitudes (1e-10 relative, E(T) and E[T]) and gives the reference's
    UCCSD(T) energy of test/h2o_triplet.jl:8 (1e-6).  The (T) of the solve's device-resident
    amplitudes and of the same amplitudes passed back as blocks agree to 1e-12 relative: the
    blocks are rebuilt exactly antisymmetric with the spin-forbidden elements zero, which the
    solver's spin-orbital T2 holds only to rounding."""
    S, fd = _oracle("H2O.FCIDUMP.gz", 2)
    ctx = E.UContext.from_fcidump(os.path.join(GOLDEN, "H2O.FCIDUMP.gz"), ms2=2)
    try:
        e, T, it, conv = ctx.solve(E.ECG_UCCSD, E.default_opts(thr=1e-14, maxit=100))
        assert conv
        et, etb = ctx.pert_triples()
        et2, etb2 = ctx.pert_triples(T)
        assert abs(et2 - et) < 1e-12 * abs(et) and abs(etb2 - etb) < 1e-12 * abs(etb)
        eto, etbo = U.calc_pertT_unrestricted(S, *S.from_blocks(*T))
        print(f"\nU(T): {et:.12f} oracle {eto:.12f}; UCCSD(T) {e[0] + et:.12f} reference -0.295388574918", flush=True)
        assert abs(et - eto) < 1e-10 * abs(eto) and abs(etb - etbo) < 1e-10 * abs(etbo)
        assert abs(e[0] + et - (-0.295388574918)) < 1e-6
    finally:
        ctx.close()


def test_ccdriver_unrestricted_reference_keys():
    """``@cc uccsd(t)`` with ``@opt wf ms2=2`` on test/files/H2O.FCIDUMP (test/h2o_triplet.jl:4-21) through
    the host mirror: HF, MP2, UCCSD and UCCSD(T) at the reference's 1e-6."""
    e = E.ccdriver_unrestricted(os.path.join(GOLDEN, "H2O.FCIDUMP.gz"), "uccsd(t)", ms2=2)
    assert e["converged"]
    ref = {"HF": -75.62407982361415, "MP2": -0.22401008330020 - 75.62407982361415,
           "UCCSD": -0.276560985613 - 75.62407982361415, "UCCSD(T)": -0.295388574918 - 75.62407982361415}
    for k, v in ref.items():
        assert abs(e[k] - v) < 1e-6, (k, e[k], v)


def test_ccdriver_dcsd_on_open_shell_is_udcsd():
    """``@cc dcsd`` on the open-shell H2O+ dump (test/h2o_cation.jl:27-28) runs UDCSD: 1e-6 of the
    reference's EDCSD."""
    e = E.ccdriver_unrestricted(os.path.join(GOLDEN, "H2O_CATION.FCIDUMP.gz"), "dcsd")
    assert e["converged"]
    assert abs(e["UDCSDc"] - (-0.243223819179)) < 1e-6
