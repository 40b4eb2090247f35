"""GPU parity of the closed-shell CCSD/DCSD path against the CPU oracle (FP64).

Bars: residuals within 1e-10 max-abs (o=5, v=19) of the oracle; converged
energies within 1e-9 Eh of the oracle and 1e-6 Eh of the reference's own
test energies (test/h2o.jl, test/h2o_st1.jl).
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from gpu_helpers import requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]

import elemco_amd as E  # noqa: E402
from elemco_amd.synthetic import conditioned_scale  # noqa: E402
from oracle import ccsd as O  # noqa: E402
from oracle.synthetic import make_df_system  # noqa: E402

C0 = (5, 19, 72, 0)


@pytest.fixture(scope="module")
def c0():
    no, nv, nq, seed = C0
    s = make_df_system(no, nv, nq, seed)
    EC = E.ECInfo.synthetic(no, nv, nq, seed)
    ECo = O.ECInfo.from_df(s)
    O.calc_fock_matrix(ECo)
    O.calc_MP2(ECo)
    return s, EC, ECo


def test_synthetic_inputs_match_oracle_generator():
    eps, h, B, e0 = E.make_system(*C0)
    s = make_df_system(*C0)
    assert np.array_equal(eps, s.eps) and np.array_equal(B, s.B)
    assert np.max(np.abs(h - s.h)) < 1e-14


def test_fock_hf_mp2(c0):
    s, EC, ECo = c0
    f, eps, ehf = EC.ctx.fock()
    assert np.max(np.abs(f - ECo.files["f_mm"])) < 1e-12
    assert abs(ehf - O.calc_HF_energy(ECo)) < 1e-11
    T1, T2, emp2 = EC.ctx.mp2()
    assert np.max(np.abs(T2 - ECo.files["T_vvoo"])) < 1e-12
    assert np.max(np.abs(T1 - ECo.files["T_vo"])) < 1e-12
    e = O.calc_doubles_energy(ECo, ECo.files["T_vvoo"])["E"] + O.calc_singles_energy(ECo, ECo.files["T_vo"], True)["E"]
    assert abs(emp2[0] - e) < 1e-11


@pytest.mark.parametrize("dc", [False, True])
@pytest.mark.parametrize("amps", ["mp2", "random"])
def test_residual_c0(c0, dc, amps):
    s, EC, ECo = c0
    T1, T2 = ECo.files["T_vo"].copy(), ECo.files["T_vvoo"].copy()
    if amps == "random":
        rng = np.random.default_rng(11)
        T1 = 0.05 * rng.standard_normal(T1.shape)
        T2 = 0.05 * rng.standard_normal(T2.shape)
        T2 = T2 + T2.transpose(1, 0, 3, 2)
    R1o, R2o = O.calc_cc_resid(ECo, T1, T2, dc=dc)
    R1, R2 = E.calc_cc_resid(EC, T1, T2, dc=dc)
    assert np.max(np.abs(R1 - R1o)) < 1e-10
    assert np.max(np.abs(R2 - R2o)) < 1e-10


@pytest.mark.parametrize("method", ["ccsd", "dcsd"])
def test_solve_c0(method):
    no, nv, nq, seed = C0
    s = make_df_system(no, nv, nq, seed)
    eo = O.ccdriver(O.ECInfo.from_df(s), method)
    EC = E.ECInfo.synthetic(no, nv, nq, seed)
    e = E.ccdriver(EC, method)
    assert e["converged"]
    assert abs(e[method.upper() + "c"] - eo[method.upper() + "c"]) < 1e-9
    assert abs(e["HF"] - eo["HF"]) < 1e-10


@pytest.mark.parametrize("method", [E.ECG_CCSD, E.ECG_DCSD])
@pytest.mark.parametrize("prec,refine_every", [(E.ECG_FP32, 0), (E.ECG_FP32, 4), (E.ECG_BF16X3, 0)])
def test_mixed_solve_converges_to_fp64(method, prec, refine_every):
    """Defect-corrected low-precision solve: same converged energy as FP64 (1e-9 Eh,
    both converged to thr = 1e-13 so that the comparison does not depend on the
    trajectory), convergence declared on an FP64 iteration, few FP64 residuals."""
    from elemco_amd.synthetic import conditioned_scale
    # BF16 hi/lo products are used only for large products (M N / (M + N) >= 256)
    no, nv, nq, seed = (6, 40, 90, 21) if prec == E.ECG_FP32 else (10, 90, 300, 5)
    EC = E.ECInfo.synthetic(no, nv, nq, seed, scale=0.7 * conditioned_scale(no, nv, nq))
    opts = E.default_opts(maxit=80, refine_every=refine_every, thr=1e-13)
    e64, _, _, it64, c64 = EC.ctx.cc_solve(method, E.ECG_FP64, opts)
    e32, _, _, it32, c32 = EC.ctx.cc_solve(method, prec, opts)
    n64 = EC.ctx.stats()["fp64_iterations"]
    assert c64 and c32
    assert abs(float(e32[0]) - float(e64[0])) < 1e-9
    assert 1 <= n64 < it32
    assert float(e32[5]) < 1e-13      # the final (exact) residual norm meets thr


def test_h2o_fcidump_ccsd():
    """Reference fixture test/files/H2O.FCIDUMP through the DF factorisation (test/h2o.jl:6-8)."""
    EC = E.ECInfo.from_fcidump(os.path.join(GOLDEN, "H2O.FCIDUMP.gz"))
    e = E.ccdriver(EC, "ccsd")
    assert abs(e["HF"] - (-75.6457645933)) < 1e-6
    assert abs(e["MP2c"] - (-0.287815830908)) < 1e-6
    from oracle.fcidump import read_fcidump
    eo = O.ccdriver(O.ECInfo.from_fcidump(read_fcidump(os.path.join(GOLDEN, "H2O.FCIDUMP.gz"))), "ccsd")
    assert abs(e["CCSDc"] - eo["CCSDc"]) < 1e-9


@pytest.mark.parametrize("name,method,eref", [("H2O.FCIDUMP.gz", E.ECG_CCSD, None),
                                              ("H2O_ST1.FCIDUMP.gz", E.ECG_DCSD, -0.082498102641)])
def test_create_from_fcidump_one_call(name, method, eref):
    """ecg_create_from_fcidump (native reader + exact DF factors + ecg_create) gives the
    energies of the host-mirror path (ECInfo.from_fcidump) and of the reference (h2o_st1.jl:9)."""
    from elemco_amd.lib import Context
    path = os.path.join(GOLDEN, name)
    ctx = Context.from_fcidump(path)
    EC = E.ECInfo.from_fcidump(path)
    assert (ctx.nocc, ctx.nvirt, ctx.naux) == (EC.ctx.nocc, EC.ctx.nvirt, EC.ctx.naux)
    opts = EC.opts()
    e1 = ctx.cc_solve(method, E.ECG_FP64, opts)[0]
    e2 = EC.ctx.cc_solve(method, E.ECG_FP64, opts)[0]
    assert abs(float(e1[0]) - float(e2[0])) < 1e-12
    if eref is not None:
        assert abs(float(e1[0]) - eref) < 1e-6
    ctx.close()


@pytest.mark.parametrize("method,eref", [("ccsd", -0.082041632192), ("dcsd", -0.082498102641)])
def test_h2o_st1_fcidump(method, eref):
    """Non-Hermitian fixture (test/h2o_st1.jl:9): two-sided DF factors."""
    EC = E.ECInfo.from_fcidump(os.path.join(GOLDEN, "H2O_ST1.FCIDUMP.gz"))
    e = E.ccdriver(EC, method)
    assert abs(e["HF"] - (-76.298014304953)) < 1e-6
    assert abs(e[method.upper() + "c"] - eref) < 1e-6


C1_GOLDEN = os.path.join(GOLDEN, "c1_oracle.json")


@pytest.mark.skipif(not os.path.exists(C1_GOLDEN), reason="c1 golden vectors not generated")
@pytest.mark.parametrize("method", ["ccsd", "dcsd"])
def test_residual_c1_golden(method):
    """o=20, v=180, naux=600: residual at MP2 amplitudes vs committed oracle values."""
    g = json.load(open(C1_GOLDEN))
    c = g["config"]
    EC = E.ECInfo.synthetic(c["nocc"], c["nvirt"], c["naux"], c["seed"])
    T1, T2, _ = EC.ctx.mp2()
    R1, R2 = E.calc_cc_resid(EC, T1, T2, dc=(method == "dcsd"))
    gr = g[f"resid_mp2_{method}"]
    R1o = np.array(gr["R1"]).reshape(R1.shape, order="F")
    assert np.max(np.abs(R1 - R1o)) < 1e-10
    chk = gr["R2"]
    assert abs(np.sum(R2 * R2) - chk["sumsq"]) < 1e-10 * max(1.0, chk["sumsq"])
    w = np.cos(np.arange(R2.size, dtype=np.float64) * 0.001).reshape(R2.shape, order="F")
    assert abs(np.sum(R2 * w) - chk["weighted"]) < 1e-9
    for idx, val in chk["samples"]:
        assert abs(R2[tuple(idx)] - val) < 1e-10


@pytest.mark.skipif(not os.path.exists(C1_GOLDEN), reason="c1 golden vectors not generated")
def test_conditioned_c1_energies_and_triples_golden():
    """o=20, v=180, naux=600 (factor width scaled by conditioned_scale): converged
    DCSD and CCSD energies and CCSD(T) against the committed oracle values."""
    g = json.load(open(C1_GOLDEN))
    c, cond = g["config"], g["conditioned"]
    EC = E.ECInfo.synthetic(c["nocc"], c["nvirt"], c["naux"], c["seed"], scale=cond["scale"])
    _, _, ehf = EC.ctx.fock()
    assert abs(ehf - cond["HF"]) < 1e-9
    for meth in ("dcsd", "ccsd"):
        e = E.ccdriver(EC, meth)
        assert e["converged"]
        assert abs(e[meth.upper() + "c"] - cond[meth.upper() + "c"]) < 1e-9
    et, etb = EC.ctx.pert_triples(EC.files["T_vo"], EC.files["T_vvoo"])
    assert abs(et - cond["(T)"]) < 1e-9
    assert abs(etb - cond["[T]"]) < 1e-9


@pytest.mark.parametrize("method,prec", [(E.ECG_CCSD, E.ECG_FP64), (E.ECG_DCSD, E.ECG_FP64),
                                         (E.ECG_CCSD, E.ECG_FP32)])
def test_graph_replayed_residual_matches_eager(method, prec):
    """Solves with the residual replayed from captured HIP graphs (opts.graphs = 1) give the
    eager solve's iterations and energies to rounding: every iteration replays the graph with
    eager updates, DIIS and host reductions in between, and the mixed solve also replays the
    FP64 refresh residual into the defect buffers.  Bar: 1e-12 Eh, same iteration count."""
    cfg = (6, 40, 90, 21)
    EC = E.ECInfo.synthetic(*cfg, scale=0.7 * conditioned_scale(*cfg[:3]))
    opts = dict(thr=1e-12, maxit=100)
    e0, T10, T20, it0, conv0 = EC.ctx.cc_solve(method, prec, E.default_opts(**opts))
    e1, T11, T21, it1, conv1 = EC.ctx.cc_solve(method, prec, E.default_opts(graphs=1, **opts))
    e2, _, _, it2, conv2 = EC.ctx.cc_solve(method, prec, E.default_opts(graphs=1, **opts))   # cached graphs
    assert conv0 and conv1 and conv2 and it0 == it1 == it2
    assert abs(e1[0] - e0[0]) < 1e-12 and abs(e2[0] - e0[0]) < 1e-12
    assert float(np.max(np.abs(T21 - T20))) < 1e-10
    # an eager residual between replays (other work on the same stream) leaves them valid
    R1, R2 = EC.ctx.cc_residual(T10, T20, method)
    e3, _, _, it3, _ = EC.ctx.cc_solve(method, prec, E.default_opts(graphs=1, **opts))
    assert it3 == it0 and abs(e3[0] - e0[0]) < 1e-12


@pytest.mark.parametrize("name,method,orbs", [("H2O.FCIDUMP.gz", E.ECG_CCSD, [1]),
                                              ("H2O_ST1.FCIDUMP.gz", E.ECG_DCSD, [1]),
                                              ("H2O.FCIDUMP.gz", E.ECG_DCSD, [1, 23])])
def test_frozen_orbitals_match_oracle(name, method, orbs):
    """@freeze_orbs (freeze_orbs_in_dump, dumptools.jl:20-62; test/h2o_st1.jl:33) before the
    CC solve: the context made from the frozen dump in memory (ecg_fcidump_freeze,
    ecg_create_from_fcidump_handle) gives the oracle's energy on the oracle-frozen dump, and
    the HF energy of the full dump."""
    from elemco_amd.lib import Context, Fcidump
    from oracle.ccsd import ECInfo as OECInfo, ccdriver as occdriver
    from oracle.fcidump import read_fcidump as oread, freeze_orbs_in_dump as ofreeze
    path = os.path.join(GOLDEN, name)
    full = Context.from_fcidump(path)
    ehf_full = full.fock()[2]
    full.close()
    fd = Fcidump(path)
    fd.freeze(orbs)
    ctx = Context.from_fcidump(fd)
    assert ctx.norb == fd.norb
    assert abs(ctx.fock()[2] - ehf_full) < 1e-10
    EC = E.ECInfo.from_fcidump(path)
    opts = EC.opts()
    e1 = float(ctx.cc_solve(method, E.ECG_FP64, opts)[0][0])
    mname = {E.ECG_CCSD: "ccsd", E.ECG_DCSD: "dcsd"}[method]
    eo = occdriver(OECInfo.from_fcidump(ofreeze(oread(path), orbs)), mname)
    assert abs(e1 - eo[mname.upper() + "c"]) < 1e-9
    ctx.close()
    fd.close()


@pytest.mark.parametrize("name,method,orbs", [("H2O.FCIDUMP.gz", E.ECG_CCSD, [1]),
                                              ("H2O_ST1.FCIDUMP.gz", E.ECG_DCSD, [1, 20])])
def test_df_freeze_energy_matches_frozen_dump(name, method, orbs):
    """Frozen core on the DF route (ecg_df_freeze, then ecg_create) gives the energy of the
    frozen four-index dump (ecg_fcidump_freeze, ecg_create_from_fcidump_handle)."""
    from elemco_amd.lib import Context, Fcidump, df_freeze
    path = os.path.join(GOLDEN, name)
    fd = Fcidump(path)
    BL, BR = fd.df_factors(1e-14)
    h, bl, br, e0, no, nv = df_freeze(fd.int1, BL, BR, fd.int0, fd.nelec // 2, orbs)
    c1 = Context(no, nv, bl.shape[2], h, bl, br, e0)
    fd.freeze(orbs)
    c2 = Context.from_fcidump(fd)
    EC = E.ECInfo.from_fcidump(path)
    opts = EC.opts()
    e1 = float(c1.cc_solve(method, E.ECG_FP64, opts)[0][0])
    e2 = float(c2.cc_solve(method, E.ECG_FP64, opts)[0][0])
    assert abs(e1 - e2) < 1e-9
    assert abs(c1.fock()[2] - c2.fock()[2]) < 1e-9
    c1.close()
    c2.close()
    fd.close()
