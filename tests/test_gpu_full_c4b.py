"""The full 64-system batch (nocc uniform in 4..16, nvirt = 9 nocc, naux = 3 (nocc + nvirt),
seeds 100..163) through ecg_cc_solve_batch.

Bars:
* every system converges, and its batched FP64 energy equals the energy of its
  own single-system ecg_cc_solve to 1e-10 Eh;
* the batch with FP32 products (defect-corrected, ending on FP64 iterations)
  is within 1e-9 Eh of the FP64 batch for every system;
* the batch of device-resident iterations (opts.graphs = 1: reductions, DIIS and the
  convergence test on the device, one captured graph per system replayed every
  iteration, one host thread driving all systems) takes the same iterations as the
  host-driven batch and gives its energies to 1e-10 Eh;
* the smallest and the largest system match the CPU oracle's calc_cc (the
  reference's cc_iterations!, cc.jl:2518-2648) to 1e-9 Eh.
"""
import numpy as np
import pytest

from gpu_helpers import requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]

import elemco_amd as E  # noqa: E402
from elemco_amd.synthetic import batch_sizes  # noqa: E402
from oracle import ccsd as O  # noqa: E402
from oracle.synthetic import make_df_system  # noqa: E402

THR = 1e-12


@pytest.fixture(scope="module")
def batch():
    sizes = batch_sizes(64)
    assert len(sizes) == 64 and {s[0] for s in sizes} <= set(range(4, 17))
    print(f"\n64-system batch, nocc {min(s[0] for s in sizes)}..{max(s[0] for s in sizes)}", flush=True)
    ecs = [E.ECInfo.synthetic(*s[:4], scale=s[4]) for s in sizes]
    yield sizes, ecs
    for ec in ecs:
        ec.ctx.close()


def test_c4b_batch_matches_single_solves(batch):
    sizes, ecs = batch
    ctxs = [ec.ctx for ec in ecs]
    opts = E.default_opts(maxit=100, thr=THR)
    e64, it64, conv64 = E.cc_solve_batch(ctxs, E.ECG_CCSD, E.ECG_FP64, opts)
    assert conv64.all(), it64
    worst = 0.0
    for k, ec in enumerate(ecs):
        e, _, _, it, conv = ec.ctx.cc_solve(E.ECG_CCSD, E.ECG_FP64, opts)
        assert conv, sizes[k]
        worst = max(worst, abs(float(e[0]) - float(e64[k, 0])))
    print(f"  max |E_batch - E_single| = {worst:.2e}", flush=True)
    assert worst < 1e-10
    e32, it32, conv32 = E.cc_solve_batch(ctxs, E.ECG_CCSD, E.ECG_FP32, opts)
    assert conv32.all(), it32
    d = float(np.max(np.abs(e32[:, 0] - e64[:, 0])))
    print(f"  max |E_mixed - E_fp64| = {d:.2e}", flush=True)
    assert d < 1e-9
    E.cc_solve_batch(ctxs, E.ECG_CCSD, E.ECG_FP64, E.default_opts(maxit=100, thr=THR, graphs=1))   # captures
    eg, itg, convg = E.cc_solve_batch(ctxs, E.ECG_CCSD, E.ECG_FP64, E.default_opts(maxit=100, thr=THR, graphs=1))
    assert convg.all()
    dg = float(np.max(np.abs(eg[:, 0] - e64[:, 0])))
    print(f"  device-resident batch: max |E - E_host| = {dg:.2e}, iterations equal: {(itg == it64).all()}", flush=True)
    assert dg < 1e-10 and (itg == it64).all()
    # the extreme systems against the oracle
    for k in (int(np.argmin([s[0] for s in sizes])), int(np.argmax([s[0] for s in sizes]))):
        o, v, q, seed, sc = sizes[k]
        ECo = O.ECInfo.from_df(make_df_system(o, v, q, seed, sc))
        ECo.options["thr"] = THR
        ECo.options["maxit"] = 100
        O.calc_fock_matrix(ECo)
        O.calc_MP2(ECo)
        Eo, conv, _ = O.calc_cc(ECo, dc=False)
        assert conv
        print(f"  system {k} (o={o}): oracle {float(Eo['E']):.12f} gpu {float(e64[k, 0]):.12f}", flush=True)
        assert abs(float(Eo["E"]) - float(e64[k, 0])) < 1e-9
