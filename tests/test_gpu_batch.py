"""Many-system batch (ecg_cc_solve_batch) on the GPU.

Bar: every system of the batch converges to the energy of its own
single-system ecg_cc_solve within 1e-10 Eh (FP64); the mixed-precision batch
(FP32 products with FP64 defect correction) within 1e-9 Eh of the FP64 batch
(both at thr = 1e-12); the smallest system matches the CPU oracle within 1e-9 Eh.  Sizes: the first
systems of the seeded batch (nocc uniform in 4..16, nvirt = 9 nocc,
naux = 3 (nocc + nvirt), seeds 100..), factor width as in batch_sizes.
"""
import numpy as np
import pytest

from gpu_helpers import requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]

import elemco_amd as E  # noqa: E402
from elemco_amd.synthetic import batch_sizes  # noqa: E402
from oracle import ccsd as O  # noqa: E402
from oracle.synthetic import make_df_system  # noqa: E402

NSYS = 6


def _systems():
    sizes = [s for s in batch_sizes(24) if s[0] <= 8][:NSYS]
    return sizes, [E.ECInfo.synthetic(*s[:4], scale=s[4]) for s in sizes]


def test_batch_matches_single_solves_and_oracle():
    sizes, ecs = _systems()
    ctxs = [ec.ctx for ec in ecs]
    opts = E.default_opts(maxit=100, thr=1e-12)
    e64, it64, conv64 = E.cc_solve_batch(ctxs, E.ECG_CCSD, E.ECG_FP64, opts)
    assert conv64.all(), it64
    for k, ec in enumerate(ecs):
        e, _, _, it, conv = ec.ctx.cc_solve(E.ECG_CCSD, E.ECG_FP64, opts)
        assert conv
        assert abs(float(e[0]) - float(e64[k, 0])) < 1e-10, (sizes[k], float(e[0]), float(e64[k, 0]))
    e32, it32, conv32 = E.cc_solve_batch(ctxs, E.ECG_CCSD, E.ECG_FP32, opts, max_threads=3)
    assert conv32.all(), it32
    assert float(np.max(np.abs(e32[:, 0] - e64[:, 0]))) < 1e-9
    # the smallest system against the oracle
    k = int(np.argmin([s[0] for s in sizes]))
    o, v, q, seed, sc = sizes[k]
    ECo = O.ECInfo.from_df(make_df_system(o, v, q, seed, sc))
    ECo.options["thr"] = 1e-12
    ECo.options["maxit"] = 100
    O.calc_fock_matrix(ECo)
    O.calc_MP2(ECo)
    Eo, conv, _ = O.calc_cc(ECo, dc=False)
    assert conv
    assert abs(float(Eo["E"]) - float(e64[k, 0])) < 1e-9


def test_batch_reports_failing_system():
    """A bad method code fails every system; the error names the system."""
    _, ecs = _systems()
    with pytest.raises(E.ElemcoError, match="system 0"):
        E.cc_solve_batch([ec.ctx for ec in ecs[:2]], 7, E.ECG_FP64)


def test_batch_with_graph_replay_matches_eager():
    """The batch with every context replaying its captured residual graph (opts.graphs = 1,
    one capture per host thread's stream) gives the eager batch's energies (1e-12 Eh)."""
    _, ecs = _systems()
    ctxs = [ec.ctx for ec in ecs]
    opts = dict(maxit=100, thr=1e-12)
    e0, it0, c0 = E.cc_solve_batch(ctxs, E.ECG_CCSD, E.ECG_FP64, E.default_opts(**opts))
    e1, it1, c1 = E.cc_solve_batch(ctxs, E.ECG_CCSD, E.ECG_FP64, E.default_opts(graphs=1, **opts))
    assert c0.all() and c1.all() and (it0 == it1).all()
    assert float(np.max(np.abs(e1[:, 0] - e0[:, 0]))) < 1e-12


def test_batch_leaves_single_solves_unchanged():
    """During a batch every product splits K for a quarter of the device (the systems share
    the CUs); afterwards each context's own solves split as before, so a single solve gives
    bit-identical energies before and after a batch (eager and with graph replay)."""
    _, ecs = _systems()
    ctxs = [ec.ctx for ec in ecs[-2:]]
    opts = E.default_opts(maxit=100, thr=1e-12)
    gopts = E.default_opts(maxit=100, thr=1e-12, graphs=1)
    before = [c.cc_solve(E.ECG_CCSD, E.ECG_FP64, opts)[0].copy() for c in ctxs]
    E.cc_solve_batch(ctxs, E.ECG_CCSD, E.ECG_FP64, gopts)
    after = [c.cc_solve(E.ECG_CCSD, E.ECG_FP64, opts)[0].copy() for c in ctxs]
    for b, a in zip(before, after):
        assert np.array_equal(a, b)
    g = [c.cc_solve(E.ECG_CCSD, E.ECG_FP64, gopts)[0].copy() for c in ctxs]
    for b, a in zip(before, g):
        assert abs(float(a[0]) - float(b[0])) < 1e-12
