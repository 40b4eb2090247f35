"""Allocation-free FP64 residual (ecg_cc_residual_dev_noalloc) against the
allocating one (ecg_cc_residual_dev), through the C ABI, on device-resident
amplitudes.

Two calls in a row, the second on a smaller system (the shared scratch array
keeps the first call's contents, so nothing the second call reads may depend on
it), each for CCSD and DCSD, on MP2 doubles plus seeded random singles (so the
T1 dressing runs).  Bar: max |R_noalloc - R_dev| <= 1e-10, the FP64 residual
parity bar against the oracle (the two differ only in summation order: the
allocation-free call uses the unpacked FP64 kernel and splits K only as far as
its scratch allows).  Also: a system beyond the scratch size and a non-FP64
precision are refused with an error, before any launch.
"""
import numpy as np
import pytest

from gpu_helpers import requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]

import elemco_amd as E  # noqa: E402
from elemco_amd.synthetic import conditioned_scale, make_system  # noqa: E402

SYSTEMS = [(12, 108, 360, 201), (6, 54, 180, 202)]   # second call smaller than the first


def _context(o, v, q, seed):
    _, h, B, _ = make_system(o, v, q, seed, conditioned_scale(o, v, q))
    return E.Context(o, v, q, h, B)


def _dev(torch, a):
    return torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).to("cuda:0")


def _host(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


def test_noalloc_residual_matches_allocating_residual():
    import torch
    for o, v, q, seed in SYSTEMS:
        ctx = _context(o, v, q, seed)
        try:
            _, T2, _ = ctx.mp2()
            rng = np.random.default_rng(seed)
            T1 = 0.02 * rng.standard_normal((v, o))
            dT1, dT2 = _dev(torch, T1), _dev(torch, T2)
            for meth in (E.ECG_CCSD, E.ECG_DCSD):
                out = {}
                for name in ("dev", "noalloc"):
                    dR1 = torch.full((v * o,), np.nan, dtype=torch.float64, device="cuda:0")
                    dR2 = torch.full((v * v * o * o,), np.nan, dtype=torch.float64, device="cuda:0")
                    torch.cuda.synchronize()
                    print(f"  o={o} v={v} naux={q} seed={seed} method={meth}: {name}", flush=True)
                    fn = ctx.cc_residual_dev if name == "dev" else ctx.cc_residual_dev_noalloc
                    fn(dT1.data_ptr(), dT2.data_ptr(), dR1.data_ptr(), dR2.data_ptr(), meth, E.ECG_FP64)
                    out[name] = (_host(dR1, (v, o)), _host(dR2, (v, v, o, o)))
                R1, R2 = out["dev"]
                S1, S2 = out["noalloc"]
                assert np.all(np.isfinite(S1)) and np.all(np.isfinite(S2))
                err = max(np.max(np.abs(S1 - R1)), np.max(np.abs(S2 - R2)))
                scale = max(np.max(np.abs(R1)), np.max(np.abs(R2)))
                print(f"    max|dR| = {err:.3e} (max|R| = {scale:.3e})", flush=True)
                assert err <= 1e-10
                assert scale > 1e-3      # a non-trivial residual was compared
        finally:
            ctx.close()


def test_noalloc_residual_refuses_what_it_cannot_do():
    import torch
    o, v, q, seed = 20, 180, 600, 203
    ctx = _context(o, v, q, seed)
    try:
        d1 = torch.zeros(v * o, dtype=torch.float64, device="cuda:0")
        d2 = torch.zeros(v * v * o * o, dtype=torch.float64, device="cuda:0")
        r1, r2 = torch.zeros_like(d1), torch.zeros_like(d2)
        torch.cuda.synchronize()
        with pytest.raises(E.ElemcoError, match="too large"):
            ctx.cc_residual_dev_noalloc(d1.data_ptr(), d2.data_ptr(), r1.data_ptr(), r2.data_ptr(), E.ECG_CCSD,
                                        E.ECG_FP64)
        with pytest.raises(E.ElemcoError, match="FP64 only"):
            ctx.cc_residual_dev_noalloc(d1.data_ptr(), d2.data_ptr(), r1.data_ptr(), r2.data_ptr(), E.ECG_CCSD,
                                        E.ECG_BF16X3)
    finally:
        ctx.close()
