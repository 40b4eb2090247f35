"""CPU-side checks of the product boundary: the C-ABI library loads and exports
every symbol declared in include/elemco_hip.h (no compute calls without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "elemco_hip.h")
LIB = os.path.join(ROOT, "elemco.jl_amd", "libelemco_hip.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(ecg_\w+)\s*\(", text, re.M)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


HOOKS = os.path.join(ROOT, "elemco.jl_amd", "libecg_testhooks.so")


@pytest.mark.skipif(not os.path.exists(HOOKS), reason="test hooks not built")
def test_test_hook_library_resolves_all_symbols():
    """The GPU tests' hook library loads (every symbol resolved at load time)."""
    h = ctypes.CDLL(HOOKS, mode=os.RTLD_NOW)
    for s in ("ecgt_contract", "ecgt_permute", "ecgt_bench_contract", "ecgt_mfma_probe", "ecgt_mfma_probe_bf16"):
        assert hasattr(h, s), s


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_covers_header():
    from elemco_amd import lib
    assert set(declared_symbols()) == set(lib.EXPORTED)


def test_default_options_match_reference():
    """CcOptions/DiisOptions defaults (src/infos/options.jl:130-138,274-276)."""
    from elemco_amd import lib
    o = lib.default_opts()
    assert (o.thr, o.conven, o.shifts, o.shiftp, o.resthr, o.maxit, o.maxdiis) == (1e-10, 0.1, 0.15, 0.2, 10.0,
                                                                                   50, 6)


def test_create_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import elemco_amd as E
    with pytest.raises(E.ElemcoError):
        E.ECInfo.synthetic(2, 5, 10, 0)


def test_df_factors_reproduce_fcidump_integrals():
    """Exact factorisation of the reference's H2O fixture (Hermitian: one factor) and ST1
    (non-Hermitian: signed factors, BR != BL)."""
    from elemco_amd.fcidump import read_fcidump
    for name, herm in (("H2O.FCIDUMP.gz", True), ("H2O_ST1.FCIDUMP.gz", False)):
        fd = read_fcidump(os.path.join(ROOT, "tests", "golden", name))
        assert fd.simtra == (not herm)
        BL, BR = fd.df_factors()
        assert (BR is None) == herm
        BRx = BL if BR is None else BR
        rec = np.einsum("pqQ,rsQ->pqrs", BL, BRx)
        assert np.max(np.abs(rec - fd.eri)) < 1e-12
        # a coarser tolerance keeps fewer factors and reconstructs to about that tolerance
        BL2, BR2 = fd.df_factors(1e-8)
        assert BL2.shape[2] < BL.shape[2]
        rec2 = np.einsum("pqQ,rsQ->pqrs", BL2, BL2 if BR2 is None else BR2)
        assert np.max(np.abs(rec2 - fd.eri)) < 1e-6


def test_product_fcidump_reader_matches_oracle():
    from elemco_amd.fcidump import read_fcidump
    from oracle.fcidump import read_fcidump as oread
    for name in ("H2O.FCIDUMP.gz", "H2O_ST1.FCIDUMP.gz"):
        p = os.path.join(ROOT, "tests", "golden", name)
        a, b = read_fcidump(p), oread(p)
        # oracle keeps <pq|rs> = (pr|qs)
        assert np.array_equal(a.eri, b.int2.transpose(0, 2, 1, 3))
        assert np.array_equal(a.int1, b.int1) and a.int0 == b.int0


def test_product_generator_matches_oracle():
    from elemco_amd.synthetic import make_system
    from oracle.synthetic import make_df_system
    eps, h, B, e0 = make_system(4, 13, 40, 3)
    s = make_df_system(4, 13, 40, 3)
    assert np.array_equal(eps, s.eps) and np.array_equal(B, s.B)
    assert np.max(np.abs(h - s.h)) < 1e-14
