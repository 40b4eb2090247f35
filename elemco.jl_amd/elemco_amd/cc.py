"""Host-side mirror of ElemCo.jl's closed-shell coupled-cluster entry points.

Names, arguments and results follow the reference (file:line relative to the
ElemCo.jl v0.14.0 source tree); all arithmetic runs in the HIP library
(``libelemco_hip.so``) on the GPU.  Arrays use the reference's index order
(``T1[a,i]``, ``T2[a,b,i,j]``) in column-major storage.

* ``ECInfo``                       -- ``src/infos/ecinfos.jl:103-155`` (integrals + options)
* ``ccdriver(EC, method)``         -- ``src/cc/ccdriver.jl:37-76,348-393`` (``@cc`` macro, ``ElemCo.jl:523-542``)
* ``calc_MP2``                     -- ``src/cc/cc.jl:1011-1024``
* ``calc_cc_resid``                -- ``src/cc/cc.jl:1343-1489``
* ``calc_cc``                      -- ``src/cc/cc.jl:2451-2648``
* ``calc_pertT``                   -- ``src/cc/cc_triples.jl:11-35``
* ``calc_lm_cc``                   -- ``src/cc/cc_lagrange.jl:1120-1223``
* ``calc_ccsd_vector_times_Jacobian`` -- ``src/cc/cc_lagrange.jl:9-212``
* ``calc_1RDM``                    -- ``src/cc/cc_lagrange.jl:793-837``
Errors are raised as ``ElemcoError`` (the reference raises ``error(...)``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import lib
from .fcidump import read_fcidump
from .synthetic import make_system


@dataclass
class CcOptions:
    """Closed-shell subset of ``CcOptions`` (``src/infos/options.jl:128-221``) plus precision."""
    thr: float = 1e-10
    conven: float = 0.1
    maxit: int = 50
    shifts: float = 0.15
    shiftp: float = 0.2
    precision: str = "fp64"      # "fp64" | "fp32" | "bf16x3"
    refine_every: int = 0        # mixed precision: FP64 residual every k-th iteration


@dataclass
class DiisOptions:
    """``DiisOptions`` (``src/infos/options.jl:272-282``)."""
    maxdiis: int = 6
    resthr: float = 10.0


@dataclass
class Options:
    cc: CcOptions = field(default_factory=CcOptions)
    diis: DiisOptions = field(default_factory=DiisOptions)


class ECInfo:
    """Integrals (as DF factors on the GPU), orbital spaces, options and results."""

    def __init__(self, nocc, h, BL, BR=None, e0=0.0, device=0, workspace_bytes=0):
        n = h.shape[0]
        self.nocc = int(nocc)
        self.nvirt = n - self.nocc
        self.norb = n
        self.options = Options()
        self.ctx = lib.Context(self.nocc, self.nvirt, BL.shape[2], h, BL, BR, e0, device, workspace_bytes)
        self.files = {}

    @classmethod
    def from_fcidump(cls, path, **kw):
        """``ccdriver(EC, method; fcidump)`` input path (``dump.jl:326``)."""
        fd = read_fcidump(path)
        BL, BR = fd.df_factors()
        return cls(fd.nelec // 2, fd.int1, BL, BR, fd.int0, **kw)

    @classmethod
    def synthetic(cls, nocc, nvirt, naux, seed, scale=1.0, **kw):
        """Seeded synthetic DF system; ``scale`` multiplies the 0.05 factor width."""
        eps, h, B, e0 = make_system(nocc, nvirt, naux, seed, scale)
        return cls(nocc, h, B, None, e0, **kw)

    def opts(self):
        o = self.options
        return lib.default_opts(thr=o.cc.thr, conven=o.cc.conven, shifts=o.cc.shifts, shiftp=o.cc.shiftp,
                                resthr=o.diis.resthr, maxit=o.cc.maxit, maxdiis=o.diis.maxdiis,
                                refine_every=o.cc.refine_every)

    @property
    def precision(self):
        return lib.PRECISIONS[self.options.cc.precision]


def _method(method: str):
    m = method.lower().replace("λ", "l").replace("Λ", "l")
    lam = m.startswith("l")
    if lam:
        m = m[1:]
    pert = m.endswith("(t)")
    base = m[:-3] if pert else m
    if base not in ("ccsd", "dcsd"):
        raise lib.ElemcoError(f"method {method} is not on the closed-shell CCSD/DCSD path")
    return base, lam, pert


def calc_fock_matrix(EC: ECInfo):
    """``calc_fock_matrix``/``calc_HF_energy`` (``src/cc/cctools.jl:35-76``)."""
    f, eps, ehf = EC.ctx.fock()
    EC.files["f_mm"], EC.files["e_m"] = f, eps
    return ehf


def calc_MP2(EC: ECInfo):
    """``calc_MP2`` (``src/cc/cc.jl:1011-1024``): stores ``T_vo``/``T_vvoo``, returns energies."""
    T1, T2, e = EC.ctx.mp2()
    EC.files["T_vo"], EC.files["T_vvoo"] = T1, T2
    return dict(E=e[0], ESS=e[1], EOS=e[2])


def calc_cc_resid(EC: ECInfo, T1, T2, dc=False):
    """``calc_cc_resid(EC, T1, T2; dc)`` (``src/cc/cc.jl:1343-1489``) -> (R1, R2)."""
    return EC.ctx.cc_residual(T1, T2, lib.ECG_DCSD if dc else lib.ECG_CCSD, EC.precision)


def calc_cc(EC: ECInfo, method: str = "ccsd"):
    """``calc_cc`` closed-shell (``src/cc/cc.jl:2451-2648``) from the MP2 guess."""
    base, _, _ = _method(method)
    e, T1, T2, it, conv = EC.ctx.cc_solve(lib.ECG_DCSD if base == "dcsd" else lib.ECG_CCSD, EC.precision,
                                          EC.opts())
    EC.files["T_vo"], EC.files["T_vvoo"] = T1, T2
    if not conv:
        print("WARNING: CC iterations did not converge!")
    return dict(E=e[0], ESS=e[1], EOS=e[2], pE=e[3], iterations=it, converged=conv)


def calc_lm_cc(EC: ECInfo, method: str = "ccsd"):
    """``calc_lm_cc`` closed-shell (``src/cc/cc_lagrange.jl:1120-1223``): stores ``U_vo``/``U_vvoo``."""
    base, _, _ = _method(method)
    U1, U2, it, conv = EC.ctx.lambda_solve(EC.files["T_vo"], EC.files["T_vvoo"],
                                           lib.ECG_DCSD if base == "dcsd" else lib.ECG_CCSD, EC.opts())
    EC.files["U_vo"], EC.files["U_vvoo"] = U1, U2
    if not conv:
        print("WARNING: CC-LM iterations did not converge!")
    return dict(iterations=it, converged=conv)


def calc_ccsd_vector_times_Jacobian(EC: ECInfo, U1, U2, dc=False):
    """``calc_ccsd_vector_times_Jacobian(EC, U1, U2; dc)`` (``cc_lagrange.jl:9-212``) at ``T_vo``/``T_vvoo``."""
    return EC.ctx.lambda_residual(EC.files["T_vo"], EC.files["T_vvoo"], U1, U2,
                                  lib.ECG_DCSD if dc else lib.ECG_CCSD, EC.precision)


def calc_1RDM(EC: ECInfo, U1, U2, T1, T2):
    """``calc_1RDM`` (``cc_lagrange.jl:793-837``) -> (D1, dD1)."""
    return EC.ctx.one_rdm(T1, T2, U1, U2)


def calc_pertT(EC: ECInfo, method: str = "ccsd(t)"):
    """``calc_pertT`` closed-shell (``src/cc/cc_triples.jl:11-35``) -> {"ET3": E(T), "ET3b": E[T]}."""
    _, lam, _ = _method(method)
    F = EC.files
    if lam:
        et, etb = EC.ctx.pert_triples(F["T_vo"], F["T_vvoo"], F["U_vo"], F["U_vvoo"], EC.precision)
    else:
        et, etb = EC.ctx.pert_triples(F["T_vo"], F["T_vvoo"], precision=EC.precision)
    return {"ET3": et, "ET3b": etb}


def ccdriver(EC: ECInfo, method: str):
    """``ccdriver``/``eval_cc_groundstate`` for the closed-shell path (``ccdriver.jl:37-76,348-393``).

    Returns the reference's energy keys: ``HF``, ``MP2``, ``MP2c``, ``<M>``, ``<M>c``
    and for (T) methods ``[T]``, ``(T)``, ``<M>(T)``, ``<M>(T)c`` (``<M>`` = CCSD,
    DCSD, ΛCCSD, ...).
    """
    base, lam, pert = _method(method)
    out = {}
    ehf = calc_fock_matrix(EC)
    out["HF"] = ehf
    emp2 = calc_MP2(EC)
    out["MP2c"] = emp2["E"]
    out["MP2"] = emp2["E"] + ehf
    ecc = calc_cc(EC, base)
    name = base.upper()
    out[name + "c"] = ecc["E"]
    out[name] = ecc["E"] + ehf
    out["iterations"] = ecc["iterations"]
    out["converged"] = ecc["converged"]
    if lam:
        calc_lm_cc(EC, base)
        name = "Λ" + name
    if pert:
        t = calc_pertT(EC, ("λ" if lam else "") + base + "(t)")
        out["[T]"] = t["ET3b"]
        out["(T)"] = t["ET3"]
        out[name + "(T)c"] = ecc["E"] + t["ET3"]
        out[name + "(T)"] = ecc["E"] + t["ET3"] + ehf
    return out


def ccdriver_unrestricted(fcidump: str, method: str, ms2=None, thr=1e-10, maxit=100):
    """``ccdriver(EC, method; fcidump)`` for the open-shell methods ``uccsd``, ``uccd``, ``udcsd``
    (also spelled ``dcsd``, which the reference runs as UDCSD on an open-shell system) and
    ``uccsd(t)`` (``src/cc/ccdriver.jl:37-76,348-393``; ``@opt wf ms2=...`` is ``ms2``, None keeps
    the file's MS2).

    Returns the reference's energy keys: ``HF``, ``MP2``, ``MP2c``, ``<M>``, ``<M>c`` and for
    ``uccsd(t)`` ``[T]``, ``(T)``, ``UCCSD(T)``, ``UCCSD(T)c``.
    """
    m = method.lower()
    pert = m.endswith("(t)")
    base = m[:-3] if pert else m
    if base == "dcsd":
        base = "udcsd"   # dcsd on an open-shell system (checkset_unrestricted_closedshell!)
    if base not in ("uccsd", "uccd", "udcsd") or (pert and base != "uccsd"):
        raise lib.ElemcoError(f"method {method} is not on the open-shell path (uccsd, uccd, udcsd, uccsd(t))")
    ctx = lib.UContext.from_fcidump(fcidump, -1 if ms2 is None else int(ms2))
    try:
        out = {}
        _, _, ehf = ctx.fock()
        out["HF"] = ehf
        _, emp2 = ctx.mp2()
        out["MP2c"], out["MP2"] = emp2, emp2 + ehf
        code = {"uccsd": lib.ECG_UCCSD, "uccd": lib.ECG_UCCD, "udcsd": lib.ECG_UDCSD}[base]
        e, _, it, conv = ctx.solve(code, lib.default_opts(thr=thr, maxit=maxit))
        if not conv:
            print("WARNING: CC iterations did not converge!")
        name = base.upper()
        out[name + "c"], out[name] = e[0], e[0] + ehf
        out["iterations"], out["converged"] = it, conv
        if pert:
            et, etb = ctx.pert_triples()
            out["[T]"], out["(T)"] = etb, et
            out["UCCSD(T)c"], out["UCCSD(T)"] = e[0] + et, e[0] + et + ehf
        return out
    finally:
        ctx.close()
