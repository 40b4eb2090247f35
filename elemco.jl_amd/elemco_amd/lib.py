"""ctypes binding of libelemco_hip.so (declarations: include/elemco_hip.h).

The library is the product path; importing this module fails loudly when the
shared object is missing or cannot be loaded -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ELEMCO_HIP_LIB", os.path.join(os.path.dirname(_HERE), "libelemco_hip.so"))

ECG_CCSD, ECG_DCSD = 0, 1
ECG_FP64, ECG_FP32, ECG_BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": ECG_FP64, "fp32": ECG_FP32, "bf16x3": ECG_BF16X3}


class ElemcoError(RuntimeError):
    """Error reported by the HIP library (status < 0 and ecg_last_error())."""


class SolverOpts(ctypes.Structure):
    _fields_ = [("thr", ctypes.c_double), ("conven", ctypes.c_double), ("shifts", ctypes.c_double),
                ("shiftp", ctypes.c_double), ("resthr", ctypes.c_double), ("maxit", ctypes.c_int),
                ("maxdiis", ctypes.c_int), ("refine_every", ctypes.c_int), ("graphs", ctypes.c_int)]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGS = {
    "ecg_default_opts": [ctypes.POINTER(SolverOpts)],
    "ecg_create": [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                   ctypes.c_double, ctypes.c_int, ctypes.c_size_t],
    "ecg_destroy": [_vp],
    "ecg_last_error": [_vp],
    "ecg_fock": [_vp, _dp, _dp, _dp],
    "ecg_mp2": [_vp, _dp, _dp, _dp],
    "ecg_cc_residual": [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp],
    "ecg_cc_residual_dev": [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp],
    "ecg_cc_residual_dev_noalloc": [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp],
    "ecg_cc_state_init": [_vp, _dp, _dp, ctypes.POINTER(SolverOpts)],
    "ecg_cc_state_step": [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SolverOpts), _dp],
    "ecg_cc_state_get": [_vp, _dp, _dp],
    "ecg_cc_solve": [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SolverOpts), _dp, _dp, _dp, _ip, _ip],
    "ecg_cc_solve_batch": [ctypes.c_int, ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.POINTER(SolverOpts),
                           ctypes.c_int, _dp, _ip, _ip, _ip],
    "ecg_pert_triples": [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp],
    "ecg_pert_triples_range": [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_long, ctypes.c_long,
                               _dp, _dp],
    "ecg_lambda_residual": [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp],
    "ecg_lambda_solve": [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SolverOpts), _dp, _dp, _dp, _dp, _ip, _ip],
    "ecg_one_rdm": [_vp, _dp, _dp, _dp, _dp, _dp, _dp],
    "ecg_energy_grad": [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp],
    "ecg_energy_grad_lr": [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp],
    "ecg_profile": [_vp, ctypes.c_int],
    "ecg_stats": [_vp, _dp],
    "ecg_fcidump_read": [ctypes.c_char_p, ctypes.POINTER(_vp)],
    "ecg_fcidump_free": [_vp],
    "ecg_fcidump_from_integrals": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                   ctypes.POINTER(_vp)],
    "ecg_fcidump_info": [_vp, _ip, _ip, _ip, _ip, _dp],
    "ecg_fcidump_header": [_vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p],
    "ecg_fcidump_header_key": [_vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int],
    "ecg_fcidump_integrals": [_vp, _dp, _dp],
    "ecg_fcidump_df": [_vp, ctypes.c_double, _ip, _dp, _dp, _ip],
    "ecg_create_from_fcidump": [ctypes.POINTER(_vp), ctypes.c_char_p, ctypes.c_double, ctypes.c_int,
                                ctypes.c_size_t],
    "ecg_create_from_fcidump_handle": [ctypes.POINTER(_vp), _vp, ctypes.c_double, ctypes.c_int, ctypes.c_size_t],
    "ecg_fcidump_freeze": [_vp, _ip, ctypes.c_int, _ip, _ip],
    "ecg_fcidump_write": [_vp, ctypes.c_char_p, ctypes.c_double],
    "ecg_df_freeze": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, _ip, ctypes.c_int,
                      _ip, _ip, _dp, _dp, _dp, _dp],
    "ecg_dims": [_vp, _ip, _ip, _ip],
    "ecg_df_ao_to_mo": [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip,
                        _dp],
    "ecg_create_from_ao": [ctypes.POINTER(_vp), ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                           ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_size_t],
}
_SIGS.update({
    "ecg_ucreate": [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                    ctypes.c_double, ctypes.c_int, ctypes.c_size_t],
    "ecg_ucreate_from_fcidump": [ctypes.POINTER(_vp), ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                 ctypes.c_size_t],
    "ecg_udestroy": [_vp],
    "ecg_ulast_error": [_vp],
    "ecg_udims": [_vp, _ip, _ip, _ip, _ip],
    "ecg_ufock": [_vp, _dp, _dp, _dp],
    "ecg_ump2": [_vp, _dp, _dp, _dp, _dp, _dp, _dp],
    "ecg_ucc_residual": [_vp, ctypes.c_int] + [_dp] * 10,
    "ecg_ucc_solve": [_vp, ctypes.c_int, ctypes.POINTER(SolverOpts)] + [_dp] * 6 + [_ip, _ip],
    "ecg_upert_triples": [_vp] + [_dp] * 7,
})
ECG_UCCSD, ECG_UCCD, ECG_UDCSD = 0, 1, 2
_RESTYPES = {"ecg_default_opts": None, "ecg_last_error": ctypes.c_char_p, "ecg_ulast_error": ctypes.c_char_p}
EXPORTED = tuple(_SIGS)

_lib = None


def load():
    """Load the shared library once; raises ElemcoError if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ElemcoError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _RESTYPES.get(name, ctypes.c_int)
    _lib = lib
    return lib


def default_opts(**kw) -> SolverOpts:
    o = SolverOpts()
    load().ecg_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown solver option {k}")
        setattr(o, k, v)
    return o


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def fortran(a, shape=None):
    """Column-major FP64 copy/view of ``a`` (the Julia layout the C ABI expects)."""
    a = np.asarray(a, dtype=np.float64)
    if shape is not None and a.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {a.shape}")
    return np.asfortranarray(a)


def df_ao_to_mo(C, AAP, PQ, tol=1e-8, device=0):
    """MO DF factors ``B[p,q,L]`` from AO three-index integrals ``(mu nu|P)`` and the metric
    ``(P|Q)`` (``ecg_df_ao_to_mo``: ``generate_3idx_integrals`` with ``sqrtinvchol``,
    ``src/integrals/dftools.jl:92-125``)."""
    lib = load()
    C = fortran(C)
    nao, nmo = C.shape
    nfit = PQ.shape[0]
    AAP = fortran(AAP, (nao, nao, nfit))
    PQ = fortran(PQ, (nfit, nfit))
    B = np.zeros((nmo, nmo, nfit), order="F")
    naux = ctypes.c_int()
    st = lib.ecg_df_ao_to_mo(_ptr(C), nao, nmo, _ptr(AAP), _ptr(PQ), nfit, float(tol), int(device),
                             ctypes.byref(naux), _ptr(B))
    if st != 0:
        raise ElemcoError(f"ecg_df_ao_to_mo failed ({st}): {lib.ecg_last_error(None).decode()}")
    return np.asfortranarray(B[:, :, :naux.value])


def df_freeze(h, BL, BR, e0, nocc, orbs):
    """Frozen orbitals on DF integrals (``ecg_df_freeze``; ``freeze_core!`` / ``freeze_nvirt!`` after
    the full Fock matrix, ``src/cc/ccdriver.jl:93-98``, as ``freeze_orbs_in_dump`` does on four-index
    integrals, ``src/integrals/dumptools.jl:20-62``).  ``orbs``: 1-based.  Returns
    ``(h, BL, BR, e0, nocc, nvirt)`` of the frozen system (``BR`` None when it was None)."""
    lib = load()
    n, _, naux = BL.shape
    h = fortran(h, (n, n))
    BL = fortran(BL, (n, n, naux))
    BR = None if BR is None else fortran(BR, (n, n, naux))
    arr = np.ascontiguousarray(np.asarray(list(orbs), dtype=np.int32).reshape(-1))
    optr = arr.ctypes.data_as(_ip) if arr.size else None
    nk, nc, e = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    args = (int(n), int(nocc), int(naux), _ptr(h), _ptr(BL), _ptr(BR) if BR is not None else None, float(e0), optr,
            int(arr.size), ctypes.byref(nk), ctypes.byref(nc))
    if lib.ecg_df_freeze(*args, None, None, None, None) != 0:
        raise ElemcoError(f"ecg_df_freeze failed: {lib.ecg_last_error(None).decode()}")
    m = nk.value
    ho = np.zeros((m, m), order="F")
    BLo = np.zeros((m, m, naux), order="F")
    BRo = None if BR is None else np.zeros((m, m, naux), order="F")
    if lib.ecg_df_freeze(*args, _ptr(ho), _ptr(BLo), _ptr(BRo) if BRo is not None else None, ctypes.byref(e)) != 0:
        raise ElemcoError(f"ecg_df_freeze failed: {lib.ecg_last_error(None).decode()}")
    no = int(nocc) - nc.value
    return ho, BLo, BRo, e.value, no, m - no


class Fcidump:
    """A closed-shell FCIDUMP read by the library's host-side reader (``ecg_fcidump_read``:
    ``read_fcidump``, ``src/integrals/dump.jl:326-346``; text or NPY2/NPY1 integrals)."""

    def __init__(self, path):
        lib = load()
        handle = _vp()
        if lib.ecg_fcidump_read(os.fsencode(path), ctypes.byref(handle)) != 0:
            raise ElemcoError(f"ecg_fcidump_read failed: {lib.ecg_last_error(None).decode()}")
        self._h, self._lib = handle, lib
        self._refresh()

    @classmethod
    def from_integrals(cls, h, eri, nelec, ms2=0, st=0, e0=0.0):
        """In-memory dump from arrays (``ecg_fcidump_from_integrals``: the reference's ``FDump``
        constructor, ``src/integrals/dump.jl:107-128``); ``eri[p,q,r,s] = (pq|rs)``."""
        lib = load()
        n = h.shape[0]
        h = fortran(h, (n, n))
        eri = fortran(eri, (n, n, n, n))
        handle = _vp()
        if lib.ecg_fcidump_from_integrals(int(n), int(nelec), int(ms2), int(st), _ptr(h), _ptr(eri), float(e0),
                                          ctypes.byref(handle)) != 0:
            raise ElemcoError(f"ecg_fcidump_from_integrals failed: {lib.ecg_last_error(None).decode()}")
        self = cls.__new__(cls)
        self._h, self._lib = handle, lib
        self._refresh()
        return self

    def _refresh(self):
        lib, handle = self._lib, self._h
        norb, nelec, ms2, st = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        e0 = ctypes.c_double()
        self._chk(lib.ecg_fcidump_info(handle, ctypes.byref(norb), ctypes.byref(nelec), ctypes.byref(ms2),
                                       ctypes.byref(st), ctypes.byref(e0)))
        self.norb, self.nelec, self.ms2, self.st, self.int0 = norb.value, nelec.value, ms2.value, st.value, e0.value
        self.head = {}
        buf = ctypes.create_string_buffer(1 << 16)
        typ = ctypes.create_string_buffer(2)
        k = 0
        while lib.ecg_fcidump_header_key(handle, k, buf, len(buf)) == 0:
            key = buf.value.decode()
            self._chk(lib.ecg_fcidump_header(handle, key.encode(), buf, len(buf), typ))
            conv = {b"i": int, b"f": float}.get(typ.value, str)
            self.head[key] = [conv(x) for x in buf.value.decode().split(",")] if buf.value else []
            k += 1
        n = self.norb
        self.int1 = np.zeros((n, n), order="F")
        self.eri = np.zeros((n, n, n, n), order="F")
        self._chk(lib.ecg_fcidump_integrals(handle, _ptr(self.int1), _ptr(self.eri)))

    def _chk(self, st):
        if st != 0:
            raise ElemcoError(self._lib.ecg_last_error(None).decode())

    def freeze(self, orbs):
        """``freeze_orbs_in_dump(EC, orbs)`` (``src/integrals/dumptools.jl:20-62``) on this dump
        (``ecg_fcidump_freeze``): 1-based orbitals; occupied ones become a frozen core folded into
        ``int1`` and ``int0``, virtual ones are dropped.  Returns ``(ncore, nfrozen_virtual)``."""
        arr = np.ascontiguousarray(np.asarray(list(orbs), dtype=np.int32).reshape(-1))
        ncore, nvirt = ctypes.c_int(), ctypes.c_int()
        ptr = arr.ctypes.data_as(_ip) if arr.size else None
        self._chk(self._lib.ecg_fcidump_freeze(self._h, ptr, int(arr.size), ctypes.byref(ncore), ctypes.byref(nvirt)))
        self._refresh()
        return ncore.value, nvirt.value

    def write(self, path, tol=1e-12):
        """``write_fcidump(fd, fcidump, tol)`` (``src/integrals/dump.jl:752-945``) via
        ``ecg_fcidump_write``."""
        self._chk(self._lib.ecg_fcidump_write(self._h, os.fsencode(path), float(tol)))

    def df_factors(self, rtol=0.0):
        """Exact DF factors ``(BL, BR)`` (``BR`` None when it equals ``BL``) of the integrals
        (``ecg_fcidump_df``)."""
        naux, has_br = ctypes.c_int(), ctypes.c_int()
        self._chk(self._lib.ecg_fcidump_df(self._h, float(rtol), ctypes.byref(naux), None, None, ctypes.byref(has_br)))
        n = self.norb
        BL = np.zeros((n, n, naux.value), order="F")
        BR = np.zeros((n, n, naux.value), order="F") if has_br.value else None
        self._chk(self._lib.ecg_fcidump_df(self._h, float(rtol), ctypes.byref(naux), _ptr(BL), _ptr(BR),
                                           ctypes.byref(has_br)))
        return BL, BR

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ecg_fcidump_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One library context: integrals resident on one GPU (``ecg_create``)."""

    def __init__(self, nocc, nvirt, naux, h, BL, BR=None, e0=0.0, device=0, workspace_bytes=0):
        lib = load()
        n = nocc + nvirt
        self.nocc, self.nvirt, self.naux, self.norb = nocc, nvirt, naux, n
        h = fortran(h, (n, n))
        BL = fortran(BL, (n, n, naux))
        BR = None if BR is None else fortran(BR, (n, n, naux))
        handle = _vp()
        st = lib.ecg_create(ctypes.byref(handle), nocc, nvirt, naux, _ptr(h), _ptr(BL), _ptr(BR), float(e0),
                            int(device), int(workspace_bytes))
        if st != 0:
            raise ElemcoError(f"ecg_create failed ({st}): {lib.ecg_last_error(None).decode()}")
        self._h = handle
        self._lib = lib

    @classmethod
    def from_fcidump(cls, path, rtol=0.0, device=0, workspace_bytes=0):
        """Context straight from an FCIDUMP file (``ecg_create_from_fcidump``: reader, exact DF
        factors and ``ecg_create`` with nocc = NELEC / 2 in one native call)."""
        lib = load()
        handle = _vp()
        if isinstance(path, Fcidump):   # a dump in memory (e.g. after Fcidump.freeze)
            st = lib.ecg_create_from_fcidump_handle(ctypes.byref(handle), path._h, float(rtol), int(device),
                                                    int(workspace_bytes))
        else:
            st = lib.ecg_create_from_fcidump(ctypes.byref(handle), os.fsencode(path), float(rtol), int(device),
                                             int(workspace_bytes))
        if st != 0:
            raise ElemcoError(f"ecg_create_from_fcidump failed ({st}): {lib.ecg_last_error(None).decode()}")
        self = cls.__new__(cls)
        self._h, self._lib = handle, lib
        o, v, q = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(lib.ecg_dims(handle, ctypes.byref(o), ctypes.byref(v), ctypes.byref(q)))
        self.nocc, self.nvirt, self.naux, self.norb = o.value, v.value, q.value, o.value + v.value
        return self

    @classmethod
    def from_ao(cls, nocc, C, h_ao, AAP, PQ, tol=1e-8, e0=0.0, device=0, workspace_bytes=0):
        """Context from AO integrals (``ecg_create_from_ao``: DF preparation on the device,
        ``h = C^T h_AO C``)."""
        lib = load()
        C = fortran(C)
        nao, nmo = C.shape
        nfit = PQ.shape[0]
        h_ao = fortran(h_ao, (nao, nao))
        AAP = fortran(AAP, (nao, nao, nfit))
        PQ = fortran(PQ, (nfit, nfit))
        handle = _vp()
        st = lib.ecg_create_from_ao(ctypes.byref(handle), int(nocc), _ptr(C), nao, nmo, _ptr(h_ao), _ptr(AAP), _ptr(PQ),
                                    nfit, float(tol), float(e0), int(device), int(workspace_bytes))
        if st != 0:
            raise ElemcoError(f"ecg_create_from_ao failed ({st}): {lib.ecg_last_error(None).decode()}")
        self = cls.__new__(cls)
        self._h, self._lib = handle, lib
        o, v, q = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(lib.ecg_dims(handle, ctypes.byref(o), ctypes.byref(v), ctypes.byref(q)))
        self.nocc, self.nvirt, self.naux, self.norb = o.value, v.value, q.value, o.value + v.value
        return self

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ecg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st != 0:
            raise ElemcoError(self._lib.ecg_last_error(self._h).decode())

    # ---- shapes
    def zeros1(self):
        return np.zeros((self.nvirt, self.nocc), order="F")

    def zeros2(self):
        return np.zeros((self.nvirt, self.nvirt, self.nocc, self.nocc), order="F")

    def _t1(self, a):
        return fortran(a, (self.nvirt, self.nocc))

    def _t2(self, a):
        return fortran(a, (self.nvirt, self.nvirt, self.nocc, self.nocc))

    # ---- entry points
    def fock(self):
        n = self.norb
        f = np.zeros((n, n), order="F")
        e = np.zeros(n)
        ehf = ctypes.c_double()
        self._chk(self._lib.ecg_fock(self._h, _ptr(f), _ptr(e), ctypes.byref(ehf)))
        return f, e, ehf.value

    def mp2(self):
        T1, T2, e = self.zeros1(), self.zeros2(), np.zeros(3)
        self._chk(self._lib.ecg_mp2(self._h, _ptr(T1), _ptr(T2), _ptr(e)))
        return T1, T2, e

    def cc_residual(self, T1, T2, method=ECG_CCSD, precision=ECG_FP64):
        T1, T2 = self._t1(T1), self._t2(T2)
        R1, R2 = self.zeros1(), self.zeros2()
        self._chk(self._lib.ecg_cc_residual(self._h, method, precision, _ptr(T1), _ptr(T2), _ptr(R1), _ptr(R2)))
        return R1, R2

    def cc_residual_dev(self, dT1, dT2, dR1, dR2, method=ECG_CCSD, precision=ECG_FP64):
        """Device-pointer variant: arguments are integer device addresses."""
        self._chk(self._lib.ecg_cc_residual_dev(self._h, method, precision, dT1, dT2, dR1, dR2))

    def cc_residual_dev_noalloc(self, dT1, dT2, dR1, dR2, method=ECG_CCSD, precision=ECG_FP64):
        """Allocation-free device-pointer variant (FP64; scratch in the library's static
        device array of ECG_NOALLOC_SCRATCH_BYTES per GPU)."""
        self._chk(self._lib.ecg_cc_residual_dev_noalloc(self._h, method, precision, dT1, dT2, dR1, dR2))

    def cc_solve(self, method=ECG_CCSD, precision=ECG_FP64, opts=None):
        opts = opts or default_opts()
        T1, T2, e = self.zeros1(), self.zeros2(), np.zeros(6)
        it, conv = ctypes.c_int(), ctypes.c_int()
        self._chk(self._lib.ecg_cc_solve(self._h, method, precision, ctypes.byref(opts), _ptr(T1), _ptr(T2),
                                         _ptr(e), ctypes.byref(it), ctypes.byref(conv)))
        return e, T1, T2, it.value, bool(conv.value)

    def state_init(self, T1=None, T2=None, opts=None):
        opts = opts or default_opts()
        if T1 is not None:
            T1, T2 = self._t1(T1), self._t2(T2)
        self._chk(self._lib.ecg_cc_state_init(self._h, _ptr(T1), _ptr(T2), ctypes.byref(opts)))

    def state_step(self, method=ECG_CCSD, precision=ECG_FP64, opts=None):
        opts = opts or default_opts()
        info = np.zeros(8)
        self._chk(self._lib.ecg_cc_state_step(self._h, method, precision, ctypes.byref(opts), _ptr(info)))
        return info

    def state_get(self):
        T1, T2 = self.zeros1(), self.zeros2()
        self._chk(self._lib.ecg_cc_state_get(self._h, _ptr(T1), _ptr(T2)))
        return T1, T2

    def pert_triples(self, T1, T2, U1=None, U2=None, precision=ECG_FP64):
        lam = U1 is not None
        T1, T2 = self._t1(T1), self._t2(T2)
        if lam:
            U1, U2 = self._t1(U1), self._t2(U2)
        et, etb = ctypes.c_double(), ctypes.c_double()
        self._chk(self._lib.ecg_pert_triples(self._h, int(lam), precision, _ptr(T1), _ptr(T2), _ptr(U1), _ptr(U2),
                                             ctypes.byref(et), ctypes.byref(etb)))
        return et.value, etb.value

    def pert_triples_range(self, T1, T2, first, count, U1=None, U2=None, precision=ECG_FP64):
        """(T) over triplets first..first+count-1 of the reference loop order (partial sums)."""
        lam = U1 is not None
        T1, T2 = self._t1(T1), self._t2(T2)
        if lam:
            U1, U2 = self._t1(U1), self._t2(U2)
        et, etb = ctypes.c_double(), ctypes.c_double()
        self._chk(self._lib.ecg_pert_triples_range(self._h, int(lam), precision, _ptr(T1), _ptr(T2), _ptr(U1),
                                                   _ptr(U2), int(first), int(count), ctypes.byref(et),
                                                   ctypes.byref(etb)))
        return et.value, etb.value

    def lambda_residual(self, T1, T2, U1, U2, method=ECG_CCSD, precision=ECG_FP64):
        T1, T2, U1, U2 = self._t1(T1), self._t2(T2), self._t1(U1), self._t2(U2)
        R1, R2 = self.zeros1(), self.zeros2()
        self._chk(self._lib.ecg_lambda_residual(self._h, method, precision, _ptr(T1), _ptr(T2), _ptr(U1), _ptr(U2),
                                                _ptr(R1), _ptr(R2)))
        return R1, R2

    def lambda_solve(self, T1, T2, method=ECG_CCSD, opts=None, precision=ECG_FP64):
        opts = opts or default_opts()
        T1, T2 = self._t1(T1), self._t2(T2)
        U1, U2 = self.zeros1(), self.zeros2()
        it, conv = ctypes.c_int(), ctypes.c_int()
        self._chk(self._lib.ecg_lambda_solve(self._h, method, precision, ctypes.byref(opts), _ptr(T1), _ptr(T2),
                                             _ptr(U1), _ptr(U2), ctypes.byref(it), ctypes.byref(conv)))
        return U1, U2, it.value, bool(conv.value)

    def one_rdm(self, T1, T2, U1, U2):
        T1, T2, U1, U2 = self._t1(T1), self._t2(T2), self._t1(U1), self._t2(U2)
        n = self.norb
        D1, dD1 = np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
        self._chk(self._lib.ecg_one_rdm(self._h, _ptr(T1), _ptr(T2), _ptr(U1), _ptr(U2), _ptr(D1), _ptr(dD1)))
        return D1, dD1

    def energy_grad(self, T1, T2, U1, U2, method=ECG_CCSD):
        """``ecg_energy_grad``: (dE/dh [n, n], dE/dB [n, n, naux]) at converged T and U."""
        T1, T2, U1, U2 = self._t1(T1), self._t2(T2), self._t1(U1), self._t2(U2)
        n = self.norb
        gh = np.zeros((n, n), order="F")
        gb = np.zeros((n, n, self.naux), order="F")
        self._chk(self._lib.ecg_energy_grad(self._h, method, _ptr(T1), _ptr(T2), _ptr(U1), _ptr(U2), _ptr(gh),
                                            _ptr(gb)))
        return gh, gb

    def energy_grad_lr(self, T1, T2, U1, U2, method=ECG_CCSD):
        """``ecg_energy_grad_lr`` (two-sided factors): (dE/dh, dE/dBL, dE/dBR)."""
        T1, T2, U1, U2 = self._t1(T1), self._t2(T2), self._t1(U1), self._t2(U2)
        n = self.norb
        gh = np.zeros((n, n), order="F")
        gl = np.zeros((n, n, self.naux), order="F")
        gr = np.zeros((n, n, self.naux), order="F")
        self._chk(self._lib.ecg_energy_grad_lr(self._h, method, _ptr(T1), _ptr(T2), _ptr(U1), _ptr(U2), _ptr(gh),
                                               _ptr(gl), _ptr(gr)))
        return gh, gl, gr

    def profile(self, enable=True):
        """Record HIP events around every MFMA GEMM launch (``ecg_profile``)."""
        self._chk(self._lib.ecg_profile(self._h, int(enable)))

    def stats(self):
        s = np.zeros(8)
        self._chk(self._lib.ecg_stats(self._h, _ptr(s)))
        return dict(residual_ms=s[0], ladder_ms=s[1], gemm_ms=s[2], gemm_launches=int(s[3]), gemm_flops=s[4],
                    fp64_iterations=int(s[5]), gemm_bytes=s[6], triples_block=int(s[7]))


class UContext:
    """Open-shell context (``ecg_ucreate``): UCCSD / UCCD on one GPU.  Amplitudes in the
    reference's spin blocks (T1a, T1b, T2a, T2b, T2ab)."""

    def __init__(self, nocca, noccb, h, B, e0=0.0, device=0, workspace_bytes=0):
        lib = load()
        n, _, naux = B.shape
        h = fortran(h, (n, n))
        B = fortran(B, (n, n, naux))
        handle = _vp()
        st = lib.ecg_ucreate(ctypes.byref(handle), int(nocca), int(noccb), int(n), int(naux), _ptr(h), _ptr(B),
                             float(e0), int(device), int(workspace_bytes))
        if st != 0:
            raise ElemcoError(f"ecg_ucreate failed ({st}): {lib.ecg_ulast_error(None).decode()}")
        self._h, self._lib = handle, lib
        self._dims()

    @classmethod
    def from_fcidump(cls, path, ms2=-1, rtol=0.0, device=0, workspace_bytes=0):
        """``ecg_ucreate_from_fcidump``: nocca/noccb from NELEC and ms2 (the header's MS2 when ms2 < 0)."""
        lib = load()
        handle = _vp()
        st = lib.ecg_ucreate_from_fcidump(ctypes.byref(handle), os.fsencode(path), int(ms2), float(rtol), int(device),
                                          int(workspace_bytes))
        if st != 0:
            raise ElemcoError(f"ecg_ucreate_from_fcidump failed ({st}): {lib.ecg_ulast_error(None).decode()}")
        self = cls.__new__(cls)
        self._h, self._lib = handle, lib
        self._dims()
        return self

    def _dims(self):
        n, a, b, q = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(self._lib.ecg_udims(self._h, ctypes.byref(n), ctypes.byref(a), ctypes.byref(b), ctypes.byref(q)))
        self.norb, self.nocca, self.noccb, self.naux = n.value, a.value, b.value, q.value
        self.nvirta, self.nvirtb = self.norb - self.nocca, self.norb - self.noccb

    def _chk(self, st):
        if st != 0:
            raise ElemcoError(self._lib.ecg_ulast_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ecg_udestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def zeros(self):
        na, nb, va, vb = self.nocca, self.noccb, self.nvirta, self.nvirtb
        return [np.zeros(s, order="F") for s in ((va, na), (vb, nb), (va, va, na, na), (vb, vb, nb, nb),
                                                 (va, vb, na, nb))]

    def fock(self):
        n = self.norb
        fa, fb = np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
        e = ctypes.c_double()
        self._chk(self._lib.ecg_ufock(self._h, _ptr(fa), _ptr(fb), ctypes.byref(e)))
        return fa, fb, e.value

    def mp2(self):
        T = self.zeros()
        e = ctypes.c_double()
        self._chk(self._lib.ecg_ump2(self._h, *[_ptr(t) for t in T], ctypes.byref(e)))
        return T, e.value

    def residual(self, T, method=ECG_UCCSD):
        T = [fortran(t, z.shape) for t, z in zip(T, self.zeros())]
        R = self.zeros()
        self._chk(self._lib.ecg_ucc_residual(self._h, int(method), *[_ptr(t) for t in T], *[_ptr(r) for r in R]))
        return R

    def solve(self, method=ECG_UCCSD, opts=None):
        """Returns (energies [E_H, E_proj, dE, NormR], amplitudes, iterations, converged)."""
        opts = opts or default_opts()
        T = self.zeros()
        e = np.zeros(4)
        it, conv = ctypes.c_int(), ctypes.c_int()
        self._chk(self._lib.ecg_ucc_solve(self._h, int(method), ctypes.byref(opts), *[_ptr(t) for t in T], _ptr(e),
                                          ctypes.byref(it), ctypes.byref(conv)))
        return e, T, it.value, bool(conv.value)

    def pert_triples(self, T=None):
        """``ecg_upert_triples``: (E(T), E[T]) at the blocks T (or the last solve's amplitudes)."""
        et, etb = ctypes.c_double(), ctypes.c_double()
        ptrs = [None] * 5 if T is None else [_ptr(fortran(t, z.shape)) for t, z in zip(T, self.zeros())]
        self._chk(self._lib.ecg_upert_triples(self._h, *ptrs, ctypes.byref(et), ctypes.byref(etb)))
        return et.value, etb.value


def cc_solve_batch(contexts, method=ECG_CCSD, precision=ECG_FP64, opts=None, max_threads=0):
    """``ecg_cc_solve_batch``: solve independent systems concurrently.

    Returns (energies [nsys, 6], iterations, converged); raises ElemcoError if any
    system failed (the message names the first failing system).
    """
    lib = load()
    opts = opts or default_opts()
    n = len(contexts)
    handles = (_vp * max(1, n))(*[c._h for c in contexts])
    e = np.zeros((n, 6))
    it = np.zeros(n, dtype=np.int32)
    conv = np.zeros(n, dtype=np.int32)
    st = np.zeros(n, dtype=np.int32)
    ip = lambda a: a.ctypes.data_as(_ip)
    rc = lib.ecg_cc_solve_batch(n, handles, method, precision, ctypes.byref(opts), int(max_threads), _ptr(e), ip(it),
                                ip(conv), ip(st))
    if rc != 0:
        k = int(np.nonzero(st)[0][0])
        raise ElemcoError(f"system {k}: {lib.ecg_last_error(contexts[k]._h).decode()}")
    return e, it.astype(int), conv.astype(bool)
