"""FCIDUMP input for the GPU path (format of ElemCo.jl ``src/integrals/dump.jl``).

The file is read by the library's host-side reader (``ecg_fcidump_read``,
``elemco.jl_amd/csrc/ecg_fcidump.cpp``): the namelist header (``read_header``,
``dump.jl:360-405``), then the integrals from the NPY files named by ``NPY2`` /
``NPY1`` (``dump.jl:442-461``) or from the integral lines (``dump.jl:553-604``)
with the permutational symmetries of ``set_int2!`` (``dump.jl:472-512``):
particle symmetry always, the full 8-fold symmetry unless the header has
``ST>0`` (similarity-transformed, non-Hermitian).

``FDump.df_factors`` brings the four-index integrals to the factorised form the
library takes, ``(pq|rs) = sum_Q BL[p,q,Q] BR[r,s,Q]``, exactly: diagonally
pivoted LDL^T of the ``(pq),(rs)`` supermatrix (symmetric by particle
symmetry), ``BL = L sqrt|D|``, ``BR = L sign(D) sqrt|D|`` (``BR = BL``, a
Cholesky factorisation, for ordinary Hermitian integrals).
"""
from __future__ import annotations

from . import lib


class FDump:
    """Closed-shell FDump (``dump.jl:89-131``): ``head`` (typed header values),
    ``int1`` = h[p,q], ``eri`` = (pq|rs) in chemists' notation, ``int0``."""

    def __init__(self, native: lib.Fcidump):
        self._native = native
        self.head = native.head
        self.int1 = native.int1
        self.eri = native.eri
        self.int0 = native.int0

    @property
    def norb(self):
        return self._native.norb

    @property
    def nelec(self):
        if self._native.nelec < 0:
            raise lib.ElemcoError("NELEC option not found in fcidump")
        return self._native.nelec

    @property
    def simtra(self):
        return self._native.st > 0

    def _sync(self):
        n = self._native
        self.head, self.int1, self.eri, self.int0 = n.head, n.int1, n.eri, n.int0

    def freeze_orbs(self, orbs):
        """``freeze_orbs_in_dump(EC, freeze_orbs)`` (``src/integrals/dumptools.jl:20-62``), in
        place: 1-based orbital numbers; returns ``(ncore, nfrozen_virtual)``."""
        out = self._native.freeze(orbs)
        self._sync()
        return out

    def write(self, path: str, tol: float = 1e-12):
        """``write_fcidump(fd, fcidump, tol)`` (``src/integrals/dump.jl:752-945``)."""
        self._native.write(path, tol)

    def df_factors(self, rtol: float = 0.0):
        """Exact DF factors ``(BL, BR)``, n x n x naux; ``BR`` is None when it equals ``BL``."""
        return self._native.df_factors(rtol)


def freeze_orbs_in_dump(fd: FDump, freeze_orbs) -> FDump:
    """``freeze_orbs_in_dump(EC, freeze_orbs)`` (``src/integrals/dumptools.jl:20-62``): ``freeze_orbs``
    is an iterable of 1-based orbital numbers or a string such as ``"1-4+6"`` / ``"1:4;6"``
    (``parse_orbstring``)."""
    if isinstance(freeze_orbs, str):
        freeze_orbs = parse_orbstring(freeze_orbs)
    fd.freeze_orbs(freeze_orbs)
    return fd


def parse_orbstring(s: str):
    """Orbital list from ``"1-4+6-8"`` or ``"1:4;6:8"`` (both ``[1,2,3,4,6,7,8]``; ``"-3"`` is
    ``1-3``): ``parse_orbstring`` (``src/infos/ecinfos.jl:606-660``) without orbital symmetry
    labels (``"orb.sym"`` raises)."""
    t = s.replace(" ", "").replace("-", ":").replace("+", ";")
    if "." in t:
        raise ValueError("orbital strings with symmetry labels are not supported: " + s)
    if not t or any(c not in "0123456789:;" for c in t):
        raise ValueError("Use only `0123456789:;+-` characters in the orbstring: " + s)
    out = []
    for part in t.split(";"):
        if not part:
            continue
        if ":" in part:
            a, b = part.split(":")
            out += list(range(int(a) if a else 1, int(b) + 1))
        else:
            out.append(int(part))
    if len(set(out)) != len(out):
        raise ValueError("Repeated orbitals found in orbstring " + s)
    return sorted(out)


def write_fcidump(fd: FDump, fcidump: str, tol: float = 1e-12):
    """``write_fcidump(fd, fcidump, tol)`` (``src/integrals/dump.jl:752-758``)."""
    fd.write(fcidump, tol)


def read_fcidump(path: str) -> FDump:
    """``read_fcidump(fcidump)`` (``dump.jl:326-346``) for IUHF=0 files (plain or .gz)."""
    return FDump(lib.Fcidump(path))
