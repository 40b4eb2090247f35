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

    def df_factors(self, rtol: float = 0.0):
        """Exact DF factors ``(BL, BR)``, n x n x naux; ``BR`` is None when it equals ``BL``."""
        return self._native.df_factors(rtol)


def read_fcidump(path: str) -> FDump:
    """``read_fcidump(fcidump)`` (``dump.jl:326-346``) for IUHF=0 files (plain or .gz)."""
    return FDump(lib.Fcidump(path))
