"""FCIDUMP input for the GPU path (format of ElemCo.jl ``src/integrals/dump.jl``).

``read_fcidump`` parses the namelist header (``dump.jl:360-405``) and the
integral lines (``dump.jl:553-604``), applying the permutational symmetries of
``set_int2!`` (``dump.jl:472-512``): particle symmetry always, the full 8-fold
symmetry unless the header has ``ST>0`` (similarity-transformed, non-Hermitian).

``df_factors`` brings the four-index integrals to the factorised form the
library takes, ``(pq|rs) = sum_Q BL[p,q,Q] BR[r,s,Q]``, exactly: the
``(pq),(rs)`` supermatrix is symmetric (particle symmetry), so its eigenpairs
give ``BL = V sqrt|w|`` and ``BR = V sign(w) sqrt|w|`` (``BR = BL`` when no
eigenvalue is negative, i.e. for ordinary Hermitian integrals).
"""
from __future__ import annotations

import gzip
import re

import numpy as np


class FDump:
    def __init__(self, head, int1, eri, int0):
        self.head = head
        self.int1 = int1      # h[p,q]
        self.eri = eri        # (pq|rs), chemists' notation
        self.int0 = int0

    @property
    def norb(self):
        return int(self.head["NORB"][0])

    @property
    def nelec(self):
        return int(self.head["NELEC"][0])

    @property
    def simtra(self):
        return int(self.head.get("ST", [0])[0]) > 0


def _parse_header(lines):
    head = {"IUHF": [0], "ST": [0]}
    toks = []
    end = None
    for k, raw in enumerate(lines):
        s = raw.strip()
        if not s:
            continue
        if s in ("/", "&END"):
            end = k
            break
        toks += [t for t in re.split(r"[ ,]", s.replace("=", " = ")) if t]
    if end is None:
        raise ValueError("FCIDUMP header not terminated by '/' or '&END'")
    name, vals = None, []
    i = 0
    while i < len(toks):
        if i + 1 < len(toks) and toks[i + 1] == "=":
            if name is not None:
                head[name] = vals
            name, vals = toks[i].upper(), []
            i += 2
            continue
        t = toks[i].strip("\"'")
        if name is not None and not t.startswith("&"):
            try:
                vals.append(int(t))
            except ValueError:
                try:
                    vals.append(float(t))
                except ValueError:
                    vals.append(t)
        i += 1
    if name is not None:
        head[name] = vals
    return head, end + 1


def read_fcidump(path: str) -> FDump:
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt") as f:
        lines = f.readlines()
    head, start = _parse_header(lines)
    if int(head.get("IUHF", [0])[0]) > 0:
        raise ValueError("unrestricted FCIDUMP: the closed-shell path needs IUHF=0")
    n = int(head["NORB"][0])
    simtra = int(head.get("ST", [0])[0]) > 0
    h = np.zeros((n, n))
    eri = np.zeros((n, n, n, n))
    int0 = None
    for raw in lines[start:]:
        p = raw.split()
        if len(p) != 5:
            continue
        val = float(p[0].replace("D", "E").replace("d", "e"))
        i, j, k, l = (int(x) - 1 for x in p[1:])
        if max(i, j, k, l) >= n:
            raise ValueError("Index larger than norb: " + raw)
        if l >= 0:
            eri[i, j, k, l] = val
            eri[k, l, i, j] = val
            if not simtra:
                eri[j, i, k, l] = eri[i, j, l, k] = eri[j, i, l, k] = val
                eri[l, k, i, j] = eri[k, l, j, i] = eri[l, k, j, i] = val
        elif j >= 0:
            h[i, j] = val
            if not simtra:
                h[j, i] = val
        elif i < 0:
            int0 = val
    if int0 is None:
        raise ValueError("No core energy found in fcidump. Incomplete file?")
    return FDump(head, h, eri, int0)


def df_factors(eri: np.ndarray, hermitian: bool = True, rtol: float = 1e-14):
    """Exact factorisation of (pq|rs); returns (BL, BR or None), each n x n x naux.

    For Hermitian integrals the supermatrix is positive semidefinite; negative
    eigenvalues above -1e-9 max|w| are rounding noise of the printed integrals
    and are dropped, larger ones are an error.
    """
    n = eri.shape[0]
    M = eri.reshape(n * n, n * n)
    asym = np.max(np.abs(M - M.T))
    if asym > 1e-10 * max(1.0, np.max(np.abs(M))):
        raise ValueError("integrals lack particle symmetry (pq|rs) = (rs|pq)")
    w, V = np.linalg.eigh(0.5 * (M + M.T))
    wmax = np.max(np.abs(w))
    if hermitian:
        if np.min(w) < -1e-9 * wmax:
            raise ValueError("Hermitian integrals with a negative (pq|rs) eigenvalue")
        w = np.where(w > 0, w, 0.0)
    keep = np.abs(w) > rtol * wmax
    w, V = w[keep], V[:, keep]
    B = (V * np.sqrt(np.abs(w))).reshape(n, n, -1)
    if np.all(w > 0):
        return B, None
    return B, B * np.sign(w)[None, None, :]
