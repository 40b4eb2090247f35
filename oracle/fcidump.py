"""FCIDUMP reader (oracle side), restating ``src/integrals/dump.jl``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference keeps ``<pq|rs>`` (physicists' notation) as ``int2[p,q,tri(r,s)]``
with ``r <= s`` (``dump.jl:100-103``); this restatement keeps the full
four-index array ``int2[p,q,r,s] = <pq|rs>``, which holds the same values.
Symmetries applied when a line ``val i1 i2 i3 i4`` (chemists' ``(i1 i2|i3 i4)``)
is read follow ``set_int2!`` (``dump.jl:472-512``): particle symmetry always,
the remaining permutations only when the header has ``ST=0`` (Hermitian).
"""
from __future__ import annotations

import re

import numpy as np


class FDump:
    """In-memory FCIDUMP: header dict, ``int1[p,q]``, ``int2[p,q,r,s]=<pq|rs>``, ``int0``."""

    def __init__(self, head: dict, int1: np.ndarray, int2: np.ndarray, int0: float):
        self.head = head
        self.int1 = int1
        self.int2 = int2
        self.int0 = int0

    @property
    def norb(self) -> int:
        return int(self.head["NORB"][0])

    @property
    def nelec(self) -> int:
        return int(self.head["NELEC"][0])

    @property
    def simtra(self) -> bool:
        """Similarity-transformed (non-Hermitian) integrals (``dump.jl:331-334``)."""
        return int(self.head.get("ST", [0])[0]) > 0


def read_header(lines):
    """Parse the namelist header (``dump.jl:360-405``); returns (head, index of first data line)."""
    head = {"IUHF": [0], "ST": [0]}
    tokens = []
    iline = 0
    for iline, raw in enumerate(lines):
        line = raw.strip()
        if not line:
            continue
        if line in ("/", "&END"):
            break
        line = line.replace("=", " = ")
        tokens += [t for t in re.split(r"[ ,]", line) if t]
    else:
        raise ValueError("FCIDUMP header not terminated")
    # variable name is the token before '='; its values run up to the next name
    name = None
    vals = []
    i = 0
    while i < len(tokens):
        if i + 1 < len(tokens) and tokens[i + 1] == "=":
            if name is not None:
                head[name] = vals
            name = tokens[i].upper()
            vals = []
            i += 2
            continue
        tok = tokens[i].strip("\"'")
        if name is not None and not tok.startswith("&"):
            for conv in (int, float):
                try:
                    vals.append(conv(tok))
                    break
                except ValueError:
                    pass
            else:
                vals.append(tok)
        i += 1
    if name is not None:
        head[name] = vals
    return head, iline + 1


def _set_int2(int2, i1, i2, i3, i4, val, simtra):
    """Store chemists' (i1 i2|i3 i4) as <i1 i3|i2 i4> with its symmetry images."""
    int2[i1, i3, i2, i4] = val
    int2[i3, i1, i4, i2] = val            # particle symmetry
    if not simtra:
        int2[i1, i4, i2, i3] = val
        int2[i2, i3, i1, i4] = val
        int2[i2, i4, i1, i3] = val
        int2[i4, i1, i3, i2] = val
        int2[i3, i2, i4, i1] = val
        int2[i4, i2, i3, i1] = val


def read_fcidump(path: str) -> FDump:
    """Read a text FCIDUMP (``dump.jl:326-353,553-604``); ``.gz`` files are decompressed."""
    if path.endswith(".gz"):
        import gzip
        with gzip.open(path, "rt") as f:
            lines = f.readlines()
    else:
        with open(path) as f:
            lines = f.readlines()
    head, start = read_header(lines)
    if head.get("IUHF", [0])[0] > 0:
        raise NotImplementedError("unrestricted FCIDUMP is outside the closed-shell path")
    norb = int(head["NORB"][0])
    simtra = int(head.get("ST", [0])[0]) > 0
    int1 = np.zeros((norb, norb))
    int2 = np.zeros((norb, norb, norb, norb))
    int0 = None
    for raw in lines[start:]:
        parts = raw.split()
        if len(parts) != 5:
            continue
        val = float(parts[0].replace("D", "E").replace("d", "e"))
        i1, i2, i3, i4 = (int(x) for x in parts[1:])
        if max(i1, i2, i3, i4) > norb:
            raise ValueError("Index larger than norb: " + raw)
        if i4 > 0:
            _set_int2(int2, i1 - 1, i2 - 1, i3 - 1, i4 - 1, val, simtra)
        elif i2 > 0:
            int1[i1 - 1, i2 - 1] = val
            if not simtra:
                int1[i2 - 1, i1 - 1] = val
        elif i1 <= 0:
            int0 = val
    if int0 is None:
        raise ValueError("No core energy found in fcidump. Incomplete file?")
    return FDump(head, int1, int2, int0)
