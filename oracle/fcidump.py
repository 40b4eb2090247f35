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


def _gen_fock(fd: FDump, occ) -> np.ndarray:
    """``gen_fock`` (``src/scf/fockfactory.jl:27-30``) over the occupied orbitals ``occ``
    (0-based): h[p,q] + sum_i 2 <pi|qi> - <pi|iq>."""
    occ = list(occ)
    g = fd.int2
    f = fd.int1.copy()
    for i in occ:
        f += 2.0 * g[:, i, :, i] - g[:, i, i, :]
    return f


def freeze_orbs_in_dump(fd: FDump, freeze_orbs) -> FDump:
    """``freeze_orbs_in_dump`` closed-shell branch (``src/integrals/dumptools.jl:20-62``),
    in place.  ``freeze_orbs``: 1-based orbital numbers; the occupied space is the lowest
    NELEC/2 orbitals (``setup_space_fd!``, ``ecinfos.jl:171-192``, default occupations)."""
    norb, nelec = fd.norb, fd.nelec
    nocc = nelec // 2
    occ = list(range(nocc))
    full_fock = _gen_fock(fd, occ)
    fr = {o - 1 for o in freeze_orbs}
    freeze_occ = [i for i in occ if i in fr]                       # intersect(SP['o'], freeze_orbs)
    freeze_virt = [a for a in range(nocc, norb) if a in fr]        # intersect(SP['v'], freeze_orbs)
    nonfrozen = [p for p in range(norb) if p not in freeze_occ and p not in freeze_virt]
    core_fock = full_fock - _gen_fock(fd, [i for i in occ if i not in freeze_occ])
    fd.int0 += 2 * sum(fd.int1[c, c] for c in freeze_occ) + sum(core_fock[c, c] for c in freeze_occ)
    ix = np.ix_(nonfrozen, nonfrozen)
    fd.int1 = fd.int1[ix] + core_fock[ix]
    fd.int2 = fd.int2[np.ix_(nonfrozen, nonfrozen, nonfrozen, nonfrozen)]
    m = len(nonfrozen)
    fd.head["NORB"] = [m]                                           # modify_header! (dump.jl:182-196)
    fd.head["NELEC"] = [nelec - 2 * len(freeze_occ)]
    fd.head["ORBSYM"] = [1] * m
    return fd


def write_fcidump(fd: FDump, path: str, tol: float = 1e-12) -> None:
    """``write_fcidump`` (``src/integrals/dump.jl:752-945``) for IUHF=0: header, two-electron
    lines in the reference's loop order (``write_integrals2_normal`` / ``_simtra``), one-electron
    lines (``write_integrals1``), core energy.  The value written on line "p q r s" is
    <pr|qs> = (pq|rs).  Header keys after NORB, NELEC, MS2, ISYM, ORBSYM go by type
    (int, float, string) and name."""
    keys = ["NORB", "NELEC", "MS2", "ISYM", "ORBSYM"]
    out = ["&FCI"]

    def fmt(v):
        return f'"{v}"' if isinstance(v, str) else (repr(v) if isinstance(v, float) else str(v))

    for k in keys:
        if k in fd.head:
            out.append(f" {k}={','.join(fmt(v) for v in fd.head[k])},")

    def group(vals):
        v = vals[0] if vals else ""
        return 0 if isinstance(v, int) else (1 if isinstance(v, float) else 2)

    rest = sorted((k for k in fd.head if k not in keys), key=lambda k: (group(fd.head[k]), k))
    for k in rest:
        vals = fd.head[k]
        if k in ("IUHF", "ST", "III") and vals and vals[0] == 0:
            continue
        out.append(f" {k}={','.join(fmt(v) for v in vals)},")
    out.append("/")
    n = fd.norb
    g = fd.int2

    def line(v, p, q, r, s):
        out.append("%23.15e %3i %3i %3i %3i" % (v, p, q, r, s))

    if fd.simtra:
        for p in range(1, n + 1):
            for q in range(1, n + 1):
                for r in range(1, p):
                    for s in range(1, n + 1):
                        v = g[p - 1, r - 1, q - 1, s - 1]
                        if abs(v) > tol:
                            line(v, p, q, r, s)
                for s in range(1, q + 1):
                    v = g[p - 1, p - 1, q - 1, s - 1]
                    if abs(v) > tol:
                        line(v, p, q, p, s)
        pairs1 = [(p, q) for p in range(1, n + 1) for q in range(1, n + 1)]
    else:
        for p in range(1, n + 1):
            for q in range(1, p + 1):
                for r in range(1, p + 1):
                    for s in range(1, r + 1):
                        if r * (r - 1) // 2 + s <= p * (p - 1) // 2 + q:
                            v = g[p - 1, r - 1, q - 1, s - 1]
                            if abs(v) > tol:
                                line(v, p, q, r, s)
        pairs1 = [(p, q) for p in range(1, n + 1) for q in range(1, p + 1)]
    for p, q in pairs1:
        v = fd.int1[p - 1, q - 1]
        if abs(v) > tol:
            line(v, p, q, 0, 0)
    line(fd.int0, 0, 0, 0, 0)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
