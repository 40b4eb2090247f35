"""Seeded synthetic density-fitted closed-shell systems (oracle side).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

A system is fully determined by ``(nocc, nvirt, naux, seed)``:

* orbital energies: occupied ~ U[-1.5, -0.3] Eh, virtual ~ U[0.2, 3.0] Eh,
  each block sorted ascending;
* DF factors ``B[p,q,Q] = B[q,p,Q] ~ N(0, 0.05^2)`` (upper triangle drawn,
  mirrored), so ``(pq|rs) = sum_Q B[p,q,Q] B[r,s,Q]`` has all 8-fold symmetries;
* Fock matrix ``F = diag(eps)`` and one-electron integrals
  ``h[p,q] = F[p,q] - sum_i (2 (pq|ii) - (pi|iq))`` so that the closed-shell
  Fock build of ``src/scf/fockfactory.jl:27-30`` returns ``diag(eps)`` exactly
  (up to rounding); core energy ``E0 = 0``.

The random numbers come from a counter-based generator (splitmix64 of a
(seed, stream, counter) key; Box-Muller for normals) so the same system can be
regenerated bit-for-bit by the product library's own generator
(``elemco.jl_amd/csrc/ecg_synth.cpp``) without shipping gigabytes of input.
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
STREAM_EOCC, STREAM_EVIRT, STREAM_B = 1, 2, 3


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser (Steele, Lea, Flood 2014), vectorised over uint64."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def _keys(seed: int, stream: int, count: int, start: int = 0) -> np.ndarray:
    base = splitmix64(np.array([np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream)],
                               dtype=np.uint64))[0]
    ctr = np.arange(start, start + count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return splitmix64(base ^ (ctr * np.uint64(0xD1B54A32D192ED03)))


def uniform01(seed: int, stream: int, count: int, start: int = 0) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits of each 64-bit draw."""
    k = _keys(seed, stream, count, start)
    return (k >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, count: int) -> np.ndarray:
    """Standard normals by Box-Muller; normal t uses draws 2t and 2t+1."""
    u = uniform01(seed, stream, 2 * count)
    u1 = 1.0 - u[0::2]          # (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


class DFSystem:
    """Closed-shell system given by orbital energies and DF factors B[p,q,Q]."""

    def __init__(self, eps: np.ndarray, B: np.ndarray, h: np.ndarray, nocc: int, e0: float = 0.0):
        self.eps = eps
        self.B = B
        self.h = h
        self.nocc = nocc
        self.norb = eps.shape[0]
        self.nvirt = self.norb - nocc
        self.naux = B.shape[2]
        self.e0 = e0

    @property
    def fock(self) -> np.ndarray:
        return np.diag(self.eps)


def conditioned_scale(nocc: int, nvirt: int, naux: int) -> float:
    """Factor on the 0.05 width that keeps o^2 v^2 Q sigma^4 at its o=5, v=19, Q=72 value.

    With sigma = 0.05 the MP2 energy grows like o^2 v^2 Q sigma^4 (-754 Eh at
    o=10, v=90, Q=300; -2.1e4 Eh at o=20, v=180, Q=600) and CCSD diverges; the
    conditioned systems keep it near the o=5 value (about -2 Eh).
    """
    return (5.0**2 * 19.0**2 * 72.0 / (float(nocc) ** 2 * float(nvirt) ** 2 * float(naux))) ** 0.25


def make_df_system(nocc: int, nvirt: int, naux: int, seed: int, scale: float = 1.0) -> DFSystem:
    """Build the synthetic system of the given sizes and seed (factor width 0.05*scale)."""
    n = nocc + nvirt
    eo = np.sort(-1.5 + 1.2 * uniform01(seed, STREAM_EOCC, nocc))
    ev = np.sort(0.2 + 2.8 * uniform01(seed, STREAM_EVIRT, nvirt))
    eps = np.concatenate([eo, ev])
    ntri = n * (n + 1) // 2
    # draw order: for Q, for q (column), for p <= q  -> flat upper-triangle index p + q(q+1)/2
    g = (0.05 * scale) * normal(seed, STREAM_B, ntri * naux).reshape(naux, ntri)
    iu = np.triu_indices(n)            # row p <= col q
    tri = iu[0] + iu[1] * (iu[1] + 1) // 2
    B = np.empty((n, n, naux))
    for Q in range(naux):
        m = np.zeros((n, n))
        m[iu] = g[Q, tri]
        m = m + np.triu(m, 1).T
        B[:, :, Q] = m
    o = slice(0, nocc)
    # (pq|ii) and (pi|iq) summed over occupied i
    J = np.einsum('pqQ,Q->pq', B, np.einsum('iiQ->Q', B[o, o, :]))
    K = np.einsum('piQ,iqQ->pq', B[:, o, :], B[o, :, :])
    h = np.diag(eps) - (2.0 * J - K)
    return DFSystem(eps, B, h, nocc)


def write_fcidump(path: str, sysm: DFSystem, thresh: float = 0.0) -> None:
    """Write the system as a text FCIDUMP (format of ``src/integrals/dump.jl:752``).

    Two-electron lines are chemist ``(i1 i2|i3 i4)`` with the 8-fold canonical
    ordering i1>=i2, i3>=i4, (i1 i2)>=(i3 i4); one-electron lines ``i1>=i2``;
    the core energy line has all indices zero.
    """
    n = sysm.norb
    B = sysm.B
    eri = np.einsum('pqQ,rsQ->pqrs', B, B)
    with open(path, "w") as f:
        f.write(f" &FCI NORB={n},NELEC={2 * sysm.nocc},MS2=0,\n")
        f.write("  ORBSYM=" + ",".join("1" for _ in range(n)) + ",\n")
        f.write("  ISYM=1,\n /\n")
        for i1 in range(n):
            for i2 in range(i1 + 1):
                ij = i1 * (i1 + 1) // 2 + i2
                for i3 in range(n):
                    for i4 in range(i3 + 1):
                        kl = i3 * (i3 + 1) // 2 + i4
                        if kl > ij:
                            continue
                        val = eri[i1, i2, i3, i4]
                        if abs(val) > thresh:
                            f.write(f" {val: .16E} {i1 + 1:4d} {i2 + 1:4d} {i3 + 1:4d} {i4 + 1:4d}\n")
        for i1 in range(n):
            for i2 in range(i1 + 1):
                val = sysm.h[i1, i2]
                if abs(val) > thresh:
                    f.write(f" {val: .16E} {i1 + 1:4d} {i2 + 1:4d} {0:4d} {0:4d}\n")
        f.write(f" {sysm.e0: .16E} {0:4d} {0:4d} {0:4d} {0:4d}\n")
