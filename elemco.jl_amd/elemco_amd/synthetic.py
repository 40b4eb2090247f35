"""Seeded synthetic density-fitted closed-shell systems (the benchmark inputs).

A system is fully determined by ``(nocc, nvirt, naux, seed)``:
occupied orbital energies ~ U[-1.5, -0.3] Eh and virtual ~ U[0.2, 3.0] Eh
(each block sorted ascending), symmetric DF factors ``B[p,q,Q] ~ N(0, 0.05^2)``
(upper triangle drawn in the order Q, q, p<=q and mirrored), Fock matrix
``diag(eps)`` and one-electron integrals ``h = F - (2J - K)`` over the occupied
orbitals, so the closed-shell Fock build reproduces ``diag(eps)``; ``E0 = 0``.
Random numbers: splitmix64 of a (seed, stream, counter) key, uniform doubles
from the top 53 bits, normals by Box-Muller -- reproducible on any machine.
"""
from __future__ import annotations

import numpy as np

_MASK = (1 << 64) - 1


def _mix(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _draws(seed, stream, count):
    with np.errstate(over="ignore"):
        base = _mix(np.array([(seed * 0x100000001B3 + stream) & _MASK], dtype=np.uint64))[0]
        ctr = np.arange(count, dtype=np.uint64)
        return _mix(base ^ (ctr * np.uint64(0xD1B54A32D192ED03)))


def _uniform(seed, stream, count):
    return (_draws(seed, stream, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _normal(seed, stream, count):
    u = _uniform(seed, stream, 2 * count)
    return np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(2.0 * np.pi * u[1::2])


def conditioned_scale(nocc: int, nvirt: int, naux: int) -> float:
    """Width factor keeping o^2 v^2 Q sigma^4 at its o=5, v=19, Q=72 value (E_MP2 near -2 Eh).

    The systems with the plain width 0.05 have E_MP2 = -2.1e4 Eh at o=20, v=180,
    Q=600 and a divergent CCSD; they serve as timing workloads, the conditioned
    ones for converged energies.
    """
    return (5.0**2 * 19.0**2 * 72.0 / (float(nocc) ** 2 * float(nvirt) ** 2 * float(naux))) ** 0.25


def make_system(nocc: int, nvirt: int, naux: int, seed: int, scale: float = 1.0):
    """Return (eps, h, B, e0) with B of shape (n, n, naux) in column-major order (width 0.05*scale)."""
    n = nocc + nvirt
    eps = np.concatenate([np.sort(-1.5 + 1.2 * _uniform(seed, 1, nocc)),
                          np.sort(0.2 + 2.8 * _uniform(seed, 2, nvirt))])
    ntri = n * (n + 1) // 2
    g = (0.05 * scale) * _normal(seed, 3, ntri * naux).reshape(naux, ntri)
    rows, cols = np.triu_indices(n)
    tri = rows + cols * (cols + 1) // 2
    B = np.empty((n, n, naux), order="F")
    for Q in range(naux):
        m = np.zeros((n, n))
        m[rows, cols] = g[Q, tri]
        B[:, :, Q] = m + np.triu(m, 1).T
    J = np.tensordot(B, np.einsum('iiQ->Q', B[:nocc, :nocc, :]), axes=([2], [0]))
    K = np.tensordot(B[:, :nocc, :], B[:nocc, :, :], axes=([1, 2], [0, 2]))
    h = np.diag(eps) - (2.0 * J - K)
    return eps, np.asfortranarray(h), B, 0.0


BATCH_WIDTH = 0.7   # factor on conditioned_scale for the many-system batch


def batch_sizes(nsys: int = 64, first_seed: int = 100):
    """Sizes of the many-system batch: system k has seed first_seed + k,
    nocc = 4 + floor(13 u) with u the first uniform draw of stream 4 of that seed
    (nocc uniform in 4..16), nvirt = 9 nocc and naux = 3 (nocc + nvirt); the
    factor width is 0.05 * BATCH_WIDTH * conditioned_scale (at the full
    conditioned width some of the o=4 systems have divergent CCSD iterations).
    Returns a list of (nocc, nvirt, naux, seed, scale)."""
    out = []
    for k in range(nsys):
        seed = first_seed + k
        o = 4 + int(13.0 * _uniform(seed, 4, 1)[0])
        v, q = 9 * o, 3 * (o + 9 * o)
        out.append((o, v, q, seed, BATCH_WIDTH * conditioned_scale(o, v, q)))
    return out
