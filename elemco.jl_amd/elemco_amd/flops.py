"""Algorithmic operation counts of the closed-shell path (leading terms).

These are the counts every reported rate divides by, for the GPU and the CPU
baseline alike; o = nocc, v = nvirt, Q = naux.

CCSD iteration:  Q v^4 + o^2 v^4 + 18 o^3 v^3 + 6 o^4 v^2 + 4 o^2 v^3 + 4 o^3 v^2 + 6 Q o^2 v^2 + 2 Q o v^3
  (Q v^4: unique half of the DF-assembled (ac|bd); o^2 v^4: i<=j-packed ladder
  contraction; 18 o^3 v^3: the five quadratic/ring contractions and the
  ladder's occupied-virtual piece; 6 o^4 v^2: hole-hole ladder; the rest are the
  Fock-like terms and the DF block assembly.)
DCSD iteration:  same with 10 o^3 v^3 and 4 o^4 v^2.
(T):             (o(o+1)(o+2)/6 - o) (12 v^4 + 12 o v^3);  Lambda-(T) twice that.
"""


def ccsd_iteration(o, v, Q, dcsd=False):
    c3 = 10 if dcsd else 18
    c4 = 4 if dcsd else 6
    return float(Q * v**4 + o**2 * v**4 + c3 * o**3 * v**3 + c4 * o**4 * v**2 + 4 * o**2 * v**3
                 + 4 * o**3 * v**2 + 6 * Q * o**2 * v**2 + 2 * Q * o * v**3)


def triplet_count(o):
    return o * (o + 1) * (o + 2) // 6 - o


def triples(o, v, lam=False):
    f = triplet_count(o) * (12.0 * v**4 + 12.0 * o * v**3)
    return 2.0 * f if lam else f
