"""Open-shell (unrestricted) CCSD oracle.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference's UCCSD (``calc_cc_resid(EC, T1a, T1b, T2a, T2b, T2ab; dc=false)``,
``src/cc/cc.jl:1496-1934``) works on spin blocks: amplitudes ``T1a[a,i]``,
``T1b[A,I]``, ``T2a[a,b,i,j]``, ``T2b[A,B,I,J]`` (both antisymmetric) and
``T2ab[a,B,i,J]``, with dressed integral files per spin case.  This oracle states
the same equations in the spin-orbital form of Stanton and Gauss (J. Chem. Phys.
94, 4334 (1991); as written out by Crawford and Schaefer, Rev. Comput. Chem. 14,
33 (2000), with the full Fock matrix so non-canonical orbitals are allowed) on
spin orbitals ordered [occupied alpha, occupied beta, virtual alpha, virtual
beta].  The residual's alpha-alpha, beta-beta and alpha-beta blocks are the
reference's ``R2a``, ``R2b``, ``R2ab`` (``R1a``, ``R1b`` likewise), with the
sign convention of the closed-shell residual (``R = 0`` at convergence, the
update ``T -= R / (e_a - e_i + shift)`` of ``update_singles`` /
``update_doubles``, ``src/cc/cctools.jl:198-297``).

Around it, restated from the reference:
  * spin Fock matrices ``gen_fock(EC, :alpha/:beta)`` and the open-shell HF energy
    (``calc_fock_matrix`` / ``calc_HF_energy``, ``src/cc/cctools.jl:35-76``);
  * the UMP2 guess and energy (``calc_UMP2``, ``src/cc/cc.jl:1034-1060``);
  * doubles, singles and Hylleraas energies per spin block (``cc.jl:123-154``,
    ``:176-193``, ``:251-308``);
  * the iterations of ``cc_iterations!`` (``cc.jl:2518-2648``) with the reference's
    DIIS (``oracle.ccsd.Diis``) on plain dots.

Pins (tests/test_oracle.py): the closed-shell limit (nocca = noccb) equals the
closed-shell oracle's CCSD and DCSD residuals block by block for random amplitudes;
the reference's open-shell energies of ``test/h2o_cation.jl`` (HF, MP2, UCCSD, UCCD,
UDCSD) and
``test/h2o_triplet.jl`` (HF, MP2, UCCSD); two same-spin electrons, where UCCSD
is full CI of the Sz = 1 space (closed form: a dense eigenproblem).
"""
from __future__ import annotations

import numpy as np

from .ccsd import Diis


def _es(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


class USystem:
    """Open-shell system on spin orbitals.

    ``h`` (n x n), ``eri[p,q,r,s] = (pq|rs)`` (chemists' notation), ``e0``; ``nocca``
    alpha and ``noccb`` beta electrons in the lowest orbitals (the reference's default
    spaces ``o``/``O``, ``setup_space_fd!``, ``src/infos/ecinfos.jl:171-192``).
    """

    def __init__(self, h, eri, e0, nocca, noccb):
        n = h.shape[0]
        self.n, self.na, self.nb, self.e0 = n, nocca, noccb, float(e0)
        occ = [(i, 0) for i in range(nocca)] + [(i, 1) for i in range(noccb)]
        vir = [(a, 0) for a in range(nocca, n)] + [(a, 1) for a in range(noccb, n)]
        so = occ + vir
        self.O, self.V = len(occ), len(vir)
        sp = np.array([p for p, _ in so])
        sg = np.array([s for _, s in so])
        same = (sg[:, None] == sg[None, :]).astype(float)
        self.h = h[np.ix_(sp, sp)] * same
        # <pq|rs> = (pr|qs) delta(s_p s_r) delta(s_q s_s);  A = <pq||rs> = <pq|rs> - <pq|sr>
        g = eri[np.ix_(sp, sp, sp, sp)].transpose(0, 2, 1, 3)
        g = g * same[:, None, :, None] * same[None, :, None, :]
        self.A = g - g.transpose(0, 1, 3, 2)
        self.g = g
        self.o = slice(0, self.O)
        self.v = slice(self.O, self.O + self.V)
        # gen_fock(EC, :alpha / :beta): f = h + sum_m <pm||qm> over occupied spin orbitals
        self.f = self.h + _es('pmqm->pq', self.A[:, self.o, :, self.o])
        self.eps = np.diag(self.f).copy()

    def blk(self, s):
        return {'o': self.o, 'v': self.v}[s]

    def ints(self, spaces):
        return self.A[tuple(self.blk(c) for c in spaces)]

    # ---- spin blocks of the reference <-> spin-orbital arrays
    def to_blocks(self, T1, T2):
        """(T1a, T1b, T2a, T2b, T2ab) in the reference's layout from spin-orbital T1[a,i], T2[a,b,i,j]."""
        na, nb, va = self.na, self.nb, self.n - self.na
        T1a, T1b = T1[:va, :na], T1[va:, na:]
        return (T1a.copy(), T1b.copy(), T2[:va, :va, :na, :na].copy(), T2[va:, va:, na:, na:].copy(),
                T2[:va, va:, :na, na:].copy())

    def from_blocks(self, T1a, T1b, T2a, T2b, T2ab):
        na, nb, va = self.na, self.nb, self.n - self.na
        T1 = np.zeros((self.V, self.O))
        T1[:va, :na], T1[va:, na:] = T1a, T1b
        T2 = np.zeros((self.V, self.V, self.O, self.O))
        T2[:va, :va, :na, :na] = T2a
        T2[va:, va:, na:, na:] = T2b
        T2[:va, va:, :na, na:] = T2ab
        T2[va:, :va, :na, na:] = -T2ab.transpose(1, 0, 2, 3)
        T2[:va, va:, na:, :na] = -T2ab.transpose(0, 1, 3, 2)
        T2[va:, :va, na:, :na] = T2ab.transpose(1, 0, 3, 2)
        return T1, T2


def hf_energy(S: USystem) -> float:
    """``calc_HF_energy`` open-shell branch (``cctools.jl:71-74``):
    0.5 (sum eps_o^a + sum eps_o^b + sum h_oo^a + sum h_oo^b) + E0."""
    return 0.5 * (float(np.sum(S.eps[S.o])) + float(np.trace(S.h[S.o, S.o]))) + S.e0


def calc_UMP2(S: USystem):
    """``calc_UMP2`` (``cc.jl:1034-1060``): T2 = -<ab||ij>/D (no shift), T1 = -f_vo/D; returns
    (T1, T2, E) with E the doubles plus the Fock singles energy."""
    eo, ev = S.eps[S.o], S.eps[S.v]
    D2 = ev[:, None, None, None] + ev[None, :, None, None] - eo[None, None, :, None] - eo[None, None, None, :]
    T2 = -S.ints('vvoo') / D2
    T1 = -S.f[S.v, S.o] / (ev[:, None] - eo[None, :])
    E = calc_doubles_energy(S, T2) + calc_singles_energy(S, T1, fock_only=True)
    return T1, T2, E


def calc_doubles_energy(S: USystem, T2) -> float:
    """``calc_doubles_energy(EC, T2a, T2b, T2ab)`` (``cc.jl:176-193``): 1/2 T2a <ij|ab> + 1/2 T2b <IJ|AB>
    + T2ab <iJ|aB>, i.e. 1/4 sum over spin orbitals of T2 <ij||ab>."""
    return 0.25 * float(np.sum(T2 * S.ints('oovv').transpose(2, 3, 0, 1)))


def calc_singles_energy(S: USystem, T1, fock_only=False) -> float:
    """``calc_singles_energy(EC, T1a, T1b)`` (``cc.jl:123-154``)."""
    E = float(np.sum(T1 * S.f[S.o, S.v].T))
    if not fock_only:
        E += 0.5 * float(_es('ai,bj,ijab->', T1, T1, S.ints('oovv')))
    return E


def calc_hylleraas(S: USystem, T1, T2, R1, R2) -> float:
    """``calc_hylleraas`` of the spin blocks (``cc.jl:251-308``): projected energy plus
    1/4 T2.R2 (1 for the alpha-beta block) and R1.T1."""
    return calc_doubles_energy(S, T2) + 0.25 * float(np.sum(T2 * R2)) + calc_singles_energy(S, T1) + \
        float(np.sum(R1 * T1))


def _p(X, axes):
    """Antisymmetriser P(pq) X = X - X with the two axes swapped."""
    perm = list(range(X.ndim))
    perm[axes[0]], perm[axes[1]] = perm[axes[1]], perm[axes[0]]
    return X - X.transpose(perm)


def calc_ucc_resid(S: USystem, T1, T2, dc=False):
    """UCCSD (dc = False) or UDCSD (dc = True) residual (R1[a,i], R2[a,b,i,j]) on spin orbitals
    (the equations of the reference's ``calc_cc_resid(EC, T1a, T1b, T2a, T2b, T2ab; dc)``,
    ``cc.jl:1496-1934``).

    The reference writes both in T1-dressed form: the doubles residual is the dressed linear part
    plus terms quadratic in T2 with the undressed <kl|cd>, and the distinguishable-cluster
    variant changes only those quadratic terms (``cc.jl:1724-1808``): the hole-hole term
    (``x_klij``, ``x_KLIJ``, ``x_kLiJ`` += oovv T2) is dropped, the quadratic parts of
    ``x_ij`` / ``x_ab`` are halved (``dcfac``), and the ring terms keep only the integral
    <kl|cd> whose pairs (k,c), (l,d) each sit in one amplitude (the exchange parts
    ``oovv[k,l,d,c]`` and the crossed ``x_DBik`` term are dropped).  Since the singles and the
    linear part are the CCSD ones, R_DC = R_CCSD - dQ with dQ the difference of those
    quadratic terms, written here in spin orbitals (``_dc_correction``)."""
    o, v = S.o, S.v
    f = S.f
    foo, fvv, fov, fvo = f[o, o], f[v, v], f[o, v], f[v, o]
    oovv = S.ints('oovv')          # <mn||ef> [m,n,e,f]
    ovvv = S.ints('ovvv')          # <ma||fe> [m,a,f,e]
    ooov = S.ints('ooov')          # <mn||ie> [m,n,i,e]
    ovvo = S.ints('ovvo')          # <mb||ej> [m,b,e,j]
    oooo = S.ints('oooo')
    vvvv = S.ints('vvvv')
    vvvo = S.ints('vvvo')          # <ab||ej>
    ovoo = S.ints('ovoo')          # <mb||ij>
    vovo = S.ints('vovo')          # <na||if> as [a,n,f,i] via <an||fi>
    tt = _es('ai,bj->abij', T1, T1)
    tau_t = T2 + 0.5 * (tt - tt.transpose(1, 0, 2, 3))
    tau = T2 + tt - tt.transpose(1, 0, 2, 3)
    Fae = fvv - 0.5 * _es('me,am->ae', fov, T1) + _es('fm,mafe->ae', T1, ovvv) \
        - 0.5 * _es('afmn,mnef->ae', tau_t, oovv)
    Fmi = foo + 0.5 * _es('ei,me->mi', T1, fov) + _es('en,mnie->mi', T1, ooov) \
        + 0.5 * _es('efin,mnef->mi', tau_t, oovv)
    Fme = fov + _es('fn,mnef->me', T1, oovv)
    Wmnij = oooo + _p(_es('ej,mnie->mnij', T1, ooov), (2, 3)) + 0.25 * _es('efij,mnef->mnij', tau, oovv)
    Wabef = vvvv - _p(_es('bm,amef->abef', T1, S.ints('vovv')), (0, 1)) \
        + 0.25 * _es('abmn,mnef->abef', tau, oovv)
    Wmbej = ovvo + _es('fj,mbef->mbej', T1, ovvv) - _es('bn,mnej->mbej', T1, S.ints('oovo')) \
        - _es('fbjn,mnef->mbej', 0.5 * T2 + _es('fj,bn->fbjn', T1, T1), oovv)
    # singles
    R1 = fvo + _es('ei,ae->ai', T1, Fae) - _es('am,mi->ai', T1, Fmi) + _es('aeim,me->ai', T2, Fme) \
        - _es('fn,anfi->ai', T1, vovo) - 0.5 * _es('efim,maef->ai', T2, ovvv) \
        - 0.5 * _es('aemn,nmei->ai', T2, S.ints('oovo'))
    # doubles
    R2 = S.ints('vvoo').copy()
    R2 += _p(_es('aeij,be->abij', T2, Fae - 0.5 * _es('bm,me->be', T1, Fme)), (0, 1))
    R2 -= _p(_es('abim,mj->abij', T2, Fmi + 0.5 * _es('ej,me->mj', T1, Fme)), (2, 3))
    R2 += 0.5 * _es('abmn,mnij->abij', tau, Wmnij)
    R2 += 0.5 * _es('efij,abef->abij', tau, Wabef)
    X = _es('aeim,mbej->abij', T2, Wmbej) - _es('ei,am,mbej->abij', T1, T1, ovvo)
    R2 += _p(_p(X, (0, 1)), (2, 3))
    R2 += _p(_es('ei,abej->abij', T1, vvvo), (2, 3))
    R2 -= _p(_es('am,mbij->abij', T1, ovoo), (0, 1))
    if dc:
        R2 -= _dc_correction(S, T2)
    return R1, R2


def _dc_correction(S: USystem, T2):
    """CCSD minus DCSD quadratic doubles terms (T2 only, undressed <mn||ef> = <mn|ef> - <mn|fe>):
    1/4 t_mn^ab t_ij^ef <mn||ef>  (hole-hole, dropped)
    - 1/4 P(ab) t_ij^ae t_mn^bf <mn||ef> - 1/4 P(ij) t_im^ab t_jn^ef <mn||ef>  (x_ab, x_ij halved)
    - 1/2 P(ij) P(ab) t_im^ae t_jn^bf <mn|fe>  (exchange ring, dropped)."""
    o, v = S.o, S.v
    oovv = S.ints('oovv')
    ex = S.g[o, o, v, v].transpose(0, 1, 3, 2)      # <mn|fe> as [m,n,e,f]
    dQ = 0.25 * _es('abmn,efij,mnef->abij', T2, T2, oovv)
    dQ -= 0.25 * _p(_es('aeij,bfmn,mnef->abij', T2, T2, oovv), (0, 1))
    dQ -= 0.25 * _p(_es('abim,efjn,mnef->abij', T2, T2, oovv), (2, 3))
    dQ -= 0.5 * _p(_p(_es('aeim,bfjn,mnef->abij', T2, T2, ex), (0, 1)), (2, 3))
    return dQ


def update(S: USystem, R1, R2, shifts, shiftp):
    """``update_singles`` / ``update_doubles`` (``cctools.jl:198-297``): -R/(e_a - e_i + shift)."""
    eo, ev = S.eps[S.o], S.eps[S.v]
    d1 = ev[:, None] - eo[None, :] + shifts
    d2 = ev[:, None, None, None] + ev[None, :, None, None] - eo[None, None, :, None] - eo[None, None, None, :] + shiftp
    return -R1 / d1, -R2 / d2


class _Opts:
    def __init__(self, thr, maxit):
        self.options = dict(thr=thr, conven=0.1, maxit=maxit, shifts=0.15, shiftp=0.2, maxdiis=6, resthr=10.0)


def calc_ucc(S: USystem, thr=1e-10, maxit=100, singles=True, dc=False):
    """``cc_iterations!`` (``cc.jl:2518-2648``) for UCCSD (UCCD with ``singles=False``: T1 = 0,
    ``exclevel[1] = :none``; UDCSD with ``dc=True``) from the UMP2 guess: residual, Hylleraas energy, update with
    shifts, DIIS, convergence on the squared residual norm (``calc_singles_norm`` /
    ``calc_doubles_norm`` of the spin blocks) and the energy change.  Returns
    (E_H, T1, T2, converged, iterations)."""
    opts = _Opts(thr, maxit)
    T1, T2, _ = calc_UMP2(S)
    if not singles:
        T1[...] = 0.0
    diis = Diis(opts)
    # the reference's unrestricted DIIS dots (calc_u_singles_dot, calc_samespin_doubles_dot,
    # calc_ab_doubles_dot, cctools.jl:453-540): sum T1 T1' and 1/4 sum T2 T2' on spin orbitals
    dots = (lambda x, y: float(np.sum(x * y)), lambda x, y: 0.25 * float(np.sum(x * y)))
    thren = np.sqrt(thr) * 0.1
    Eh, conv, it = 0.0, False, 0
    for it in range(1, maxit + 1):
        R1, R2 = calc_ucc_resid(S, T1, T2, dc=dc)
        if not singles:
            R1[...] = 0.0
        Eh = calc_hylleraas(S, T1, T2, R1, R2)
        d1, d2 = update(S, R1, R2, 0.15, 0.2)
        T1 += d1
        T2 += d2
        diis.perform([T1, T2], [R1, R2], dots)
        En = calc_doubles_energy(S, T2) + calc_singles_energy(S, T1)
        NormR = float(np.sum(R1 * R1)) + 0.25 * float(np.sum(R2 * R2))
        if NormR < thr and abs(En - Eh) < thren:
            conv = True
            break
    return Eh, T1, T2, conv, it


def ucc_driver(h, eri, e0, nocca, noccb, thr=1e-10, method="uccsd"):
    """Energies with the reference's keys for ``@cc uccsd`` / ``@cc uccd`` / ``@cc udcsd``
    (``ccdriver``, ``src/cc/ccdriver.jl:37-76``; ``dcsd`` on an open-shell system is UDCSD,
    ``checkset_unrestricted_closedshell!``, ``ccdriver.jl:202-214``): HF, MP2 (total), UCCSDc
    (correlation) and UCCSD (total)."""
    S = USystem(h, eri, e0, nocca, noccb)
    EHF = hf_energy(S)
    _, _, Emp2 = calc_UMP2(S)
    name = method.upper()
    E, T1, T2, conv, it = calc_ucc(S, thr=thr, singles=(method != "uccd"), dc=(method == "udcsd"))
    return {'HF': EHF, 'MP2c': Emp2, 'MP2': EHF + Emp2, name + 'c': E, name: EHF + E, 'converged': conv,
            'iterations': it}


def _pijk(X):
    """P(i/jk) on the last three axes of X[a,b,c,i,j,k]: X - X(i<->j) - X(i<->k)."""
    return X - X.transpose(0, 1, 2, 4, 3, 5) - X.transpose(0, 1, 2, 5, 4, 3)


def _pabc(X):
    """P(a/bc) on the first three axes: X - X(a<->b) - X(a<->c)."""
    return X - X.transpose(1, 0, 2, 3, 4, 5) - X.transpose(2, 1, 0, 3, 4, 5)


def calc_pertT_unrestricted(S: USystem, T1, T2):
    """(T) of UCCSD, ``calc_pertT_unrestricted`` (``src/cc/cc_triples.jl:297-320``) with its same-spin
    and mixed-spin parts (``calc_pertT_samespin``, ``calc_pertT_mixedspin``, ``:357-640``) in
    spin-orbital form: connected triples
        D W_ijk^abc = P(i/jk) P(a/bc) [sum_e t_jk^ae <ei||bc> - sum_m t_im^bc <ma||jk>],
    D = e_i + e_j + e_k - e_a - e_b - e_c (diagonal Fock, as ``orbital_energies``),
    E[T] = 1/36 sum W D W, and the singles and Fock parts of the reference's IntX / IntY terms
    E(T) - E[T] = 1/36 sum W P(i/jk) P(a/bc) (t_i^a <jk||bc> + f_ia t_jk^bc).
    Returns (E(T), E[T])."""
    o, v = S.o, S.v
    eo, ev = S.eps[o], S.eps[v]
    vvvo = S.ints('vvvo')                 # <ab||ej>; <ei||bc> = <bc||ei>
    ovoo = S.ints('ovoo')                 # <ma||jk>
    oovv = S.ints('oovv')                 # <jk||bc>
    Z = _es('aejk,bcei->abcijk', T2, vvvo) - _es('bcim,majk->abcijk', T2, ovoo)
    W = _pabc(_pijk(Z))
    D = (eo[None, None, None, :, None, None] + eo[None, None, None, None, :, None] + eo[None, None, None, None, None, :]
         - ev[:, None, None, None, None, None] - ev[None, :, None, None, None, None] - ev[None, None, :, None, None, None])
    X = W / D
    Etb = float(np.sum(W * X)) / 36.0
    fov = S.f[o, v]
    Y = _es('ai,jkbc->abcijk', T1, oovv) + _es('ia,bcjk->abcijk', fov, T2)
    Ed = float(np.sum(X * _pabc(_pijk(Y)))) / 36.0
    return Etb + Ed, Etb
