"""Closed-shell CCSD/DCSD oracle: a numpy restatement of ElemCo.jl's path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Every function names the reference function and lines it restates (paths are
relative to the ElemCo.jl v0.14.0 source tree).  Tensor index order follows the
reference exactly (``T2[a,b,i,j]``, ``int2[p,q,r,s] = <pq|rs>``); only the
storage order differs (numpy row-major instead of Julia column-major), which
does not change any value.  Contractions are written as ``einsum`` with the same
index letters as the reference's ``@mtensor`` lines.
"""
from __future__ import annotations

import numpy as np

from .fcidump import FDump
from .synthetic import DFSystem


def _es(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


# --------------------------------------------------------------------------
# integral access
# --------------------------------------------------------------------------
class FourIndexInts:
    """Integral source holding the full ``<pq|rs>`` array (FCIDUMP back-end)."""

    def __init__(self, int2: np.ndarray):
        self.int2 = int2

    def block(self, s1, s2, s3, s4):
        return np.ascontiguousarray(self.int2[s1, s2, s3, s4])

    def kext(self, D2: np.ndarray) -> np.ndarray:
        """``K[p,r,i,j] = sum_{q,s} <pr|qs> D2[q,s,i,j]`` over all orbitals."""
        return _es('prqs,qsij->prij', self.int2, D2)


class DFInts:
    """Integral source from DF factors: ``<pq|rs> = (pr|qs) = sum_Q B[p,r,Q] B[q,s,Q]``."""

    def __init__(self, B: np.ndarray, batch: int = 32):
        self.B = B
        self.batch = batch

    def block(self, s1, s2, s3, s4):
        return _es('prQ,qsQ->pqrs', self.B[s1, s3, :], self.B[s2, s4, :])

    def kext(self, D2: np.ndarray) -> np.ndarray:
        """``K[p,r,i,j] = sum_{q,s} <pr|qs> D2[q,s,i,j]`` with ``<pr|qs> = sum_Q B[p,q,Q] B[r,s,Q]``.

        Evaluated one p at a time as two matrix products (no four-index
        transposes): ``V_p[(r,s),q] = B[(r,s),:] B[p,q,:]^T`` then
        ``K[p] = V_p[r,(s,q)] D2[q,s,(ij)]``.
        """
        n, nq = self.B.shape[0], self.B.shape[2]
        B2 = np.ascontiguousarray(self.B).reshape(n * n, nq)
        Dt = np.ascontiguousarray(D2.transpose(1, 0, 2, 3)).reshape(n * n, -1)   # [(s,q), ij]
        out = np.empty((n, n) + D2.shape[2:])
        for p in range(n):
            Vp = B2 @ self.B[p].T                                  # [(r,s), q]
            out[p] = (Vp.reshape(n, n * n) @ Dt).reshape((n,) + D2.shape[2:])
        return out


class ECInfo:
    """Minimal counterpart of ``ECInfo`` (``src/infos/ecinfos.jl:103-155``).

    ``files`` stands in for the scratch files the reference writes and re-reads
    (``T_vo``, ``T_vvoo``, ``d_*``, ``df_mm`` ...).  Option defaults are those
    of ``CcOptions``/``DiisOptions`` (``src/infos/options.jl:128-221,272-282``).
    """

    def __init__(self, ints, int1: np.ndarray, int0: float, nocc: int, simtra: bool = False):
        self.ints = ints
        self.int1 = np.array(int1, dtype=np.float64)
        self.int0 = float(int0)
        self.norb = self.int1.shape[0]
        self.nocc = nocc
        self.nvirt = self.norb - nocc
        self.simtra = simtra
        self.o = slice(0, nocc)
        self.v = slice(nocc, self.norb)
        self.m = slice(0, self.norb)
        self.files = {}
        self.options = dict(thr=1e-10, conven=0.1, maxit=50, shifts=0.15, shiftp=0.2,
                            maxdiis=6, resthr=10.0, use_kext=True, calc_d_vovv=False)
        self.verbose = False

    @classmethod
    def from_fcidump(cls, fd: FDump):
        return cls(FourIndexInts(fd.int2), fd.int1, fd.int0, fd.nelec // 2, fd.simtra)

    @classmethod
    def from_df(cls, sysm: DFSystem):
        return cls(DFInts(sysm.B), sysm.h, sysm.e0, sysm.nocc)

    def sp(self, c):
        return {'o': self.o, 'v': self.v, 'm': self.m, ':': self.m}[c]

    def ints2(self, spaces: str) -> np.ndarray:
        """``ints2(EC, spaces)`` (``src/tools/tensortools.jl:295-306``): block ``<s1 s2|s3 s4>``."""
        return self.ints.block(*(self.sp(c) for c in spaces))

    def ints1(self, spaces: str) -> np.ndarray:
        """``ints1`` (``src/tools/tensortools.jl:237-248``)."""
        return self.int1[self.sp(spaces[0]), self.sp(spaces[1])].copy()


# --------------------------------------------------------------------------
# Fock matrix, HF energy, orbital energies
# --------------------------------------------------------------------------
def gen_fock(EC: ECInfo) -> np.ndarray:
    """``gen_fock`` (``src/scf/fockfactory.jl:27-30``)."""
    pi_qi = EC.ints2(':o:o')
    pi_iq = EC.ints2(':oo:')
    return EC.int1 + 2.0 * _es('piqi->pq', pi_qi) - _es('piiq->pq', pi_iq)


def calc_fock_matrix(EC: ECInfo) -> None:
    """``calc_fock_matrix`` closed-shell branch (``src/cc/cctools.jl:35-45``)."""
    fock = gen_fock(EC)
    EC.files['f_mm'] = fock
    EC.files['e_m'] = np.diag(fock).copy()


def calc_HF_energy(EC: ECInfo) -> float:
    """``calc_HF_energy`` closed-shell (``src/cc/cctools.jl:65-70``)."""
    eo = EC.files['e_m'][EC.o]
    return float(np.sum(eo) + np.sum(np.diag(EC.ints1('oo'))) + EC.int0)


def orbital_energies(EC: ECInfo):
    """``orbital_energies`` (``src/scf/orbtools.jl:197-209``)."""
    eps = EC.files['e_m']
    return eps[EC.o], eps[EC.v]


# --------------------------------------------------------------------------
# amplitude updates, norms, dots (src/cc/cctools.jl)
# --------------------------------------------------------------------------
def update_singles(R1, eo, ev, shift):
    """``update_singles`` (``cctools.jl:198-210``): dT1 = -R1/(e_a - e_i + shift)."""
    return -R1 / (ev[:, None] - eo[None, :] + shift)


def update_doubles(R2, eo, ev, shift):
    """``update_doubles`` (``cctools.jl:245-258``): dT2 = -R2/(e_a+e_b-e_i-e_j+shift)."""
    den = (ev[:, None, None, None] + ev[None, :, None, None]
           - eo[None, None, :, None] - eo[None, None, None, :] + shift)
    return -R2 / den


def update_singles_EC(EC, R1, use_shift=True):
    eo, ev = orbital_energies(EC)
    return update_singles(R1, eo, ev, EC.options['shifts'] if use_shift else 0.0)


def update_doubles_EC(EC, R2, use_shift=True):
    eo, ev = orbital_energies(EC)
    return update_doubles(R2, eo, ev, EC.options['shiftp'] if use_shift else 0.0)


def calc_singles_norm(T1):
    """``calc_singles_norm`` (``cctools.jl:410-413``)."""
    return 2.0 * float(np.sum(T1 * T1))


def calc_contra_singles_norm(T1):
    """``calc_contra_singles_norm`` (``cctools.jl:420-423``)."""
    return 0.5 * float(np.sum(T1 * T1))


def calc_doubles_norm(T2):
    """``calc_doubles_norm`` (``cctools.jl:477-480``)."""
    return float(np.sum((2.0 * T2 - T2.transpose(1, 0, 2, 3)) * T2))


def calc_contra_doubles_norm(T2):
    """``calc_contra_doubles_norm`` (``cctools.jl:487-490``)."""
    return float(np.sum((2.0 * T2 + T2.transpose(1, 0, 2, 3)) * T2)) / 3.0


def calc_cs_singles_dot(T1, T1_):
    """``calc_cs_singles_dot`` (``cctools.jl:453-456``)."""
    return 2.0 * float(np.sum(T1 * T1_))


def calc_cs_doubles_dot(T2, T2_):
    """``calc_cs_doubles_dot`` (``cctools.jl:511-514``)."""
    return float(np.sum((2.0 * T2 - T2.transpose(1, 0, 2, 3)) * T2_))


def plain_dot(X, Y):
    """``LinearAlgebra.dot`` of one component (``diis.jl:170-172``)."""
    return float(np.sum(X * Y))


# --------------------------------------------------------------------------
# DIIS (src/solvers/diis.jl)
# --------------------------------------------------------------------------
class Diis:
    """Pulay DIIS of ``src/solvers/diis.jl:34-305`` with in-memory vectors."""

    def __init__(self, EC: ECInfo, maxdiis=None, resthr=None):
        self.maxdiis = EC.options['maxdiis'] if maxdiis is None else maxdiis
        self.resthr = EC.options['resthr'] if resthr is None else resthr
        self.amps = [None] * self.maxdiis
        self.res = [None] * self.maxdiis
        self.next = 0            # 0-based counterpart of diis.next
        self.nDim = 0
        self.bmat = np.zeros((self.maxdiis + 1, self.maxdiis + 1))

    def _dot(self, X, Y, dots):
        if not dots:
            return sum(plain_dot(x, y) for x, y in zip(X, Y))
        return sum(f(x, y) for f, x, y in zip(dots, X, Y))

    def _update_bmat(self, nDim, Res, ithis, dots):
        """``update_Bmat`` (``diis.jl:215-237``)."""
        this = self._dot(Res, Res, dots)
        for i in range(nDim):
            if i != ithis:
                d = self._dot(Res, self.res[i], dots)
                self.bmat[i, ithis] = d
                self.bmat[ithis, i] = d
            else:
                self.bmat[ithis, ithis] = this
            self.bmat[i, nDim] = -1.0
            self.bmat[nDim, i] = -1.0
        self.bmat[nDim, nDim] = 0.0
        return this

    def perform(self, Amps, Res, dots=()):
        """``perform!`` (``diis.jl:246-305``, CROP branch omitted: ``crop=false`` default)."""
        if self.nDim < self.maxdiis:
            self.nDim += 1
        ithis = self.next
        nDim = self.nDim
        self.amps[ithis] = [a.copy() for a in Amps]
        self.res[ithis] = [r.copy() for r in Res]
        this = self._update_bmat(nDim, Res, ithis, dots)
        rhs = np.zeros(nDim + 1)
        rhs[nDim] = -1.0
        bmat = self.bmat[:nDim + 1, :nDim + 1]
        # svd(bmat) \ rhs : minimum-norm least-squares solution through the SVD
        U, s, Vt = np.linalg.svd(bmat)
        # LinearAlgebra keeps singular values >= eps*S[1] (ldiv! of an SVD)
        tol = np.finfo(np.float64).eps * s[0]
        sinv = np.where(s >= tol, 1.0 / np.where(s >= tol, s, 1.0), 0.0)
        coeffs = Vt.T @ (sinv * (U.T @ rhs))
        if nDim == 1 and self.next == 0 and this > self.resthr:
            self.nDim = 0
            return Amps
        elif nDim < self.maxdiis:
            self.next = nDim
        else:
            coefmin = 1.e10
            for i in range(nDim):
                if abs(coeffs[i]) < coefmin and (i != ithis or nDim == 1):
                    coefmin = abs(coeffs[i])
                    self.next = i
        for k in range(len(Amps)):
            opt = coeffs[0] * self.amps[0][k]
            for i in range(1, nDim):
                opt = opt + coeffs[i] * self.amps[i][k]
            Amps[k][...] = opt
        return Amps


# --------------------------------------------------------------------------
# energies (src/cc/cc.jl:97-241)
# --------------------------------------------------------------------------
def calc_singles_energy(EC, T1, fock_only=False):
    """``calc_singles_energy`` closed-shell (``cc.jl:97-114``)."""
    ET1 = ET1SS = ET1OS = 0.0
    if T1.size > 0:
        if not fock_only:
            oovv = EC.ints2('oovv')
            ET1d = float(_es('ai,bj,ijab->', T1, T1, oovv))
            ET1ex = float(_es('bi,aj,ijab->', T1, T1, oovv))
            ET1SS = ET1d - ET1ex
            ET1OS = ET1d
            ET1 = ET1SS + ET1OS
        ET1 += 2.0 * float(np.sum(T1 * EC.files['f_mm'][EC.o, EC.v].T))
    return dict(E=ET1, ESS=ET1SS, EOS=ET1OS, EO=0.0)


def calc_doubles_energy(EC, T2):
    """``calc_doubles_energy`` closed-shell (``cc.jl:157-167``)."""
    oovv = EC.ints2('oovv')
    ET2d = float(_es('abij,ijab->', T2, oovv))
    ET2ex = float(_es('baij,ijab->', T2, oovv))
    ET2SS = ET2d - ET2ex
    ET2OS = ET2d
    return dict(E=ET2SS + ET2OS, ESS=ET2SS, EOS=ET2OS, EO=0.0)


def calc_hylleraas(EC, T1, T2, R1, R2):
    """``calc_hylleraas`` closed-shell (``cc.jl:195-241``)."""
    int2 = EC.ints2('oovv')
    ET1 = ET1SS = ET1OS = 0.0
    if T1.size > 0:
        ET1d = float(_es('ai,bj,ijab->', T1, T1, int2))
        ET1ex = float(_es('bi,aj,ijab->', T1, T1, int2))
        ET1SS = ET1d - ET1ex
        ET1OS = ET1d
        ET1 = ET1SS + ET1OS
    pET2d = float(_es('abij,ijab->', T2, int2))
    pET2ex = float(_es('baij,ijab->', T2, int2))
    rET2d = float(np.sum(T2 * R2))
    rET2ex = float(np.sum(T2.transpose(1, 0, 2, 3) * R2))
    ET2d = pET2d + rET2d
    ET2ex = pET2ex + rET2ex
    ET2SS = ET2d - ET2ex
    ET2OS = ET2d
    ET2 = ET2SS + ET2OS
    pET2SS = pET2d - pET2ex
    pET2OS = pET2d
    pET2 = pET2SS + pET2OS
    if T1.size > 0:
        fov = EC.files['f_mm'][EC.o, EC.v]
        pET1 = 2.0 * float(np.sum(fov.T * T1))
        rET1 = 2.0 * float(np.sum(R1 * T1))
        ET1 += pET1 + rET1
    else:
        pET1 = 0.0
    pET1 += ET1SS + ET1OS
    ET2 += ET1
    ET2SS += ET1SS
    ET2OS += ET1OS
    pET2 += pET1
    pET2SS += ET1SS
    pET2OS += ET1OS
    return dict(pE=pET2, pESS=pET2SS, pEOS=pET2OS, pEO=0.0,
                E=ET2, ESS=ET2SS, EOS=ET2OS, EO=0.0)


# --------------------------------------------------------------------------
# T1 dressing (src/cc/cc.jl:453-813, closed-shell, not mixed)
# --------------------------------------------------------------------------
def calc_dressed_ints(EC: ECInfo, T1, calc_d_vovv=None):
    """``calc_dressed_ints(EC,T1)`` (``cc.jl:924-929`` -> ``:453-770`` and ``:777-813``)."""
    if calc_d_vovv is None:
        calc_d_vovv = EC.options['calc_d_vovv']
    F = EC.files
    # <ik|j \hat l>  (cc.jl:483-495)
    hd_oooo = EC.ints2('oooo')
    ooov = EC.ints2('ooov')
    hd_oooo += _es('ijkd,dl->ijkl', ooov, T1)
    # <\hat a k| \hat j l>  (cc.jl:520-534)
    hd_vooo = EC.ints2('vooo')
    hd_vooo -= _es('kilj,ai->akjl', hd_oooo, T1)
    vovo = EC.ints2('vovo')
    hd_vooo += _es('akbl,bj->akjl', vovo, T1)
    # <ki\hat|dj>  (cc.jl:552-559)
    d_oovo = EC.ints2('oovo')
    oovv = EC.ints2('oovv')
    d_oovo += _es('kidb,bj->kidj', oovv, T1)
    F['d_oovo'] = d_oovo
    # <ij\hat|kl>  (cc.jl:561-563)
    hd_oooo += _es('ikbl,bj->ikjl', d_oovo, T1)
    F['d_oooo'] = hd_oooo
    # <ak\hat|jd>  (cc.jl:568-596)
    vovv = EC.ints2('vovv')
    d_voov = EC.ints2('voov')
    d_voov -= _es('kidj,ai->akjd', d_oovo, T1)
    d_voov += _es('akbd,bj->akjd', vovv, T1)
    F['d_voov'] = d_voov
    # <ak|b \hat l>  (cc.jl:603-608)
    hd_vovo = EC.ints2('vovo')
    hd_vovo += _es('akbd,dl->akbl', vovv, T1)
    # <ak\hat|bl>  (cc.jl:658-660)
    hd_vovo -= _es('ikbl,ai->akbl', d_oovo, T1)
    F['d_vovo'] = hd_vovo
    if calc_d_vovv:
        # <ak\hat|bd>  (cc.jl:669-676)
        d_vovv = EC.ints2('vovv')
        d_vovv -= _es('ikbd,ai->akbd', oovv, T1)
        F['d_vovv'] = d_vovv
    # <aj\hat|kl>  (cc.jl:713-718)
    d_vooo = hd_vooo
    d_vooo += _es('akjd,dl->akjl', d_voov, T1)
    F['d_vooo'] = d_vooo
    dress_fock_closedshell(EC, T1)


def dress_fock_closedshell(EC: ECInfo, T1):
    """``dress_fock_closedshell`` (``cc.jl:777-813``)."""
    F = EC.files
    o, v = EC.o, EC.v
    d_int1 = EC.int1.copy()
    dinter = EC.ints1(':v')
    d_int1[:, o] += dinter @ T1
    dinter = d_int1[o, :].copy()
    d_int1[v, :] -= T1 @ dinter
    F['dh_mm'] = d_int1.copy()
    dfock = d_int1
    d_oooo = F['d_oooo']
    d_vooo = F['d_vooo']
    d_oovo = F['d_oovo']
    foo = 2.0 * _es('ikjk->ij', d_oooo) - _es('ikkj->ij', d_oooo)
    fvo = 2.0 * _es('akik->ai', d_vooo) - _es('akki->ai', d_vooo)
    fov = 2.0 * _es('ikak->ia', d_oovo) - _es('kiak->ia', d_oovo)
    fvv = 2.0 * _es('akbk->ab', F['d_vovo']) - _es('akkb->ab', F['d_voov'])
    dfock[o, o] += foo
    dfock[v, o] += fvo
    dfock[o, v] += fov
    dfock[v, v] += fvv
    F['df_mm'] = dfock


def pseudo_dressed_ints(EC: ECInfo):
    """``pseudo_dressed_ints`` (``cc.jl:875-895``): undressed blocks (no singles)."""
    F = EC.files
    for blk in ('oovo', 'voov', 'vovo', 'vooo', 'oooo'):
        F['d_' + blk] = EC.ints2(blk)
    if EC.options['calc_d_vovv']:
        F['d_vovv'] = EC.ints2('vovv')
    F['dh_mm'] = EC.int1.copy()
    F['df_mm'] = F['f_mm'].copy()


# --------------------------------------------------------------------------
# residual (src/cc/cc.jl:1085-1111, 1343-1489)
# --------------------------------------------------------------------------
def calc_D2(EC: ECInfo, T1, T2, scalepp=False):
    """``calc_D2`` (``cc.jl:1085-1111``): ``D[p,q,i,j]`` over all orbitals."""
    n, no = EC.norb, EC.nocc
    o, v = EC.o, EC.v
    D2 = np.zeros((n, n, no, no))
    D2[v, v] = T2
    I = np.eye(no)
    D2[o, o] = _es('ij,lk->ikjl', I, I)
    if T1.size > 0:
        D2[v, v] += _es('ai,bj->abij', T1, T1)
        D2[o, v] = _es('ij,ak->jaik', I, T1)
        D2[v, o] = _es('ij,ak->ajki', I, T1)
    if scalepp:
        idx = np.arange(n)
        D2[idx, idx] *= 0.5
    return D2


def calc_kext(EC: ECInfo, T1, T2):
    """Particle-particle ladder ``K2pq`` of ``cc.jl:1400-1412``.

    The reference contracts the upper triangle ``q<=s`` with the halved
    diagonal of ``D2`` and symmetrises, ``K[p,r,i,j] = rK[p,r,i,j] + rK[r,p,j,i]``;
    by particle symmetry ``<pr|qs> = <rp|sq>`` and ``D[q,s,j,i] = D[s,q,i,j]``
    that equals the full sum taken here.
    """
    D2 = calc_D2(EC, T1, T2, False)
    return EC.ints.kext(D2)


def calc_cc_resid(EC: ECInfo, T1, T2, dc=False, linearized=False):
    """``calc_cc_resid`` closed-shell CCSD (``dc=False``) / DCSD (``dc=True``), ``cc.jl:1343-1489``."""
    o, v = EC.o, EC.v
    nocc, nvirt = EC.nocc, EC.nvirt
    F = EC.files
    if T1.size > 0:
        calc_dressed_ints(EC, T1)
    else:
        pseudo_dressed_ints(EC)
    T2t = 2.0 * T2 - T2.transpose(1, 0, 2, 3)
    dfock = F['df_mm']
    if T1.size > 0:
        R1 = F['dh_mm'][v, o].copy()
        int2 = F['d_oovo']
        fov = dfock[o, v]
        R1 += _es('abij,jb->ai', T2t, fov)
        R1 -= _es('kjci,cakj->ai', int2, T2t)
    else:
        R1 = np.zeros_like(T1)
    R2 = np.zeros((nvirt, nvirt, nocc, nocc))
    klcd = EC.ints2('oovv')
    int2 = F['d_oooo'].copy()
    if not dc and not linearized:
        int2 += _es('klcd,cdij->klij', klcd, T2)
    R2 += _es('klij,abkl->abij', int2, T2)
    # kext (cc.jl:1400-1425)
    K2pq = calc_kext(EC, T1, T2)
    R2 += K2pq[v, v]
    if T1.size > 0:
        R2 -= _es('kbij,ak->abij', K2pq[o, v], T1)
        R2 -= _es('akij,bk->abij', K2pq[v, o], T1)
        R2 += _es('klij,ak,bl->abij', K2pq[o, o], T1, T1)
        R1 += 2.0 * _es('akik->ai', K2pq[v, o]) - _es('akki->ai', K2pq[v, o])
        x1 = 2.0 * _es('klil->ki', K2pq[o, o]) - _es('klli->ki', K2pq[o, o])
        R1 -= _es('ki,ak->ai', x1, T1)
    if not dc and not linearized:
        R2 += _es('klcd,adkj,cbil->abij', klcd, T2, T2)
    fac = 0.5 if dc else 1.0
    xad = dfock[v, v].copy()
    xki = dfock[o, o].copy()
    if not linearized:
        xad -= fac * _es('klcd,cakl->ad', klcd, T2t)
        xki += fac * _es('klcd,cdil->ki', klcd, T2t)
    R2r = _es('ad,dbij->abij', xad, T2)
    R2r -= _es('ki,abkj->abij', xki, T2)
    int2 = F['d_voov'].copy()
    if not linearized:
        int2 += 0.5 * _es('klcd,adil->akic', klcd, T2t)
    R2r += _es('akic,cbkj->abij', int2, T2t)
    if not dc and not linearized:
        T2t = T2t - T2
        R2r -= _es('klcd,daki,cblj->abij', klcd, T2, T2t)
    int2 = F['d_vovo']
    R2r -= _es('akci,cbkj->abij', int2, T2)
    R2r -= _es('bkci,ackj->abij', int2, T2)
    R2 += R2r + R2r.transpose(1, 0, 3, 2)
    return R1, R2


# --------------------------------------------------------------------------
# MP2 guess and CC iterations
# --------------------------------------------------------------------------
def calc_MP2(EC: ECInfo, addsingles=True):
    """``calc_MP2`` (``cc.jl:1011-1024``)."""
    T2 = update_doubles_EC(EC, EC.ints2('vvoo'), use_shift=False)
    EMp2 = calc_doubles_energy(EC, T2)
    EC.files['T_vvoo'] = T2
    if addsingles:
        eo, ev = orbital_energies(EC)
        T1 = update_singles(EC.files['f_mm'][EC.v, EC.o], eo, ev, 0.0)
        EMp2s = calc_singles_energy(EC, T1, fock_only=True)
        EC.files['T_vo'] = T1
        EMp2 = {k: EMp2[k] + EMp2s[k] for k in EMp2}
    return EMp2


def cc_iterations(EC: ECInfo, T1, T2, dc=False, log=None):
    """``cc_iterations!`` for closed-shell CCSD/DCSD (``cc.jl:2518-2648``).

    Returns (Eh dict, converged, iterations); ``T1``/``T2`` are updated in place.
    """
    diis = Diis(EC)
    dots = (calc_cs_singles_dot, calc_cs_doubles_dot)
    thren = np.sqrt(EC.options['thr']) * EC.options['conven']
    Eh = dict(E=0.0)
    converged = False
    it = 0
    for it in range(1, EC.options['maxit'] + 1):
        R1, R2 = calc_cc_resid(EC, T1, T2, dc=dc)
        NormT2 = calc_doubles_norm(T2)
        NormR2 = calc_doubles_norm(R2)
        Eh = calc_hylleraas(EC, T1, T2, R1, R2)
        T2 += update_doubles_EC(EC, R2)
        NormT1 = calc_singles_norm(T1)
        NormR1 = calc_singles_norm(R1)
        T1 += update_singles_EC(EC, R1)
        diis.perform([T1, T2], [R1, R2], dots)
        EC.files['T_vvoo'] = T2.copy()
        En = calc_doubles_energy(EC, T2)['E']
        EC.files['T_vo'] = T1.copy()
        En += calc_singles_energy(EC, T1)['E']
        dE = En - Eh['E']
        NormR = NormR1 + NormR2
        if log is not None:
            log.append((it, NormR, 1.0 + NormT1 + NormT2, Eh['E'], dE))
        if NormR < EC.options['thr'] and abs(dE) < thren:
            converged = True
            break
    return Eh, converged, it


def calc_cc(EC: ECInfo, dc=False, log=None):
    """``calc_cc`` closed-shell branch (``cc.jl:2451-2516``); starts from ``T_vo``/``T_vvoo``."""
    T1 = EC.files.get('T_vo', np.zeros((EC.nvirt, EC.nocc))).copy()
    T2 = EC.files.get('T_vvoo', np.zeros((EC.nvirt, EC.nvirt, EC.nocc, EC.nocc))).copy()
    Eh, conv, it = cc_iterations(EC, T1, T2, dc=dc, log=log)
    EC.files['T_vo'] = T1
    EC.files['T_vvoo'] = T2
    return Eh, conv, it


def ccdriver(EC: ECInfo, method: str):
    """Closed-shell path of ``ccdriver``/``eval_cc_groundstate`` (``src/cc/ccdriver.jl:37-76,348-393``).

    ``method`` is one of ``ccsd``, ``dcsd``, ``ccsd(t)``, ``λccsd``, ``λccsd(t)``
    (``Λ`` or ``l`` prefix accepted).  Returns the energy dict with the
    reference's keys (``HF``, ``MP2c``, ``CCSD``/``DCSD`` ... ``(T)``, ``[T]``).
    """
    from . import lagrange, triples
    m = method.lower().replace('λ', 'l').replace('Λ', 'l')
    lam = m.startswith('l')
    if lam:
        m = m[1:]
    pert = m.endswith('(t)')
    base = m[:-3] if pert else m
    if base not in ('ccsd', 'dcsd'):
        raise ValueError(f"method {method} is outside the closed-shell CCSD/DCSD path")
    dc = base == 'dcsd'
    name = base.upper()
    out = {}
    calc_fock_matrix(EC)
    EHF = calc_HF_energy(EC)
    out['HF'] = EHF
    EMp2 = calc_MP2(EC)
    out['MP2c'] = EMp2['E']
    out['MP2'] = EMp2['E'] + EHF
    Eh, conv, it = calc_cc(EC, dc=dc)
    out[name + 'c'] = Eh['E']
    out[name] = Eh['E'] + EHF
    out['converged'] = conv
    out['iterations'] = it
    if lam:
        lagrange.calc_lm_cc(EC, dc=dc)
        name = 'Λ' + name
    if pert:
        if lam:
            ET3, ET3b = triples.calc_LpertT_closed_shell(EC)
        else:
            ET3, ET3b = triples.calc_pertT_closed_shell(EC)
        out['[T]'] = ET3b
        out['(T)'] = ET3
        out[name + '(T)c'] = Eh['E'] + ET3
        out[name + '(T)'] = Eh['E'] + ET3 + EHF
    return out
