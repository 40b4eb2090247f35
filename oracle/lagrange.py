"""Closed-shell Lambda (CC Lagrange multiplier) oracle, restating ``src/cc/cc_lagrange.jl``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``U1``/``U2`` are the reference's contravariant multipliers (``U_vo``/``U_vvoo``).
"""
from __future__ import annotations

import numpy as np

from .ccsd import (ECInfo, Diis, calc_dressed_ints, update_doubles_EC, update_singles_EC,
                   calc_contra_doubles_norm, calc_contra_singles_norm)


def _es(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


def calc_1RDM(EC: ECInfo, U1, U2, T1, T2):
    """``calc_1RDM`` closed-shell (``cc_lagrange.jl:793-837``); returns (D1, dD1)."""
    Doo = -2.0 * _es('cdik,cdjk->ij', U2, T2)
    Dvv = 2.0 * _es('bckl,ackl->ab', U2, T2)
    if U1.size > 0:
        Dov = U1.T.copy()
        Dvo = 2.0 * _es('ck,acik->ai', U1, T2) - _es('ck,acki->ai', U1, T2)
    else:
        Dov = np.zeros((U2.shape[2], U2.shape[0]))
        Dvo = np.zeros((U2.shape[0], U2.shape[2]))
    if T1.size > 0:
        dDov = Dov
        dDoo = Doo - Dov @ T1
        dDvv = Dvv + T1 @ Dov
        dDvo = Dvo + 2.0 * T1 - Dvv @ T1 + T1 @ dDoo
    else:
        dDov, dDoo, dDvv, dDvo = Dov, Doo, Dvv, Dvo
    n = EC.norb
    o, v = EC.o, EC.v
    D1 = np.zeros((n, n))
    D1[o, o] = Doo
    D1[v, v] = Dvv
    D1[o, v] = Dov
    D1[v, o] = Dvo
    dD1 = np.zeros((n, n))
    dD1[o, o] = dDoo
    dD1[v, v] = dDvv
    dD1[o, v] = dDov
    dD1[v, o] = dDvo
    return D1, dD1


def calc_dU2(EC: ECInfo, T1, U2):
    """``calc_dU2`` closed-shell (``cc_lagrange.jl:222-236``): ``dU2[p,q,m,n]``."""
    n = EC.norb
    o, v = EC.o, EC.v
    dU2 = np.zeros((n, n) + U2.shape[2:])
    dU2[v, v] = U2
    if T1.size > 0:
        dUovoo = _es('abmn,ai->ibmn', U2, T1)
        dU2[v, o] = -_es('abmn,bj->ajmn', U2, T1)
        dU2[o, o] = _es('ibmn,bj->ijmn', dUovoo, T1)
        dU2[o, v] = -dUovoo
    return dU2


def calc_3ext_times_T2(EC: ECInfo, T2):
    """``calc_3ext_times_T2`` closed-shell (``cc_lagrange.jl:948-952``): ``vT_ovoo``."""
    int2 = EC.files['d_vovv']
    EC.files['vT_ovoo'] = _es('bmdc,dcji->mbij', int2, T2)


def _kext_lambda(EC: ECInfo, dU2):
    """``Kmmoo[r,s,m,n] = sum_{pq} <pq|rs> dU2[p,q,m,n]`` (``cc_lagrange.jl:41-58``).

    The reference builds the r<=s triangle and fills r>s from ``Kxoo[x,n,m]``;
    by particle symmetry that equals the full sum.  ``<pq|rs> = <rs|pq>`` only for
    Hermitian integrals, so the sum is taken over the first two indices here as
    in the reference.
    """
    n = EC.norb
    ints = EC.ints
    if hasattr(ints, 'int2'):
        return _es('pqrs,pqmn->rsmn', ints.int2, dU2)
    # DF: <pq|rs> = sum_Q B[p,r,Q] B[q,s,Q]; one r at a time as two matrix products:
    # V_r[(s,q),p] = B[q,s,:] B[p,r,:]^T,  K[r,s,mn] = V_r[s,(q,p)] dU2[p,q,(mn)]
    B = ints.B
    nq = B.shape[2]
    B2T = np.ascontiguousarray(B.transpose(1, 0, 2)).reshape(n * n, nq)      # [(s,q), Q]
    dUt = np.ascontiguousarray(dU2.transpose(1, 0, 2, 3)).reshape(n * n, -1)  # [(q,p), mn]
    out = np.empty((n, n) + dU2.shape[2:])
    for r in range(n):
        Vr = B2T @ B[:, r, :].T
        out[r] = (Vr.reshape(n, n * n) @ dUt).reshape((n,) + dU2.shape[2:])
    return out


def calc_ccsd_vector_times_Jacobian(EC: ECInfo, U1, U2, dc=False):
    """``calc_ccsd_vector_times_Jacobian`` closed-shell (``cc_lagrange.jl:9-212``)."""
    F = EC.files
    o, v = EC.o, EC.v
    T1 = F['T_vo']
    T2 = F['T_vvoo']
    D1, dD1 = calc_1RDM(EC, U1, U2, T1, T2)
    fock = F['f_mm']
    dfock = F['df_mm']
    fov = fock[o, v]
    dfov = dfock[o, v]
    R1 = 2.0 * fov.T.copy() if U1.size > 0 else U1.copy()
    oovv = EC.ints2('oovv')
    R2 = 2.0 * _es('mnef->efmn', oovv) - _es('nmef->efmn', oovv)
    int2 = F['d_oooo'].copy()
    if not dc:
        int2 += _es('mncd,cdij->mnij', oovv, T2)
    R2 += _es('mnij,efij->efmn', int2, U2)
    # kext (cc_lagrange.jl:39-69)
    dU2 = calc_dU2(EC, T1, U2)
    Kmmoo = _kext_lambda(EC, dU2)
    R2 += Kmmoo[v, v]
    if U1.size > 0:
        R1 += 2.0 * _es('ejmj->em', Kmmoo[v, o])
        if T1.size > 0:
            R1 += 2.0 * _es('ebmj,bj->em', Kmmoo[v, v], T1)
    fac = 0.25 if dc else 0.5
    tR2 = fac * _es('mk,knef->efmn', D1[o, o], oovv)
    tR2 -= fac * _es('ce,mncf->efmn', D1[v, v], oovv)
    tT2 = 2.0 * T2 - T2.transpose(0, 1, 3, 2)
    xoo = _es('cdil,mlcd->mi', tT2, oovv)
    xvv = _es('ackl,klec->ae', tT2, oovv)
    int2 = F['d_voov'].copy()
    int2 += _es('kmce,acik->amie', oovv, tT2)
    tR2 += _es('afin,amie->efmn', U2, int2)
    if U1.size > 0:
        tR2 += 0.5 * _es('em,nf->efmn', U1, dfov)
    int2 = F['d_vovv']
    if U1.size > 0:
        tR2 += 0.5 * _es('an,amfe->efmn', U1, int2)
    oovo = F['d_oovo']
    if U1.size > 0:
        tR2 -= 0.5 * _es('fi,mnei->efmn', U1, oovo)
    pR2 = 2.0 * tR2 - tR2.transpose(0, 1, 3, 2)
    del tR2
    D2 = _es('acik,cbkj->abij', U2, tT2)
    del tT2
    if U1.size > 0:
        R1 += 2.0 * _es('adml,aled->em', D2, int2)
        R1 -= 2.0 * _es('edil,lmdi->em', D2, oovo)
    if not dc:
        pR2 -= _es('fcnl,mlce->efmn', D2, oovv)
    D2 = _es('acik,cbkj->abij', U2, T2)
    D2 += _es('caik,bckj->abij', U2, T2)
    if not dc:
        pR2 += _es('ednk,kmfd->efmn', D2, oovv)
    if U1.size > 0:
        R1 -= 2.0 * _es('acmk,akce->em', D2, int2)
        R1 += 2.0 * _es('ecik,mkci->em', D2, oovo)
    del D2
    D2 = _es('cdij,cdkl->ijkl', U2, T2)
    if U1.size > 0:
        R1 += 2.0 * _es('mjkl,klej->em', D2, oovo)
    if not dc:
        R2 += _es('mnkl,klef->efmn', D2, oovv)
    del D2
    if U1.size > 0:
        vT_ovoo = F['vT_ovoo']
        R1 -= 2.0 * _es('ebij,mbij->em', U2, vT_ovoo)
    fac = 0.5 if dc else 1.0
    pR2 += _es('afmn,ae->efmn', U2, dfock[v, v] - fac * xvv)
    pR2 -= _es('efin,mi->efmn', U2, dfock[o, o] + fac * xoo)
    int2 = F['d_vovo']
    pR2 -= _es('afin,amei->efmn', U2, int2)
    pR2 -= _es('ebin,bmfi->efmn', U2, int2)
    R2 += pR2 + pR2.transpose(1, 0, 3, 2)
    del pR2
    if U1.size > 0:
        momm = EC.ints2('momm')
        R1 += _es('pq,qmpe->em', dD1, 2.0 * momm[:, :, :, v]) - _es('pq,qmep->em', dD1, momm[:, :, v, :])
        int2 = F['d_vooo']
        R1 -= 2.0 * _es('ebij,bmji->em', U2, int2)
        R1 += _es('mk,ke->em', D1[o, o], dfov)
        R1 -= _es('de,md->em', D1[v, v], dfov)
        R1 += _es('am,ae->em', U1, dfock[v, v] - xvv)
        R1 -= _es('ei,mi->em', U1, dfock[o, o] + xoo)
    return R1, R2


def calc_correlation_norm(EC: ECInfo, U1, U2):
    """``calc_correlation_norm`` closed-shell (``cc_lagrange.jl:1065-1077``)."""
    T2 = EC.files['T_vvoo']
    Norm = float(np.sum(U2 * T2))
    if U1.size > 0:
        T1 = EC.files['T_vo']
        if T1.size > 0:
            Norm += float(_es('abij,ai,bj->', U2, T1, T1))
            Norm += float(np.sum(U1 * T1))
    return Norm


def add_singles2doubles(T2, T1, make_contravariant=True):
    """``add_singles2doubles!`` closed-shell (``src/cc/cctools.jl:919-928``), returns new arrays."""
    T2 = T2.copy()
    T1 = T1.copy()
    if T1.size > 0:
        T2 += _es('ai,bj->abij', T1, T1)
    if make_contravariant:
        tT2 = T2 - T2.transpose(0, 1, 3, 2)
        T2 += tT2
        T1 += T1
    return T1, T2


def lm_cc_iterations(EC: ECInfo, dc=False, log=None):
    """``lm_cc_iterations!`` closed-shell (``cc_lagrange.jl:1155-1223``); returns (converged, it)."""
    T1 = EC.files['T_vo']
    T2 = EC.files['T_vvoo']
    calc_dressed_ints(EC, T1, calc_d_vovv=True)
    calc_3ext_times_T2(EC, T2)
    diis = Diis(EC)
    U1, U2 = add_singles2doubles(T2, T1)
    thren = np.sqrt(EC.options['thr']) * EC.options['conven']
    LTNorm = 0.0
    converged = False
    it = 0
    for it in range(1, EC.options['maxit'] + 1):
        R1, R2 = calc_ccsd_vector_times_Jacobian(EC, U1, U2, dc=dc)
        U2 += update_doubles_EC(EC, R2)
        NormLM2 = calc_contra_doubles_norm(U2)
        NormR2 = calc_contra_doubles_norm(R2)
        NormLM1 = calc_contra_singles_norm(U1)
        NormR1 = calc_contra_singles_norm(R1)
        U1 += update_singles_EC(EC, R1)
        diis.perform([U1, U2], [R1, R2])
        EC.files['U_vo'] = U1.copy()
        EC.files['U_vvoo'] = U2.copy()
        prev = LTNorm
        LTNorm = calc_correlation_norm(EC, U1, U2)
        NormR = NormR1 + NormR2
        if log is not None:
            log.append((it, NormR, 1.0 + NormLM1 + NormLM2, LTNorm))
        if NormR < EC.options['thr'] and abs(LTNorm - prev) < thren:
            converged = True
            break
    return converged, it


def calc_lm_cc(EC: ECInfo, dc=False, log=None):
    """``calc_lm_cc`` closed-shell (``cc_lagrange.jl:1120-1153``)."""
    return lm_cc_iterations(EC, dc=dc, log=log)
