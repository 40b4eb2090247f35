"""Closed-shell (T) and Lambda-(T) oracle, restating ``src/cc/cc_triples.jl:39-288``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
from __future__ import annotations

import numpy as np

from .ccsd import ECInfo, orbital_energies


def _es(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


def triplets(nocc):
    """Loop order of ``cc_triples.jl:61-71``: k, j<=k, i<=j, skipping i=j=k; yields (i,j,k,fac)."""
    for k in range(nocc):
        for j in range(k + 1):
            prefac = 1.0 if j == k else 2.0
            for i in range(j + 1):
                fac = prefac
                if i == j:
                    if j == k:
                        continue
                    fac = 1.0
                yield i, j, k, fac


def _kijk(T2, vvvo, ovoo, i, j, k):
    """``Kijk`` of ``cc_triples.jl:87-102``."""
    K = _es('ad,bcd->abc', T2[:, :, i, j], vvvo[:, :, :, k])
    K += _es('db,acd->abc', T2[:, :, i, j], vvvo[:, :, :, k])
    K += _es('ad,cbd->abc', T2[:, :, i, k], vvvo[:, :, :, j])
    K += _es('dc,abd->abc', T2[:, :, i, k], vvvo[:, :, :, j])
    K += _es('bd,cad->abc', T2[:, :, j, k], vvvo[:, :, :, i])
    K += _es('dc,bad->abc', T2[:, :, j, k], vvvo[:, :, :, i])
    K -= _es('bal,lc->abc', T2[:, :, :, i], ovoo[:, :, j, k])
    K -= _es('abl,lc->abc', T2[:, :, :, j], ovoo[:, :, i, k])
    K -= _es('cal,lb->abc', T2[:, :, :, i], ovoo[:, :, k, j])
    K -= _es('acl,lb->abc', T2[:, :, :, k], ovoo[:, :, i, j])
    K -= _es('cbl,la->abc', T2[:, :, :, j], ovoo[:, :, k, i])
    K -= _es('bcl,la->abc', T2[:, :, :, k], ovoo[:, :, j, i])
    return K


def _xijk(K, eo, ev, i, j, k):
    """``X`` of ``cc_triples.jl:112-116``."""
    X = (4.0 * K - 2.0 * K.transpose(0, 2, 1) - 2.0 * K.transpose(2, 1, 0)
         - 2.0 * K.transpose(1, 0, 2) + K.transpose(1, 2, 0) + K.transpose(2, 0, 1))
    den = eo[i] + eo[j] + eo[k] - ev[:, None, None] - ev[None, :, None] - ev[None, None, :]
    return X / den


def calc_pertT_closed_shell(EC: ECInfo, T1=None, T2=None, triplet_subset=None):
    """``calc_pertT_closed_shell`` (``cc_triples.jl:39-148``); returns (E(T), E[T]).

    ``triplet_subset`` (optional list of (i,j,k,fac)) restricts the triplet loop,
    which the CPU-baseline timing uses to time a bounded sample.
    """
    if T1 is None:
        T1 = EC.files['T_vo']
    if T2 is None:
        T2 = EC.files['T_vvoo']
    vv_oo = EC.ints2('oovv').transpose(2, 3, 0, 1)
    vvvo = EC.ints2('vvvo')
    ovoo = EC.ints2('ovoo')
    nocc, nvir = EC.nocc, EC.nvirt
    eo, ev = orbital_energies(EC)
    Enb3 = 0.0
    IntX = np.zeros((nvir, nocc))
    IntY = np.zeros((nvir, nocc))
    trip = triplets(nocc) if triplet_subset is None else triplet_subset
    for i, j, k, fac in trip:
        K = _kijk(T2, vvvo, ovoo, i, j, k)
        X = _xijk(K, eo, ev, i, j, k)
        Enb3 += fac * float(np.sum(K * X))
        IntX[:, i] += fac * _es('abc,bc->a', X, vv_oo[:, :, j, k])
        IntX[:, j] += fac * _es('abc,ac->b', X, vv_oo[:, :, i, k])
        IntX[:, k] += fac * _es('abc,ab->c', X, vv_oo[:, :, i, j])
        IntY[:, i] += fac * _es('abc,bc->a', X, T2[:, :, j, k])
        IntY[:, j] += fac * _es('abc,ac->b', X, T2[:, :, i, k])
        IntY[:, k] += fac * _es('abc,ab->c', X, T2[:, :, i, j])
    En3 = float(np.sum(T1 * IntX))
    fov = EC.files['f_mm'][EC.o, EC.v]
    En3 += float(np.sum(fov.T * IntY))
    En3 += Enb3
    return En3, Enb3


def contra2covariant(T2):
    """``contra2covariant`` (``src/cc/cctools.jl:935-938``)."""
    return (2.0 * T2 + T2.transpose(1, 0, 2, 3)) / 3.0


def calc_LpertT_closed_shell(EC: ECInfo):
    """``calc_ΛpertT_closed_shell`` (``cc_triples.jl:159-288``); returns (E(T), E[T])."""
    T1 = EC.files['T_vo']
    T2 = EC.files['T_vvoo']
    U1 = EC.files['U_vo']
    U2 = contra2covariant(EC.files['U_vvoo'])
    vv_oo = EC.ints2('oovv').transpose(2, 3, 0, 1)
    vvvo = EC.ints2('vvvo')
    ovoo = EC.ints2('ovoo')
    vv_vo = EC.ints2('vovv').transpose(2, 3, 0, 1)
    ov_oo = EC.ints2('ooov').transpose(2, 3, 0, 1)
    nocc, nvir = EC.nocc, EC.nvirt
    eo, ev = orbital_energies(EC)
    Enb3 = 0.0
    IntX = np.zeros((nvir, nocc))
    IntY = np.zeros((nvir, nocc))
    for i, j, k, fac in triplets(nocc):
        K = _kijk(T2, vvvo, ovoo, i, j, k)
        X = _xijk(K, eo, ev, i, j, k)
        KU = _kijk(U2, vv_vo, ov_oo, i, j, k)
        Enb3 += fac * float(np.sum(KU * X))
        IntX[:, i] += fac * _es('abc,bc->a', X, vv_oo[:, :, j, k])
        IntX[:, j] += fac * _es('abc,ac->b', X, vv_oo[:, :, i, k])
        IntX[:, k] += fac * _es('abc,ab->c', X, vv_oo[:, :, i, j])
        IntY[:, i] += fac * _es('abc,bc->a', X, U2[:, :, j, k])
        IntY[:, j] += fac * _es('abc,ac->b', X, U2[:, :, i, k])
        IntY[:, k] += fac * _es('abc,ab->c', X, U2[:, :, i, j])
    En3 = 0.5 * float(np.sum(U1 * IntX))
    fov = EC.files['f_mm'][EC.o, EC.v]
    En3 += float(np.sum(fov.T * IntY))
    En3 += Enb3
    return En3, Enb3
