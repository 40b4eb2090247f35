"""Pins of the CPU oracle (no GPU).

Each test checks the oracle against something it did not compute itself:
the reference's own test energies, a closed form, an independent brute-force
formula, or finite differences.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import lagrange, triples
from oracle.ccsd import (ECInfo, DFInts, ccdriver, calc_cc_resid, calc_dressed_ints,
                         calc_doubles_energy, calc_singles_energy, calc_fock_matrix,
                         calc_HF_energy)
from oracle.fcidump import read_fcidump
from oracle.synthetic import make_df_system, splitmix64, write_fcidump

H2O = os.path.join(GOLDEN, "H2O.FCIDUMP.gz")
H2O_ST1 = os.path.join(GOLDEN, "H2O_ST1.FCIDUMP.gz")
EPS_REF = 1e-6          # tolerance of test/h2o.jl:5 and test/h2o_st1.jl:5


@pytest.fixture(scope="module")
def h2o():
    return read_fcidump(H2O)


@pytest.fixture(scope="module")
def h2o_st1():
    return read_fcidump(H2O_ST1)


# ---------------------------------------------------------------- reference energies
def test_h2o_ccsd_t(h2o):
    """test/h2o.jl:6-8,19-22."""
    e = ccdriver(ECInfo.from_fcidump(h2o), "ccsd(t)")
    assert abs(e["HF"] - (-75.6457645933)) < EPS_REF
    assert abs(e["MP2c"] - (-0.287815830908)) < EPS_REF
    assert abs(e["CCSD(T)c"] - (-0.329259440500)) < EPS_REF


def test_h2o_lambda_ccsd_t(h2o):
    """test/h2o.jl:9,24-25."""
    e = ccdriver(ECInfo.from_fcidump(h2o), "λccsd(t)")
    assert abs(e["ΛCCSD(T)c"] - (-0.326915143863)) < EPS_REF


def test_h2o_dcsd(h2o):
    """test/h2o.jl:10,27-28."""
    e = ccdriver(ECInfo.from_fcidump(h2o), "dcsd")
    assert abs(e["DCSDc"] - (-0.328754956597)) < EPS_REF


@pytest.mark.parametrize("method,eref", [("ccsd", -0.082041632192), ("dcsd", -0.082498102641)])
def test_h2o_st1(h2o_st1, method, eref):
    """Non-Hermitian FCIDUMP, test/h2o_st1.jl:6-9,16-22."""
    e = ccdriver(ECInfo.from_fcidump(h2o_st1), method)
    assert abs(e["HF"] - (-76.298014304953)) < EPS_REF
    assert abs(e["MP2c"] - (-0.069545740864)) < EPS_REF
    assert abs(e[method.upper() + "c"] - eref) < EPS_REF


# ---------------------------------------------------------------- closed forms
def _fci_two_electron(h, eri):
    """Singlet FCI for two electrons: Psi = sum_pq C_pq phi_p(1) phi_q(2), C symmetric.

    (H C)_rs = sum_p h_rp C_ps + sum_q h_sq C_rq + sum_pq (rp|sq) C_pq.
    """
    n = h.shape[0]
    I = np.eye(n)
    H = (np.einsum('rp,sq->rspq', h, I) + np.einsum('sq,rp->rspq', h, I)
         + np.einsum('rpsq->rspq', eri)).reshape(n * n, n * n)
    iu = np.triu_indices(n)
    basis = []
    for p, q in zip(*iu):
        c = np.zeros((n, n))
        c[p, q] = c[q, p] = 1.0
        basis.append(c.ravel() / np.linalg.norm(c))
    S = np.array(basis).T
    Hs = S.T @ H @ S
    return np.linalg.eigvalsh(0.5 * (Hs + Hs.T))[0]


@pytest.mark.parametrize("method", ["ccsd", "dcsd"])
def test_two_electron_equals_fci(method):
    """CCSD and DCSD are exact for two electrons (DCSD: Kats & Manby, JCP 139, 021102)."""
    s = make_df_system(1, 9, 40, 5)
    EC = ECInfo.from_df(s)
    EC.options["thr"] = 1e-20
    EC.options["maxit"] = 200
    e = ccdriver(EC, method)
    eri = np.einsum('pqQ,rsQ->pqrs', s.B, s.B)
    efci = _fci_two_electron(s.h, eri) + s.e0
    assert abs(e[method.upper()] - efci) < 1e-10


def test_two_electron_triples_vanish():
    """o=1: every triplet has i=j=k and is skipped (cc_triples.jl:66-69), so E(T)=0."""
    s = make_df_system(1, 6, 20, 2)
    e = ccdriver(ECInfo.from_df(s), "ccsd(t)")
    assert e["(T)"] == 0.0 and e["[T]"] == 0.0


# ---------------------------------------------------------------- brute force [T]
def test_bracket_t_bruteforce():
    """[T] from an explicit T3 over all (i,j,k), as in cc_tests.jl:54-75.

    The restricted i<=j<=k loop with factors 2/1 must equal the unrestricted
    sum  -sum W (4/3 T3 - 2 T3[a,i,b,k,c,j] + 2/3 T3[c,i,a,j,b,k]),
    W = T3 (e_a+e_b+e_c-e_i-e_j-e_k), T3 = K/(e_i+e_j+e_k-e_a-e_b-e_c).
    """
    s = make_df_system(3, 6, 25, 7)
    EC = ECInfo.from_df(s)
    e = ccdriver(EC, "ccsd(t)")
    T2 = EC.files["T_vvoo"]
    eo, ev = EC.files["e_m"][:3], EC.files["e_m"][3:]
    vvvo, ovoo = EC.ints2("vvvo"), EC.ints2("ovoo")
    no, nv = 3, 6
    T3 = np.zeros((nv, no, nv, no, nv, no))
    for i in range(no):
        for j in range(no):
            for k in range(no):
                K = triples._kijk(T2, vvvo, ovoo, i, j, k)
                D = eo[i] + eo[j] + eo[k] - ev[:, None, None] - ev[None, :, None] - ev[None, None, :]
                T3[:, i, :, j, :, k] = K / D
    Dv = ev[:, None, None, None, None, None] + ev[None, None, :, None, None, None] + ev[None, None, None, None, :, None]
    Do = eo[None, :, None, None, None, None] + eo[None, None, None, :, None, None] + eo[None, None, None, None, None, :]
    W = T3 * (Dv - Do)
    brute = -np.sum(W * (4.0 / 3.0 * T3 - 2.0 * T3.transpose(0, 1, 2, 5, 4, 3)
                         + 2.0 / 3.0 * T3.transpose(2, 1, 4, 3, 0, 5)))
    assert abs(brute - e["[T]"]) < 1e-12 * max(1.0, abs(brute))


# ---------------------------------------------------------------- Lambda residual = dL/dT
def _converged(no, nv, nq, seed, method="ccsd"):
    s = make_df_system(no, nv, nq, seed)
    EC = ECInfo.from_df(s)
    EC.options["thr"] = 1e-22
    EC.options["maxit"] = 300
    ccdriver(EC, method)
    return s, EC


@pytest.mark.parametrize("dc", [False, True])
def test_lambda_residual_is_lagrangian_gradient(dc):
    """R_Lambda(U) = d/dT [E(T) + U1.R1(T) + U2.R2(T)] (equations.tex:755-801)."""
    s, EC = _converged(3, 7, 30, 11, "dcsd" if dc else "ccsd")
    T1 = EC.files["T_vo"].copy()
    T2 = EC.files["T_vvoo"].copy()
    rng = np.random.default_rng(3)
    U1 = 0.01 * rng.standard_normal(T1.shape)
    U2 = 0.01 * rng.standard_normal(T2.shape)
    U2 = U2 + U2.transpose(1, 0, 3, 2)
    calc_dressed_ints(EC, T1, calc_d_vovv=True)
    lagrange.calc_3ext_times_T2(EC, T2)
    RL1, RL2 = lagrange.calc_ccsd_vector_times_Jacobian(EC, U1, U2, dc=dc)
    dT1 = rng.standard_normal(T1.shape)
    dT2 = rng.standard_normal(T2.shape)
    dT2 = dT2 + dT2.transpose(1, 0, 3, 2)

    def L(t1, t2):
        files = dict(EC.files)
        R1, R2 = calc_cc_resid(EC, t1, t2, dc=dc)
        EC.files.update(files)
        E = calc_doubles_energy(EC, t2)["E"] + calc_singles_energy(EC, t1)["E"]
        return E + np.sum(U1 * R1) + np.sum(U2 * R2)

    h = 1e-4
    fd = lambda h: (L(T1 + h * dT1, T2 + h * dT2) - L(T1 - h * dT1, T2 - h * dT2)) / (2 * h)
    rich = (4.0 * fd(h / 2) - fd(h)) / 3.0
    an = np.sum(RL1 * dT1) + np.sum(RL2 * dT2)
    assert abs(rich - an) < 1e-8 * abs(an)


def test_one_rdm_is_energy_derivative():
    """dE_corr/dh[p,q] = dD1[q,p] (cc_lagrange.jl:793-837), by Richardson-extrapolated FD."""
    s = make_df_system(3, 7, 30, 11)

    def solve(h1):
        EC = ECInfo(DFInts(s.B), h1, 0.0, s.nocc)
        EC.options["thr"] = 1e-22
        EC.options["maxit"] = 300
        return ccdriver(EC, "ccsd"), EC

    _, EC = solve(s.h)
    lagrange.calc_lm_cc(EC)
    F = EC.files
    D1, dD1 = lagrange.calc_1RDM(EC, F["U_vo"], F["U_vvoo"], F["T_vo"], F["T_vvoo"])
    H = np.random.default_rng(1).standard_normal(s.h.shape)
    fd = lambda h: (solve(s.h + h * H)[0]["CCSDc"] - solve(s.h - h * H)[0]["CCSDc"]) / (2 * h)
    rich = (4.0 * fd(5e-4) - fd(1e-3)) / 3.0
    an = np.sum(dD1.T * H)
    assert abs(rich - an) < 1e-7 * abs(an)


# ---------------------------------------------------------------- synthetic systems, FCIDUMP
def test_splitmix64_published_vectors():
    """First outputs of Vigna's splitmix64 with state 0."""
    g = np.uint64(0x9E3779B97F4A7C15)
    x = np.array([0, g, (2 * int(g)) % 2**64], dtype=np.uint64)
    out = [int(v) for v in splitmix64(x)]
    assert out == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_synthetic_fock_is_diagonal():
    s = make_df_system(4, 11, 30, 0)
    EC = ECInfo.from_df(s)
    calc_fock_matrix(EC)
    assert np.max(np.abs(EC.files["f_mm"] - np.diag(s.eps))) < 1e-13
    assert np.all(np.diff(s.eps[:4]) >= 0) and np.all(np.diff(s.eps[4:]) >= 0)
    assert np.allclose(s.B, s.B.transpose(1, 0, 2))


def test_fcidump_roundtrip(tmp_path):
    """Config-0 style FCIDUMP: the 4-index path reproduces the DF path."""
    s = make_df_system(2, 5, 12, 0)
    p = str(tmp_path / "sys.FCIDUMP")
    write_fcidump(p, s)
    fd = read_fcidump(p)
    e4 = ccdriver(ECInfo.from_fcidump(fd), "ccsd(t)")
    edf = ccdriver(ECInfo.from_df(s), "ccsd(t)")
    assert abs(e4["HF"] - edf["HF"]) < 1e-12
    assert abs(e4["CCSD(T)"] - edf["CCSD(T)"]) < 1e-11


@pytest.mark.parametrize("rank", [50, 13])
def test_dfprep_oracle_closed_form(rank):
    """oracle.dfprep (generate_3idx_integrals + sqrtinvchol, dftools.jl:92-125,
    tensortools.jl:377-391): sum_L mmL mmL = C^T C^T (mu nu|P) pinv((P|Q)) (P|ls) C C, for a
    full-rank and a rank-deficient metric; M M^T = pinv((P|Q))."""
    from oracle.dfprep import df_ao_to_mo, sqrtinvchol
    rng = np.random.default_rng(rank)
    nao, nmo, nfit = 8, 6, 20
    G = rng.standard_normal((nfit, rank))
    PQ = G @ G.T
    AAP = rng.standard_normal((nao, nao, nfit))
    AAP = AAP + AAP.transpose(1, 0, 2)
    C = rng.standard_normal((nao, nmo))
    Pinv = np.linalg.pinv(PQ, rcond=1e-12)
    M = sqrtinvchol(PQ, 1e-9)
    assert M.shape == (nfit, min(rank, nfit))
    assert np.max(np.abs(M @ M.T - Pinv)) < 1e-9 * np.max(np.abs(Pinv))
    B = df_ao_to_mo(C, AAP, PQ, 1e-9)
    ref = np.einsum("mp,nq,mnP,PQ,lsQ,lr,st->pqrt", C, C, AAP, Pinv, AAP, C, C, optimize=True)
    assert np.max(np.abs(np.einsum("pqL,rsL->pqrs", B, B) - ref)) < 1e-9 * np.max(np.abs(ref))



# ---------------------------------------------------------------- frozen orbitals, writer
@pytest.mark.parametrize("name,orbs", [("H2O.FCIDUMP.gz", [1]), ("H2O.FCIDUMP.gz", [1, 2, 23]),
                                       ("H2O_ST1.FCIDUMP.gz", [1]), ("H2O_ST1.FCIDUMP.gz", [2, 20])])
def test_freeze_orbs_oracle_invariants(name, orbs):
    """freeze_orbs_in_dump (dumptools.jl:20-62) against what freezing must leave unchanged:
    the HF energy (asserted by the reference itself after @freeze_orbs, test/h2o_st1.jl:33-35;
    against the reference's EHF for these files) and the Fock matrix (gen_fock over the
    remaining occupied orbitals of the frozen dump = the full Fock on the kept orbitals)."""
    from oracle.fcidump import _gen_fock, freeze_orbs_in_dump
    ehf_ref = {"H2O.FCIDUMP.gz": -75.6457645933, "H2O_ST1.FCIDUMP.gz": -76.298014304953}[name]
    full = read_fcidump(os.path.join(GOLDEN, name))
    nocc = full.nelec // 2
    f_full = _gen_fock(full, range(nocc))
    kept = [p for p in range(full.norb) if p + 1 not in orbs]
    fd = freeze_orbs_in_dump(read_fcidump(os.path.join(GOLDEN, name)), orbs)
    ncore = sum(1 for o in orbs if o <= nocc)
    assert fd.norb == full.norb - len(orbs) and fd.nelec == full.nelec - 2 * ncore
    f_fr = _gen_fock(fd, range(nocc - ncore))
    assert np.max(np.abs(f_fr - f_full[np.ix_(kept, kept)])) < 1e-12
    ehf = lambda d, f, no: float(np.sum(np.diag(f)[:no]) + np.sum(np.diag(d.int1)[:no]) + d.int0)
    e_full, e_fr = ehf(full, f_full, nocc), ehf(fd, f_fr, nocc - ncore)
    assert abs(e_full - ehf_ref) < EPS_REF
    assert abs(e_fr - e_full) < 1e-11
    # a dropped core-Fock term in int0 breaks the energy invariant (planted in a copy)
    bad = read_fcidump(os.path.join(GOLDEN, name))
    freeze_orbs_in_dump(bad, orbs)
    bad.int0 -= sum(f_full[c, c] - full.int1[c, c] for c in range(nocc) if c + 1 in orbs)
    assert abs(ehf(bad, f_fr, nocc - ncore) - e_full) > 1e-3


@pytest.mark.parametrize("name", ["H2O.FCIDUMP.gz", "H2O_ST1.FCIDUMP.gz"])
def test_write_fcidump_oracle_roundtrip(tmp_path, name):
    """write_fcidump (dump.jl:752-945) -> read_fcidump gives back every integral: the
    written set (8-fold unique, or particle-symmetric for ST > 0) covers the array."""
    from oracle.fcidump import write_fcidump as write_fd
    fd = read_fcidump(os.path.join(GOLDEN, name))
    for tol in (0.0, 1e-12):
        out = str(tmp_path / f"out{tol}.FCIDUMP")
        write_fd(fd, out, tol)
        back = read_fcidump(out)
        # IUHF, ST and III are not written when 0 (FDUMP_OPTIONAL, dump.jl:19)
        opt0 = lambda h: {k: v for k, v in h.items() if not (k in ("IUHF", "ST", "III") and v and v[0] == 0)}
        assert opt0(back.head) == opt0(fd.head)
        keep2 = np.where(np.abs(fd.int2) > tol, fd.int2, 0.0)
        keep1 = np.where(np.abs(fd.int1) > tol, fd.int1, 0.0)
        assert np.max(np.abs(back.int2 - keep2)) < 1e-15 and np.max(np.abs(back.int1 - keep1)) < 1e-15
        assert back.int0 == fd.int0


# ---------------------------------------------------------------- open-shell (UCCSD) oracle
def _uccsd_system(name, ms2):
    from oracle import uccsd as U
    fd = read_fcidump(os.path.join(GOLDEN, name))
    na, nb = (fd.nelec + ms2) // 2, (fd.nelec - ms2) // 2
    return U, U.USystem(fd.int1, fd.int2.transpose(0, 2, 1, 3), fd.int0, na, nb), fd


@pytest.mark.parametrize("dc", [False, True])
def test_uccsd_closed_shell_limit_equals_ccsd_residual(dc):
    """With nocca = noccb the spin-orbital UCCSD (UDCSD) residual's blocks are the closed-shell
    CCSD (DCSD) residual (calc_cc_resid, cc.jl:1343-1489): R1a = R1b = R1, R2ab = R2,
    R2a = R2b = R2 - R2[b,a,i,j], for random amplitudes (every term of both residuals
    contributes; a dropped or mis-signed term of the distinguishable-cluster correction fails)."""
    U, S, fd = _uccsd_system("H2O.FCIDUMP.gz", 0)
    EC = ECInfo.from_fcidump(fd)
    calc_fock_matrix(EC)
    assert abs(U.hf_energy(S) - calc_HF_energy(EC)) < 1e-12
    rng = np.random.default_rng(3)
    no, nv = EC.nocc, EC.nvirt
    T1 = 0.05 * rng.standard_normal((nv, no))
    T2 = 0.05 * rng.standard_normal((nv, nv, no, no))
    T2 = T2 + T2.transpose(1, 0, 3, 2)
    R1c, R2c = calc_cc_resid(EC, T1, T2, dc=dc)
    Ts = S.from_blocks(T1, T1, T2 - T2.transpose(1, 0, 2, 3), T2 - T2.transpose(1, 0, 2, 3), T2)
    R1a, R1b, R2a, R2b, R2ab = S.to_blocks(*U.calc_ucc_resid(S, *Ts, dc=dc))
    R2s = R2c - R2c.transpose(1, 0, 2, 3)
    for got, ref in ((R1a, R1c), (R1b, R1c), (R2ab, R2c), (R2a, R2s), (R2b, R2s)):
        assert np.max(np.abs(got - ref)) < 1e-12


@pytest.mark.parametrize("name,ms2,method,ref", [
    ("H2O_CATION.FCIDUMP.gz", 1, "uccsd", {"HF": -75.337282954481, "MP2c": -0.207727619864, "UCCSDc": -0.232320803220}),
    ("H2O_CATION.FCIDUMP.gz", 1, "uccd", {"UCCDc": -0.213495204514}),
    ("H2O_CATION.FCIDUMP.gz", 1, "udcsd", {"UDCSDc": -0.243223819179}),
    ("H2O.FCIDUMP.gz", 2, "uccsd", {"HF": -75.62407982361415, "MP2c": -0.22401008330020, "UCCSDc": -0.276560985613}),
])
def test_uccsd_reference_energies(name, ms2, method, ref):
    """test/h2o_cation.jl:4-29 (H2O+ doublet on ROHF orbitals: HF, MP2, UCCSD, UCCD, and DCSD,
    which the driver runs as UDCSD on an open-shell system) and
    test/h2o_triplet.jl:4-21 (H2O triplet, ms2 = 2: HF, MP2, UCCSD), at the reference's 1e-6."""
    from oracle import uccsd as U
    fd = read_fcidump(os.path.join(GOLDEN, name))
    na, nb = (fd.nelec + ms2) // 2, (fd.nelec - ms2) // 2
    e = U.ucc_driver(fd.int1, fd.int2.transpose(0, 2, 1, 3), fd.int0, na, nb, thr=1e-14, method=method)
    assert e["converged"]
    for k, v in ref.items():
        assert abs(e[k] - v) < EPS_REF, (k, e[k], v)


def test_uccsd_two_same_spin_electrons_is_fci():
    """Two alpha electrons (Sz = 1): UCCSD is exact, so E_corr = E_FCI - E_HF with E_FCI the lowest
    eigenvalue of H in the two-electron determinant space |pq>, p < q (Slater-Condon rules:
    H_(pq),(rs) = h_pr d_qs - h_ps d_qr - h_qr d_ps + h_qs d_pr + <pq||rs>)."""
    from oracle import uccsd as U
    from oracle.synthetic import make_df_system
    s = make_df_system(2, 6, 20, 17, scale=2.0)
    n = s.h.shape[0]
    eri = np.einsum('pqQ,rsQ->pqrs', s.B, s.B)
    S = U.USystem(s.h, eri, s.e0, 2, 0)
    E, _, _, conv, _ = U.calc_ucc(S, thr=1e-20, maxit=300)
    assert conv
    pairs = [(p, q) for p in range(n) for q in range(p + 1, n)]
    h = s.h
    A = lambda p, q, r, t: eri[p, r, q, t] - eri[p, t, q, r]
    d = lambda a, b: 1.0 if a == b else 0.0
    H = np.array([[h[p, r] * d(q, t) - h[p, t] * d(q, r) - h[q, r] * d(p, t) + h[q, t] * d(p, r) + A(p, q, r, t)
                   for (r, t) in pairs] for (p, q) in pairs])
    efci = float(np.linalg.eigvalsh(H)[0]) + s.e0
    ehf = H[0, 0] + s.e0
    assert abs(U.hf_energy(S) - ehf) < 1e-12
    assert abs(E - (efci - ehf)) < 1e-10


def test_unrestricted_triples_closed_shell_limit_and_reference():
    """(T) of UCCSD in spin-orbital form (oracle/uccsd.py, cc_triples.jl:297-640): with nocca = noccb
    it equals the closed-shell (T) oracle (calc_pertT_closed_shell, cc_triples.jl:39-148) for
    perturbed MP2 amplitudes (1e-14), and on the H2O triplet it gives the reference's
    UCCSD(T) energy of test/h2o_triplet.jl:8 (1e-6)."""
    from oracle import uccsd as U
    fd = read_fcidump(H2O)
    eri = fd.int2.transpose(0, 2, 1, 3)
    no = fd.nelec // 2
    S = U.USystem(fd.int1, eri, fd.int0, no, no)
    EC = ECInfo.from_fcidump(fd)
    calc_fock_matrix(EC)
    from oracle.ccsd import calc_MP2
    calc_MP2(EC)
    T1 = EC.files["T_vo"] + 0.01 * np.random.default_rng(2).standard_normal(EC.files["T_vo"].shape)
    T2 = EC.files["T_vvoo"]
    et, etb = triples.calc_pertT_closed_shell(EC, T1, T2)
    Ts = S.from_blocks(T1, T1, T2 - T2.transpose(1, 0, 2, 3), T2 - T2.transpose(1, 0, 2, 3), T2)
    eu, ebu = U.calc_pertT_unrestricted(S, *Ts)
    assert abs(eu - et) < 1e-14 and abs(ebu - etb) < 1e-14
    St = U.USystem(fd.int1, eri, fd.int0, (fd.nelec + 2) // 2, (fd.nelec - 2) // 2)
    E, T1t, T2t, conv, _ = U.calc_ucc(St, thr=1e-14)
    assert conv
    e3, _ = U.calc_pertT_unrestricted(St, T1t, T2t)
    assert abs(E + e3 - (-0.295388574918)) < EPS_REF
