"""Host-side FCIDUMP reader of the library (ecg_fcidump_*; reference: src/integrals/dump.jl).

The reference's own fixtures (test/files/H2O.FCIDUMP, H2O_ST1.FCIDUMP, stored
gzip-compressed under tests/golden) pin the text path against the oracle reader;
the NPY path (NPY2/NPY1 + ENUC, dump.jl:442-461) is pinned against the text path
on the same integrals written in the reference's NPY layout (<pq|rs> at
[p, q, r + s(s+1)/2], r <= s, Fortran order).  No GPU needed.
"""
import gzip
import os

import numpy as np
import pytest

from elemco_amd import lib
from elemco_amd.fcidump import read_fcidump

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _npy_layout(eri):
    """(pq|rs) -> the reference's NPY int2: <pq|rs> = (pr|qs) at [p, q, r + s(s+1)/2], r <= s."""
    n = eri.shape[0]
    int2 = np.zeros((n, n, n * (n + 1) // 2), order="F")
    for s in range(n):
        for r in range(s + 1):
            int2[:, :, r + s * (s + 1) // 2] = eri[:, r, :, s]
    return int2


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


@pytest.mark.parametrize("name", ["H2O.FCIDUMP.gz", "H2O_ST1.FCIDUMP.gz"])
def test_npy_integrals_match_text(tmp_path, name):
    fd = read_fcidump(os.path.join(GOLDEN, name))
    n = fd.norb
    np.save(tmp_path / "int2.npy", np.asfortranarray(_npy_layout(fd.eri)))
    np.save(tmp_path / "int1.npy", np.asfortranarray(fd.int1))
    st = fd.head["ST"][0]
    _write(tmp_path / "FCIDUMP", f" &FCI NORB={n},NELEC={fd.nelec},MS2=0,\n  ST={st},\n"
           f"  NPY2='int2.npy', NPY1=\"int1.npy\",\n  ENUC={fd.int0!r},\n /\n")
    got = read_fcidump(str(tmp_path / "FCIDUMP"))
    assert got.head["NPY2"] == ["int2.npy"] and got.head["NPY1"] == ["int1.npy"]
    assert np.array_equal(got.eri, fd.eri)
    assert np.array_equal(got.int1, fd.int1)
    assert got.int0 == fd.int0


def test_npy_missing_file_falls_back_to_text(tmp_path):
    """mmap_integrals (dump.jl:731-745): a missing NPY file means the integral lines are read."""
    with gzip.open(os.path.join(GOLDEN, "H2O.FCIDUMP.gz"), "rt") as f:
        text = f.read()
    head, rest = text.split("/", 1) if text.lstrip().startswith("&") else (None, None)
    assert head is not None
    _write(tmp_path / "FCIDUMP", head.rstrip() + "\n NPY2='absent.npy', NPY1='absent1.npy', ENUC=0.0,\n/" + rest)
    ref = read_fcidump(os.path.join(GOLDEN, "H2O.FCIDUMP.gz"))
    got = read_fcidump(str(tmp_path / "FCIDUMP"))
    assert np.array_equal(got.eri, ref.eri) and got.int0 == ref.int0


def test_npy_c_order_is_refused(tmp_path):
    n = 2
    np.save(tmp_path / "int2.npy", np.ascontiguousarray(np.ones((n, n, 3))))
    np.save(tmp_path / "int1.npy", np.asfortranarray(np.eye(n)))
    _write(tmp_path / "FCIDUMP", f"&FCI NORB={n},NELEC=2,\nNPY2='int2.npy',NPY1='int1.npy',ENUC=1.0,\n/\n")
    with pytest.raises(lib.ElemcoError, match="C-ordered"):
        read_fcidump(str(tmp_path / "FCIDUMP"))


def test_header_typing_and_case(tmp_path):
    """read_header (dump.jl:360-405): values typed by the first one, keys upper-cased."""
    _write(tmp_path / "FCIDUMP", "&FCI norb=2,NELEC=2,MS2=0,\n ORBSYM=1,1,\n isym=1,\n NAME='h2 test', X=0.5,1,\n"
           "&END\n 0.5 1 1 1 1\n 0.25 2 2 1 1\n -1.0 1 1 0 0\n -0.5 2 1 0 0\n 0.7 0 0 0 0\n")
    fd = read_fcidump(str(tmp_path / "FCIDUMP"))
    assert fd.head["NORB"] == [2] and fd.head["ORBSYM"] == [1, 1] and fd.head["ISYM"] == [1]
    assert fd.head["NAME"] == ["h2", "test"]          # blanks split values, quotes are stripped
    assert fd.head["X"] == [0.5, 1.0]
    assert fd.head["ST"] == [0] and fd.head["IUHF"] == [0]   # defaults
    assert fd.int0 == 0.7
    # 8-fold symmetry (ST = 0) of the two-electron lines, symmetric h
    assert fd.eri[1, 1, 0, 0] == fd.eri[0, 0, 1, 1] == 0.25
    assert fd.int1[1, 0] == fd.int1[0, 1] == -0.5 and fd.int1[0, 0] == -1.0


def test_simtra_keeps_particle_symmetry_only(tmp_path):
    _write(tmp_path / "FCIDUMP", "&FCI NORB=2,NELEC=2,ST=1,\n/\n 0.3 1 2 1 1\n 0.4 2 1 0 0\n 0.1 0 0 0 0\n")
    fd = read_fcidump(str(tmp_path / "FCIDUMP"))
    assert fd.simtra
    assert fd.eri[0, 1, 0, 0] == fd.eri[0, 0, 0, 1] == 0.3     # (12|11) = (11|12)
    assert fd.eri[1, 0, 0, 0] == 0.0                            # (21|11) not implied
    assert fd.int1[1, 0] == 0.4 and fd.int1[0, 1] == 0.0


@pytest.mark.parametrize("body,msg", [
    ("&FCI NORB=2,NELEC=2,\n/\n 0.5 3 1 1 1\n 0.1 0 0 0 0\n", "Index larger than norb"),
    ("&FCI NORB=2,NELEC=2,\n/\n 0.5 1 1 1 1\n", "No core energy"),
    ("&FCI NORB=2,NELEC=2,IUHF=1,\n/\n 0.1 0 0 0 0\n", "IUHF"),
    ("&FCI NORB=2,NELEC=2,\n 0.1 0 0 0 0\n", "not terminated"),
    ("&FCI NELEC=2,\n/\n 0.1 0 0 0 0\n", "NORB"),
])
def test_reader_errors(tmp_path, body, msg):
    _write(tmp_path / "FCIDUMP", body)
    with pytest.raises(lib.ElemcoError, match=msg):
        read_fcidump(str(tmp_path / "FCIDUMP"))


def test_df_of_indefinite_integrals_needing_two_by_two_pivots_is_an_error(tmp_path):
    """A supermatrix with zero diagonal and nonzero off-diagonal entries, (11|22) = (22|11) = 1
    and everything else 0 (eigenvalues +-1), has no 1x1 pivot: an error, not a wrong factor."""
    _write(tmp_path / "FCIDUMP", "&FCI NORB=2,NELEC=2,ST=1,\n/\n 1.0 1 1 2 2\n 0.0 0 0 0 0\n")
    fd = read_fcidump(str(tmp_path / "FCIDUMP"))
    with pytest.raises(lib.ElemcoError, match="2x2 pivots"):
        fd.df_factors()


# ---------------------------------------------------------------- freeze_orbs_in_dump, write_fcidump
@pytest.mark.parametrize("name,orbs", [("H2O.FCIDUMP.gz", [1]), ("H2O.FCIDUMP.gz", [1, 2, 23]),
                                       ("H2O_ST1.FCIDUMP.gz", [1]), ("H2O_ST1.FCIDUMP.gz", [2, 20]),
                                       ("H2O.FCIDUMP.gz", [1, 1, 0, 99])])
def test_freeze_matches_oracle(name, orbs):
    """ecg_fcidump_freeze = the oracle's freeze_orbs_in_dump (dumptools.jl:20-62); numbers
    outside 1..NORB and repeats match no orbital, as with the reference's intersections."""
    from oracle.fcidump import read_fcidump as oread, freeze_orbs_in_dump as ofreeze
    path = os.path.join(GOLDEN, name)
    fd = read_fcidump(path)
    nocc = fd.nelec // 2
    ncore, nvirt = fd.freeze_orbs(orbs)
    valid = {o for o in orbs if 1 <= o <= 23 + (name != "H2O.FCIDUMP.gz")}
    assert (ncore, nvirt) == (sum(o <= nocc for o in valid), sum(o > nocc for o in valid))
    od = ofreeze(oread(path), orbs)
    n = od.norb
    assert fd.norb == n and fd.nelec == od.nelec
    assert fd.head["ORBSYM"] == [1] * n and fd.head["NORB"] == [n]
    assert abs(fd.int0 - od.int0) < 1e-12
    assert np.max(np.abs(fd.int1 - od.int1)) < 1e-13
    # oracle int2[p,q,r,s] = <pq|rs> = (pr|qs)
    assert np.array_equal(fd.eri, od.int2.transpose(0, 2, 1, 3))


@pytest.mark.parametrize("name", ["H2O.FCIDUMP.gz", "H2O_ST1.FCIDUMP.gz"])
def test_write_matches_oracle(tmp_path, name):
    """ecg_fcidump_write writes the oracle writer's integral lines byte for byte, and a
    frozen dump read back from its file equals the dump in memory."""
    from oracle.fcidump import read_fcidump as oread, write_fcidump as owrite
    path = os.path.join(GOLDEN, name)
    fd = read_fcidump(path)
    a, b = str(tmp_path / "native.FCIDUMP"), str(tmp_path / "oracle.FCIDUMP")
    fd.write(a)
    owrite(oread(path), b)
    ta, tb = open(a).read(), open(b).read()
    assert ta.split("/\n", 1)[1] == tb.split("/\n", 1)[1]
    ha, hb = oread(a).head, oread(b).head
    assert ha == hb
    fd.freeze_orbs("1")
    fd.write(a, 0.0)
    back = read_fcidump(a)
    assert back.norb == fd.norb and back.nelec == fd.nelec and back.int0 == fd.int0
    assert np.max(np.abs(back.eri - fd.eri)) < 1e-15 and np.max(np.abs(back.int1 - fd.int1)) < 1e-15


def test_parse_orbstring():
    """parse_orbstring (src/infos/ecinfos.jl:606-660) without symmetry labels."""
    from elemco_amd.fcidump import parse_orbstring
    assert parse_orbstring("1-4+6-8") == parse_orbstring("1:4;6:8") == [1, 2, 3, 4, 6, 7, 8]
    assert parse_orbstring("-3+7") == [1, 2, 3, 7]
    assert parse_orbstring("5+2") == [2, 5]
    for bad in ("1-3+2", "1,2", "1.1"):
        with pytest.raises(ValueError):
            parse_orbstring(bad)


@pytest.mark.parametrize("name,orbs", [("H2O.FCIDUMP.gz", [1]), ("H2O.FCIDUMP.gz", [1, 2, 23]),
                                       ("H2O_ST1.FCIDUMP.gz", [1]), ("H2O_ST1.FCIDUMP.gz", [2, 20])])
def test_df_freeze_matches_four_index_freeze(name, orbs):
    """ecg_df_freeze on the exact DF factors of a dump = freeze_orbs_in_dump on its four-index
    integrals (the oracle's, itself pinned by the HF-energy and Fock invariants): h, core
    energy, and the integrals rebuilt from the frozen factors."""
    from oracle.fcidump import read_fcidump as oread, freeze_orbs_in_dump as ofreeze
    path = os.path.join(GOLDEN, name)
    fd = read_fcidump(path)
    BL, BR = fd.df_factors(1e-14)
    nocc = fd.nelec // 2
    h, bl, br, e0, no, nv = lib.df_freeze(fd.int1, BL, BR, fd.int0, nocc, orbs)
    od = ofreeze(oread(path), orbs)
    assert (no + nv, 2 * no) == (od.norb, od.nelec)
    assert abs(e0 - od.int0) < 1e-10
    assert np.max(np.abs(h - od.int1)) < 1e-10
    eri = np.einsum("pqQ,rsQ->pqrs", bl, bl if br is None else br)
    assert np.max(np.abs(eri - od.int2.transpose(0, 2, 1, 3))) < 1e-10
    assert (br is None) == (BR is None)


@pytest.mark.parametrize("name", ["H2O.FCIDUMP.gz", "H2O_ST1.FCIDUMP.gz"])
def test_dump_from_integrals_matches_file(tmp_path, name):
    """ecg_fcidump_from_integrals (the in-memory FDump of the reference, dump.jl:107-128) on the
    integrals read from a file: same header values the path uses, bitwise the same integrals,
    DF factors and written integral lines."""
    path = os.path.join(GOLDEN, name)
    f1 = lib.Fcidump(path)
    f2 = lib.Fcidump.from_integrals(f1.int1, f1.eri, f1.nelec, f1.ms2, f1.st, f1.int0)
    try:
        assert (f2.norb, f2.nelec, f2.ms2, f2.st, f2.int0) == (f1.norb, f1.nelec, f1.ms2, f1.st, f1.int0)
        assert np.array_equal(f2.int1, f1.int1) and np.array_equal(f2.eri, f1.eri)
        BL1, BR1 = f1.df_factors()
        BL2, BR2 = f2.df_factors()
        assert np.array_equal(BL1, BL2) and (BR1 is None) == (BR2 is None)
        if BR1 is not None:
            assert np.array_equal(BR1, BR2)
        a, b = str(tmp_path / "a.FCIDUMP"), str(tmp_path / "b.FCIDUMP")
        f1.write(a)
        f2.write(b)
        assert open(a).read().split("/\n", 1)[1] == open(b).read().split("/\n", 1)[1]
    finally:
        f1.close()
        f2.close()
