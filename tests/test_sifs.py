"""SIFS integral files (the reference's readmoints, source/readmoints.f90) through the C ABI.

The fixtures under tests/golden/sifs/ were written by colib's own header writer and
label / dword packers (oracle/_ref/sifs_ref, see tests/golden/make_sifs_fixtures.py), so
the bit layout the reader decodes is the reference library's: 8-bit labels (ifmt 0) and
10-bit labels (ifmt 1), 1-e arrays over several records with frozen-core values, an
overlap array that must stay out of H1, two symmetry blocks, dipole arrays, and the
14-orbital synthetic integrals, which must come back bit for bit.  Host-only file I/O.
"""
import ctypes
import os
import shutil

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sifs")
CASES = ["a_fmt8", "b_fmt10_synth14", "c_sym2", "d_dipole"]


def _fixture(name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    return os.path.join(GOLD, name + ".sifs"), d


@pytest.mark.parametrize("name", CASES)
def test_fixture_arrays_bitexact(J, name):
    path, d = _fixture(name)
    n = int(d["nbft"])
    m1, m2, nuc, fc = J.read_moints(path, n, 1)
    assert np.array_equal(m1, d["h1"])
    assert np.array_equal(m2, d["moints2"])
    assert nuc == d["en"][0] and fc == d["en"][1]


def test_overlap_and_frozen_core_semantics(J):
    # a_fmt8: the overlap array carries fcore 0.125 and must not count; kinetic (-2.5) and
    # potential (1.75) contribute the value of their last record each
    path, d = _fixture("a_fmt8")
    _, _, nuc, fc = J.read_moints(path, 6, 1)
    assert (nuc, fc) == (9.25, -0.75)


def test_synthetic_integrals_round_trip(J):
    path, _ = _fixture("b_fmt10_synth14")
    m1, m2, nuc, fc = J.read_moints(path, 14, 1)
    h, e = J.synth_integrals(14, 7)
    assert np.array_equal(m1, h) and np.array_equal(m2, e)
    assert (nuc, fc) == (10.0, 0.0)


def test_dipole_arrays(J):
    path, d = _fixture("d_dipole")
    m1, _, nuc, fc = J.read_moints(path, 5, 3)
    assert np.array_equal(m1, d["dipoles"])
    assert nuc == 7.5 and fc == 0.0


def test_other_type_reads_only_two_electron(J):
    path, d = _fixture("a_fmt8")
    m1, m2, nuc, fc = J.read_moints(path, 6, 2)
    assert not m1.any() and np.array_equal(m2, d["moints2"]) and fc == 0.0


def test_errors(J, tmp_path):
    path, _ = _fixture("a_fmt8")
    raw = open(path, "rb").read()
    with pytest.raises(J.JciError, match="Could not open"):
        J.read_moints(str(tmp_path / "none"), 6, 1)
    # truncated inside the 2-e records
    p = tmp_path / "trunc"
    p.write_bytes(raw[:-100])
    with pytest.raises(J.JciError):
        J.read_moints(str(p), 6, 1)
    # header 1 only
    p.write_bytes(raw[:4 + 56 + 4])
    with pytest.raises(J.JciError, match="header 2"):
        J.read_moints(str(p), 6, 1)
    # a record marker that does not match its trailer
    bad = bytearray(raw)
    bad[0] ^= 1
    p.write_bytes(bytes(bad))
    with pytest.raises(J.JciError, match="header 1"):
        J.read_moints(str(p), 6, 1)
    # moints arrays too short for the file's orbitals
    with pytest.raises(J.JciError, match="shorter"):
        J.read_moints(path, 5, 1)


def test_readmointegrals_dropin(J, tmp_path):
    """The reference's C entry (source/ioutil.c:427) with its output arguments."""
    src, d = _fixture("a_fmt8")
    dst = tmp_path / "moints"
    shutil.copy(src, dst)
    L = J.lib()
    f = L.readmointegrals
    f.restype = None
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                  ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                  ctypes.POINTER(ctypes.c_double)]
    m1 = np.zeros(21)
    m2 = np.zeros(231)
    nuc, fc = ctypes.c_double(), ctypes.c_double()
    f(m1.ctypes.data, m2.ctypes.data, 1, 6, os.fsencode(str(dst)), 21, 231, ctypes.byref(nuc),
      ctypes.byref(fc))
    assert np.array_equal(m1, d["h1"]) and np.array_equal(m2, d["moints2"])
    assert (nuc.value, fc.value) == (9.25, -0.75)
