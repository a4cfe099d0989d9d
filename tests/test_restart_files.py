"""The MPI driver's binary vector files (civec, ci.restart, hvvec) through the C ABI.

Layout, from mpi_source/mpi_utilities.c:353-415 (writer) and :426-531 (reader): a 4-byte
int dim, a 4-byte int len, then dim vectors of len doubles in native byte order, with no
Fortran record markers.  The checks decode the bytes independently with numpy (dtype
int32 then float64), compare the file size with 8 + 8*dim*len, and exercise the reader's
error codes (2 no file, 3/4 short header, 5 dimensions differ, 6 short data).  These
calls are host-only file I/O; no GPU is used.
"""
import os

import numpy as np
import pytest


def test_layout_and_round_trip(J, tmp_path):
    v = np.random.default_rng(3).standard_normal((3, 11))
    f = tmp_path / "civec"
    J.write_vectors_ufmt(str(f), v)
    raw = f.read_bytes()
    assert len(raw) == 8 + 8 * 3 * 11
    assert np.frombuffer(raw[:8], dtype=np.int32).tolist() == [3, 11]
    assert np.array_equal(np.frombuffer(raw[8:], dtype=np.float64).reshape(3, 11), v)
    assert np.array_equal(J.read_vectors_ufmt(str(f), 11, 3), v)


def test_reader_error_codes(J, tmp_path):
    v = np.arange(12, dtype=np.float64).reshape(2, 6)
    f = tmp_path / "ci.restart"
    J.write_vectors_ufmt(str(f), v)
    raw = f.read_bytes()
    cases = [(tmp_path / "missing", raw, None, 2),
             (tmp_path / "h0", raw[:2], None, 3),
             (tmp_path / "h1", raw[:6], None, 4),
             (f, raw, (6, 3), 5),
             (f, raw, (5, 2), 5),
             (tmp_path / "short", raw[:8 + 8 * 7], None, 6)]
    for path, data, dims, code in cases:
        if path.name not in ("missing", "ci.restart"):
            path.write_bytes(data)
        length, dim = dims if dims else (6, 2)
        with pytest.raises(J.JciError) as e:
            J.read_vectors_ufmt(str(path), length, dim)
        assert e.value.code == code, (path.name, e.value)


def test_empty_and_single(J, tmp_path):
    f = tmp_path / "hvvec"
    J.write_vectors_ufmt(str(f), np.zeros((2, 5)))
    assert os.path.getsize(f) == 8 + 80
    assert not J.read_vectors_ufmt(str(f), 5, 2).any()
    g = tmp_path / "one"
    J.write_vectors_ufmt(str(g), np.array([1.5]))
    assert np.frombuffer(g.read_bytes()[:8], dtype=np.int32).tolist() == [1, 1]
