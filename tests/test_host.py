"""CPU checks of the product library: it loads without a GPU, exports every function the
header declares, and its host-only pieces (integral generator, reference-format
determinant list, single matrix elements, namelist reader, civector.dat / ciwvfcn.dat
writers) match the oracle, the reference build and the reference's file formats."""
import ctypes
import hashlib
import os
import subprocess

import numpy as np
import pytest


def test_library_exports_header(J):
    out = subprocess.run(["nm", "-D", "--defined-only", J.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    declared = J.header_functions()
    assert len(declared) >= 40
    missing = [f for f in declared if f not in syms]
    assert not missing, missing


def test_gpu_entry_points_fail_loudly_without_gpu(J):
    if J.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(J.JciError):
        J.Context(0)


def test_synth_integrals_recipe(J):
    n = 12
    m1, m2 = J.synth_integrals(n, 100)
    a1, a2 = J.synth_integrals(n, 100)
    assert np.array_equal(m1, a1) and np.array_equal(m2, a2)
    b1, _ = J.synth_integrals(n, 101)
    assert not np.array_equal(m1, b1)
    # h_ii = -2.0 + 0.25 i (1-based), packed index i(i+1)/2 - 1 for (i,i)
    for i in range(1, n + 1):
        assert m1[i * (i + 1) // 2 - 1] == -2.0 + 0.25 * i
    off = np.array([m1[i * (i - 1) // 2 + j - 1] for i in range(1, n + 1) for j in range(1, i)])
    assert 0.02 < off.std() < 0.08 and abs(off.mean()) < 0.03
    # (ij|kl) = sum_P L_ij L_kl is a Gram matrix over pairs: positive semidefinite
    p = n * (n + 1) // 2
    G = np.zeros((p, p))
    iu = np.tril_indices(p)
    G[iu] = m2
    G = G + G.T - np.diag(np.diag(G))
    assert np.linalg.eigvalsh(G).min() > -1e-12
    # rank <= 2n (number of L^P factors)
    assert np.sum(np.linalg.eigvalsh(G) > 1e-10) <= 2 * n
    # pinned bytes (generator stability across builds and machines)
    h = hashlib.sha256(m1.tobytes() + m2.tobytes()).hexdigest()
    golden = os.path.join(os.path.dirname(__file__), "golden", "synth_n12_seed100.sha256")
    if os.path.exists(golden):
        assert h == open(golden).read().strip()


@pytest.mark.parametrize("spec", [(2, 2, 4, 0, 4, 0), (3, 3, 7, 1, 3, 2), (3, 3, 8, 2, 2, 2),
                                  (4, 3, 9, 1, 4, 1), (4, 4, 10, 2, 4, 2), (4, 4, 12, 0, 12, 0)])
def test_citrunc_rtnlist_bitexact_vs_reference(J, spec):
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    from jayci_amd import dropin
    na, nb, n, nd, nact, xl = spec
    dl, info = dropin.citrunc_rtnlist(na, nb, n, 0, nd, nact, 0, xl)
    rs = ref.RefSpace(na, nb, n, nd, nact, xl)
    assert info["ndets"] == rs.ndets
    assert info["astr_len"] == rs.astr_len and info["bstr_len"] == rs.bstr_len
    mine = dropin.det_bytes(dl, rs.ndets)
    theirs = ctypes.string_at(ctypes.cast(rs.dlist, ctypes.c_void_p),
                              ctypes.sizeof(ref.Det) * rs.ndets)
    # the reference leaves byte2 and unused virtx slots as set by init_detlist; compare
    # the fields it defines: byte1, virtx[0..nvrtx), nvrtx, cas
    ours = (dropin.Det * rs.ndets).from_buffer_copy(mine)
    refs = (ref.Det * rs.ndets).from_buffer_copy(theirs)
    for k in range(rs.ndets):
        for s in ("astr", "bstr"):
            a, b = getattr(ours[k], s), getattr(refs[k], s)
            assert a.byte1 == b.byte1 and a.nvrtx == b.nvrtx
            assert list(a.virtx)[:a.nvrtx] == list(b.virtx)[:b.nvrtx]
        assert ours[k].cas == refs[k].cas
    dropin.free(dl)


def test_hmatels_host_vs_oracle(J, orc):
    from jayci_amd import dropin
    na, nb, n, nd, nact, xl = 3, 3, 8, 2, 2, 2
    m1, m2 = J.synth_integrals(n, 5)
    dl, info = dropin.citrunc_rtnlist(na, nb, n, 0, nd, nact, 0, xl)
    sp = orc.Space(na, nb, n, nd, nact, xl, 0)
    a, b = sp.dets()
    rng = np.random.default_rng(0)
    for i in rng.integers(0, info["ndets"], 25):
        cols, vals, _, _ = sp.row(int(i), m1, m2)
        for c, v in zip(cols[:40], vals[:40]):
            got = J.lib().hmatels(dl[int(i)], dl[int(c)], m1.ctypes.data_as(ctypes.c_void_p),
                                  m2.ctypes.data_as(ctypes.c_void_p), na, nb, nd + nact)
            assert abs(got - v) < 1e-13
        d = J.lib().hmatels(dl[int(i)], dl[int(i)], m1.ctypes.data_as(ctypes.c_void_p),
                            m2.ctypes.data_as(ctypes.c_void_p), na, nb, nd + nact)
        assert abs(d - sp.hij(a[i], b[i], a[i], b[i], m1, m2)[0]) < 1e-13
    dropin.free(dl)


def test_namelists(J, tmp_path, monkeypatch):
    from jayci_amd import dropin
    monkeypatch.chdir(tmp_path)
    assert dropin.readgeninput()["err"] == 1          # no jayci.in
    (tmp_path / "jayci.in").write_text(
        "&general\n electrons = 16, orbitals = 60,\n nfrozen=0, ndocc = 4, nactive = 8,\n"
        " xlevel = 2, printlvl = 1, printwvf = 1\n/\n"
        "&dalginfo\n maxiter=50, krymin=4, KRYMAX=12, nroots=3,\n"
        " prediagr=1, restol=1.0d-6, refdim=64 ! comment\n/\n")
    g = dropin.readgeninput()
    assert g == {"electrons": 16, "orbitals": 60, "nfrozen": 0, "ndocc": 4, "nactive": 8,
                 "xlevel": 2, "nfrzvirt": 0, "printlvl": 1, "printwvf": 1, "err": 0}
    d = dropin.readdaiinput()
    assert d == {"maxiter": 50, "krymin": 4, "krymax": 12, "nroots": 3, "prediagr": 1,
                 "refdim": 64, "restol": 1e-6, "err": 0}
    # defaults of source/readnamelist.f90:79-85 when keys are absent
    (tmp_path / "jayci.in").write_text("&general electrons=2, orbitals=2 /\n&dalginfo /\n")
    d = dropin.readdaiinput()
    assert d == {"maxiter": 20, "krymin": 3, "krymax": 8, "nroots": 1, "prediagr": 1,
                 "refdim": 3, "restol": 1e-5, "err": 0}


def test_civector_file_format(J, tmp_path, monkeypatch):
    """civector.dat layout of source/ioutil.c:200-209: "%d %d %d\\n", "%d\\n\\n", then per
    root "%15.8lf\\n\\n" and "%15.12lf\\n" per coefficient, and a blank line."""
    from jayci_amd import dropin
    monkeypatch.chdir(tmp_path)
    vec = np.array([[0.5, -0.25, 1.0 / 3.0], [1e-13, -0.9999999999994, 0.0]])
    ev = np.array([-1.5, 2.123456789])
    rc = J.lib().print_civectors(2, 1, 5, dropin.row_pointers(vec), 3, 2,
                                 ev.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    exp = ("3 5 3\n2\n\n" + "%15.8f\n\n" % -1.5 + "".join("%15.12f\n" % x for x in vec[0]) + "\n"
           + "%15.8f\n\n" % 2.123456789 + "".join("%15.12f\n" % x for x in vec[1]) + "\n")
    assert (tmp_path / "civector.dat").read_text() == exp
    back = np.zeros((2, 3))
    ev2 = np.zeros(2)
    rc = J.lib().read_civectors(2, 1, 5, dropin.row_pointers(back), 3, 2,
                                ev2.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    np.testing.assert_allclose(back, vec, atol=1e-12)
    # error codes of source/ioutil.c:215-258
    assert J.lib().read_civectors(2, 1, 6, dropin.row_pointers(back), 3, 2,
                                  ev2.ctypes.data_as(ctypes.c_void_p)) == 2
    assert J.lib().read_civectors(2, 1, 5, dropin.row_pointers(np.zeros((3, 3))), 3, 3,
                                  np.zeros(3).ctypes.data_as(ctypes.c_void_p)) == 12
    os.remove(tmp_path / "civector.dat")
    assert J.lib().read_civectors(2, 1, 5, dropin.row_pointers(back), 3, 2,
                                  ev2.ctypes.data_as(ctypes.c_void_p)) == 1


def test_ciwvfcn_file_format(J, tmp_path, monkeypatch):
    """ciwvfcn.dat layout of source/write_wavefunction.c:39-64."""
    from jayci_amd import dropin
    monkeypatch.chdir(tmp_path)
    (tmp_path / "jayci.in").write_text("&general electrons=4, orbitals=5, ndocc=1, nactive=2,"
                                       " xlevel=2 /\n")
    dl, info = dropin.citrunc_rtnlist(2, 2, 5, 0, 1, 2, 0, 2)
    N = info["ndets"]
    vec = np.linspace(-1, 1, N).reshape(1, N)
    ev = np.array([-3.25])
    J.lib().write_wavefunction(dl, N, 1, dropin.row_pointers(vec),
                               ev.ctypes.data_as(ctypes.c_void_p), 5, 3, 4)
    lines = (tmp_path / "ciwvfcn.dat").read_text().split("\n")
    assert lines[0] == "4 5" and lines[1] == "0 1 2 0" and lines[2] == "2"
    assert lines[3] == "4 5 3 %d" % N and lines[4] == ""
    d0 = dl[0]
    assert lines[5] == "%d %d %d %d %d  %d %d %d %d %d  %d" % (
        d0.astr.byte1, d0.astr.byte2, d0.astr.virtx[0], d0.astr.virtx[1], d0.astr.nvrtx,
        d0.bstr.byte1, d0.bstr.byte2, d0.bstr.virtx[0], d0.bstr.virtx[1], d0.bstr.nvrtx, d0.cas)
    assert lines[5 + N] == "" and lines[6 + N] == "" and lines[7 + N] == "1"
    assert lines[9 + N] == "%15.8f" % -3.25
    assert lines[11 + N] == "%15.12f" % vec[0, 0]
    dropin.free(dl)


def test_jacobi_host(J):
    rng = np.random.default_rng(1)
    for n in (1, 2, 9, 64):
        A = rng.standard_normal((n, n))
        A = A + A.T
        ev, V = J.jacobi(A)
        np.testing.assert_allclose(ev, np.linalg.eigvalsh(A), atol=1e-12 * n)
        np.testing.assert_allclose(V @ V.T, np.eye(n), atol=1e-12 * n)
