"""The reference's driver sequence through the drop-in C ABI, on the GPU.

source/execute_ci_calculation.c:206-262 (with run_jayci.c:77 upstream):
citrunc_rtnlist -> generate_cimap -> compute_hdgls -> compute_hv -> dvdalg (predr 1, then
predr 2 from the civector.dat that print_civectors wrote) -> print_civectors /
write_wavefunction -> deallocate_cimap.

Checks, each against something independent of the product:
* determinant list: the oracle's determinants in the same order (serial truncation rule);
* generate_cimap: the section lists equal the runs of a numpy popcount screen
  (excitation degree <= 2 over j >= i, the diagonal first; source/cimapping.c:23-65,92-141);
* compute_hdgls: the oracle's diagonal;
* compute_hv: c += H v (source/action_util.c:121-155) from a nonzero c, twice, against
  c0 + k * oracle sigma;
* dvdalg: eigenvalues + totfrze against dense (or sparse Lanczos) eigenvalues of the
  oracle's H, and residual norms of the returned vectors through the oracle's sigma;
* print_civectors / write_wavefunction: the files equal the layouts of source/ioutil.c:
  184-211 and source/write_wavefunction.c:39-64 filled with the returned numbers.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (alpha, beta, orbitals, ndocc, nactv, xlvl): serial-rule spaces as citrunc_rtnlist builds them
CHAIN_SPACES = [
    (2, 2, 4, 0, 4, 0),        # FCI, 36 dets
    (3, 3, 7, 1, 3, 2),        # docc / active / virtual, serial rule
    (4, 3, 9, 1, 4, 1),        # odd electron count, xlvl 1
    (3, 3, 8, 0, 8, 0),        # FCI 3a3b/8, 3136 dets
    (4, 4, 10, 0, 10, 0),      # FCI 4a4b/10, 44,100 dets
]

ENUC = 10.0


def _c(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def popcount_sections(a, b, rows):
    """Section lists by brute force: for row i, the j >= i whose alpha and beta strings
    differ by at most two electrons in all, as maximal runs of offsets j - i."""
    out = []
    for i in rows:
        i = int(i)
        deg = (np.bitwise_count(a[i:] ^ a[i]).astype(np.int32)
               + np.bitwise_count(b[i:] ^ b[i]).astype(np.int32))
        js = np.nonzero(deg <= 4)[0]                  # offsets, 0 (the diagonal) first
        brk = np.nonzero(np.diff(js) != 1)[0]
        starts = js[np.r_[0, brk + 1]]
        ends = js[np.r_[brk, len(js) - 1]]
        out.append(list(zip(starts.tolist(), ends.tolist())))
    return out


def sections_of(hmap, rows):
    from jayci_amd import dropin
    rm = (dropin.RowMap * (int(max(rows)) + 1)).from_address(hmap)
    out = []
    for i in rows:
        s = rm[int(i)].sec
        lst = []
        while s:
            lst.append((s.contents.first, s.contents.last))
            s = s.contents.next
        out.append(lst)
    return out


def reference_eigs(osp, m1, m2, k):
    """Lowest k eigenvalues of the oracle's H: dense for small spaces, Lanczos otherwise."""
    N = osp.ndets
    if N <= 4000:
        return np.linalg.eigvalsh(osp.dense_h(m1, m2))[:k]
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl
    rows, cols, vals = [], [], []
    for i in range(N):
        c, v, _, _ = osp.row(i, m1, m2)
        rows.append(np.full(len(c), i))
        cols.append(c)
        vals.append(v)
    hd = osp.hdiag(m1, m2)
    H = sps.csr_matrix((np.concatenate(vals + [hd]),
                        (np.concatenate(rows + [np.arange(N)]),
                         np.concatenate(cols + [np.arange(N)]))), shape=(N, N))
    w = spl.eigsh(H, k=k, which="SA", tol=1e-13, ncv=max(4 * k, 24))[0]
    return np.sort(w)


def civector_text(elec, orbs, N, ev, vec):
    s = "%d %d %d\n%d\n\n" % (elec, orbs, N, len(ev))
    for r in range(len(ev)):
        s += "%15.8f\n\n" % ev[r] + "".join("%15.12f\n" % x for x in vec[r]) + "\n"
    return s


@pytest.mark.parametrize("spec", CHAIN_SPACES, ids=lambda s: "%da%db_%d_d%d_a%d_x%d" % s)
def test_reference_driver_chain(J, orc, ctx, spec, tmp_path, monkeypatch):
    from jayci_amd import dropin
    L = J.lib()
    na, nb, norbs, ndocc, nactv, xlvl = spec
    ninto = ndocc + nactv
    monkeypatch.chdir(tmp_path)
    print("chain %s" % (spec,), flush=True)
    seed = 7000 + norbs
    m1, m2 = J.synth_integrals(norbs, seed)

    # citrunc_rtnlist (source/citruncate.c:123): the oracle's list, same order
    dl, info = dropin.citrunc_rtnlist(na, nb, norbs, 0, ndocc, nactv, 0, xlvl)
    N = info["ndets"]
    osp = orc.Space(na, nb, norbs, ndocc, nactv, xlvl, 0)
    assert osp.ndets == N
    oa, ob = osp.dets()
    arr = (dropin.Det * N).from_address(ctypes.addressof(dl.contents))
    ga = np.array([dropin.occ_mask(d.astr) for d in arr], dtype=np.uint64)
    gb = np.array([dropin.occ_mask(d.bstr) for d in arr], dtype=np.uint64)
    assert np.array_equal(ga, oa) and np.array_equal(gb, ob)

    # generate_cimap: section lists bit-exact against the popcount screen
    rc, hmap = dropin.generate_cimap(dl, N, ninto)
    assert rc == 0
    # every row of the small spaces, 3000 sampled rows (and the first and last) of the larger
    rows = np.arange(N) if N <= 4000 else np.unique(
        np.r_[0, N - 1, np.random.default_rng(3).integers(0, N, size=3000)])
    secs = sections_of(hmap, rows)
    assert secs == popcount_sections(oa, ob, rows)
    rm = (dropin.RowMap * N).from_address(hmap)
    assert all(rm[int(i)].nsec == len(sx) for i, sx in zip(rows, secs))

    # compute_hdgls
    hd = np.zeros(N)
    L.compute_hdgls(dl, N, _c(m1), _c(m2), na, nb, _c(hd), ninto)
    np.testing.assert_allclose(hd, osp.hdiag(m1, m2), rtol=0, atol=1e-12 * np.abs(hd).max())

    # compute_hv accumulates: c = c0 + H v, then c0 + 2 H v
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(N)
    c0 = rng.standard_normal(N)
    c = c0.copy()
    sig = osp.sigma(m1, m2, v)
    L.compute_hv(dl, N, _c(m1), _c(m2), na, nb, _c(v), _c(c), ninto, hmap)
    scale = np.abs(sig).max()
    np.testing.assert_allclose(c, c0 + sig, rtol=0, atol=1e-12 * scale)
    L.compute_hv(dl, N, _c(m1), _c(m2), na, nb, _c(v), _c(c), ninto, hmap)
    np.testing.assert_allclose(c, c0 + 2 * sig, rtol=0, atol=2e-12 * scale)

    # dvdalg predr 1: krymin roots so that civector.dat can seed predr 2
    krymin = min(3, N)
    nroots = krymin
    krymax = min(12, N)
    refdim = min(64, N)
    wref = reference_eigs(osp, m1, m2, nroots) + ENUC
    civ = np.zeros((nroots, N))
    cival = np.zeros(nroots)
    restol = 1e-7
    rc = L.dvdalg(dl, N, _c(m1), _c(m2), na, nb, _c(hd), ninto, ENUC, 300, krymin, krymax,
                  nroots, restol, hmap, dropin.row_pointers(civ), _c(cival), 1, 0, norbs, refdim)
    assert rc == 0
    np.testing.assert_allclose(cival, wref, rtol=0, atol=1e-8)
    for r in range(nroots):
        res = osp.sigma(m1, m2, civ[r]) - (cival[r] - ENUC) * civ[r]
        assert np.linalg.norm(res) < 10 * restol
        assert abs(np.linalg.norm(civ[r]) - 1.0) < 1e-10

    # print_civectors: the reference layout with these numbers
    assert L.print_civectors(na, nb, norbs, dropin.row_pointers(civ), N, nroots, _c(cival)) == 0
    assert (tmp_path / "civector.dat").read_text() == civector_text(na + nb, norbs, N, cival, civ)

    # dvdalg predr 2 (source/initguess_roldv.c:20): krymin vectors from that file
    n2 = max(1, nroots - 1)
    civ2 = np.zeros((n2, N))
    cival2 = np.zeros(n2)
    rc = L.dvdalg(dl, N, _c(m1), _c(m2), na, nb, _c(hd), ninto, ENUC, 300, krymin, krymax,
                  n2, restol, hmap, dropin.row_pointers(civ2), _c(cival2), 2, 0, norbs, refdim)
    assert rc == 0
    np.testing.assert_allclose(cival2, wref[:n2], rtol=0, atol=1e-8)
    # a file with fewer roots than krymin: the reader's code 10 + roots becomes 100
    L.print_civectors(na, nb, norbs, dropin.row_pointers(civ2), N, n2, _c(cival2))
    if n2 < krymin:
        rc = L.dvdalg(dl, N, _c(m1), _c(m2), na, nb, _c(hd), ninto, ENUC, 300, krymin, krymax,
                      n2, restol, hmap, dropin.row_pointers(civ2), _c(cival2), 2, 0, norbs,
                      refdim)
        assert rc == 100

    # write_wavefunction (reads &general from jayci.in)
    (tmp_path / "jayci.in").write_text(
        "&general\n electrons = %d, orbitals = %d, nfrozen = 0, ndocc = %d, nactive = %d,\n"
        " xlevel = %d, printwvf = 1\n/\n" % (na + nb, norbs, ndocc, nactv, xlvl))
    L.write_wavefunction(dl, N, nroots, dropin.row_pointers(civ), _c(cival), norbs, ninto,
                         na + nb)
    head = "%d %d\n0 %d %d 0\n%d\n%d %d %d %d\n\n" % (na + nb, norbs, ndocc, nactv, xlvl,
                                                   na + nb, norbs, ninto, N)
    body = "".join("%d %d %d %d %d  %d %d %d %d %d  %d\n" % (
        d.astr.byte1, d.astr.byte2, d.astr.virtx[0], d.astr.virtx[1], d.astr.nvrtx,
        d.bstr.byte1, d.bstr.byte2, d.bstr.virtx[0], d.bstr.virtx[1], d.bstr.nvrtx, d.cas)
        for d in arr)
    tail = "\n\n%d\n\n" % nroots + "".join(
        "%15.8f\n\n" % cival[r] + "".join("%15.12f\n" % x for x in civ[r]) + "\n"
        for r in range(nroots))
    assert (tmp_path / "ciwvfcn.dat").read_text() == head + body + tail

    # deallocate_cimap frees the lists and hmap, returns the row count
    assert dropin.deallocate_cimap(hmap, N) == N
    dropin.free(dl)


def test_compute_hv_sees_integrals_changed_in_place(J, orc, ctx):
    """The matrix elements cached for an hmap are rebuilt when the integrals change in
    place (same arrays, new values)."""
    from jayci_amd import dropin
    L = J.lib()
    na, nb, norbs = 3, 3, 8
    m1, m2 = J.synth_integrals(norbs, 11)
    dl, info = dropin.citrunc_rtnlist(na, nb, norbs, 0, 0, norbs, 0, 0)
    N = info["ndets"]
    osp = orc.Space(na, nb, norbs, 0, norbs, 0, 0)
    rc, hmap = dropin.generate_cimap(dl, N, norbs)
    assert rc == 0
    v = np.random.default_rng(1).standard_normal(N)
    c = np.zeros(N)
    L.compute_hv(dl, N, _c(m1), _c(m2), na, nb, _c(v), _c(c), norbs, hmap)
    np.testing.assert_allclose(c, osp.sigma(m1, m2, v), rtol=0, atol=1e-12 * np.abs(c).max())
    # change one two-electron integral far from the start of the array, in place
    k = len(m2) - 3
    m2[k] += 0.125
    c[:] = 0.0
    L.compute_hv(dl, N, _c(m1), _c(m2), na, nb, _c(v), _c(c), norbs, hmap)
    ref = osp.sigma(m1, m2, v)
    np.testing.assert_allclose(c, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    dropin.deallocate_cimap(hmap, N)
    dropin.free(dl)


def test_pdavidson_restart_with_buffer_roots(J, orc, ctx, tmp_path):
    """prediagr 3 with more tracked pairs (nroots + nbuffer = 4) than the krymin = 3 vectors
    ci.restart holds: the solver takes the file's 3 vectors and fills the fourth from the
    block guess (mpi_source/pdavidson.c:434-448 reads exactly krymin vectors)."""
    spec = (3, 3, 8, 0, 8, 0, 1)
    na, nb, n, nd, nact, xl, rule = spec
    m1, m2 = J.synth_integrals(n, 23)
    gs = J.Space(ctx, na, nb, n, nd, nact, xl, rule)
    h = J.Hamiltonian(gs, m1, m2, J.DIRECT_TABLES)
    w = np.linalg.eigvalsh(orc.Space(*spec).dense_h(m1, m2))[:2]
    kw = dict(nroots=2, krymin=3, krymax=12, refdim=32, restol=1e-8, nbuffer=2)
    ev, st = h.pdavidson(1, str(tmp_path), maxiter=2, **kw)
    assert st["rc"] == 1 and (tmp_path / "ci.restart").exists()
    back = J.read_vectors_ufmt(str(tmp_path / "ci.restart"), gs.ndets, 3)
    assert back.shape == (3, gs.ndets)
    ev, st = h.pdavidson(3, str(tmp_path), maxiter=300, **kw)
    assert st["rc"] == 0
    np.testing.assert_allclose(ev, w, rtol=0, atol=1e-8)
    # the Python wrapper rejects a user guess with fewer than krymin vectors
    with pytest.raises(ValueError):
        h.solve(nroots=2, krymin=3, guess=J.GUESS_USER, guess_vectors=back[:2])
    h.close()
    gs.close()


def test_foreign_list_takes_the_pair_list_path(J, orc, ctx):
    """A determinant list citrunc_rtnlist did not return (here a copy of one in a buffer of
    the caller's) is not matched to a registered space: compute_hv and dvdalg build its
    connectivity from the list itself and give the same results."""
    from jayci_amd import dropin
    L = J.lib()
    na, nb, norbs, nd, nact, xl = 3, 3, 7, 1, 3, 2
    m1, m2 = J.synth_integrals(norbs, 41)
    dl, info = dropin.citrunc_rtnlist(na, nb, norbs, 0, nd, nact, 0, xl)
    N = info["ndets"]
    copy = (dropin.Det * N)()
    ctypes.memmove(copy, dl, ctypes.sizeof(dropin.Det) * N)
    own = ctypes.cast(copy, ctypes.POINTER(dropin.Det))
    osp = orc.Space(na, nb, norbs, nd, nact, xl, 0)
    v = np.random.default_rng(5).standard_normal(N)
    ref = osp.sigma(m1, m2, v)
    for lst in (dl, own):
        rc, hmap = dropin.generate_cimap(lst, N, nd + nact)
        assert rc == 0
        c = np.zeros(N)
        L.compute_hv(lst, N, _c(m1), _c(m2), na, nb, _c(v), _c(c), nd + nact, hmap)
        np.testing.assert_allclose(c, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
        civ = np.zeros((1, N))
        cival = np.zeros(1)
        hd = osp.hdiag(m1, m2)
        rc = L.dvdalg(lst, N, _c(m1), _c(m2), na, nb, _c(hd), nd + nact, ENUC, 300, 3, 12, 1,
                      1e-8, hmap, dropin.row_pointers(civ), _c(cival), 1, 0, norbs, 64)
        assert rc == 0
        np.testing.assert_allclose(cival, np.linalg.eigvalsh(osp.dense_h(m1, m2))[:1] + ENUC,
                                   rtol=0, atol=1e-8)
        assert dropin.deallocate_cimap(hmap, N) == N
    dropin.free(dl)
