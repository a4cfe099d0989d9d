"""GPU parity: the HIP path (through the C ABI) against the oracle on the same inputs.

Integer work (strings, determinant order, connected pairs, phases, degrees) is compared
bit for bit; FP64 matrix elements and sigma to 1e-12 relative; FP32 sigma to 2e-5
relative of the vector scale; eigenvalues to 1e-8 Ha (restol 1e-7 gives errors ~1e-13).
"""
import numpy as np
import pytest

from conftest import SMALL_SPACES, rand_vectors

pytestmark = pytest.mark.gpu


def _spaces(J, orc, ctx, spec, seed=11):
    na, nb, n, nd, nact, xl, rule = spec
    m1, m2 = J.synth_integrals(n, seed)
    gs = J.Space(ctx, na, nb, n, nd, nact, xl, rule)
    os_ = orc.Space(na, nb, n, nd, nact, xl, rule)
    return gs, os_, m1, m2


@pytest.mark.parametrize("spec", SMALL_SPACES)
def test_strings_groups_blocks_dets(J, orc, ctx, spec):
    gs, os_, _, _ = _spaces(J, orc, ctx, spec)
    assert gs.ndets == os_.ndets
    for spin in (0, 1):
        assert np.array_equal(gs.strings(spin), os_.strings(spin))
        assert np.array_equal(gs.groups(spin), os_.groups(spin))
    assert np.array_equal(gs.blocks(), os_.blocks())
    a, b = gs.dets()
    oa, ob = os_.dets()
    assert np.array_equal(a, oa) and np.array_equal(b, ob)
    # addressing round trip on a sample
    rng = np.random.default_rng(0)
    for I in rng.integers(0, gs.ndets, size=min(20, gs.ndets)):
        assert gs.index(a[I], b[I]) == I


@pytest.mark.parametrize("spec", SMALL_SPACES)
def test_hdiag(J, orc, ctx, spec):
    gs, os_, m1, m2 = _spaces(J, orc, ctx, spec)
    h = J.Hamiltonian(gs, m1, m2)
    np.testing.assert_allclose(h.hdiag(), os_.hdiag(m1, m2), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("spec", SMALL_SPACES)
def test_rows_bitexact(J, orc, ctx, spec):
    gs, os_, m1, m2 = _spaces(J, orc, ctx, spec)
    h = J.Hamiltonian(gs, m1, m2)
    N = gs.ndets
    rows = np.unique(np.r_[0, N - 1, np.random.default_rng(1).integers(0, N, size=min(40, N))])
    got = h.rows(rows)
    for I, (c, v, ph, dg) in zip(rows, got):
        oc, ov, oph, odg = os_.row(int(I), m1, m2)
        assert np.array_equal(c, oc), "columns differ in row %d" % I
        assert np.array_equal(ph, oph), "phases differ in row %d" % I
        assert np.array_equal(dg, odg), "degrees differ in row %d" % I
        np.testing.assert_allclose(v, ov, rtol=1e-12, atol=1e-13)


METHODS64 = ["SIGMA_PAIRS64", "SIGMA_PAIRS_SYM64", "SIGMA_DIRECT64", "SIGMA_FAST64", "SIGMA_FCI64",
             "SIGMA_GEMM64"]
METHODS32 = ["SIGMA_PAIRS32", "SIGMA_DIRECT32", "SIGMA_FAST32", "SIGMA_FCI32", "SIGMA_GEMM32"]
FCI_ONLY = ("SIGMA_FCI64", "SIGMA_FCI32", "SIGMA_GEMM64", "SIGMA_GEMM32")


@pytest.mark.parametrize("spec", SMALL_SPACES)
def test_sigma_all_methods(J, orc, ctx, spec):
    gs, os_, m1, m2 = _spaces(J, orc, ctx, spec)
    full = J.Hamiltonian(gs, m1, m2, J.PAIRS_FP64 | J.PAIRS_FP32 | J.DIRECT_TABLES)
    upper = J.Hamiltonian(gs, m1, m2, J.PAIRS_FP64 | J.PAIRS_UPPER)
    N = gs.ndets
    X = rand_vectors(N, 3, 5)
    ref = os_.sigma(m1, m2, X)
    scale = np.abs(ref).max()
    fci = spec[3] == 0 and spec[4] == spec[2]
    for name in METHODS64:
        if name in FCI_ONLY and not fci:
            continue
        meth = getattr(J, name)
        h = upper if name == "SIGMA_PAIRS_SYM64" else full
        k = 1 if name == "SIGMA_PAIRS_SYM64" else 3
        y = h.sigma(X[:k], method=meth)
        np.testing.assert_allclose(y, ref[:k], rtol=0, atol=1e-12 * scale, err_msg=name)
    for name in METHODS32:
        if name in FCI_ONLY and not fci:
            continue
        y = full.sigma(X, method=getattr(J, name))
        np.testing.assert_allclose(y, ref, rtol=0, atol=2e-5 * scale, err_msg=name)


@pytest.mark.parametrize("spec,nvec", [((4, 4, 10, 0, 10, 0, 1), 1),   # 44,100 dets
                                       ((5, 4, 11, 0, 11, 0, 1), 3),   # open shell, 152,460
                                       ((5, 5, 11, 0, 11, 0, 1), 2),   # 213,444
                                       ((1, 6, 15, 0, 15, 0, 1), 2)])  # 75,075
def test_sigma_dense_string_larger(J, orc, ctx, spec, nvec):
    """The dense-string MFMA sigma on full-CI spaces whose string counts and orbital-pair
    counts span several 64-row product tiles and many column tiles of the intermediate.
    1a6b/15: 5,005 beta strings (rows longer than the kernel's LDS row buffer, so the
    gathers read the vector in device memory), 120 orbital pairs padded to 128 (the deepest
    reduction) and 60-entry beta lists; 15 alpha strings."""
    gs, os_, m1, m2 = _spaces(J, orc, ctx, spec, seed=21)
    h = J.Hamiltonian(gs, m1, m2, J.DIRECT_TABLES)
    X = rand_vectors(gs.ndets, nvec, 8)
    ref = os_.sigma(m1, m2, X)
    scale = np.abs(ref).max()
    y = h.sigma(X, method=J.SIGMA_GEMM64)
    np.testing.assert_allclose(y, ref, rtol=0, atol=1e-12 * scale)
    y = h.sigma(X, method=J.SIGMA_GEMM32)
    np.testing.assert_allclose(y, ref, rtol=0, atol=2e-5 * scale)


@pytest.mark.parametrize("spec,nvec", [((3, 3, 8, 0, 8, 0, 1), 3),     # 56 beta strings: tiles wrap rows
                                       ((5, 4, 11, 0, 11, 0, 1), 2)])  # 152,460
def test_sigma_dense_tile_widths(J, orc, ctx, spec, nvec):
    """Every E-kernel column-tile width (FP64 16 and 32, FP32 32 and 64) against the oracle;
    invalid widths are refused."""
    gs, os_, m1, m2 = _spaces(J, orc, ctx, spec, seed=5)
    h = J.Hamiltonian(gs, m1, m2, J.DIRECT_TABLES)
    X = rand_vectors(gs.ndets, nvec, 9)
    ref = os_.sigma(m1, m2, X)
    scale = np.abs(ref).max()
    for bn64, bn32 in ((16, 32), (32, 64)):
        assert h.gemm_tiles(bn64, bn32) == (bn64, bn32)
        y = h.sigma(X, method=J.SIGMA_GEMM64)
        np.testing.assert_allclose(y, ref, rtol=0, atol=1e-12 * scale, err_msg="bn64=%d" % bn64)
        y = h.sigma(X, method=J.SIGMA_GEMM32)
        np.testing.assert_allclose(y, ref, rtol=0, atol=2e-5 * scale, err_msg="bn32=%d" % bn32)
    with pytest.raises(J.JciError):
        h.gemm_tiles(64, 0)
    with pytest.raises(J.JciError):
        h.gemm_tiles(0, 16)


def test_pair_counts_fci(J, orc, ctx):
    # FCI 3a3b/8: connections per row = 2*3*5 singles + 2*C(3,2)*C(5,2) doubles + (3*5)^2
    gs, os_, m1, m2 = _spaces(J, orc, ctx, (3, 3, 8, 0, 8, 0, 1))
    h = J.Hamiltonian(gs, m1, m2, J.PAIRS_FP64)
    per_row = 2 * 15 + 2 * 3 * 10 + 15 * 15
    assert h.npairs == gs.ndets * per_row
    hu = J.Hamiltonian(gs, m1, m2, J.PAIRS_FP64 | J.PAIRS_UPPER)
    assert hu.npairs == gs.ndets * per_row // 2 == os_.count_upper()


@pytest.mark.parametrize("spec", [s for s in SMALL_SPACES if s[0] + s[1] > 1])
def test_davidson_vs_dense(J, orc, ctx, spec):
    gs, os_, m1, m2 = _spaces(J, orc, ctx, spec)
    N = gs.ndets
    if N > 3200:
        pytest.skip("dense reference too large")
    H = os_.dense_h(m1, m2)
    w = np.linalg.eigh(H)[0]
    nroots = min(3, N - 1)
    h = J.Hamiltonian(gs, m1, m2, J.PAIRS_FP64 | J.PAIRS_FP32)
    ev, vec, st = h.solve(nroots=nroots, restol=1e-7, krymin=nroots + 1, krymax=16,
                          refdim=min(32, N))
    np.testing.assert_allclose(ev, w[:nroots], rtol=0, atol=1e-8)
    for r in range(nroots):
        res = H @ vec[r] - ev[r] * vec[r]
        assert np.linalg.norm(res) < 1e-6
        assert abs(np.linalg.norm(vec[r]) - 1.0) < 1e-10


@pytest.mark.parametrize("precision", ["PREC_FP64", "PREC_MIXED"])
@pytest.mark.parametrize("flags,sig64,sig32", [
    ("pairs", "SIGMA_PAIRS64", "SIGMA_PAIRS32"),
    ("none", "SIGMA_DIRECT64", "SIGMA_DIRECT32"),
    ("tables", "SIGMA_FAST64", "SIGMA_FAST32"),
    ("gemm", "SIGMA_GEMM64", "SIGMA_GEMM32"),
])
def test_davidson_precision_and_methods(J, orc, ctx, precision, flags, sig64, sig32):
    spec = (3, 3, 8, 0, 8, 0, 1) if flags == "gemm" else (3, 3, 8, 2, 2, 2, 1)
    gs, os_, m1, m2 = _spaces(J, orc, ctx, spec, seed=3)
    H = os_.dense_h(m1, m2)
    w = np.linalg.eigh(H)[0]
    fl = {"pairs": J.PAIRS_FP64 | J.PAIRS_FP32, "none": 0, "tables": J.DIRECT_TABLES,
          "gemm": J.DIRECT_TABLES}[flags]
    h = J.Hamiltonian(gs, m1, m2, fl)
    ev, vec, st = h.solve(nroots=3, restol=1e-7, krymin=4, krymax=12, refdim=40,
                          precision=getattr(J, precision), sigma_fp64=getattr(J, sig64),
                          sigma_fp32=getattr(J, sig32), guess=J.GUESS_DIAG)
    np.testing.assert_allclose(ev, w[:3], rtol=0, atol=1e-8)
    if precision == "PREC_MIXED":
        assert st["nsigma32"] > 0 and st["nsigma64"] > 0


@pytest.mark.parametrize("precision", ["PREC_FP64", "PREC_MIXED"])
@pytest.mark.parametrize("spec", [(3, 3, 8, 0, 8, 0, 1), (3, 3, 8, 2, 2, 2, 1)])
def test_davidson_gdk_restart_and_buffer_tolerance(J, orc, ctx, precision, spec):
    """GD+k restarts (the previous iteration's Ritz vectors kept next to the lowest ones) and
    buffer pairs that stop expanding at residual 1e-3: a subspace of 10 for 3 roots + 1
    buffer collapses every two iterations, and the roots still match dense eigenvalues to
    1e-8 with residuals below the tolerance."""
    gs, os_, m1, m2 = _spaces(J, orc, ctx, spec, seed=8)
    H = os_.dense_h(m1, m2)
    w = np.linalg.eigh(H)[0]
    h = J.Hamiltonian(gs, m1, m2, J.DIRECT_TABLES)
    ev, vec, st = h.solve(nroots=3, restol=1e-7, krymin=5, krymax=10, refdim=40,
                          precision=getattr(J, precision), guess=J.GUESS_DIAG, nbuffer=1,
                          buftol=1e-3, restart=J.RESTART_GDK)
    np.testing.assert_allclose(ev, w[:3], rtol=0, atol=1e-8)
    for r in range(3):
        assert np.linalg.norm(H @ vec[r] - ev[r] * vec[r]) < 2e-7
        for q in range(r):
            assert abs(vec[r] @ vec[q]) < 1e-10


@pytest.mark.parametrize("n,a,nb", [(10, 4, 64), (8, 3, 200), (6, 2, 225), (12, 4, 64)])
def test_lowest_diag_selection(J, ctx, n, a, nb):
    """The chunked device selection of the guess determinants equals a stable sort of the
    diagonal (value, then index): one chunk and several (44,100 and 245,025 determinants
    span 22 and 120 chunks of 2048), nb up to the whole space (6 orbitals, 2 electrons:
    225 determinants), and a diagonal with ties (the integer-valued diagonal below)."""
    m1, m2 = J.synth_integrals(n, 31)
    sp = J.Space(ctx, a, a, n)
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    hd = h.hdiag()
    want = np.argsort(hd, kind="stable")[:nb]
    np.testing.assert_array_equal(h.lowest_diag(nb), want)
    h.close()
    # ties: one-electron integrals only, on the diagonal, integer valued
    m1t = np.zeros_like(m1)
    for i in range(n):
        m1t[i * (i + 1) // 2 + i] = float(i % 3)
    ht = J.Hamiltonian(sp, m1t, np.zeros_like(m2), J.DIRECT_TABLES)
    hdt = ht.hdiag()
    assert len(np.unique(hdt)) < len(hdt) // 4
    np.testing.assert_array_equal(ht.lowest_diag(nb), np.argsort(hdt, kind="stable")[:nb])
    ht.close()
    sp.close()


def test_user_guess_and_maxiter(J, orc, ctx):
    gs, os_, m1, m2 = _spaces(J, orc, ctx, (3, 3, 8, 0, 8, 0, 1), seed=4)
    h = J.Hamiltonian(gs, m1, m2, J.PAIRS_FP64)
    ev0, vec0, _ = h.solve(nroots=2, restol=1e-8, krymin=3, krymax=12)
    # restart from the converged vectors: converges at once
    g = np.vstack([vec0, rand_vectors(gs.ndets, 1, 9)])
    ev1, _, st = h.solve(nroots=2, restol=1e-8, krymin=3, krymax=12, guess=J.GUESS_USER,
                         guess_vectors=g)
    np.testing.assert_allclose(ev1, ev0, atol=1e-10)
    assert st["iterations"] <= 2
    # maxiter exhausted: nonzero code 1, no exception with check=False
    _, _, st2 = h.solve(nroots=2, restol=1e-14, krymin=3, krymax=6, maxiter=3, check=False)
    assert st2["rc"] == 1 and not st2["converged"]


def test_gram_mfma_layout(J, orc, ctx):
    """The subspace matrix V^T (H V) goes through the MFMA Gram kernel; a 1-root solve on a
    space whose H is diagonal-dominant exercises it, and the Rayleigh quotient of the
    returned vector must equal the eigenvalue."""
    gs, os_, m1, m2 = _spaces(J, orc, ctx, (4, 4, 10, 2, 4, 2, 1), seed=6)
    h = J.Hamiltonian(gs, m1, m2, J.PAIRS_FP64)
    ev, vec, st = h.solve(nroots=5, restol=1e-7, krymin=8, krymax=40, refdim=64,
                          guess=J.GUESS_DIAG)
    hv = h.sigma(vec)
    for r in range(5):
        assert abs(vec[r] @ hv[r] - ev[r]) < 1e-10
        for q in range(r):
            assert abs(vec[r] @ vec[q]) < 1e-10


def test_buffer_root_finds_lowest_state(J, ctx):
    """Full CI 5a5b/10 (63,504 determinants, seed 400060) has its two lowest states 0.012 Ha
    apart and a guess that favours the second; with one buffer root the mixed-precision solve
    lands on the lowest.  Reference: Lanczos (scipy eigsh on the oracle sigma) gives
    E0 = 36.10636502, E1 = 36.11837624 (E_nuc = 10 included)."""
    m1, m2 = J.synth_integrals(10, 400060)
    sp = J.Space(ctx, 5, 5, 10)
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    for prec in (J.PREC_MIXED, J.PREC_FP64):
        for restart, buftol in ((J.RESTART_RITZ, 0.0), (J.RESTART_GDK, 1e-3)):
            ev, _, st = h.solve(nroots=1, restol=1e-6, krymin=3, krymax=12, refdim=64,
                                guess=J.GUESS_DIAG, totfrze=J.SYNTH_ENUC, vectors=False,
                                precision=prec, nbuffer=1, buftol=buftol, restart=restart)
            assert abs(ev[0] - 36.10636502) < 1e-7, (prec, restart, ev, st)


def test_mrcisd_16e60o_sampled_rows_bitexact(J, orc, ctx):
    """MRCI-SD space of 16 electrons in 60 orbitals (4 doubly occupied + CAS(8e,8o) + 48
    external, excitation level 2): 603,651,972 determinants in 36 blocks.  String tables and
    blocks equal the oracle's; 24 sampled rows (first, last, random) have bit-identical
    columns, phases and degrees and matrix elements within 1e-12."""
    m1, m2 = J.synth_integrals(60, 1)
    gs = J.Space(ctx, 8, 8, 60, 4, 8, 2, 1)
    os_ = orc.Space(8, 8, 60, 4, 8, 2, 1)
    assert gs.ndets == os_.ndets == 603651972
    for spin in (0, 1):
        assert np.array_equal(gs.strings(spin), os_.strings(spin))
        assert np.array_equal(gs.groups(spin), os_.groups(spin))
    assert np.array_equal(gs.blocks(), os_.blocks())
    h = J.Hamiltonian(gs, m1, m2)
    rows = np.unique(np.r_[0, gs.ndets - 1,
                           np.random.default_rng(5).integers(0, gs.ndets, size=22)])
    got = h.rows(rows)
    for I, (c, v, ph, dg) in zip(rows, got):
        oc, ov, oph, odg = os_.row(int(I), m1, m2)
        assert np.array_equal(c, oc), "columns differ in row %d" % I
        assert np.array_equal(ph, oph), "phases differ in row %d" % I
        assert np.array_equal(dg, odg), "degrees differ in row %d" % I
        np.testing.assert_allclose(v, ov, rtol=1e-12, atol=1e-13)


def test_pdavidson_guess_modes_and_restart_files(J, orc, ctx, tmp_path):
    """The MPI driver's Davidson step: prediagr 1 (block guess), 2 (unit vectors) and 3
    (ci.restart) reach the dense eigenvalues; civec holds the Ritz vectors; an unconverged
    run writes ci.restart (krymin vectors) and hvvec (zeros, as the reference does), and a
    run restarted from that file converges to the same energies."""
    spec = (3, 3, 8, 0, 8, 0, 1)
    gs, os_, m1, m2 = _spaces(J, orc, ctx, spec, seed=6)
    N = gs.ndets
    H = os_.dense_h(m1, m2)
    w = np.linalg.eigh(H)[0]
    h = J.Hamiltonian(gs, m1, m2, J.PAIRS_FP64)
    kw = dict(nroots=2, restol=1e-7, krymin=4, krymax=12, refdim=40)
    for prediagr in (1, 2):
        d = tmp_path / ("p%d" % prediagr)
        d.mkdir()
        ev, st = h.pdavidson(prediagr, str(d), **kw)
        assert st["rc"] == 0
        np.testing.assert_allclose(ev, w[:2], rtol=0, atol=1e-8)
        vec = J.read_vectors_ufmt(str(d / "civec"), N, 2)
        for r in range(2):
            assert np.linalg.norm(H @ vec[r] - ev[r] * vec[r]) < 1e-6
        assert not (d / "ci.restart").exists() and not (d / "hvvec").exists()
    # unconverged: restart files written
    d = tmp_path / "short"
    d.mkdir()
    ev0, st = h.pdavidson(1, str(d), nroots=2, restol=1e-7, krymin=4, krymax=8, refdim=40,
                          maxiter=2)
    assert st["rc"] == 1
    rs = J.read_vectors_ufmt(str(d / "ci.restart"), N, 4)
    np.testing.assert_allclose(rs @ rs.T, np.eye(4), atol=1e-10)
    assert not J.read_vectors_ufmt(str(d / "hvvec"), N, 2).any()
    # prediagr 3 continues from ci.restart
    ev3, st3 = h.pdavidson(3, str(d), **kw)
    assert st3["rc"] == 0
    np.testing.assert_allclose(ev3, w[:2], rtol=0, atol=1e-8)
    ev1, st1 = h.pdavidson(1, str(tmp_path / "p1"), **kw)
    assert st3["iterations"] < st1["iterations"]
    # a restart file of other dimensions is refused with the reader's code
    J.write_vectors_ufmt(str(d / "ci.restart"), np.zeros((3, N)))
    with pytest.raises(J.JciError):
        h.pdavidson(3, str(d), **kw)
