"""Pinning the oracle (CPU restatement) before trusting it.

Each check is chosen so that a plausible slip (wrong sign, dropped term, wrong index,
swapped operand) fails:
  * brute-force second quantization (ladder operators) on small spaces: every element;
  * closed forms: no two-electron integrals -> E0 = sum of the lowest orbital energies;
    one electron -> H = h; orbital relabelling leaves the FCI spectrum unchanged;
  * closed-form counts: C(12,4)^2 determinants and 1,424 connections per FCI row;
    603,651,972 (MRCI-SD rule) and 1,770,874,308 (serial rule) determinants for
    16e/60o with 4 docc + CAS(8,8) + 48 external at xlvl 2;
  * the reference's own serial sources (oracle/_ref): determinant list and order, CAS
    flags, connectivity map sections, and the matrix elements where the serial rules are
    right (diagonal, singles, alpha-beta doubles).
"""
import itertools
import os
from math import comb

import numpy as np
import pytest

from conftest import SMALL_SPACES


def rand_packed(n, seed):
    """Random packed integrals from dense symmetric tensors (independent of the product's
    generator): h symmetric, (ij|kl) = sum_P L^P_ij L^P_kl."""
    from oracle import brute  # noqa: F401
    rng = np.random.default_rng(seed)
    L = rng.normal(size=(2 * n, n, n)) * 0.1
    L = L + L.transpose(0, 2, 1)
    h = rng.normal(size=(n, n)) * 0.1
    h = h + h.T + np.diag(np.arange(n) * 0.3 - 2)
    g = np.einsum("pij,pkl->ijkl", L, L)

    def i1(i, j):
        return i * (i - 1) // 2 + j if i >= j else j * (j - 1) // 2 + i

    m1n = n * (n + 1) // 2
    m1 = np.zeros(m1n)
    m2 = np.zeros(m1n * (m1n + 1) // 2)
    for i in range(1, n + 1):
        for j in range(1, i + 1):
            m1[i1(i, j) - 1] = h[i - 1, j - 1]
            for k in range(1, n + 1):
                for l in range(1, k + 1):
                    m2[i1(i1(i, j), i1(k, l)) - 1] = g[i - 1, j - 1, k - 1, l - 1]
    return m1, m2, h, g


@pytest.mark.parametrize("spec", [s for s in SMALL_SPACES if s[2] <= 8 and s[0] + s[1] <= 6])
def test_oracle_vs_brute_force(orc, spec):
    from oracle import brute
    na, nb, n, nd, nact, xl, rule = spec
    sp = orc.Space(na, nb, n, nd, nact, xl, rule)
    if sp.ndets > 700:
        pytest.skip("brute force too slow")
    m1, m2, h, g = rand_packed(n, 7)
    a, b = sp.dets()
    Hb = brute.hamiltonian(a, b, n, h, g)
    H = sp.dense_h(m1, m2)
    np.testing.assert_allclose(H, Hb, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(H, H.T)


def test_no_two_electron_integrals_closed_form(orc):
    # (ij|kl) = 0: determinants are products of orbitals, E0 = lowest 2 + 2 eigenvalues of h
    n = 6
    m1, _, h, _ = rand_packed(n, 3)
    m2 = np.zeros((n * (n + 1) // 2) * (n * (n + 1) // 2 + 1) // 2)
    sp = orc.Space(2, 2, n)
    w = np.linalg.eigh(sp.dense_h(m1, m2))[0]
    eh = np.linalg.eigvalsh(h)
    assert abs(w[0] - 2 * (eh[0] + eh[1])) < 1e-12
    # first excitation: one electron promoted 1 -> 2 in either spin
    assert abs(w[1] - (2 * eh[0] + eh[1] + eh[2])) < 1e-12


def test_one_electron_is_h(orc):
    n = 5
    m1, m2, h, _ = rand_packed(n, 4)
    sp = orc.Space(1, 0, n)
    H = sp.dense_h(m1, m2)
    np.testing.assert_allclose(H, h, rtol=0, atol=1e-15)


def test_orbital_relabelling_invariance(orc):
    from oracle import brute  # noqa: F401
    n = 6
    m1, m2, h, g = rand_packed(n, 5)
    perm = np.array([3, 0, 5, 1, 4, 2])
    hp = h[np.ix_(perm, perm)]
    gp = g[np.ix_(perm, perm, perm, perm)]

    def pack(h, g):
        def i1(i, j):
            return i * (i - 1) // 2 + j if i >= j else j * (j - 1) // 2 + i
        m1n = n * (n + 1) // 2
        a = np.zeros(m1n)
        b = np.zeros(m1n * (m1n + 1) // 2)
        for i in range(1, n + 1):
            for j in range(1, i + 1):
                a[i1(i, j) - 1] = h[i - 1, j - 1]
                for k in range(1, n + 1):
                    for l in range(1, k + 1):
                        b[i1(i1(i, j), i1(k, l)) - 1] = g[i - 1, j - 1, k - 1, l - 1]
        return a, b
    p1, p2 = pack(hp, gp)
    sp = orc.Space(3, 2, n)
    w1 = np.linalg.eigvalsh(sp.dense_h(m1, m2))
    w2 = np.linalg.eigvalsh(sp.dense_h(p1, p2))
    np.testing.assert_allclose(w1, w2, rtol=0, atol=1e-11)


def test_counts_fci12(orc):
    sp = orc.Space(4, 4, 12)
    assert sp.ndets == comb(12, 4) ** 2 == 245025
    m1, m2 = np.zeros(78), np.zeros(3081)
    rng = np.random.default_rng(0)
    for I in rng.integers(0, sp.ndets, 5):
        cols, _, ph, dg = sp.row(int(I), m1, m2)
        # 2*4*8 singles + 2*C(4,2)*C(8,2) same-spin doubles + (4*8)^2 alpha-beta doubles
        assert len(cols) == 2 * 32 + 2 * 6 * 28 + 32 * 32 == 1424
        assert set(np.unique(dg)) == {1, 2}
        assert set(np.unique(ph)) <= {-1, 1}


def test_counts_mrci_16e_60o(orc):
    # MRCI-SD rule (mpi_source/citruncate.c:414-471) and serial rule (source/citruncate.c:291)
    sp1 = orc.Space(8, 8, 60, 4, 8, 2, 1)
    assert sp1.nstr(0) == 790734
    assert sp1.ndets == 603651972
    assert len(sp1.blocks()) == 36
    sp0 = orc.Space(8, 8, 60, 4, 8, 2, 0)
    assert sp0.ndets == 1770874308
    assert len(sp0.blocks()) == 54
    # group sizes are products of binomials
    for d, a, v, start, cnt in sp1.groups(0):
        assert cnt == comb(4, d) * comb(8, a) * comb(48, v)


def test_lexicographic_string_order(orc):
    sp = orc.Space(3, 3, 7)
    s = sp.strings(0)
    combos = list(itertools.combinations(range(1, 8), 3))
    exp = [sum(1 << (o - 1) for o in c) for c in combos]
    assert [int(x) for x in s] == exp


def test_jacobi_vs_numpy(orc):
    rng = np.random.default_rng(2)
    for n in (1, 3, 10, 40):
        A = rng.standard_normal((n, n))
        A = A + A.T
        ev, V = orc.jacobi(A)
        np.testing.assert_allclose(ev, np.linalg.eigvalsh(A), atol=1e-12 * n)
        for i in range(n):
            assert np.linalg.norm(A @ V[i] - ev[i] * V[i]) < 1e-11 * n


def test_oracle_davidson_small(orc):
    n = 7
    m1, m2, _, _ = rand_packed(n, 8)
    sp = orc.Space(3, 3, n)
    w = np.linalg.eigvalsh(sp.dense_h(m1, m2))
    rc, ev, vec, rn, ns = sp.davidson(m1, m2, maxiter=300, krymin=4, krymax=12, nroots=3,
                                      restol=1e-9, refdim=30)
    assert rc == 0
    np.testing.assert_allclose(ev, w[:3], atol=1e-10)


# ---- against the reference's own sources (oracle/_ref) ----------------------------------
REF_SPACES = [(2, 2, 4, 0, 4, 0), (3, 3, 7, 1, 3, 2), (3, 3, 8, 2, 2, 2), (4, 3, 9, 1, 4, 1),
              (4, 4, 10, 2, 4, 2), (3, 3, 8, 0, 8, 0)]


def _ref_or_skip():
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return ref


@pytest.mark.parametrize("spec", REF_SPACES)
def test_detlist_matches_reference(orc, spec):
    ref = _ref_or_skip()
    na, nb, n, nd, nact, xl = spec
    rs = ref.RefSpace(na, nb, n, nd, nact, xl)
    ra, rb = rs.masks()
    sp = orc.Space(na, nb, n, nd, nact, xl, 0)
    a, b = sp.dets()
    assert rs.ndets == sp.ndets
    assert np.array_equal(ra, a) and np.array_equal(rb, b)
    # CAS flag: no virtual occupation in either string
    inner = (1 << (nd + nact)) - 1
    exp = np.array([int((int(x) & ~inner) == 0 and (int(y) & ~inner) == 0) for x, y in zip(a, b)])
    assert np.array_equal(rs.cas_flags(), exp)


@pytest.mark.parametrize("spec", REF_SPACES)
def test_cimap_matches_reference(orc, spec):
    ref = _ref_or_skip()
    na, nb, n, nd, nact, xl = spec
    rs = ref.RefSpace(na, nb, n, nd, nact, xl)
    hmap, secs = rs.cimap()
    rs.free_cimap(hmap)
    sp = orc.Space(na, nb, n, nd, nact, xl, 0)
    z1, z2 = np.zeros(n * (n + 1) // 2), np.zeros((n * (n + 1) // 2) * (n * (n + 1) // 2 + 1) // 2)
    for i in range(sp.ndets):
        cols = sp.row(i, z1, z2)[0]
        conn = [i] + [int(c) for c in cols if c > i]
        runs, start = [], None
        for j in conn:
            if start is None:
                start, prev = j, j
            elif j == prev + 1:
                prev = j
            else:
                runs.append((start - i, prev - i))
                start, prev = j, j
        runs.append((start - i, prev - i))
        assert secs[i] == runs, "row %d" % i


@pytest.mark.parametrize("spec", REF_SPACES)
def test_matrix_elements_vs_reference(orc, spec):
    """Diagonal, singles and alpha-beta doubles equal the serial reference's hmatels;
    same-spin doubles follow the corrected rule (the serial one is wrong, its values are
    not compared)."""
    ref = _ref_or_skip()
    na, nb, n, nd, nact, xl = spec
    m1, m2, _, _ = rand_packed(n, 9)
    rs = ref.RefSpace(na, nb, n, nd, nact, xl)
    sp = orc.Space(na, nb, n, nd, nact, xl, 0)
    a, b = sp.dets()
    np.testing.assert_allclose(rs.hdgls(m1, m2), sp.hdiag(m1, m2), rtol=0, atol=1e-12)
    rng = np.random.default_rng(3)
    checked = {1: 0, 2: 0}
    for i in rng.integers(0, sp.ndets, size=min(60, sp.ndets)):
        cols, vals, ph, dg = sp.row(int(i), m1, m2)
        for c, v, d in zip(cols, vals, dg):
            same_spin_double = d == 2 and (a[i] == a[c] or b[i] == b[c])
            if same_spin_double:
                continue
            r = rs.hmatels(int(i), int(c), m1, m2)
            assert abs(r - v) < 1e-12, (i, c, d, r, v)
            checked[int(d)] += 1
    assert checked[1] > 0 and checked[2] > 0


@pytest.mark.parametrize("spec", REF_SPACES[:3])
def test_reference_sigma_differs_only_by_same_spin_doubles(orc, spec):
    """The reference's compute_hv equals the oracle's sigma once the same-spin double
    contributions (which the serial reference gets wrong) are replaced by the
    reference's own values; i.e. the only difference is that rule."""
    ref = _ref_or_skip()
    na, nb, n, nd, nact, xl = spec
    m1, m2, _, _ = rand_packed(n, 10)
    rs = ref.RefSpace(na, nb, n, nd, nact, xl)
    hmap, _ = rs.cimap()
    v = np.random.default_rng(1).standard_normal(rs.ndets)
    c_ref = rs.compute_hv(m1, m2, v, hmap)
    rs.free_cimap(hmap)
    sp = orc.Space(na, nb, n, nd, nact, xl, 0)
    a, b = sp.dets()
    c = sp.sigma(m1, m2, v)
    for i in range(sp.ndets):
        cols, vals, _, dg = sp.row(i, m1, m2)
        for cc, vv, d in zip(cols, vals, dg):
            if d == 2 and (a[i] == a[cc] or b[i] == b[cc]):
                lo, hi = min(i, int(cc)), max(i, int(cc))   # compute_hv evaluates (lo, hi)
                c[i] += (rs.hmatels(lo, hi, m1, m2) - vv) * v[cc]
    np.testing.assert_allclose(c, c_ref, rtol=0, atol=1e-11)
