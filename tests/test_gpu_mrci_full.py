"""MRCI-SD 16 electrons in 60 orbitals (4 doubly occupied + CAS(8e,8o) + 48 external,
excitation level 2, MRCI-SD rule: 603,651,972 determinants, 790,734 strings per spin).

* The per-string excitation tables the matrix-free sigma reads (built on the GPU by
  k_exc), for sampled strings of every group: each single / double of the string whose
  target is a string of the space, with the target's index and group, the orbitals, the
  sign and the one-spin element, against an enumeration here and the oracle's
  Slater-Condon element of the same string pair (the other spin empty);
* one matrix-free FP64 sigma pass of a seeded random vector against the oracle's rows
  (owner computes, ascending columns) for 16 sampled determinants, to 1e-12 relative.
"""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N_MRCI = 603651972


@pytest.fixture(scope="module")
def mrci(J, orc, ctx):
    m1, m2 = J.synth_integrals(60, 1)
    gs = J.Space(ctx, 8, 8, 60, 4, 8, 2, 1)
    assert gs.ndets == N_MRCI
    h = J.Hamiltonian(gs, m1, m2, J.DIRECT_TABLES)
    print("tables built", flush=True)
    osp = orc.Space(8, 8, 60, 4, 8, 2, 1)
    yield m1, m2, gs, h, osp
    h.close()
    gs.close()


def _sign(m, i, f):
    lo, hi = min(i, f), max(i, f)
    between = ((1 << hi) - 1) & ~((1 << (lo + 1)) - 1)
    return -1 if bin(m & between).count("1") % 2 else 1


def test_mrcisd_16e60o_excitation_tables(J, mrci):
    m1, m2, gs, h, osp = mrci
    strings = gs.strings(0)
    assert np.array_equal(strings, osp.strings(0))
    index = {int(s): k for k, s in enumerate(strings)}
    groups = gs.groups(0)                       # (docc, actv, virt, start, count)
    starts = groups[:, 3]

    def group_of(J_):
        return int(np.searchsorted(starts, J_, side="right") - 1)
    rng = np.random.default_rng(4)
    sample = [int(g[3] + rng.integers(0, g[4])) for g in groups] + \
        [int(x) for x in rng.integers(0, len(strings), size=3)]
    n = 60
    for I in sample:
        m = int(strings[I])
        occ = [k for k in range(n) if m >> k & 1]
        emp = [k for k in range(n) if not m >> k & 1]
        # singles
        exp = {}
        for i in occ:
            for f in emp:
                t = m ^ (1 << i) ^ (1 << f)
                if t in index:
                    exp[index[t]] = (i, f, _sign(m, i, f))
        tg, info, val = h.string_excitations(0, I, 1)
        assert len(tg) == len(exp), (I, len(tg), len(exp))
        for Jt, inf, v in zip(tg, info, val):
            i, f, s = exp[int(Jt)]
            assert (inf & 63, (inf >> 6) & 63) == (i, f)
            assert ((inf >> 12) & 1) == (s < 0)
            assert ((inf >> 13) & 31) == group_of(int(Jt))
            assert ((inf >> 18) & 4095) == max(i, f) * (max(i, f) + 1) // 2 + min(i, f)
            hij, deg, ph = osp.hij(m, 0, int(strings[Jt]), 0, m1, m2)
            assert deg == 1 and ph == s
            assert abs(v - s * hij) <= 1e-12 * max(1.0, abs(hij))
        # doubles
        exp = {}
        for a in range(len(occ)):
            for b in range(a + 1, len(occ)):
                i0, i1 = occ[a], occ[b]
                for c in range(len(emp)):
                    for d in range(c + 1, len(emp)):
                        f0, f1 = emp[c], emp[d]
                        t = m ^ (1 << i0) ^ (1 << i1) ^ (1 << f0) ^ (1 << f1)
                        if t in index:
                            s = _sign(m, i0, f0) * _sign(m ^ (1 << i0) ^ (1 << f0), i1, f1)
                            exp[index[t]] = (i0, i1, f0, f1, s)
        tg, info, val = h.string_excitations(0, I, 2)
        assert len(tg) == len(exp), (I, len(tg), len(exp))
        for Jt, inf, v in zip(tg, info, val):
            i0, i1, f0, f1, s = exp[int(Jt)]
            assert (inf & 63, (inf >> 6) & 63, (inf >> 12) & 63, (inf >> 18) & 63) == \
                (i0, i1, f0, f1)
            assert ((inf >> 24) & 1) == (s < 0)
            assert ((inf >> 25) & 31) == group_of(int(Jt))
            hij, deg, ph = osp.hij(m, 0, int(strings[Jt]), 0, m1, m2)
            assert deg == 2 and ph == s
            assert abs(v - hij) <= 1e-12 * max(1.0, abs(hij))
        print("string %d: tables match" % I, flush=True)


def test_mrcisd_16e60o_sigma_rows(J, mrci):
    m1, m2, gs, h, osp = mrci
    N = gs.ndets
    x = np.random.default_rng(1).standard_normal(N)
    print("vector ready", flush=True)
    y = h.sigma(x, method=J.SIGMA_FAST64)
    print("sigma done", flush=True)
    rows = np.unique(np.r_[0, N - 1, np.random.default_rng(2).integers(0, N, size=14)])
    ref = osp.sigma_rows(m1, m2, x, rows)
    print("oracle rows done", flush=True)
    np.testing.assert_allclose(y[rows], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
