"""Batched (lockstep, one sigma launch for all instances) solves against standalone solves
and dense eigensolutions: every instance's eigenvalues within 1e-8 Ha, alone and in the
mix, in FP64 and in mixed precision."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (na, nb, norbs, ndocc, nactv, xlvl, rule, nroots)
INSTANCES = [(2, 2, 6, 0, 6, 0, 1, 1), (3, 3, 8, 0, 8, 0, 1, 2), (3, 2, 7, 0, 7, 0, 1, 3),
             (3, 3, 8, 2, 2, 2, 1, 2), (4, 4, 9, 0, 9, 0, 1, 4), (3, 3, 7, 1, 3, 2, 0, 1),
             (4, 3, 9, 1, 4, 1, 0, 2), (2, 2, 5, 0, 5, 0, 1, 2)]


def _build(J, orc, ctx, spec, seed):
    na, nb, n, nd, nact, xl, rule, nr = spec
    m1, m2 = J.synth_integrals(n, seed)
    sp = J.Space(ctx, na, nb, n, nd, nact, xl, rule)
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    w = None
    if sp.ndets <= 3000:
        w = np.linalg.eigvalsh(orc.Space(na, nb, n, nd, nact, xl, rule).dense_h(m1, m2))
    return sp, h, w, nr


@pytest.mark.parametrize("precision", ["PREC_FP64", "PREC_MIXED"])
@pytest.mark.parametrize("graphs", [True, False])
def test_batch_matches_dense_and_alone(J, orc, ctx, precision, graphs):
    prec = getattr(J, precision)
    items = [_build(J, orc, ctx, spec, 100 + i) for i, spec in enumerate(INSTANCES)]
    b = J.Batch(ctx, 16, 16, graphs=graphs)
    slots = []
    for sp, h, w, nr in items:
        p = h.params(nroots=nr, restol=1e-7, krymin=nr + 1, krymax=min(16, sp.ndets),
                     refdim=min(48, sp.ndets), precision=prec, guess=J.GUESS_DIAG)
        slots.append(b.add(h, p))
    b.run(400)
    for (sp, h, w, nr), s in zip(items, slots):
        ev, vec, st = b.result(s, vectors=True)
        assert st["done"] == 1, st
        alone, _, _ = h.solve(nroots=nr, restol=1e-7, krymin=nr + 1, krymax=min(16, sp.ndets),
                              refdim=min(48, sp.ndets), precision=prec, guess=J.GUESS_DIAG)
        np.testing.assert_allclose(ev, alone, rtol=0, atol=1e-8)
        if w is not None:
            np.testing.assert_allclose(ev, w[:nr], rtol=0, atol=1e-8)
        hv = h.sigma(vec, method=J.SIGMA_FAST64)
        for r in range(nr):
            assert np.linalg.norm(hv[r] - ev[r] * vec[r]) < 2e-7
        if precision == "PREC_MIXED":
            assert st["nsigma32"] > 0
        b.release(s)
    b.close()


@pytest.mark.parametrize("precision", ["PREC_FP64", "PREC_MIXED"])
def test_batch_gdk_restart_and_buffer_tolerance(J, orc, ctx, precision):
    """Batched solves with GD+k restarts and buffer pairs stopping at residual 1e-3 (small
    subspaces, so every instance restarts several times) against dense eigenvalues and the
    standalone solver with the same settings."""
    prec = getattr(J, precision)
    items = [_build(J, orc, ctx, spec, 300 + i) for i, spec in enumerate(INSTANCES)]
    b = J.Batch(ctx, 16, 16, graphs=True)
    kws, slots = [], []
    for sp, h, w, nr in items:
        kw = dict(nroots=nr, restol=1e-7, krymin=min(nr + 2, sp.ndets),
                  krymax=min(2 * (nr + 1) + 2, sp.ndets), refdim=min(48, sp.ndets),
                  precision=prec, guess=J.GUESS_DIAG, nbuffer=1, buftol=1e-3,
                  restart=J.RESTART_GDK)
        kws.append(kw)
        slots.append(b.add(h, h.params(**kw)))
    b.run(600)
    for (sp, h, w, nr), s, kw in zip(items, slots, kws):
        ev, vec, st = b.result(s, vectors=True)
        assert st["done"] == 1, st
        alone, _, _ = h.solve(**kw)
        np.testing.assert_allclose(ev, alone, rtol=0, atol=1e-8)
        if w is not None:
            np.testing.assert_allclose(ev, w[:nr], rtol=0, atol=1e-8)
        hv = h.sigma(vec, method=J.SIGMA_FAST64)
        for r in range(nr):
            assert np.linalg.norm(hv[r] - ev[r] * vec[r]) < 2e-7
        b.release(s)
    b.close()


def test_batch_continuous_admission(J, orc, ctx):
    """Instances added while others are mid-solve give the same result as alone."""
    items = [_build(J, orc, ctx, spec, 200 + i) for i, spec in enumerate(INSTANCES)]
    b = J.Batch(ctx, 4, 16)
    pending = list(range(len(items)))
    active = {}
    results = {}
    steps = 0
    while pending or active:
        while pending and len(active) < 4:
            i = pending.pop(0)
            sp, h, w, nr = items[i]
            p = h.params(nroots=nr, restol=1e-7, krymin=nr + 1, krymax=min(16, sp.ndets),
                         refdim=min(48, sp.ndets), precision=J.PREC_MIXED, guess=J.GUESS_DIAG)
            active[b.add(h, p)] = i
        b.step()
        steps += 1
        for s in list(active):
            ev, _, st = b.result(s)
            if st["done"]:
                results[active.pop(s)] = ev
                b.release(s)
        assert steps < 2000
    for i, (sp, h, w, nr) in enumerate(items):
        alone, _, _ = h.solve(nroots=nr, restol=1e-7, krymin=nr + 1, krymax=min(16, sp.ndets),
                              refdim=min(48, sp.ndets), precision=J.PREC_MIXED, guess=J.GUESS_DIAG)
        np.testing.assert_allclose(results[i], alone, rtol=0, atol=1e-8)


@pytest.mark.parametrize("tiles", [(0, 0), (16, 64)])
def test_batch_pool_rounds_and_buffer_roots(J, orc, ctx, tiles):
    """A pool smaller than two vectors' intermediates forces one vector per sigma round; with
    buffer roots the results still match the standalone solves and the dense spectrum.  Run
    with the default E tile widths and with the other ones (FP64 16, FP32 64)."""
    specs = [(3, 3, 8, 0, 8, 0, 1, 2), (4, 4, 9, 0, 9, 0, 1, 1), (3, 2, 7, 0, 7, 0, 1, 2),
             (3, 3, 8, 2, 2, 2, 1, 1)]
    items = [_build(J, orc, ctx, spec, 300 + i) for i, spec in enumerate(specs)]
    b = J.Batch(ctx, 8, 16, pool_bytes=2 << 20)
    if tiles[0]:
        b.set_tiles(*tiles)
    slots = []
    for sp, h, w, nr in items:
        p = h.params(nroots=nr, restol=1e-7, krymin=nr + 1, krymax=min(16, sp.ndets),
                     refdim=min(48, sp.ndets), precision=J.PREC_MIXED, guess=J.GUESS_DIAG,
                     nbuffer=1)
        slots.append(b.add(h, p))
    b.run(400)
    for (sp, h, w, nr), s in zip(items, slots):
        ev, _, st = b.result(s)
        assert st["done"] == 1, st
        alone, _, _ = h.solve(nroots=nr, restol=1e-7, krymin=nr + 1, krymax=min(16, sp.ndets),
                              refdim=min(48, sp.ndets), precision=J.PREC_MIXED,
                              guess=J.GUESS_DIAG, nbuffer=1)
        np.testing.assert_allclose(ev, alone, rtol=0, atol=1e-8)
        if w is not None:
            np.testing.assert_allclose(ev, w[:nr], rtol=0, atol=1e-8)
        b.release(s)
    b.close()


@pytest.mark.parametrize("kmax", [28, 44, 64])
def test_batch_basis_sizes(J, orc, ctx, kmax):
    """The combine and Gram kernels are instantiated for bases of 16, 32, 48 and 64 vectors,
    and the instances of a phase are launched in groups by basis size.  A 12-vector instance
    beside one of up to kmax vectors against dense eigenvalues and the standalone solver
    with the same settings."""
    specs = [(3, 3, 7, 0, 7, 0, 1, 3), (3, 2, 8, 0, 8, 0, 1, 2)]   # 1225 and 1568 determinants
    items = [_build(J, orc, ctx, spec, 700 + i) for i, spec in enumerate(specs)]
    b = J.Batch(ctx, 16, 16)
    slots, kws = [], []
    for (sp, h, w, nr), km in zip(items, [kmax, 12]):
        kw = dict(nroots=nr, restol=1e-9, krymin=nr + 1, krymax=min(km, sp.ndets),
                  refdim=min(48, sp.ndets), precision=J.PREC_FP64, guess=J.GUESS_DIAG)
        kws.append(kw)
        slots.append(b.add(h, h.params(**kw)))
    b.run(600)
    for (sp, h, w, nr), s, kw in zip(items, slots, kws):
        ev, _, st = b.result(s)
        assert st["done"] == 1, st
        alone, _, _ = h.solve(**kw)
        np.testing.assert_allclose(ev, alone, rtol=0, atol=1e-9)
        np.testing.assert_allclose(ev, w[:nr], rtol=0, atol=1e-9)
        b.release(s)
    b.close()


@pytest.mark.parametrize("precision", ["PREC_FP64", "PREC_MIXED"])
def test_batch_result_independent_of_mix(J, orc, ctx, precision):
    """An instance's eigenvalues and iteration count are bit-identical whether it runs alone
    or beside other instances (of other basis sizes, so other kernel instantiations and
    launch groupings)."""
    prec = getattr(J, precision)
    target = (3, 3, 8, 0, 8, 0, 1, 2)
    others = [(4, 4, 9, 0, 9, 0, 1, 4), (2, 2, 6, 0, 6, 0, 1, 1), (3, 3, 7, 1, 3, 2, 0, 1)]
    runs = []
    for mix in ([], others):
        items = [_build(J, orc, ctx, target, 900)] + [_build(J, orc, ctx, s, 901 + i)
                                                       for i, s in enumerate(mix)]
        b = J.Batch(ctx, 16, 16)
        slots = []
        for (sp, h, w, nr), km in zip(items, [40, 44, 12, 16]):
            slots.append(b.add(h, h.params(nroots=nr, restol=1e-8, krymin=nr + 1,
                                           krymax=min(km, sp.ndets), refdim=min(48, sp.ndets),
                                           precision=prec, guess=J.GUESS_DIAG)))
        b.run(600)
        ev, _, st = b.result(slots[0])
        assert st["done"] == 1, st
        runs.append((ev.copy(), st["iterations"], st["nsigma64"], st["nsigma32"]))
        b.close()
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    assert runs[0][1:] == runs[1][1:]


def test_batch_tuned_tiles_bit_identical(J, orc, ctx):
    """E tile widths autotuned on a representative full-CI Hamiltonian and set on the batch;
    the batch gives bit-identical energies and iteration counts with the tuned widths and with
    every other width (a width only partitions the columns of the E product)."""
    specs = [(3, 3, 8, 0, 8, 0, 1, 2), (4, 4, 9, 0, 9, 0, 1, 1), (3, 2, 7, 0, 7, 0, 1, 2)]
    items = [_build(J, orc, ctx, spec, 700 + i) for i, spec in enumerate(specs)]
    rep = items[1][1]
    runs = []
    for tiles in ("tuned", (16, 32), (32, 64)):
        b = J.Batch(ctx, 8, 16)
        if tiles == "tuned":
            bn64, bn32 = b.tune_tiles(rep, nvec=1, reps=3)
            assert bn64 in (16, 32) and bn32 in (32, 64)
        else:
            b.set_tiles(*tiles)
        slots = []
        for sp, h, w, nr in items:
            p = h.params(nroots=nr, restol=1e-7, krymin=nr + 1, krymax=min(16, sp.ndets),
                         refdim=min(48, sp.ndets), precision=J.PREC_MIXED, guess=J.GUESS_DIAG,
                         nbuffer=1)
            slots.append(b.add(h, p))
        b.run(400)
        out = []
        for (sp, h, w, nr), s in zip(items, slots):
            ev, _, st = b.result(s)
            assert st["done"] == 1, st
            if w is not None:
                np.testing.assert_allclose(ev, w[:nr], rtol=0, atol=1e-8)
            out.append((ev.copy(), st["iterations"]))
        runs.append(out)
        b.close()
    for other in runs[1:]:
        for (e0, i0), (e1, i1) in zip(runs[0], other):
            assert np.array_equal(e0, e1) and i0 == i1
