"""The serving loop: requests built on a worker thread (own context) and admitted into the
running batch, against standalone solves of the same requests; and a batch slot whose
initial block was prepared ahead against one that computes it at admission."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spec(n, a, roots):
    return {"norbs": n, "na": a, "nb": a, "ndocc": 0, "nactv": n, "xlvl": 0, "rule": 1,
            "nroots": roots}


def test_server_matches_standalone_solves(J, ctx):
    from jayci_amd import serve
    # 12 requests over 30 ms, 4 slots: some wait for a slot, admissions overlap steps
    specs = [_spec(8, 3, 1), _spec(9, 4, 2), _spec(10, 3, 3), _spec(8, 4, 1), _spec(9, 3, 2),
             _spec(10, 4, 1), _spec(8, 3, 2), _spec(9, 4, 1), _spec(10, 3, 2), _spec(8, 4, 3),
             _spec(9, 3, 1), _spec(10, 4, 2)]
    stream = [(0.0025 * i, s, 900 + i) for i, s in enumerate(specs)]
    for precision in (J.PREC_MIXED, J.PREC_FP64):
        srv = serve.Server(ctx, capacity=4, budget_bytes=60e9, max_norbs=16, precision=precision,
                           restol=1e-6, pool_bytes=4 << 30)
        try:
            done = srv.run(stream)
        finally:
            srv.close()
        assert sorted(r.idx for r in done) == list(range(len(specs)))
        for r in done:
            assert r.stats["done"] == 1, (r.idx, r.stats)
            s = r.spec
            m1, m2 = J.synth_integrals(s["norbs"], r.seed)
            sp = J.Space(ctx, s["na"], s["nb"], s["norbs"])
            h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
            kw = serve.solve_params(s, sp.ndets, 1e-8)
            kw.update(buftol=0.0, nbuffer=2, restart=J.RESTART_RITZ, maxiter=1000)
            ev, _, _ = h.solve(precision=J.PREC_FP64, vectors=False, **kw)
            np.testing.assert_allclose(r.energies, ev, rtol=0, atol=1e-8)
            h.close()
            sp.close()


def test_prepared_guess_same_as_computed_at_admission(J, ctx):
    """Slot A: initial block prepared ahead on a second context; slot B: computed by the add.
    Same Hamiltonian settings, so the same iterations and energies."""
    from jayci_amd import serve
    ctx2 = J.Context(ctx.device)
    b = J.Batch(ctx, 4, 16, graphs=True)
    m1, m2 = J.synth_integrals(10, 77)
    spa = J.Space(ctx2, 4, 4, 10)
    ha = J.Hamiltonian(spa, m1, m2, J.DIRECT_TABLES)
    spb = J.Space(ctx, 4, 4, 10)
    hb = J.Hamiltonian(spb, m1, m2, J.DIRECT_TABLES)
    kw = serve.solve_params(_spec(10, 4, 2), spa.ndets)
    pa = ha.params(precision=J.PREC_MIXED, **kw)
    ha.prepare_guess(pa)
    sa = b.add(ha, pa)
    sb = b.add(hb, hb.params(precision=J.PREC_MIXED, **kw))
    b.run(400)
    eva, _, sta = b.result(sa)
    evb, _, stb = b.result(sb)
    assert sta["done"] == 1 and stb["done"] == 1
    assert sta["iterations"] == stb["iterations"]
    np.testing.assert_allclose(eva, evb, rtol=0, atol=1e-10)
    b.close()
    ha.close(); spa.close(); hb.close(); spb.close()
    ctx2.close()


def test_server_lane_for_truncated_spaces(J, orc, ctx):
    """Requests with more orbitals than the batch is sized for (here MRCI-SD spaces of 12
    orbitals, 3600 determinants, batch sized for 10) run on the lane beside the batched
    full-CI requests; every result against the dense eigenvalues of the oracle's H, in both
    precisions."""
    from jayci_amd import serve
    mr = lambda roots: {"norbs": 12, "na": 3, "nb": 3, "ndocc": 1, "nactv": 3, "xlvl": 2,
                        "rule": 1, "nroots": roots}
    specs = [_spec(8, 3, 1), mr(2), _spec(9, 4, 2), _spec(10, 3, 1), mr(1), _spec(8, 4, 3),
             mr(3)]
    stream = [(0.002 * i, s, 950 + i) for i, s in enumerate(specs)]
    ref = {}
    for i, s in enumerate(specs):
        m1, m2 = J.synth_integrals(s["norbs"], 950 + i)
        H = orc.Space(s["na"], s["nb"], s["norbs"], s["ndocc"], s["nactv"], s["xlvl"],
                      s["rule"]).dense_h(m1, m2)
        ref[i] = np.linalg.eigvalsh(H)[:s["nroots"]] + J.SYNTH_ENUC
    for precision in (J.PREC_MIXED, J.PREC_FP64):
        srv = serve.Server(ctx, capacity=4, budget_bytes=60e9, max_norbs=10, precision=precision,
                           restol=1e-6, pool_bytes=4 << 30)
        try:
            done = srv.run(stream)
        finally:
            srv.close()
        assert sorted(r.idx for r in done) == list(range(len(specs)))
        for r in done:
            assert r.stats["done"] == 1, (r.idx, r.stats)
            np.testing.assert_allclose(r.energies, ref[r.idx], rtol=0, atol=1e-8)
            assert r.t_done >= r.arrival
