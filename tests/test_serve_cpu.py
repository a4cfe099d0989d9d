"""Serving loop host logic (no GPU): solver settings of a request, the seeded stream, and the
admission error for a request that cannot fit the budget."""
import numpy as np
import pytest


def test_solve_params_served_settings(J):
    from jayci_amd import serve
    spec = {"nroots": 3}
    kw = serve.solve_params(spec, 4008004)
    assert kw["nbuffer"] == 1 and kw["buftol"] == 1e-3 and kw["restart"] == J.RESTART_GDK
    assert kw["krymin"] == 5 and kw["krymax"] == 32 and kw["etol"] == 1e-9
    small = serve.solve_params(spec, 10)
    assert small["krymax"] == 10 and small["refdim"] == 10
    p = J.make_params(10, **{k: v for k, v in small.items() if k != "nroots"}, nroots=3)
    assert p.buftol == 1e-3 and p.restart == J.RESTART_GDK


def test_stream_is_seeded():
    from jayci_amd import serve
    a = serve.make_stream(16, 200.0, 400)
    b = serve.make_stream(16, 200.0, 400)
    assert a == b
    t = [x[0] for x in a]
    assert all(t2 > t1 for t1, t2 in zip(t, t[1:]))
    for _, spec, _ in a:
        assert 8 <= spec["norbs"] <= 14 and 3 <= spec["na"] <= 5 and 1 <= spec["nroots"] <= 4


def test_request_above_budget_is_an_error():
    from jayci_amd import serve
    srv = serve.Server.__new__(serve.Server)   # host state only: no batch, no context
    srv.capacity, srv.budget, srv.restol, srv.nbuffer, srv.max_norbs = 4, 1.0, 1e-6, 1, 16
    stream = [(0.0, serve.instance_spec(np.random.default_rng(1)), 1)]
    with pytest.raises(RuntimeError, match="above the budget"):
        srv.run(stream)


def test_stream_mrci_share_and_lane_settings():
    """The config-4 stream: the MRCI-SD share replaces requests of the same seeded stream
    (arrival times and the other requests unchanged); the lane's settings and device bytes
    for the 16e/60o space (603,651,972 determinants) stay within a 200 GB budget less an
    8 GiB sigma pool for 1, 2 and 3 roots."""
    from jayci_amd import serve
    a = serve.make_stream(40, 200.0, 400)
    b = serve.make_stream(40, 200.0, 400, mrci_share=0.1)
    assert [x[0] for x in a] == [x[0] for x in b]
    mr = [i for i, (_, s, _) in enumerate(b) if s["norbs"] == 60]
    assert len(mr) == 4
    for i in range(40):
        if i not in mr:
            assert a[i] == b[i]
        else:
            s = b[i][1]
            assert (s["na"], s["nb"], s["ndocc"], s["nactv"], s["xlvl"]) == (8, 8, 4, 8, 2)
            assert 1 <= s["nroots"] <= 3
    N = 603651972
    for r in (1, 2, 3):
        spec = dict(serve.mrci_spec(np.random.default_rng(0)), nroots=r)
        kw = serve.lane_solve_params(spec, N)
        assert kw["nroots"] == r and kw["krymax"] >= kw["krymin"] > r
        nt = r + (1 if r == 1 else 0)
        assert kw["krymax"] == max(4 * nt, 8)
        # within a budget the subspace shrinks, to twice the tracked pairs at the least
        bud = 200e9 - (8 << 30)
        kb = serve.lane_solve_params(spec, N, budget=bud)
        assert max(2 * nt, 6) <= kb["krymax"] <= kw["krymax"]
        assert (serve.lane_bytes(spec, N, budget=bud) <= bud
                or kb["krymax"] == max(2 * nt, 6))
        assert serve.lane_bytes(spec, N, budget=bud) == serve.lane_vector_bytes(
            N, kb["krymax"], nt)


def test_admission_order_budget_and_lane():
    """serve.admit: a lane request waiting for the busy lane does not hold up the batch
    requests behind it; the budget and the batch's slots stop the pass in queue order."""
    from types import SimpleNamespace as R
    from jayci_amd import serve
    lane = lambda r: r.lane
    q = [R(lane=True, bytes=50), R(lane=False, bytes=10), R(lane=False, bytes=10)]
    # lane busy: the two batch requests behind the lane request go ahead
    assert serve.admit(q, lane, True, 0, 100, 8) == (None, [1, 2], 20)
    # lane free: the lane request starts, both batch requests too
    assert serve.admit(q, lane, False, 0, 100, 8) == (0, [1, 2], 70)
    # budget: the first request that does not fit stops the pass, even if later ones fit
    q2 = [R(lane=False, bytes=90), R(lane=False, bytes=5)]
    assert serve.admit(q2, lane, False, 20, 100, 8) == (None, [], 20)
    # slots: one free slot takes the first batch request only
    assert serve.admit(q2, lane, False, 0, 100, 1) == (None, [0], 90)
    # only one lane request starts per pass; a second waits behind it
    q3 = [R(lane=False, bytes=1), R(lane=True, bytes=1), R(lane=True, bytes=1),
          R(lane=False, bytes=1)]
    assert serve.admit(q3, lane, False, 0, 100, 8) == (1, [0, 3], 3)
