"""The oracle's eigenvalues of the default served stream's small full-CI requests
(tests/golden/stream400_energies.json, from the oracle's CPU Davidson at residual 1e-9),
pinned by dense diagonalisation of the oracle's H on three of the smallest requests, and
the fixture's requests checked against the stream generator they were drawn from."""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "stream400_energies.json")


def _gold():
    return json.load(open(GOLD))


def test_fixture_matches_stream(J):
    from jayci_amd import serve
    stream = serve.make_stream(64, 200.0, 400)
    g = _gold()
    assert len(g["requests"]) == 31
    for r in g["requests"]:
        _, spec, seed = stream[r["request"]]
        assert spec == r["spec"] and seed == r["seed"]
        assert r["rc"] == 0 and r["rnorm_last"] < 1e-9
        assert len(r["energies"]) == spec["nroots"]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_fixture_vs_dense_eigh(J, orc, k):
    small = [r for r in _gold()["requests"] if r["ndets"] == 3136][k]
    s = small["spec"]
    m1, m2 = J.synth_integrals(s["norbs"], small["seed"])
    H = orc.Space(s["na"], s["nb"], s["norbs"], 0, s["norbs"], 0, 1).dense_h(m1, m2)
    ev = np.linalg.eigvalsh(H)[:s["nroots"]] + J.SYNTH_ENUC
    np.testing.assert_allclose(small["energies"], ev, rtol=0, atol=1e-9)
