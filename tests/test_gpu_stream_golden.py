"""The default served stream's small full-CI requests solved on the GPU with the served
settings, FP64 and mixed precision, against the oracle's eigenvalues
(tests/golden/stream400_energies.json): every root within 1e-8 Ha."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "stream400_energies.json")


def test_stream_requests_vs_oracle(J, ctx):
    from jayci_amd import serve
    for r in json.load(open(GOLD))["requests"]:
        s = r["spec"]
        m1, m2 = J.synth_integrals(s["norbs"], r["seed"])
        sp = J.Space(ctx, s["na"], s["nb"], s["norbs"])
        h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
        kw = serve.solve_params(s, sp.ndets, 1e-6)
        for prec in (J.PREC_FP64, J.PREC_MIXED):
            ev, _, st = h.solve(precision=prec, vectors=False, **kw)
            assert st["converged"], (r["request"], prec, st)
            np.testing.assert_allclose(ev, r["energies"], rtol=0, atol=1e-8,
                                       err_msg="request %d precision %d" % (r["request"], prec))
        h.close()
        sp.close()
