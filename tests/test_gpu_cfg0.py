"""Full CI 4a4b / 12 orbitals (245,025 determinants), synthetic integrals of seed 100: the
reference's CPU-runnable configuration, solved on the GPU.

* one seeded sigma of three vectors, FP64 and FP32 dense-string paths, against the
  oracle's sigma (1e-12 / 2e-5 of the vector scale);
* Davidson, 1 and 3 roots, plain FP64 and mixed precision (FP32 sigma until the residual
  is below 1e-4, then FP64), against tests/golden/cfg0_energies.json: the oracle's
  Davidson (the control flow of source/davidson.c:91-288) converged to residual 1e-9.
  Target 1e-8 Ha.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cfg0_energies.json")))


@pytest.fixture(scope="module")
def cfg0(J, ctx):
    assert GOLD["seed"] == 100
    m1, m2 = J.synth_integrals(12, GOLD["seed"])
    sp = J.Space(ctx, 4, 4, 12)
    assert sp.ndets == 245025
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    yield m1, m2, sp, h
    h.close()
    sp.close()


def test_cfg0_sigma_against_oracle(J, orc, cfg0):
    m1, m2, sp, h = cfg0
    X = np.random.default_rng(100).standard_normal((3, sp.ndets))
    print("oracle sigma (245,025 determinants, 3 vectors)", flush=True)
    ref = orc.Space(4, 4, 12).sigma(m1, m2, X)
    scale = np.abs(ref).max()
    for meth, tol in ((J.SIGMA_GEMM64, 1e-12), (J.SIGMA_FAST64, 1e-12), (J.SIGMA_GEMM32, 2e-5),
                      (J.SIGMA_FAST32, 2e-5)):
        y = h.sigma(X, method=meth)
        assert np.abs(y - ref).max() / scale < tol, meth


@pytest.mark.parametrize("nroots", [1, 3])
@pytest.mark.parametrize("precision", ["PREC_FP64", "PREC_MIXED"])
def test_cfg0_davidson_against_golden(J, cfg0, nroots, precision):
    m1, m2, sp, h = cfg0
    gold = np.array(GOLD["energies"][:nroots])
    ev, _, st = h.solve(nroots=nroots, restol=1e-6, krymin=nroots + 1,
                        krymax=max(12, 8 * (nroots + 1)), refdim=64,
                        precision=getattr(J, precision), guess=J.GUESS_DIAG, nbuffer=1,
                        etol=1e-9, buftol=1e-3, restart=J.RESTART_GDK, totfrze=J.SYNTH_ENUC,
                        vectors=False)
    print("cfg0 %s %d roots: %s (%d iterations, sigma %d fp64 / %d fp32)"
          % (precision, nroots, ev, st["iterations"], st["nsigma64"], st["nsigma32"]), flush=True)
    assert st["converged"]
    assert max(st["rnorm"]) < 1e-6
    np.testing.assert_allclose(ev, gold, rtol=0, atol=1e-8)
    if precision == "PREC_MIXED":
        assert st["nsigma32"] > 0 and st["nsigma64"] > 0
