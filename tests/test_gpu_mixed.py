"""Mixed precision at the FP32 noise floor: with the switch residual set far below what FP32
sigma can reach, the FP32 phase must hand over to FP64 when it stalls (no new 10 % minimum
of the largest residual for 6 iterations, below 1e-2), and the solve then converge to the
dense eigenvalues of the oracle's H, in the single-instance solver and in the batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPEC = (3, 3, 12, 1, 3, 2, 1)      # MRCI-SD space, 3600 determinants


def _ref(J, orc, nroots):
    na, nb, n, nd, nact, xl, rule = SPEC
    m1, m2 = J.synth_integrals(n, 31)
    H = orc.Space(na, nb, n, nd, nact, xl, rule).dense_h(m1, m2)
    return m1, m2, np.linalg.eigvalsh(H)[:nroots]


def test_solver_hands_over_at_fp32_floor(J, orc, ctx):
    m1, m2, ref = _ref(J, orc, 2)
    sp = J.Space(ctx, *SPEC)
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    ev, _, st = h.solve(nroots=2, restol=1e-8, maxiter=400, krymin=3, krymax=12,
                        precision=J.PREC_MIXED, mixed_switch=1e-13, vectors=False,
                        check=False)
    assert st["converged"], st
    assert st["nsigma32"] > 0 and st["nsigma64"] > 0, st
    np.testing.assert_allclose(ev, ref, rtol=0, atol=1e-9)
    h.close()
    sp.close()


def test_batch_hands_over_at_fp32_floor(J, orc, ctx):
    m1, m2, ref = _ref(J, orc, 2)
    sp = J.Space(ctx, *SPEC)
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    b = J.Batch(ctx, 2, 12, graphs=True)
    p = h.params(nroots=2, restol=1e-8, maxiter=400, krymin=3, krymax=12,
                 precision=J.PREC_MIXED, mixed_switch=1e-13)
    s = b.add(h, p)
    b.run(400)
    ev, _, st = b.result(s)
    assert st["done"] == 1, st
    assert st["nsigma32"] > 0 and st["nsigma64"] > 0, st
    np.testing.assert_allclose(ev, ref, rtol=0, atol=1e-9)
    b.close()
    h.close()
    sp.close()
