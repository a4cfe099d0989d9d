"""Unit checks of the Davidson vector kernels against numpy (FP64)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,w,n", [(1, 1, 1), (3, 5, 37), (16, 16, 64), (17, 4, 1000),
                                   (64, 64, 4097), (40, 8, 100003)])
def test_gram_mfma(J, ctx, m, w, n):
    rng = np.random.default_rng(m * 1000 + w)
    V = rng.standard_normal((m, n))
    W = rng.standard_normal((w, n))
    G = ctx.gram(V, W)
    np.testing.assert_allclose(G, V @ W.T, rtol=1e-12, atol=1e-12 * np.sqrt(n))
