"""The matrix-free table sigma in every launch shape against the oracle: one fused kernel and
the two-pass form (one spin's side and the alpha-beta products, then the other spin's side,
with every block in its own orientation, short-row blocks transposed or every block
transposed), beta tiles of 64 / 128 / 256 strings or per block (64 .. 512,
one launch per width), alpha singles staged 1, 2,
4 or 8 at a time, FP64 (1e-12 of the vector scale) and FP32 (2e-5), 1 and 3 vectors; tiles
in blockIdx order and dealt to the XCDs in runs, string-major and chunk-major."""
import numpy as np
import pytest

from conftest import SMALL_SPACES, rand_vectors

pytestmark = pytest.mark.gpu

# the truncated spaces (several blocks) plus one full-CI space
SPACES = [s for s in SMALL_SPACES if s[3] > 0 or s[4] != s[2]] + [(3, 3, 8, 0, 8, 0, 1)]


@pytest.mark.parametrize("spec", SPACES)
def test_table_sigma_shapes(J, orc, ctx, spec):
    na, nb, n, nd, nact, xl, rule = spec
    m1, m2 = J.synth_integrals(n, 17)
    gs = J.Space(ctx, na, nb, n, nd, nact, xl, rule)
    h = J.Hamiltonian(gs, m1, m2, J.DIRECT_TABLES)
    X = rand_vectors(gs.ndets, 3, 9)
    ref = orc.Space(na, nb, n, nd, nact, xl, rule).sigma(m1, m2, X)
    scale = np.abs(ref).max()
    for split in (0, 1, 2, 3):
        assert h.fast_split(split) == split
        for width in (64, 128, 256, 0):
            for chunk in (1, 2, 4, 8):
                assert h.fast_tiles(width, chunk) == (width, chunk)
                for k in (1, 3):
                    y64 = h.sigma(X[:k], method=J.SIGMA_FAST64)
                    y32 = h.sigma(X[:k], method=J.SIGMA_FAST32)
                    e64 = np.abs(y64 - ref[:k]).max() / scale
                    e32 = np.abs(y32 - ref[:k]).max() / scale
                    assert e64 < 1e-12, (split, width, chunk, k, e64)
                    assert e32 < 2e-5, (split, width, chunk, k, e32)
    # tiles dealt to the XCDs in runs of 3 / 16 (grid padded past the last tile), in
    # string-major and chunk-major order
    for xg, cm in ((3, 0), (16, 0), (0, 1), (16, 1)):
        assert h.fast_xcd(xg, cm) == (xg, cm)
        for split in (0, 1, 2, 3):
            h.fast_split(split)
            h.fast_tiles(0, 0)
            y64 = h.sigma(X, method=J.SIGMA_FAST64)
            assert np.abs(y64 - ref).max() / scale < 1e-12, (xg, cm, split)
    h.fast_xcd(0, 0)
    h.close()
    gs.close()
