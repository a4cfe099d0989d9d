"""The allocation-free sigma entry point (jci_sigma_dev_noalloc) against jci_sigma_dev on the
same device buffers: bit-identical outputs on two calls in a row, the second on a smaller
space, for every method that stages through the library's scratch array; a space whose
dense-string intermediate does not fit the array is rejected with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libjayci_gpu.so itself is linked against (same process instance)."""
    path = None
    for line in open("/proc/self/maps"):
        if "libamdhip64.so" in line:
            path = line.split()[-1]
            break
    assert path, "libamdhip64 not loaded"
    h = ctypes.CDLL(path)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return h


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    def get(self, shape):
        out = np.empty(shape)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def fill(self, byte):
        assert self.hip.hipMemset(self.p, byte, self.n) == 0

    def free(self):
        self.hip.hipFree(self.p)


def _entry_points(J):
    L = J.lib()
    for name in ("jci_sigma_dev", "jci_sigma_dev_noalloc"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.jci_sigma_noalloc_capacity.restype = ctypes.c_longlong
    return L


# (alpha, beta, orbitals), vectors, methods; the second case is the smaller input
CASES = [
    ((4, 4, 12), 3, ["SIGMA_GEMM64", "SIGMA_GEMM32", "SIGMA_FAST32", "SIGMA_PAIRS32",
                     "SIGMA_FCI32", "SIGMA_FAST64"]),
    ((3, 3, 10), 2, ["SIGMA_GEMM64", "SIGMA_GEMM32", "SIGMA_FAST32", "SIGMA_PAIRS32",
                     "SIGMA_FCI32", "SIGMA_FAST64"]),
]


def test_sigma_noalloc_matches_sigma_dev(J, ctx):
    L = _entry_points(J)
    hip = _hip()
    assert L.jci_sigma_noalloc_capacity() == 128 << 20
    for (na, nb, n), nvec, methods in CASES:
        print("noalloc: full CI %da%db/%d, %d vectors" % (na, nb, n, nvec), flush=True)
        m1, m2 = J.synth_integrals(n, 31 + n)
        sp = J.Space(ctx, na, nb, n)
        h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES | J.PAIRS_FP32)
        N = sp.ndets
        x = np.random.default_rng(n).standard_normal((N, nvec))     # determinant-major
        bx, b1, b2 = DevBuf(hip, x.nbytes), DevBuf(hip, x.nbytes), DevBuf(hip, x.nbytes)
        bx.put(x)
        for name in methods:
            m = getattr(J, name)
            b1.fill(0x7f)
            b2.fill(0x11)
            assert L.jci_sigma_dev(h._h, m, nvec, bx.p, b1.p) == 0, J.last_error()
            ctx.sync()
            assert L.jci_sigma_dev_noalloc(h._h, m, nvec, bx.p, b2.p) == 0, J.last_error()
            ctx.sync()
            y1, y2 = b1.get((N, nvec)), b2.get((N, nvec))
            assert np.array_equal(y1, y2), name
            assert np.isfinite(y1).all()
        for b in (bx, b1, b2):
            b.free()
        h.close()
        sp.close()


def test_sigma_noalloc_rejects_oversized(J, ctx):
    """5a5b/13: the FP64 dense-string intermediate of one vector (41 rows x 1,656,369
    determinants x 8 bytes, 543 MB) exceeds the 128 MiB array."""
    L = _entry_points(J)
    hip = _hip()
    m1, m2 = J.synth_integrals(13, 5)
    sp = J.Space(ctx, 5, 5, 13)
    h = J.Hamiltonian(sp, m1, m2, J.DIRECT_TABLES)
    N = sp.ndets
    bx, by = DevBuf(hip, 8 * N), DevBuf(hip, 8 * N)
    bx.put(np.ones(N))
    assert L.jci_sigma_dev_noalloc(h._h, J.SIGMA_GEMM64, 1, bx.p, by.p) == -1
    assert "scratch array" in J.last_error()
    # a method without scratch needs still runs, and matches the allocating entry point
    assert L.jci_sigma_dev_noalloc(h._h, J.SIGMA_FAST64, 1, bx.p, by.p) == 0
    ctx.sync()
    y2 = by.get((N,))
    assert L.jci_sigma_dev(h._h, J.SIGMA_FAST64, 1, bx.p, by.p) == 0
    ctx.sync()
    assert np.array_equal(y2, by.get((N,)))
    bx.free()
    by.free()
    h.close()
    sp.close()
