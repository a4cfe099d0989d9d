"""ctypes bindings of the oracle (CPU restatement of the JayCI sigma / Davidson path).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check and by bench.py's
cpu_baseline leg, and there only as the checker.  The product (jayci_amd) never imports
it.  The C source, oracle/jayci_oracle.c, cites the reference routine each function
restates.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libjayci_ref.so")

_lib = None


def build(ref=False):
    """Compile liboracle.so (and, when ref=True, oracle/_ref/libjayci_ref.so and the SIFS
    fixture writer oracle/_ref/sifs_ref)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if ref:
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])
        subprocess.check_call(["sh", os.path.join(HERE, "sifs", "build.sh")],
                              stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        c_ll, c_int, c_dbl = ctypes.c_longlong, ctypes.c_int, ctypes.c_double
        P = ctypes.c_void_p
        L.orc_space_new.restype = P
        L.orc_space_new.argtypes = [c_int] * 7
        L.orc_space_free.argtypes = [P]
        L.orc_space_ndets.restype = c_ll
        L.orc_space_ndets.argtypes = [P]
        L.orc_space_nstr.restype = c_ll
        L.orc_space_nstr.argtypes = [P, c_int]
        L.orc_space_ngroups.restype = c_int
        L.orc_space_ngroups.argtypes = [P, c_int]
        L.orc_space_nblocks.restype = c_int
        L.orc_space_nblocks.argtypes = [P]
        L.orc_space_strings.argtypes = [P, c_int, P]
        L.orc_space_groups.argtypes = [P, c_int, P]
        L.orc_space_blocks.argtypes = [P, P]
        L.orc_space_dets.argtypes = [P, P, P]
        L.orc_space_index.restype = c_ll
        L.orc_space_index.argtypes = [P, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_index1e.restype = ctypes.c_long
        L.orc_index1e.argtypes = [ctypes.c_long] * 2
        L.orc_index2e.restype = ctypes.c_long
        L.orc_index2e.argtypes = [ctypes.c_long] * 4
        L.orc_hij.restype = c_dbl
        L.orc_hij.argtypes = [ctypes.c_uint64] * 4 + [c_int, P, P, P, P]
        L.orc_hdiag.argtypes = [P, P, P, P, c_int]
        L.orc_row.restype = c_ll
        L.orc_row.argtypes = [P, c_ll, P, P, P, P, P, P, c_ll]
        L.orc_sigma.argtypes = [P, P, P, c_int, P, P, c_int]
        L.orc_sigma_rows.argtypes = [P, P, P, c_int, P, P, c_ll, P, c_int]
        L.orc_count_upper.restype = c_ll
        L.orc_count_upper.argtypes = [P, c_int]
        L.orc_jacobi.restype = c_int
        L.orc_jacobi.argtypes = [c_int, P, P, P]
        L.orc_davidson.restype = c_int
        L.orc_davidson.argtypes = [P, P, P, c_dbl, c_int, c_int, c_int, c_int, c_dbl, c_int,
                                   P, P, P, P, P, c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Space:
    """Determinant space: groups, strings and (p,q) blocks in reference order.

    rule 0 = serial truncation (source/citruncate.c:291-312),
    rule 1 = MRCI-SD truncation (mpi_source/citruncate.c:414-471).
    FCI is ndocc=0, nactv=norbs, xlvl=0 (either rule).
    """

    def __init__(self, na, nb, norbs, ndocc=0, nactv=None, xlvl=0, rule=1):
        if nactv is None:
            nactv = norbs - ndocc
        self.na, self.nb, self.norbs = na, nb, norbs
        self.ndocc, self.nactv, self.xlvl, self.rule = ndocc, nactv, xlvl, rule
        self._h = lib().orc_space_new(na, nb, norbs, ndocc, nactv, xlvl, rule)
        if not self._h:
            raise ValueError("oracle space construction failed")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_space_free(self._h)
            self._h = None

    @property
    def ndets(self):
        return lib().orc_space_ndets(self._h)

    def nstr(self, spin):
        return lib().orc_space_nstr(self._h, spin)

    def strings(self, spin):
        out = np.zeros(self.nstr(spin), dtype=np.uint64)
        lib().orc_space_strings(self._h, spin, _p(out))
        return out

    def groups(self, spin):
        n = lib().orc_space_ngroups(self._h, spin)
        out = np.zeros((n, 5), dtype=np.int64)
        lib().orc_space_groups(self._h, spin, _p(out))
        return out

    def blocks(self):
        n = lib().orc_space_nblocks(self._h)
        out = np.zeros((n, 3), dtype=np.int64)
        lib().orc_space_blocks(self._h, _p(out))
        return out

    def dets(self):
        a = np.zeros(self.ndets, dtype=np.uint64)
        b = np.zeros(self.ndets, dtype=np.uint64)
        lib().orc_space_dets(self._h, _p(a), _p(b))
        return a, b

    def index(self, ma, mb):
        return lib().orc_space_index(self._h, int(ma), int(mb))

    def hij(self, ia, ib, ja, jb, m1, m2):
        d = ctypes.c_int()
        ph = ctypes.c_int()
        v = lib().orc_hij(int(ia), int(ib), int(ja), int(jb), self.norbs, _p(m1), _p(m2),
                          ctypes.byref(d), ctypes.byref(ph))
        return v, d.value, ph.value

    def hdiag(self, m1, m2, nthreads=0):
        out = np.zeros(self.ndets)
        lib().orc_hdiag(self._h, _p(m1), _p(m2), _p(out), nthreads)
        return out

    def row(self, I, m1, m2):
        """Off-diagonal connected entries of row I (both triangles), sorted by column."""
        n = lib().orc_row(self._h, I, _p(m1), _p(m2), None, None, None, None, 0)
        cols = np.zeros(n, dtype=np.int64)
        vals = np.zeros(n)
        ph = np.zeros(n, dtype=np.int32)
        dg = np.zeros(n, dtype=np.int32)
        lib().orc_row(self._h, I, _p(m1), _p(m2), _p(cols), _p(vals), _p(ph), _p(dg), n)
        return cols, vals, ph, dg

    def sigma(self, m1, m2, v, nthreads=0):
        v = np.ascontiguousarray(v, dtype=np.float64)
        single = v.ndim == 1
        V = v.reshape(1, -1) if single else v
        out = np.zeros_like(V)
        lib().orc_sigma(self._h, _p(m1), _p(m2), V.shape[0], _p(V), _p(out), nthreads)
        return out[0] if single else out

    def sigma_rows(self, m1, m2, v, rows, nthreads=0):
        v = np.ascontiguousarray(v, dtype=np.float64)
        single = v.ndim == 1
        V = v.reshape(1, -1) if single else v
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        out = np.zeros((V.shape[0], len(rows)))
        lib().orc_sigma_rows(self._h, _p(m1), _p(m2), V.shape[0], _p(V), _p(rows), len(rows),
                             _p(out), nthreads)
        return out[0] if single else out

    def count_upper(self, nthreads=0):
        return lib().orc_count_upper(self._h, nthreads)

    def dense_h(self, m1, m2):
        """Dense H from the Slater-Condon rules (small spaces only)."""
        a, b = self.dets()
        N = len(a)
        H = np.zeros((N, N))
        for i in range(N):
            H[i, i] = self.hij(a[i], b[i], a[i], b[i], m1, m2)[0]
            cols, vals, _, _ = self.row(i, m1, m2)
            H[i, cols] = vals
        return H

    def davidson(self, m1, m2, totfrze=0.0, maxiter=200, krymin=4, krymax=12, nroots=1,
                 restol=1e-6, refdim=None, guess=None, nthreads=0):
        N = self.ndets
        if refdim is None:
            refdim = max(krymin, 16)
        cival = np.zeros(nroots)
        civec = np.zeros((nroots, N))
        rn = ctypes.c_double()
        ns = ctypes.c_longlong()
        g = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64)
        rc = lib().orc_davidson(self._h, _p(m1), _p(m2), totfrze, maxiter, krymin, krymax,
                                nroots, restol, refdim, _p(g), _p(cival), _p(civec),
                                ctypes.byref(rn), ctypes.byref(ns), nthreads)
        return rc, cival, civec, rn.value, ns.value


def jacobi(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    ev = np.zeros(n)
    vec = np.zeros((n, n))
    lib().orc_jacobi(n, _p(a), _p(ev), _p(vec))
    return ev, vec
