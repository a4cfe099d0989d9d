"""jayci_amd: MI355X sigma (H v) build and Davidson eigensolver of JayCI.

Python mirror of the C ABI in include/jayci_gpu.h (libjayci_gpu.so, built in-tree by
jayci_amd.build).  The reference path this replaces is source/davidson.c:91 (dvdalg) and
source/action_util.c:121 (compute_hv) with their determinant-space and integral inputs.
Everything here runs on the GPU through the library; there is no CPU fallback: if the
library is missing, importing raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjayci_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "jayci_gpu.h")

RULE_SERIAL, RULE_MRCISD = 0, 1
PAIRS_FP64, PAIRS_FP32, PAIRS_UPPER, DIRECT_TABLES = 1, 2, 4, 8
(SIGMA_AUTO, SIGMA_PAIRS64, SIGMA_PAIRS32, SIGMA_PAIRS_SYM64, SIGMA_DIRECT64, SIGMA_DIRECT32,
 SIGMA_FAST64, SIGMA_FAST32, SIGMA_FCI64, SIGMA_FCI32, SIGMA_GEMM64, SIGMA_GEMM32) = range(12)
NMETHODS = 12
PREC_FP64, PREC_MIXED = 0, 1
RESTART_RITZ, RESTART_GDK = 0, 1
GUESS_BLOCK, GUESS_USER, GUESS_DIAG, GUESS_UNIT = 1, 2, 3, 4
MAXROOTS = 8
SYNTH_ENUC = 10.0

if not os.path.exists(LIB_PATH):
    raise ImportError("jayci_amd: %s is not built (run python jayci_amd/build.py)" % LIB_PATH)

_L = ctypes.CDLL(LIB_PATH)
_P = ctypes.c_void_p
_i, _ll, _d = ctypes.c_int, ctypes.c_longlong, ctypes.c_double


class SolveParams(ctypes.Structure):
    _fields_ = [("maxiter", _i), ("krymin", _i), ("krymax", _i), ("nroots", _i), ("refdim", _i),
                ("restol", _d), ("precision", _i), ("mixed_switch", _d), ("sigma_fp64", _i),
                ("sigma_fp32", _i), ("guess", _i), ("totfrze", _d), ("verbose", _i),
                ("nbuffer", _i), ("etol", _d), ("buftol", _d), ("restart", _i)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", _i), ("converged", _i), ("nbasis", _i), ("nsigma64", _ll),
                ("nsigma32", _ll), ("rnorm", _d * MAXROOTS), ("seconds", _d),
                ("sigma_seconds", _d)]


def _sig(name, res, args):
    f = getattr(_L, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("jci_last_error", ctypes.c_char_p, [])
_sig("jci_device_count", _i, [ctypes.POINTER(_i)])
_sig("jci_device_memory", _i, [_i, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)])
_sig("jci_ctx_create", _i, [_i, ctypes.POINTER(_P)])
_sig("jci_ctx_destroy", None, [_P])
_sig("jci_ctx_sync", _i, [_P])
_sig("jci_ctx_stream", _P, [_P])
_sig("jci_space_create", _i, [_P, _i, _i, _i, _i, _i, _i, _i, ctypes.POINTER(_P)])
_sig("jci_space_destroy", None, [_P])
_sig("jci_space_ndets", _ll, [_P])
_sig("jci_space_nstrings", _ll, [_P, _i])
_sig("jci_space_ngroups", _i, [_P, _i])
_sig("jci_space_groups", _i, [_P, _i, _P])
_sig("jci_space_nblocks", _i, [_P])
_sig("jci_space_blocks", _i, [_P, _P])
_sig("jci_space_strings", _i, [_P, _i, _P])
_sig("jci_space_dets", _i, [_P, _ll, _ll, _P, _P])
_sig("jci_space_index", _ll, [_P, ctypes.c_uint64, ctypes.c_uint64])
_sig("jci_ham_create", _i, [_P, _P, _P, _i, ctypes.POINTER(_P)])
_sig("jci_ham_destroy", None, [_P])
_sig("jci_ham_npairs", _ll, [_P])
_sig("jci_ham_hdiag", _i, [_P, _P])
_sig("jci_ham_lowest_diag", _i, [_P, _i, _P])
_sig("jci_ham_rows", _i, [_P, _P, _i, _ll, _P, _P, _P, _P, _P])
_sig("jci_ham_string_excitations", _ll, [_P, _i, _ll, _i, _ll, _P, _P, _P])
_sig("jci_sigma_dev_noalloc", _i, [_P, _i, _i, _P, _P])
_sig("jci_sigma_noalloc_capacity", _ll, [])
_sig("jci_sigma_dev", _i, [_P, _i, _i, _P, _P])
_sig("jci_sigma", _i, [_P, _i, _i, _P, _P])
_sig("jci_solve_defaults", None, [ctypes.POINTER(SolveParams)])
_sig("jci_solve", _i, [_P, ctypes.POINTER(SolveParams), _P, _P, _P, ctypes.POINTER(SolveStats)])
_sig("jci_synth_integrals", _i, [_i, ctypes.c_uint64, _P, _P])
_sig("jci_jacobi", _i, [_i, _P, _P, _P])
_sig("jci_gram", _i, [_P, _i, _i, _ll, _P, _P, _P])
_sig("jci_bench_sigma", _i, [_P, _i, _i, _i, ctypes.POINTER(_d)])
_sig("jci_bench_copy", _i, [_P, _ll, _i, ctypes.POINTER(_d)])
_sig("jci_ham_autotune", _i, [_P, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), _P])
_sig("jci_ham_fast_width", _i, [_P])
_sig("jci_ham_fast_tiles", _i, [_P, ctypes.POINTER(_i), ctypes.POINTER(_i)])
_sig("jci_ham_fast_split", _i, [_P, ctypes.POINTER(_i)])
_sig("jci_ham_fast_xcd", _i, [_P, ctypes.POINTER(_i), ctypes.POINTER(_i)])
_sig("jci_batch_create", _i, [_P, _i, _i, ctypes.POINTER(_P)])
_sig("jci_batch_destroy", None, [_P])
_sig("jci_batch_set_graphs", _i, [_P, _i])
_sig("jci_batch_set_pool", _i, [_P, _ll])
_sig("jci_batch_set_tiles", _i, [_P, _i, _i])
_sig("jci_batch_tune_tiles", _i, [_P, _P, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i)])
_sig("jci_ham_gemm_tiles", _i, [_P, ctypes.POINTER(_i), ctypes.POINTER(_i)])
_sig("jci_batch_bytes", _ll, [_ll, _i])
_sig("jci_batch_add", _i, [_P, _P, ctypes.POINTER(SolveParams), ctypes.POINTER(_i)])
_sig("jci_batch_step", _i, [_P, ctypes.POINTER(_i)])
_sig("jci_batch_run", _i, [_P, _i])
_sig("jci_batch_result", _i, [_P, _i, ctypes.POINTER(_i), _P, _P, ctypes.POINTER(SolveStats)])
_sig("jci_batch_release", _i, [_P, _i])
_sig("jci_batch_finished", _i, [_P, _P, _i])
_sig("jci_ham_prepare_guess", _i, [_P, _P])
_sig("jci_batch_sigma_seconds", _d, [_P])
_sig("jci_batch_phase_seconds", _i, [_P, _P])
_sig("jci_write_vectors_ufmt", _i, [ctypes.c_char_p, _i, _i, _P])
_sig("jci_readmoints", _i, [ctypes.c_char_p, _i, _i, _P, _ll, _P, _ll, _P])
_sig("jci_read_vectors_ufmt", _i, [ctypes.c_char_p, _i, _i, _P])
_sig("jci_pdavidson", _i, [_P, ctypes.POINTER(SolveParams), _i, ctypes.c_char_p, _P,
                           ctypes.POINTER(SolveStats)])


class JciError(RuntimeError):
    pass


def last_error():
    return _L.jci_last_error().decode(errors="replace")


def _check(rc, what):
    if rc != 0:
        raise JciError("%s failed (%d): %s" % (what, rc, last_error()))


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def lib():
    return _L


def header_functions():
    """Names of the functions include/jayci_gpu.h declares."""
    import re
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt)
    skip = {"if", "sizeof", "defined", "struct", "return"}
    decl = []
    for m in re.finditer(r"^[A-Za-z_][^;#{}]*?\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", txt,
                         flags=re.M | re.S):
        decl.append(m.group(1))
    return sorted(set(d for d in decl if d not in skip))


def m1_len(n):
    return n * (n + 1) // 2


def m2_len(n):
    a = m1_len(n)
    return a * (a + 1) // 2


def synth_integrals(norbs, seed):
    """Synthetic packed integrals (see jci_synth_integrals in include/jayci_gpu.h)."""
    m1 = np.zeros(m1_len(norbs))
    m2 = np.zeros(m2_len(norbs))
    _check(_L.jci_synth_integrals(norbs, seed, _p(m1), _p(m2)), "jci_synth_integrals")
    return m1, m2


def jacobi(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    ev = np.zeros(n)
    vec = np.zeros((n, n))
    _check(_L.jci_jacobi(n, _p(a), _p(ev), _p(vec)), "jci_jacobi")
    return ev, vec


def device_memory(device=0):
    """(free, total) device memory in bytes."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    _check(_L.jci_device_memory(device, ctypes.byref(f), ctypes.byref(t)), "jci_device_memory")
    return f.value, t.value


def device_count():
    n = _i()
    _L.jci_device_count(ctypes.byref(n))
    return n.value


class Context:
    def __init__(self, device=0):
        h = _P()
        _check(_L.jci_ctx_create(device, ctypes.byref(h)), "jci_ctx_create")
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            _L.jci_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def sync(self):
        _check(_L.jci_ctx_sync(self._h), "jci_ctx_sync")

    @property
    def stream(self):
        return _L.jci_ctx_stream(self._h)

    def bench_copy(self, nbytes, reps=20):
        """Mean ms of a device copy of nbytes (HIP events on the context stream)."""
        ms = _d()
        _check(_L.jci_bench_copy(self._h, nbytes, reps, ctypes.byref(ms)), "jci_bench_copy")
        return ms.value

    def gram(self, V, W):
        """V^T W on the FP64 matrix cores; V (m, n), W (w, n)."""
        V = np.ascontiguousarray(V, dtype=np.float64)
        W = np.ascontiguousarray(W, dtype=np.float64)
        G = np.zeros((V.shape[0], W.shape[0]))
        _check(_L.jci_gram(self._h, V.shape[0], W.shape[0], V.shape[1], _p(V), _p(W), _p(G)),
               "jci_gram")
        return G


class Space:
    """Determinant space: alpha/beta string tables on the device (reference order)."""

    def __init__(self, ctx, na, nb, norbs, ndocc=0, nactv=None, xlvl=0, rule=RULE_MRCISD):
        if nactv is None:
            nactv = norbs - ndocc
        self.ctx = ctx
        self.na, self.nb, self.norbs = na, nb, norbs
        self.ndocc, self.nactv, self.xlvl, self.rule = ndocc, nactv, xlvl, rule
        h = _P()
        _check(_L.jci_space_create(ctx._h, na, nb, norbs, ndocc, nactv, xlvl, rule, ctypes.byref(h)),
               "jci_space_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _L.jci_space_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def ndets(self):
        return _L.jci_space_ndets(self._h)

    def nstrings(self, spin):
        return _L.jci_space_nstrings(self._h, spin)

    def groups(self, spin):
        n = _L.jci_space_ngroups(self._h, spin)
        out = np.zeros((n, 5), dtype=np.int64)
        _L.jci_space_groups(self._h, spin, _p(out))
        return out

    def blocks(self):
        n = _L.jci_space_nblocks(self._h)
        out = np.zeros((n, 3), dtype=np.int64)
        _L.jci_space_blocks(self._h, _p(out))
        return out

    def strings(self, spin):
        out = np.zeros(self.nstrings(spin), dtype=np.uint64)
        _check(_L.jci_space_strings(self._h, spin, _p(out)), "jci_space_strings")
        return out

    def dets(self, first=0, count=None):
        if count is None:
            count = self.ndets - first
        a = np.zeros(count, dtype=np.uint64)
        b = np.zeros(count, dtype=np.uint64)
        _check(_L.jci_space_dets(self._h, first, count, _p(a), _p(b)), "jci_space_dets")
        return a, b

    def index(self, a, b):
        return _L.jci_space_index(self._h, int(a), int(b))


class Hamiltonian:
    """Integrals + diagonal on the device, with an optional materialised pair list."""

    def __init__(self, space, m1, m2, flags=0):
        self.space = space
        self.m1 = np.ascontiguousarray(m1, dtype=np.float64)
        self.m2 = np.ascontiguousarray(m2, dtype=np.float64)
        if len(self.m1) < m1_len(space.norbs) or len(self.m2) < m2_len(space.norbs):
            raise ValueError("integral arrays too short for %d orbitals" % space.norbs)
        h = _P()
        _check(_L.jci_ham_create(space._h, _p(self.m1), _p(self.m2), flags, ctypes.byref(h)),
               "jci_ham_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _L.jci_ham_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def npairs(self):
        return _L.jci_ham_npairs(self._h)

    def hdiag(self):
        out = np.zeros(self.space.ndets)
        _check(_L.jci_ham_hdiag(self._h, _p(out)), "jci_ham_hdiag")
        return out

    def prepare_guess(self, params):
        """Compute now the initial block of the next Batch.add with these parameters (on
        this Hamiltonian's context), so that the add only uploads it."""
        _check(_L.jci_ham_prepare_guess(self._h, ctypes.byref(params)), "jci_ham_prepare_guess")

    def lowest_diag(self, nb):
        """Indices of the nb lowest diagonal elements, ascending by (value, index)."""
        out = np.zeros(nb, dtype=np.int64)
        _check(_L.jci_ham_lowest_diag(self._h, nb, _p(out)), "jci_ham_lowest_diag")
        return out

    def rows(self, rows, cap=None):
        """Connected entries of each row: list of (cols, vals, phase, deg) sorted by column."""
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        if cap is None:
            na, nb, n = self.space.na, self.space.nb, self.space.norbs
            from math import comb
            sa = na * (n - na)
            sb = nb * (n - nb)
            cap = sa + sb + comb(na, 2) * comb(n - na, 2) + comb(nb, 2) * comb(n - nb, 2) + sa * sb + 1
        nr = len(rows)
        cols = np.zeros(nr * cap, dtype=np.int64)
        vals = np.zeros(nr * cap)
        ph = np.zeros(nr * cap, dtype=np.int32)
        dg = np.zeros(nr * cap, dtype=np.int32)
        cnt = np.zeros(nr, dtype=np.int64)
        _check(_L.jci_ham_rows(self._h, _p(rows), nr, cap, _p(cols), _p(vals), _p(ph), _p(dg),
                               _p(cnt)), "jci_ham_rows")
        out = []
        for t in range(nr):
            c = int(cnt[t])
            if c > cap:
                raise JciError("row capacity exceeded")
            sl = slice(t * cap, t * cap + c)
            o = np.argsort(cols[sl], kind="stable")
            out.append((cols[sl][o], vals[sl][o], ph[sl][o], dg[sl][o]))
        return out

    def string_excitations(self, spin, string, kind):
        """Excitation-table entries of one string: (targets, info, values) of its singles
        (kind 1) or doubles (kind 2); see jci_ham_string_excitations."""
        n = _L.jci_ham_string_excitations(self._h, spin, string, kind, 0, None, None, None)
        if n < 0:
            raise JciError("jci_ham_string_excitations failed: %s" % last_error())
        t = np.zeros(n, dtype=np.int64)
        info = np.zeros(n, dtype=np.int32)
        v = np.zeros(n)
        _L.jci_ham_string_excitations(self._h, spin, string, kind, n, _p(t), _p(info), _p(v))
        return t, info, v

    def bench_sigma(self, method=SIGMA_AUTO, nvec=1, reps=20):
        """Mean ms per sigma launch (HIP events on the context stream)."""
        ms = _d()
        _check(_L.jci_bench_sigma(self._h, method, nvec, reps, ctypes.byref(ms)), "jci_bench_sigma")
        return ms.value

    def autotune(self, nvec=1, reps=5):
        """Time all sigma variants; returns (best64, best32, {method: ms}, tile width)."""
        b64, b32 = _i(), _i()
        t = np.zeros(NMETHODS)
        _check(_L.jci_ham_autotune(self._h, nvec, reps, ctypes.byref(b64), ctypes.byref(b32), _p(t)),
               "jci_ham_autotune")
        return b64.value, b32.value, {m: t[m] for m in range(NMETHODS) if t[m] > 0}, \
            _L.jci_ham_fast_width(self._h)

    def fast_tiles(self, width=-1, chunk=-1):
        """Tile shape of the table sigma (jci_ham_fast_tiles): width 64 / 128 / 256, or 0 for
        per-block widths; -1 keeps.  Returns (width, chunk)."""
        w, c = _i(width), _i(chunk)
        _check(_L.jci_ham_fast_tiles(self._h, ctypes.byref(w), ctypes.byref(c)), "jci_ham_fast_tiles")
        return w.value, c.value

    def fast_split(self, split=-1):
        """Table sigma as one fused kernel (0) or two passes with every block in its own
        orientation (1), short-row blocks transposed (2) or every block transposed (3); -1
        queries (jci_ham_fast_split)."""
        v = _i(split)
        _check(_L.jci_ham_fast_split(self._h, ctypes.byref(v)), "jci_ham_fast_split")
        return v.value

    def fast_xcd(self, group=-1, chunk_major=-1):
        """Tile order of the table sigma (jci_ham_fast_xcd): tiles per XCD run and chunk-major
        order inside a block; -1 keeps.  Returns (group, chunk_major)."""
        v, c = _i(group), _i(chunk_major)
        _check(_L.jci_ham_fast_xcd(self._h, ctypes.byref(v), ctypes.byref(c)), "jci_ham_fast_xcd")
        return v.value, c.value

    def gemm_tiles(self, bn64=0, bn32=0):
        """E tile widths of the dense-string sigma; positive arguments set them."""
        a, b = _i(bn64), _i(bn32)
        _check(_L.jci_ham_gemm_tiles(self._h, ctypes.byref(a), ctypes.byref(b)),
               "jci_ham_gemm_tiles")
        return a.value, b.value

    def sigma(self, x, method=SIGMA_AUTO):
        """H x for x of shape (N,) or (k, N) (vector-major); returns the same shape."""
        x = np.asarray(x, dtype=np.float64)
        single = x.ndim == 1
        X = x.reshape(1, -1) if single else x
        k, N = X.shape
        if N != self.space.ndets:
            raise ValueError("vector length %d != ndets %d" % (N, self.space.ndets))
        Xi = np.ascontiguousarray(X.T)          # determinant-major [N][k]
        Yi = np.zeros_like(Xi)
        _check(_L.jci_sigma(self._h, method, k, _p(Xi), _p(Yi)), "jci_sigma")
        Y = np.ascontiguousarray(Yi.T)
        return Y[0] if single else Y

    def params(self, nroots=1, restol=1e-6, maxiter=200, krymin=None, krymax=None, refdim=64,
               precision=PREC_FP64, mixed_switch=1e-4, guess=GUESS_BLOCK, totfrze=0.0,
               sigma_fp64=SIGMA_AUTO, sigma_fp32=SIGMA_AUTO, verbose=False, nbuffer=0, etol=0.0,
               buftol=0.0, restart=RESTART_RITZ):
        return make_params(self.space.ndets, nroots, restol, maxiter, krymin, krymax, refdim,
                           precision, mixed_switch, guess, totfrze, sigma_fp64, sigma_fp32, verbose,
                           nbuffer, etol, buftol, restart)

    def solve(self, nroots=1, restol=1e-6, maxiter=200, krymin=None, krymax=None, refdim=64,
              precision=PREC_FP64, mixed_switch=1e-4, guess=GUESS_BLOCK, guess_vectors=None,
              totfrze=0.0, sigma_fp64=SIGMA_AUTO, sigma_fp32=SIGMA_AUTO, vectors=True,
              verbose=False, check=True, nbuffer=0, etol=0.0, buftol=0.0, restart=RESTART_RITZ):
        p = make_params(self.space.ndets, nroots, restol, maxiter, krymin, krymax, refdim,
                        precision, mixed_switch, guess, totfrze, sigma_fp64, sigma_fp32, verbose,
                        nbuffer, etol, buftol, restart)
        N = self.space.ndets
        ev = np.zeros(nroots)
        vec = np.zeros((nroots, N)) if vectors else None
        g = None
        if guess == GUESS_USER:
            g = np.ascontiguousarray(guess_vectors, dtype=np.float64)
            if g.ndim != 2 or g.shape[1] != N or g.shape[0] < p.krymin:
                raise ValueError("guess_vectors must hold krymin=%d vectors of %d, got shape %s"
                                 % (p.krymin, N, g.shape))
        st = SolveStats()
        rc = _L.jci_solve(self._h, ctypes.byref(p), _p(g), _p(ev), _p(vec), ctypes.byref(st))
        if check and rc != 0:
            raise JciError("jci_solve failed (%d): %s" % (rc, last_error()))
        stats = {"rc": rc, "iterations": st.iterations, "converged": bool(st.converged),
                 "nbasis": st.nbasis, "nsigma64": st.nsigma64, "nsigma32": st.nsigma32,
                 "rnorm": [st.rnorm[i] for i in range(nroots)], "seconds": st.seconds,
                 "sigma_seconds": st.sigma_seconds}
        return ev, vec, stats


    def pdavidson(self, prediagr=1, directory=None, nroots=1, restol=1e-6, maxiter=200,
                  krymin=None, krymax=None, refdim=64, precision=PREC_FP64, totfrze=0.0,
                  verbose=False, nbuffer=0, etol=0.0):
        """The MPI driver's Davidson step (jci_pdavidson): prediagr 1 block guess, 2 unit
        vectors, 3 directory/ci.restart; writes civec (and ci.restart, hvvec when not
        converged) into `directory` (None = working directory).  Returns (energies, stats)
        with stats["rc"] 0 converged or 1 not converged."""
        p = self.params(nroots=nroots, restol=restol, maxiter=maxiter, krymin=krymin,
                        krymax=krymax, refdim=refdim, precision=precision, totfrze=totfrze,
                        verbose=verbose, nbuffer=nbuffer, etol=etol)
        ev = np.zeros(nroots)
        st = SolveStats()
        d = None if directory is None else os.fsencode(directory)
        rc = _L.jci_pdavidson(self._h, ctypes.byref(p), prediagr, d, _p(ev), ctypes.byref(st))
        if rc not in (0, 1):
            raise JciError("jci_pdavidson failed (%d): %s" % (rc, last_error()))
        return ev, {"rc": rc, "iterations": st.iterations, "converged": bool(st.converged),
                    "rnorm": [st.rnorm[i] for i in range(nroots)]}


def _rows(a):
    """ctypes array of row pointers into a C-contiguous float64 matrix."""
    return (_P * a.shape[0])(*[a[i:].ctypes.data_as(_P) for i in range(a.shape[0])])


def write_vectors_ufmt(fname, vecs):
    """Write (dim, len) vectors in the MPI driver's binary layout (jci_write_vectors_ufmt)."""
    v = np.ascontiguousarray(np.atleast_2d(vecs), dtype=np.float64)
    rc = _L.jci_write_vectors_ufmt(os.fsencode(fname), v.shape[1], v.shape[0],
                                   ctypes.cast(_rows(v), _P))
    if rc:
        raise JciError("jci_write_vectors_ufmt failed (%d): %s" % (rc, last_error()))


def read_vectors_ufmt(fname, length, dim):
    """Read dim vectors of `length` doubles (jci_read_vectors_ufmt).  Raises JciError with
    the reference's code in .code on failure."""
    v = np.zeros((dim, length))
    rc = _L.jci_read_vectors_ufmt(os.fsencode(fname), length, dim, ctypes.cast(_rows(v), _P))
    if rc:
        e = JciError("jci_read_vectors_ufmt failed (%d): %s" % (rc, last_error()))
        e.code = rc
        raise e
    return v


def read_moints(path, norbs, type1n=1):
    """Read a SIFS integral file (jci_readmoints): returns (moints1, moints2, nuc_rep,
    frozen_core) with moints1 of m1_len(norbs) (3 arrays for type1n 3) and moints2 of
    m2_len(norbs), packed as the reference's index1e / index2e."""
    m1 = np.zeros(m1_len(norbs) * (3 if type1n == 3 else 1))
    m2 = np.zeros(m2_len(norbs))
    en = np.zeros(2)
    rc = _L.jci_readmoints(os.fsencode(path), type1n, norbs, _p(m1), m1_len(norbs), _p(m2),
                           m2_len(norbs), _p(en))
    if rc:
        raise JciError("jci_readmoints failed (%d): %s" % (rc, last_error()))
    return m1, m2, float(en[0]), float(en[1])


def make_params(ndets, nroots=1, restol=1e-6, maxiter=200, krymin=None, krymax=None, refdim=64,
                precision=PREC_FP64, mixed_switch=1e-4, guess=GUESS_BLOCK, totfrze=0.0,
                sigma_fp64=SIGMA_AUTO, sigma_fp32=SIGMA_AUTO, verbose=False, nbuffer=0,
                etol=0.0, buftol=0.0, restart=RESTART_RITZ):
    p = SolveParams()
    _L.jci_solve_defaults(ctypes.byref(p))
    p.nroots = nroots
    p.restol = restol
    p.maxiter = maxiter
    p.krymin = krymin if krymin is not None else max(nroots + 1, min(4, ndets))
    p.krymax = krymax if krymax is not None else min(max(p.krymin + 2 * nroots, 16), 64)
    p.refdim = refdim
    p.precision = precision
    p.mixed_switch = mixed_switch
    p.guess = guess
    p.totfrze = totfrze
    p.sigma_fp64 = sigma_fp64
    p.sigma_fp32 = sigma_fp32
    p.verbose = 1 if verbose else 0
    p.nbuffer = nbuffer
    p.etol = etol
    p.buftol = buftol
    p.restart = restart
    return p


class Batch:
    """Lockstep batched Davidson over many instances (continuous batching)."""

    def __init__(self, ctx, capacity, max_norbs=64, graphs=True, pool_bytes=None):
        self.ctx = ctx
        self.capacity = capacity
        h = _P()
        _check(_L.jci_batch_create(ctx._h, capacity, max_norbs, ctypes.byref(h)), "jci_batch_create")
        self._h = h
        self._keep = {}
        if not graphs:
            _L.jci_batch_set_graphs(h, 0)
        if pool_bytes is not None:
            _check(_L.jci_batch_set_pool(h, int(pool_bytes)), "jci_batch_set_pool")

    def tune_tiles(self, ham, nvec=1, reps=5):
        """E tile widths autotuned on a representative full-CI Hamiltonian
        (jci_batch_tune_tiles); returns (bn64, bn32)."""
        a, b = _i(), _i()
        _check(_L.jci_batch_tune_tiles(self._h, ham._h, nvec, reps, ctypes.byref(a),
                                       ctypes.byref(b)), "jci_batch_tune_tiles")
        return a.value, b.value

    def set_tiles(self, bn64=0, bn32=0):
        _check(_L.jci_batch_set_tiles(self._h, bn64, bn32), "jci_batch_set_tiles")

    def close(self):
        if getattr(self, "_h", None):
            _L.jci_batch_destroy(self._h)
            self._h = None
            self._keep = {}

    def __del__(self):
        self.close()

    @staticmethod
    def bytes_for(ndets, krymax):
        return _L.jci_batch_bytes(ndets, krymax)

    def add(self, ham, params):
        slot = _i()
        _check(_L.jci_batch_add(self._h, ham._h, ctypes.byref(params), ctypes.byref(slot)),
               "jci_batch_add")
        self._keep[slot.value] = (ham, params)
        return slot.value

    def step(self):
        n = _i()
        _check(_L.jci_batch_step(self._h, ctypes.byref(n)), "jci_batch_step")
        return n.value

    def run(self, max_steps=1000):
        _check(_L.jci_batch_run(self._h, max_steps), "jci_batch_run")

    def result(self, slot, vectors=False):
        ham, params = self._keep[slot]
        done = _i()
        ev = np.zeros(params.nroots)
        vec = np.zeros((params.nroots, ham.space.ndets)) if vectors else None
        st = SolveStats()
        _check(_L.jci_batch_result(self._h, slot, ctypes.byref(done), _p(ev), _p(vec),
                                   ctypes.byref(st)), "jci_batch_result")
        stats = {"done": done.value, "iterations": st.iterations, "converged": bool(st.converged),
                 "nsigma64": st.nsigma64, "nsigma32": st.nsigma32,
                 "rnorm": [st.rnorm[i] for i in range(params.nroots)], "seconds": st.seconds}
        return ev, vec, stats

    def finished(self):
        """Slots whose solve has ended and that are not released yet."""
        out = np.zeros(self.capacity, dtype=np.int32)
        n = _L.jci_batch_finished(self._h, _p(out), self.capacity)
        return [int(x) for x in out[:min(n, self.capacity)]]

    def release(self, slot):
        _check(_L.jci_batch_release(self._h, slot), "jci_batch_release")
        self._keep.pop(slot, None)

    @property
    def sigma_seconds(self):
        return _L.jci_batch_sigma_seconds(self._h)

    @property
    def phase_seconds(self):
        """Host seconds per step phase: sigma, Gram/Rayleigh-Ritz, residuals/decisions,
        selection/collapse/hand-over, orthonormalisation, append."""
        out = np.zeros(6)
        _check(_L.jci_batch_phase_seconds(self._h, _p(out)), "jci_batch_phase_seconds")
        return dict(zip(("sigma", "rayleigh_ritz", "residual", "collapse", "orthonormalise",
                         "append"), out.tolist()))
