"""ctypes bindings of oracle/_ref/libjayci_ref.so, the reference's own serial sources
(source/{citruncate,straddress,combinatorial,binarystr,bitutil,action_util,cimapping,
moindex,arrayutil,errorlib,iminmax,allocate_mem,binary,abecalc}.c) compiled as they lie
by oracle/Makefile (target `ref`).

TEST INFRASTRUCTURE ONLY.  Used to pin the oracle (determinant order, connectivity map,
matrix elements where the serial rules are right) and as the reference CPU timing.
The Davidson driver (source/davidson.c) needs LAPACK, which is not in this image, so it
is not part of this build.
"""
import ctypes
import os

import numpy as np

from . import REF_LIB_PATH


class OccStr(ctypes.Structure):
    # source/include/binarystr.h:6-11
    _fields_ = [("byte1", ctypes.c_longlong), ("byte2", ctypes.c_longlong),
                ("virtx", ctypes.c_int * 2), ("nvrtx", ctypes.c_int)]


class Det(ctypes.Structure):
    # source/include/binarystr.h:14-19
    _fields_ = [("astr", OccStr), ("bstr", OccStr), ("cas", ctypes.c_int)]


class Section(ctypes.Structure):
    pass


Section._fields_ = [("first", ctypes.c_int), ("last", ctypes.c_int),
                    ("next", ctypes.POINTER(Section)), ("prev", ctypes.POINTER(Section))]


class RowMap(ctypes.Structure):
    # source/include/cimapping.h:7-17
    _fields_ = [("nsec", ctypes.c_int), ("sec", ctypes.POINTER(Section))]


_lib = None
_libc = ctypes.CDLL(None)
_libc.malloc.restype = ctypes.c_void_p
_libc.malloc.argtypes = [ctypes.c_size_t]
_libc.free.argtypes = [ctypes.c_void_p]


def available():
    return os.path.exists(REF_LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(REF_LIB_PATH)
        ip = ctypes.POINTER(ctypes.c_int)
        L.citrunc_rtnlist.restype = ctypes.POINTER(Det)
        L.citrunc_rtnlist.argtypes = [ctypes.c_int] * 8 + [ip] * 6
        L.hmatels.restype = ctypes.c_double
        L.hmatels.argtypes = [Det, Det, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int]
        L.generate_cimap.restype = ctypes.c_int
        L.generate_cimap.argtypes = [ctypes.POINTER(Det), ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p]
        L.deallocate_cimap.restype = ctypes.c_int
        L.deallocate_cimap.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.compute_hv.argtypes = [ctypes.POINTER(Det), ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.c_void_p]
        L.compute_hdgls.argtypes = [ctypes.POINTER(Det), ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def occstr_mask(o):
    """Full orbital mask of a reference occstr (internal bits + virtual orbital list)."""
    m = int(o.byte1) & 0xFFFFFFFFFFFFFFFF
    for k in range(o.nvrtx):
        m |= 1 << (o.virtx[k] - 1)
    return m


class RefSpace:
    """Determinant list from the reference's citrunc_rtnlist (source/citruncate.c:123)."""

    def __init__(self, aelec, belec, orbs, ndocc, nactv, xlvl, nfrzc=0, nfrzv=0):
        L = lib()
        out = [ctypes.c_int() for _ in range(6)]
        self.dlist = L.citrunc_rtnlist(aelec, belec, orbs, nfrzc, ndocc, nactv, nfrzv, xlvl,
                                       *[ctypes.byref(x) for x in out])
        self.astr_len, self.bstr_len, self.ndets, self.ci_orbs, self.ci_aelec, self.ci_belec = \
            [x.value for x in out]
        self.aelec, self.belec = self.ci_aelec, self.ci_belec
        self.ninto = ndocc + nactv

    def masks(self):
        a = np.zeros(self.ndets, dtype=np.uint64)
        b = np.zeros(self.ndets, dtype=np.uint64)
        for i in range(self.ndets):
            d = self.dlist[i]
            a[i] = occstr_mask(d.astr)
            b[i] = occstr_mask(d.bstr)
        return a, b

    def cas_flags(self):
        return np.array([self.dlist[i].cas for i in range(self.ndets)], dtype=np.int32)

    def hmatels(self, i, j, m1, m2):
        return lib().hmatels(self.dlist[i], self.dlist[j], _p(m1), _p(m2), self.aelec,
                             self.belec, self.ninto)

    def hdgls(self, m1, m2):
        out = np.zeros(self.ndets)
        lib().compute_hdgls(self.dlist, self.ndets, _p(m1), _p(m2), self.aelec, self.belec,
                            _p(out), self.ninto)
        return out

    def cimap(self):
        """Run generate_cimap; returns (handle, sections) with sections[i] a list of
        (first, last) offsets relative to row i, exactly as the reference stores them."""
        L = lib()
        hmap = _libc.malloc(ctypes.sizeof(RowMap) * self.ndets)
        rc = L.generate_cimap(self.dlist, self.ndets, self.ninto, hmap)
        if rc != 0:
            raise RuntimeError("generate_cimap failed")
        rows = (RowMap * self.ndets).from_address(hmap)
        secs = []
        for i in range(self.ndets):
            s = rows[i].sec
            lst = []
            while s:
                lst.append((s.contents.first, s.contents.last))
                s = s.contents.next
            secs.append(lst)
        return hmap, secs

    def free_cimap(self, hmap):
        lib().deallocate_cimap(hmap, self.ndets)

    def compute_hv(self, m1, m2, v, hmap):
        c = np.zeros(self.ndets)
        lib().compute_hv(self.dlist, self.ndets, _p(m1), _p(m2), self.aelec, self.belec,
                         _p(np.ascontiguousarray(v)), _p(c), self.ninto, hmap)
        return c
