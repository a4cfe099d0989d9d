"""ctypes view of the reference-named drop-in entry points of libjayci_gpu.

The structures follow source/include/binarystr.h:6-19 and source/include/cimapping.h:7-17
(the layouts the C ABI in include/jayci_gpu.h keeps).  These wrappers exist for tests and
for Python callers that drive the reference-style flow:
citrunc_rtnlist -> generate_cimap -> compute_hdgls -> dvdalg -> print_civectors.
"""
import ctypes

import numpy as np

from . import lib

_L = lib()
_P = ctypes.c_void_p


class OccStr(ctypes.Structure):
    _fields_ = [("byte1", ctypes.c_longlong), ("byte2", ctypes.c_longlong),
                ("virtx", ctypes.c_int * 2), ("nvrtx", ctypes.c_int)]


class Det(ctypes.Structure):
    _fields_ = [("astr", OccStr), ("bstr", OccStr), ("cas", ctypes.c_int)]


class Section(ctypes.Structure):
    pass


Section._fields_ = [("first", ctypes.c_int), ("last", ctypes.c_int),
                    ("next", ctypes.POINTER(Section)), ("prev", ctypes.POINTER(Section))]


class RowMap(ctypes.Structure):
    _fields_ = [("nsec", ctypes.c_int), ("sec", ctypes.POINTER(Section))]


_libc = ctypes.CDLL(None)
_libc.malloc.restype = _P
_libc.malloc.argtypes = [ctypes.c_size_t]
_libc.free.argtypes = [_P]

_ip = ctypes.POINTER(ctypes.c_int)
_L.citrunc_rtnlist.restype = ctypes.POINTER(Det)
_L.citrunc_rtnlist.argtypes = [ctypes.c_int] * 8 + [_ip] * 6
_L.generate_cimap.restype = ctypes.c_int
_L.generate_cimap.argtypes = [ctypes.POINTER(Det), ctypes.c_int, ctypes.c_int, _P]
_L.deallocate_cimap.restype = ctypes.c_int
_L.deallocate_cimap.argtypes = [_P, ctypes.c_int]
_L.compute_hdgls.restype = None
_L.compute_hdgls.argtypes = [ctypes.POINTER(Det), ctypes.c_int, _P, _P, ctypes.c_int,
                             ctypes.c_int, _P, ctypes.c_int]
_L.hmatels.restype = ctypes.c_double
_L.hmatels.argtypes = [Det, Det, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
_L.compute_hv.restype = None
_L.compute_hv.argtypes = [ctypes.POINTER(Det), ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int,
                          _P, _P, ctypes.c_int, _P]
_L.dvdalg.restype = ctypes.c_int
_L.dvdalg.argtypes = [ctypes.POINTER(Det), ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int, _P,
                      ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      ctypes.c_int, ctypes.c_double, _P, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)),
                      _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
_L.print_civectors.restype = ctypes.c_int
_L.print_civectors.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.c_int,
                               ctypes.c_int, _P]
_L.read_civectors.restype = ctypes.c_int
_L.read_civectors.argtypes = _L.print_civectors.argtypes
_L.write_wavefunction.restype = None
_L.write_wavefunction.argtypes = [ctypes.POINTER(Det), ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), _P,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int]
_L.readgeninput.restype = None
_L.readgeninput.argtypes = [_ip] * 10
_L.readdaiinput.restype = None
_L.readdaiinput.argtypes = [_ip] * 6 + [ctypes.POINTER(ctypes.c_double), _ip]


def _p(a):
    return a.ctypes.data_as(_P)


def occ_mask(o):
    m = int(o.byte1) & 0xFFFFFFFFFFFFFFFF
    for k in range(o.nvrtx):
        m |= 1 << (o.virtx[k] - 1)
    return m


def citrunc_rtnlist(aelec, belec, orbs, nfrzc, ndocc, nactv, nfrzv, xlvl):
    out = [ctypes.c_int() for _ in range(6)]
    dl = _L.citrunc_rtnlist(aelec, belec, orbs, nfrzc, ndocc, nactv, nfrzv, xlvl,
                            *[ctypes.byref(x) for x in out])
    vals = [x.value for x in out]
    return dl, dict(zip(["astr_len", "bstr_len", "ndets", "ci_orbs", "ci_aelec", "ci_belec"], vals))


def det_bytes(dl, n):
    """Raw bytes of a struct det array (for bit-exact comparisons)."""
    return ctypes.string_at(ctypes.cast(dl, _P), ctypes.sizeof(Det) * n)


def free(ptr):
    _libc.free(ctypes.cast(ptr, _P))


def row_pointers(mat):
    """double** view of a C-contiguous 2-D array (rows)."""
    rows = (ctypes.POINTER(ctypes.c_double) * mat.shape[0])()
    for i in range(mat.shape[0]):
        rows[i] = mat[i].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return rows


def generate_cimap(dl, ndets, ninto):
    hmap = _libc.malloc(ctypes.sizeof(RowMap) * ndets)
    rc = _L.generate_cimap(dl, ndets, ninto, hmap)
    return rc, hmap


def cimap_sections(hmap, ndets):
    rows = (RowMap * ndets).from_address(hmap)
    out = []
    for i in range(ndets):
        s = rows[i].sec
        lst = []
        while s:
            lst.append((s.contents.first, s.contents.last))
            s = s.contents.next
        out.append(lst)
    return out


def deallocate_cimap(hmap, ndets):
    return _L.deallocate_cimap(hmap, ndets)


def readgeninput():
    v = [ctypes.c_int() for _ in range(10)]
    _L.readgeninput(*[ctypes.byref(x) for x in v])
    keys = ["electrons", "orbitals", "nfrozen", "ndocc", "nactive", "xlevel", "nfrzvirt",
            "printlvl", "printwvf", "err"]
    return dict(zip(keys, [x.value for x in v]))


def readdaiinput():
    v = [ctypes.c_int() for _ in range(6)]
    r = ctypes.c_double()
    e = ctypes.c_int()
    _L.readdaiinput(*[ctypes.byref(x) for x in v], ctypes.byref(r), ctypes.byref(e))
    keys = ["maxiter", "krymin", "krymax", "nroots", "prediagr", "refdim"]
    d = dict(zip(keys, [x.value for x in v]))
    d["restol"] = r.value
    d["err"] = e.value
    return d
