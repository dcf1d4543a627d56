"""ctypes binding of libparagraph_host.so (include/paragraph_host.h): the host
stages that complete an index around the GPU stage.

Reference routines (tests/test_t2iend.cu): MatchNN's entry point
(:179-221), MatchSup (:365-488), FusionNN (:490-537), FusionFinal
(:539-613), SaveGraph (:615-633).  Lists come back as CSR (off, ids).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libparagraph_host.so")
_lib = None

_P = ctypes.c_void_p
_U32 = ctypes.c_uint32

EXPORTED = (
    "pg_host_params_init", "pg_host_entry_point", "pg_host_match_sup",
    "pg_host_bipartite", "pg_host_final_width", "pg_host_fusion_final",
    "pg_host_save_index", "pg_host_bin_shape", "pg_host_read_bin",
    "pg_host_pad_gt", "pg_host_last_error",
)


class HostParams(ctypes.Structure):
    _fields_ = [(n, _U32) for n in (
        "n_base", "dim", "n_query", "gt_k", "M_nn", "rounds", "M_final",
        "threads")]


class HostError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise HostError(f"{_LIB_PATH} is not built (run "
                            "python -m paragraph_amd.build)")
        L = ctypes.CDLL(_LIB_PATH)
        L.pg_host_params_init.argtypes = [ctypes.POINTER(HostParams), _U32,
                                          _U32, _U32]
        L.pg_host_entry_point.argtypes = [ctypes.POINTER(HostParams), _P,
                                          ctypes.POINTER(_U32)]
        L.pg_host_match_sup.argtypes = [ctypes.POINTER(HostParams), _P, _P,
                                        _U32, _P, _P]
        L.pg_host_bipartite.argtypes = [ctypes.POINTER(HostParams), _P, _P,
                                        _P, _P]
        L.pg_host_final_width.argtypes = [ctypes.POINTER(HostParams)]
        L.pg_host_final_width.restype = _U32
        L.pg_host_fusion_final.argtypes = [ctypes.POINTER(HostParams), _P, _P,
                                           _P, _P, _U32, _P, _U32, _P, _P]
        L.pg_host_save_index.argtypes = [ctypes.c_char_p, _U32, _U32, _P, _P]
        L.pg_host_bin_shape.argtypes = [ctypes.c_char_p, ctypes.POINTER(_U32),
                                        ctypes.POINTER(_U32)]
        L.pg_host_read_bin.argtypes = [ctypes.c_char_p, ctypes.c_int, _U32,
                                       ctypes.c_uint64, ctypes.c_uint64, _P]
        L.pg_host_pad_gt.argtypes = [_P, ctypes.c_uint64, _U32, _U32, _P]
        L.pg_host_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise HostError(f"rc={rc}: {lib().pg_host_last_error().decode()}")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    a = np.asarray(a)
    if a.dtype == np.int32:
        a = a.view(np.uint32)
    return np.ascontiguousarray(a, dtype=np.uint32)


def make_params(n_base, dim, n_query, **overrides):
    """The reference's values (M_nn 40, 6 rounds, M_final 70, gt_k 128),
    then overrides."""
    p = HostParams()
    _check(lib().pg_host_params_init(ctypes.byref(p), n_base, dim, n_query))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, int(v))
    return p


def entry_point(p, gt):
    gt = _u32(gt)
    ep = _U32(0)
    _check(lib().pg_host_entry_point(ctypes.byref(p), _ptr(gt),
                                     ctypes.byref(ep)))
    return int(ep.value)


def match_sup(p, base, gt, topn):
    base, gt = _f32(base), _u32(gt)
    off = np.empty(p.n_base + 1, np.uint32)
    ids = np.empty(max(p.n_base * p.M_nn, 1), np.uint32)
    _check(lib().pg_host_match_sup(ctypes.byref(p), _ptr(base), _ptr(gt),
                                   topn, _ptr(off), _ptr(ids)))
    return off, ids[:off[-1]].copy()


def bipartite(p, base, gt):
    base, gt = _f32(base), _u32(gt)
    off = np.empty(p.n_base + 1, np.uint32)
    ids = np.empty(max(p.n_base * 2 * p.M_nn, 1), np.uint32)
    _check(lib().pg_host_bipartite(ctypes.byref(p), _ptr(base), _ptr(gt),
                                   _ptr(off), _ptr(ids)))
    return off, ids[:off[-1]].copy()


def fusion_final(p, base, bi_off, bi_ids, link, proj):
    base = _f32(base)
    bi_off, bi_ids = _u32(bi_off), _u32(bi_ids)
    link, proj = _u32(link), _u32(proj)
    width = int(lib().pg_host_final_width(ctypes.byref(p)))
    off = np.empty(p.n_base + 1, np.uint32)
    ids = np.empty(max(p.n_base * width, 1), np.uint32)
    _check(lib().pg_host_fusion_final(
        ctypes.byref(p), _ptr(base), _ptr(bi_off), _ptr(bi_ids), _ptr(link),
        link.shape[1], _ptr(proj), proj.shape[1], _ptr(off), _ptr(ids)))
    return off, ids[:off[-1]].copy()


def save_index(path, ep, off, ids):
    off, ids = _u32(off), _u32(ids)
    _check(lib().pg_host_save_index(os.fsencode(path), ep, off.shape[0] - 1,
                                    _ptr(off), _ptr(ids)))


def bin_shape(path):
    """(rows, cols) from a Big-ANN .fbin / .ibin header."""
    r, c = _U32(), _U32()
    _check(lib().pg_host_bin_shape(os.fsencode(path), ctypes.byref(r),
                                   ctypes.byref(c)))
    return int(r.value), int(c.value)


def read_bin(path, dtype, headered=True, cols=0, row0=0, nrows=None,
             out=None):
    """Rows [row0, row0 + nrows) of a .fbin / .ibin file (headered) or of a
    headerless array of `cols` columns, read by the C++ reader into `out`
    (a C-contiguous array, e.g. pinned staging memory) or a new array."""
    dtype = np.dtype(dtype)
    if dtype.itemsize != 4:
        raise ValueError("4-byte elements (float32 / int32 / uint32)")
    if headered:
        rows, cols_h = bin_shape(path)
        cols = cols or cols_h
    else:
        if not cols:
            raise ValueError("a headerless file needs cols")
        rows = os.path.getsize(path) // (4 * cols)
    if nrows is None:
        nrows = rows - row0
    if out is None:
        out = np.empty((nrows, cols), dtype)
    if out.size != nrows * cols or not out.flags.c_contiguous:
        raise ValueError("out must be C-contiguous with nrows x cols elements")
    _check(lib().pg_host_read_bin(os.fsencode(path), 1 if headered else 0,
                                  cols, row0, nrows, _ptr(out)))
    return out


def pad_gt(gt, width=128):
    """gt rows padded (or truncated) to `width` columns with -1 (int32), by
    the C++ routine."""
    g = np.ascontiguousarray(gt, dtype=np.int32)
    out = np.empty((g.shape[0], width), np.int32)
    _check(lib().pg_host_pad_gt(_ptr(g), g.shape[0], g.shape[1], width,
                                _ptr(out)))
    return out


def default_threads():
    """OpenMP threads the host stages use with threads = 0."""
    v = os.environ.get("OMP_NUM_THREADS")
    return int(v) if v and v.isdigit() else (os.cpu_count() or 1)
