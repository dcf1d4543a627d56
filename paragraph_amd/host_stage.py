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
    "pg_host_save_index", "pg_host_last_error",
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


def default_threads():
    """OpenMP threads the host stages use with threads = 0."""
    v = os.environ.get("OMP_NUM_THREADS")
    return int(v) if v and v.isdigit() else (os.cpu_count() or 1)
