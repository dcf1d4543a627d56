"""ctypes wrappers around oracle/libcpu_ref.so (the C++ restatement).

TEST INFRASTRUCTURE ONLY: see oracle/__init__.py.  Every wrapper takes and
returns numpy arrays; the C functions they call cite the reference routine
they restate (oracle/cpu_ref.cpp).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcpu_ref.so")
_lib = None

TOMB = 0xFFFFFFFF
MASKED_TOMB = 0x7FFFFFFF


class OrcStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "prune_dist", "prune_pairs", "merge_dist", "merge_pairs",
        "search_dist", "search_iters", "search_capped", "rev_overflow")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "n_base", "dim", "n_query", "gt_k", "proj_deg", "proj_pr", "proj_rev",
        "M", "link_rev", "Km", "Kp", "Kd", "topk", "reset_iter", "rounds",
        "entry_seed", "max_iters")]


def build():
    """Compile libcpu_ref.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64


def _declare(L):
    L.orc_chain_dist.restype = ctypes.c_float
    L.orc_chain_dist.argtypes = [_P, _P, _U32]
    L.orc_search_dist.restype = ctypes.c_float
    L.orc_search_dist.argtypes = [_P, _P, _U32]
    L.orc_match_winner.argtypes = [_P, _U64, _U32, _U32, _P]
    L.orc_top1_rows.argtypes = [_P, _U32, _P, _U32, _P]
    L.orc_prune.argtypes = [_P, _U32, _U32, _P, _U32, _U32, _P, _P, _P]
    L.orc_reverse_csr.argtypes = [_P, _P, _U32, _U32, _P, _P]
    L.orc_merge.argtypes = [_P, _U32, _U32, _P, _P, _U32, _P, _P, _U32, _U32,
                            ctypes.c_int, _P, _P]
    L.orc_search.argtypes = [_P, _U64, _P, _U32, _U32, _P, _U32, _U32, _U32,
                             _U32, _U32, _U32, _U32, _U32, _P, _P]
    L.orc_prune_nodes.argtypes = [_P, _U32, _U32, _P, _U32, _P, _U32, _U32,
                                  _P, _P]
    L.orc_merge_nodes.argtypes = [_P, _U32, _U32, _P, _U32, _P, _P, _U32, _P,
                                  _P, _U32, _U32, ctypes.c_int, _P]
    L.orc_gpu_stage.argtypes = [ctypes.POINTER(OrcParams), _P, _P, _P, _P, _P]
    L.orc_gpu_stage.restype = ctypes.c_int
    L.orc_max_threads.restype = ctypes.c_int
    L.orc_set_threads.argtypes = [ctypes.c_int]
    # host stages (oracle/cpu_fusion.cpp)
    L.orc_cpu_compare.restype = ctypes.c_float
    L.orc_cpu_compare.argtypes = [_P, _P, _U32]
    L.orc_rng_prune.restype = _U32
    L.orc_rng_prune.argtypes = [_P, _U32, _U32, _P, _P, _U32, _U32, _U32,
                                ctypes.c_int, _P]
    L.orc_entry_point.restype = _U32
    L.orc_entry_point.argtypes = [_P, _U32, _U32, _U32]
    L.orc_match_sup.argtypes = [_P, _U32, _U32, _P, _U32, _U32, _U32, _U32,
                                _P, _P]
    L.orc_bipartite.argtypes = [_P, _U32, _U32, _P, _U32, _U32, _U32, _U32,
                                _P, _P]
    L.orc_fusion_final.argtypes = [_P, _U32, _U32, _P, _P, _P, _U32, _P, _U32,
                                   _U32, _P, _P]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def chain_dist(x, y):
    x, y = _f32(x), _f32(y)
    return float(lib().orc_chain_dist(_ptr(x), _ptr(y), x.shape[0]))


def search_dist(q, x):
    q, x = _f32(q), _f32(x)
    return float(lib().orc_search_dist(_ptr(q), _ptr(x), q.shape[0]))


def match_winner(gt, n_base):
    gt = _u32(gt)
    w = np.empty(n_base, np.uint32)
    lib().orc_match_winner(_ptr(gt), gt.shape[0], gt.shape[1], n_base, _ptr(w))
    return w


def top1_rows(gt, winner):
    gt, winner = _u32(gt), _u32(winner)
    rows = np.empty((winner.shape[0], gt.shape[1]), np.uint32)
    lib().orc_top1_rows(_ptr(gt), gt.shape[1], _ptr(winner), winner.shape[0],
                        _ptr(rows))
    return rows


def prune(base, rows, mp, stats=None):
    base, rows = _f32(base), _u32(rows)
    n, dim = base.shape
    pruned = np.full((n, mp), TOMB, np.uint32)
    pcnt = np.zeros(n, np.uint32)
    lib().orc_prune(_ptr(base), n, dim, _ptr(rows), rows.shape[1], mp,
                    _ptr(pruned), _ptr(pcnt),
                    ctypes.byref(stats) if stats is not None else None)
    return pruned, pcnt


def reverse_csr(pruned, pcnt):
    pruned, pcnt = _u32(pruned), _u32(pcnt)
    n, mp = pruned.shape
    off = np.zeros(n + 1, np.uint32)
    srcs = np.zeros(max(int(pcnt.sum()), 1), np.uint32)
    lib().orc_reverse_csr(_ptr(pruned), _ptr(pcnt), n, mp, _ptr(off),
                          _ptr(srcs))
    return off, srcs[: int(off[-1])]


def merge(base, off, srcs, cap, pruned, pcnt, deg, strict, stats=None):
    base, off, pruned, pcnt = _f32(base), _u32(off), _u32(pruned), _u32(pcnt)
    srcs = _u32(srcs) if len(srcs) else np.zeros(1, np.uint32)
    n, dim = base.shape
    out = np.empty((n, deg), np.uint32)
    lib().orc_merge(_ptr(base), n, dim, _ptr(off), _ptr(srcs), cap,
                    _ptr(pruned), _ptr(pcnt), pruned.shape[1], deg,
                    1 if strict else 0, _ptr(out),
                    ctypes.byref(stats) if stats is not None else None)
    return out


def search(queries, base, graph, Km, Kp, Kd, topk, reset_iter=15,
           entry_seed=3030517, max_iters=100000, stats=None):
    queries, base, graph = _f32(queries), _f32(base), _u32(graph)
    nq = queries.shape[0]
    n, dim = base.shape
    out = np.empty((nq, topk), np.uint32)
    lib().orc_search(_ptr(queries), nq, _ptr(base), n, dim, _ptr(graph),
                     graph.shape[1], Km, Kp, Kd, topk, reset_iter, entry_seed,
                     max_iters, _ptr(out),
                     ctypes.byref(stats) if stats is not None else None)
    return out


def prune_nodes(base, nodes, rows_s, mp):
    """orc_prune restricted to `nodes` (rows_s: their candidate rows)."""
    base, nodes, rows_s = _f32(base), _u32(nodes), _u32(rows_s)
    n, dim = base.shape
    pruned = np.full((nodes.shape[0], mp), TOMB, np.uint32)
    pcnt = np.zeros(nodes.shape[0], np.uint32)
    lib().orc_prune_nodes(_ptr(base), n, dim, _ptr(nodes), nodes.shape[0],
                          _ptr(rows_s), rows_s.shape[1], mp, _ptr(pruned),
                          _ptr(pcnt))
    return pruned, pcnt


def merge_nodes(base, nodes, off, srcs, cap, pruned, pcnt, deg, strict):
    """orc_merge restricted to `nodes` (CSR and pruned lists cover all N)."""
    base, nodes, off = _f32(base), _u32(nodes), _u32(off)
    pruned, pcnt = _u32(pruned), _u32(pcnt)
    srcs = _u32(srcs) if len(srcs) else np.zeros(1, np.uint32)
    n, dim = base.shape
    out = np.empty((nodes.shape[0], deg), np.uint32)
    lib().orc_merge_nodes(_ptr(base), n, dim, _ptr(nodes), nodes.shape[0],
                          _ptr(off), _ptr(srcs), cap, _ptr(pruned), _ptr(pcnt),
                          pruned.shape[1], deg, 1 if strict else 0, _ptr(out))
    return out


def gpu_stage(params, base, gt, stats=None):
    """Whole stage; `params` is a dict with OrcParams field names."""
    base, gt = _f32(base), _u32(gt)
    p = OrcParams(**{k: int(params[k]) for k, _ in OrcParams._fields_})
    n = p.n_base
    proj = np.empty((n, p.proj_deg), np.uint32)
    link = np.empty((n, p.M), np.uint32)
    rc = lib().orc_gpu_stage(ctypes.byref(p), _ptr(base), _ptr(gt),
                             _ptr(proj), _ptr(link),
                             ctypes.byref(stats) if stats is not None else None)
    if rc != 0:
        raise RuntimeError(f"orc_gpu_stage failed: {rc}")
    return proj, link


# ------------------------------------------------------------ host stages
# Restatements of the reference's CPU routines around the GPU stage
# (oracle/cpu_fusion.cpp).  Lists are returned as CSR (off, ids).

def cpu_compare(x, y):
    """The reference CPU distance `compare` (tests/test_t2iend.cu:31-78)."""
    x, y = _f32(x), _f32(y)
    return float(lib().orc_cpu_compare(_ptr(x), _ptr(y), x.shape[0]))


def rng_prune(base, ids, dists, target, M, strict):
    """RNGPrune (tests/test_t2iend.cu:99-177) of candidates (ids, dists),
    sorted here by (distance, id)."""
    base = _f32(base)
    ids, dists = _u32(ids), _f32(dists)
    out = np.empty(max(int(M), 1), np.uint32)
    n = lib().orc_rng_prune(_ptr(base), base.shape[0], base.shape[1],
                            _ptr(ids), _ptr(dists), ids.shape[0], target, M,
                            1 if strict else 0, _ptr(out))
    return out[:n].copy()


def entry_point(gt, n_base):
    gt = _u32(gt)
    return int(lib().orc_entry_point(_ptr(gt), gt.shape[0], gt.shape[1],
                                     n_base))


def match_sup(base, gt, topn, M_nn):
    base, gt = _f32(base), _u32(gt)
    n = base.shape[0]
    off = np.empty(n + 1, np.uint32)
    ids = np.empty(max(n * M_nn, 1), np.uint32)
    lib().orc_match_sup(_ptr(base), n, base.shape[1], _ptr(gt), gt.shape[0],
                        gt.shape[1], topn, M_nn, _ptr(off), _ptr(ids))
    return off, ids[:off[-1]].copy()


def bipartite(base, gt, M_nn, rounds):
    base, gt = _f32(base), _u32(gt)
    n = base.shape[0]
    off = np.empty(n + 1, np.uint32)
    ids = np.empty(max(n * 2 * M_nn, 1), np.uint32)
    lib().orc_bipartite(_ptr(base), n, base.shape[1], _ptr(gt), gt.shape[0],
                        gt.shape[1], M_nn, rounds, _ptr(off), _ptr(ids))
    return off, ids[:off[-1]].copy()


def fusion_final(base, bi_off, bi_ids, link, proj, M_final):
    base = _f32(base)
    bi_off, bi_ids = _u32(bi_off), _u32(bi_ids)
    link, proj = _u32(link), _u32(proj)
    n = base.shape[0]
    width = int(np.max(np.diff(bi_off.astype(np.int64)), initial=0)) + M_final
    off = np.empty(n + 1, np.uint32)
    ids = np.empty(max(n * width, 1), np.uint32)
    lib().orc_fusion_final(_ptr(base), n, base.shape[1], _ptr(bi_off),
                           _ptr(bi_ids), _ptr(link), link.shape[1],
                           _ptr(proj), proj.shape[1], M_final, _ptr(off),
                           _ptr(ids))
    return off, ids[:off[-1]].copy()
