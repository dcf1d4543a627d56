"""ctypes binding of libparagraph_hip.so (the C ABI in include/paragraph.h).

The library is the only compute path: if it is missing, or no GPU is present
when a compute call is made, these functions raise instead of falling back to
any CPU implementation.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libparagraph_hip.so")

PG_FP32, PG_BF16, PG_FP8E4M3 = 0, 1, 2
TOMB = 0xFFFFFFFF

_U32 = ctypes.c_uint32
_P = ctypes.c_void_p


class PgParams(ctypes.Structure):
    _fields_ = [(n, _U32) for n in (
        "n_base", "dim", "n_query", "gt_k", "proj_deg", "proj_pr", "proj_rev",
        "M", "link_rev", "Km", "Kp", "Kd", "topk", "reset_iter", "table_size",
        "rounds", "entry_seed", "max_iters", "precision", "rerank",
        "tune_flags")] + [("reserved", _U32 * 7)]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_
                if n != "reserved"}


class PgStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "prune_dist", "prune_pairs", "merge_dist", "merge_pairs",
        "search_dist", "search_iters", "search_capped", "rev_overflow",
        "screen_fallback")] + [("reserved", ctypes.c_uint64 * 7),
                               ("stage_ms", ctypes.c_float * 16),
                               ("n_stages", _U32)]

    def as_dict(self):
        d = {n: int(getattr(self, n)) for n, _ in self._fields_[:9]}
        d["stage_ms"] = {stage_name(i): float(self.stage_ms[i])
                         for i in range(self.n_stages)}
        return d


EXPORTED = (
    "pg_params_init", "pg_params_check", "pg_create", "pg_destroy",
    "pg_gpu_stage_async", "pg_gpu_stage_device_async", "pg_wait", "pg_query",
    "pg_get_stats", "pg_reset_stats", "pg_get_stage_counters", "pg_last_error", "pg_stage_name",
    "pg_run_match", "pg_run_prune", "pg_run_reverse", "pg_run_merge",
    "pg_run_search", "pg_run_pair_distances", "pg_sync", "pg_run_fusion",
    "pg_run_bipartite_async", "pg_run_fusion_final", "pg_search_noalloc",
    "pg_host_alloc", "pg_host_free", "pg_exact_topk_ip",
)

_lib = None


class PgError(RuntimeError):
    pass


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PgError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m paragraph_amd.build` (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    L.pg_params_init.argtypes = [ctypes.POINTER(PgParams), _U32, _U32, _U32,
                                 _U32, _U32]
    L.pg_params_check.argtypes = [ctypes.POINTER(PgParams), ctypes.c_char_p,
                                  _U32]
    L.pg_create.restype = _P
    L.pg_create.argtypes = [ctypes.c_int, ctypes.POINTER(PgParams)]
    L.pg_destroy.argtypes = [_P]
    L.pg_destroy.restype = None
    L.pg_gpu_stage_async.argtypes = [_P, _P, _P, _P, _P]
    L.pg_gpu_stage_device_async.argtypes = [_P, _P, _P, _P, _P]
    for f in ("pg_wait", "pg_query", "pg_reset_stats", "pg_sync"):
        getattr(L, f).argtypes = [_P]
    L.pg_get_stats.argtypes = [_P, ctypes.POINTER(PgStats)]
    L.pg_get_stage_counters.argtypes = [_P, _U32,
                                        ctypes.POINTER(ctypes.c_uint64)]
    L.pg_last_error.restype = ctypes.c_char_p
    L.pg_last_error.argtypes = [_P]
    L.pg_stage_name.restype = ctypes.c_char_p
    L.pg_stage_name.argtypes = [_U32]
    L.pg_run_match.argtypes = [_P, _P, _U32, _U32, _P]
    L.pg_run_prune.argtypes = [_P, _P, _P, _U32, _P, _P, _U32, _U32, _P, _P]
    L.pg_run_reverse.argtypes = [_P, _P, _P, _U32, _P, _P]
    L.pg_run_merge.argtypes = [_P, _P, _P, _P, _U32, _P, _P, _U32, _U32,
                               ctypes.c_int, _P]
    L.pg_run_search.argtypes = [_P, _P, _U32, _P, _P, _U32, _P]
    L.pg_run_fusion.argtypes = [_P, _P, _P, _P, _P, _U32, _P, _U32, _P, _P]
    L.pg_run_bipartite_async.argtypes = [_P, _P, _P, _P]
    L.pg_run_fusion_final.argtypes = [_P, _P, _P, _P, _U32, _P, _U32, _P, _P]
    L.pg_run_pair_distances.argtypes = [_P, _P, _U32, _P, _P, _U32,
                                        ctypes.c_int, _P]
    L.pg_search_noalloc.argtypes = [ctypes.POINTER(PgParams), _P, _U32, _P,
                                    _P, _U32, _P, _P]
    L.pg_exact_topk_ip.argtypes = [_P, _U32, _P, _U32, _U32, _U32, _P,
                                   ctypes.POINTER(_U32), _P]
    L.pg_host_alloc.argtypes = [ctypes.c_uint64, ctypes.POINTER(_P)]
    L.pg_host_free.argtypes = [_P]


def stage_name(i):
    return lib().pg_stage_name(i).decode()


def make_params(n_base, dim, n_query, M, L, **overrides):
    """Project mapping of (N, dim, M, L) onto the build knobs, then overrides."""
    p = PgParams()
    rc = lib().pg_params_init(ctypes.byref(p), n_base, dim, n_query, M, L)
    if rc != 0:
        raise PgError(f"pg_params_init failed ({rc})")
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, int(v))
    return p


def search_noalloc(params, queries, base, graph, out):
    """pg_search_noalloc: the exact FP32 search without a context, on
    torch's current stream of the tensors' device (no device allocation;
    its scratch is a fixed array of the library)."""
    import torch
    stream = torch.cuda.current_stream(queries.device).cuda_stream
    rc = lib().pg_search_noalloc(
        ctypes.byref(params), _dptr(queries), queries.shape[0], _dptr(base),
        _dptr(graph), graph.shape[1], _dptr(out), ctypes.c_void_p(stream))
    if rc != 0:
        raise PgError(f"rc={rc}: {lib().pg_last_error(None).decode()}")


def exact_topk_ip(queries, base, k, out):
    """pg_exact_topk_ip: out[q] = the k base ids with the largest inner
    product (FP32 chain scores, ties to the lower id) for device tensors
    queries (n x dim f32), base (N x dim f32), out (n x k int32).  Returns
    the number of rows computed by the exact pass (crowded scores)."""
    import torch
    stream = torch.cuda.current_stream(queries.device).cuda_stream
    n_exact = _U32(0)
    rc = lib().pg_exact_topk_ip(_dptr(queries), queries.shape[0],
                                _dptr(base), base.shape[0], base.shape[1], k,
                                _dptr(out), ctypes.byref(n_exact),
                                ctypes.c_void_p(stream))
    if rc != 0:
        raise PgError(f"rc={rc}: {lib().pg_last_error(None).decode()}")
    return int(n_exact.value)


class PinnedArray:
    """A numpy array over pinned host memory from pg_host_alloc (freed with
    pg_host_free when this object goes away); .a is the array."""

    def __init__(self, shape, dtype):
        import numpy as np
        dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * dtype.itemsize
        ptr = _P()
        rc = lib().pg_host_alloc(max(n, 1), ctypes.byref(ptr))
        if rc != 0:
            raise PgError(f"rc={rc}: {lib().pg_last_error(None).decode()}")
        self._ptr = ptr.value
        buf = (ctypes.c_uint8 * max(n, 1)).from_address(self._ptr)
        self.a = np.frombuffer(buf, dtype=np.uint8, count=n).view(
            dtype).reshape(shape)

    def __del__(self):
        if getattr(self, "_ptr", None):
            self.a = None
            lib().pg_host_free(_P(self._ptr))
            self._ptr = None


def check_params(p):
    buf = ctypes.create_string_buffer(256)
    rc = lib().pg_params_check(ctypes.byref(p), buf, 256)
    return rc, buf.value.decode()


def _check(rc, ctx):
    if rc != 0:
        raise PgError(f"rc={rc}: {lib().pg_last_error(ctx).decode()}")


def _dptr(t):
    """Device pointer of a torch tensor.  The library works on its own
    stream and requires device inputs (and output buffers a caller filled)
    to be complete when a call is made, so the tensor's producing stream
    (torch's current stream on its device) is synchronized first."""
    if t is None:
        return None
    if not t.is_cuda:
        raise PgError("device tensor expected")
    if not t.is_contiguous():
        raise PgError("contiguous tensor expected")
    import torch
    torch.cuda.current_stream(t.device).synchronize()
    return ctypes.c_void_p(t.data_ptr())


class Context:
    """Owns a pg_ctx (device workspace, stream, events)."""

    def __init__(self, params, device=0):
        self.params = params
        L = lib()
        self._ctx = L.pg_create(device, ctypes.byref(params))
        if not self._ctx:
            raise PgError(L.pg_last_error(None).decode())

    def close(self):
        if getattr(self, "_ctx", None):
            lib().pg_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- whole stage
    def stage_device(self, base, gt, proj, link):
        _check(lib().pg_gpu_stage_device_async(
            self._ctx, _dptr(base), _dptr(gt), _dptr(proj), _dptr(link)),
            self._ctx)

    def stage_host(self, base, gt, proj, link):
        """numpy buffers (C-contiguous float32 / uint32)."""
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _check(lib().pg_gpu_stage_async(self._ctx, P(base), P(gt), P(proj),
                                        P(link)), self._ctx)

    def wait(self):
        _check(lib().pg_wait(self._ctx), self._ctx)

    def query(self):
        return lib().pg_query(self._ctx)

    def stats(self):
        s = PgStats()
        _check(lib().pg_get_stats(self._ctx, ctypes.byref(s)), self._ctx)
        return s

    def stage_counters(self, stage):
        buf = (ctypes.c_uint64 * 16)()
        _check(lib().pg_get_stage_counters(self._ctx, stage, buf), self._ctx)
        names = [n for n, _ in PgStats._fields_[:9]]
        return {n: int(buf[i]) for i, n in enumerate(names)}

    def reset_stats(self):
        _check(lib().pg_reset_stats(self._ctx), self._ctx)

    # ---- single stages (device tensors)
    def run_match(self, gt, winner):
        _check(lib().pg_run_match(self._ctx, _dptr(gt), gt.shape[0],
                                  gt.shape[1], _dptr(winner)), self._ctx)

    def run_prune_rows(self, base, rows, mp, pruned, pcnt):
        _check(lib().pg_run_prune(self._ctx, _dptr(base), _dptr(rows),
                                  rows.shape[1], None, None, 0, mp,
                                  _dptr(pruned), _dptr(pcnt)), self._ctx)

    def run_prune_top1(self, base, gt, winner, mp, pruned, pcnt):
        _check(lib().pg_run_prune(self._ctx, _dptr(base), None, 0, _dptr(gt),
                                  _dptr(winner), gt.shape[1], mp,
                                  _dptr(pruned), _dptr(pcnt)), self._ctx)

    def run_reverse(self, pruned, pcnt, off, srcs):
        _check(lib().pg_run_reverse(self._ctx, _dptr(pruned), _dptr(pcnt),
                                    pruned.shape[1], _dptr(off), _dptr(srcs)),
               self._ctx)

    def run_merge(self, base, off, srcs, cap, pruned, pcnt, deg, strict, out):
        _check(lib().pg_run_merge(self._ctx, _dptr(base), _dptr(off),
                                  _dptr(srcs), cap, _dptr(pruned), _dptr(pcnt),
                                  pruned.shape[1], deg, 1 if strict else 0,
                                  _dptr(out)), self._ctx)

    def run_search(self, queries, base, graph, out):
        _check(lib().pg_run_search(self._ctx, _dptr(queries), queries.shape[0],
                                   _dptr(base), _dptr(graph), graph.shape[1],
                                   _dptr(out)), self._ctx)

    def run_pair_distances(self, queries, base, ids, out, chain=False):
        """out[q, t] = distance of queries[q] to base[ids[q, t]]; chain=False:
        the search evaluator at this context's precision, chain=True: the
        FP32 prune chain."""
        _check(lib().pg_run_pair_distances(
            self._ctx, _dptr(queries), queries.shape[0], _dptr(base),
            _dptr(ids), ids.shape[1], 1 if chain else 0, _dptr(out)),
            self._ctx)

    def run_fusion(self, hp, base, gt, proj, link, off, ids):
        """The host stages on the GPU (pg_run_fusion): hp is a
        host_stage.HostParams; off (N + 1) and ids (N x final width) are
        device tensors."""
        _check(lib().pg_run_fusion(
            self._ctx, ctypes.byref(hp), _dptr(base), _dptr(gt), _dptr(proj),
            proj.shape[1], _dptr(link), link.shape[1], _dptr(off),
            _dptr(ids)), self._ctx)

    def run_bipartite_async(self, hp, base, gt):
        """MatchSup rounds + FusionNN on the context's second stream
        (pg_run_bipartite_async)."""
        _check(lib().pg_run_bipartite_async(self._ctx, ctypes.byref(hp),
                                            _dptr(base), _dptr(gt)),
               self._ctx)

    def run_fusion_final(self, hp, base, proj, link, off, ids):
        """FusionFinal after the bipartite stage and the main stream's work
        (pg_run_fusion_final)."""
        _check(lib().pg_run_fusion_final(
            self._ctx, ctypes.byref(hp), _dptr(base), _dptr(proj),
            proj.shape[1], _dptr(link), link.shape[1], _dptr(off),
            _dptr(ids)), self._ctx)

    def sync(self):
        _check(lib().pg_sync(self._ctx), self._ctx)
