"""The allocation-free search entry point (pg_search_noalloc) against the
context's pg_run_search through the C ABI: identical rows on two calls in a
row, the second with a smaller batch, for one workgroup per query, the
persistent queue, and the persistent queue in the locality order; and its
host-side rejections (screened precision, a batch beyond its scratch, a
graph narrower than Kd), which return before anything is launched."""
import ctypes

import pytest

from paragraph_amd import _native, datagen

pytestmark = pytest.mark.gpu

N, DIM, M, L = 4000, 48, 32, 64
_w = {}


def _built(cuda):
    """One FP32 build (projection and link rows) of a small workload."""
    import torch
    if not _w:
        w = datagen.make_workload(N, DIM, 16, n_eval=3000, device="cpu")
        p = _native.make_params(N, DIM, N, M, L)
        base_d = torch.from_numpy(w["base"]).to(cuda)
        gt_d = torch.from_numpy(w["gt"]).to(cuda)
        proj_d = torch.empty((N, p.proj_deg), dtype=torch.int32, device=cuda)
        link_d = torch.empty((N, M), dtype=torch.int32, device=cuda)
        with _native.Context(p) as ctx:
            ctx.stage_device(base_d, gt_d, proj_d, link_d)
            ctx.wait()
        _w.update(base=base_d, proj=proj_d, link=link_d,
                  queries=torch.from_numpy(w["eval_queries"]).to(cuda))
    return _w


@pytest.mark.parametrize("flags", [0x0, 0x20, 0x220],
                         ids=["grid", "persistent", "persistent-ordered"])
def test_noalloc_search_equals_run_search(flags, cuda):
    import torch
    w = _built(cuda)
    p = _native.make_params(N, DIM, N, M, L, tune_flags=flags)
    with _native.Context(p) as ctx:
        # two calls in a row, the second with a smaller batch and the
        # projection graph (degree 16), so stale scratch would show
        for q, g in ((w["queries"], w["link"]), (w["queries"][:700], w["proj"]),
                     (w["queries"][:2500], w["link"])):
            print(f"noalloc flags={flags:#x} n_query={q.shape[0]} "
                  f"deg={g.shape[1]}", flush=True)
            ref = torch.full((q.shape[0], p.topk), -7, dtype=torch.int32,
                             device=cuda)
            ctx.run_search(q, w["base"], g, ref)
            ctx.sync()
            out = torch.full_like(ref, -9)
            _native.search_noalloc(p, q, w["base"], g, out)
            torch.cuda.synchronize()
            assert torch.equal(out, ref)
            assert int((ref == -7).sum()) == 0


def test_noalloc_search_rejections(cuda):
    import torch
    w = _built(cuda)
    L_ = _native.lib()
    out = torch.empty((10, 2 * L), dtype=torch.int32, device=cuda)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())

    def call(p, n_query, graph_deg):
        return L_.pg_search_noalloc(ctypes.byref(p), ptr(w["queries"]),
                                    n_query, ptr(w["base"]), ptr(w["link"]),
                                    graph_deg, ptr(out), None)

    pb = _native.make_params(N, DIM, N, M, L, precision=_native.PG_BF16)
    assert call(pb, 10, M) == -1
    assert b"FP32" in L_.pg_last_error(None)
    po = _native.make_params(N, DIM, N, M, L, tune_flags=0x220)
    assert call(po, (1 << 20) + 1, M) == -1
    assert b"exceeds" in L_.pg_last_error(None)
    assert call(po, 10, 8) == -1
    assert b"Kd" in L_.pg_last_error(None)
