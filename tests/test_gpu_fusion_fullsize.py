"""The whole index build at the 1M x 200 workload (inner product, M 32,
L 128, exact top-128 gt), BF16-screened GPU stage: the host stages on the
GPU (pg_run_fusion, and the bipartite stage beside the GPU stage on the
second stream) must give the same final lists as the host library on the
same GPU-stage outputs, id for id.  The two are independent
implementations of MatchSup, FusionNN and FusionFinal (the host one is
checked against the oracle in test_host_stage.py), so at full size their
agreement is the size-independent check of the GPU host stages."""
import numpy as np
import pytest

from paragraph_amd import _native, datagen, host_stage

pytestmark = pytest.mark.gpu
TOMB = 0xFFFFFFFF


def test_fullsize_gpu_host_stages_equal_host_library(cuda):
    import torch
    w = datagen.make_config("cfg2_1m_d200", device=cuda)
    base = np.ascontiguousarray(w["base"])
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    n, dim = base.shape
    print(f"[fullsize fusion] n={n} dim={dim} M={w['M']} L={w['L']}",
          flush=True)
    p = _native.make_params(n, dim, n, w["M"], w["L"],
                            precision=_native.PG_BF16)
    hp = host_stage.make_params(n, dim, n, gt_k=gt.shape[1])
    width = int(host_stage.lib().pg_host_final_width(
        host_stage.ctypes.byref(hp)))
    i32 = dict(dtype=torch.int32, device=cuda)
    base_d = torch.from_numpy(base).to(cuda)
    gt_d = torch.from_numpy(gt.view(np.int32)).to(cuda)
    proj_d = torch.empty((n, p.proj_deg), **i32)
    link_d = torch.empty((n, p.M), **i32)
    outs = []
    with _native.Context(p) as ctx:
        for overlapped in (False, True):
            off_d = torch.full((n + 1,), -7, **i32)
            ids_d = torch.full((n * width,), -7, **i32)
            if overlapped:
                ctx.run_bipartite_async(hp, base_d, gt_d)
                ctx.stage_device(base_d, gt_d, proj_d, link_d)
                ctx.run_fusion_final(hp, base_d, proj_d, link_d, off_d, ids_d)
            else:
                ctx.stage_device(base_d, gt_d, proj_d, link_d)
                ctx.run_fusion(hp, base_d, gt_d, proj_d, link_d, off_d, ids_d)
            ctx.wait()
            off = off_d.cpu().numpy().view(np.uint32)
            outs.append((off.copy(),
                         ids_d[:int(off[-1])].cpu().numpy().view(np.uint32)))
            print(f"[fullsize fusion] overlapped={overlapped} "
                  f"entries={int(off[-1])}", flush=True)
    proj = proj_d.cpu().numpy().view(np.uint32)
    link = link_d.cpu().numpy().view(np.uint32)
    bo, bi = host_stage.bipartite(hp, base, gt)
    fo, fi = host_stage.fusion_final(hp, base, bo, bi, link, proj)
    for off, ids in outs:
        np.testing.assert_array_equal(off, fo)
        np.testing.assert_array_equal(ids, fi)
    # index invariants: CSR offsets non-decreasing, rows within the final
    # width, every id a base row, every node with at least one neighbour
    lens = np.diff(fo.astype(np.int64))
    assert fo[0] == 0 and (lens >= 0).all() and lens.max() <= width
    assert (fi < n).all()
    assert (lens > 0).mean() > 0.999
