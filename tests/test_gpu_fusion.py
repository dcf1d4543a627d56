"""The host stages on the GPU (pg_run_fusion: MatchSup rounds, FusionNN,
FusionFinal) against the host library and the oracle, id for id, through
the C ABI.  The distances use the reference CPU `compare` order in IEEE FP32
on both sides, so the lists must be identical."""
import numpy as np
import pytest

from oracle import cpu_ref
from paragraph_amd import _native, datagen, host_stage

pytestmark = pytest.mark.gpu

TOMB = 0xFFFFFFFF


def _dev(a, torch):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).cuda()


def _gpu_fusion(hp, base, gt, proj, link, torch):
    n = base.shape[0]
    p = _native.make_params(n, base.shape[1], n, 32, 64)
    width = host_stage.lib().pg_host_final_width(host_stage.ctypes.byref(hp))
    off_d = torch.full((n + 1,), -7, dtype=torch.int32, device="cuda")
    ids_d = torch.full((max(n * width, 1),), -7, dtype=torch.int32,
                       device="cuda")
    with _native.Context(p) as ctx:
        for _ in range(2):  # the workspace is reused on the second call
            ctx.run_fusion(hp, _dev(base, torch), _dev(gt, torch),
                           _dev(proj, torch), _dev(link, torch), off_d, ids_d)
            ctx.sync()
    off = off_d.cpu().numpy().view(np.uint32)
    ids = ids_d.cpu().numpy().view(np.uint32)
    return off, ids[:off[-1]]


def _case(n, dim, nc, norm, seed):
    w = datagen.make_workload(n, dim, nc, normalize=norm, device="cpu")
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    rng = np.random.default_rng(seed)
    link = rng.integers(0, n, (n, 32)).astype(np.uint32)
    link[:, 24:] = TOMB
    link[::7, 3] = np.arange(0, n, 7)  # the node itself in its link row
    proj = rng.integers(0, n, (n, 16)).astype(np.uint32)
    proj[::5, 15] = n + 3              # an out-of-range id
    return w["base"], gt, link, proj


@pytest.mark.parametrize("n,dim,nc,norm", [(1500, 64, 16, False),
                                           (1200, 200, 64, False),
                                           (900, 37, 8, True),
                                           (1000, 512, 32, True)])
def test_gpu_fusion_equals_host_stages(n, dim, nc, norm, cuda):
    import torch
    base, gt, link, proj = _case(n, dim, nc, norm, n)
    hp = host_stage.make_params(n, dim, n, gt_k=gt.shape[1])
    print(f"[fusion] n={n} dim={dim}", flush=True)
    off, ids = _gpu_fusion(hp, base, gt, proj, link, torch)
    bo, bi = host_stage.bipartite(hp, base, gt)
    fo, fi = host_stage.fusion_final(hp, base, bo, bi, link, proj)
    np.testing.assert_array_equal(off, fo)
    np.testing.assert_array_equal(ids, fi)


def test_gpu_fusion_ragged_zipf_gt(cuda):
    """Zipf query components, per-query gt lengths in [16, 128] padded with
    -1, dim 512 normalised, 48-wide link rows (the skewed workload's shape)."""
    import torch
    n, dim = 1200, 512
    w = datagen.make_workload(n, dim, 64, normalize=True, zipf=1.1,
                              gt_len_range=(16, 128), device="cpu")
    base = w["base"]
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    rng = np.random.default_rng(9)
    link = rng.integers(0, n, (n, 48)).astype(np.uint32)
    proj = rng.integers(0, n, (n, 16)).astype(np.uint32)
    hp = host_stage.make_params(n, dim, n, gt_k=gt.shape[1])
    off, ids = _gpu_fusion(hp, base, gt, proj, link, torch)
    bo, bi = host_stage.bipartite(hp, base, gt)
    fo, fi = host_stage.fusion_final(hp, base, bo, bi, link, proj)
    np.testing.assert_array_equal(off, fo)
    np.testing.assert_array_equal(ids, fi)


def test_gpu_fusion_small_degrees_equal_oracle(cuda):
    """M_nn 60 (reverse cap 68, overflowing), 4 rounds, M_final 12 (the
    unsigned room: bipartite lists longer than M_final keep growing)."""
    import torch
    n, dim = 800, 48
    base, gt, link, proj = _case(n, dim, 8, True, 3)
    hp = host_stage.make_params(n, dim, n, gt_k=gt.shape[1], M_nn=60,
                                rounds=4, M_final=12)
    off, ids = _gpu_fusion(hp, base, gt, proj, link, torch)
    bo, bi = cpu_ref.bipartite(base, gt, 60, 4)
    fo, fi = cpu_ref.fusion_final(base, bo, bi, link, proj, 12)
    np.testing.assert_array_equal(off, fo)
    np.testing.assert_array_equal(ids, fi)
    assert np.diff(bo.astype(np.int64)).max() > 12


def test_gpu_fusion_rejects_out_of_range_shapes(cuda):
    import torch
    n, dim = 200, 16
    base, gt, link, proj = _case(n, dim, 4, False, 1)
    hp = host_stage.make_params(n, dim, n, gt_k=gt.shape[1], M_nn=80)
    with pytest.raises(_native.PgError, match="M_nn"):
        _gpu_fusion(hp, base, gt, proj, link, torch)
    hp = host_stage.make_params(n + 1, dim, n + 1, gt_k=gt.shape[1])
    with pytest.raises(_native.PgError, match="n_base"):
        _gpu_fusion(hp, base, gt, proj, link, torch)


def test_bipartite_beside_gpu_stage_equals_host_stages(cuda):
    """The bipartite stage on the context's second stream, beside the GPU
    stage on its main stream, then FusionFinal on the GPU stage's own
    outputs: the same lists as the host stages on those outputs, and as
    pg_run_fusion."""
    import torch
    n, dim = 1200, 200
    w = datagen.make_workload(n, dim, 64, device="cpu")
    base = w["base"]
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    hp = host_stage.make_params(n, dim, n, gt_k=gt.shape[1])
    p = _native.make_params(n, dim, n, 32, 64)
    width = host_stage.lib().pg_host_final_width(host_stage.ctypes.byref(hp))
    base_d, gt_d = _dev(base, torch), _dev(gt, torch)
    i32 = dict(dtype=torch.int32, device="cuda")
    proj_d = torch.empty((n, p.proj_deg), **i32)
    link_d = torch.empty((n, p.M), **i32)
    outs = []
    with _native.Context(p) as ctx:
        with pytest.raises(_native.PgError, match="bipartite"):
            ctx.run_fusion_final(hp, base_d, proj_d, link_d,
                                 torch.empty(n + 1, **i32),
                                 torch.empty(n * width, **i32))
        for overlapped in (True, False, True):
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
            outs.append((off, ids_d.cpu().numpy().view(np.uint32)[:off[-1]]))
    proj = proj_d.cpu().numpy().view(np.uint32)
    link = link_d.cpu().numpy().view(np.uint32)
    bo, bi = host_stage.bipartite(hp, base, gt)
    fo, fi = host_stage.fusion_final(hp, base, bo, bi, link, proj)
    for off, ids in outs:
        np.testing.assert_array_equal(off, fo)
        np.testing.assert_array_equal(ids, fi)


def test_gpu_fusion_hub_with_long_reverse_list(cuda):
    """Every query's gt row holds base 7 and most hold base 11, so their
    MatchSup forward lists (non-strict, filled to M_nn) keep them: reverse
    segments of thousands of sources, past the merge's 512-source LDS chunk,
    cut to the cap of the lowest query ids (rule C6)."""
    import torch
    n, dim, K = 3000, 64, 32
    rng = np.random.default_rng(21)
    base = rng.standard_normal((n, dim)).astype(np.float32)
    gt = np.empty((n, K), np.uint32)
    for q in range(n):
        gt[q] = rng.permutation(n)[:K]
    gt[:, 5] = 7
    gt[::2, 9] = 11
    link = rng.integers(0, n, (n, 32)).astype(np.uint32)
    proj = rng.integers(0, n, (n, 16)).astype(np.uint32)
    hp = host_stage.make_params(n, dim, n, gt_k=K, rounds=4)
    off, ids = _gpu_fusion(hp, base, gt, proj, link, torch)
    bo, bi = host_stage.bipartite(hp, base, gt)
    fo, fi = host_stage.fusion_final(hp, base, bo, bi, link, proj)
    np.testing.assert_array_equal(off, fo)
    np.testing.assert_array_equal(ids, fi)
    # the hub's supply list really is a capped merge (> M_nn candidates)
    so, si = host_stage.match_sup(hp, base, gt, 1)
    assert so[8] - so[7] > 0
