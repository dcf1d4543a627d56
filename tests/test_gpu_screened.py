"""GPU tests of the BF16-screened path.

The screened search is not bit-exact with anything (its list order uses BF16
MFMA distances); it is checked against what the arithmetic bounds: on a
complete graph it must return the exact top-Km up to pairs whose FP64
distances differ by less than the BF16 screening error, and a whole screened
build must give well-formed rows whose neighbour sets largely agree with the
FP32 build.  The prune / merge steps of the screened build re-rank in FP32
and are the same kernels as the FP32 path (checked id-for-id there).
"""
import numpy as np
import pytest

from oracle import cpu_ref
from paragraph_amd import _native, datagen

pytestmark = pytest.mark.gpu

TOMB = 0xFFFFFFFF


def _dev(a, torch):
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def _host(t):
    a = t.cpu().numpy()
    return a.view(np.uint32) if a.dtype == np.int32 else a


def bf16_round(x):
    """Round float32 to bfloat16 (nearest even) and back, in numpy."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(
        np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("dim", [64, 200, 512])
def test_bf16_search_complete_graph_is_topk_within_bound(dim, cuda):
    import torch
    rng = np.random.default_rng(dim)
    n = 48
    base = rng.standard_normal((n, dim)).astype(np.float32)
    qs = rng.standard_normal((16, dim)).astype(np.float32)
    graph = np.full((n, n - 1), TOMB, np.uint32)
    for v in range(n):
        graph[v] = [u for u in range(n) if u != v]
    p = _native.make_params(n, dim, n, 2, 16, Kd=n - 1, Kp=1, topk=16,
                            proj_deg=n - 1, rounds=1,
                            precision=_native.PG_BF16)
    assert _native.check_params(p)[0] == 0, _native.check_params(p)
    out_d = torch.empty((qs.shape[0], p.topk), dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_search(_dev(qs, torch), _dev(base, torch), _dev(graph, torch),
                       out_d)
        ctx.sync()
    out = _host(out_d)
    # exact distances and the BF16 screening bound (Cauchy-Schwarz form)
    d64 = -(qs.astype(np.float64) @ base.astype(np.float64).T)
    norms = np.linalg.norm(base.astype(np.float64), axis=1)
    qn = np.linalg.norm(qs.astype(np.float64), axis=1)
    eps = (2 * 2.0 ** -8 + 2.0 ** -16 + 4 * dim * 2.0 ** -24) * \
        qn[:, None] * norms[None, :]
    # the search only ever visits all n nodes here (complete graph)
    for q in range(qs.shape[0]):
        got = out[q, :16]
        assert len(set(got.tolist())) == 16
        order = np.argsort(d64[q], kind="stable")
        kth = d64[q, order[15]]
        for i in got:   # everything returned is within bound of the top-16
            assert d64[q, i] <= kth + eps[q, i] + eps[q, order[15]]
        # screened distances reproduce the BF16 model of the product
        qb, bb = bf16_round(qs[q]).astype(np.float64), \
            bf16_round(base[got]).astype(np.float64)
        model = -(bb @ qb)
        assert np.all(np.diff(model) >= -1e-3 * (np.abs(model).max() + 1))


def test_bf16_build_rows_well_formed_and_close_to_fp32(cuda):
    import torch
    n, dim, M, L = 3000, 200, 32, 64
    w = datagen.make_workload(n, dim, 64, device="cpu", n_eval=300)
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    base_d, gt_d = _dev(w["base"], torch), _dev(gt, torch)
    links = {}
    for prec in (_native.PG_FP32, _native.PG_BF16):
        p = _native.make_params(n, dim, n, M, L, precision=prec)
        proj_d = torch.empty((n, p.proj_deg), dtype=torch.int32, device=cuda)
        link_d = torch.empty((n, M), dtype=torch.int32, device=cuda)
        with _native.Context(p) as ctx:
            ctx.stage_device(base_d, gt_d, proj_d, link_d)
            ctx.wait()
            st = ctx.stats()
        assert st.search_capped == 0
        links[prec] = (_host(proj_d), _host(link_d))
    # projection does not depend on the search: identical
    np.testing.assert_array_equal(links[_native.PG_FP32][0],
                                  links[_native.PG_BF16][0])
    a, b = links[_native.PG_FP32][1], links[_native.PG_BF16][1]
    for g in (a, b):
        valid = g != TOMB
        assert (g[valid] < n).all()
        assert (np.diff(valid.astype(np.int8), axis=1) <= 0).all()
    overlap = np.mean([len(set(x[x != TOMB]) & set(y[y != TOMB])) /
                       max(1, len(set(x[x != TOMB]))) for x, y in zip(a, b)])
    assert overlap > 0.8, overlap


def test_screened_prune_equals_fp32_prune(cuda):
    """The screened build re-ranks the search rows in FP32 with the same
    prune kernel: on identical rows it returns the oracle's lists."""
    import torch
    n, dim = 1500, 200
    w = datagen.make_workload(n, dim, 64, device="cpu")
    rng = np.random.default_rng(1)
    rows = rng.integers(0, n, (n, 256)).astype(np.uint32)
    p = _native.make_params(n, dim, n, 32, 128, precision=_native.PG_BF16)
    pr_d = torch.full((n, 32), -1, dtype=torch.int32, device=cuda)
    pc_d = torch.zeros(n, dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_prune_rows(_dev(w["base"], torch), _dev(rows, torch), 32,
                           pr_d, pc_d)
        ctx.sync()
    pr, pc = cpu_ref.prune(w["base"], rows, 32)
    np.testing.assert_array_equal(_host(pc_d), pc)
    mask = np.arange(32)[None, :] < pc[:, None]
    np.testing.assert_array_equal(_host(pr_d)[mask], pr[mask])


def _stage_inputs(n, dim, seed=0, integer=False):
    from paragraph_amd import datagen
    if integer:
        # small-integer rows with many duplicates: exact distance ties, so
        # the screen cannot decide and the exact fallback must run
        rng = np.random.default_rng(seed)
        protos = rng.integers(-2, 3, (n // 8, dim)).astype(np.float32)
        base = protos[rng.integers(0, protos.shape[0], n)]
        return np.ascontiguousarray(base)
    return datagen.base_vectors(n, dim, 64)


@pytest.mark.parametrize("scan", [False, True])
@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("integer", [False, True])
@pytest.mark.parametrize("dim,mp", [(200, 32), (512, 64)])
def test_screened_prune_and_merge_equal_oracle(scan, split, integer, dim, mp,
                                               cuda):
    """BF16 screen + exact fallback must reproduce the FP32 decisions exactly,
    including on data full of exact ties (fallbacks exercised), with the
    kept-major prune sweep and the candidate-major prune (scan).  A split
    screen context runs its prune sweep on exact chains."""
    import torch
    n = 1200
    base = _stage_inputs(n, dim, integer=integer)
    rng = np.random.default_rng(dim + int(integer))
    rows = rng.integers(0, n, (n, 256)).astype(np.uint32)
    rows[rng.random(rows.shape) < 0.05] = 0x7FFFFFFF
    p = _native.make_params(n, dim, n, mp, 128, precision=_native.PG_BF16,
                            tune_flags=0x20 | (0x80 if split else 0) |
                            (0x100 if scan else 0))
    base_d = _dev(base, torch)
    pr_d = torch.full((n, mp), -1, dtype=torch.int32, device=cuda)
    pc_d = torch.zeros(n, dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_prune_rows(base_d, _dev(rows, torch), mp, pr_d, pc_d)
        ctx.sync()
        pr, pc = cpu_ref.prune(base, rows, mp)
        np.testing.assert_array_equal(_host(pc_d), pc)
        mask = np.arange(mp)[None, :] < pc[:, None]
        np.testing.assert_array_equal(_host(pr_d)[mask], pr[mask])
        off, srcs = cpu_ref.reverse_csr(pr, pc)
        cap = 126 - mp
        out_d = torch.empty((n, mp), dtype=torch.int32, device=cuda)
        ctx.run_merge(base_d, _dev(off, torch),
                      _dev(srcs if len(srcs) else np.zeros(1, np.uint32),
                           torch), cap, pr_d, pc_d, mp, True, out_d)
        ctx.sync()
        st = ctx.stats()
    out = cpu_ref.merge(base, off, srcs, cap, pr, pc, mp, True)
    np.testing.assert_array_equal(_host(out_d), out)
    if integer and (scan or not split):
        assert st.screen_fallback > 0


def test_screened_whole_stage_matches_oracle_given_its_search(cuda):
    """In the screened build only the search differs from the FP32 path:
    running the oracle's prune / merge on the screened search rows gives the
    screened link graph exactly (a build of one link round)."""
    import torch
    from paragraph_amd import datagen
    n, dim, M, L = 1500, 200, 32, 64
    w = datagen.make_workload(n, dim, 64, device="cpu")
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    p = _native.make_params(n, dim, n, M, L, precision=_native.PG_BF16,
                            rounds=1)
    base_d, gt_d = _dev(w["base"], torch), _dev(gt, torch)
    proj_d = torch.empty((n, 16), dtype=torch.int32, device=cuda)
    link_d = torch.empty((n, M), dtype=torch.int32, device=cuda)
    res_d = torch.empty((n, p.topk), dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.stage_device(base_d, gt_d, proj_d, link_d)
        ctx.wait()
        ctx.run_search(base_d, base_d, proj_d, res_d)
        ctx.sync()
    res = _host(res_d)
    pr, pc = cpu_ref.prune(w["base"], res, M)
    off, srcs = cpu_ref.reverse_csr(pr, pc)
    link = cpu_ref.merge(w["base"], off, srcs, 126 - M, pr, pc, M, True)
    np.testing.assert_array_equal(_host(link_d), link)


@pytest.mark.parametrize("dim", [200, 512])
def test_fp8_search_complete_graph_recall(dim, cuda):
    """FP8 e4m3 screening (per-vector scales): on a complete graph the
    returned top-16 must mostly be the exact top-16, and always among the
    exact top-32."""
    import torch
    rng = np.random.default_rng(100 + dim)
    n = 64
    base = rng.standard_normal((n, dim)).astype(np.float32)
    qs = rng.standard_normal((32, dim)).astype(np.float32)
    graph = np.full((n, n - 1), TOMB, np.uint32)
    for v in range(n):
        graph[v] = [u for u in range(n) if u != v]
    p = _native.make_params(n, dim, n, 2, 16, Kd=n - 1, Kp=1, topk=16,
                            proj_deg=n - 1, rounds=1,
                            precision=_native.PG_FP8E4M3)
    out_d = torch.empty((qs.shape[0], p.topk), dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_search(_dev(qs, torch), _dev(base, torch), _dev(graph, torch),
                       out_d)
        ctx.sync()
    out = _host(out_d)
    d64 = -(qs.astype(np.float64) @ base.astype(np.float64).T)
    hits = []
    for q in range(qs.shape[0]):
        order = np.argsort(d64[q], kind="stable")
        got = set(out[q, :16].tolist())
        assert len(got) == 16
        assert got <= set(order[:32].tolist())
        hits.append(len(got & set(order[:16].tolist())) / 16)
    assert np.mean(hits) >= 0.8, np.mean(hits)


def test_fp8_build_and_prune(cuda):
    import torch
    n, dim, M, L = 3000, 512, 64, 160
    w = datagen.make_workload(n, dim, 256, normalize=True, device="cpu")
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    base_d, gt_d = _dev(w["base"], torch), _dev(gt, torch)
    links = {}
    for prec in (_native.PG_FP32, _native.PG_FP8E4M3):
        p = _native.make_params(n, dim, n, M, L, precision=prec)
        proj_d = torch.empty((n, p.proj_deg), dtype=torch.int32, device=cuda)
        link_d = torch.empty((n, M), dtype=torch.int32, device=cuda)
        with _native.Context(p) as ctx:
            ctx.stage_device(base_d, gt_d, proj_d, link_d)
            ctx.wait()
            assert ctx.stats().search_capped == 0
        links[prec] = _host(link_d)
    a, b = links[_native.PG_FP32], links[_native.PG_FP8E4M3]
    valid = b != TOMB
    assert (b[valid] < n).all()
    overlap = np.mean([len(set(x[x != TOMB]) & set(y[y != TOMB])) /
                       max(1, len(set(x[x != TOMB]))) for x, y in zip(a, b)])
    assert overlap > 0.6, overlap
    # the FP8 build's prune is the BF16-screened exact prune
    rng = np.random.default_rng(2)
    rows = rng.integers(0, n, (n, 320)).astype(np.uint32)
    p = _native.make_params(n, dim, n, M, L, precision=_native.PG_FP8E4M3)
    pr_d = torch.full((n, M), -1, dtype=torch.int32, device=cuda)
    pc_d = torch.zeros(n, dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_prune_rows(base_d, _dev(rows, torch), M, pr_d, pc_d)
        ctx.sync()
    pr, pc = cpu_ref.prune(w["base"], rows, M)
    np.testing.assert_array_equal(_host(pc_d), pc)
    mask = np.arange(M)[None, :] < pc[:, None]
    np.testing.assert_array_equal(_host(pr_d)[mask], pr[mask])
