"""GPU parity: every stage of the HIP path against the CPU oracle, bit-exact
on neighbour ids, at sizes the oracle finishes in seconds.

Calls go through the C ABI (libparagraph_hip.so via ctypes); inputs are
seeded synthetic workloads of the benchmark shapes, scaled down in N.
"""
import numpy as np
import pytest

from oracle import cpu_ref
from paragraph_amd import _native, datagen

pytestmark = pytest.mark.gpu

CASES = {
    # name: (n_base, dim, n_comp, M, L, normalize, zipf, gt_len_range)
    "d200_m32_l100": (2000, 200, 64, 32, 100, False, None, None),
    "d512_m64_l160_norm": (1500, 512, 256, 64, 160, True, None, None),
    "d200_m48_l128_zipf_ragged": (1500, 200, 1024, 48, 128, True, 1.1,
                                  (16, 128)),
    "d64_m32_l64": (1200, 64, 16, 32, 64, False, None, None),
}

_cache = {}


def workload(name):
    if name not in _cache:
        n, dim, nc, M, L, norm, zipf, gl = CASES[name]
        w = datagen.make_workload(n, dim, nc, normalize=norm, zipf=zipf,
                                  gt_len_range=gl, device="cpu")
        w["gt_u32"] = np.ascontiguousarray(w["gt"].view(np.uint32))
        w["params"] = _native.make_params(n, dim, n, M, L)
        _cache[name] = w
    return _cache[name]


def to_dev(a, torch):
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def from_dev(t):
    a = t.cpu().numpy()
    return a.view(np.uint32) if a.dtype == np.int32 else a


def orc_params(p):
    d = p.as_dict()
    return {k: d[k] for k in (
        "n_base", "dim", "n_query", "gt_k", "proj_deg", "proj_pr", "proj_rev",
        "M", "link_rev", "Km", "Kp", "Kd", "topk", "reset_iter", "rounds",
        "entry_seed", "max_iters")}


@pytest.mark.parametrize("name", list(CASES))
def test_stages_match_oracle(name, cuda):
    import torch
    w = workload(name)
    p = w["params"]
    n, dim = w["base"].shape
    print(f"[parity] stages {name} n={n} dim={dim}", flush=True)
    base_d = to_dev(w["base"], torch)
    gt_d = to_dev(w["gt_u32"], torch)
    i32 = dict(dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        # K0 match
        winner_d = torch.empty(n, **i32)
        ctx.run_match(gt_d, winner_d)
        winner = cpu_ref.match_winner(w["gt_u32"], n)
        ctx.sync()
        np.testing.assert_array_equal(from_dev(winner_d), winner)
        # K2 on top-1 rows
        pr_d = torch.full((n, p.proj_pr), -1, **i32)
        pc_d = torch.zeros(n, **i32)
        ctx.run_prune_top1(base_d, gt_d, winner_d, p.proj_pr, pr_d, pc_d)
        rows = cpu_ref.top1_rows(w["gt_u32"], winner)
        pr, pc = cpu_ref.prune(w["base"], rows, p.proj_pr)
        ctx.sync()
        np.testing.assert_array_equal(from_dev(pc_d), pc)
        mask = np.arange(p.proj_pr)[None, :] < pc[:, None]
        np.testing.assert_array_equal(from_dev(pr_d)[mask], pr[mask])
        # reverse CSR: same offsets, same multiset per segment
        off_d = torch.zeros(n + 1, **i32)
        srcs_d = torch.zeros(max(1, int(pc.sum())), **i32)
        ctx.run_reverse(pr_d, pc_d, off_d, srcs_d)
        off, srcs = cpu_ref.reverse_csr(pr, pc)
        ctx.sync()
        off_g = from_dev(off_d)
        np.testing.assert_array_equal(off_g, off)
        srcs_g = from_dev(srcs_d)[: off[-1]]
        for c in range(n):
            a, b = off[c], off[c + 1]
            if b - a > 1:
                srcs_g[a:b] = np.sort(srcs_g[a:b])
        np.testing.assert_array_equal(srcs_g, srcs)
        # K3 projection (non-strict)
        proj_d = torch.empty((n, p.proj_deg), **i32)
        ctx.run_merge(base_d, off_d, srcs_d, p.proj_rev, pr_d, pc_d,
                      p.proj_deg, False, proj_d)
        proj = cpu_ref.merge(w["base"], off, srcs, p.proj_rev, pr, pc,
                             p.proj_deg, False)
        ctx.sync()
        np.testing.assert_array_equal(from_dev(proj_d), proj)
        # K4 search over the projection graph
        res_d = torch.empty((n, p.topk), **i32)
        ctx.run_search(base_d, base_d, proj_d, res_d)
        res = cpu_ref.search(w["base"], w["base"], proj, p.Km, p.Kp, p.Kd,
                             p.topk)
        ctx.sync()
        res_g = from_dev(res_d)
        bad = np.nonzero((res_g != res).any(axis=1))[0]
        assert bad.size == 0, f"search rows differ: {bad[:10]} of {n}"
        # K2 on search rows, reverse, K3 strict
        pr2_d = torch.full((n, p.M), -1, **i32)
        pc2_d = torch.zeros(n, **i32)
        ctx.run_prune_rows(base_d, res_d, p.M, pr2_d, pc2_d)
        pr2, pc2 = cpu_ref.prune(w["base"], res, p.M)
        ctx.sync()
        np.testing.assert_array_equal(from_dev(pc2_d), pc2)
        mask = np.arange(p.M)[None, :] < pc2[:, None]
        np.testing.assert_array_equal(from_dev(pr2_d)[mask], pr2[mask])
        off2, srcs2 = cpu_ref.reverse_csr(pr2, pc2)
        srcs2_d = torch.zeros(max(1, int(pc2.sum())), **i32)
        ctx.run_reverse(pr2_d, pc2_d, off_d, srcs2_d)
        link_d = torch.empty((n, p.M), **i32)
        ctx.run_merge(base_d, off_d, srcs2_d, p.link_rev, pr2_d, pc2_d, p.M,
                      True, link_d)
        link = cpu_ref.merge(w["base"], off2, srcs2, p.link_rev, pr2, pc2,
                             p.M, True)
        ctx.sync()
        np.testing.assert_array_equal(from_dev(link_d), link)


@pytest.mark.parametrize("name", list(CASES))
def test_whole_stage_matches_oracle(name, cuda):
    import torch
    w = workload(name)
    p = w["params"]
    n = p.n_base
    print(f"[parity] whole stage {name}", flush=True)
    st_o = cpu_ref.OrcStats()
    proj_o, link_o = cpu_ref.gpu_stage(orc_params(p), w["base"], w["gt_u32"],
                                       stats=st_o)
    base_d = to_dev(w["base"], torch)
    gt_d = to_dev(w["gt_u32"], torch)
    proj_d = torch.empty((n, p.proj_deg), dtype=torch.int32, device=cuda)
    link_d = torch.empty((n, p.M), dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.stage_device(base_d, gt_d, proj_d, link_d)
        ctx.wait()
        st = ctx.stats().as_dict()
    np.testing.assert_array_equal(from_dev(proj_d), proj_o)
    np.testing.assert_array_equal(from_dev(link_d), link_o)
    so = st_o.as_dict()
    for k in ("prune_dist", "prune_pairs", "merge_dist", "merge_pairs",
              "search_dist", "search_iters", "search_capped", "rev_overflow"):
        assert st[k] == so[k], (k, st[k], so[k])


# (configuration, tune_flags besides the library's default)
WHOLE_STAGE = [("cfg1_100k_d200", 0x624), ("cfg2_1m_d200", 0x224)]


@pytest.mark.parametrize("cfg,flags2", WHOLE_STAGE,
                         ids=[c for c, _ in WHOLE_STAGE])
def test_config_whole_stage_matches_oracle(cfg, flags2, cuda):
    """The 100K x 200 and 1M x 200 configurations at their stated sizes
    (64-component mixture, rotated + offset training queries, exact
    top-128 inner-product gt, M 32, L 100 / 128): the whole FP32 GPU
    stage, with the library's default knobs and with the locality-ordered
    queue (at 1M, the bench's tuned FP32 knobs), equals the oracle's
    projection and link rows id for id over every node, counters
    included."""
    import time

    import torch
    t0 = time.time()
    w = datagen.make_config(cfg, device="cuda")
    base = w["base"]
    n, dim = base.shape
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    print(f"[parity] {cfg} whole stage n={n} dim={dim}: workload after "
          f"{time.time() - t0:.0f} s", flush=True)
    p = _native.make_params(n, dim, n, w["M"], w["L"])
    st_o = cpu_ref.OrcStats()
    proj_o, link_o = cpu_ref.gpu_stage(orc_params(p), base, gt, stats=st_o)
    print(f"[parity] {cfg}: oracle done after {time.time() - t0:.0f} s",
          flush=True)
    base_d = to_dev(base, torch)
    gt_d = to_dev(gt, torch)
    for flags in (p.tune_flags, flags2):
        pp = _native.PgParams.from_buffer_copy(p)
        pp.tune_flags = flags
        proj_d = torch.empty((n, pp.proj_deg), dtype=torch.int32, device=cuda)
        link_d = torch.empty((n, pp.M), dtype=torch.int32, device=cuda)
        with _native.Context(pp) as ctx:
            ctx.stage_device(base_d, gt_d, proj_d, link_d)
            ctx.wait()
            st = ctx.stats().as_dict()
        np.testing.assert_array_equal(from_dev(proj_d), proj_o)
        np.testing.assert_array_equal(from_dev(link_d), link_o)
        for k, v in st_o.as_dict().items():
            assert st[k] == v, (flags, k, st[k], v)


def test_host_entry_point_matches_device_entry_point(cuda):
    import torch
    w = workload("d64_m32_l64")
    p = w["params"]
    n = p.n_base
    proj_h = np.empty((n, p.proj_deg), np.uint32)
    link_h = np.empty((n, p.M), np.uint32)
    with _native.Context(p) as ctx:
        ctx.stage_host(w["base"], w["gt_u32"], proj_h, link_h)
        ctx.wait()
        proj_d = torch.empty((n, p.proj_deg), dtype=torch.int32, device=cuda)
        link_d = torch.empty((n, p.M), dtype=torch.int32, device=cuda)
        ctx.stage_device(to_dev(w["base"], torch), to_dev(w["gt_u32"], torch),
                         proj_d, link_d)
        ctx.wait()
    np.testing.assert_array_equal(proj_h, from_dev(proj_d))
    np.testing.assert_array_equal(link_h, from_dev(link_d))


@pytest.mark.parametrize("prec,flags", [
    (_native.PG_FP32, 0x20), (_native.PG_BF16, 0x221),
    (_native.PG_FP8E4M3, 0x20), (_native.PG_BF16, 0x60)],
    ids=["fp32", "bf16-order", "fp8", "bf16-graph"])
@pytest.mark.parametrize("pinned", [True, False], ids=["pinned", "pageable"])
def test_streamed_host_entry_matches_device_entry(prec, flags, pinned, cuda):
    """The host entry point streams its inputs in several chunks (gt chunks
    matched and base chunks screen-converted as they land) and its outputs
    out: the rows equal the device entry's, on two builds in a row from the
    same context (the second must wait for the first), from pinned
    (pg_host_alloc) and pageable buffers, and in graph mode."""
    import torch
    w = workload("d200_m32_l100")
    p = _native.PgParams.from_buffer_copy(w["params"])
    p.precision = prec
    p.tune_flags = flags
    n = p.n_base
    print(f"streamed host entry prec={prec} flags={flags:#x} "
          f"pinned={pinned}", flush=True)
    keep = []

    def buf(shape, dtype, src=None):
        if pinned:
            pa = _native.PinnedArray(shape, dtype)
            keep.append(pa)
            a = pa.a
        else:
            a = np.empty(shape, dtype)
        if src is not None:
            a[:] = src
        return a

    base_h = buf(w["base"].shape, np.float32, w["base"])
    gt_h = buf(w["gt_u32"].shape, np.uint32, w["gt_u32"])
    with _native.Context(p) as ctx:
        proj_d = torch.empty((n, p.proj_deg), dtype=torch.int32, device=cuda)
        link_d = torch.empty((n, p.M), dtype=torch.int32, device=cuda)
        ctx.stage_device(to_dev(w["base"], torch), to_dev(w["gt_u32"], torch),
                         proj_d, link_d)
        ctx.wait()
        for rep in range(2):
            proj_h = buf((n, p.proj_deg), np.uint32)
            link_h = buf((n, p.M), np.uint32)
            proj_h[:] = 7
            link_h[:] = 7
            ctx.stage_host(base_h, gt_h, proj_h, link_h)
            if rep == 0:
                # a second build enqueued behind the first one
                proj_h2 = buf((n, p.proj_deg), np.uint32)
                link_h2 = buf((n, p.M), np.uint32)
                ctx.stage_host(base_h, gt_h, proj_h2, link_h2)
            ctx.wait()
            np.testing.assert_array_equal(proj_h, from_dev(proj_d))
            np.testing.assert_array_equal(link_h, from_dev(link_d))
            if rep == 0:
                np.testing.assert_array_equal(proj_h2, from_dev(proj_d))
                np.testing.assert_array_equal(link_h2, from_dev(link_d))
    del keep


def test_table_size_changes_nothing(cuda):
    """The visited table is an exact set: the reference's 4096 slots and the
    auto-sized table give the same search rows."""
    import torch
    w = workload("d200_m32_l100")
    n = w["params"].n_base
    base_d = to_dev(w["base"], torch)
    rng = np.random.default_rng(3)
    graph = rng.integers(0, n, (n, 16)).astype(np.uint32)
    graph_d = to_dev(graph, torch)
    outs = []
    for ts in (0, 4096, 1789):
        p = _native.make_params(n, w["base"].shape[1], n, 32, 100,
                                table_size=ts)
        out_d = torch.empty((n, p.topk), dtype=torch.int32, device=cuda)
        with _native.Context(p) as ctx:
            ctx.run_search(base_d, base_d, graph_d, out_d)
            ctx.sync()
        outs.append(from_dev(out_d))
    p = _native.make_params(n, w["base"].shape[1], n, 32, 100)
    ref = cpu_ref.search(w["base"], w["base"], graph, p.Km, p.Kp, p.Kd,
                         p.topk)
    for o in outs:
        np.testing.assert_array_equal(o, ref)


def test_build_is_deterministic(cuda):
    import torch
    w = workload("d200_m32_l100")
    p = w["params"]
    n = p.n_base
    outs = []
    with _native.Context(p) as ctx:
        for _ in range(2):
            proj_d = torch.empty((n, p.proj_deg), dtype=torch.int32,
                                 device=cuda)
            link_d = torch.empty((n, p.M), dtype=torch.int32, device=cuda)
            ctx.stage_device(to_dev(w["base"], torch),
                             to_dev(w["gt_u32"], torch), proj_d, link_d)
            ctx.wait()
            outs.append((from_dev(proj_d), from_dev(link_d)))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_persistent_search_matches_oracle(cuda):
    """The persistent work-queue grid (tune bit 0x20) computes the same rows
    as one workgroup per query."""
    import torch
    w = workload("d200_m32_l100")
    n = w["params"].n_base
    rng = np.random.default_rng(5)
    graph = rng.integers(0, n, (n, 16)).astype(np.uint32)
    base_d, graph_d = to_dev(w["base"], torch), to_dev(graph, torch)
    p = _native.make_params(n, w["base"].shape[1], n, 32, 100,
                            tune_flags=0x20)
    out_d = torch.empty((n, p.topk), dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_search(base_d, base_d, graph_d, out_d)
        ctx.sync()
    ref = cpu_ref.search(w["base"], w["base"], graph, p.Km, p.Kp, p.Kd,
                         p.topk)
    np.testing.assert_array_equal(from_dev(out_d), ref)


WAVE_CASES = [(prec, flags, L)
              for prec in (_native.PG_FP32, _native.PG_BF16,
                           _native.PG_FP8E4M3)
              for flags in (0x800, 0x820, 0xA20, 0x802)
              for L in (48, 100)]


@pytest.mark.parametrize("prec,flags,L", WAVE_CASES,
                         ids=[f"p{c[0]}-{c[1]:#x}-L{c[2]}" for c in WAVE_CASES])
def test_two_waves_per_query_same_rows(prec, flags, L, cuda):
    """Two waves per query (tune bit 11: the second wave shares the
    distance evaluations) give the same search rows and counters as one
    wave, with one workgroup per query, the persistent grid, the locality
    order and a wider gather; FP32 rows also equal the oracle's."""
    import torch
    w = workload("d200_m32_l100")
    n = w["params"].n_base
    rng = np.random.default_rng(7)
    graph = rng.integers(0, n, (n, 16)).astype(np.uint32)
    graph[rng.random((n, 16)) < 0.1] = 0xFFFFFFFF
    base_d, graph_d = to_dev(w["base"], torch), to_dev(graph, torch)
    print(f"two waves prec={prec} flags={flags:#x} L={L}", flush=True)
    outs, stats = [], []
    for f in (flags & ~0x800, flags):
        p = _native.make_params(n, w["base"].shape[1], n, 32, L,
                                precision=prec, tune_flags=f)
        out_d = torch.full((n, p.topk), -7, dtype=torch.int32, device=cuda)
        with _native.Context(p) as ctx:
            ctx.run_search(base_d, base_d, graph_d, out_d)
            ctx.sync()
            st = ctx.stats()
        outs.append(from_dev(out_d))
        stats.append((st.search_dist, st.search_iters))
    np.testing.assert_array_equal(outs[1], outs[0])
    assert stats[1] == stats[0]
    if prec == _native.PG_FP32:
        ref = cpu_ref.search(w["base"], w["base"], graph, p.Km, p.Kp, p.Kd,
                             p.topk)
        np.testing.assert_array_equal(outs[1], ref)


@pytest.mark.parametrize("n_query", [2000, 1003, 5, 1])
def test_locality_order_search_matches_oracle(n_query, cuda):
    """The persistent grid taking its queries in the locality order (tune
    bit 0x200: nearest-pivot groups, one queue segment per XCD) computes
    every query exactly as the oracle, for query counts that leave some of
    the 8 segments short or empty; rows are pre-filled, so a query the
    queue skipped would show."""
    import torch
    w = workload("d200_m32_l100")
    n = w["params"].n_base
    rng = np.random.default_rng(6)
    graph = rng.integers(0, n, (n, 16)).astype(np.uint32)
    queries = np.ascontiguousarray(w["base"][rng.permutation(n)[:n_query]])
    base_d, graph_d = to_dev(w["base"], torch), to_dev(graph, torch)
    q_d = to_dev(queries, torch)
    p = _native.make_params(n, w["base"].shape[1], n, 32, 100,
                            tune_flags=0x20 | 0x200)
    out_d = torch.full((n_query, p.topk), -7, dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        for _ in range(2):  # the second call reuses the order
            ctx.run_search(q_d, base_d, graph_d, out_d)
            ctx.sync()
    ref = cpu_ref.search(queries, w["base"], graph, p.Km, p.Kp, p.Kd, p.topk)
    np.testing.assert_array_equal(from_dev(out_d), ref)


@pytest.mark.parametrize("prec", [_native.PG_FP32, _native.PG_BF16])
def test_locality_order_build_equals_plain_build(prec, cuda):
    """A whole build with the locality order for the search and the prune
    sweep (bit 9) and the merge (bit 10), eager and captured in a hipGraph
    and replayed, gives the same rows as without it; the FP32 rows also
    equal the oracle."""
    import torch
    w = workload("d200_m32_l100")
    p0 = w["params"]
    n = p0.n_base
    base_d, gt_d = to_dev(w["base"], torch), to_dev(w["gt_u32"], torch)
    outs = []
    for tf in (0x20, 0x20 | 0x200, 0x20 | 0x200 | 0x40,
               0x20 | 0x200 | 0x400, 0x20 | 0x200 | 0x400 | 0x40):
        p = _native.make_params(n, p0.dim, n, p0.M, p0.Km, precision=prec,
                                tune_flags=tf)
        proj_d = torch.full((n, 16), -7, dtype=torch.int32, device=cuda)
        link_d = torch.full((n, p0.M), -7, dtype=torch.int32, device=cuda)
        with _native.Context(p) as ctx:
            for _ in range(2 if tf & 0x40 else 1):
                torch.cuda.synchronize()
                ctx.stage_device(base_d, gt_d, proj_d, link_d)
                ctx.wait()
        outs.append((from_dev(proj_d), from_dev(link_d)))
    for o in outs[1:]:
        np.testing.assert_array_equal(o[0], outs[0][0])
        np.testing.assert_array_equal(o[1], outs[0][1])
    if prec == _native.PG_FP32:
        proj_o, link_o = cpu_ref.gpu_stage(orc_params(p0), w["base"],
                                           w["gt_u32"])
        np.testing.assert_array_equal(outs[1][0], proj_o)
        np.testing.assert_array_equal(outs[1][1], link_o)


@pytest.mark.parametrize("Km,Kp,max_iters", [(64, 20, 100000), (100, 7, 1),
                                              (100, 7, 3)])
def test_search_wide_expansion_and_watchdog_match_oracle(Km, Kp, max_iters,
                                                         cuda):
    """Kp*Kd = 320 candidate slots (past the compile-time slices: the
    generic gather / visited loop and the full-sort merge), and the watchdog
    stopping after 1 and 3 iterations (the forced final sort)."""
    import torch
    w = workload("d200_m32_l100")
    n = w["params"].n_base
    rng = np.random.default_rng(11)
    graph = rng.integers(0, n, (n, 16)).astype(np.uint32)
    base_d, graph_d = to_dev(w["base"], torch), to_dev(graph, torch)
    p = _native.make_params(n, w["base"].shape[1], n, 32, Km, Kp=Kp,
                            max_iters=max_iters)
    out_d = torch.empty((n, p.topk), dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_search(base_d, base_d, graph_d, out_d)
        ctx.sync()
    ref = cpu_ref.search(w["base"], w["base"], graph, p.Km, p.Kp, p.Kd,
                         p.topk, max_iters=max_iters)
    np.testing.assert_array_equal(from_dev(out_d), ref)


@pytest.mark.parametrize("prec", [_native.PG_FP32, _native.PG_BF16])
def test_graph_replays_match_eager(prec, cuda):
    """hipGraph mode (tune bit 0x40): every replay, with eager work and
    synchronisations in between, equals the eager build."""
    import torch
    w = workload("d200_m32_l100")
    p0 = w["params"]
    n = p0.n_base
    base_d, gt_d = to_dev(w["base"], torch), to_dev(w["gt_u32"], torch)
    kw = dict(precision=prec)
    pe = _native.make_params(n, p0.dim, n, p0.M, p0.Km, **kw)
    pg = _native.make_params(n, p0.dim, n, p0.M, p0.Km,
                             tune_flags=0x40 | 0x20, **kw)
    proj_e = torch.empty((n, 16), dtype=torch.int32, device=cuda)
    link_e = torch.empty((n, p0.M), dtype=torch.int32, device=cuda)
    with _native.Context(pe) as ctx:
        ctx.stage_device(base_d, gt_d, proj_e, link_e)
        ctx.wait()
    ref = (from_dev(proj_e), from_dev(link_e))
    proj_g = torch.empty_like(proj_e)
    link_g = torch.empty_like(link_e)
    with _native.Context(pg) as ctx:
        for r in range(3):
            proj_g.fill_(-7)
            link_g.fill_(-7)
            torch.cuda.synchronize()
            ctx.stage_device(base_d, gt_d, proj_g, link_g)
            ctx.wait()
            np.testing.assert_array_equal(from_dev(proj_g), ref[0])
            np.testing.assert_array_equal(from_dev(link_g), ref[1])
            _ = (base_d * 2.0).sum().item()  # eager work between replays


@pytest.mark.parametrize("order", ["ascending", "descending", "shuffled",
                                   "gpu_reverse"])
@pytest.mark.parametrize("prec", [_native.PG_FP32, _native.PG_BF16])
def test_merge_hub_segments_match_oracle(order, prec, cuda):
    """Reverse segments past the 256-id chunk (four hubs with about 420 to
    6 000 sources; the other nodes have a few): the merge keeps the cap smallest
    sources whatever the order inside a segment (ascending: nothing
    staged; descending: every id staged; shuffled; the GPU's own reverse
    fill order), strict and non-strict, equal to the oracle's merge of the
    ascending CSR over every node."""
    import torch
    rng = np.random.default_rng(21)
    n, dim, mp = 6000, 16, 12
    base = rng.standard_normal((n, dim)).astype(np.float32)
    hubs = np.array([5, 77, 4001, 2999], np.uint32)
    pr = np.full((n, mp), 0xFFFFFFFF, np.uint32)
    pc = np.zeros(n, np.uint32)
    for b in range(n):
        row = []
        for h, pin in zip(hubs, (1.0, 0.5, 0.14, 0.07)):
            if h != b and rng.random() < pin:
                row.append(int(h))
        k = int(rng.integers(0, mp - len(row) + 1))
        others = rng.choice(n, size=k + 4, replace=False)
        for o in others:
            if o != b and o not in row and len(row) < mp:
                row.append(int(o))
        pr[b, :len(row)] = row
        pc[b] = len(row)
    off, srcs = cpu_ref.reverse_csr(pr, pc)
    seg = np.diff(off.astype(np.int64))
    assert seg.max() > 20 * 256 and (seg[hubs] > 256).all()
    if order in ("descending", "shuffled"):
        s2 = srcs.copy()
        for c in range(n):
            a, e = int(off[c]), int(off[c + 1])
            s2[a:e] = (s2[a:e][::-1] if order == "descending"
                       else rng.permutation(s2[a:e]))
        srcs_gpu = s2
    else:
        srcs_gpu = srcs
    # the context's own M / L are unused by run_merge (cap, deg and the row
    # width are its arguments)
    p = _native.make_params(n, dim, n, 32, 32, precision=prec)
    base_d = to_dev(base, torch)
    pr_d, pc_d = to_dev(pr, torch), to_dev(pc, torch)
    off_d = to_dev(off, torch)
    srcs_d = to_dev(np.ascontiguousarray(srcs_gpu), torch)
    with _native.Context(p) as ctx:
        if order == "gpu_reverse":
            srcs_d = torch.zeros(int(pc.sum()), dtype=torch.int32,
                                 device=cuda)
            ctx.run_reverse(pr_d, pc_d, off_d, srcs_d)
            ctx.sync()
            assert np.array_equal(from_dev(off_d), off)
        for cap, deg, strict in ((100, 24, True), (64, 40, False),
                                 (1, 8, True)):
            out_d = torch.empty((n, deg), dtype=torch.int32, device=cuda)
            ctx.run_merge(base_d, off_d, srcs_d, cap, pr_d, pc_d, deg, strict,
                          out_d)
            ctx.sync()
            exp = cpu_ref.merge(base, off, srcs, cap, pr, pc, deg, strict)
            np.testing.assert_array_equal(from_dev(out_d), exp,
                                          err_msg=f"cap {cap}")
