"""Every benchmark configuration at its stated size, through
size-independent checks: the stage-by-stage pipeline through the C ABI,
with the oracle re-running the match (in full) and the prune, reverse CSR,
merge and (FP32) search on a random node sample from the same inputs (rows
must match id-for-id); the whole-stage entry equal to the stage-by-stage
composition; determinism; row invariants.

  10M x 200 (cfg4) and 5M x 512 Zipf with ragged gt (cfg5): FP32 and the
  BF16-screened build;
  1M x 200 (cfg2): the BF16 and FP8 value paths (tuned beams and knobs);
  1M x 512 (cfg3): FP32 (locality-ordered queue and merge), BF16 (split
  screen, candidate-major prune) and the FP8 e4m3 leg.

A screened search is not the FP32 one, so on the screened legs the oracle
checks the match, prune, reverse CSR and merge rows (every edge is decided
at FP32 there).  The 10M and 5M builds take their gt from 200K-row shards
(the exact top-128 within the shard a query is assigned to): an exact N x N
gt at those sizes is a bench input, and any well-formed gt exercises the
build the same way.  PG_SCALE_SMALL=1 runs the two large configurations
scaled to 1M and 100K rows with exact gt (a quick run)."""
import os

import numpy as np
import pytest

from oracle import cpu_ref
from paragraph_amd import _native, datagen

pytestmark = pytest.mark.gpu
TOMB = 0xFFFFFFFF
SMALL = os.environ.get("PG_SCALE_SMALL") == "1"
LARGE_SCALE = {"cfg4_10m_d200": 0.1 if SMALL else 1.0,
               "cfg5_5m_d512_zipf": 0.02 if SMALL else 1.0}
BF16 = _native.PG_BF16
FP8 = _native.PG_FP8E4M3
# (config, precision, parameter overrides, beam)
CASES = [("cfg4_10m_d200", "fp32", {}, None),
         ("cfg4_10m_d200", "bf16",
          dict(precision=BF16, tune_flags=0x221, table_size=768), 48),
         ("cfg5_5m_d512_zipf", "fp32", {}, None),
         ("cfg5_5m_d512_zipf", "bf16", dict(precision=BF16, tune_flags=0x221),
          None),
         ("cfg2_1m_d200", "bf16",
          dict(precision=BF16, tune_flags=0x221, table_size=960), 56),
         ("cfg2_1m_d200", "fp8",
          dict(precision=FP8, tune_flags=0x221, table_size=768), 48),
         ("cfg3_1m_d512", "fp32", dict(tune_flags=0x624), None),
         ("cfg3_1m_d512", "bf16",
          dict(precision=BF16, tune_flags=0x7A1, table_size=1024), 64),
         ("cfg3_1m_d512", "fp8",
          dict(precision=FP8, tune_flags=0x221, table_size=1024), 64)]

_workloads = {}


def _workload(cfg):
    """One generated workload per configuration, shared by its legs (the
    cases of a configuration are adjacent, so one is kept at a time)."""
    if cfg not in _workloads:
        _workloads.clear()
        scale = LARGE_SCALE.get(cfg, 1.0)
        large = cfg in LARGE_SCALE and not SMALL
        _workloads[cfg] = datagen.make_config(
            cfg, scale=scale, device="cuda",
            gt_subset=200_000 if large else None)
    return _workloads[cfg]


def _u(t):
    return t.cpu().numpy().view(np.uint32)


def _sorted_segments(off, srcs):
    """CSR sources sorted within each segment (the GPU fills segments in
    atomic order; the merge sorts them)."""
    seg = np.repeat(np.arange(off.shape[0] - 1), np.diff(off).astype(np.int64))
    return srcs[np.lexsort((srcs, seg))]


@pytest.mark.parametrize("cfg,prec,over,beam", CASES,
                         ids=[f"{c[0]}-{c[1]}" for c in CASES])
def test_scale_sampled_parity(cfg, prec, over, beam, cuda):
    import time

    import torch
    t0 = time.time()
    print(f"[scale] {cfg} {prec}: generating", flush=True)
    w = _workload(cfg)
    print(f"[scale] {cfg} {prec}: workload ready after {time.time() - t0:.0f}"
          " s", flush=True)
    base = w["base"]
    n, dim = base.shape
    M, L = w["M"], beam or w["L"]
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    print(f"{cfg} {prec} n={n} dim={dim} M={M} L={L}", flush=True)
    p = _native.make_params(n, dim, n, M, L, **over)
    i32 = dict(dtype=torch.int32, device=cuda)
    base_d = torch.from_numpy(base).to(cuda)
    gt_d = torch.from_numpy(gt.view(np.int32)).to(cuda)
    rng = np.random.default_rng(5)
    S = np.sort(rng.choice(n, size=min(300, n), replace=False)).astype(
        np.uint32)
    with _native.Context(p) as ctx:
        # stage by stage, checking a node sample against the oracle
        win_d = torch.empty(n, **i32)
        ctx.run_match(gt_d, win_d)
        pr_d = torch.full((n, p.proj_pr), -1, **i32)
        pc_d = torch.zeros(n, **i32)
        ctx.run_prune_top1(base_d, gt_d, win_d, p.proj_pr, pr_d, pc_d)
        ctx.sync()
        win = _u(win_d)
        assert np.array_equal(win, cpu_ref.match_winner(gt, n))
        rows = cpu_ref.top1_rows(gt, win)
        opr, opc = cpu_ref.prune_nodes(base, S, rows[S], p.proj_pr)
        pr, pc = _u(pr_d), _u(pc_d)
        assert np.array_equal(pc[S], opc)
        mask = np.arange(p.proj_pr)[None, :] < opc[:, None]
        assert np.array_equal(pr[S][mask], opr[mask])
        off_d = torch.zeros(n + 1, **i32)
        srcs_d = torch.zeros(max(1, int(pc.sum())), **i32)
        ctx.run_reverse(pr_d, pc_d, off_d, srcs_d)
        graph_d = torch.empty((n, p.proj_deg), **i32)
        ctx.run_merge(base_d, off_d, srcs_d, p.proj_rev, pr_d, pc_d,
                      p.proj_deg, False, graph_d)
        ctx.sync()
        off, srcs = _u(off_d), _u(srcs_d)[:int(pc.sum())]
        ooff, osrcs = cpu_ref.reverse_csr(pr, pc)
        assert np.array_equal(off, ooff)
        assert np.array_equal(_sorted_segments(off, srcs), osrcs)
        om = cpu_ref.merge_nodes(base, S, ooff, osrcs, p.proj_rev, pr, pc,
                                 p.proj_deg, False)
        assert np.array_equal(_u(graph_d)[S], om)
        print(f"[scale] {cfg} {prec}: projection checked after "
              f"{time.time() - t0:.0f} s", flush=True)
        for r in range(p.rounds):
            graph = _u(graph_d)
            res_d = torch.empty((n, p.topk), **i32)
            ctx.run_search(base_d, base_d, graph_d, res_d)
            ctx.sync()
            res = _u(res_d)
            if prec == "fp32":
                osr = cpu_ref.search(base[S], base, graph, p.Km, p.Kp, p.Kd,
                                     p.topk)
                assert np.array_equal(res[S], osr), f"search round {r + 1}"
            pr_d = torch.full((n, p.M), -1, **i32)
            pc_d = torch.zeros(n, **i32)
            ctx.run_prune_rows(base_d, res_d, p.M, pr_d, pc_d)
            ctx.sync()
            pr, pc = _u(pr_d), _u(pc_d)
            opr, opc = cpu_ref.prune_nodes(base, S, res[S], p.M)
            assert np.array_equal(pc[S], opc)
            mask = np.arange(p.M)[None, :] < opc[:, None]
            assert np.array_equal(pr[S][mask], opr[mask])
            srcs_d = torch.zeros(max(1, int(pc.sum())), **i32)
            ctx.run_reverse(pr_d, pc_d, off_d, srcs_d)
            link_d = torch.empty((n, p.M), **i32)
            ctx.run_merge(base_d, off_d, srcs_d, p.link_rev, pr_d, pc_d, p.M,
                          True, link_d)
            ctx.sync()
            off, srcs = _u(off_d), _u(srcs_d)[:int(pc.sum())]
            ooff, osrcs = cpu_ref.reverse_csr(pr, pc)
            assert np.array_equal(off, ooff)
            assert np.array_equal(_sorted_segments(off, srcs), osrcs)
            om = cpu_ref.merge_nodes(base, S, ooff, osrcs, p.link_rev, pr, pc,
                                     p.M, True)
            assert np.array_equal(_u(link_d)[S], om), f"merge round {r + 1}"
            graph_d = link_d
            print(f"[scale] {cfg} {prec}: round {r + 1} checked after "
                  f"{time.time() - t0:.0f} s", flush=True)
        # whole stage: equal to the composition, and deterministic
        proj_d = torch.empty((n, p.proj_deg), **i32)
        link2_d = torch.empty((n, p.M), **i32)
        ctx.stage_device(base_d, gt_d, proj_d, link2_d)
        ctx.wait()
        assert torch.equal(link2_d, graph_d)
        first = link2_d.clone()
        ctx.stage_device(base_d, gt_d, proj_d, link2_d)
        ctx.wait()
        assert torch.equal(link2_d, first)
        st = ctx.stats()
        assert st.search_capped == 0
    link = _u(first)
    valid = link != TOMB
    assert (link[valid] < n).all()
    assert not (link == np.arange(n, dtype=np.uint32)[:, None]).any()
    # rows are packed: no valid entry after a tomb
    assert (np.diff(valid.astype(np.int8), axis=1) <= 0).all()
    assert valid[:, 0].mean() > 0.99
