"""The large configurations (10M x 200; 5M x 512 with Zipf queries and
ragged gt) through size-independent checks: the stage-by-stage pipeline
through the C ABI, with the oracle re-running search, prune and merge on a
random node sample from the same inputs (FP32 rows must match id-for-id);
the whole-stage entry equal to the stage-by-stage composition; determinism;
row invariants.  The 1M x 200 and 1M x 512 builds also run at full size
with their tuned settings: FP32 at 1M x 512 (persistent locality-ordered
queue, merge in the same order), the BF16-screened path at both (beam
56 and 64, visited caches of 960 and 1024 slots; at 1M x 512 the split
screen and the candidate-major prune), and the FP8-screened path at
1M x 200 (beam 48, visited cache of 768 slots).  A screened search is not the
FP32 one, so there only the match, prune, reverse CSR and merge, which
decide every edge at FP32, are checked against the oracle on the sample.

Default sizes are scaled down (1M and 100K base rows, exact gt).  With
PG_SCALE_FULL=1 the configurations run at their full sizes; gt is then the
exact top-128 within a fixed random 200K-row subset of the base (an N x N
exact gt is out of reach for a test, and any well-formed gt exercises the
build the same way)."""
import os

import numpy as np
import pytest

from oracle import cpu_ref
from paragraph_amd import _native, datagen

pytestmark = pytest.mark.gpu
TOMB = 0xFFFFFFFF
FULL = os.environ.get("PG_SCALE_FULL") == "1"
# (config, scale, precision, parameter overrides, beam)
CASES = [("cfg4_10m_d200", 1.0 if FULL else 0.1, "fp32", {}, None),
         ("cfg5_5m_d512_zipf", 1.0 if FULL else 0.02, "fp32", {}, None),
         ("cfg2_1m_d200", 1.0, "bf16",
          dict(precision=_native.PG_BF16, tune_flags=0x221, table_size=960),
          56),
         ("cfg2_1m_d200", 1.0, "fp8",
          dict(precision=_native.PG_FP8E4M3, tune_flags=0x221,
               table_size=768), 48),
         ("cfg3_1m_d512", 1.0, "fp32", dict(tune_flags=0x624), None),
         ("cfg3_1m_d512", 1.0, "bf16",
          dict(precision=_native.PG_BF16, tune_flags=0x7A1, table_size=1024),
          64)]


def _u(t):
    return t.cpu().numpy().view(np.uint32)


def _sorted_segments(off, srcs):
    """CSR sources sorted within each segment (the GPU fills segments in
    atomic order; the merge sorts them)."""
    seg = np.repeat(np.arange(off.shape[0] - 1), np.diff(off).astype(np.int64))
    return srcs[np.lexsort((srcs, seg))]


@pytest.mark.parametrize("cfg,scale,prec,over,beam", CASES,
                         ids=[f"{c[0]}-{c[2]}" for c in CASES])
def test_scale_sampled_parity(cfg, scale, prec, over, beam, cuda):
    import torch
    w = datagen.make_config(cfg, scale=scale, device="cuda",
                            gt_subset=200_000 if FULL else None)
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
