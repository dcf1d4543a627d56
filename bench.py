"""Benchmark of the GPU index-build stage on one MI355X.

One step = one whole GPU stage (match, projection prune/reverse/merge, two
link rounds of search/prune/reverse/merge) over the 1M x 200 workload:
1,000,000 base vectors from a 64-component Gaussian mixture, 1,000,000
rotated+offset training queries, exact top-128 gt, M = 32, L = 128, inner
product.  Inputs are resident in HBM when the timed region starts.

Prints one JSON line: throughput (base points indexed / s), the roofline
of the dominant kernel (algorithmic bytes per launch / its average duration
measured with HIP events on the build stream), recall@10 of the link graph on
10,000 held-out queries, and the CPU oracle timed on a bounded node sample
of the same workload on this machine's host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def profiled_traffic(kernel, workload, precision):
    """HBM bytes per launch recorded from rocprofv3 FETCH_SIZE/WRITE_SIZE
    runs (scaled by a 1 GiB copy) in profiles/*.md, or None."""
    import glob
    import re
    key = f"traffic_bytes_per_launch[{kernel},{workload},{precision}]"
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            if line.startswith(key):
                m = re.search(r"=\s*([0-9]+)", line)
                if m:
                    return int(m.group(1))
    return None


def stage_bytes(kind, c, p, n, D=None, mp=None, deg=None):
    """Algorithmic HBM bytes of one launch from its counters (FP32 rows)."""
    row = 4 * p.dim
    if kind == "search":
        return (row * (n + c["search_dist"]) + 4 * p.Kp * p.Kd *
                c["search_iters"] + 4 * p.topk * n)
    if kind == "prune":
        return row * (n + c["prune_dist"]) + 4 * D * n + 4 * mp * n
    if kind == "merge":
        return (row * (n + c["merge_dist"]) + 4 * c["merge_dist"] +
                4 * deg * n)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg2_1m_d200")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--precision", default="fp32", choices=["fp32"])
    ap.add_argument("--cpu-sample", type=int, default=4000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-recall", action="store_true")
    ap.add_argument("--n-eval", type=int, default=10000)
    ap.add_argument("--tune-flags", type=int, default=0,
                    help="perf-only knobs passed to pg_params.tune_flags")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 1 GiB device copy first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; "
                          "--gpus must be 1"}))
        sys.exit(2)

    import torch
    from paragraph_amd import _native, datagen

    dev = torch.device("cuda:0")
    cfgd = datagen.CONFIGS[args.config]
    t0 = time.time()
    w = datagen.make_config(args.config, scale=args.scale,
                            n_eval=0 if args.no_recall else args.n_eval,
                            device="cuda")
    n, dim = w["base"].shape
    M, L = w["M"], w["L"]
    log(f"workload {args.config} n={n} dim={dim} M={M} L={L} generated in "
        f"{time.time() - t0:.1f}s")
    gt_u32 = np.ascontiguousarray(w["gt"].view(np.uint32))
    p = _native.make_params(n, dim, n, M, L, tune_flags=args.tune_flags)
    base_d = torch.from_numpy(w["base"]).to(dev)
    gt_d = torch.from_numpy(w["gt"]).to(dev)
    proj_d = torch.empty((n, p.proj_deg), dtype=torch.int32, device=dev)
    link_d = torch.empty((n, M), dtype=torch.int32, device=dev)
    ctx = _native.Context(p)
    if args.calib_copy:
        # a plain 1 GiB device copy, the reference point for HBM counters
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for i in range(args.warmup):
        ctx.stage_device(base_d, gt_d, proj_d, link_d)
        ctx.wait()
        log(f"warmup {i} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.stage_device(base_d, gt_d, proj_d, link_d)
        ctx.wait()
        log(f"step {i} done")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    value = n / (ms_per_step / 1000.0)
    log(f"timed {args.steps} steps: {ms_per_step:.1f} ms/step, "
        f"{value:.0f} pts/s")

    st = ctx.stats()
    stage_ms = [float(st.stage_ms[i]) for i in range(st.n_stages)]
    names = [_native.stage_name(i) for i in range(st.n_stages)]
    counters = [ctx.stage_counters(i) for i in range(st.n_stages)]
    kinds = {}
    for i, nm in enumerate(names):
        kind = nm.split("_")[0]
        if kind == "match":
            continue
        D = p.gt_k if nm == "prune_proj" else p.topk
        mp = p.proj_pr if nm == "prune_proj" else p.M
        deg = p.proj_deg if nm == "merge_proj" else p.M
        b = stage_bytes(kind, counters[i], p, n, D=D, mp=mp, deg=deg)
        k = kinds.setdefault(kind, {"ms": 0.0, "bytes": 0, "launches": 0})
        k["ms"] += stage_ms[i]
        k["bytes"] += b
        k["launches"] += 1
    dom = max(kinds, key=lambda k: kinds[k]["ms"])
    kd = kinds[dom]
    avg_ms = kd["ms"] / kd["launches"]
    bytes_per_launch = kd["bytes"] / kd["launches"]
    achieved = bytes_per_launch / (avg_ms / 1000.0) / 1e9
    kname = {"search": "search_kernel", "prune": "prune_kernel",
             "merge": "merge_kernel", "reverse": "reverse_*"}.get(dom, dom)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(kname, args.config,
                                            args.precision),
                "kernel": kname,
                "avg_launch_ms": round(avg_ms, 3),
                "alg_bytes_per_launch": int(bytes_per_launch)}
    log("stage ms:", {nm: round(x, 2) for nm, x in zip(names, stage_ms)})

    extra = {"stage_ms": {nm: round(x, 3) for nm, x in zip(names, stage_ms)},
             "stage_counters": {nm: c for nm, c in zip(names, counters)},
             "counters": st.as_dict()}
    extra["counters"].pop("stage_ms", None)

    # ---- recall@10 of the link graph on held-out queries (GPU search)
    if not args.no_recall:
        eq = w["eval_queries"]
        eq_d = torch.from_numpy(eq).to(dev)
        out_d = torch.empty((eq.shape[0], p.topk), dtype=torch.int32,
                            device=dev)
        ctx.run_search(eq_d, base_d, link_d, out_d)
        ctx.sync()
        found = out_d[:, :10].cpu().numpy()
        # exact top-10 in float64
        B64 = base_d.double()
        truth = []
        for s in range(0, eq.shape[0], 1000):
            sc = eq_d[s:s + 1000].double() @ B64.T
            truth.append(torch.topk(sc, 10, dim=1).indices.cpu().numpy())
        truth = np.concatenate(truth)
        hits = sum(len(set(a) & set(b)) for a, b in zip(found, truth))
        extra["recall_at_10_link_graph"] = round(hits / truth.size, 4)
        del B64
        log(f"recall@10 (link graph, {eq.shape[0]} held-out queries): "
            f"{extra['recall_at_10_link_graph']}")

    # ---- CPU baseline: the oracle on a bounded node sample
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(ctx, p, w, gt_u32, base_d, gt_d, args.cpu_sample,
                           log)

    line = {
        "metric": "base points indexed/sec, 1 MI355X, FP32 vs BF16-MFMA at "
                  "recall@10 within 0.5 pt",
        "value": round(value, 1),
        "unit": "base points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 64-component Gaussian mixture; exact "
                "top-128 inner-product gt)",
        "config": {"workload": f"{args.config}: {n} base x {dim} f32, {n} "
                   f"training queries, gt k=128, M={M}, L={L}, inner "
                   "product, precision fp32",
                   "n_base": n, "dim": dim, "M": M, "L": L,
                   "precision": args.precision},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def cpu_baseline(ctx, p, w, gt_u32, base_d, gt_d, sample, log):
    """Oracle throughput on this host: per-node stages timed on `sample`
    random nodes and scaled to N; the global stages (match, reverse) run in
    full.  Inputs of each sampled stage are the GPU pipeline's intermediates
    (identical to the oracle's by the parity tests)."""
    import torch
    from oracle import cpu_ref
    from paragraph_amd import _native

    n, dim = w["base"].shape
    base = w["base"]
    dev = base_d.device
    i32 = dict(dtype=torch.int32, device=dev)
    u = lambda t: t.cpu().numpy().view(np.uint32)
    # GPU intermediates, stage by stage
    winner_d = torch.empty(n, **i32)
    ctx.run_match(gt_d, winner_d)
    pr1_d = torch.full((n, p.proj_pr), -1, **i32)
    pc1_d = torch.zeros(n, **i32)
    ctx.run_prune_top1(base_d, gt_d, winner_d, p.proj_pr, pr1_d, pc1_d)
    ctx.sync()
    pr1, pc1 = u(pr1_d), u(pc1_d)
    off_d = torch.zeros(n + 1, **i32)
    srcs_d = torch.zeros(max(1, int(pc1.sum())), **i32)
    ctx.run_reverse(pr1_d, pc1_d, off_d, srcs_d)
    proj_d = torch.empty((n, p.proj_deg), **i32)
    ctx.run_merge(base_d, off_d, srcs_d, p.proj_rev, pr1_d, pc1_d,
                  p.proj_deg, False, proj_d)
    ctx.sync()
    rounds = []
    graph_d = proj_d
    for r in range(p.rounds):
        res_d = torch.empty((n, p.topk), **i32)
        ctx.run_search(base_d, base_d, graph_d, res_d)
        prr_d = torch.full((n, p.M), -1, **i32)
        pcr_d = torch.zeros(n, **i32)
        ctx.run_prune_rows(base_d, res_d, p.M, prr_d, pcr_d)
        ctx.sync()
        srcs2_d = torch.zeros(max(1, int(u(pcr_d).sum())), **i32)
        ctx.run_reverse(prr_d, pcr_d, off_d, srcs2_d)
        link_d = torch.empty((n, p.M), **i32)
        ctx.run_merge(base_d, off_d, srcs2_d, p.link_rev, prr_d, pcr_d, p.M,
                      True, link_d)
        ctx.sync()
        rounds.append(dict(graph=u(graph_d), res=u(res_d), pr=u(prr_d),
                           pc=u(pcr_d)))
        graph_d = link_d

    threads = cpu_ref.max_threads()
    rng = np.random.default_rng(7)
    S = np.sort(rng.choice(n, size=min(sample, n), replace=False)).astype(
        np.uint32)
    scale = n / S.shape[0]
    t = {}

    def timed(name, fn, per_node):
        t0 = time.perf_counter()
        r = fn()
        dt = time.perf_counter() - t0
        t[name] = dt * (scale if per_node else 1.0)
        return r

    winner = timed("match", lambda: cpu_ref.match_winner(gt_u32, n), False)
    rows = cpu_ref.top1_rows(gt_u32, winner)
    timed("prune_proj", lambda: cpu_ref.prune_nodes(base, S, rows[S],
                                                    p.proj_pr), True)
    off, srcs = timed("reverse_proj", lambda: cpu_ref.reverse_csr(pr1, pc1),
                      False)
    timed("merge_proj", lambda: cpu_ref.merge_nodes(
        base, S, off, srcs, p.proj_rev, pr1, pc1, p.proj_deg, False), True)
    for r, rd in enumerate(rounds, 1):
        timed(f"search_round{r}", lambda: cpu_ref.search(
            base[S], base, rd["graph"], p.Km, p.Kp, p.Kd, p.topk), True)
        timed(f"prune_round{r}", lambda: cpu_ref.prune_nodes(
            base, S, rd["res"][S], p.M), True)
        off2, srcs2 = timed(f"reverse_round{r}",
                            lambda: cpu_ref.reverse_csr(rd["pr"], rd["pc"]),
                            False)
        timed(f"merge_round{r}", lambda: cpu_ref.merge_nodes(
            base, S, off2, srcs2, p.link_rev, rd["pr"], rd["pc"], p.M, True),
            True)
    total = sum(t.values())
    log("cpu oracle stage seconds (extrapolated to N):",
        {k: round(v, 2) for k, v in t.items()})
    return {"value": round(n / total, 1), "unit": "base points/s",
            "cores": threads, "kind": "port",
            "sample": f"oracle (cpu_ref, C++/OpenMP) per-node stages timed on "
                      f"{S.shape[0]} random nodes of {n} and scaled by "
                      f"{scale:.0f}; match and reverse-list stages run on all "
                      f"nodes; extrapolated stage time {total:.1f}s",
            "stage_s": {k: round(v, 3) for k, v in t.items()}}


if __name__ == "__main__":
    main()
