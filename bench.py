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
# MI355X: 256 CUs x 4 SIMDs at 2.4 GHz; a wave64 VALU instruction occupies
# its SIMD for 4 cycles (16 lanes per cycle), so the issue peak is
# 1024 x 2.4e9 / 4 wave instructions per second.
N_SIMDS = 1024
CLOCK_HZ = 2.4e9
VALU_PEAK_GIPS = N_SIMDS * CLOCK_HZ / 4 / 1e9


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def profiled(name, kernel, workload, precision):
    """A per-launch figure `name[kernel,workload,precision] = value`
    recorded in profiles/*.md from rocprofv3 runs (HBM bytes scaled by a
    1 GiB copy, VALU instructions, MFMA busy cycles, mean launch ms), or
    None."""
    import glob
    import re
    key = f"{name}[{kernel},{workload},{precision}]"
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            if line.startswith(key):
                m = re.search(r"=\s*([0-9.eE+-]+)", line)
                if m:
                    return float(m.group(1))
    return None


def stage_bytes(kind, c, p, n, D=None, mp=None, deg=None, search_row=None):
    """Algorithmic bytes of one launch from its counters: each distinct
    (node-task, row) pair once, rows charged at s * dim bytes (no padding).
    FP32 rows for prune/merge; the search reads FP32 rows, or the screened
    copy's rows (search_row bytes)."""
    row = 4 * p.dim
    if kind == "search":
        srow = search_row or row
        return (4 * p.dim * n + srow * c["search_dist"] + 4 * p.Kp * p.Kd *
                c["search_iters"] + 4 * p.topk * n)
    if kind == "prune":
        return row * (n + c["prune_dist"]) + 4 * D * n + 4 * mp * n
    if kind == "merge":
        return (row * (n + c["merge_dist"]) + 4 * c["merge_dist"] +
                4 * deg * n)
    return 0


def time_build(ctx, base_d, gt_d, proj_d, link_d, steps, warmup, label):
    import torch
    for i in range(warmup):
        ctx.stage_device(base_d, gt_d, proj_d, link_d)
        ctx.wait()
        log(f"{label} warmup {i} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        ctx.stage_device(base_d, gt_d, proj_d, link_d)
        ctx.wait()
        log(f"{label} step {i} done")
    torch.cuda.synchronize()
    return 1000.0 * (time.perf_counter() - t0) / steps


def roofline_of(ctx, p, n, workload, precision):
    """Dominant kernel of the last build against the resource that binds
    it.  Three fractions over its mean launch time (HIP events on the build
    stream): HBM (rocprof FETCH_SIZE + WRITE_SIZE bytes per launch, scaled
    by a 1 GiB copy, over the chip's 8 TB/s), VALU issue (rocprof
    SQ_INSTS_VALU per launch over 1024 SIMDs x 2.4 GHz / 4 cycles) and MFMA
    (SQ_VALU_MFMA_BUSY_CYCLES per launch over 1024 SIMDs x 2.4 GHz); the
    largest is `bound`.  The algorithmic bytes (rows at s * dim bytes,
    cache-inclusive) over the same time are reported beside them."""
    from paragraph_amd import _native
    st = ctx.stats()
    stage_ms = [float(st.stage_ms[i]) for i in range(st.n_stages)]
    names = [_native.stage_name(i) for i in range(st.n_stages)]
    counters = [ctx.stage_counters(i) for i in range(st.n_stages)]
    # bytes of one base row as the search reads it (no padding): BF16,
    # e4m3 plus its FP32 scale, or FP32
    row = {"bf16": 2 * p.dim, "fp8": p.dim + 4}.get(precision, 4 * p.dim)
    kinds = {}
    for i, nm in enumerate(names):
        kind = nm.split("_")[0]
        if kind == "match":
            continue
        D = p.gt_k if nm == "prune_proj" else p.topk
        mp = p.proj_pr if nm == "prune_proj" else p.M
        deg = p.proj_deg if nm == "merge_proj" else p.M
        b = stage_bytes(kind, counters[i], p, n, D=D, mp=mp, deg=deg,
                        search_row=row)
        # link rounds only: the projection launches are small and of
        # another shape
        if nm.endswith("_proj"):
            continue
        k = kinds.setdefault(kind, {"ms": 0.0, "bytes": 0, "launches": 0})
        k["ms"] += stage_ms[i]
        k["bytes"] += b
        k["launches"] += 1
    dom = max(kinds, key=lambda k: kinds[k]["ms"])
    kd = kinds[dom]
    avg_ms = kd["ms"] / kd["launches"]
    avg_s = avg_ms / 1000.0
    bytes_per_launch = kd["bytes"] / kd["launches"]
    alg_gbs = bytes_per_launch / avg_s / 1e9
    # the link-round prune runs as the kept-major sweep unless tune_flags
    # bit 8 selects the candidate-major kernel
    prune_k = "prune_kernel" if p.tune_flags & 0x100 else "prune_sweep_kernel"
    kname = {"search": "search_kernel", "prune": prune_k,
             "merge": "merge_kernel", "reverse": "reverse_*"}.get(dom, dom)
    prof = lambda what: profiled(what, kname, workload, precision)
    traffic = prof("traffic_bytes_per_launch")
    valu = prof("valu_insts_per_launch")
    mfma = prof("mfma_busy_cycles_per_launch")
    fr = {}
    if traffic:
        fr["hbm"] = (traffic / avg_s / 1e9, HBM_PEAK_GBS, "GB/s")
    if valu:
        fr["valu"] = (valu / avg_s / 1e9, VALU_PEAK_GIPS,
                      "G wave64 VALU instructions/s")
    if mfma:
        fr["mfma"] = (100.0 * mfma / (N_SIMDS * CLOCK_HZ * avg_s), 100.0,
                      "% of MFMA cycles busy")
    if fr:
        bound = max(fr, key=lambda k: fr[k][0] / fr[k][1])
        a, pk, unit = fr[bound]
        frac = round(a / pk, 4)
    else:
        # no profile of this kernel on this workload: only the algorithmic
        # (cache-inclusive) byte rate, which names no bound
        bound, a, pk, unit, frac = None, alg_gbs, HBM_PEAK_GBS, "GB/s", None
    roofline = {"bound": bound, "achieved": round(a, 2), "peak": pk,
                "unit": unit, "frac": frac,
                "traffic": int(traffic) if traffic else None,
                "kernel": kname, "avg_launch_ms": round(avg_ms, 3),
                "profile_avg_launch_ms": prof("profile_avg_launch_ms"),
                "alg_bytes_per_launch": int(bytes_per_launch),
                "alg_gbs": round(alg_gbs, 1),
                "alg_frac": round(alg_gbs / HBM_PEAK_GBS, 4),
                "fractions": {k: round(v[0] / v[1], 4)
                              for k, v in fr.items()}}
    if traffic:
        roofline["hbm_gbs"] = round(traffic / avg_s / 1e9, 1)
        roofline["hbm_frac"] = round(traffic / avg_s / 1e9 / HBM_PEAK_GBS,
                                     4)
    detail = {"stage_ms": {nm: round(x, 3) for nm, x in zip(names, stage_ms)},
              "stage_counters": {nm: c for nm, c in zip(names, counters)}}
    return roofline, detail


def recall_at_10(eval_ctx, eq_d, base_d, link_d, truth, topk):
    import torch
    out_d = torch.empty((eq_d.shape[0], topk), dtype=torch.int32,
                        device=eq_d.device)
    eval_ctx.run_search(eq_d, base_d, link_d, out_d)
    eval_ctx.sync()
    found = out_d[:, :10].cpu().numpy()
    hits = sum(len(set(a) & set(b)) for a, b in zip(found, truth))
    return hits / truth.size


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg2_1m_d200")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--precision", default="fp32,bf16",
                    help="comma list of builds to time (fp32, bf16, fp8); "
                         "value = the first screened one (bf16 / fp8), or "
                         "fp32 when it is the only one")
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="CPU baseline on this many random nodes, scaled "
                         "to N (default: the whole workload in full up to "
                         "1M base rows, about 30 s on 16 threads; a "
                         "2000-node sample above)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-recall", action="store_true")
    ap.add_argument("--no-dist-error", action="store_true")
    ap.add_argument("--no-host-window", action="store_true",
                    help="skip the builds from host buffers (PCIe-"
                         "inclusive rates, reported beside value)")
    ap.add_argument("--n-eval", type=int, default=10000)
    ap.add_argument("--tune-flags", type=int, default=None,
                    help="perf-only knobs for the FP32 build (default: the "
                         "library's)")
    ap.add_argument("--tune-flags-bf16", type=int, default=None,
                    help="perf-only knobs for the BF16 build")
    ap.add_argument("--bf16-params", default="",
                    help="overrides of the BF16 build's knobs, e.g. "
                         "'Kp=16,table_size=2048' (beam knobs change recall; "
                         "the FP32 build always uses the reference mapping)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the autotuner (library defaults; the "
                         "screened builds then use the FP32 beam mapping)")
    ap.add_argument("--beam", type=int, default=None,
                    help="beam width L of the screened builds (Km = L, "
                         "Kp = ceil(L/16), topk = 2L), skipping the beam "
                         "tuner")
    ap.add_argument("--n-tune", type=int, default=50000,
                    help="tuning queries the beam autotuner scores recall "
                         "on (a seed disjoint from the eval queries)")
    ap.add_argument("--no-controls", action="store_true",
                    help="skip the control builds (FP32 at the tuned "
                         "beam, the screened path at the workload's beam, "
                         "the final index without link rows)")
    ap.add_argument("--gt-subset", type=int, default=None,
                    help="gt = exact top-128 within shards of this many base "
                         "rows (a quick stand-in; default: exact top-128 "
                         "over all rows, by torch's FP32 product up to 3M "
                         "base rows and by the library's exact top-k "
                         "kernel above, about 8 minutes at 10M x 200)")
    ap.add_argument("--end-to-end", action="store_true",
                    help="also time the whole index build of the value "
                         "path, twice: with the host stages on the host "
                         "(bipartite stage on a CPU thread beside the GPU "
                         "stage, then FusionFinal and the entry point) and "
                         "with them on the GPU (bipartite stage on a second "
                         "stream); reported in extra, never as value")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 1 GiB device copy and a 1 GiB sum first "
                         "(counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; "
                          "--gpus must be 1"}))
        sys.exit(2)

    import torch
    from paragraph_amd import _native, datagen

    dev = torch.device("cuda:0")
    t0 = time.time()
    n_cfg = int(datagen.CONFIGS[args.config]["n_base"] * args.scale)
    gt_subset = args.gt_subset
    w = datagen.make_config(args.config, scale=args.scale,
                            n_eval=0 if args.no_recall else args.n_eval,
                            device="cuda", gt_subset=gt_subset,
                            exact_native=n_cfg > 3_000_000,
                            log=lambda t: log(t))
    n, dim = w["base"].shape
    M, L = w["M"], w["L"]
    log(f"workload {args.config} n={n} dim={dim} M={M} L={L} generated in "
        f"{time.time() - t0:.1f}s")
    gt_u32 = np.ascontiguousarray(w["gt"].view(np.uint32))
    base_d = torch.from_numpy(w["base"]).to(dev)
    gt_d = torch.from_numpy(w["gt"]).to(dev)
    if args.calib_copy:
        # reference points for the HBM counters: a plain 1 GiB device copy
        # (reads 1 GiB, writes 1 GiB) and a 1 GiB sum (reads 1 GiB only)
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        total = float(src.sum())
        log(f"calibration: 1 GiB copy and 1 GiB sum ({total:.0f})")
        del src, dst

    precs = [x.strip() for x in args.precision.split(",") if x.strip()]
    if args.precision == "both":
        precs = ["fp32", "bf16"]
    for x in precs:
        if x not in ("fp32", "bf16", "fp8"):
            ap.error(f"unknown precision {x}")
    results = {}
    tf32 = {} if args.tune_flags is None else {"tune_flags": args.tune_flags}
    tf16 = {} if args.tune_flags_bf16 is None else {
        "tune_flags": args.tune_flags_bf16}
    p32 = _native.make_params(n, dim, n, M, L, **tf32)

    def exact_top10(q):
        q_d = torch.from_numpy(np.ascontiguousarray(q)).to(dev)
        B64 = base_d.double()
        tl = []
        for s in range(0, q.shape[0], 1000):
            sc = q_d[s:s + 1000].double() @ B64.T
            tl.append(torch.topk(sc, 10, dim=1).indices.cpu().numpy())
            del sc
        del B64
        torch.cuda.empty_cache()
        return q_d, np.concatenate(tl)

    # the evaluation searcher: reference mapping at the workload's L
    p_eval = _native.make_params(n, dim, n, M, L)
    eval_ctx = _native.Context(p_eval)
    # the final index of a build (its GPU stage rows through the host stages
    # on the GPU, on the evaluation context)
    from paragraph_amd import host_stage
    hp_final = host_stage.make_params(n, dim, gt_u32.shape[0],
                                      gt_k=gt_u32.shape[1])
    fin_width = int(host_stage.lib().pg_host_final_width(
        host_stage.ctypes.byref(hp_final)))
    fin_off_d = torch.empty(n + 1, dtype=torch.int32, device=dev)
    fin_ids_d = torch.empty(n * fin_width, dtype=torch.int32, device=dev)

    def final_index(proj, link):
        eval_ctx.run_fusion(hp_final, base_d, gt_d, proj, link, fin_off_d,
                            fin_ids_d)
        eval_ctx.sync()
        off = fin_off_d.cpu().numpy().view(np.uint32)
        return off, fin_ids_d[:int(off[-1])].cpu().numpy().view(np.uint32)

    truth = None
    if not args.no_recall:
        eq_d, truth = exact_top10(w["eval_queries"])
    screened_req = [x for x in precs if x != "fp32"]
    tune_beam = args.autotune and bool(screened_req) and args.beam is None
    if tune_beam:
        tq_d, tq_truth = exact_top10(datagen.tune_queries(args.config,
                                                          args.n_tune))

        # the screened beam is scored on the per-query recall of the index
        # it builds
        def tune_hits(proj, link):
            return final_index_hits(final_index(proj, link), base_d, tq_d,
                                    tq_truth, n, dim)
    ref_tune_hits = None
    ref_link = None
    if tune_beam and "fp32" not in precs:
        # the FP32 build's recall on the tuning queries is the reference
        with _native.Context(p32) as c32:
            pr_ = torch.empty((n, p32.proj_deg), dtype=torch.int32,
                              device=dev)
            lk_ = torch.empty((n, M), dtype=torch.int32, device=dev)
            c32.stage_device(base_d, gt_d, pr_, lk_)
            c32.wait()
            ref_tune_hits = tune_hits(pr_, lk_)
            ref_link = lk_
            del pr_
    # FP32 first: its recall anchors the beam tuning of the screened builds
    order = (["fp32"] if "fp32" in precs else []) + screened_req
    for prec in order:
        if prec == "fp32":
            p = p32
        else:
            ov = dict(kv.split("=") for kv in args.bf16_params.split(",")
                      if kv)
            kw = dict(tf16)
            kw.update({k: int(v) for k, v in ov.items()})
            p = _native.make_params(
                n, dim, n, M, L, precision=_native.PG_BF16 if prec == "bf16"
                else _native.PG_FP8E4M3, **kw)
        tuned = None
        from paragraph_amd import autotune
        if prec != "fp32" and args.beam is not None:
            p = autotune.beam_params(p, args.beam)
            # explicit knobs win over the beam mapping's defaults
            for k, v in ov.items():
                setattr(p, k, int(v))
        if args.autotune:
            tuned = {}
            if prec != "fp32" and tune_beam:
                p, brep = autotune.tune_beam(base_d, gt_d, p, tune_hits,
                                             ref_tune_hits, log=log,
                                             ref_link=ref_link)
                tuned["beam"] = brep
                log(f"{prec}: autotuned beam L={brep['L']} in "
                    f"{brep['seconds']} s")
            best, rep = autotune.autotune(base_d, gt_d, p, use_cache=False,
                                          log=log)
            p = autotune.apply(p, best)
            if prec == "fp32":
                p32 = p
            tuned["knobs"] = best
            tuned["knob_seconds"] = rep["seconds"]
            log(f"{prec}: autotuned knobs {best} in {rep['seconds']} s")
        proj_d = torch.empty((n, p.proj_deg), dtype=torch.int32, device=dev)
        link_d = torch.empty((n, M), dtype=torch.int32, device=dev)
        ctx = _native.Context(p)
        ms = time_build(ctx, base_d, gt_d, proj_d, link_d, args.steps,
                        args.warmup, prec)
        if p.tune_flags & 0x40:
            # graph replays carry no stage events: one eager build of the
            # same knobs gives the per-stage breakdown
            pe = _native.PgParams.from_buffer_copy(p)
            pe.tune_flags = p.tune_flags & ~0x40
            with _native.Context(pe) as ectx:
                time_build(ectx, base_d, gt_d, proj_d, link_d, 1, 0,
                           prec + " (eager breakdown)")
                roof, detail = roofline_of(ectx, pe, n, args.config, prec)
        else:
            roof, detail = roofline_of(ctx, p, n, args.config, prec)
        r = {"ms_per_step": ms, "value": n / (ms / 1000.0),
             "roofline": roof, **detail,
             "counters": {k: v for k, v in ctx.stats().as_dict().items()
                          if k != "stage_ms"}}
        log(f"{prec}: {ms:.1f} ms/step, {r['value']:.0f} pts/s; stage ms "
            f"{ {k: round(v, 1) for k, v in detail['stage_ms'].items()} }")
        if prec == "fp32" and tune_beam:
            ref_tune_hits = tune_hits(proj_d, link_d)
            ref_link = link_d
        if truth is not None:
            r["link_graph_recall_at_10"] = round(recall_at_10(
                eval_ctx, eq_d, base_d, link_d, truth, p_eval.topk), 4)
            log(f"{prec}: recall@10 of the link graph "
                f"({truth.shape[0]} held-out queries) = "
                f"{r['link_graph_recall_at_10']}")
            r["final_index_recall_at_10"] = final_index_recall(
                final_index(proj_d, link_d), base_d, eq_d, truth, n, dim)
            log(f"{prec}: recall@10 of the final index (the reference "
                "evaluation searcher's knobs: Km 128, Kp 6, Kd 64, reset "
                "every 4; visited table cleared per query) = "
                f"{r['final_index_recall_at_10']}")
        if tuned is not None:
            r["autotune"] = tuned
        r["tune_flags"] = int(p.tune_flags)
        r["table_size"] = int(p.table_size)
        r["beam"] = {"L": int(p.Km), "Kp": int(p.Kp), "Kd": int(p.Kd),
                     "topk": int(p.topk)}
        r["link"] = link_d
        r["proj"] = proj_d
        r["ctx"] = ctx
        results[prec] = r

    # ---- controls: where the screened speed-up comes from
    controls = None
    if (not args.no_controls and "fp32" in results and truth is not None
            and any(x != "fp32" for x in results)):
        controls = beam_controls(results, p32, base_d, gt_d, n, M, L, dim,
                                 args.steps, eval_ctx, p_eval, eq_d, truth,
                                 final_index, log)

    # ---- host window: whole builds from host buffers through the host
    # entry point (inputs streamed in on the copy stream, outputs streamed
    # out), for every precision: pinned buffers (pg_host_alloc), timed over
    # the same steps, and one build from pageable buffers; reported beside
    # `value`, never as it
    host_window = None
    if not args.no_host_window:
        host_window = time_host_window(results, w, gt_u32, n, M,
                                       max(1, args.steps), log)

    # ---- end-to-end window: the reference's whole build
    # (tests/test_t2iend.cu:850-1243), inputs in host memory for the host
    # stages and in HBM for the GPU stage
    end_to_end = None
    if args.end_to_end:
        end_to_end = time_end_to_end(results, precs, w, gt_u32, base_d, gt_d,
                                     n, M, log)
        end_to_end.pop("_final", None)

    # ---- CPU baseline: the oracle, in full up to 1M base rows, else on a
    # bounded node sample
    cpu = None
    if not args.no_cpu_baseline:
        if args.cpu_sample is None and n <= 1_000_000:
            cpu = cpu_full(p32, w, gt_u32, results.get("fp32"), log)
        else:
            ctx32 = results["fp32"]["ctx"] if "fp32" in results else \
                _native.Context(p32)
            cpu = cpu_baseline(ctx32, p32, w, gt_u32, base_d, gt_d,
                               args.cpu_sample or 2000, log)

    screened = [x for x in precs if x != "fp32"]
    dist_err = None
    if not args.no_dist_error:
        ev = {}
        ctx32 = results["fp32"]["ctx"] if "fp32" in results else \
            _native.Context(p32)
        ev["fp32_search"] = (ctx32, False)
        ev["fp32_chain"] = (ctx32, True)
        for sp in screened:
            ev[sp + "_screen"] = (results[sp]["ctx"], False)
        dist_err = distance_error(ev, w, base_d)
        log(f"distance error vs FP64: {dist_err}")
    head_prec = screened[0] if screened else "fp32"
    head = results[head_prec]
    extra = {}
    for prec, r in results.items():
        extra[prec] = {k: (round(v, 3) if isinstance(v, float) else v)
                       for k, v in r.items()
                       if k not in ("link", "proj", "ctx")}
    if dist_err is not None:
        extra["distance_error_vs_fp64"] = dist_err
    if host_window is not None:
        extra["host_window"] = host_window
    if end_to_end is not None:
        extra["end_to_end"] = end_to_end
    if controls is not None:
        extra["controls"] = controls
    if truth is not None:
        extra["final_index_searcher"] = (
            "pg_run_search with search_shared_hashtable's knobs (Km 128, Kp "
            "6, Kd 64, topk 128, reset every 4, entry ids (3030517 + i) mod "
            "N) over the final lists padded to rows; each query starts from "
            "an empty visited table, where the reference's warp keeps its "
            "previous query's table until the first reset")
    for sp in screened:
        if "fp32" not in results:
            break
        extra[f"speedup_{sp}_vs_fp32"] = round(
            results["fp32"]["ms_per_step"] / results[sp]["ms_per_step"], 3)
        if truth is not None:
            # the index's recall (the tolerance of the metric) and the GPU
            # stage's link graph alone (a proxy, searched by the build's own
            # searcher at the workload's L)
            extra[f"link_graph_recall_delta_pt_{sp}"] = round(
                100 * (results[sp]["link_graph_recall_at_10"] -
                       results["fp32"]["link_graph_recall_at_10"]), 2)
            extra[f"final_index_recall_delta_pt_{sp}"] = round(
                100 * (results[sp]["final_index_recall_at_10"] -
                       results["fp32"]["final_index_recall_at_10"]), 2)
        a = results["fp32"]["link"].cpu().numpy()
        b = results[sp]["link"].cpu().numpy()
        # per-node neighbour-set overlap of the link rows (first 20k nodes)
        extra[f"link_overlap_{sp}_vs_fp32"] = round(float(np.mean([
            len(set(x[x >= 0]) & set(y[y >= 0])) / max(1, len(set(x[x >= 0])))
            for x, y in zip(a[:20000], b[:20000])])), 4)
    line = {
        "metric": "base points indexed/sec, 1 MI355X, FP32 vs BF16-MFMA at "
                  "recall@10 within 0.5 pt",
        "value": round(head["value"], 1),
        "unit": "base points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16": "bf16", "fp8": "fp8e4m3"}[head_prec],
        "data": "synthetic (seeded Gaussian-mixture generator of the "
                "workload; " + ("exact top-128 inner-product gt" if not
                                gt_subset else
                                f"gt = exact top-128 inner product within "
                                f"{gt_subset}-row base shards") + ")",
        "config": {"workload": f"{args.config}: {n} base x {dim} f32, {n} "
                   f"training queries, gt k=128, M={M}, L={L}, inner "
                   f"product; value = {head_prec} path"
                   + ({"bf16": " (BF16 MFMA search screening, FP32 "
                       "re-rank, BF16-screened exact prune)",
                       "fp8": " (FP8 e4m3 per-vector-scaled MFMA search "
                       "screening, FP32 re-rank, BF16-screened exact "
                       "prune)"}.get(head_prec, ""))
                   + (f"; screened beam L={head['beam']['L']}, re-rank "
                      f"width {head['beam']['topk']}"
                      + (" (autotuned: beams walked down from L in steps "
                         "of 8, stopping at the first whose final-index "
                         "recall@10 change against the FP32 build has a "
                         "lower confidence bound (mean - 2 SE, paired, on "
                         "separate tuning queries) below -min(0.5 pt, 5%); "
                         "the fastest accepted beam)"
                         if tune_beam else "")
                      if head_prec != "fp32" else ""),
                   "n_base": n, "dim": dim, "M": M, "L": L,
                   "screened_L": head["beam"]["L"],
                   "precision": head_prec},
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    for r in results.values():
        r["ctx"].close()


def time_host_window(results, w, gt_u32, n, M, steps, log):
    """Host window of each built precision: `steps` builds from pinned host
    buffers (after one untimed), and one from pageable buffers; the link
    rows are checked against the device entry's."""
    import torch
    from paragraph_amd import _native
    dim = w["base"].shape[1]
    base_p = _native.PinnedArray((n, dim), np.float32)
    gt_p = _native.PinnedArray(gt_u32.shape, np.uint32)
    base_p.a[:] = w["base"]
    gt_p.a[:] = gt_u32
    base_h = np.ascontiguousarray(w["base"])
    out = {}
    for prec, r in results.items():
        ctx = r["ctx"]
        proj_p = _native.PinnedArray((n, ctx.params.proj_deg), np.uint32)
        link_p = _native.PinnedArray((n, M), np.uint32)
        ctx.stage_host(base_p.a, gt_p.a, proj_p.a, link_p.a)
        ctx.wait()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.stage_host(base_p.a, gt_p.a, proj_p.a, link_p.a)
            ctx.wait()
        ms = 1000.0 * (time.perf_counter() - t0) / steps
        same = bool(np.array_equal(link_p.a,
                                   r["link"].cpu().numpy().view(np.uint32)))
        hproj = np.empty((n, ctx.params.proj_deg), np.uint32)
        hlink = np.empty((n, M), np.uint32)
        t1 = time.perf_counter()
        ctx.stage_host(base_h, gt_u32, hproj, hlink)
        ctx.wait()
        ms_page = 1000.0 * (time.perf_counter() - t1)
        same = same and bool(np.array_equal(hlink, link_p.a))
        out[prec] = {"pinned_ms": round(ms, 2),
                     "pinned_base_points_per_s": round(n / (ms / 1000.0), 1),
                     "pageable_ms": round(ms_page, 2),
                     "device_ms": round(r["ms_per_step"], 2),
                     "same_rows_as_device_entry": same}
        log(f"host window {prec}: {out[prec]}")
        del proj_p, link_p
    screened = [x for x in out if x != "fp32"]
    if "fp32" in out and screened:
        out[f"speedup_{screened[0]}_vs_fp32_pinned"] = round(
            out["fp32"]["pinned_ms"] / out[screened[0]]["pinned_ms"], 3)
    return out


def beam_controls(results, p32, base_d, gt_d, n, M, L, dim, steps,
                  eval_ctx, p_eval, eq_d, truth, final_index, log):
    """Control builds beside the value path, to separate what precision and
    what the beam give: the FP32 path at the screened path's tuned beam,
    the screened path at the workload's beam L (the precision-only
    speed-up), and the final index of the FP32 build with its link rows
    (and its projection rows) replaced by tombs, i.e. how much of the final
    index's recall the GPU stage's rows carry."""
    import torch
    from paragraph_amd import _native, autotune
    head = next(x for x in results if x != "fp32")
    hp = results[head]["ctx"].params
    f32 = results["fp32"]
    i32 = dict(dtype=torch.int32, device=base_d.device)
    out = {"screened_path": head}

    def build(p, label):
        proj = torch.empty((n, p.proj_deg), **i32)
        link = torch.empty((n, M), **i32)
        with _native.Context(p) as ctx:
            ms = time_build(ctx, base_d, gt_d, proj, link, steps, 1, label)
        r = {"L": int(p.Km), "ms_per_step": round(ms, 2),
             "value": round(n / (ms / 1000.0), 1),
             "link_graph_recall_at_10": round(recall_at_10(
                 eval_ctx, eq_d, base_d, link, truth, p_eval.topk), 4),
             "final_index_recall_at_10": final_index_recall(
                 final_index(proj, link), base_d, eq_d, truth, n, dim),
             "link_overlap_vs_fp32": autotune.link_overlap(f32["link"],
                                                           link)}
        log(f"control {label}: {r}")
        return r

    Lt = int(hp.Km)
    if Lt != L:
        out["fp32_at_tuned_beam"] = build(autotune.beam_params(p32, Lt),
                                          f"fp32 L={Lt}")
        pl = autotune.beam_params(hp, L)
        out[f"{head}_at_workload_beam"] = build(pl, f"{head} L={L}")
        out["speedup_precision_only"] = round(
            f32["ms_per_step"] / out[f"{head}_at_workload_beam"]
            ["ms_per_step"], 3)
        out["speedup_same_beam_tuned"] = round(
            out["fp32_at_tuned_beam"]["ms_per_step"] /
            results[head]["ms_per_step"], 3)
    out["speedup_beam_tuned"] = round(
        f32["ms_per_step"] / results[head]["ms_per_step"], 3)
    # the GPU stage's share of the final index: FusionFinal without the
    # link rows, and without the link and projection rows
    proj32 = results["fp32"]["proj"]
    tomb_link = torch.full((n, M), -1, **i32)
    tomb_proj = torch.full_like(proj32, -1)
    out["fp32_final_index_recall_at_10"] = \
        results["fp32"]["final_index_recall_at_10"]
    out["fp32_final_index_recall_without_link_rows"] = final_index_recall(
        final_index(proj32, tomb_link), base_d, eq_d, truth, n, dim)
    out["fp32_final_index_recall_without_link_or_projection_rows"] = \
        final_index_recall(final_index(tomb_proj, tomb_link), base_d, eq_d,
                           truth, n, dim)
    log(f"controls: {out}")
    return out


def time_end_to_end(results, precs, w, gt_u32, base_d, gt_d, n, M, log):
    """One whole index build of the value path: the GPU stage (device
    entry) with the host bipartite stage on a CPU thread beside it, then
    FusionFinal and the entry point on the host."""
    import threading

    import torch
    from paragraph_amd import host_stage
    prec = next((x for x in precs if x != "fp32"), "fp32")
    ctx = results[prec]["ctx"]
    base_h = np.ascontiguousarray(w["base"])
    hp = host_stage.make_params(n, base_h.shape[1], gt_u32.shape[0],
                                gt_k=gt_u32.shape[1])
    proj_d = torch.empty((n, ctx.params.proj_deg), dtype=torch.int32,
                         device=base_d.device)
    link_d = torch.empty((n, M), dtype=torch.int32, device=base_d.device)
    torch.cuda.synchronize()
    bip, t_host = {}, {}

    def host_work():
        t0 = time.perf_counter()
        bip["csr"] = host_stage.bipartite(hp, base_h, gt_u32)
        t_host["bipartite"] = time.perf_counter() - t0

    t0 = time.perf_counter()
    th = threading.Thread(target=host_work)
    th.start()
    ctx.stage_device(base_d, gt_d, proj_d, link_d)
    ctx.wait()
    t_gpu = time.perf_counter() - t0
    proj = proj_d.cpu().numpy().view(np.uint32)
    link = link_d.cpu().numpy().view(np.uint32)
    th.join()
    t1 = time.perf_counter()
    ep = host_stage.entry_point(hp, gt_u32)
    off, ids = host_stage.fusion_final(hp, base_h, *bip["csr"], link, proj)
    t_fin = time.perf_counter() - t1
    total = time.perf_counter() - t0
    lens = np.diff(off.astype(np.int64))
    r = {"precision": prec, "ms": round(1000 * total, 1),
         "base_points_per_s": round(n / total, 1),
         "gpu_stage_ms": round(1000 * t_gpu, 1),
         "host_bipartite_ms": round(1000 * t_host["bipartite"], 1),
         "fusion_final_and_ep_ms": round(1000 * t_fin, 1),
         "host_threads": host_stage.default_threads(),
         "entry_point": ep, "final_degree_mean": round(float(lens.mean()), 2),
         "final_degree_max": int(lens.max())}
    log(f"end-to-end build, host stages on the host ({prec}): {r}")
    # the same build with the host stages on the GPU: the bipartite stage
    # beside the GPU stage on a second stream, FusionFinal after it; only
    # the sequential entry-point scan stays on the host, beside both
    width = int(host_stage.lib().pg_host_final_width(
        host_stage.ctypes.byref(hp)))
    off_d = torch.empty(n + 1, dtype=torch.int32, device=base_d.device)
    ids_d = torch.empty(n * width, dtype=torch.int32, device=base_d.device)
    ctx.run_fusion(hp, base_d, gt_d, proj_d, link_d, off_d, ids_d)  # warm-up
    ctx.sync()
    gtimes = []
    for _ in range(2):
        torch.cuda.synchronize()
        eps = {}
        t0 = time.perf_counter()
        th = threading.Thread(
            target=lambda: eps.setdefault("ep", host_stage.entry_point(
                hp, gt_u32)))
        th.start()
        # the bipartite stage on the context's second stream beside the GPU
        # stage, then FusionFinal on the main stream
        ctx.run_bipartite_async(hp, base_d, gt_d)
        ctx.stage_device(base_d, gt_d, proj_d, link_d)
        ctx.run_fusion_final(hp, base_d, proj_d, link_d, off_d, ids_d)
        ctx.wait()
        off_g = off_d.cpu().numpy().view(np.uint32)
        ids_g = ids_d[:int(off_g[-1])].cpu().numpy().view(np.uint32)
        th.join()
        gtimes.append(time.perf_counter() - t0)
    tg = min(gtimes)
    r_gpu = {"precision": prec, "ms": round(1000 * tg, 1),
             "base_points_per_s": round(n / tg, 1),
             "same_lists_as_host_stages": bool(
                 np.array_equal(off_g, off) and np.array_equal(ids_g, ids)
                 and eps["ep"] == ep)}
    log(f"end-to-end build, host stages on the GPU ({prec}): {r_gpu}")
    return {"host_stages_on_host": r, "host_stages_on_gpu": r_gpu,
            "_final": (off, ids)}


def final_index_hits(final, base_d, eq_d, truth, n, dim):
    """Per-query recall@10 hits (0..10) of the final index with the
    reference's evaluation searcher knobs (search_shared_hashtable,
    Gsearcher.cuh:108-296: Km 128, Kp 6, Kd 64, reset every 4) over its
    lists padded to rows.  Unlike the reference's searcher, each query
    starts from an empty visited table (the reference's warp keeps the
    table of its previous query, Gsearcher.cuh:206-210)."""
    import torch
    from paragraph_amd import _native, io
    off, ids = final
    width = max(64, int(np.diff(off.astype(np.int64)).max()))
    rows_d = torch.from_numpy(io.csr_to_rows(off, ids, width).view(
        np.int32)).to(base_d.device)
    p = io.eval_search_params(n, dim, eq_d.shape[0])
    out_d = torch.empty((eq_d.shape[0], p.topk), dtype=torch.int32,
                        device=base_d.device)
    with _native.Context(p) as ctx:
        ctx.run_search(eq_d, base_d, rows_d, out_d)
        ctx.sync()
    found = out_d[:, :10].cpu().numpy()
    # the search emits distinct ids, so hits = matches against the truth row
    return (found[:, :, None] == truth[:, None, :10]).any(axis=2).sum(axis=1)


def final_index_recall(final, base_d, eq_d, truth, n, dim):
    """recall@10 of the final index (final_index_hits over all queries)."""
    return round(float(final_index_hits(final, base_d, eq_d, truth, n,
                                        dim).sum()) / truth.size, 4)


def distance_error(ctxs, w, base_d, n_sample=2000):
    """Error of each distance evaluator against FP64 on (training query,
    exact top-128 id) pairs: |d - d64| / (|x| |y|) (mean, p99, max), and the
    top-10 overlap inside each 128-id row between the evaluator's order and
    the FP64 order."""
    import torch
    nq = min(n_sample, w["queries"].shape[0])
    q_d = torch.from_numpy(np.ascontiguousarray(w["queries"][:nq])).to(
        base_d.device)
    ids = np.ascontiguousarray(w["gt"][:nq])
    ids_d = torch.from_numpy(ids).to(base_d.device)
    valid = ids_d >= 0
    safe = torch.where(valid, ids_d, torch.zeros_like(ids_d)).long()
    rows = base_d[safe].double()                    # nq x k x dim
    q64 = q_d.double()
    d64 = -(rows @ q64[:, :, None]).squeeze(2)
    norm = rows.norm(dim=2) * q64.norm(dim=1, keepdim=True)
    big = torch.tensor(float("inf"), dtype=torch.float64, device=d64.device)
    top64 = torch.topk(torch.where(valid, -d64, -big), 10, dim=1).indices
    out = {}
    for name, (ctx, chain) in ctxs.items():
        got = torch.empty(ids.shape, dtype=torch.float32, device=base_d.device)
        ctx.run_pair_distances(q_d, base_d, ids_d, got, chain=chain)
        ctx.sync()
        g = got.double()
        rel = ((g - d64).abs() / norm.clamp_min(1e-300))[valid]
        topg = torch.topk(torch.where(valid, -g, -big), 10, dim=1).indices
        ov = (topg[:, :, None] == top64[:, None, :]).any(2).double().mean()
        out[name] = {"rel_err_mean": float(rel.mean()),
                     "rel_err_p99": float(torch.quantile(
                         rel[:min(rel.numel(), 1 << 24)].float(), 0.99)),
                     "rel_err_max": float(rel.max()),
                     "top10_overlap_in_gt128": round(float(ov), 4)}
    out["pairs"] = int(valid.sum())
    return out


def cpu_full(p, w, gt_u32, r32, log):
    """Oracle throughput on this host over the whole workload: the FP32
    stage (match, projection, both link rounds) of every node, timed in
    full; its rows are compared with the FP32 GPU build's (r32)."""
    from oracle import cpu_ref
    d = p.as_dict()
    op = {k: d[k] for k, _ in cpu_ref.OrcParams._fields_}
    n = w["base"].shape[0]
    t0 = time.perf_counter()
    proj, link = cpu_ref.gpu_stage(op, w["base"], gt_u32)
    dt = time.perf_counter() - t0
    same = None
    if r32 is not None:
        same = bool(np.array_equal(
            proj, r32["proj"].cpu().numpy().view(np.uint32)) and
            np.array_equal(link, r32["link"].cpu().numpy().view(np.uint32)))
    log(f"cpu oracle, whole stage on all {n} nodes: {dt:.1f} s; rows equal "
        f"to the FP32 GPU build's: {same}")
    return {"value": round(n / dt, 1), "unit": "base points/s",
            "cores": cpu_ref.max_threads(), "kind": "port",
            "sample": f"the whole workload in full: oracle (cpu_ref, "
                      f"C++/OpenMP) FP32 stage over all {n} nodes, "
                      f"{dt:.1f} s",
            "seconds": round(dt, 2), "rows_equal_fp32_gpu": same}


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
