"""Writes the golden fixtures of the index-build stage into tests/golden/.

  hand_cases.json   tiny cases whose expected outputs are derived by hand or
                    by FP64 brute force (never by the oracle): top-1 match,
                    a 2-D prune, the merge duplicate / non-strict fill case,
                    and a complete-graph search.
  stage_*.npz       small whole-stage workloads (dim 8 / 200 / 512) from the
                    project generator; inputs stored exactly (values rounded
                    to float16 before use, ids as uint16 with 0xFFFF = tomb)
                    and the oracle's projection / link rows and counters as
                    the expected outputs (regression vectors).

Run from the repository root: python tests/golden/make_golden.py
(--hand-only rewrites hand_cases.json alone).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import cpu_ref  # noqa: E402
from paragraph_amd import datagen  # noqa: E402

TOMB = 0xFFFFFFFF
INT_MAX = 0x7FFFFFFF

# name: n_base, dim, n_query, n_comp, M, L, normalize, gt_len_range
STAGE_CASES = {
    "stage_d8": (1200, 8, 400, 16, 16, 32, False, None),
    "stage_d200": (600, 200, 600, 16, 32, 64, False, (16, 128)),
    "stage_d512": (256, 512, 256, 8, 24, 48, True, None),
}


def hand_cases():
    cases = {}
    # top-1 match: base b is claimed by the lowest query whose gt[q][0] == b
    gt = [[3, 1, 2], [0, 2, 1], [3, 0, 1], [TOMB, 1, 2], [7, 1, 2],
          [0, 1, 3]]
    cases["match"] = {"gt": gt, "n_base": 5,
                      "winner": [1, TOMB, TOMB, 0, TOMB]}
    # 2-D prune of node 0, IP distance -<b, c>:
    #   1: d=-3 kept first; 2: -<2,1> = -6 < -2 occluded; 3: -<3,1> = 0 < 0
    #   false -> kept; 4: -<4,3> = 0 < 1 occluded.  Other nodes: all-tomb.
    base = [[1.0, 0.0], [3.0, 0.0], [2.0, 2.0], [0.0, 1.0], [-1.0, 0.0]]
    rows = [[TOMB] * 6 for _ in range(5)]
    rows[0][:5] = [4, 3, 2, 1, 0]
    cases["prune_2d"] = {"base": base, "rows": rows, "mp": 4,
                         "pruned0": [1, 3], "pcnt": [2, 0, 0, 0, 0]}
    # merge of node 0: reverse {2}, pruned {1, 2}; sorted 1 (d=-2), 2 (rev),
    # 2 (pruned); strict keeps 1, 2 (-<2,1> = 0 < 0 false) and occludes the
    # second copy (-<2,2> = -1 < 0); the non-strict fill marks only the first
    # copy of 2 and appends the second.
    cases["merge_dup"] = {
        "base": [[1.0, 0.0], [2.0, 0.0], [0.0, 1.0], [5.0, 5.0]],
        "pruned": [[1, 2, TOMB], [TOMB] * 3, [TOMB] * 3, [TOMB] * 3],
        "pcnt": [2, 0, 0, 0], "off": [0, 1, 1, 1, 1], "srcs": [2],
        "cap": 3, "deg": 4,
        "strict0": [1, 2, TOMB, TOMB], "nonstrict0": [1, 2, 2, TOMB]}
    # complete-graph search: the result is the exact top-Km by distance; the
    # expected rows come from FP64 distances on points whose FP64 gaps are far
    # above float32 rounding
    rng = np.random.default_rng(21)
    n, dim, Km = 24, 16, 8
    while True:
        b = rng.integers(-8, 9, (n, dim)).astype(np.float32) / 4
        q = rng.integers(-8, 9, (4, dim)).astype(np.float32) / 4
        d = -(q.astype(np.float64) @ b.astype(np.float64).T)
        ok = True
        for row in d:
            s = np.sort(row)
            if np.min(np.diff(s[:Km + 1])) < 0.5:
                ok = False
        if ok:
            break
    expect = [sorted(range(n), key=lambda i: (row[i], i))[:Km] for row in d]
    graph = [[u for u in range(n) if u != v] for v in range(n)]
    cases["search_complete"] = {
        "base": b.tolist(), "queries": q.tolist(), "graph": graph,
        "Km": Km, "Kp": 1, "Kd": n - 1, "topk": Km, "entry_seed": 0,
        "rows": expect}
    cases["search_reset15"] = search_reset_case()
    cases["search_collect_kp2"] = search_collect_case()
    return cases


def search_reset_case():
    """K4 over a chain where the visited-table reset every 15 iterations
    re-admits nodes evaluated before it, which changes the output tail.

    18 nodes on a line, x_i = [i + 1], query [1]: d(i) = -(i + 1), exact in
    float32.  Degree 3: row k = [k+1, k-2, k-3] (tomb below 0) for k <= 16,
    row 17 = [16, 15, 14].  Km 2, Kp 1, Kd 3, topk = Km + Kp*Kd = 5 (the
    whole list), entry points 0, 1, 2 (entry_seed 0).  Derivation (F =
    FLT_MAX, * = expanded; the list is sorted by (distance, id)):
      init: [(-3,2) (-2,1) (-1,0) (F,T) (F,T)], visited = {2, 1} (the top
            Km only);
      it 1: expand 2, row [3, 0, T]: 3 and 0 are new ->
            [(-4,3) (-3,2*) (-2,1) (-1,0) (F,T)];
      it k = 2..14: expand k+1, row [k+2, k-1, k-2]: only k+2 is new ->
            [(-(k+3),k+2) (-(k+2),(k+1)*) (-(k+1),k*) (F,k-2) (F,k-1)];
      it 15: expand 16, row [17, 14, 13]: 17 new ->
            [(-18,17) (-17,16*) (-16,15*) (F,13) (F,14)];
            15th iteration: visited = the top Km ids {17, 16};
      it 16: expand 17, row [16, 15, 14]: 16 is in the re-seeded set (F),
            15 and 14 were dropped by the reset and are evaluated again ->
            [(-18,17*) (-17,16*) (-16,15) (-15,14) (F,16)];
      it 17: nothing unexpanded in the top 2: stop after 16 iterations,
            21 distance evaluations (3 + 2 + 14 + 2).
    Without the reset (or with it every 14 or 16 iterations) 15 and 14 stay
    visited in iteration 16 and the tail is ordered by id: [17, 16, 14, 15,
    16] (every 14: the reset after iteration 14 lets 14 and 13 be
    evaluated in iteration 15, and iteration 16 then finds 16, 15, 14 all
    visited)."""
    n = 18
    graph = []
    for k in range(n - 1):
        graph.append([k + 1 if k + 1 < n else TOMB,
                      k - 2 if k >= 2 else TOMB,
                      k - 3 if k >= 3 else TOMB])
    graph.append([16, 15, 14])
    return {"base": [[float(i + 1)] for i in range(n)], "queries": [[1.0]],
            "graph": graph, "Km": 2, "Kp": 1, "Kd": 3, "topk": 5,
            "entry_seed": 0,
            "rows_by_reset_iter": {"15": [[17, 16, 15, 14, 16]],
                                   "14": [[17, 16, 14, 15, 16]],
                                   "16": [[17, 16, 14, 15, 16]],
                                   "1000": [[17, 16, 14, 15, 16]]},
            "search_iters": 16, "search_dist": 21}


def search_collect_case():
    """K4 collect order with Kp = 2: the first two unexpanded entries of the
    top Km in list order, skipping expanded ones.

    12 nodes, dim 1, query [1], d(i) = -x_i with x = 1, 2, 4, 6, 5, 3, 7,
    0.5, 8, 0.25, 5.5, 0.125 (exact).  Degree 2: row 3 = [4, 5], row 2 =
    [6, 7], row 6 = [8, 9], row 4 = [10, 11], row 8 = [6, 3], row 10 =
    [4, 2], the others tomb.  Km 4, Kp 2, Kd 2, topk 8 (the whole list),
    entry points 0..3.  Derivation:
      init: [3 2 1 0 | T T T T] (d -6 -4 -2 -1), visited {3, 2, 1, 0};
      it 1: expand 3 and 2; 4, 5, 6, 7 are new (d -5 -3 -7 -0.5) ->
            top [6, 3*, 4, 2*], tail [5, 1, 0, 7];
      it 2: expand 6 (position 0) and 4 (position 2; 3* at position 1 is
            skipped); 8, 9, 10, 11 are new (d -8 -0.25 -5.5 -0.125) ->
            top [8, 6*, 3*, 10], tail [4*, 2*, 9, 11];
      it 3: expand 8 (position 0) and 10 (position 3); 6, 3, 4, 2 are all
            visited (F) -> top [8*, 6*, 3*, 10*], tail by id [2, 3, 4, 6];
      it 4: nothing unexpanded: stop after 3 iterations, 12 evaluations.
    Taking the first Kp entries regardless of the expanded bit would expand
    3 again in iteration 2 and change the rest."""
    x = [1, 2, 4, 6, 5, 3, 7, 0.5, 8, 0.25, 5.5, 0.125]
    graph = [[TOMB, TOMB] for _ in x]
    graph[3], graph[2] = [4, 5], [6, 7]
    graph[6], graph[4] = [8, 9], [10, 11]
    graph[8], graph[10] = [6, 3], [4, 2]
    return {"base": [[float(v)] for v in x], "queries": [[1.0]],
            "graph": graph, "Km": 4, "Kp": 2, "Kd": 2, "topk": 8,
            "entry_seed": 0,
            "rows_by_reset_iter": {"15": [[8, 6, 3, 10, 2, 3, 4, 6]]},
            "search_iters": 3, "search_dist": 12}


def stage_case(name, spec):
    n, dim, nq, ncomp, M, L, norm, gl = spec
    w = datagen.make_workload(n, dim, ncomp, n_query=nq, normalize=norm,
                              gt_len_range=gl, device="cpu")
    base = w["base"].astype(np.float16).astype(np.float32)
    queries = w["queries"].astype(np.float16).astype(np.float32)
    gt = datagen.exact_topk_ip(queries, base, 128, device="cpu")
    if gl is not None:
        gt = datagen.ragged_gt(gt, *gl)
    gt_u = gt.view(np.uint32)
    params = {"n_base": n, "dim": dim, "n_query": nq, "gt_k": 128,
              "proj_deg": 16, "proj_pr": 15, "proj_rev": 111, "M": M,
              "link_rev": 126 - M, "Km": L, "Kp": (L + 15) // 16, "Kd": 16,
              "topk": 2 * L, "reset_iter": 15, "rounds": 2,
              "entry_seed": 3030517, "max_iters": 100000}
    st = cpu_ref.OrcStats()
    proj, link = cpu_ref.gpu_stage(params, base, gt_u, stats=st)
    names = sorted(params)
    ids16 = lambda a: np.where(a == TOMB, 0xFFFF, a).astype(np.uint16)
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"),
        base_f16=base.astype(np.float16), gt_u16=ids16(gt_u),
        param_names=np.array(names), param_values=np.array(
            [params[k] for k in names], np.int64),
        proj_u16=ids16(proj), link_u16=ids16(link),
        stat_names=np.array(list(st.as_dict())),
        stat_values=np.array(list(st.as_dict().values()), np.int64))
    print(name, "proj", proj.shape, "link", link.shape, st.as_dict())


def main():
    with open(os.path.join(HERE, "hand_cases.json"), "w") as f:
        json.dump(hand_cases(), f, indent=None, separators=(",", ":"))
    if "--hand-only" in sys.argv:
        return
    for name, spec in STAGE_CASES.items():
        stage_case(name, spec)


if __name__ == "__main__":
    main()
