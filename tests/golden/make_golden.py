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
    return cases


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
    for name, spec in STAGE_CASES.items():
        stage_case(name, spec)


if __name__ == "__main__":
    main()
