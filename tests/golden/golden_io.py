"""Loads the golden fixtures written by make_golden.py (no pickle)."""
import glob
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TOMB = 0xFFFFFFFF


def _ids32(a16):
    a = a16.astype(np.uint32)
    a[a16 == 0xFFFF] = TOMB
    return a


def stage_names():
    return sorted(os.path.splitext(os.path.basename(p))[0]
                  for p in glob.glob(os.path.join(HERE, "stage_*.npz")))


def load_stage(name):
    """dict(base f32, gt u32, params dict, proj u32, link u32, stats dict)."""
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        params = {str(k): int(v) for k, v in zip(z["param_names"],
                                                 z["param_values"])}
        stats = {str(k): int(v) for k, v in zip(z["stat_names"],
                                                z["stat_values"])}
        return {"base": z["base_f16"].astype(np.float32),
                "gt": _ids32(z["gt_u16"]), "params": params,
                "proj": _ids32(z["proj_u16"]), "link": _ids32(z["link_u16"]),
                "stats": stats}


def load_hand():
    with open(os.path.join(HERE, "hand_cases.json")) as f:
        return json.load(f)
