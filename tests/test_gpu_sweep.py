"""GPU tests of the kept-major link-row prune (the default) against the
candidate-major prune kernel (tune_flags bit 8) and the oracle.

The rows hit every rule the sweep must keep: unsorted rows (the first
candidate is often not the smallest key, so the first sweep is repeated
against the right entry), self entries before and after the first stop id,
duplicates, INT_MAX / tomb / out-of-range ids, rows without a stop id, exact
distance ties (integer-valued base), and caps from 2 to 128.  Sorted rows
(exact inner-product order, like the search's) take the speculative path.
The exact FP32 sweep must also reproduce the candidate-major pair counter.
"""
import numpy as np
import pytest

from oracle import cpu_ref
from paragraph_amd import _native, datagen

pytestmark = pytest.mark.gpu

TOMB = 0xFFFFFFFF
SCAN = 0x100
SPLIT = 0x80
ORDER = 0x200  # locality order: the sweep takes its nodes in that order


def _dev(a, torch):
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def _host(t):
    a = t.cpu().numpy()
    return a.view(np.uint32) if a.dtype == np.int32 else a


def _base(kind, n, dim, seed=3):
    if kind == "integer":
        rng = np.random.default_rng(seed)
        protos = rng.integers(-2, 3, (n // 8, dim)).astype(np.float32)
        return np.ascontiguousarray(protos[rng.integers(0, n // 8, n)])
    return datagen.base_vectors(n, dim, 16, normalize=(kind == "normal"))


def _rows(base, D, seed, order):
    n = base.shape[0]
    rng = np.random.default_rng(seed)
    if order == "sorted":
        ip = base.astype(np.float64) @ base.astype(np.float64).T
        rows = np.argsort(-ip, axis=1, kind="stable")[:, :D].astype(np.uint32)
    else:
        rows = rng.integers(0, n, (n, D)).astype(np.uint32)
    u = rng.random(rows.shape)
    rows[u < 0.02] = 0x7FFFFFFF
    rows[(u >= 0.02) & (u < 0.03)] = TOMB
    rows[(u >= 0.03) & (u < 0.035)] = n + 7
    selfm = (u >= 0.035) & (u < 0.06)
    ids = np.broadcast_to(np.arange(n, dtype=np.uint32)[:, None], rows.shape)
    rows[selfm] = ids[selfm]
    # a third of the rows carry no stop id at all
    clean = rng.random(n) < 0.33
    cr = rows[clean]
    cr[cr >= n] = 0
    rows[clean] = cr
    return np.ascontiguousarray(rows)


def _prune(base_d, rows_d, n, dim, mp, prec, flags, torch):
    p = _native.make_params(n, dim, n, 32, 128, precision=prec,
                            tune_flags=flags)
    pr_d = torch.full((n, mp), -7, dtype=torch.int32, device="cuda")
    pc_d = torch.zeros(n, dtype=torch.int32, device="cuda")
    with _native.Context(p) as ctx:
        ctx.run_prune_rows(base_d, rows_d, mp, pr_d, pc_d)
        ctx.sync()
        st = ctx.stats().as_dict()
    return _host(pr_d), _host(pc_d), st


CASES = [
    # data, dim, D, mp, row order
    ("gauss", 200, 256, 32, "sorted"),
    ("gauss", 200, 256, 2, "random"),
    ("gauss", 200, 256, 128, "random"),
    ("integer", 200, 256, 8, "random"),
    ("integer", 64, 256, 32, "sorted"),
    ("normal", 512, 256, 64, "sorted"),
    ("normal", 96, 300, 32, "random"),
    ("gauss", 37, 200, 16, "random"),
]


@pytest.mark.parametrize("prec,flags", [
    (_native.PG_FP32, 0x20), (_native.PG_BF16, 0x20),
    (_native.PG_BF16, 0x20 | SPLIT), (_native.PG_FP8E4M3, 0x20),
    (_native.PG_FP32, 0x20 | ORDER), (_native.PG_BF16, 0x20 | ORDER)],
    ids=["fp32", "bf16", "bf16-split", "fp8", "fp32-order", "bf16-order"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
def test_sweep_equals_scan_and_oracle(case, prec, flags, cuda):
    import torch
    kind, dim, D, mp, order = case
    n = 1024
    base = _base(kind, n, dim)
    rows = _rows(base, D, seed=dim + mp, order=order)
    print(f"[sweep] {kind} dim={dim} D={D} mp={mp} {order} prec={prec} "
          f"flags={flags:#x}", flush=True)
    base_d, rows_d = _dev(base, torch), _dev(rows, torch)
    pr_s, pc_s, st_s = _prune(base_d, rows_d, n, dim, mp, prec, flags, torch)
    pr_c, pc_c, st_c = _prune(base_d, rows_d, n, dim, mp, prec,
                              flags | SCAN, torch)
    st_o = cpu_ref.OrcStats()
    pr_o, pc_o = cpu_ref.prune(base, rows, mp, stats=st_o)
    mask = np.arange(mp)[None, :] < pc_o[:, None]
    np.testing.assert_array_equal(pc_s, pc_o)
    np.testing.assert_array_equal(pr_s[mask], pr_o[mask])
    np.testing.assert_array_equal(pc_c, pc_o)
    np.testing.assert_array_equal(pr_c[mask], pr_o[mask])
    # rows past the count are untouched
    assert (pr_s[~mask] == np.uint32(0xFFFFFFF9)).all()
    assert st_s["prune_dist"] == st_c["prune_dist"] == st_o.as_dict()[
        "prune_dist"]
    if prec == _native.PG_FP32:
        assert st_s["prune_pairs"] == st_c["prune_pairs"] == st_o.as_dict()[
            "prune_pairs"]

