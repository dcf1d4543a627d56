"""GPU tests of the distance evaluators, through pg_run_pair_distances.

FP32: the search evaluator and the prune chain must equal the oracle's
search_dist / chain_dist bit for bit.  BF16 / FP8: the MFMA result must equal
the FP64 dot product of the rounded operands (bf16 / per-vector-scaled
e4m3fn, rounded on the host by paragraph_amd.numerics, which is pinned
against torch's conversions) within the accumulation bound: BF16,
gamma(dpad) * sum |a_i b_i| (FP32 accumulation); FP8, 2^-10 * sum |a_i b_i|
for the instruction's truncating 8-term group sums plus the same gamma term
and two FP32 roundings of the scale products.  A wrong lane layout or
scale fails these bounds by orders of magnitude, a wrong operand rounding
mode by a few times (systematic half-ulp errors over dim terms).
"""
import numpy as np
import pytest

from oracle import cpu_ref
from paragraph_amd import _native, numerics

pytestmark = pytest.mark.gpu

TOMB = 0xFFFFFFFF


def _case(dim, n=300, nq=7, k=100, seed=0):
    rng = np.random.default_rng(seed + dim)
    base = rng.standard_normal((n, dim)).astype(np.float32)
    base[5] = 0.0  # a zero row (FP8 scale 1)
    q = rng.standard_normal((nq, dim)).astype(np.float32)
    ids = rng.integers(0, n, size=(nq, k)).astype(np.uint32)
    ids[0, 3] = TOMB
    ids[1, 7] = n + 5          # out of range -> NaN
    ids[2, :] = 5
    return base, q, ids


def _run(p, base, q, ids, chain, torch):
    out = torch.empty(ids.shape, dtype=torch.float32, device="cuda")
    with _native.Context(p) as ctx:
        ctx.run_pair_distances(torch.from_numpy(q).cuda(),
                               torch.from_numpy(base).cuda(),
                               torch.from_numpy(ids.view(np.int32)).cuda(),
                               out, chain=chain)
        ctx.sync()
    return out.cpu().numpy()


def _params(n, dim, prec):
    return _native.make_params(n, dim, n, 32, 64, precision=prec)


@pytest.mark.parametrize("dim", [37, 200, 512])
def test_fp32_evaluators_bit_exact(dim, cuda):
    import torch
    base, q, ids = _case(dim)
    n = base.shape[0]
    p = _params(n, dim, _native.PG_FP32)
    got_s = _run(p, base, q, ids, False, torch)
    got_c = _run(p, base, q, ids, True, torch)
    for i in range(q.shape[0]):
        for t in range(ids.shape[1]):
            v = int(ids[i, t])
            if v >= n:
                assert np.isnan(got_s[i, t]) and np.isnan(got_c[i, t])
                continue
            es = np.float32(cpu_ref.search_dist(q[i], base[v]))
            ec = np.float32(cpu_ref.chain_dist(q[i], base[v]))
            assert got_s[i, t].view(np.uint32) == es.view(np.uint32), (i, t)
            assert got_c[i, t].view(np.uint32) == ec.view(np.uint32), (i, t)


def _gamma(nterms):
    u = 2.0 ** -23
    return nterms * u / (1 - nterms * u)


@pytest.mark.parametrize("dim", [64, 200, 512])
def test_bf16_evaluator_matches_rounded_operands(dim, cuda):
    import torch
    base, q, ids = _case(dim, seed=1)
    n = base.shape[0]
    p = _params(n, dim, _native.PG_BF16)
    got = _run(p, base, q, ids, False, torch)
    dpad = (dim + 31) // 32 * 32
    bb = numerics.bf16_round(base).astype(np.float64)
    qb = numerics.bf16_round(q).astype(np.float64)
    for i in range(q.shape[0]):
        ok = ids[i] < n
        assert np.all(np.isnan(got[i, ~ok]))
        rows = bb[ids[i, ok]]
        exact = -(rows @ qb[i])
        absum = np.abs(rows) @ np.abs(qb[i])
        err = np.abs(got[i, ok].astype(np.float64) - exact)
        assert np.all(err <= _gamma(dpad) * absum * 1.01 + 1e-30), (
            i, float(np.max(err / np.maximum(absum, 1e-30))))


@pytest.mark.parametrize("dim", [64, 200, 512])
def test_fp8_evaluator_matches_rounded_operands(dim, cuda):
    import torch
    base, q, ids = _case(dim, seed=2)
    n = base.shape[0]
    p = _params(n, dim, _native.PG_FP8E4M3)
    got = _run(p, base, q, ids, False, torch)
    dpad = (dim + 31) // 32 * 32
    bq, bs = numerics.fp8_quantize(base)
    qq, qs = numerics.fp8_quantize(q)
    for i in range(q.shape[0]):
        ok = ids[i] < n
        assert np.all(np.isnan(got[i, ~ok]))
        rows = bq[ids[i, ok]]
        sc = qs[i].astype(np.float64) * bs[ids[i, ok]].astype(np.float64)
        exact = -(rows @ qq[i]) * sc
        absum = (np.abs(rows) @ np.abs(qq[i])) * sc
        # v_mfma_f32_16x16x32_fp8_fp8 sums each group of 8 consecutive k
        # aligned to the group's largest product, dropping bits below
        # 2^(e_max - 13) (measured on gfx950): at most 7 truncated terms of
        # < 2^-13 max|p| each, i.e. <= 2^-10 of the group's sum |p|; group
        # sums and chunks then add in FP32.
        tol = (2.0 ** -10 + _gamma(dpad) * 1.01 + 3 * 2.0 ** -24) * absum \
            + 1e-30
        err = np.abs(got[i, ok].astype(np.float64) - exact)
        assert np.all(err <= tol), (
            i, float(np.max(err / np.maximum(absum, 1e-30))))


def test_pair_distances_rejects_wide_rows(cuda):
    import torch
    p = _params(100, 16, _native.PG_FP32)
    q = torch.zeros((1, 16), dtype=torch.float32, device=cuda)
    ids = torch.zeros((1, 1025), dtype=torch.int32, device=cuda)
    out = torch.empty((1, 1025), dtype=torch.float32, device=cuda)
    with _native.Context(p) as ctx:
        with pytest.raises(_native.PgError):
            ctx.run_pair_distances(q, q, ids, out)
