"""C ABI checks that need no GPU: the library loads, exports every entry
point include/paragraph.h declares, validates parameters, and fails cleanly
(no crash, no CPU fallback) when no device is present."""
import ctypes
import os
import re

import numpy as np
import pytest

from paragraph_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "paragraph.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _native.lib()
    syms = declared_symbols()
    assert len(syms) >= 18
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(_native.EXPORTED) <= set(syms)


def test_params_init_mapping():
    p = _native.make_params(1_000_000, 200, 1_000_000, 32, 128)
    assert (p.M, p.link_rev, p.Km, p.Kd, p.Kp, p.topk) == (32, 94, 128, 16,
                                                            8, 256)
    assert (p.proj_deg, p.proj_pr, p.proj_rev) == (16, 15, 111)
    assert (p.reset_iter, p.table_size, p.rounds, p.entry_seed) == (
        15, 0, 2, 3030517)
    assert p.tune_flags == 0x20  # persistent search grid
    p = _native.make_params(100_000, 200, 100_000, 32, 100)
    assert (p.Kp, p.topk) == (7, 200)
    assert _native.check_params(p)[0] == 0


@pytest.mark.parametrize("field,value,needle", [
    ("n_base", 0, "n_base"),
    ("n_base", 1 << 31, "n_base"),
    ("dim", 0, "dim"),
    ("M", 1, "M must"),
    ("link_rev", 100, "link_rev"),
    ("table_size", 2048, "exceed"),
    ("Kp", 4, "Kp*Kd"),
    ("topk", 1000, "topk"),
    ("Kd", 17, "Kd"),
    ("reset_iter", 0, "reset_iter"),
])
def test_params_check_rejects(field, value, needle):
    p = _native.make_params(1000, 64, 1000, 32, 128)
    setattr(p, field, value)
    rc, msg = _native.check_params(p)
    assert rc == -1 and needle in msg, msg


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    p = _native.make_params(1000, 64, 1000, 32, 64)
    with pytest.raises(_native.PgError):
        _native.Context(p)


def test_invalid_params_rejected_before_device():
    p = _native.make_params(1000, 64, 1000, 32, 64, table_size=900)
    L = _native.lib()
    ctx = L.pg_create(0, ctypes.byref(p))
    assert not ctx
    assert b"invalid params" in L.pg_last_error(None)


def test_stage_names():
    assert _native.stage_name(0) == "match"
    assert _native.stage_name(4) == "search_round1"
    assert _native.stage_name(11) == "merge_round2"


def test_autotune_candidates_valid():
    """Every knob set the autotuner tries passes the parameter check and
    changes only perf knobs."""
    from paragraph_amd import autotune
    for prec in (_native.PG_FP32, _native.PG_BF16, _native.PG_FP8E4M3):
        p = _native.make_params(100000, 200, 100000, 32, 128, precision=prec)
        cands = autotune.candidates(p)
        assert len(cands) >= 4
        for kn in cands:
            pp = autotune.apply(p, {**kn, "tune_flags": kn["tune_flags"]})
            assert _native.check_params(pp)[0] == 0, kn
            for f in ("M", "Km", "Kp", "Kd", "topk", "rerank", "precision",
                      "rounds", "reset_iter", "entry_seed"):
                assert getattr(pp, f) == getattr(p, f)
        pp = autotune.apply(p, {"tune_flags": 0xA4, "table_size": 0})
        assert pp.tune_flags == 0xA4
        # the prune knob (bit 8) is replaced, not or-ed in
        p.tune_flags = 0x120
        pp = autotune.apply(p, {"tune_flags": 0x21, "table_size": 0})
        assert pp.tune_flags == 0x21


def test_autotune_whole_build_candidates():
    """The knob sets the tuner times on whole builds: both queue orders
    with every visited-table size it knows for the precision (the sampled
    best included), the chosen gather batch and prune variant kept, only
    perf knobs changed, every set valid."""
    from paragraph_amd import autotune
    for prec, gather in ((_native.PG_FP32, 4), (_native.PG_BF16, 1),
                         (_native.PG_FP8E4M3, 0x801)):
        p = _native.make_params(100000, 200, 100000, 32, 48, precision=prec)
        sampled = {kn["table_size"] for kn in autotune.candidates(p)}
        for best_table in sorted(sampled):
            sets = autotune.build_candidates(p, gather, 0x80, best_table)
            got = {(pp.tune_flags & 0x220, pp.table_size) for pp in sets}
            assert got == {(g, t) for g in (0x20, 0x220) for t in sampled}
            for pp in sets:
                assert _native.check_params(pp)[0] == 0
                assert pp.tune_flags & 0x81F == gather
                assert pp.tune_flags & 0x80
                for f in ("M", "Km", "Kp", "Kd", "topk", "rerank",
                          "precision", "rounds", "reset_iter", "entry_seed"):
                    assert getattr(pp, f) == getattr(p, f)
        # FP32 keeps the exact visited set: one (automatic) table size
        if prec == _native.PG_FP32:
            assert sampled == {0}


def test_builder_host_stages_choice():
    """BuildIndex runs the host stages on the GPU only when the gt has one
    row per base row; "auto" falls back to the host stages otherwise."""
    from paragraph_amd.builder import resolve_host_stages
    assert resolve_host_stages("auto", 100, 100) == "gpu"
    assert resolve_host_stages("auto", 100, 60) == "host"
    assert resolve_host_stages("host", 100, 100) == "host"
    with pytest.raises(ValueError):
        resolve_host_stages("gpu", 100, 60)
    with pytest.raises(ValueError):
        resolve_host_stages("cpu", 100, 100)


def test_builder_rejects_unknown_precision():
    from paragraph_amd import Builder
    b = Builder(np.zeros((4, 8), np.float32), np.zeros((4, 128), np.int32))
    with pytest.raises(ValueError):
        b.LaunchBuild(M=16, L=16, precision="fp16")


def test_beam_acceptance_rule():
    """beam_accepts: paired per-query differences, lower confidence bound
    against 0.5 pt (or 5% of a small recall)."""
    from paragraph_amd import autotune
    rng = np.random.default_rng(3)
    ref = rng.integers(0, 11, 40000)
    # identical hits: delta 0, bound 0, accepted
    ok, d, lo, tol = autotune.beam_accepts(ref, ref)
    assert ok and d == 0.0 and lo == 0.0 and abs(tol - 0.005) < 1e-12
    # one hit lost on 2% of the queries: -0.2 pt, tight bound, accepted
    h = ref.copy()
    idx = np.flatnonzero(h > 0)[:800]
    h[idx] -= 1
    ok, d, lo, _ = autotune.beam_accepts(h, ref)
    assert ok and abs(d - (-800 / 40000 / 10)) < 1e-12 and -0.005 < lo < d
    # one hit lost on 6%: -0.6 pt, rejected
    h = ref.copy()
    idx = np.flatnonzero(h > 0)[:2400]
    h[idx] -= 1
    assert not autotune.beam_accepts(h, ref)[0]
    # a -0.45 pt mean on few queries is rejected by its wide bound, and
    # accepted on many queries with the same per-query spread
    small = np.array([10] * 1000)
    hs = small.copy()
    hs[:9] -= 5
    assert not autotune.beam_accepts(hs, small)[0]
    big = np.tile(small, 400)
    hb = np.tile(hs, 400)
    assert autotune.beam_accepts(hb, big)[0]
    # small recall: 5% relative cap
    ref = np.zeros(1000, np.int64)
    ref[:66] = 1
    assert abs(autotune.beam_accepts(ref, ref)[3] - 0.05 * 0.0066) < 1e-12
    for L in (128, 96, 64):
        p = _native.make_params(1000, 64, 1000, 32, 128,
                                precision=_native.PG_BF16)
        pp = autotune.beam_params(p, L)
        assert (pp.Km, pp.Kp, pp.topk) == (L, (L + 15) // 16, 2 * L)
        assert _native.check_params(pp)[0] == 0
