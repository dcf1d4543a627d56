"""Host stages around the GPU stage: the entry point, MatchSup, FusionNN,
FusionFinal and the index file (reference: tests/test_t2iend.cu:31-633).

Pinning of the oracle (oracle/cpu_fusion.cpp), which has no reference
binary or golden vectors to check against:
  - `compare`: an exact-rational emulation of its summation order (16 FMA
    lanes, folds to 8 and 4, masked remainder, horizontal adds), and inputs
    on which that order or the remainder mask (rule C8) changes the float;
  - RNGPrune: hand-derived cases (strict, non-strict fill from position 1,
    the node itself at the front and inside the list);
  - the entry point: a hand-counted scan (rule C4);
  - MatchSup / FusionNN / FusionFinal: an independent Python restatement on
    small-integer data, where every distance is exact in FP32 in any order,
    covering the first-come match, the reverse cap (rule C6), the merge
    order (rule C7), and FusionFinal's unsigned room computation.
Then the product (libparagraph_host.so) equals the oracle id for id on
float workloads, with any thread count, and writes the reference's index
layout.
"""
import ctypes
from collections import defaultdict
from fractions import Fraction

import numpy as np
import pytest

from oracle import cpu_ref
from paragraph_amd import datagen, host_stage, io

from test_oracle import fma32, round_f32

TOMB = 0xFFFFFFFF


# --------------------------------------------------- compare (distances)

def add32(a, b):
    return round_f32(Fraction(float(a)) + Fraction(float(b)))


def compare_emulated(a, b):
    """The summation order of tests/test_t2iend.cu:31-78, exactly rounded
    step by step; remainder lanes 1..r with the element past the end as 0."""
    n = len(a)
    acc = [np.float32(0.0)] * 16
    i = 0
    while n - i >= 16:
        acc = [fma32(a[i + l], b[i + l], acc[l]) for l in range(16)]
        i += 16
    for w in (8, 4):
        acc = [add32(acc[l + w], acc[l]) for l in range(w)]
        if n - i >= w:
            acc = [fma32(a[i + l], b[i + l], acc[l]) for l in range(w)]
            i += w
    r = n - i
    if r:
        for l in range(4):
            on = 1 <= l <= r and i + l < n
            acc[l] = fma32(a[i + l] if on else 0.0, b[i + l] if on else 0.0,
                           acc[l])
    return np.float32(-add32(add32(acc[0], acc[1]), add32(acc[2], acc[3])))


@pytest.mark.parametrize("dim", [1, 3, 4, 5, 7, 8, 12, 15, 16, 17, 24, 31, 40,
                                 200, 203])
def test_compare_matches_exact_emulation(dim):
    rng = np.random.default_rng(dim)
    for _ in range(4):
        a = rng.standard_normal(dim).astype(np.float32)
        b = rng.standard_normal(dim).astype(np.float32)
        assert cpu_ref.cpu_compare(a, b) == compare_emulated(a, b)


def test_compare_order_differs_from_a_sequential_chain():
    # the 16-lane order gives a different float from the prune chain on
    # some inputs: the emulation is not a restatement of the chain
    rng = np.random.default_rng(3)
    diff = 0
    for _ in range(200):
        a = rng.standard_normal(200).astype(np.float32)
        b = rng.standard_normal(200).astype(np.float32)
        diff += cpu_ref.cpu_compare(a, b) != cpu_ref.chain_dist(a, b)
    assert diff > 20


def test_compare_remainder_mask():
    # dim 5: blocks of 4 + remainder 1, whose mask takes element 5 (past
    # the end, 0 here) and never element 4; dim 6 takes element 5 only
    ones = np.ones(6, np.float32)
    a = np.array([1, 1, 1, 1, 100, 10], np.float32)
    assert cpu_ref.cpu_compare(a[:5], ones[:5]) == -4.0
    assert cpu_ref.cpu_compare(a, ones) == -14.0
    # dim 8: no remainder, every element counts
    a8 = np.arange(1, 9, dtype=np.float32)
    assert cpu_ref.cpu_compare(a8, np.ones(8, np.float32)) == -36.0


# ------------------------------------------------------------ RNGPrune

def test_rng_prune_hand_cases():
    # IP distance -<x, y>; node 0.  Candidates by distance: 1 (-3),
    # 2 (-2, occluded by 1: -<2,1> = -6 < -2), 3 (0, kept: -<3,1> = 0 is not
    # < 0), 4 (1: -<4,1> = 3 is not < 1, -<4,3> = -0.5 < 1, occluded)
    # (dim 4: rows padded with zeros, so that the remainder mask of rule C8
    # does not apply)
    X = np.array([[1, 0, 0, 0], [3, 0, 0, 0], [2, 2, 0, 0], [0, 1, 0, 0],
                  [-1, 0.5, 0, 0]], np.float32)
    ids = np.array([4, 3, 2, 1], np.uint32)
    d = np.array([cpu_ref.cpu_compare(X[0], X[i]) for i in ids], np.float32)
    assert cpu_ref.rng_prune(X, ids, d, 0, 8, True).tolist() == [1, 3]
    # non-strict: filled from sorted position 1 on (2, then 4)
    assert cpu_ref.rng_prune(X, ids, d, 0, 8, False).tolist() == [1, 3, 2, 4]
    assert cpu_ref.rng_prune(X, ids, d, 0, 3, False).tolist() == [1, 3, 2]
    assert cpu_ref.rng_prune(X, ids, d, 0, 1, True).tolist() == [1]
    # the node itself first in the list is skipped; later it is never kept
    ids2 = np.array([0, 1, 3], np.uint32)
    d2 = np.array([-10.0, -3.0, 0.0], np.float32)
    assert cpu_ref.rng_prune(X, ids2, d2, 0, 8, True).tolist() == [1, 3]
    d3 = np.array([-1.0, -3.0, 0.5], np.float32)
    assert cpu_ref.rng_prune(X, ids2, d3, 0, 8, False).tolist() == [1, 3]
    # only the node: nothing kept
    assert cpu_ref.rng_prune(X, ids2[:1], d2[:1], 0, 8, False).tolist() == []


# ----------------------------------------------------------- entry point

def test_entry_point_hand_count():
    N = 6
    gt = np.full((5, 4), TOMB, np.uint32)
    gt[0, :1] = [2]           # 2:1            takes 2
    gt[1, :2] = [2, 3]        # 2:2 3:1        takes 3
    gt[2, :3] = [2, 3, 1]     # 2:3 3:2 1:1    takes 1
    gt[3, :4] = [3, 2, 1, 4]  # 3:3 2:4 1:2 4:1 takes 4
    gt[4, :2] = [9, 3]        # id >= N ends the row
    assert cpu_ref.entry_point(gt, N) == 2
    p = host_stage.make_params(N, 4, 5, gt_k=4, rounds=2)
    assert host_stage.entry_point(p, gt) == 2
    # a tie goes to the smallest id (rule C4)
    gt2 = np.array([[4, TOMB], [1, TOMB]], np.uint32)
    assert cpu_ref.entry_point(gt2, N) == 1
    assert host_stage.entry_point(
        host_stage.make_params(N, 4, 2, gt_k=2, rounds=1), gt2) == 1


# ------------------------------------- independent restatement (exact data)

def py_dist(X, a, b):
    return -int(np.dot(X[a].astype(np.int64), X[b].astype(np.int64)))


def py_rng(X, full, target, M, strict):
    N = len(X)
    res = []
    s = 0
    while s < len(full) and full[s][1] == target:
        s += 1
    if s == len(full):
        return res
    res.append(full[s][1])
    for d, c in full[s + 1:]:
        if len(res) >= M:
            break
        if c == target or any(c == k or py_dist(X, c, k) < d for k in res):
            continue
        res.append(c)
    if not strict:
        for d, c in full[1:]:
            if len(res) >= M:
                break
            if c not in res and c < N and c != target:
                res.append(c)
    return res


def py_match_sup(X, gt, topn, M):
    N, (Q, K) = len(X), gt.shape
    match, taken = {}, set()
    for q in range(Q):
        nn = int(gt[q, topn])
        if nn < N and nn not in taken:
            taken.add(nn)
            match[q] = nn
    fwd = {}
    for q, m in match.items():
        seen, full = {m}, []
        for j in range(K):
            i = int(gt[q, j])
            if i >= N:
                break
            if i in seen:
                continue
            seen.add(i)
            full.append((py_dist(X, i, m), i))
        fwd[q] = py_rng(X, sorted(full), m, M, False)
    rev = defaultdict(list)
    py_match_sup.capped = 0
    for q in sorted(match):
        for des in fwd[q]:
            if len(rev[des]) < 128 - M:  # C6: lowest query ids
                rev[des].append(match[q])
            else:
                py_match_sup.capped += 1
    graph = {match[q]: fwd[q] for q in match}
    out = []
    for b in range(N):
        u = sorted(set(graph.get(b, [])) | set(rev[b]))  # C7
        if len(u) > M:
            u = py_rng(X, sorted((py_dist(X, i, b), i) for i in u), b, M,
                       True)
        out.append(u)
    return out


def py_bipartite(X, gt, M, R):
    supply = [py_match_sup(X, gt, t, M) for t in range(1, R + 1)]
    out = []
    for b in range(len(X)):
        seen, full = set(), []
        for g in supply:
            for i in g[b]:
                if i in seen or i == b or i >= len(X):
                    continue
                seen.add(i)
                full.append((py_dist(X, b, i), i))
        out.append(py_rng(X, sorted(full), b, 2 * M, True))
    return out


def py_fusion_final(X, bip, link, proj, Mf):
    N = len(X)
    out = []
    for b in range(N):
        seen, full, fin = set(), [], []
        for i in bip[b]:
            if i in seen or i == b or i >= N:
                continue
            seen.add(i)
            full.append((py_dist(X, b, i), i))
            fin.append(i)
        for i in map(int, link[b]):
            if i < N and i not in seen:  # the node itself is not excluded
                seen.add(i)
                full.append((py_dist(X, b, i), i))
        for i in map(int, proj[b]):
            if i < N and i not in seen and i != b:
                seen.add(i)
                full.append((py_dist(X, b, i), i))
        pruned = py_rng(X, sorted(full), b, Mf, True)
        room = Mf - len(fin) if len(fin) <= Mf else float("inf")
        extra = [c for c in pruned if c not in fin][:max(room, 0)
                                                    if room != float("inf")
                                                    else None]
        out.append(fin + extra)
    return out


def csr_lists(off, ids):
    return [ids[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)]


def exact_case(seed, N=90, dim=8, K=24):
    rng = np.random.default_rng(seed)
    X = rng.integers(-3, 4, (N, dim)).astype(np.float32)
    gt = np.full((N, K), TOMB, np.uint32)
    for q in range(N):
        L = int(rng.integers(K // 2, K + 1))
        row = rng.permutation(N)[:L].astype(np.uint32)
        if L > 3:
            row[int(rng.integers(1, L))] = row[0]  # a duplicate id
        if rng.random() < 0.1:
            row[int(rng.integers(0, L))] = N + 5   # an id >= N ends the row
        gt[q, :L] = row
    # hubs: many queries' chosen columns point at a few bases
    gt[::3, 1] = 7
    gt[1::5, 2] = 11
    return X, gt


@pytest.mark.parametrize("seed,M", [(0, 6), (1, 6), (2, 124), (3, 124),
                                    (4, 3)])
def test_match_sup_and_bipartite_match_restatement(seed, M):
    X, gt = exact_case(seed)
    capped = 0
    for topn in (1, 2, 3):
        off, ids = cpu_ref.match_sup(X, gt, topn, M)
        assert csr_lists(off, ids) == py_match_sup(X, gt, topn, M), topn
        capped += py_match_sup.capped
    if M > 100:  # reverse cap 128 - M: the overflow rule is exercised
        assert capped > 0
    R = 3
    off, ids = cpu_ref.bipartite(X, gt, M, R)
    assert csr_lists(off, ids) == py_bipartite(X, gt, M, R)


@pytest.mark.parametrize("seed,M,Mf", [(5, 6, 5), (6, 6, 70), (7, 3, 2)])
def test_fusion_final_matches_restatement(seed, M, Mf):
    X, gt = exact_case(seed)
    N = len(X)
    rng = np.random.default_rng(seed + 100)
    link = rng.integers(0, N, (N, 8)).astype(np.uint32)
    link[:, 6:] = TOMB
    link[::4, 0] = np.arange(0, N, 4)  # the node itself in its link row
    proj = rng.integers(0, N, (N, 4)).astype(np.uint32)
    bo, bi = cpu_ref.bipartite(X, gt, M, 3)
    bip = csr_lists(bo, bi)
    off, ids = cpu_ref.fusion_final(X, bo, bi, link, proj, Mf)
    got = csr_lists(off, ids)
    assert got == py_fusion_final(X, bip, link, proj, Mf)
    # the unsigned room: a bipartite list longer than M_final keeps growing
    long_rows = [b for b in range(N) if len(bip[b]) > Mf]
    if Mf < 2 * M:
        assert long_rows and any(len(got[b]) > len(bip[b]) for b in long_rows)


# ------------------------------------------------ product against oracle

@pytest.mark.parametrize("n,dim,nc,norm", [(1500, 64, 16, False),
                                           (1200, 200, 64, False),
                                           (900, 37, 8, True),
                                           (1000, 96, 32, True)])
def test_product_equals_oracle(n, dim, nc, norm):
    w = datagen.make_workload(n, dim, nc, normalize=norm, device="cpu")
    base = w["base"]
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    p = host_stage.make_params(n, dim, gt.shape[0], gt_k=gt.shape[1])
    assert host_stage.entry_point(p, gt) == cpu_ref.entry_point(gt, n)
    for topn in (1, 4):
        a = host_stage.match_sup(p, base, gt, topn)
        b = cpu_ref.match_sup(base, gt, topn, p.M_nn)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
    bo, bi = host_stage.bipartite(p, base, gt)
    oo, oi = cpu_ref.bipartite(base, gt, p.M_nn, p.rounds)
    np.testing.assert_array_equal(bo, oo)
    np.testing.assert_array_equal(bi, oi)
    rng = np.random.default_rng(n)
    link = rng.integers(0, n, (n, 32)).astype(np.uint32)
    link[:, 24:] = TOMB
    link[::7, 3] = np.arange(0, n, 7)
    proj = rng.integers(0, n, (n, 16)).astype(np.uint32)
    fo, fi = host_stage.fusion_final(p, base, bo, bi, link, proj)
    go, gi = cpu_ref.fusion_final(base, oo, oi, link, proj, p.M_final)
    np.testing.assert_array_equal(fo, go)
    np.testing.assert_array_equal(fi, gi)


def test_product_equals_oracle_ragged_zipf_gt():
    """The skewed workload's shape: Zipf query components, per-query gt
    lengths in [16, 128] padded with -1 (MatchSup's rows end at the first
    id >= N), L2-normalised dim 512."""
    n, dim = 1200, 512
    w = datagen.make_workload(n, dim, 64, normalize=True, zipf=1.1,
                              gt_len_range=(16, 128), device="cpu")
    base = w["base"]
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    assert (gt == TOMB).any()
    p = host_stage.make_params(n, dim, n, gt_k=gt.shape[1])
    bo, bi = host_stage.bipartite(p, base, gt)
    oo, oi = cpu_ref.bipartite(base, gt, p.M_nn, p.rounds)
    np.testing.assert_array_equal(bo, oo)
    np.testing.assert_array_equal(bi, oi)
    rng = np.random.default_rng(9)
    link = rng.integers(0, n, (n, 48)).astype(np.uint32)
    proj = rng.integers(0, n, (n, 16)).astype(np.uint32)
    fo, fi = host_stage.fusion_final(p, base, bo, bi, link, proj)
    go, gi = cpu_ref.fusion_final(base, oo, oi, link, proj, p.M_final)
    np.testing.assert_array_equal(fo, go)
    np.testing.assert_array_equal(fi, gi)


def test_product_small_degrees_and_threads():
    """M_nn / M_final small enough that the reverse cap, the strict merge
    and FusionFinal's unsigned room all occur; 1 and 8 threads agree."""
    n, dim = 800, 48
    w = datagen.make_workload(n, dim, 8, normalize=True, device="cpu")
    base = w["base"]
    gt = np.ascontiguousarray(w["gt"].view(np.uint32))
    outs = []
    for threads in (1, 8):
        p = host_stage.make_params(n, dim, n, gt_k=gt.shape[1], M_nn=100,
                                   M_final=12, rounds=3, threads=threads)
        bo, bi = host_stage.bipartite(p, base, gt)
        proj = gt[:, :16]
        link = gt[:, 16:48]
        fo, fi = host_stage.fusion_final(p, base, bo, bi, link, proj)
        outs.append((bo, bi, fo, fi))
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)
    oo, oi = cpu_ref.bipartite(base, gt, 100, 3)
    np.testing.assert_array_equal(outs[0][0], oo)
    np.testing.assert_array_equal(outs[0][1], oi)
    go, gi = cpu_ref.fusion_final(base, oo, oi, gt[:, 16:48], gt[:, :16], 12)
    np.testing.assert_array_equal(outs[0][2], go)
    np.testing.assert_array_equal(outs[0][3], gi)
    assert np.diff(oo.astype(np.int64)).max() > 12  # unsigned room occurred


def test_index_file_layout(tmp_path):
    off = np.array([0, 2, 2, 5], np.uint32)
    ids = np.array([1, 2, 0, 1, 2], np.uint32)
    path = tmp_path / "index.bin"
    host_stage.save_index(str(path), 1, off, ids)
    raw = np.fromfile(path, np.uint32)
    assert raw.tolist() == [1, 3, 2, 1, 2, 0, 3, 0, 1, 2]
    ep, lists = io.load_index(str(path))
    assert ep == 1 and [l.tolist() for l in lists] == [[1, 2], [], [0, 1, 2]]


def test_bin_reader_headered_headerless_and_chunks(tmp_path):
    """The C++ .fbin / .ibin reader against numpy: whole files, row ranges
    (streaming chunks), headerless arrays, and its size checks."""
    rng = np.random.default_rng(14)
    a = rng.standard_normal((37, 5)).astype(np.float32)
    g = rng.integers(-1, 1000, (37, 9)).astype(np.int32)
    fa, fg, raw = (str(tmp_path / x) for x in ("a.fbin", "g.ibin", "a.raw"))
    io.write_bin(fa, a)
    io.write_bin(fg, g)
    io.write_bin(raw, a, headered=False)
    assert host_stage.bin_shape(fa) == (37, 5)
    np.testing.assert_array_equal(host_stage.read_bin(fa, np.float32), a)
    np.testing.assert_array_equal(host_stage.read_bin(fg, np.int32), g)
    parts = [host_stage.read_bin(fa, np.float32, row0=r, nrows=min(8, 37 - r))
             for r in range(0, 37, 8)]
    np.testing.assert_array_equal(np.concatenate(parts), a)
    np.testing.assert_array_equal(
        host_stage.read_bin(raw, np.float32, headered=False, cols=5,
                            row0=30, nrows=7), a[30:])
    np.testing.assert_array_equal(io.read_bin(fa, np.float32), a)
    with pytest.raises(host_stage.HostError, match="beyond"):
        host_stage.read_bin(fa, np.float32, row0=30, nrows=8)
    with pytest.raises(host_stage.HostError, match="columns"):
        host_stage.read_bin(fa, np.float32, cols=4, nrows=1)
    for width in (128, 9, 4):
        np.testing.assert_array_equal(host_stage.pad_gt(g, width),
                                      io.pad_gt(g, width))


def test_build_tool_usage():
    """pg_build_index (the C++ front end's command-line build) rejects a
    bad command line before touching a GPU."""
    import os
    import subprocess
    from paragraph_amd import build
    if not os.path.exists(build.TOOL):
        pytest.skip("pg_build_index is not built")
    r = subprocess.run([build.TOOL, "only-one-arg"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_host_library_exports_and_errors():
    L = host_stage.lib()
    for name in host_stage.EXPORTED:
        assert hasattr(L, name), name
    p = host_stage.make_params(10, 4, 10, gt_k=4, M_nn=128)
    base = np.zeros((10, 4), np.float32)
    gt = np.zeros((10, 4), np.uint32)
    with pytest.raises(host_stage.HostError, match="M_nn"):
        host_stage.bipartite(p, base, gt)
    p = host_stage.make_params(10, 4, 10, gt_k=4, rounds=4)
    with pytest.raises(host_stage.HostError, match="rounds"):
        host_stage.bipartite(p, base, gt)
    assert L.pg_host_entry_point(None, None, None) != 0


def test_cpp_builder_header_compiles(tmp_path):
    """include/paragraph/Builder.hpp with BuildIndex / SaveIndex
    instantiated compiles against both C ABI headers."""
    import os
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "use_builder.cpp"
    src.write_text(
        '#include "paragraph/Builder.hpp"\n'
        "int main() {\n"
        "  std::vector<float> b(8 * 4);\n"
        "  std::vector<uint32_t> g(8 * 128, 0xFFFFFFFFu);\n"
        "  Gbuilder::Builder<float> x(b, g, 4);\n"
        "  Gbuilder::Index idx = x.BuildIndex(32, 64);\n"
        "  Gbuilder::Builder<float>::SaveIndex(idx, \"/dev/null\");\n"
        "  return int(idx.ep);\n"
        "}\n")
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-c", str(src), "-o",
                        str(tmp_path / "use_builder.o"), "-I",
                        os.path.join(root, "include")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
