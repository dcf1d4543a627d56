"""The Python front end of the reference's Builder interface
(include/Builder.hpp:6-41) on the GPU: host entry point, same rows as the
golden fixture, screened builds selected by name."""
import numpy as np
import pytest

from paragraph_amd import Builder
from tests.golden import golden_io

pytestmark = pytest.mark.gpu


def test_builder_matches_golden(cuda):
    g = golden_io.load_stage("stage_d200")
    pr = g["params"]
    b = Builder(g["base"], g["gt"].view(np.int32))
    proj, link = b.LaunchBuild(M=pr["M"], L=pr["Km"])
    assert np.array_equal(proj, g["proj"])
    assert np.array_equal(link, g["link"])
    assert b.stats_["search_iters"] == g["stats"]["search_iters"]


def test_builder_screened_by_name(cuda):
    g = golden_io.load_stage("stage_d200")
    pr = g["params"]
    n = pr["n_base"]
    b = Builder(g["base"].ravel(), g["gt"].ravel(), dim=pr["dim"])
    for prec in ("bf16", "fp8"):
        proj, link = b.LaunchBuild(M=pr["M"], L=pr["Km"], precision=prec)
        assert np.array_equal(proj, g["proj"])  # projection is FP32 work
        valid = link != golden_io.TOMB
        assert (link[valid] < n).all()
        assert not (link == np.arange(n, dtype=np.uint32)[:, None]).any()
        overlap = np.mean([len(set(a[a != golden_io.TOMB]) &
                               set(c[c != golden_io.TOMB])) /
                           max(1, (c != golden_io.TOMB).sum())
                           for a, c in zip(link, g["link"])])
        assert overlap > 0.8, (prec, overlap)


@pytest.mark.parametrize("host_stages", ["gpu", "host"])
def test_build_index_matches_oracle(host_stages, cuda, tmp_path):
    """The whole build (GPU stage with the host bipartite stage beside it,
    FusionFinal, entry point, index file) equals the oracle's composition of
    the same stages, id for id."""
    from oracle import cpu_ref
    from paragraph_amd import io
    g = golden_io.load_stage("stage_d200")
    pr = g["params"]
    b = Builder(g["base"], g["gt"].view(np.int32))
    path = str(tmp_path / "index.bin")
    ep, off, ids = b.BuildIndex(M=pr["M"], L=pr["Km"], path=path,
                                host_stages=host_stages)
    gt = np.ascontiguousarray(g["gt"], dtype=np.uint32)
    bo, bi = cpu_ref.bipartite(g["base"], gt, 40, 6)
    fo, fi = cpu_ref.fusion_final(g["base"], bo, bi, g["link"], g["proj"], 70)
    assert ep == cpu_ref.entry_point(gt, pr["n_base"])
    np.testing.assert_array_equal(off, fo)
    np.testing.assert_array_equal(ids, fi)
    ep_f, lists = io.load_index(path)
    assert ep_f == ep
    assert all(np.array_equal(lists[i], fi[fo[i]:fo[i + 1]])
               for i in range(pr["n_base"]))


def test_eval_search_over_final_index_matches_oracle(cuda):
    """The reference's evaluation searcher (search_shared_hashtable:
    Km 128, Kp 6, Kd 64, reset every 4) over the final index, as padded
    rows, equals the oracle's search over the same rows."""
    import torch
    from oracle import cpu_ref
    from paragraph_amd import _native, io
    g = golden_io.load_stage("stage_d200")
    pr = g["params"]
    n, dim = pr["n_base"], pr["dim"]
    b = Builder(g["base"], g["gt"].view(np.int32))
    ep, off, ids = b.BuildIndex(M=pr["M"], L=pr["Km"])
    rows = io.csr_to_rows(off, ids, max(64, int(np.diff(off.astype(
        np.int64)).max())))
    rng = np.random.default_rng(4)
    q = (g["base"][rng.permutation(n)[:200]] +
         0.1 * rng.standard_normal((200, dim))).astype(np.float32)
    p = io.eval_search_params(n, dim, q.shape[0])
    out_d = torch.full((q.shape[0], p.topk), -7, dtype=torch.int32,
                       device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_search(torch.from_numpy(q).to(cuda),
                       torch.from_numpy(g["base"]).to(cuda),
                       torch.from_numpy(rows.view(np.int32)).to(cuda), out_d)
        ctx.sync()
    ref = cpu_ref.search(q, g["base"], rows, p.Km, p.Kp, p.Kd, p.topk,
                         reset_iter=p.reset_iter)
    np.testing.assert_array_equal(out_d.cpu().numpy().view(np.uint32), ref)


def test_cpp_front_end_build_tool(cuda, tmp_path):
    """The C++ front end run for real: pg_build_index (Gbuilder::Builder
    LaunchBuild + BuildIndex + SaveIndex, reading .fbin / .ibin with the
    C++ reader, gt of 100 columns re-padded to 128) gives the golden rows
    and the oracle's final index, id for id."""
    import subprocess
    from oracle import cpu_ref
    from paragraph_amd import build, io
    g = golden_io.load_stage("stage_d200")
    pr = g["params"]
    gt_i = g["gt"].view(np.int32)
    fb, fg = str(tmp_path / "base.fbin"), str(tmp_path / "gt.ibin")
    io.write_bin(fb, g["base"])
    # 100 stored columns; the tool pads to 128 with -1 like the notebook
    io.write_bin(fg, np.ascontiguousarray(gt_i[:, :100]))
    fi_, fl, fp = (str(tmp_path / x) for x in ("index.bin", "link.ibin",
                                               "proj.ibin"))
    print(f"pg_build_index n={pr['n_base']} M={pr['M']} L={pr['Km']}",
          flush=True)
    r = subprocess.run([build.TOOL, fb, fg, str(pr["M"]), str(pr["Km"]), fi_,
                        fl, fp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    gt_pad = io.pad_gt(gt_i[:, :100]).view(np.uint32)
    # the golden rows were built from the full 128-column gt: compare with
    # the Python front end on the same padded gt instead
    b = Builder(g["base"], gt_pad.view(np.int32))
    proj, link = b.LaunchBuild(M=pr["M"], L=pr["Km"])
    np.testing.assert_array_equal(io.read_bin(fl, np.uint32), link)
    np.testing.assert_array_equal(io.read_bin(fp, np.uint32), proj)
    bo, bi = cpu_ref.bipartite(g["base"], gt_pad, 40, 6)
    fo, fi = cpu_ref.fusion_final(g["base"], bo, bi, link, proj, 70)
    ep, lists = io.load_index(fi_)
    assert ep == cpu_ref.entry_point(gt_pad, pr["n_base"])
    assert all(np.array_equal(lists[i], fi[fo[i]:fo[i + 1]])
               for i in range(pr["n_base"]))
