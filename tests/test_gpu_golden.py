"""The HIP path against the committed golden fixtures, through the C ABI:
the hand-derived cases stage by stage, and the whole stage of each
generated workload id-for-id with its counters."""
import numpy as np
import pytest

from paragraph_amd import _native
from tests.golden import golden_io

pytestmark = pytest.mark.gpu
TOMB = golden_io.TOMB


def _dev(a, torch):
    if a.dtype == np.uint32:
        return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    a = t.cpu().numpy()
    return a.view(np.uint32) if a.dtype == np.int32 else a


@pytest.fixture(scope="module")
def hand():
    return golden_io.load_hand()


def test_hand_match(hand, cuda):
    import torch
    c = hand["match"]
    gt = np.array(c["gt"], np.uint32)
    p = _native.make_params(c["n_base"], 2, gt.shape[0], 16, 16,
                            gt_k=gt.shape[1])
    win = torch.empty(c["n_base"], dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_match(_dev(gt, torch), win)
        ctx.sync()
    assert _host(win).tolist() == c["winner"]


def test_hand_prune(hand, cuda):
    import torch
    c = hand["prune_2d"]
    base = np.array(c["base"], np.float32)
    rows = np.array(c["rows"], np.uint32)
    n = base.shape[0]
    p = _native.make_params(n, base.shape[1], n, 16, 16)
    pr = torch.empty((n, c["mp"]), dtype=torch.int32, device=cuda)
    pc = torch.empty(n, dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_prune_rows(_dev(base, torch), _dev(rows, torch), c["mp"], pr,
                           pc)
        ctx.sync()
    pcn = _host(pc)
    assert pcn.tolist() == c["pcnt"]
    assert _host(pr)[0, :pcn[0]].tolist() == c["pruned0"]


def test_hand_merge(hand, cuda):
    import torch
    c = hand["merge_dup"]
    base = np.array(c["base"], np.float32)
    n = base.shape[0]
    p = _native.make_params(n, base.shape[1], n, 16, 16)
    pruned = np.array(c["pruned"], np.uint32)
    for strict, key in ((True, "strict0"), (False, "nonstrict0")):
        out = torch.empty((n, c["deg"]), dtype=torch.int32, device=cuda)
        with _native.Context(p) as ctx:
            ctx.run_merge(_dev(base, torch),
                          _dev(np.array(c["off"], np.uint32), torch),
                          _dev(np.array(c["srcs"], np.uint32), torch),
                          c["cap"], _dev(pruned, torch),
                          _dev(np.array(c["pcnt"], np.uint32), torch),
                          c["deg"], strict, out)
            ctx.sync()
        assert _host(out)[0].tolist() == c[key]


def test_hand_search(hand, cuda):
    import torch
    c = hand["search_complete"]
    base = np.array(c["base"], np.float32)
    q = np.array(c["queries"], np.float32)
    graph = np.array(c["graph"], np.uint32)
    n = base.shape[0]
    p = _native.make_params(n, base.shape[1], n, 2, c["Km"], Kp=c["Kp"],
                            Kd=c["Kd"], topk=c["topk"], proj_deg=graph.shape[1],
                            rounds=1, entry_seed=c["entry_seed"])
    assert _native.check_params(p)[0] == 0, _native.check_params(p)
    out = torch.empty((q.shape[0], c["topk"]), dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_search(_dev(q, torch), _dev(base, torch), _dev(graph, torch),
                       out)
        ctx.sync()
    assert _host(out).tolist() == c["rows"]


SEARCH_HAND = [("search_reset15", r) for r in ("15", "14", "16", "1000")] \
    + [("search_collect_kp2", "15")]


@pytest.mark.parametrize("name,reset", SEARCH_HAND,
                         ids=[f"{n}-reset{r}" for n, r in SEARCH_HAND])
@pytest.mark.parametrize("flags", [0x0, 0x20], ids=["grid", "persistent"])
def test_hand_search_traces(hand, name, reset, flags, cuda):
    """The hand-derived K4 traces (reset period, Kp = 2 collect order, the
    FLT_MAX tail) through pg_run_search, rows and counters."""
    import torch
    c = hand[name]
    base = np.array(c["base"], np.float32)
    q = np.array(c["queries"], np.float32)
    graph = np.array(c["graph"], np.uint32)
    n = base.shape[0]
    p = _native.make_params(n, base.shape[1], n, 2, c["Km"], Kp=c["Kp"],
                            Kd=c["Kd"], topk=c["topk"],
                            proj_deg=graph.shape[1], rounds=1,
                            reset_iter=int(reset), tune_flags=flags,
                            entry_seed=c["entry_seed"])
    assert _native.check_params(p)[0] == 0, _native.check_params(p)
    out = torch.empty((q.shape[0], c["topk"]), dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_search(_dev(q, torch), _dev(base, torch), _dev(graph, torch),
                       out)
        ctx.sync()
        st = ctx.stats()
    assert _host(out).tolist() == c["rows_by_reset_iter"][reset]
    if reset == "15":
        assert st.search_iters == c["search_iters"]
        assert st.search_dist == c["search_dist"]


@pytest.mark.parametrize("name", golden_io.stage_names())
def test_stage_golden(name, cuda):
    import torch
    g = golden_io.load_stage(name)
    pr = g["params"]
    p = _native.make_params(pr["n_base"], pr["dim"], pr["n_query"], pr["M"],
                            pr["Km"])
    for k, v in pr.items():
        assert getattr(p, k) == v, (k, getattr(p, k), v)
    n = pr["n_base"]
    proj = torch.empty((n, pr["proj_deg"]), dtype=torch.int32, device=cuda)
    link = torch.empty((n, pr["M"]), dtype=torch.int32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.stage_device(_dev(g["base"], torch), _dev(g["gt"], torch), proj,
                         link)
        ctx.wait()
        st = ctx.stats().as_dict()
    assert np.array_equal(_host(proj), g["proj"])
    assert np.array_equal(_host(link), g["link"])
    for k, v in g["stats"].items():
        assert st[k] == v, (k, st[k], v)
