"""The oracle against the committed golden fixtures (tests/golden/):
hand-derived tiny cases (expected outputs derived by hand or FP64 brute
force) and whole-stage regression vectors of small generated workloads."""
import numpy as np
import pytest

from oracle import cpu_ref
from tests.golden import golden_io

TOMB = golden_io.TOMB


@pytest.fixture(scope="module")
def hand():
    return golden_io.load_hand()


def test_hand_match(hand):
    c = hand["match"]
    w = cpu_ref.match_winner(np.array(c["gt"], np.uint32), c["n_base"])
    assert w.tolist() == c["winner"]


def test_hand_prune(hand):
    c = hand["prune_2d"]
    pr, pc = cpu_ref.prune(np.array(c["base"], np.float32),
                           np.array(c["rows"], np.uint32), c["mp"])
    assert pc.tolist() == c["pcnt"]
    assert pr[0, :pc[0]].tolist() == c["pruned0"]


def test_hand_merge(hand):
    c = hand["merge_dup"]
    args = (np.array(c["base"], np.float32), np.array(c["off"], np.uint32),
            np.array(c["srcs"], np.uint32), c["cap"],
            np.array(c["pruned"], np.uint32), np.array(c["pcnt"], np.uint32),
            c["deg"])
    assert cpu_ref.merge(*args, strict=True)[0].tolist() == c["strict0"]
    assert cpu_ref.merge(*args, strict=False)[0].tolist() == c["nonstrict0"]


def test_hand_search(hand):
    c = hand["search_complete"]
    out = cpu_ref.search(np.array(c["queries"], np.float32),
                         np.array(c["base"], np.float32),
                         np.array(c["graph"], np.uint32), c["Km"], c["Kp"],
                         c["Kd"], c["topk"], entry_seed=c["entry_seed"])
    assert out.tolist() == c["rows"]


SEARCH_HAND = [("search_reset15", r) for r in ("15", "14", "16", "1000")] \
    + [("search_collect_kp2", "15")]


@pytest.mark.parametrize("name,reset", SEARCH_HAND,
                         ids=[f"{n}-reset{r}" for n, r in SEARCH_HAND])
def test_hand_search_traces(hand, name, reset):
    """Hand-derived K4 traces: the reset period of the visited table, the
    collect order with Kp = 2, and the FLT_MAX tail order (topk = Km +
    Kp*Kd shows the whole list)."""
    c = hand[name]
    st = cpu_ref.OrcStats()
    out = cpu_ref.search(np.array(c["queries"], np.float32),
                         np.array(c["base"], np.float32),
                         np.array(c["graph"], np.uint32), c["Km"], c["Kp"],
                         c["Kd"], c["topk"], reset_iter=int(reset),
                         entry_seed=c["entry_seed"], stats=st)
    assert out.tolist() == c["rows_by_reset_iter"][reset]
    if reset == "15":
        assert st.search_iters == c["search_iters"]
        assert st.search_dist == c["search_dist"]


@pytest.mark.parametrize("name", golden_io.stage_names())
def test_stage_regression(name):
    g = golden_io.load_stage(name)
    st = cpu_ref.OrcStats()
    proj, link = cpu_ref.gpu_stage(g["params"], g["base"], g["gt"], stats=st)
    assert np.array_equal(proj, g["proj"])
    assert np.array_equal(link, g["link"])
    assert st.as_dict() == g["stats"]


def test_fixture_inputs_exact():
    """Stored inputs round-trip exactly (values are float16-representable,
    ids fit uint16)."""
    for name in golden_io.stage_names():
        g = golden_io.load_stage(name)
        assert np.array_equal(g["base"].astype(np.float16).astype(np.float32),
                              g["base"])
        assert g["params"]["n_base"] < 0xFFFF
