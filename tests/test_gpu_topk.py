"""The exact top-k gt kernel (pg_exact_topk_ip) through the C ABI.

Definition checked: per query, the k base ids with the largest inner
product scored by the FP32 sequential chain, ties to the lower id.
  - float data: the expected rows are the top k, by (chain score desc, id
    asc), of the FP64 top k + 64 candidates, their chain scores taken from
    pg_run_pair_distances' chain kind (pinned to the oracle elsewhere);
  - small-integer data (every score exact in FP32 and FP64, many ties):
    the expected rows come from int64 scores and a stable sort by id;
  - fewer base rows than k: -1 padding; query and base counts that are not
    multiples of the 64-query / 64-row tiles;
  - scores crowded around the k-th beyond the candidate buffer: the exact
    pass gives the same definition."""
import numpy as np
import pytest

from paragraph_amd import _native

pytestmark = pytest.mark.gpu


def _chain_scores(q, b, ids, cuda):
    import torch
    n, dim = b.shape
    p = _native.make_params(n, dim, q.shape[0], 16, 16)
    out = torch.empty(ids.shape, dtype=torch.float32, device=cuda)
    with _native.Context(p) as ctx:
        ctx.run_pair_distances(torch.from_numpy(q).to(cuda),
                               torch.from_numpy(b).to(cuda),
                               torch.from_numpy(ids.astype(np.int32)).to(cuda),
                               out, chain=True)
        ctx.sync()
    return -out.cpu().numpy()  # the chain distance is -dot


def _run(q, b, k, cuda, n_exact=None):
    import torch
    out = torch.full((q.shape[0], k), -7, dtype=torch.int32, device=cuda)
    ne = _native.exact_topk_ip(torch.from_numpy(q).to(cuda),
                               torch.from_numpy(b).to(cuda), k, out)
    torch.cuda.synchronize()
    if n_exact is not None:
        n_exact.append(ne)
    return out.cpu().numpy()


@pytest.mark.parametrize("nq,nb,dim,k", [(1000, 30001, 200, 128),
                                         (333, 20000, 512, 128),
                                         (130, 5000, 64, 32)])
def test_float_data_matches_chain_order(nq, nb, dim, k, cuda):
    rng = np.random.default_rng(nq + dim)
    c = rng.standard_normal((32, dim)).astype(np.float32)
    b = (c[rng.integers(0, 32, nb)] +
         0.3 * rng.standard_normal((nb, dim))).astype(np.float32)
    q = (c[rng.integers(0, 32, nq)] +
         0.3 * rng.standard_normal((nq, dim))).astype(np.float32)
    print(f"exact top-k nq={nq} nb={nb} dim={dim} k={k}", flush=True)
    got = _run(q, b, k, cuda)
    s64 = q.astype(np.float64) @ b.astype(np.float64).T
    cand = np.argsort(-s64, axis=1, kind="stable")[:, :k + 64]
    sc = _chain_scores(q, b, cand, cuda)
    for i in range(nq):
        order = np.lexsort((cand[i], -sc[i]))[:k]
        assert got[i].tolist() == cand[i][order].tolist(), i


def test_integer_data_ties_to_lower_id(cuda):
    rng = np.random.default_rng(3)
    nq, nb, dim, k = 200, 5003, 16, 128
    b = rng.integers(-4, 5, (nb, dim)).astype(np.float32)
    q = rng.integers(-4, 5, (nq, dim)).astype(np.float32)
    got = _run(q, b, k, cuda)
    s = q.astype(np.int64) @ b.astype(np.int64).T
    for i in range(nq):
        exp = np.lexsort((np.arange(nb), -s[i]))[:k]
        assert got[i].tolist() == exp.tolist(), i


def test_fewer_rows_than_k_pads(cuda):
    rng = np.random.default_rng(4)
    b = rng.standard_normal((100, 40)).astype(np.float32)
    q = rng.standard_normal((65, 40)).astype(np.float32)
    got = _run(q, b, 128, cuda)
    assert (got[:, 100:] == -1).all()
    s = q.astype(np.float64) @ b.astype(np.float64).T
    for i in range(65):
        assert sorted(got[i, :100].tolist()) == list(range(100))
        assert (np.diff(s[i][got[i, :100]]) <= 1e-5).all()


def test_crowded_scores_take_the_exact_pass(cuda):
    """Rows whose scores crowd around the k-th (here: many equal and
    near-equal base rows) overflow the screen's buffer and are computed
    by the exact pass; the result is still the exact top k."""
    rng = np.random.default_rng(5)
    nb, dim, k = 3000, 32, 16
    # 2000 equal rows, then rows of small integers: for an all-ones query
    # the top 16 ends inside a 2000-way tie
    b = np.ones((nb, dim), np.float32)
    b[2000:] = rng.integers(-2, 3, (nb - 2000, dim)).astype(np.float32)
    q = np.ones((70, dim), np.float32)
    q[1::2] = rng.integers(-3, 4, (35, dim)).astype(np.float32)
    ne = []
    got = _run(q, b, k, cuda, ne)
    assert ne[0] > 0  # the constant queries take the exact pass
    s = q.astype(np.int64) @ b.astype(np.int64).T
    for i in range(q.shape[0]):
        exp = np.lexsort((np.arange(nb), -s[i]))[:k]
        assert got[i].tolist() == exp.tolist(), i
