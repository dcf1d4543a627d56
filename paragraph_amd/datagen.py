"""Seeded synthetic workloads for the index build (the five benchmark shapes).

Generator (numpy default_rng; seeds fixed here):
  centers  C_k ~ N(0, 1)^dim                                   seed 11
  base     x = C_k(x) + 0.3 N(0, I)                             seed 12
  rotation R = Q of QR(Gaussian), columns sign-fixed by diag(R) seed 13
  offset   o ~ N(0, 0.5^2)^dim, drawn once                      seed 14
  training queries y = R (C_k + 0.3 N(0, I)) + o                seed 15
  held-out eval queries: same distribution                      seed 16
  autotuner queries: same distribution                          seed 18
Optional: L2 normalisation (base, and queries after rotation + offset), Zipf
component choice for queries, ragged gt lengths padded with -1.

Ground truth: exact top-k base ids by inner product (ties to the lower id),
computed in float32 (on the GPU when one is present) and stored as int32 with
-1 padding, the layout the reference reads (transform_gt_to_128.ipynb pads
to 128 columns with -1): torch's float32 product for up to a few million
rows, the library's exact top-k kernel (pg_exact_topk_ip, FP32 chain
scores) for any size.
"""
import numpy as np

CONFIGS = {
    # name: n_base, dim, n_comp, M, L, normalize, zipf, gt_len_range
    "cfg1_100k_d200": dict(n_base=100_000, dim=200, n_comp=64, M=32, L=100),
    "cfg2_1m_d200": dict(n_base=1_000_000, dim=200, n_comp=64, M=32, L=128),
    "cfg3_1m_d512": dict(n_base=1_000_000, dim=512, n_comp=256, M=64, L=160,
                         normalize=True),
    "cfg4_10m_d200": dict(n_base=10_000_000, dim=200, n_comp=64, M=32, L=128),
    "cfg5_5m_d512_zipf": dict(n_base=5_000_000, dim=512, n_comp=1024, M=48,
                              L=128, normalize=True, zipf=1.1,
                              gt_len_range=(16, 128)),
}


def _normal(rng, shape):
    return rng.standard_normal(shape, dtype=np.float32)


def centers(dim, n_comp):
    return _normal(np.random.default_rng(11), (n_comp, dim))


def rotation(dim):
    g = np.random.default_rng(13).standard_normal((dim, dim))
    q, r = np.linalg.qr(g)
    q = q * np.sign(np.diag(r))[None, :]
    return q.astype(np.float32)


def offset(dim):
    return (0.5 * np.random.default_rng(14).standard_normal(dim)).astype(
        np.float32)


def _components(rng, n, n_comp, zipf):
    if zipf is None:
        return rng.integers(0, n_comp, n)
    w = 1.0 / np.arange(1, n_comp + 1, dtype=np.float64) ** zipf
    return rng.choice(n_comp, size=n, p=w / w.sum())


def _normalize(x):
    n = np.linalg.norm(x, axis=1, keepdims=True)
    n[n == 0] = 1.0
    return (x / n).astype(np.float32)


def base_vectors(n_base, dim, n_comp, normalize=False, chunk=1 << 20):
    C = centers(dim, n_comp)
    rng = np.random.default_rng(12)
    out = np.empty((n_base, dim), np.float32)
    for s in range(0, n_base, chunk):
        e = min(n_base, s + chunk)
        comp = rng.integers(0, n_comp, e - s)
        out[s:e] = C[comp] + 0.3 * _normal(rng, (e - s, dim))
    return _normalize(out) if normalize else out


def query_vectors(n, dim, n_comp, seed, normalize=False, zipf=None,
                  chunk=1 << 20):
    C = centers(dim, n_comp)
    R = rotation(dim)
    o = offset(dim)
    rng = np.random.default_rng(seed)
    out = np.empty((n, dim), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = _components(rng, e - s, n_comp, zipf)
        y = C[comp] + 0.3 * _normal(rng, (e - s, dim))
        out[s:e] = y @ R.T + o
    return _normalize(out) if normalize else out


def exact_topk_ip(queries, base, k, device=None, q_chunk=None):
    """Exact top-k ids by inner product, ties to the lower id (int32).

    Float32 scores; with torch on a GPU when `device` is a CUDA device.
    """
    import torch
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    dev = torch.device(device)
    B = torch.from_numpy(np.ascontiguousarray(base)).to(dev)
    nq = queries.shape[0]
    out = np.empty((nq, k), np.int32)
    if q_chunk is None:
        q_chunk = max(1, min(4096, (1 << 31) // max(1, base.shape[0])))
    for s in range(0, nq, q_chunk):
        e = min(nq, s + q_chunk)
        Q = torch.from_numpy(np.ascontiguousarray(queries[s:e])).to(dev)
        scores = Q @ B.T
        kk = min(k, base.shape[0])
        v, i = torch.topk(scores, kk, dim=1, largest=True, sorted=True)
        # deterministic order inside the top-k: score desc, id asc
        i = i.to(torch.int64)
        key_v = -v.double()
        order = torch.argsort(i, dim=1, stable=True)
        key_v = torch.gather(key_v, 1, order)
        i = torch.gather(i, 1, order)
        order2 = torch.argsort(key_v, dim=1, stable=True)
        i = torch.gather(i, 1, order2)
        res = np.full((e - s, k), -1, np.int32)
        res[:, :kk] = i.cpu().numpy().astype(np.int32)
        out[s:e] = res
    return out


def exact_topk_ip_native(queries, base, k, device="cuda", q_chunk=1 << 20,
                         log=None):
    """Exact top-k ids by inner product with the library's GPU kernel
    (pg_exact_topk_ip: BF16 MFMA screen with a rigorous bound, survivors
    re-scored with the FP32 sequential chain, ties to the lower id), for
    any N x N size; int32 with -1 padding.  `log(text)` is called after
    every chunk of q_chunk queries."""
    import time

    import torch
    from . import _native
    t0 = time.time()
    dev = torch.device(device)
    B = torch.from_numpy(np.ascontiguousarray(base)).to(dev)
    nq = queries.shape[0]
    out = np.empty((nq, k), np.int32)
    o = torch.empty((min(nq, q_chunk), k), dtype=torch.int32, device=dev)
    for s in range(0, nq, q_chunk):
        e = min(nq, s + q_chunk)
        Q = torch.from_numpy(np.ascontiguousarray(queries[s:e])).to(dev)
        exact_topk_ip_native.exact_rows += _native.exact_topk_ip(
            Q, B, k, o[:e - s])
        torch.cuda.synchronize(dev)
        out[s:e] = o[:e - s].cpu().numpy()
        del Q
        if log is not None:
            log(f"exact gt: {e} of {nq} queries after {time.time() - t0:.0f}"
                " s")
    del B, o
    torch.cuda.empty_cache()
    return out


exact_topk_ip_native.exact_rows = 0  # rows the kernel's exact pass computed


def subset_topk_ip(queries, base, k, subset, seed=19, device=None):
    """Sharded top-k by inner product (ties to the lower id): the base is
    split into random shards of about `subset` rows and query q takes the
    exact top-k of shard q mod #shards.  Every base row can appear in gt,
    at the cost of an N x subset product instead of N x N; a well-formed gt
    for builds whose full exact gt is out of reach (it is only a build
    input)."""
    rng = np.random.default_rng(seed)
    n = base.shape[0]
    n_sh = max(1, -(-n // subset))
    perm = rng.permutation(n)
    shards = np.array_split(perm, n_sh)
    nq = queries.shape[0]
    out = np.full((nq, k), -1, np.int32)
    for s, ids in enumerate(shards):
        ids = np.sort(ids)
        qsel = np.arange(s, nq, n_sh)
        if qsel.size == 0:
            continue
        local = exact_topk_ip(queries[qsel], base[ids], k, device)
        out[qsel] = np.where(local >= 0, ids[np.maximum(local, 0)], -1)
    return out


def ragged_gt(gt, lo, hi, seed=17):
    """Per-row length uniform in [lo, hi], rest padded with -1."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, gt.shape[0])
    g = gt.copy()
    cols = np.arange(g.shape[1])[None, :]
    g[cols >= lens[:, None]] = -1
    return g


def make_workload(n_base, dim, n_comp, n_query=None, gt_k=128,
                  normalize=False, zipf=None, gt_len_range=None,
                  n_eval=0, device=None, with_gt=True, gt_subset=None,
                  exact_native=False, log=None):
    """Returns dict(base, queries, gt (int32, -1 padded), eval_queries).
    gt_subset: exact top-k within that many random base rows instead of all;
    exact_native: the exact gt from the library's GPU kernel (any size;
    torch's float32 product otherwise); log: its progress callback."""
    n_query = n_base if n_query is None else n_query
    base = base_vectors(n_base, dim, n_comp, normalize)
    queries = query_vectors(n_query, dim, n_comp, 15, normalize, zipf)
    w = dict(base=base, queries=queries)
    if with_gt:
        if gt_subset is not None and gt_subset < n_base:
            gt = subset_topk_ip(queries, base, gt_k, gt_subset, device=device)
        elif exact_native:
            gt = exact_topk_ip_native(queries, base, gt_k,
                                      device=device or "cuda", log=log)
        else:
            gt = exact_topk_ip(queries, base, gt_k, device)
        if gt_len_range is not None:
            gt = ragged_gt(gt, *gt_len_range)
        w["gt"] = gt
    if n_eval:
        w["eval_queries"] = query_vectors(n_eval, dim, n_comp, 16, normalize,
                                          zipf)
    return w


def make_config(name, scale=1.0, n_eval=0, device=None, with_gt=True,
                gt_subset=None, exact_native=False, log=None):
    """One of CONFIGS, optionally scaled down in n_base (for tests)."""
    c = dict(CONFIGS[name])
    n = max(64, int(c["n_base"] * scale))
    w = make_workload(n, c["dim"], c["n_comp"], normalize=c.get("normalize",
                      False), zipf=c.get("zipf"),
                      gt_len_range=c.get("gt_len_range"), n_eval=n_eval,
                      device=device, with_gt=with_gt, gt_subset=gt_subset,
                      exact_native=exact_native, log=log)
    w["M"], w["L"] = c["M"], c["L"]
    w["cfg"] = c
    return w


def tune_queries(name, n):
    """Queries the autotuner scores recall on (disjoint seed from the
    training and held-out eval queries)."""
    c = CONFIGS[name]
    return query_vectors(n, c["dim"], c["n_comp"], 18,
                         c.get("normalize", False), c.get("zipf"))
