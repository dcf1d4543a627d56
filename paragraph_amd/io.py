"""Vector / gt / index file formats of the reference.

- Big-ANN .fbin / .ibin: uint32 n, uint32 dim header, then n*dim float32 /
  int32 row-major (README.md:15-21).
- Headerless raw arrays, as the reference's loader reads them
  (include/utils.hpp:18-34 reads vec.size() elements with no header skip;
  transform_data_to_gpu.ipynb strips the header).
- gt re-padded to 128 columns with -1 (transform_gt_to_128.ipynb).
- Index layout written by SaveGraph (tests/test_t2iend.cu:615-633):
  uint32 ep, uint32 n, then per node uint32 count followed by count ids.
"""
import numpy as np


def read_bin(path, dtype, headered=True, n=None, dim=None, mmap=True):
    dtype = np.dtype(dtype)
    if headered:
        hdr = np.fromfile(path, dtype=np.uint32, count=2)
        n_h, dim_h = int(hdr[0]), int(hdr[1])
        if mmap:
            a = np.memmap(path, dtype=dtype, mode="r", offset=8,
                          shape=(n_h, dim_h))
        else:
            a = np.fromfile(path, dtype=dtype, offset=8).reshape(n_h, dim_h)
        return a
    if n is None or dim is None:
        raise ValueError("headerless files need n and dim")
    if mmap:
        return np.memmap(path, dtype=dtype, mode="r", shape=(n, dim))
    return np.fromfile(path, dtype=dtype, count=n * dim).reshape(n, dim)


def write_bin(path, a, headered=True):
    a = np.ascontiguousarray(a)
    with open(path, "wb") as f:
        if headered:
            np.array(a.shape, dtype=np.uint32).tofile(f)
        a.tofile(f)


def pad_gt(gt, width=128):
    """Pad or truncate gt rows to `width` columns with -1 (int32)."""
    gt = np.asarray(gt)
    out = np.full((gt.shape[0], width), -1, np.int32)
    k = min(width, gt.shape[1])
    out[:, :k] = gt[:, :k]
    return out


def rows_to_lists(rows, tomb=0xFFFFFFFF):
    """Tomb-padded rows -> per-node id lists."""
    rows = np.asarray(rows).view(np.uint32) if np.asarray(rows).dtype == \
        np.int32 else np.asarray(rows, dtype=np.uint32)
    return [r[r != tomb] for r in rows]


def save_index(path, lists, ep):
    """SaveGraph layout: ep, n, then (count, ids...) per node."""
    n = len(lists)
    with open(path, "wb") as f:
        np.array([ep, n], np.uint32).tofile(f)
        for l in lists:
            l = np.asarray(l, dtype=np.uint32)
            np.array([l.shape[0]], np.uint32).tofile(f)
            l.tofile(f)


def load_index(path):
    raw = np.fromfile(path, dtype=np.uint32)
    ep, n = int(raw[0]), int(raw[1])
    lists, pos = [], 2
    for _ in range(n):
        c = int(raw[pos])
        lists.append(raw[pos + 1:pos + 1 + c].copy())
        pos += 1 + c
    if pos != raw.shape[0]:
        raise ValueError("trailing data in index file")
    return ep, lists


def csr_to_rows(off, ids, width=None, tomb=0xFFFFFFFF):
    """CSR lists -> fixed-width rows padded with tomb (uint32), the graph
    layout the searcher reads (max_degree stride, Gsearcher.cuh:108-296).
    width defaults to the longest list."""
    off = np.asarray(off, dtype=np.int64)
    ids = np.asarray(ids).view(np.uint32) if np.asarray(ids).dtype == \
        np.int32 else np.asarray(ids, dtype=np.uint32)
    n = off.shape[0] - 1
    lens = np.diff(off)
    if width is None:
        width = int(lens.max(initial=1))
    if (lens > width).any():
        raise ValueError("a list is longer than width")
    rows = np.full((n, width), tomb, np.uint32)
    col = np.arange(ids.shape[0], dtype=np.int64) - np.repeat(off[:-1], lens)
    rows[np.repeat(np.arange(n), lens), col] = ids
    return rows


def eval_search_params(n_base, dim, n_query, Km=128, Kp=6, Kd=64, topk=128,
                       reset_iter=4):
    """Build-parameter block for the reference's evaluation searcher
    (search_shared_hashtable's defaults, Gsearcher.cuh:108-113: Km 128,
    Kp 6, Kd 64, reset every 4 iterations) over an index of row width >= Kd;
    only pg_run_search reads it."""
    from . import _native
    return _native.make_params(n_base, dim, n_query, 64, Km, Kp=Kp, Kd=Kd,
                               topk=topk, reset_iter=reset_iter, rounds=1,
                               proj_deg=max(Kd, 16), table_size=0)
