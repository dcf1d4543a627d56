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
