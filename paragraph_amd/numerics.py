"""Host restatements of the operand roundings the screened evaluators apply,
used to state their error bounds (tests, bench).  numpy only.

bf16_round: float32 -> bfloat16, round to nearest even (the (__bf16) cast).
fp8_quantize: per-vector scaled OCP e4m3fn rows as the FP8 path stores them:
  s = max|x| / 448 (float32 division; 1 for a zero row), q = e4m3(x * (1/s)),
  round to nearest even, |q| <= 448, subnormal step 2^-9.
"""
import numpy as np


def bf16_round(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(
        np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).reshape(np.shape(x))


def e4m3_round(v):
    """float32 -> nearest e4m3fn value (ties to even), as float64.  Inputs
    are expected within [-448 (1 + 2^-4), 448 (1 + 2^-4)); larger magnitudes
    raise."""
    v = np.asarray(v, dtype=np.float64)
    a = np.abs(v)
    if np.any(a >= 464.0):
        raise ValueError("value beyond the e4m3fn range")
    e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    e = np.maximum(e, -6.0)                 # subnormals share the 2^-6 binade
    step = np.exp2(e - 3.0)
    r = np.rint(a / step) * step            # np.rint: ties to even
    return np.sign(v) * np.minimum(r, 448.0)


def fp8_quantize(x):
    """Rows of x (float32, 2-D) -> (q values as float64, scales float32)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    mx = np.max(np.abs(x), axis=1).astype(np.float32)
    s = np.where(mx > 0, mx / np.float32(448.0), np.float32(1.0)).astype(
        np.float32)
    inv = (np.float32(1.0) / s).astype(np.float32)
    v = (x * inv[:, None]).astype(np.float32)
    return e4m3_round(v), s
