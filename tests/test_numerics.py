"""Host roundings used by the screened-path error bounds, pinned against
independent implementations (torch's dtype conversions)."""
import numpy as np
import torch

from paragraph_amd import numerics


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(50000) * 10.0 ** k
                        for k in (-20, -3, 0, 3, 20)]).astype(np.float32)
    x = np.concatenate([x, np.float32([0.0, -0.0, 1.0 + 2 ** -8,
                                       1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9])])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(numerics.bf16_round(x), ref)


def test_e4m3_round_matches_torch():
    rng = np.random.default_rng(4)
    v = np.concatenate([rng.standard_normal(100000) * 100.0,
                        rng.standard_normal(100000) * 0.01,
                        np.array([448.0, -448.0, 0.0, 2 ** -9, 2 ** -10,
                                  3 * 2 ** -10, 2 ** -6, 1.0625, 1.1875,
                                  449.0, 463.9])]).astype(np.float32)
    v = v[np.abs(v) < 464.0]
    ref = torch.from_numpy(v).to(torch.float8_e4m3fn).to(torch.float64)
    assert np.array_equal(numerics.e4m3_round(v), ref.numpy())


def test_e4m3_special_values():
    # ties to even: 1 + 1/16 lies between 1 and 1.125 -> 1; 1 + 3/16 -> 1.25
    assert numerics.e4m3_round(np.float32([1.0625, 1.1875])).tolist() == [
        1.0, 1.25]
    # smallest subnormal 2^-9; half of it ties to 0
    assert numerics.e4m3_round(np.float32([2 ** -9, 2 ** -10])).tolist() == [
        2 ** -9, 0.0]


def test_fp8_quantize_scales():
    x = np.float32([[1.0, -2.0, 0.5], [0.0, 0.0, 0.0], [448.0, 1.0, -3.0]])
    q, s = numerics.fp8_quantize(x)
    assert s.tolist() == [np.float32(2.0) / np.float32(448.0), 1.0, 1.0]
    assert np.max(np.abs(q[0])) == 448.0 and np.all(q[1] == 0)
    assert q[2].tolist() == [448.0, 1.0, -3.0]
