"""The CPU reference of the surrogate's backward pass against central finite
differences of its own loss (float64, no bfloat16 rounding): pins the
gradient formulas the device kernels are then compared with."""
import numpy as np

from tests import surrogate_ref as S


def test_reference_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = np.zeros(S.NPARAM)
    p[S.oW1:S.ob1] = rng.uniform(-0.2, 0.2, S.H * S.F)
    p[S.ob1:S.oW2] = rng.uniform(0.0, 0.1, S.H)          # most units active
    p[S.oW2:S.ob2] = rng.uniform(-0.2, 0.2, S.H * S.H)
    p[S.ob2:S.oW3] = rng.uniform(0.0, 0.1, S.H)
    p[S.oW3:S.oW3 + 3 * S.H] = rng.uniform(-0.2, 0.2, 3 * S.H)
    p[S.ob3:S.ob3 + 3] = rng.uniform(-0.1, 0.1, 3)
    X = rng.uniform(-1, 0, (32, S.F))
    T = np.concatenate([rng.normal(size=(32, 3)), (rng.random((32, 1)) > 0.2).astype(float)], 1)
    loss, g, _ = S.forward_backward(p, X, T, emulate_bf16=False, dtype=np.float64)
    idx = np.concatenate([rng.integers(S.oW1, S.ob1, 12), rng.integers(S.ob1, S.oW2, 4),
                          rng.integers(S.oW2, S.ob2, 12), rng.integers(S.ob2, S.oW3, 4),
                          rng.integers(S.oW3, S.oW3 + 3 * S.H, 8), np.arange(S.ob3, S.ob3 + 3)])
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp, _, _ = S.forward_backward(q, X, T, emulate_bf16=False, dtype=np.float64)
        q[i] -= 2 * eps
        lm, _, _ = S.forward_backward(q, X, T, emulate_bf16=False, dtype=np.float64)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.0e-5], np.float32)
    r = S.bf16(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125) and r[3] == np.float32(1.0078125)
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-5) <= 3.0e-5 * 2 ** -8
