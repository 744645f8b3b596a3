"""Pins the network checker (oracle/mlp_oracle.py): its backward pass in
exact (FP64, unrounded) mode equals central finite differences of its loss,
and its BF16 rounding is round-to-nearest-even on the BF16 grid."""
import numpy as np

import mlp_oracle as mo


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-39, 1e30],
                 dtype=np.float32)
    y = mo.bf16(x)
    assert y[0] == 1.0
    assert y[1] == 1.0  # 1 + 2^-8 is half-way between 1 and 1 + 2^-7: ties to the even 1
    assert y[2] == 1.0
    assert y[3] == 1.0 + 2 ** -7
    assert y[4] == -2.5
    # every result lies on the BF16 grid (low 16 bits zero) within half an ulp
    bits = y.view(np.uint32)
    assert not (bits & 0xFFFF).any()
    normal = np.isfinite(y) & (np.abs(x) > 1.2e-38)  # subnormals keep fewer bits
    rel = np.abs(y[normal] - x[normal]) / np.abs(x[normal])
    assert (rel <= 2 ** -8 + 1e-12).all()


def test_exact_backward_equals_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mo.NUM_PARAMS)
    X = rng.normal(0, 1, (7, mo.N_IN))
    T = rng.normal(0, 1, (7, mo.N_OUT))
    g, loss = mo.gradient(p, X, T, exact=True)
    assert loss == mo.loss_exact(p, X, T)
    idx = rng.choice(mo.NUM_PARAMS, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(mo.NUM_PARAMS)
        e[i] = h
        fd = (mo.loss_exact(p + e, X, T) - mo.loss_exact(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_rounded_gradient_close_to_exact():
    """The BF16 emulation differs from the exact gradient only by rounding."""
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.2, mo.NUM_PARAMS).astype(np.float32)
    X = rng.normal(0, 1, (256, mo.N_IN)).astype(np.float32)
    T = rng.normal(0, 1, (256, mo.N_OUT)).astype(np.float32)
    g32, l32 = mo.gradient(p, X, T)
    g64, l64 = mo.gradient(p.astype(np.float64), X, T, exact=True)
    assert np.linalg.norm(g32 - g64) <= 0.05 * np.linalg.norm(g64)
    assert abs(l32 - l64) <= 0.02 * l64
