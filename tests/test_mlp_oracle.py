"""Pins of the CPU reference of the relaxation surrogate network
(oracle/mlp.py): the analytic backward against central finite differences
of the exact FP64 loss, and the BF16 rounding helper against closed forms."""
import numpy as np

from fooskagpu_amd import surrogate
from oracle import mlp


def test_bf16_rounding_closed_forms():
    assert mlp.bf16(1.0) == 1.0
    assert mlp.bf16(1.0 + 2 ** -8) == 1.0  # tie -> even mantissa
    assert mlp.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert mlp.bf16(-3.0) == -3.0
    x = np.random.default_rng(0).normal(size=1000)
    assert np.all(np.abs(mlp.bf16(x) - x) <= np.abs(x) * 2 ** -8)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = surrogate.init_weights(5).astype(np.float64)
    w += rng.normal(scale=0.05, size=w.size)  # non-zero biases
    x = rng.normal(size=(7, 3))
    t = rng.normal(size=(7, 3))
    g = mlp.backward(w, x, t, bf16_round=False)
    idx = rng.choice(w.size, 60, replace=False)
    idx = np.concatenate([idx, [0, 191, 192, 255, 256, 4351, 4352, 4415, 4416, 8511, 8512, 8575, 8576, 8767, 8768,
                                8770]])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mlp.loss(wp, x, t) - mlp.loss(wm, x, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_parameter_layout_matches_the_c_abi():
    assert sum(int(np.prod(s)) for _, s in mlp.SHAPES) == 8771 == surrogate.N_PARAMS
    assert [n for n, _ in mlp.SHAPES] == [n for n, _ in surrogate.SHAPES]
