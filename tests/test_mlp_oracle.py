"""Pins the MLP oracle: analytic gradients vs central finite differences,
bf16 rounding vs its definition, Adam vs its closed form on a scalar."""
import numpy as np

from oracle import mlp_oracle as mo


def small_net(seed=0, sizes=(13, 24, 24, 24, 40)):
    rng = np.random.default_rng(seed)
    Ws = [rng.standard_normal((sizes[i + 1], sizes[i])) * np.sqrt(2.0 / sizes[i]) for i in range(len(sizes) - 1)]
    bs = [rng.standard_normal(sizes[i + 1]) * 0.1 for i in range(len(sizes) - 1)]
    return Ws, bs


def test_gradients_match_finite_differences():
    Ws, bs = small_net()
    rng = np.random.default_rng(1)
    X = rng.standard_normal((8, 13))
    T = rng.standard_normal((8, 40))
    assert mo.check_gradients(Ws, bs, X, T) < 1e-6


def test_bf16_rounding():
    # bf16 keeps 7 fraction bits: the ulp at 1.0 is 2^-7
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -10, -3.1415926, 65504.0, 1e-30], np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                  # exactly half an ulp: tie -> even
    assert r[2] == 1.0 + 2 ** -7        # above half an ulp -> up
    assert abs(r[3] - (-3.140625)) < 1e-12
    # every result is representable in 8 significant bits
    m, _ = np.frexp(r[r != 0])
    assert np.all(np.abs(m * 256 - np.round(m * 256)) < 1e-9)


def test_adam_scalar_closed_form():
    w, m, v = np.array([1.0]), np.zeros(1), np.zeros(1)
    # first step moves by lr * sign(g) (bias-corrected moments are g and g^2)
    w1, m, v = mo.adam_step(w, np.array([0.3]), m, v, 1, 0.01, eps=0.0)
    assert abs(w1[0] - (1.0 - 0.01)) < 1e-15
