"""Pins of the checker's learned-guiding network (oracle/oracle.cpp, guiding section).

- the exact gradient of the objective against central finite differences of the exact loss
  (double precision): a wrong index, transposition, ReLU mask or softmax term fails;
- the BF16 restatement (the device's rounding steps) against the exact one;
- closed forms: the frequency encoding, a zero output layer gives the uniform bin table and the
  mixing probability of its bias, the objective's value for a known distribution.
"""
import math

import numpy as np
import pytest

import oracle

N_PARAMS = 64 * 80 + 4 * 64 * 64 + 144 * 64 + 5 * 64 + 144
W6 = 64 * 80 + 4 * 64 * 64           # offset of the output layer's weights
B6 = 64 * 80 + 4 * 64 * 64 + 144 * 64 + 5 * 64  # offset of the output layer's biases


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, dtype=np.float32)
    off = 0
    for rows, cols in [(64, 80)] + [(64, 64)] * 4 + [(144, 64)]:
        lim = math.sqrt(6.0 / cols) * scale
        p[off:off + rows * cols] = rng.uniform(-lim, lim, rows * cols)
        off += rows * cols
    p[off:] = rng.uniform(-0.1, 0.1, N_PARAMS - off)
    return p


def _records(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), dtype=np.float32)
    r[:, :5] = rng.uniform(0, 1, size=(n, 5))
    r[:, 5] = rng.integers(0, 128, size=n)
    r[:, 6] = rng.uniform(0.01, 2.0, size=n)
    r[:, 7] = rng.exponential(1.0, size=n)
    r[::7, 7] = 0.0  # records that carried no radiance
    return r


def test_param_count():
    assert oracle.guide_num_params() == N_PARAMS


def test_exact_gradient_matches_finite_differences():
    p = _params(1).astype(np.float64)
    rec = _records(24, 2)
    g, loss = oracle.guide_gradient(p.astype(np.float32), rec, rounded=False)
    assert abs(loss - oracle.guide_loss64(p, rec)) < 1e-4 * max(1.0, abs(loss))
    rng = np.random.default_rng(3)
    # parameters of every layer and every kind (weights, biases, the mixing row and bias)
    picks = list(rng.choice(N_PARAMS, size=120, replace=False)) + [W6 + 128 * 64 + 5, B6 + 128, B6 + 3, 0]
    h = 1e-4
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle.guide_loss64(pp, rec) - oracle.guide_loss64(pm, rec)) / (2 * h)
        # the float32 parameter copy and ReLU kinks limit agreement to about 1e-4 of the scale
        assert abs(fd - g[i]) <= 2e-3 * max(abs(fd), abs(g[i])) + 2e-5, (i, fd, g[i])


def test_rounded_gradient_tracks_exact():
    """BF16 steps change each gradient entry by about 2^-8 relative, not its direction."""
    p = _params(4)
    rec = _records(256, 5)
    ge, le = oracle.guide_gradient(p, rec, rounded=False)
    gr, lr = oracle.guide_gradient(p, rec, rounded=True)
    assert abs(le - lr) < 0.02 * abs(le)
    cos = float(ge @ gr / (np.linalg.norm(ge) * np.linalg.norm(gr)))
    assert cos > 0.999
    assert np.linalg.norm(gr - ge) < 0.05 * np.linalg.norm(ge)


def test_zero_output_layer_gives_uniform_table_and_bias_ratio():
    p = _params(6)
    p[W6:W6 + 144 * 64] = 0.0
    p[B6:B6 + 144] = 0.0
    zc = math.log(0.3 / 0.7)
    p[B6 + 128] = zc
    out = oracle.guide_forward(p, np.random.default_rng(7).uniform(size=(16, 5)))
    np.testing.assert_allclose(out[:, :128], 1.0 / 128, rtol=1e-6)
    np.testing.assert_allclose(out[:, 128], 0.05 + 0.85 / (1 + math.exp(-zc)), rtol=1e-6)


def test_table_is_a_distribution_with_full_support():
    out = oracle.guide_forward(_params(8, scale=3.0), np.random.default_rng(9).uniform(size=(64, 5)))
    np.testing.assert_allclose(out[:, :128].sum(axis=1), 1.0, rtol=1e-5)
    assert (out[:, :128] > 0).all()
    assert ((out[:, 128] >= 0.05) & (out[:, 128] <= 0.9)).all()
    assert out[:, :128].max() > 4.0 / 128  # a non-trivial network is not uniform


def test_uniform_network_loss_closed_form():
    """With zero output weights and biases (except the mixing bias): q = 1 / (4 pi) for every bin,
    so loss = -sum (w / sum w) log(c / 4 pi + (1 - c) p_b)."""
    p = _params(10)
    p[W6:W6 + 144 * 64] = 0.0
    p[B6:B6 + 144] = 0.0
    rec = _records(50, 11)
    _, loss = oracle.guide_gradient(p, rec, rounded=False)
    c = 0.05 + 0.85 * 0.5
    w = rec[:, 7] / rec[:, 7].sum()
    exp = -(w * np.log(c / (4 * np.pi) + (1 - c) * rec[:, 6].astype(np.float64))).sum()
    assert abs(loss - exp) < 1e-5 * abs(exp)
