"""Pins of the oracle's warm-start network (no reference counterpart): BF16
rounding against numpy's float32 bit pattern, the loss against a numpy
forward pass, and the gradient against central differences taken on
BF16-exact weights with BF16-exact steps (the loss is piecewise quadratic in
each parameter, so the differences are exact away from ReLU kinks)."""
import numpy as np
import pytest

NI, NH, NO = 24, 128, 12
OW1, OB1 = 0, NI * NH
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH * NO
NP = OB3 + NO


def _bf16(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _weights(rng):
    w = np.zeros(NP, np.float32)
    # multiples of 2^-6 below 2 in magnitude: exact in BF16 (8 significant bits)
    q = lambda n, s: (np.round(rng.normal(size=n) * s * 64) / 64).clip(-1.9, 1.9)
    w[OW1:OB1] = q(NI * NH, 0.4)
    w[OB1:OW2] = q(NH, 0.1)
    w[OW2:OB2] = q(NH * NH, 0.15)
    w[OB2:OW3] = q(NH, 0.1)
    w[OW3:OB3] = q(NH * NO, 0.15)
    w[OB3:] = q(NO, 0.1)
    return w


def _forward(w, X):
    W1 = _bf16(w[OW1:OB1]).reshape(NI, NH).astype(np.float64)
    W2 = _bf16(w[OW2:OB2]).reshape(NH, NH).astype(np.float64)
    W3 = _bf16(w[OW3:OB3]).reshape(NH, NO).astype(np.float64)
    h1 = np.maximum(X.astype(np.float64) @ W1 + w[OB1:OW2], 0)
    h2 = np.maximum(h1 @ W2 + w[OB2:OW3], 0)
    return h2 @ W3 + w[OB3:]


def test_bf16_rounding(oracle):
    rng = np.random.default_rng(0)
    for f in rng.normal(size=50).astype(np.float32) * 10:
        assert oracle.lib().orc_bf16_round(float(f)) == float(_bf16(f))
    # ties go to even
    assert oracle.lib().orc_bf16_round(1.0 + 2 ** -8) == 1.0
    assert oracle.lib().orc_bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6


def test_mlp_loss_matches_numpy(oracle):
    rng = np.random.default_rng(1)
    w = rng.normal(size=NP).astype(np.float32) * 0.1
    X = rng.normal(size=(7, NI)).astype(np.float32)
    T = rng.normal(size=(7, NO)).astype(np.float32)
    L, _, Y = oracle.mlp_loss_grad(w, X, T, want_y=True)
    Yn = _forward(w, X)
    np.testing.assert_allclose(Y, Yn, rtol=1e-6, atol=1e-6)
    assert L == pytest.approx(((Yn - T) ** 2).mean(), rel=1e-12)


def test_mlp_gradient_central_differences(oracle):
    rng = np.random.default_rng(2)
    w = _weights(rng)
    X = (np.round(rng.normal(size=(6, NI)) * 16) / 16).astype(np.float32)
    T = (np.round(rng.normal(size=(6, NO)) * 16) / 16).astype(np.float32)
    _, g = oracle.mlp_loss_grad(w, X, T)
    h = 2.0 ** -12
    picks = [OB3, OB3 + 2, OB2 + 5, OB1 + 7, OW3 + 11, OW2 + 300, OW1 + 77, OW2 + 9000]
    checked = 0
    for p in picks:
        hp = h if p >= OB3 or (OB1 <= p < OW2) or (OB2 <= p < OW3) else 2.0 ** -6
        wp, wm = w.copy(), w.copy()
        wp[p] += hp
        wm[p] -= hp
        fd = (oracle.mlp_loss_grad(wp, X, T)[0] - oracle.mlp_loss_grad(wm, X, T)[0]) / (2 * hp)
        if abs(fd - g[p]) <= 1e-9 + 1e-7 * abs(g[p]):
            checked += 1
    # a perturbation that crosses a ReLU kink breaks the piecewise-quadratic
    # identity; nearly all picks must match exactly
    assert checked >= len(picks) - 1
