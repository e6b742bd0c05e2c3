"""Pins the CPU MLP reference: its gradient equals central finite differences
of its loss, its forward pass equals a plain per-sample loop, and the
parameter count is the 5057 of a 12 -> 64 -> 64 -> 1 network."""
import numpy as np

import _mlp_ref as ref


def test_parameter_count():
    assert ref.P == 12 * 64 + 64 + 64 * 64 + 64 + 64 + 1 == 5057


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, ref.P)
    x = rng.normal(0, 1, (40, 12))
    y = (rng.random(40) < 0.4).astype(float)
    wt = rng.random(40) + 0.5
    _, g = ref.loss_and_grad(p, x, y, wt, w_pos=2.0, w_neg=0.7)
    eps = 1e-6
    for i in list(rng.choice(ref.P, 60, replace=False)) + [ref.OB3, ref.OB1, ref.OW3]:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        num = (ref.loss_and_grad(pp, x, y, wt, 2.0, 0.7)[0] -
               ref.loss_and_grad(pm, x, y, wt, 2.0, 0.7)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), (i, num, g[i])


def test_forward_matches_loop():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, ref.P)
    x = rng.normal(0, 1, (5, 12))
    W1, b1, W2, b2, w3, b3 = ref.unpack(p)
    for n in range(5):
        h1 = [max(0.0, b1[i] + sum(W1[i, k] * x[n, k] for k in range(12))) for i in range(64)]
        h2 = [max(0.0, b2[j] + sum(W2[j, i] * h1[i] for i in range(64))) for j in range(64)]
        z = b3 + sum(w3[j] * h2[j] for j in range(64))
        assert abs(ref.logits(p, x[n:n + 1])[0][0] - z) < 1e-9
