"""CPU reference (float64) of the residency MLP: 12 -> 64 -> 64 -> 1, ReLU,
weighted logistic loss averaged over the batch, ReLU'(0) = 0.  Test
infrastructure: the checker of the device forward / backward kernels."""
import numpy as np

F, H = 12, 64
OW1, OB1 = 0, H * F
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H
P = OB3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OW1:OB1].reshape(H, F), p[OB1:OW2], p[OW2:OB2].reshape(H, H), p[OB2:OW3],
            p[OW3:OB3], p[OB3])


def logits(p, x):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ W1.T + b1
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0.0)
    return a2 @ w3 + b3, (z1, a1, z2, a2)


def loss_and_grad(p, x, y, wt=None, w_pos=1.0, w_neg=1.0):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = y.size
    sw = (np.ones(n) if wt is None else np.asarray(wt, np.float64)) * np.where(y > 0.5, w_pos,
                                                                              w_neg)
    z3, (z1, a1, z2, a2) = logits(p, x)
    loss = np.sum(sw * (np.maximum(z3, 0) - z3 * y + np.log1p(np.exp(-np.abs(z3))))) / n
    d3 = sw * (1.0 / (1.0 + np.exp(-z3)) - y) / n
    g = np.zeros(P)
    g[OW3:OB3] = a2.T @ d3
    g[OB3] = d3.sum()
    d2 = np.outer(d3, w3) * (z2 > 0)
    g[OW2:OB2] = (d2.T @ a1).ravel()
    g[OB2:OW3] = d2.sum(0)
    d1 = (d2 @ W2) * (z1 > 0)
    g[OW1:OB1] = (d1.T @ x).ravel()
    g[OB1:OW2] = d1.sum(0)
    return loss, g


def prob(p, x):
    return 1.0 / (1.0 + np.exp(-logits(p, x)[0]))
