"""FP64 numpy reference of the learned selector's network and loss --
TEST INFRASTRUCTURE ONLY (checks the GPU training kernels' gradients).

Network: x[16] -> relu(W1 x + b1)[64] -> relu(W2 h1 + b2)[64] -> W3 h2 + b3 [3]
Loss: (1/n) sum_i sum_m mask[i,m] * (o[i,m] - y[i,m])^2
Packed layout (floats): W1 [16 in][64 out], b1[64], W2 [64 in][64 out],
b2[64], W3 [3][64], b3[3] -- 5,443 values.
"""
from __future__ import annotations

import numpy as np

F, H, O = 16, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + F * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + O * H
N_PARAMS = OFF_B3 + O


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[OFF_W1:OFF_B1].reshape(F, H), w[OFF_B1:OFF_W2], w[OFF_W2:OFF_B2].reshape(H, H),
            w[OFF_B2:OFF_W3], w[OFF_W3:OFF_B3].reshape(O, H), w[OFF_B3:N_PARAMS])


def forward(x, w):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h2 @ W3.T + b3


def loss_and_grad(x, y, mask, w):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    mask = np.asarray(mask, dtype=np.float64)
    n = x.shape[0]
    a1 = x @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    o = h2 @ W3.T + b3
    e = o - y
    loss = float((mask * e * e).sum() / n)
    do = 2.0 * mask * e / n              # [n,3]
    gW3 = do.T @ h2                      # [3,64]
    gb3 = do.sum(0)
    dh2 = (do @ W3) * (a2 > 0)           # [n,64]
    gW2 = h1.T @ dh2                     # [64 in, 64 out]
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (a1 > 0)
    gW1 = x.T @ dh1                      # [16, 64]
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
