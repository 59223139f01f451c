"""float64 numpy restatement of the cost-model network (test infrastructure).

Layout and semantics follow include/uhs_tune.h: parameters W1[20][128], b1,
W2[128][128], b2, W3[128][128], b3, W4[128], b4; inputs standardised as
(x - mean) / scale, target as (t - tmean) / tscale; ReLU on the three hidden
layers; loss 0.5 * mean((y - t)^2) in standardised units."""
import numpy as np

IN, H = 20, 128
SIZES = [(IN, H), (H, H), (H, H), (H, 1)]


def unpack(p):
    p = np.asarray(p, np.float64)
    out, k = [], 0
    for i, o in SIZES:
        W = p[k:k + i * o].reshape(i, o); k += i * o
        b = p[k:k + o]; k += o
        out.append((W, b))
    assert k == len(p)
    return out


def forward(p, xs):
    acts = [xs]
    h = xs
    for l, (W, b) in enumerate(unpack(p)):
        h = h @ W + b
        if l < 3:
            h = np.maximum(h, 0.0)
        acts.append(h)
    return acts


def loss_grad(p, xs, ts):
    layers = unpack(p)
    acts = forward(p, xs)
    y = acts[-1][:, 0]
    n = len(ts)
    loss = 0.5 * np.mean((y - ts) ** 2)
    d = ((y - ts) / n)[:, None]
    grads = []
    for l in range(3, -1, -1):
        W, b = layers[l]
        grads.append((acts[l].T @ d, d.sum(0)))
        if l:
            d = (d @ W.T) * (acts[l] > 0)
    grads.reverse()
    return loss, np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def adam(p, g, m, v, lr, t, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
