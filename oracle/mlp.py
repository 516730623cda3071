"""CPU oracle of the learned selector's MLP (6 -> 64 -> 64 -> 48, ReLU,
masked softmax cross-entropy, Adam) in float64 numpy.  TEST INFRASTRUCTURE:
used by tests/ to check the HIP forward, backward and Adam kernels.  The
selector has no counterpart in the reference; the pins here are the
mathematics: finite differences of the loss and the closed-form Adam rule."""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 6, 64, 48
SHAPES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, OUT)), ("b3", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([d[name].reshape(-1) for name, _ in SHAPES])


def forward(flat, x, mask=None):
    p = unpack(flat)
    x = np.asarray(x, dtype=np.float64)
    a1 = x @ p["W1"] + p["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ p["W2"] + p["b2"]
    h2 = np.maximum(a2, 0)
    z = h2 @ p["W3"] + p["b3"]
    if mask is not None:
        z = np.where(mask.astype(bool), z, -np.inf)
    z = z - z.max(axis=1, keepdims=True)
    e = np.exp(z)
    prob = e / e.sum(axis=1, keepdims=True)
    return prob, (x, h1, h2)


def loss_and_grad(flat, x, y, mask=None):
    prob, (x, h1, h2) = forward(flat, x, mask)
    p = unpack(flat)
    B = x.shape[0]
    loss = -np.mean(np.log(np.maximum(prob[np.arange(B), y], 1e-300)))
    dz = prob.copy()
    dz[np.arange(B), y] -= 1
    dz /= B
    g = {"W3": h2.T @ dz, "b3": dz.sum(0)}
    dh2 = (dz @ p["W3"].T) * (h2 > 0)
    g["W2"] = h1.T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ p["W2"].T) * (h1 > 0)
    g["W1"] = x.T @ dh1
    g["b1"] = dh1.sum(0)
    return loss, pack(g)


def adam(flat, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return flat - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0):
    """He-uniform initialisation (seeded)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shape[0])
            d[name] = rng.uniform(-lim, lim, shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)
