"""TEST INFRASTRUCTURE ONLY. NOT PART OF THE PRODUCT.

CPU FP64 restatement of the histogram surrogate network of
irs_amd/csrc/surrogate.hip, the checker for its GPU forward and backward
passes and its Adam step. Network: 6 inputs zero-padded to 16 -> 256 -> 256 -> 256 -> 128,
SiLU hidden activations, softmax output, loss = mean over rows of
-sum_j t_j log p_j. Parameters are flat, layer by layer: W [out x in]
row-major, then b [out].

With bf16=True the operands of every matrix product are rounded to BF16
(round to nearest even) exactly where the GPU rounds them:
  * the inputs, the weights and the hidden activations in the forward pass;
  * the output gradient (p - t) / M and the hidden-layer gradients dZ in
    the backward pass.
The bias gradients use the unrounded dZ, as on the GPU.

The backward pass is pinned by central finite differences of the loss in
FP64 (tests/test_surrogate.py). The Adam step is pinned by closed forms of
the published algorithm: the first step moves every weight with a non-zero
gradient by lr against the gradient's sign; a constant gradient moves it by
lr g / (|g| + eps) on every step (the bias corrections make the moment
estimates exact); with a zero gradient the weights shrink by (1 - lr wd) per
step and the biases stay.
"""

import numpy as np

DIMS = (16, 256, 256, 256, 128)
NIN = 6


def bf16(x):
    """Round float64 values to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout():
    """(w_offset, b_offset) per layer and the total parameter count."""
    offs, off = [], 0
    for l in range(4):
        w = off
        off += DIMS[l + 1] * DIMS[l]
        b = off
        off += DIMS[l + 1]
        offs.append((w, b))
    return offs, off


def unpack(params):
    offs, _ = layout()
    Ws, bs = [], []
    for l, (w, b) in enumerate(offs):
        Ws.append(params[w:w + DIMS[l + 1] * DIMS[l]].reshape(DIMS[l + 1], DIMS[l]).astype(np.float64))
        bs.append(params[b:b + DIMS[l + 1]].astype(np.float64))
    return Ws, bs


def pad_inputs(x):
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros((x.shape[0], DIMS[0]))
    out[:, :NIN] = x
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, bf16_ops=False):
    """Softmax probabilities and the cached pre-activations / activations."""
    r = bf16 if bf16_ops else (lambda a: a)
    Ws, bs = unpack(params)
    h = r(pad_inputs(x))
    hs, zs = [h], []
    for l in range(4):
        z = h @ r(Ws[l]).T + bs[l]
        zs.append(z)
        if l < 3:
            h = r(silu(z))
            hs.append(h)
    z = zs[-1]
    e = np.exp(z - z.max(axis=1, keepdims=True))
    p = e / e.sum(axis=1, keepdims=True)
    return p, zs, hs


def loss(params, x, t, bf16_ops=False):
    p, _, _ = forward(params, x, bf16_ops)
    return float(np.mean(-np.sum(np.asarray(t) * np.log(p), axis=1)))


def gradients(params, x, t, bf16_ops=False):
    """(loss, flat gradient) of the mean cross-entropy."""
    r = bf16 if bf16_ops else (lambda a: a)
    t = np.asarray(t, dtype=np.float64)
    p, zs, hs = forward(params, x, bf16_ops)
    M = x.shape[0]
    L = float(np.mean(-np.sum(t * np.log(p), axis=1)))
    Ws, _ = unpack(params)
    offs, n = layout()
    g = np.zeros(n)
    d = (p - t) / M
    for l in range(3, -1, -1):
        dr = r(d)
        gw = dr.T @ hs[l]
        g[offs[l][0]:offs[l][0] + gw.size] = gw.ravel()
        g[offs[l][1]:offs[l][1] + d.shape[1]] = d.sum(axis=0)
        if l == 0:
            break
        dh = dr @ r(Ws[l])
        d = dh * dsilu(zs[l - 1])
    return L, g


def decay_mask():
    """1 for the weight-matrix entries, 0 for the biases: the GPU applies the
    decoupled weight decay to the weight matrices only."""
    offs, n = layout()
    mask = np.zeros(n)
    for l, (w, b) in enumerate(offs):
        mask[w:b] = 1.0
    return mask


def adam_step(params, grads, m, v, step, lr, wd, b1=0.9, b2=0.999, eps=1e-8):
    """One step of Adam (Kingma & Ba 2015, Algorithm 1) with decoupled weight
    decay (Loshchilov & Hutter 2019, AdamW) in FP64. `step` counts from 1.
    Returns (params, m, v); the inputs are not modified.

        m <- b1 m + (1 - b1) g,   v <- b2 v + (1 - b2) g^2
        w <- w (1 - lr wd [weights only]) - lr (m / (1 - b1^step)) / (sqrt(v / (1 - b2^step)) + eps)
    """
    g = np.asarray(grads, dtype=np.float64)
    m = b1 * np.asarray(m, dtype=np.float64) + (1.0 - b1) * g
    v = b2 * np.asarray(v, dtype=np.float64) + (1.0 - b2) * g * g
    mhat = m / (1.0 - b1 ** step)
    vhat = v / (1.0 - b2 ** step)
    w = np.asarray(params, dtype=np.float64) * (1.0 - lr * wd * decay_mask())
    w = w - lr * mhat / (np.sqrt(vhat) + eps)
    return w, m, v
