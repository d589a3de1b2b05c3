"""CPU restatement of the spin-up surrogate's arithmetic (numpy).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of libsedsurr.so.
No reference counterpart exists (the reference has no learned component);
each function is pinned by the tests against a library routine or the
mathematics:
  to_bf16        torch's float32 -> bfloat16 conversion (round to nearest even)
  gelu, gelu'    torch.nn.functional.gelu(approximate='tanh') and autograd
  forward/grad   central finite differences in float64 (rounding off)
  adam_step      torch.optim.Adam

The network: inputs x (n x n_in); hidden layers h_l = GELU(h_{l-1} W_l + b_l);
output y = h_L W_out + b_out; loss = mean over rows and outputs of
(y - t)^2.  With rounding on, the stored tensors are rounded to bfloat16 at
the points where the kernels store them (inputs, weights, activations,
output-gradient and hidden-layer gradients); products accumulate in float64
here (float32 on the GPU), so results agree to a tolerance, not bitwise.
"""
from __future__ import annotations

import numpy as np

GC = np.sqrt(2.0 / np.pi)
GA = 0.044715


def to_bf16(a):
    """float32 values rounded to the nearest bfloat16 (ties to even),
    returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32)
    return (r << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GC * (z + GA * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GC * (z + GA * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GC * (1.0 + 3.0 * GA * z * z)


def unflatten(params, n_in, n_hidden, n_layers, n_out):
    """Flat vector (W_l row-major fan_in x fan_out, then b_l, per layer) ->
    list of (W, b)."""
    p = np.asarray(params, dtype=np.float64)
    dims = [n_in] + [n_hidden] * n_layers + [n_out]
    out, o = [], 0
    for fi, fo in zip(dims[:-1], dims[1:]):
        W = p[o:o + fi * fo].reshape(fi, fo)
        o += fi * fo
        b = p[o:o + fo]
        o += fo
        out.append((W, b))
    assert o == p.size
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def num_params(n_in, n_hidden, n_layers, n_out):
    dims = [n_in] + [n_hidden] * n_layers + [n_out]
    return sum(fi * fo + fo for fi, fo in zip(dims[:-1], dims[1:]))


def loss_grad(params, X, T, shape, rounding=True):
    """(loss, gradient flat vector, output) of one batch."""
    n_in, n_hidden, n_layers, n_out = shape
    r = to_bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unflatten(params, *shape)
    Wr = [r(W) for W, _ in layers]
    bs = [np.asarray(b, dtype=np.float32).astype(np.float64) if rounding else b for _, b in layers]
    h = r(X)
    hs, zs = [h], []
    for l in range(n_layers):
        z = h @ Wr[l] + bs[l]
        if rounding:
            z = z.astype(np.float32).astype(np.float64)
        zs.append(z)
        h = r(gelu(z))
        hs.append(h)
    y = h @ Wr[-1] + bs[-1]
    if rounding:
        y = y.astype(np.float32).astype(np.float64)
    n = X.shape[0]
    d = y - T
    loss = float(np.mean(d * d))
    dz = r(2.0 * d / (n * n_out))
    grads = [None] * (n_layers + 1)
    for l in range(n_layers, -1, -1):
        gW = hs[l].T @ dz
        gb = dz.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            dz = r((dz @ Wr[l].T) * gelu_grad(zs[l - 1]))
    return loss, flatten(grads), y


def forward(params, X, shape, rounding=True):
    n_in, n_hidden, n_layers, n_out = shape
    return loss_grad(params, X, np.zeros((X.shape[0], n_out)), shape, rounding)[2]


def adam_step(w, g, m, v, t, lr, b1, b2, eps):
    """One Adam step (Kingma & Ba 2015, as torch.optim.Adam without weight
    decay): returns new (w, m, v); t is the 1-based step count."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
