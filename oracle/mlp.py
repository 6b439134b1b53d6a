"""TEST INFRASTRUCTURE ONLY -- CPU definition of the surrogate network.

The learned surrogate has no counterpart in the reference; this module is
the build's own float64 statement of the network the GPU kernels train and
run (optika_amd/csrc/optk_mlp.hip): 5 -> 256 -> 256 -> 256 -> 256 -> 3,
SiLU after every hidden layer, a weighted mean-square loss, and its
gradients by hand-written backpropagation.  ``bf16=True`` rounds weights and
activations to bfloat16 (round to nearest even) at the points where the
kernels hold them in bfloat16, so that the GPU gradients can be compared
tightly; ``bf16=False`` is the exact float64 network.

Pinned by central finite differences of the loss (tests/test_mlp_oracle.py):
every parameter block's analytic gradient against the numerical one.
Imported only by tests/ and never by the package.
"""

from __future__ import annotations

import numpy as np

WIDTHS = (5, 256, 256, 256, 256, 3)


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float64 values to bfloat16 (through float32, nearest even)."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    out = (u.astype(np.uint32) << 16).view(np.float32)
    return out.astype(np.float64)


def param_shapes(widths=WIDTHS):
    shapes = []
    for a, b in zip(widths[:-1], widths[1:]):
        shapes.append((b, a))  # W [out][in]
        shapes.append((b,))    # bias
    return shapes


def unflatten(flat: np.ndarray, widths=WIDTHS):
    out, o = [], 0
    for s in param_shapes(widths):
        n = int(np.prod(s))
        out.append(np.asarray(flat[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def silu(x):
    return x / (1.0 + np.exp(-x))


def silu_grad(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def forward(params, x, bf16=False):
    """Returns (y, cache).  params: flat vector; x: [rows][5]."""
    P = unflatten(params)
    L = len(P) // 2
    r = bf16_round if bf16 else (lambda v: v)
    acts, pres = [np.asarray(x, np.float64)], []
    h = acts[0]
    for layer in range(L):
        W, b = P[2 * layer], P[2 * layer + 1]
        # hidden GEMM layers use bf16 weights and bf16 inputs; the first and
        # last layers keep float32 weights
        if 0 < layer < L - 1:
            W = r(W)
        pre = h @ W.T + b
        if layer == L - 1:
            return pre, (P, acts, pres)
        pres.append(pre)
        h = r(silu(pre))
        acts.append(h)
    raise AssertionError("unreachable")


def loss_and_grad(params, x, target, weight, bf16=False):
    """Weighted MSE loss mean_r sum_j w_j (y - t)^2 and its gradient (flat,
    the parameter layout).  With ``bf16``, the hidden GEMM layers consume
    bf16-rounded gradients for their weight gradients and back-propagation
    (as the kernels do) while bias sums use the unrounded ones."""
    y, (P, acts, pres) = forward(params, x, bf16)
    rows = y.shape[0]
    d = y - target
    loss = float((weight * d * d).sum() / rows)
    g32 = 2.0 * weight * d / rows
    r = bf16_round if bf16 else (lambda v: v)
    L = len(P) // 2
    grads = [None] * len(P)
    for layer in range(L - 1, -1, -1):
        gemm = 0 < layer < L - 1
        g_use = r(g32) if gemm else g32
        grads[2 * layer] = g_use.T @ acts[layer]
        grads[2 * layer + 1] = g32.sum(0)
        if layer == 0:
            break
        W = P[2 * layer]
        dact = g_use @ (r(W) if gemm else W)
        g32 = dact * silu_grad(pres[layer - 1])
    return loss, np.concatenate([v.ravel() for v in grads])


def init_params(rng: np.random.Generator, widths=WIDTHS) -> np.ndarray:
    """He-uniform weights for SiLU layers, zero biases."""
    parts = []
    for a, b in zip(widths[:-1], widths[1:]):
        bound = np.sqrt(6.0 / a)
        parts.append(rng.uniform(-bound, bound, (b, a)).ravel())
        parts.append(np.zeros(b))
    return np.concatenate(parts)


def poly_features(u, exponents):
    """Monomials prod_k u[:, k] ** e[k] of float64 inputs u [n][5], one
    column per exponent tuple (the surrogate's polynomial baseline)."""
    u = np.asarray(u, dtype=np.float64)
    cols = []
    for e in exponents:
        c = np.ones(u.shape[0])
        for k, p in enumerate(e):
            if p:
                c = c * u[:, k] ** p
        cols.append(c)
    return np.stack(cols, 1)
