"""CPU oracle of the warm-start network (test infrastructure only).

float64 restatement of the forward pass, loss and gradients that the HIP
kernels compute (nushellx_amd/csrc/nsx_mlp.hip): layers n_in -> hidden (SiLU)
-> ... -> n_out (linear), weights rounded to BF16 (round to nearest even) in
every product when ``quantize`` is set, loss L = 1/(2 rows) sum ||y_p - y||^2.
The reference has no such network (its Lanczos starts from a random vector,
src/NuTRL/NuCoreX.f90:299-312), so the oracle is pinned by finite differences
of the loss (unquantised, float64) and by closed forms in tests/.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def layout(widths: Sequence[int]) -> List[Tuple[int, int, int, int]]:
    """(w_off, b_off, n_out, n_in) per layer in the flat parameter vector."""
    out, off = [], 0
    for ni, no in zip(widths[:-1], widths[1:]):
        out.append((off, off + ni * no, no, ni))
        off += ni * no + no
    return out


def n_params(widths: Sequence[int]) -> int:
    w_off, b_off, no, ni = layout(widths)[-1]
    return b_off + no


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def weights(params: np.ndarray, widths: Sequence[int], quantize: bool = True):
    ws = []
    for w_off, b_off, no, ni in layout(widths):
        W = params[w_off:w_off + no * ni].reshape(no, ni)
        if quantize:
            W = bf16(W.astype(np.float32))
        ws.append((np.asarray(W, dtype=np.float64), np.asarray(params[b_off:b_off + no], dtype=np.float64)))
    return ws


def forward(params, widths, X, quantize: bool = True):
    """Returns the output and the (pre-activation, input) pair of every layer."""
    ws = weights(params, widths, quantize)
    a = np.asarray(X, dtype=np.float64)
    cache = []
    for li, (W, b) in enumerate(ws):
        z = a @ W.T + b
        cache.append((z, a))
        a = _silu(z) if li + 1 < len(ws) else z
    return a, cache


def loss_grad(params, widths, X, Y, quantize: bool = True):
    """Loss and its gradient w.r.t. the parameters (w.r.t. the BF16 weight
    values when quantised), in the flat layout."""
    ws = weights(params, widths, quantize)
    yp, cache = forward(params, widths, X, quantize)
    rows = X.shape[0]
    d = (yp - np.asarray(Y, dtype=np.float64)) / rows
    loss = 0.5 * rows * np.sum(d * d)
    g = np.zeros(n_params(widths))
    lay = layout(widths)
    for li in range(len(ws) - 1, -1, -1):
        w_off, b_off, no, ni = lay[li]
        z, a_in = cache[li]
        g[w_off:w_off + no * ni] = (d.T @ a_in).ravel()
        g[b_off:b_off + no] = d.sum(axis=0)
        if li == 0:
            break
        d = (d @ ws[li][0]) * _silu_grad(cache[li - 1][0])
    return loss, g
