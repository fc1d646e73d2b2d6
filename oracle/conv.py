"""FP64 NumPy reference of the 3x3x3 convolutions of the learned preconditioner.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). No reference counterpart
exists (the reference has no learned preconditioner); these restate the definitions:

  y[b, x, co]  = bias[co] + sum_{t, ci} W[co, t, ci] x[b, x + o_t, ci]   (zero padding)
  dx[b, x, ci] = sum_{t, co} W[co, t, ci] dy[b, x - o_t, co]
  dW[co, t, ci] = sum_{b, x} dy[b, x, co] x[b, x + o_t, ci],  db[co] = sum_{b,x} dy[b,x,co]

Layout: activations [B, D, H, W, C] (channels last), weights [Cout, 27, Cin],
tap t = (dz+1)*9 + (dy+1)*3 + (dx+1) with offsets o_t = (dz, dy, dx) in (D, H, W).
"""
from __future__ import annotations

import numpy as np

TAPS = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]


def _shifted(x, o):
    """x[b, p + o] with zero padding, for all positions p."""
    B, D, H, W, C = x.shape
    out = np.zeros_like(x)
    dz, dy, dx = o
    src = (slice(max(dz, 0), D + min(dz, 0)), slice(max(dy, 0), H + min(dy, 0)), slice(max(dx, 0), W + min(dx, 0)))
    dst = (slice(max(-dz, 0), D + min(-dz, 0)), slice(max(-dy, 0), H + min(-dy, 0)), slice(max(-dx, 0), W + min(-dx, 0)))
    out[(slice(None),) + dst] = x[(slice(None),) + src]
    return out


def conv_fwd(x, W, bias=None):
    x = np.asarray(x, np.float64)
    W = np.asarray(W, np.float64)
    y = np.zeros(x.shape[:4] + (W.shape[0],))
    for t, o in enumerate(TAPS):
        y += np.einsum("bdhwc,oc->bdhwo", _shifted(x, o), W[:, t, :])
    if bias is not None:
        y += np.asarray(bias, np.float64)
    return y


def conv_bwd_data(dy, W):
    dy = np.asarray(dy, np.float64)
    W = np.asarray(W, np.float64)
    dx = np.zeros(dy.shape[:4] + (W.shape[2],))
    for t, o in enumerate(TAPS):
        dx += np.einsum("bdhwo,oc->bdhwc", _shifted(dy, tuple(-v for v in o)), W[:, t, :])
    return dx


def conv_bwd_weight(x, dy):
    x = np.asarray(x, np.float64)
    dy = np.asarray(dy, np.float64)
    dW = np.zeros((dy.shape[-1], 27, x.shape[-1]))
    for t, o in enumerate(TAPS):
        dW[:, t, :] = np.einsum("bdhwo,bdhwc->oc", dy, _shifted(x, o))
    return dW, dy.sum(axis=(0, 1, 2, 3))
