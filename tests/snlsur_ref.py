"""Float32 reference of the W3SNL1 surrogate network for its tests (torch on the CPU).

The activation is GELU in its tanh form.  The device network rounds its weights, its input features and every hidden activation to
bfloat16 (round to nearest even) and accumulates in fp32; the reference rounds at the same
places in the forward pass (a straight-through rounding, so autograd sees the identity) and
runs the backward pass in fp32.  The device also rounds the backward's intermediate
gradients to bfloat16, which the reference does not: comparisons use a tolerance.
"""
from __future__ import annotations

import numpy as np


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 to the nearest bfloat16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).copy()
    nan = np.isnan(x)
    out[nan] = np.nan
    return out


def _round(t):
    import torch
    return t + (t.to(torch.bfloat16).to(torch.float32) - t).detach()


def forward_backward(layers, x: np.ndarray, y: np.ndarray):
    """Per-point loss |o-y|^2/|y|^2, the network output and the gradients of the mean loss
    [(dW1, db1), ..., (dW4, db4)] for bf16-rounded weights and inputs."""
    import torch
    params = []
    for w, b in layers:
        params.append((torch.tensor(bf16(w), requires_grad=True), torch.tensor(np.asarray(b, np.float32),
                                                                              requires_grad=True)))
    h = torch.tensor(bf16(x))
    for i, (w, b) in enumerate(params):
        z = h @ w.T + b
        h = _round(torch.nn.functional.gelu(z, approximate="tanh")) if i < 3 else z
    o = h
    yt = torch.tensor(np.asarray(y, np.float32))
    yy = (yt * yt).sum(1)
    wgt = torch.where(yy > 0, 1.0 / yy, torch.zeros_like(yy))
    loss = ((o - yt) ** 2).sum(1) * wgt
    loss.mean().backward()
    grads = [(w.grad.numpy(), b.grad.numpy()) for w, b in params]
    return loss.detach().numpy(), o.detach().numpy(), grads
