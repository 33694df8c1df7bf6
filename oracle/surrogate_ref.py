"""CPU oracle of the BF16 surrogate network (torch, float32 on the CPU).

TEST INFRASTRUCTURE ONLY: the checker for slm_amd/surrogate.py's GPU
forward and backward.  The same U-Net (layer list slm_amd.surrogate.LAYERS)
is written with torch.nn.functional ops in float32, NCHW, and every place
where the GPU stores a bf16 tensor is a round trip through bfloat16: layer
outputs in the forward pass and, through the backward of that cast, the
gradients of those outputs.  So the two differ only by accumulation order
inside the products (fp32 on MFMA against fp32 on the CPU).
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


class RefUNet:
    def __init__(self, params, layers):
        """params: {name: (w [cout, taps, cin], b [cout])} (fp32 numpy);
        weights are used as bf16 values (the GPU computes with the bf16 copy)."""
        self.layers = {name: (cin, cout, taps, relu) for name, cin, cout, taps, relu in layers}
        self.w, self.b = {}, {}
        for name, (cin, cout, taps, _) in self.layers.items():
            w, b = params[name]
            k = 3 if taps == 9 else 1
            wt = torch.from_numpy(np.ascontiguousarray(w)).reshape(cout, k, k, cin).permute(0, 3, 1, 2).contiguous()
            self.w[name] = _bf(wt).requires_grad_(True)
            self.b[name] = torch.from_numpy(np.ascontiguousarray(b)).clone().requires_grad_(True)

    def conv(self, name, x):
        cin, cout, taps, relu = self.layers[name]
        y = F.conv2d(x, self.w[name], self.b[name], padding=1 if taps == 9 else 0)
        if relu:
            y = F.relu(y)
        return _bf(y)

    def forward(self, x_nchw):
        x = _bf(x_nchw)
        e0 = self.conv("e0b", self.conv("e0a", x))
        e1 = self.conv("e1b", self.conv("e1a", F.max_pool2d(e0, 2)))
        e2 = self.conv("e2b", self.conv("e2a", F.max_pool2d(e1, 2)))
        bn = self.conv("bnb", self.conv("bna", F.max_pool2d(e2, 2)))
        up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
        d2 = self.conv("d2b", self.conv("d2a", torch.cat([up(bn), e2], 1)))
        d1 = self.conv("d1b", self.conv("d1a", torch.cat([up(d2), e1], 1)))
        d0 = self.conv("d0b", self.conv("d0a", torch.cat([up(d1), e0], 1)))
        return self.conv("head", d0)

    def loss(self, head, target, valid, w_bce=1.0, w_len=1.0):
        """head [N, 32, H, W]; target [N, H, W, 3]; valid [N, H, W] bool."""
        v = torch.from_numpy(valid)
        t = torch.from_numpy(target)
        a, b, z = head[:, 0][v], head[:, 1][v], head[:, 2][v]
        n = max(1, int(v.sum()))
        l0 = ((a - t[..., 0][v]) ** 2).sum()
        l1 = ((b - t[..., 1][v]) ** 2).sum()
        l2 = F.binary_cross_entropy_with_logits(z, t[..., 2][v], reduction="sum")
        return (w_len * l0 + l1 + w_bce * l2) / n, (l0, l1, l2)

    def grads(self):
        """{name: (dW [cout, taps, cin], db [cout])} as fp32 numpy."""
        out = {}
        for name, (cin, cout, taps, _) in self.layers.items():
            g = self.w[name].grad.permute(0, 2, 3, 1).reshape(cout, taps, cin)
            out[name] = (g.numpy().copy(), self.b[name].grad.numpy().copy())
        return out
