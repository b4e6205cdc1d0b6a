"""Independent torch (CPU, float64) statement of the U-Net, used to pin the
NumPy oracle's forward pass and to give full reference gradients.  Down- and
up-sampling are written as fixed-weight strided convolution / transposed
convolution (a different construction from the oracle's explicit loops)."""
from __future__ import annotations

import numpy as np


def _torch():
    import torch
    return torch


def torch_forward(theta, feats, nx, ny, requires_grad=False):
    torch = _torch()
    F = torch.nn.functional
    from oracle.unet_ref import SPEC
    th = torch.tensor(np.asarray(theta, dtype=np.float64), requires_grad=requires_grad)
    n = feats.shape[0]
    x = torch.zeros((n, 16, ny, nx), dtype=torch.float64)
    x[:, :4] = torch.tensor(np.asarray(feats, dtype=np.float64).reshape(n, ny, nx, 4)).permute(0, 3, 1, 2)
    ws, off = [], 0
    for ci, co in SPEC:
        W = th[off:off + 9 * co * ci].reshape(3, 3, co, ci).permute(2, 3, 0, 1)
        off += 9 * co * ci
        b = th[off:off + co]
        off += co
        ws.append((W, b))
    hw = th[off:off + 32].reshape(2, 16)
    hb = th[off + 32:off + 34]
    k1 = torch.tensor([0.25, 0.5, 0.25], dtype=torch.float64)
    kd = torch.outer(k1, k1)
    ku = torch.outer(2 * k1, 2 * k1)

    def conv(x, l):
        return torch.relu(F.conv2d(x, ws[l][0], ws[l][1], padding=1))

    def down(x):
        c = x.shape[1]
        return F.conv2d(x, kd.expand(c, 1, 3, 3), stride=2, padding=1, groups=c)

    def up(x):
        c = x.shape[1]
        return F.conv_transpose2d(x, ku.expand(c, 1, 3, 3), stride=2, padding=1, groups=c)

    s0 = conv(conv(x, 0), 1)
    s1 = conv(conv(down(s0), 2), 3)
    s2 = conv(conv(down(s1), 4), 5)
    b3 = conv(conv(down(s2), 6), 7)
    b4 = conv(conv(torch.cat([up(b3), s2], 1), 8), 9)
    b5 = conv(conv(torch.cat([up(b4), s1], 1), 10), 11)
    b6 = conv(conv(torch.cat([up(b5), s0], 1), 12), 13)
    y = torch.einsum("nchw,oc->nhwo", b6, hw) + hb
    return y.reshape(n, nx * ny, 2), th


def torch_loss_grad(theta, feats, target, nx, ny):
    torch = _torch()
    y, th = torch_forward(theta, feats, nx, ny, requires_grad=True)
    loss = torch.mean((y - torch.tensor(np.asarray(target, dtype=np.float64))) ** 2)
    loss.backward()
    return float(loss), th.grad.numpy().copy()
