"""CPU oracle for the learned coarse-grid correction network (float64 NumPy).

TEST INFRASTRUCTURE ONLY (the checker for the GPU U-Net; never called by the
product path).  The network has no counterpart in Helmholtz.jl; this is the
textbook definition of the architecture the GPU implements:

* 3x3 convolutions with zero padding on a node grid (x fastest), taps
  ordered t = (dj + 1) * 3 + (di + 1), weights W[t][co][ci], bias b[co], ReLU;
* encoder 16 -> 16 -> (down) 32 -> 32 -> (down) 64 -> 64 -> (down) 64 -> 64,
  decoder (up, concat[up, skip]) 128 -> 64 -> 64, 96 -> 32 -> 32, 48 -> 16 -> 16,
  1x1 head 16 -> 2 (no activation);
* down = nodal full weighting ([1 2 1]/4 per axis, stride 2, zero padding),
  up = nodal bilinear interpolation;
* input: 4 feature channels zero-padded to 16.

Parameter vector layout (matches hh_unet_get_params): per conv layer W then
b, then the head's W[2][16] and b[2].
"""
from __future__ import annotations

import numpy as np

SPEC = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 64), (64, 64), (128, 64), (64, 64),
        (96, 32), (32, 32), (48, 16), (16, 16)]


def num_params() -> int:
    return sum(9 * co * ci + co for ci, co in SPEC) + 34


def unpack(theta):
    layers, off = [], 0
    for ci, co in SPEC:
        W = theta[off:off + 9 * co * ci].reshape(9, co, ci)
        off += 9 * co * ci
        b = theta[off:off + co]
        off += co
        layers.append((W, b))
    hw = theta[off:off + 32].reshape(2, 16)
    hb = theta[off + 32:off + 34]
    return layers, hw, hb


def conv3x3(x, W, b, relu=True):
    """x: [n, ny, nx, ci] -> [n, ny, nx, co]."""
    n, ny, nx, ci = x.shape
    xp = np.zeros((n, ny + 2, nx + 2, ci))
    xp[:, 1:-1, 1:-1, :] = x
    out = np.broadcast_to(b, (n, ny, nx, len(b))).copy()
    for t in range(9):
        di, dj = t % 3 - 1, t // 3 - 1
        sl = xp[:, 1 + dj:1 + dj + ny, 1 + di:1 + di + nx, :]
        out += sl @ W[t].T
    return np.maximum(out, 0.0) if relu else out


def down(x):
    n, ny, nx, c = x.shape
    ncy, ncx = (ny + 1) // 2, (nx + 1) // 2
    w = {-1: 0.25, 0: 0.5, 1: 0.25}
    out = np.zeros((n, ncy, ncx, c))
    for J in range(ncy):
        for I in range(ncx):
            s = 0.0
            for b in (-1, 0, 1):
                j = 2 * J + b
                if not 0 <= j < ny:
                    continue
                for a in (-1, 0, 1):
                    i = 2 * I + a
                    if 0 <= i < nx:
                        s = s + w[a] * w[b] * x[:, j, i, :]
            out[:, J, I, :] = s
    return out


def up(x, ny, nx):
    n, ncy, ncx, c = x.shape
    out = np.zeros((n, ny, nx, c))

    def at(i, j):
        return x[:, j, i, :] if (i < ncx and j < ncy) else 0.0

    for j in range(ny):
        for i in range(nx):
            i0, j0 = i // 2, j // 2
            if i % 2 and j % 2:
                out[:, j, i, :] = 0.25 * (at(i0, j0) + at(i0 + 1, j0) + at(i0, j0 + 1) + at(i0 + 1, j0 + 1))
            elif i % 2:
                out[:, j, i, :] = 0.5 * (at(i0, j0) + at(i0 + 1, j0))
            elif j % 2:
                out[:, j, i, :] = 0.5 * (at(i0, j0) + at(i0, j0 + 1))
            else:
                out[:, j, i, :] = at(i0, j0)
    return out


def forward(theta, feats, nx, ny):
    """feats: [n, nx*ny, 4] (pixel index x fastest) -> y: [n, nx*ny, 2]."""
    L, hw, hb = unpack(np.asarray(theta, dtype=np.float64))
    n = feats.shape[0]
    x = np.zeros((n, ny, nx, 16))
    x[..., :4] = np.asarray(feats, dtype=np.float64).reshape(n, ny, nx, 4)
    a0 = conv3x3(x, *L[0]); s0 = conv3x3(a0, *L[1])
    a1 = conv3x3(down(s0), *L[2]); s1 = conv3x3(a1, *L[3])
    a2 = conv3x3(down(s1), *L[4]); s2 = conv3x3(a2, *L[5])
    a3 = conv3x3(down(s2), *L[6]); b3 = conv3x3(a3, *L[7])
    c2 = np.concatenate([up(b3, *s2.shape[1:3]), s2], axis=-1)
    b4 = conv3x3(conv3x3(c2, *L[8]), *L[9])
    c1 = np.concatenate([up(b4, *s1.shape[1:3]), s1], axis=-1)
    b5 = conv3x3(conv3x3(c1, *L[10]), *L[11])
    c0 = np.concatenate([up(b5, *s0.shape[1:3]), s0], axis=-1)
    b6 = conv3x3(conv3x3(c0, *L[12]), *L[13])
    y = b6 @ hw.T + hb
    return y.reshape(n, nx * ny, 2)


def mse(theta, feats, target, nx, ny):
    y = forward(theta, feats, nx, ny)
    return float(np.mean((y - target) ** 2))


def adam_step(theta, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1**t)
    vh = v / (1 - b2**t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)
