"""CPU restatement of the learned absorbing layer (test infrastructure only).

Used by tests/ as the checker for the library's layer kernels; nothing in the
library imports it.  The learned layer has no counterpart in the reference
(which absorbs with the random-velocity shell of src/boundary.c:7-104), so
this file is pinned by mathematics instead of reference vectors:
  * `gradient(..., fp16=False)` against central finite differences of
    `loss(..., fp16=False)` in float64 (tests/test_layer_oracle.py);
  * `features` against closed forms on fields that are polynomials of the
    coordinates (orientation, normal sign, normalisation, depth one-hot);
  * `forward` of a network built to be the identity / a fixed linear map.

Conventions (the library's): inputs x[64] = 54 patch time differences (p's
27 then q's 27, patch index (i+1)*9 + (j+1)*3 + (k+1) with i along the
outward normal of the nearest face, j and k along the two other axes in
increasing order) divided by their largest magnitude s, a one-hot of the
layer depth at 54 + depth (width <= 8), and the Courant numbers
sqrt(v2px) * c, sqrt(v2pz) * c (c = dt * sqrt(1/dx^2) in float32) at 62, 63.  Network 64 -> 64 -> 64 -> 2 with ReLU;
parameters W1 [64][64], b1, W2 [64][64], b2, W3 [2][64], b3.
"""
from __future__ import annotations

import numpy as np

NIN, NHID, NOUT, NPATCH = 64, 64, 2, 54
NPARAM = NHID * NIN + NHID + NHID * NHID + NHID + NOUT * NHID + NOUT
STENCIL = 4


def unpack(params):
    p = np.asarray(params, np.float64)
    o = 0
    out = []
    for shape in ((NHID, NIN), (NHID,), (NHID, NHID), (NHID,), (NOUT, NHID), (NOUT,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    return out


def _h(a, fp16):
    """Round to half precision (the GPU's MFMA operands) when fp16."""
    return np.asarray(a, np.float16).astype(np.float64) if fp16 else np.asarray(a, np.float64)


def forward(params, x, fp16=True):
    """y (n, 2) and the intermediate values.  With fp16 the inputs, weights
    and hidden activations are rounded to half as the MFMA operands are;
    accumulation and biases stay in wide precision."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    xh = _h(x, fp16)
    z1 = xh @ _h(W1, fp16).T + b1
    h1 = _h(np.maximum(z1, 0.0), fp16)
    z2 = h1 @ _h(W2, fp16).T + b2
    h2 = _h(np.maximum(z2, 0.0), fp16)
    y = h2 @ _h(W3, fp16).T + b3
    return y, (xh, z1, h1, z2, h2)


FREE, DAMPING = 0, 1
CENTRE = (13, 40)   # inputs of the cell's own p and q time differences


def predict(params, x, fp16=True, mode=FREE, sigma_max=0.0):
    """The correction divided by s: y (free form) or
    -sigma_max * sigmoid(y) * x_c (damping form)."""
    y, _ = forward(params, x, fp16)
    if mode == FREE:
        return y
    xc = _h(np.asarray(x)[:, CENTRE], fp16)
    return -sigma_max / (1.0 + np.exp(-y)) * xc


def loss(params, x, t, fp16=True, mode=FREE, sigma_max=0.0):
    """Mean over samples of (u0 - t0)^2 + (u1 - t1)^2, u = predict(...)."""
    u = predict(params, x, fp16, mode, sigma_max)
    return float(np.mean(np.sum((u - np.asarray(t, np.float64)) ** 2, axis=1)))


def gradient(params, x, t, fp16=True, mode=FREE, sigma_max=0.0):
    """d loss / d params, flattened in parameter order.  With fp16 the
    backward operands are rounded to half where the GPU's MFMA operands are
    (dL/dy, dZ2, dZ1) while the bias gradients sum the unrounded values."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    y, (xh, z1, h1, z2, h2) = forward(params, x, fp16)
    n = y.shape[0]
    t = np.asarray(t, np.float64)
    if mode == FREE:
        e2 = 2.0 * (y - t)
    else:
        xc = _h(np.asarray(x)[:, CENTRE], fp16)
        sg = 1.0 / (1.0 + np.exp(-y))
        e = -sigma_max * sg * xc - t
        e2 = 2.0 * e * (-sigma_max * xc) * sg * (1.0 - sg)
    dy = _h(e2, fp16)
    dh2 = dy @ _h(W3, fp16)
    dz2f = dh2 * (z2 > 0)
    dz2 = _h(dz2f, fp16)
    dh1 = dz2 @ _h(W2, fp16)
    dz1f = dh1 * (z1 > 0)
    dz1 = _h(dz1f, fp16)
    g = [dz1.T @ xh / n, dz1f.sum(0) / n, dz2.T @ h1 / n, dz2f.sum(0) / n,
         dy.T @ h2 / n, e2.sum(0) / n]
    return np.concatenate([a.reshape(-1) for a in g])


def cell_frame(c, n):
    """Nearest face of compute cell c (lowest axis, low side first on
    ties): (axis, outward sign, tangent axes, depth)."""
    best, axis, sign = 1 << 30, 0, -1
    for a in range(3):
        lo, hi = c[a], n[a] - 1 - c[a]
        if lo < best:
            best, axis, sign = lo, a, -1
        if hi < best:
            best, axis, sign = hi, a, 1
    tang = [a for a in range(3) if a != axis]
    return axis, sign, tang, best


def courant_factor(dt, dx):
    """dt * sqrt(1/dx^2) in float32, the factor of the Courant inputs."""
    f = np.float32
    return f(dt) * np.sqrt(f(1.0) / (f(dx) * f(dx)), dtype=np.float32)


def features(pc, pp, qc, qp, c, v2px, v2pz, cdt):
    """(x[64] float32, s) of compute cell c = (cx, cy, cz) of arrays shaped
    (sz, sy, sx) with the 4-cell frame; float32 arithmetic as the GPU."""
    sz, sy, sx = pc.shape
    n = (sx - 2 * STENCIL, sy - 2 * STENCIL, sz - 2 * STENCIL)
    axis, sign, (t1, t2), depth = cell_frame(c, n)
    raw = np.zeros(NIN, np.float32)
    for q, (cur, prev) in enumerate(((pc, pp), (qc, qp))):
        for i in (-1, 0, 1):
            for j in (-1, 0, 1):
                for k in (-1, 0, 1):
                    pos = [c[0] + STENCIL, c[1] + STENCIL, c[2] + STENCIL]
                    pos[axis] += i * sign
                    pos[t1] += j
                    pos[t2] += k
                    idx = 27 * q + (i + 1) * 9 + (j + 1) * 3 + (k + 1)
                    raw[idx] = np.float32(cur[pos[2], pos[1], pos[0]]) - \
                        np.float32(prev[pos[2], pos[1], pos[0]])
    s = np.float32(np.max(np.abs(raw[:NPATCH])))
    x = np.zeros(NIN, np.float32)
    if s > 0:
        x[:NPATCH] = raw[:NPATCH] / s
    x[NPATCH + depth] = 1.0
    cz, cy, cx = c[2] + STENCIL, c[1] + STENCIL, c[0] + STENCIL
    x[62] = np.sqrt(np.float32(v2px[cz, cy, cx]), dtype=np.float32) * np.float32(cdt)
    x[63] = np.sqrt(np.float32(v2pz[cz, cy, cx]), dtype=np.float32) * np.float32(cdt)
    return x, s


def layer_cells(n, width):
    """Compute coordinates of the layer cells, x fastest, then y, then z."""
    out = []
    for z in range(n[2]):
        for y in range(n[1]):
            for x in range(n[0]):
                if min(x, y, z, n[0] - 1 - x, n[1] - 1 - y, n[2] - 1 - z) < width:
                    out.append((x, y, z))
    return out


def unpack_code(code):
    code = int(code)
    return (code & 1023, (code >> 10) & 1023, (code >> 20) & 1023)


def taper(depth, width):
    """Tapered target state factor at layer depth d:
    sin^2(pi (d + 1) / (2 (width + 1)))."""
    return np.sin(np.pi * (depth + 1) / (2.0 * (width + 1))) ** 2
