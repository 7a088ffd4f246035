"""CPU restatement of the learned warm start's network -- TEST INFRASTRUCTURE ONLY.

Only tests/ may use it, as the checker of the HIP training and inference
kernels (fargo_thorin_amd/csrc/thorin_cnn.hip); the product never imports it.
The network has no counterpart in the reference.

A 4-level, 32-channel U-Net on an H x W polar grid (zero padding across the
rings, periodic padding along the sectors), NHWC float arrays:

    a1 = relu(conv(x, 16->32))   a2 = relu(conv(a1))          level 0
    a3 = relu(conv(pool(a2)))                                  level 1
    a4 = relu(conv(pool(a3)))                                  level 2
    a5 = relu(conv(pool(a4)))                                  level 3
    a6 = relu(conv(up(a5) + a4))                               level 2
    a7 = relu(conv(up(a6) + a3))                               level 1
    a8 = relu(conv(up(a7) + a2))                               level 0
    y  = conv(a8, 32->16)[..., 0]                              level 0
    loss = mean over rings 1..H-2 of (y - t)^2

Coarse network (f = 2 or 4): x and the prediction y live on the H/f x W/f
grid; y is interpolated bilinearly to the H x W grid (cell centres, clamped
across the rings, periodic along the sectors), where the residual loss is
taken, and its gradient returns through the transpose of the interpolation.

With bf16=True every stored activation, activation gradient, the input and
the weights used by the convolutions are rounded to BF16 (round to nearest
even) where the GPU stores them; sums are formed in float64.  With
bf16=False it is the plain float64 network, whose backward pass is pinned
against central finite differences in tests/test_cnn_oracle.py.
"""
from __future__ import annotations

import numpy as np

LAYERS = [(16, 32), (32, 32), (32, 32), (32, 32), (32, 32), (32, 32), (32, 32), (32, 32), (32, 16)]


def to_bf16(a):
    """float -> nearest BF16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def n_params() -> int:
    return sum(9 * ci * co + co for ci, co in LAYERS)


def unpack(params):
    """flat parameter vector -> [(W[9][cin][cout], b[cout])]"""
    p = np.asarray(params, dtype=np.float64)
    out, off = [], 0
    for ci, co in LAYERS:
        w = p[off:off + 9 * ci * co].reshape(9, ci, co)
        off += 9 * ci * co
        b = p[off:off + co]
        off += co
        out.append((w, b))
    return out


def _shift(x, dh, dw):
    """s[h][w] = x[h + dh][(w + dw) mod W], zero for rings outside the grid"""
    H = x.shape[0]
    s = np.zeros_like(x)
    lo, hi = max(0, -dh), min(H, H - dh)
    s[lo:hi] = x[lo + dh:hi + dh]
    return np.roll(s, -dw, axis=1)


def conv(x, w, b=None):
    out = np.zeros(x.shape[:2] + (w.shape[2],))
    for t in range(9):
        out += _shift(x, t // 3 - 1, t % 3 - 1) @ w[t]
    return out if b is None else out + b


def conv_input_grad(dz, w):
    """gradient of conv with respect to its input"""
    out = np.zeros(dz.shape[:2] + (w.shape[1],))
    for t in range(9):
        out += _shift(dz, -(t // 3 - 1), -(t % 3 - 1)) @ w[t].T
    return out


def conv_weight_grad(x, dz):
    ci, co = x.shape[2], dz.shape[2]
    dw = np.zeros((9, ci, co))
    for t in range(9):
        xs = _shift(x, t // 3 - 1, t % 3 - 1)
        dw[t] = xs.reshape(-1, ci).T @ dz.reshape(-1, co)
    return dw, dz.sum(axis=(0, 1))


def pool(x):
    return 0.25 * (((x[0::2, 0::2] + x[0::2, 1::2]) + x[1::2, 0::2]) + x[1::2, 1::2])


def up(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up_grad(du):
    return ((du[0::2, 0::2] + du[0::2, 1::2]) + du[1::2, 0::2]) + du[1::2, 1::2]


def residual(y, aux):
    """rho = -f0 + y + wip y[i+1] + wim y[i-1] + wjp y[j+1] + wjm y[j-1] on rings
    1..H-2 (y = 0 on rings 0 and H-1), aux = (f0, wip, wim, wjp, wjm)"""
    H = y.shape[0]
    yv = np.zeros_like(y)
    yv[1:H - 1] = y[1:H - 1]
    f0, wip, wim, wjp, wjm = aux
    rho = np.zeros_like(y)
    a = (yv + wip * np.roll(yv, -1, axis=0) + wim * np.roll(yv, 1, axis=0) + wjp * np.roll(yv, -1, axis=1)
         + wjm * np.roll(yv, 1, axis=1))
    rho[1:H - 1] = (-f0 + a)[1:H - 1]
    return rho


def residual_grad(rho, aux):
    """d(sum rho^2)/dy / 2 (the transposed stencil), zero on rings 0 and H-1"""
    H = rho.shape[0]
    f0, wip, wim, wjp, wjm = aux
    g = (rho + np.roll(wip * rho, 1, axis=0) + np.roll(wim * rho, -1, axis=0) + np.roll(wjp * rho, 1, axis=1)
         + np.roll(wjm * rho, -1, axis=1))
    g[0] = 0.0
    g[H - 1] = 0.0
    return g


def block_mean(a, f):
    """mean over f x f cells (the coarse network's restriction)"""
    H, W = a.shape[:2]
    return a.reshape(H // f, f, W // f, f, *a.shape[2:]).mean(axis=(1, 3))


def bilinear_axis(n, f, nc, periodic):
    """for the n fine cells of an axis: the coarse cells c0, c1 whose centres
    bracket the fine cell's centre u = (i + 0.5)/f - 0.5 and the weight w of
    c1; clamped to 0..nc-1 or taken modulo nc"""
    u = (np.arange(n) + 0.5) / f - 0.5
    c0 = np.floor(u).astype(int)
    w = u - c0
    c1 = c0 + 1
    if periodic:
        c0, c1 = c0 % nc, c1 % nc
    else:
        c0, c1 = np.clip(c0, 0, nc - 1), np.clip(c1, 0, nc - 1)
    return c0, c1, w


def upsample(yc, f):
    """bilinear interpolation of the coarse plane yc to the f-times finer grid"""
    Hc, Wc = yc.shape
    i0, i1, wi = bilinear_axis(Hc * f, f, Hc, False)
    j0, j1, wj = bilinear_axis(Wc * f, f, Wc, True)
    r0 = yc[i0][:, j0] * (1 - wj) + yc[i0][:, j1] * wj
    r1 = yc[i1][:, j0] * (1 - wj) + yc[i1][:, j1] * wj
    return r0 * (1 - wi)[:, None] + r1 * wi[:, None]


def upsample_adj(g, f):
    """transpose of upsample: each fine value spread back to its four coarse
    cells with its interpolation weights"""
    H, W = g.shape
    Hc, Wc = H // f, W // f
    i0, i1, wi = bilinear_axis(H, f, Hc, False)
    j0, j1, wj = bilinear_axis(W, f, Wc, True)
    out = np.zeros((Hc, Wc))
    for ci, wr in ((i0, 1 - wi), (i1, wi)):
        for cj, wc in ((j0, 1 - wj), (j1, wj)):
            np.add.at(out, (ci[:, None], cj[None, :]), wr[:, None] * wc[None, :] * g)
    return out


def forward_backward(params, x, t, bf16=True, aux=None, f=1):
    """loss, flat gradient, prediction y (H x W) and the activations; with
    aux the residual loss mean(rho^2) over rings 1..H-2 replaces the
    correction loss; with f > 1 (and aux on the f-times finer grid) the
    residual loss is that of the interpolated prediction"""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    relu = lambda a: np.maximum(a, 0.0)
    L = unpack(params)
    Wq = [r(w) for w, _ in L]
    B = [b for _, b in L]
    H, W = x.shape[:2]
    x = r(x)
    a = {}
    a[1] = r(relu(conv(x, Wq[0], B[0])))
    a[2] = r(relu(conv(a[1], Wq[1], B[1])))
    p1 = r(pool(a[2]))
    a[3] = r(relu(conv(p1, Wq[2], B[2])))
    p2 = r(pool(a[3]))
    a[4] = r(relu(conv(p2, Wq[3], B[3])))
    p3 = r(pool(a[4]))
    a[5] = r(relu(conv(p3, Wq[4], B[4])))
    u2 = r(up(a[5]) + a[4])
    a[6] = r(relu(conv(u2, Wq[5], B[5])))
    u1 = r(up(a[6]) + a[3])
    a[7] = r(relu(conv(u1, Wq[6], B[6])))
    u0 = r(up(a[7]) + a[2])
    a[8] = r(relu(conv(u0, Wq[7], B[7])))
    y = conv(a[8], Wq[8], B[8])[..., 0]
    n = (H - 2) * W
    dy = np.zeros((H, W, 16))
    if aux is None:
        d = np.zeros((H, W))
        d[1:H - 1] = y[1:H - 1] - t[1:H - 1]
        loss = float(np.sum(d * d) / n)
        dy[..., 0] = r(2.0 * d / n)
    elif f == 1:
        rho = residual(y, aux)
        loss = float(np.sum(rho * rho) / n)
        dy[..., 0] = r(2.0 * residual_grad(rho, aux) / n)
    else:
        yf = upsample(y, f)
        nf = (H * f - 2) * W * f
        rho = residual(yf, aux)
        loss = float(np.sum(rho * rho) / nf)
        dy[..., 0] = r(upsample_adj(2.0 * residual_grad(rho, aux) / nf, f))
    g = [None] * 9
    g[8] = conv_weight_grad(a[8], dy)
    dz8 = r(conv_input_grad(dy, Wq[8]) * (a[8] > 0))
    g[7] = conv_weight_grad(u0, dz8)
    du0 = r(conv_input_grad(dz8, Wq[7]))
    dz7 = r(up_grad(du0) * (a[7] > 0))
    g[6] = conv_weight_grad(u1, dz7)
    du1 = r(conv_input_grad(dz7, Wq[6]))
    dz6 = r(up_grad(du1) * (a[6] > 0))
    g[5] = conv_weight_grad(u2, dz6)
    du2 = r(conv_input_grad(dz6, Wq[5]))
    dz5 = r(up_grad(du2) * (a[5] > 0))
    g[4] = conv_weight_grad(p3, dz5)
    dp3 = r(conv_input_grad(dz5, Wq[4]))
    dz4 = r((0.25 * up(dp3) + du2) * (a[4] > 0))
    g[3] = conv_weight_grad(p2, dz4)
    dp2 = r(conv_input_grad(dz4, Wq[3]))
    dz3 = r((0.25 * up(dp2) + du1) * (a[3] > 0))
    g[2] = conv_weight_grad(p1, dz3)
    dp1 = r(conv_input_grad(dz3, Wq[2]))
    dz2 = r((0.25 * up(dp1) + du0) * (a[2] > 0))
    g[1] = conv_weight_grad(a[1], dz2)
    dz1 = r(conv_input_grad(dz2, Wq[1]) * (a[1] > 0))
    g[0] = conv_weight_grad(x, dz1)
    flat = np.concatenate([np.concatenate([dw.ravel(), db]) for dw, db in g])
    return loss, flat, y, a


def loss_only(params, x, t, aux=None, f=1):
    """float64 loss without rounding (for finite differences)"""
    return forward_backward(params, x, t, bf16=False, aux=aux, f=f)[0]
