"""The CPU restatement of the warm-start network (oracle/cnn_oracle.py) is
itself pinned here: its float64 backward pass against central finite
differences of its forward pass, its BF16 rounding against exact values, and
the convolution's boundary handling against a direct sum."""
import numpy as np

from oracle import cnn_oracle as C


def test_bf16_rounding_is_nearest_even():
    ulp = 2.0 ** -7
    cases = [(1.0, 1.0), (1.0 + ulp / 2, 1.0), (1.0 + 1.5 * ulp, 1.0 + 2 * ulp), (1.0 + 0.75 * ulp, 1.0 + ulp),
             (-3.0, -3.0), (0.0, 0.0), (1.0 + 0.25 * ulp, 1.0)]
    for x, want in cases:
        assert C.to_bf16(np.array([x]))[0] == want, x


def test_conv_boundaries_match_a_direct_sum():
    rng = np.random.default_rng(3)
    H, W = 4, 6
    x = rng.standard_normal((H, W, 16))
    w = rng.standard_normal((9, 16, 32))
    y = C.conv(x, w)
    for h in range(H):
        for j in range(W):
            s = np.zeros(32)
            for t in range(9):
                hh, jj = h + t // 3 - 1, (j + t % 3 - 1) % W     # zero across rings, periodic in sectors
                if 0 <= hh < H:
                    s += x[hh, jj] @ w[t]
            assert np.allclose(y[h, j], s, rtol=1e-12, atol=1e-12)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(7)
    H, W = 8, 16
    p = rng.standard_normal(C.n_params()) * 0.2
    x = rng.standard_normal((H, W, 16))
    t = rng.standard_normal((H, W))
    loss, g, _, _ = C.forward_backward(p, x, t, bf16=False)
    assert loss > 0
    idx = rng.choice(C.n_params(), 80, replace=False)
    # include every bias of the first and last layers (boundary cases of the layout)
    idx = np.concatenate([idx, np.arange(9 * 16 * 32, 9 * 16 * 32 + 32), np.arange(C.n_params() - 16, C.n_params())])
    bad = []
    for k in idx:
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (C.loss_only(pp, x, t) - C.loss_only(pm, x, t)) / (2 * h)
        if abs(fd - g[k]) > 1e-6 * max(1.0, abs(fd)) + 1e-9:
            bad.append((int(k), fd, g[k]))
    # output channels 1..15 do not reach the loss: their gradient is exactly 0
    assert not bad, bad[:5]


def test_residual_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(9)
    H, W = 8, 16
    p = rng.standard_normal(C.n_params()) * 0.2
    x = rng.standard_normal((H, W, 16))
    aux = [rng.standard_normal((H, W)), -0.3 * rng.random((H, W)), -0.3 * rng.random((H, W)),
           -0.2 * rng.random((H, W)), -0.2 * rng.random((H, W))]
    loss, g, y, _ = C.forward_backward(p, x, None, bf16=False, aux=aux)
    # the residual itself against a direct loop
    rho = C.residual(y, aux)
    for i in range(1, H - 1):
        for j in range(W):
            yv = lambda ii, jj: y[ii, jj % W] if 1 <= ii <= H - 2 else 0.0
            want = -aux[0][i, j] + yv(i, j) + aux[1][i, j] * yv(i + 1, j) + aux[2][i, j] * yv(i - 1, j) + \
                aux[3][i, j] * yv(i, j + 1) + aux[4][i, j] * yv(i, j - 1)
            assert abs(rho[i, j] - want) < 1e-12
    idx = rng.choice(C.n_params(), 60, replace=False)
    for k in idx:
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (C.loss_only(pp, x, None, aux) - C.loss_only(pm, x, None, aux)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(fd)) + 1e-9, (k, fd, g[k])


def test_block_mean_keeps_the_sum():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((16, 24, 3))
    for f in (2, 4):
        m = C.block_mean(a, f)
        assert m.shape == (16 // f, 24 // f, 3)
        assert np.allclose(m.sum(axis=(0, 1)) * f * f, a.sum(axis=(0, 1)), rtol=1e-12)
        assert np.allclose(m[1, 2], a[f:2 * f, 2 * f:3 * f].mean(axis=(0, 1)), rtol=1e-12)


def test_upsample_is_bilinear_through_cell_centres():
    """What interpolation between cell centres must give: constants stay
    constant; a field linear in the coarse ring centre is reproduced at every
    fine cell whose centre lies between the first and last coarse centres, and
    held at the end values beyond them; along the sectors the same inside the
    ring of sectors, and a shift by one coarse sector shifts the result by f
    fine sectors (periodic)."""
    rng = np.random.default_rng(5)
    Hc, Wc = 8, 12
    for f in (2, 4):
        assert np.allclose(C.upsample(np.full((Hc, Wc), 2.5), f), 2.5, rtol=0, atol=1e-15)
        # linear in the ring: coarse centre I at fine coordinate f (I + 1/2)
        yc = np.repeat((3.0 + 0.7 * np.arange(Hc))[:, None], Wc, axis=1)
        yf = C.upsample(yc, f)
        rf = np.arange(Hc * f) + 0.5                     # fine centres in fine units
        want = np.clip(3.0 + 0.7 * (rf / f - 0.5), 3.0, 3.0 + 0.7 * (Hc - 1))
        assert np.allclose(yf, want[:, None], rtol=0, atol=1e-13)
        # linear in the sector away from the wrap
        yc = np.repeat((1.0 - 0.3 * np.arange(Wc))[None, :], Hc, axis=0)
        yf = C.upsample(yc, f)
        jf = np.arange(Wc * f) + 0.5
        inside = (jf / f - 0.5 >= 0) & (jf / f - 0.5 <= Wc - 1)
        assert np.allclose(yf[:, inside], (1.0 - 0.3 * (jf[inside] / f - 0.5))[None, :], rtol=0, atol=1e-13)
        # periodic: shift equivariance
        yc = rng.standard_normal((Hc, Wc))
        assert np.allclose(C.upsample(np.roll(yc, 1, axis=1), f), np.roll(C.upsample(yc, f), f, axis=1),
                           rtol=0, atol=1e-14)


def test_upsample_adjoint_is_its_transpose():
    rng = np.random.default_rng(6)
    for f, (Hc, Wc) in ((2, (8, 12)), (4, (6, 10))):
        c = rng.standard_normal((Hc, Wc))
        g = rng.standard_normal((Hc * f, Wc * f))
        lhs = float(np.sum(C.upsample(c, f) * g))
        rhs = float(np.sum(c * C.upsample_adj(g, f)))
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_coarse_residual_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(10)
    f, H, W = 2, 8, 16                                   # network grid; the system's is 16 x 32
    p = rng.standard_normal(C.n_params()) * 0.2
    x = rng.standard_normal((H, W, 16))
    aux = [rng.standard_normal((H * f, W * f)), -0.3 * rng.random((H * f, W * f)), -0.3 * rng.random((H * f, W * f)),
           -0.2 * rng.random((H * f, W * f)), -0.2 * rng.random((H * f, W * f))]
    loss, g, y, _ = C.forward_backward(p, x, None, bf16=False, aux=aux, f=f)
    rho = C.residual(C.upsample(y, f), aux)
    assert abs(loss - np.sum(rho * rho) / ((H * f - 2) * W * f)) <= 1e-14 * loss
    idx = rng.choice(C.n_params(), 50, replace=False)
    idx = np.concatenate([idx, np.arange(C.n_params() - 16, C.n_params() - 14)])
    for k in idx:
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (C.loss_only(pp, x, None, aux, f) - C.loss_only(pm, x, None, aux, f)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(fd)) + 1e-9, (k, fd, g[k])
