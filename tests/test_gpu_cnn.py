"""The learned warm start of the energy solve on the GPU (thorin_cnn_*).

* The HIP forward and backward kernels (BF16 implicit-GEMM convolutions on
  MFMA, FP32 accumulation) against the CPU restatement oracle/cnn_oracle.py,
  which rounds to BF16 at the same storage points: prediction, loss and every
  layer's parameter gradient.  Grids 16x32 and 24x48 (the coarse levels are
  narrower than one 16-pixel MFMA tile, and 48 is not a multiple of 16).
* Training on pairs the solver generates lowers the loss, and the trained
  network's guess lowers the initial residual and the number of FP32 sweeps
  to the same 1e-10 target, with the same converged temperature.
"""
import numpy as np
import pytest

from fargo_thorin_amd import Disk, Params, WarmStart, configs, make_planets, seeded_disk
from oracle import cnn_oracle as CO

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("H,W", [(16, 32), (24, 48), (16, 320)])
def test_gpu_cnn_gradients_match_the_oracle(H, W):
    print(f"[gpu-cnn] gradients {H}x{W}", flush=True)
    net = WarmStart(seed=11)
    rng = np.random.default_rng(H)
    p = net.params()
    p[-16:] = rng.standard_normal(16).astype(np.float32) * 0.1      # nonzero output biases too
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    x = rng.standard_normal((H, W, 16)).astype(np.float32)
    x[..., 6:] = 0.0
    t = (rng.standard_normal((H, W)) * 0.5).astype(np.float32)
    L, g, y = net.grad(x, t)
    Lo, go, yo, _ = CO.forward_backward(p.astype(np.float64), x, t, bf16=True)
    # the sums run in another order (FP32 on MFMA, float64 here), so a few
    # stored activations round to the neighbouring BF16 value
    assert abs(L - Lo) <= 1e-3 * Lo
    assert rel(y, yo) <= 1e-3
    off = 0
    for k, (ci, co) in enumerate(CO.LAYERS):
        n = 9 * ci * co + co
        err = rel(g[off:off + n], go[off:off + n])
        assert err <= 1e-2, f"layer {k}: gradient rel L2 {err:.3e}"
        off += n
    # output channels 1..15 do not reach the loss
    w9 = g[off - n:off - 16].reshape(9, 32, 16)
    assert np.all(w9[:, :, 1:] == 0.0)


@pytest.mark.parametrize("H,W", [(16, 32), (24, 48), (16, 320)])
def test_gpu_cnn_residual_loss_gradients_match_the_oracle(H, W):
    print(f"[gpu-cnn] residual-loss gradients {H}x{W}", flush=True)
    net = WarmStart(seed=13)
    rng = np.random.default_rng(100 + H)
    p = net.params()
    x = rng.standard_normal((H, W, 16)).astype(np.float32)
    x[..., 6:] = 0.0
    aux = np.stack([rng.standard_normal((H, W)), -0.3 * rng.random((H, W)), -0.3 * rng.random((H, W)),
                    -0.2 * rng.random((H, W)), -0.2 * rng.random((H, W))]).astype(np.float32)
    L, g, y = net.grad(x, aux=aux)
    Lo, go, yo, _ = CO.forward_backward(p.astype(np.float64), x, None, bf16=True, aux=list(aux.astype(np.float64)))
    assert abs(L - Lo) <= 1e-3 * Lo
    assert rel(y, yo) <= 1e-3
    off = 0
    for k, (ci, co) in enumerate(CO.LAYERS):
        n = 9 * ci * co + co
        a, b = g[off:off + n].astype(float), go[off:off + n]
        err = rel(a, b)
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
        # the first layer collects the most BF16 rounding flips of the pass
        assert err <= 3e-2 and cos >= 0.999, f"layer {k}: gradient rel L2 {err:.3e}, cosine {cos:.6f}"
        off += n


@pytest.mark.parametrize("f,H,W", [(2, 16, 32), (2, 24, 48), (4, 16, 32)])
def test_gpu_coarse_cnn_residual_loss_gradients_match_the_oracle(f, H, W):
    """The network on the system's grid coarsened f times: the prediction on
    the H x W network grid, the residual loss of its bilinear interpolation
    on the fH x fW system grid, and the gradients through the transposed
    interpolation, against the oracle."""
    print(f"[gpu-cnn] coarse f={f} residual-loss gradients {H}x{W}", flush=True)
    net = WarmStart(seed=17)
    net.set_coarsen(f)
    assert net.coarsen() == f
    rng = np.random.default_rng(200 + 10 * f + H)
    p = net.params()
    p[-16:] = rng.standard_normal(16).astype(np.float32) * 0.1
    net.set_params(p)
    x = rng.standard_normal((H, W, 16)).astype(np.float32)
    x[..., 6:] = 0.0
    F, G = H * f, W * f
    aux = np.stack([rng.standard_normal((F, G)), -0.3 * rng.random((F, G)), -0.3 * rng.random((F, G)),
                    -0.2 * rng.random((F, G)), -0.2 * rng.random((F, G))]).astype(np.float32)
    L, g, y = net.grad(x, aux=aux)
    Lo, go, yo, _ = CO.forward_backward(p.astype(np.float64), x, None, bf16=True, aux=list(aux.astype(np.float64)),
                                        f=f)
    assert y.shape == (H, W)
    assert abs(L - Lo) <= 1e-3 * Lo
    assert rel(y, yo) <= 1e-3
    off = 0
    for k, (ci, co) in enumerate(CO.LAYERS):
        n = 9 * ci * co + co
        a, b = g[off:off + n].astype(float), go[off:off + n]
        err = rel(a, b)
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
        assert err <= 3e-2 and cos >= 0.999, f"layer {k}: gradient rel L2 {err:.3e}, cosine {cos:.6f}"
        off += n
    net.close()


def test_gpu_coarsen_is_1_2_or_4():
    net = WarmStart(seed=1)
    with pytest.raises(RuntimeError):
        net.set_coarsen(3)
    net.set_coarsen(4)
    net.set_coarsen(1)
    net.close()


def test_gpu_warm_start_lowers_residual_and_sweeps():
    print("[gpu-cnn] warm start 128x192", flush=True)
    NR, NS = 128, 192
    P = Params(**configs.with_grid(configs.DAMPED_1024x1536, NR, NS), OutputDir="/nonexistent-thorin-out/")
    d = Disk(P)
    sysp = make_planets([1e-13], [1.0])

    def assemble(seed, steps):
        disk = seeded_disk(NR, NS, seed=seed)
        d.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
        for _ in range(steps):
            d.algogas(sysp)
        d.rad_system(0.0)

    net = WarmStart(seed=3)
    for seed in range(6):
        for steps in (1, 3):
            assemble(100 + seed, steps)
            r = d.rad_solve(1, tol=1e-12, omega=1.8, inner_sweeps=32)
            assert r["converged"]
            net.add_pair(d)
    assert net.n_pairs() == 12
    lf, ll = net.train(300, 1e-3)
    assert ll < 0.01 * lf
    assemble(999, 2)                       # a disk the network has not seen
    cold = d.rad_solve(1, tol=1e-10, omega=1.8, inner_sweeps=32)
    t_cold = d.field("temper")
    warm = d.rad_solve(1, tol=1e-10, omega=1.8, inner_sweeps=32, warm=net)
    t_warm = d.field("temper")
    assert cold["converged"] and warm["converged"] and warm["warm_used"]
    assert warm["resid0"] == cold["resid0"]                 # same target: 1e-10 of T0's residual
    assert warm["resid_warm"] < 0.05 * warm["resid0"]
    assert warm["sweeps"] < cold["sweeps"]
    assert rel(t_warm[1:-1], t_cold[1:-1]) < 1e-9
    print(f"[gpu-cnn] cold {cold['sweeps']} sweeps, warm {warm['sweeps']} sweeps, "
          f"guess residual {warm['resid_warm'] / warm['resid0']:.2e} of T0's", flush=True)


def test_gpu_coarse_warm_start_lowers_residual_and_sweeps():
    """The network on the half-resolution grid, trained on the residual of its
    interpolated guess: the guess is kept, lowers the initial residual and the
    sweeps to the same 1e-10 target, and the solve ends at the same
    temperature."""
    print("[gpu-cnn] coarse (f=2) warm start 128x192", flush=True)
    NR, NS = 128, 192
    P = Params(**configs.with_grid(configs.DAMPED_1024x1536, NR, NS), OutputDir="/nonexistent-thorin-out/")
    d = Disk(P)
    sysp = make_planets([1e-13], [1.0])

    def assemble(seed, steps):
        disk = seeded_disk(NR, NS, seed=seed)
        d.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
        for _ in range(steps):
            d.algogas(sysp)
        d.rad_system(0.0)

    net = WarmStart(seed=3)
    net.set_coarsen(2)
    net.set_loss(1)
    for seed in range(6):
        for steps in (1, 3):
            assemble(100 + seed, steps)
            assert d.rad_solve(1, tol=1e-12, omega=1.8, inner_sweeps=32)["converged"]
            net.add_pair(d)
    with pytest.raises(RuntimeError):
        net.set_coarsen(1)                 # the pairs hold half-resolution features
    lf, ll = net.train(400, 1e-3)
    assert ll < 0.5 * lf
    assemble(999, 2)
    cold = d.rad_solve(1, tol=1e-10, omega=1.8, inner_sweeps=32)
    t_cold = d.field("temper")
    warm = d.rad_solve(1, tol=1e-10, omega=1.8, inner_sweeps=32, warm=net)
    t_warm = d.field("temper")
    print(f"[gpu-cnn] coarse: cold {cold['sweeps']} sweeps, warm {warm['sweeps']} sweeps, "
          f"guess residual {warm['resid_warm'] / warm['resid0']:.2e} of T0's, loss {lf:.3e} -> {ll:.3e}", flush=True)
    assert cold["converged"] and warm["converged"] and warm["warm_used"]
    assert warm["resid0"] == cold["resid0"]
    assert warm["resid_warm"] < 0.5 * warm["resid0"]
    assert warm["sweeps"] < cold["sweeps"]
    assert rel(t_warm[1:-1], t_cold[1:-1]) < 1e-9
    net.close()


def test_gpu_coarse_warm_start_saves_time_at_2048x3072():
    """The energy solve of the 2048x3072 disk to 1e-10 (FP32 sweeps in FP64
    refinement, 64 sweeps a round): from the half-resolution network's guess
    it takes less GPU time than from T0, inference and the guess's residual
    included.  The network is trained here on pairs of other seeds and
    steps; the timed disk is not among them."""
    from fargo_thorin_amd.seed import sor_disk
    print("[gpu-cnn] coarse warm start time gain 2048x3072", flush=True)
    NR, NS = 2048, 3072
    d = Disk(Params(**configs.with_grid(configs.SOR_2048x3072, NR, NS), OutputDir="/nonexistent-thorin-out/"))

    def assemble(seed, mult=1.0):
        disk = sor_disk(NR, NS, seed=seed)
        d.init(disk["dens"], disk["vrad"], disk["vtheta"], disk["temper"])
        dt0 = d.rad_system(0.0)
        if mult != 1.0:
            d.rad_system(dt0 * mult)

    net = WarmStart(seed=3)
    net.set_coarsen(2)
    net.set_loss(1)
    for seed in range(3):
        for mult in (0.5, 1.0, 2.0):
            assemble(200 + seed, mult)
            d.rad_solve(1, tol=1e-11, omega=1.8, inner_sweeps=64)
            net.add_pair(d)
    lf, ll = net.train(3000, 3e-4)
    assemble(12345)
    cold = min((d.rad_solve(1, tol=1e-10, omega=1.8, inner_sweeps=64) for _ in range(2)), key=lambda r: r["ms"])
    t_cold = d.field("temper")
    warm = min((d.rad_solve(1, tol=1e-10, omega=1.8, inner_sweeps=64, warm=net) for _ in range(2)),
               key=lambda r: r["ms"])
    t_warm = d.field("temper")
    print(f"[gpu-cnn] 2048x3072: cold {cold['ms']:.2f} ms {cold['rounds']} rounds, warm {warm['ms']:.2f} ms "
          f"{warm['rounds']} rounds, guess residual {warm['resid_warm'] / warm['resid0']:.3e}, "
          f"loss {lf:.3e} -> {ll:.3e}", flush=True)
    assert cold["converged"] and warm["converged"] and warm["warm_used"]
    assert warm["rounds"] < cold["rounds"]
    assert warm["ms"] < cold["ms"]
    assert rel(t_warm[1:-1], t_cold[1:-1]) < 1e-9
    net.close()
    d.close()
