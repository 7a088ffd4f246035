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
