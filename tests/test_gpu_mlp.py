"""The gfx950 warm-start network against its numpy restatement (oracle/mlp.py).

Forward outputs and parameter gradients are compared with the restatement evaluated with
the same BF16 rounding points; the GPU accumulates in FP32 inside the matrix cores in a
different order, which can move an activation across a BF16 rounding boundary, so the
tolerance is 1e-2 of each quantity's scale.  The tile layout of v_mfma_f32_16x16x16_bf16
is exercised by every product (forward, transposed backward)."""
import numpy as np
import pytest
import torch

from oracle import mlp

pytestmark = pytest.mark.gpu


def _net(gpu_device, seed=0):
    from ostrich_amd.learned import WarmStartNet
    return WarmStartNet(gpu_device, seed=seed)


def test_forward_matches_restatement(gpu_device):
    net = _net(gpu_device, 3)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(200, 12)).astype(np.float32)  # 200 rows: a partial last block
    y = net.forward(torch.from_numpy(x)).cpu().numpy()
    p = net.params().cpu().numpy()
    yo, _ = mlp.forward(p, x)
    scale = np.abs(yo).max()
    print(f"[mlp] forward max err {np.abs(y - yo).max():.3e} scale {scale:.3e}", flush=True)
    assert np.abs(y - yo).max() <= 1e-2 * scale


def test_gradient_matches_restatement(gpu_device):
    net = _net(gpu_device, 4)
    rng = np.random.default_rng(1)
    x = rng.normal(size=(300, 12)).astype(np.float32)
    t = rng.normal(size=(300, 3)).astype(np.float32)
    g, loss = net.gradient(torch.from_numpy(x), torch.from_numpy(t))
    g = g.cpu().numpy()
    lo, go = mlp.loss_and_grad(net.params().cpu().numpy(), x, t)
    assert loss.item() == pytest.approx(lo, rel=1e-2)
    for name, a, b in (("W1", mlp.OW1, mlp.OB1), ("b1", mlp.OB1, mlp.OW2), ("W2", mlp.OW2, mlp.OB2),
                       ("b2", mlp.OB2, mlp.OW3), ("W3", mlp.OW3, mlp.OB3), ("b3", mlp.OB3, mlp.NPARAM)):
        scale = max(np.abs(go[a:b]).max(), 1e-8)
        err = np.abs(g[a:b] - go[a:b]).max()
        print(f"[mlp] grad {name} err {err:.3e} scale {scale:.3e}", flush=True)
        assert err <= 1e-2 * scale, name


def test_training_fits_a_smooth_map(gpu_device):
    """Adam on y = (sin x0, x1 x2, relu(x3)) from 8192 samples: loss falls by 10x."""
    net = _net(gpu_device, 5)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(8192, 12)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0]), x[:, 1] * x[:, 2], np.maximum(x[:, 3], 0)], 1).astype(np.float32)
    hist = net.fit(torch.from_numpy(x), torch.from_numpy(t), epochs=30, batch=512, lr=3e-3)
    print(f"[mlp] loss {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    assert hist[-1] < 0.1 * hist[0]
