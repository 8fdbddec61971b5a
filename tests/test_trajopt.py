"""Trajectory Adam (optim/trajectory_adam.py): the numpy restatement pinned by closed
forms, then the gfx950 kernel against it, and a short control optimisation through the
checkpointed adjoint."""
import numpy as np
import pytest
import torch

from oracle import trajopt


def test_first_step_moves_by_lr_times_sign():
    rng = np.random.default_rng(0)
    p0 = rng.normal(size=(4, 3, 5)).astype(np.float32)
    g = rng.normal(size=p0.shape).astype(np.float32)
    p = p0.copy()
    trajopt.adam_step(p, g, np.zeros_like(p), np.zeros_like(p), 0, lr=0.1, eps=1e-7)
    # m_hat = g, v_hat = g^2 after one step: the update is lr g / (|g| + eps)
    np.testing.assert_allclose(p, p0 - 0.1 * g / (np.abs(g) + 1e-7), rtol=0, atol=2e-6)


def test_constant_gradient_keeps_unit_steps_and_clipping_and_slice():
    p = np.zeros((2, 2, 6), np.float32)
    g = np.full(p.shape, 3.0, np.float32)
    g[..., 4] = -500.0  # clipped to -100: still one unit step per iteration
    m, v = np.zeros_like(p), np.zeros_like(p)
    for t in range(7):
        trajopt.adam_step(p, g, m, v, t, lr=0.01, clip=100.0, dof_offset=1, num_dofs=4)
    np.testing.assert_allclose(p[..., 1:4], -0.07, atol=1e-6)
    np.testing.assert_allclose(p[..., 4], 0.07, atol=1e-6)
    assert np.all(p[..., 0] == 0) and np.all(p[..., 5] == 0)
    np.testing.assert_allclose(m[..., 4], -100.0 * (1 - 0.9 ** 7), rtol=1e-5)


@pytest.mark.gpu
def test_gpu_adam_matches_restatement(gpu_device):
    from ostrich_amd.differentiable import TrajectoryAdam
    rng = np.random.default_rng(1)
    p0 = rng.normal(size=(50, 16, 9)).astype(np.float32)
    params = torch.as_tensor(p0, device=gpu_device).contiguous()
    opt = TrajectoryAdam(params, lr=0.05, dof_offset=6, num_dofs=3, clip_grad=2.0)
    ref, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for t in range(5):
        g = (rng.normal(size=p0.shape) * 3).astype(np.float32)
        opt.step(torch.as_tensor(g, device=gpu_device))
        trajopt.adam_step(ref, g, m, v, t, lr=0.05, dof_offset=6, num_dofs=3, clip=2.0)
    torch.cuda.synchronize()
    out = params.cpu().numpy()
    np.testing.assert_allclose(out, ref, rtol=1e-6, atol=1e-6)
    assert np.array_equal(out[..., :6], p0[..., :6])
    with pytest.raises(RuntimeError, match="outside"):
        TrajectoryAdam(params, dof_offset=8, num_dofs=3).step(torch.zeros_like(params))


@pytest.mark.gpu
def test_control_optimisation_lowers_the_loss(gpu_device):
    """Skid-steer robots on a flat plane: Adam on the wheel-velocity targets of a 10-step
    checkpointed rollout, gradients from the adjoint, lowers the loss on the final
    chassis x."""
    from ostrich_amd import OstrichEngine, collide, scenes
    from ostrich_amd.containers import Contacts, Control, State
    from ostrich_amd.differentiable import CheckpointedRollout, TrajectoryAdam
    from tests.helpers import config
    W, N, dt = 16, 10, 0.05
    model, state, control = scenes.skid_steer(W, seed=3, max_obstacles=0, mu_range=(0.8, 0.8))
    eng = OstrichEngine(model, config=config(K=32), device=gpu_device, differentiable_simulation=True)
    col = collide.PrimitiveCollider(model, device=gpu_device, gap=0.02, max_contacts_per_world=32)
    D, B = model.dofs_per_world, model.bodies_per_world
    tv = torch.zeros(N, W, D, device=gpu_device)
    tv[:, :, 6:9] = 1.0
    opt = TrajectoryAdam(tv, lr=0.5, dof_offset=6, num_dofs=3)
    s0 = State.from_numpy(state, gpu_device)
    tpos = torch.as_tensor(control["joint_target_pos"], device=gpu_device)
    losses = []
    for it in range(6):
        ro = CheckpointedRollout(eng, lambda q: Contacts.from_dict(col.collide(q), gpu_device), dt, checkpoint_every=5)
        fin = ro.forward(s0, [Control(tpos, tv[t].reshape(-1).contiguous()) for t in range(N)])
        x = fin.body_q.reshape(W, B, 7)[:, 0, 0]
        loss = ((x - 0.5) ** 2).sum()
        losses.append(float(loss))
        gq = torch.zeros_like(fin.body_q).reshape(W, B, 7)
        gq[:, 0, 0] = 2.0 * (x - 0.5)
        g = ro.backward(gq.reshape(-1, 7), torch.zeros_like(fin.body_qd))
        opt.step(g["joint_target_vel"].reshape(N, W, D).contiguous())
    print(f"[trajopt] losses {np.round(losses, 4)}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
