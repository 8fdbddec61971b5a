"""Parity of the gfx950 step_backward (through the C ABI) with the oracle's adjoint.

The GPU runs a forward step, then its adjoint with a random loss seed (dL/dq+, dL/du+).
The oracle's adjoint is evaluated at the GPU's own step output (q+, u+, lambda,
lambda_prev_iter) and the same contacts, so the comparison isolates the backward pass.
Tolerance: every gradient within 1e-3 of its largest magnitude (per quantity) of the
FP32 oracle (same pose-difference step), plus four times the FP32-vs-FP64 oracle spread
of that quantity (the FP64 adjoint differences poses with step 1e-6).  The spread is ~1e-5 on well-conditioned worlds (skid-steer); on box
stacks the coplanar contacts make the adjoint system near-singular and a 16-iteration
PCR leaves an FP32 truncation of ~1e-2 in every implementation, the oracle's included.
"""
import numpy as np
import pytest
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.containers import Contacts, Control, State
from tests.helpers import config, no_contacts, single_box, zero_control

pytestmark = pytest.mark.gpu

DT = 0.05
KEYS = ("body_q", "body_qd", "body_f", "joint_target_pos", "joint_target_vel", "joint_target_ke",
        "joint_target_kd", "shape_mu", "w_lambda", "w_u")


def _engine(model, cfg, device):
    from ostrich_amd import OstrichEngine
    return OstrichEngine(model, config=cfg, device=device, differentiable_simulation=True)


def _advance(oracle_mod, model, cfg, state, control, col, steps):
    st = {k: np.asarray(v, np.float64) for k, v in state.items()}
    for _ in range(steps):
        c = col.collide(torch.as_tensor(st["body_q"], dtype=torch.float32))
        o = oracle_mod.step(model, cfg, st, control, c, DT, precision=32)
        st = dict(st, body_q=o["body_q"], body_qd=o["body_qd"])
    return {k: np.asarray(v, np.float32) for k, v in st.items()}


def _compare(oracle_mod, model, cfg, state, control, col, device, seed=0, tol=1e-3):
    eng = _engine(model, cfg, device)
    c = col.collide(torch.as_tensor(state["body_q"], dtype=torch.float32))
    c = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}
    s_in = State.from_numpy(state, device)
    s_out = s_in.clone()
    ctl = Control.from_numpy(control, device)
    con = Contacts.from_dict(c, device)
    eng.step(s_in, s_out, ctl, con, DT)
    lam, lam_prev = eng.constraint_impulses(), eng.constraint_impulses_prev()
    rng = np.random.default_rng(seed)
    WB = model.num_worlds * model.bodies_per_world
    a_q = rng.normal(size=(WB, 7)).astype(np.float32)
    a_u = rng.normal(size=(WB, 6)).astype(np.float32)
    print(f"[adjoint parity] W={model.num_worlds} B={model.bodies_per_world} precond="
          f"{cfg.linear.preconditioner_type}", flush=True)
    g = eng.step_backward(s_in, s_out, ctl, con, DT, lam, lam_prev, torch.from_numpy(a_q), torch.from_numpy(a_u),
                          adjoint_multipliers=True)
    g = {k: v.cpu().numpy().astype(np.float64) for k, v in g.items()}
    flags = eng.stats()["flags"].cpu().numpy()
    out = dict(body_q=s_out.body_q.cpu().numpy(), body_qd=s_out.body_qd.cpu().numpy())
    lam_n, lp_n = lam.cpu().numpy(), lam_prev.cpu().numpy()
    o32 = oracle_mod.adjoint(model, cfg, state, out, control, c, DT, lam_n, lp_n, a_q, a_u, precision=32)
    o64 = oracle_mod.adjoint(model, cfg, state, out, control, c, DT, lam_n, lp_n, a_q, a_u, precision=64,
                             fd_eps=1e-6)
    np.testing.assert_array_equal(flags & 16, o32["flags"] & 16)
    for k in KEYS:
        if o64[k].size == 0 or g[k].size == 0 or k == "w_lambda":
            continue  # w_lambda is not unique where contact rows are redundant
        gk = g[k].reshape(o64[k].shape)
        scale = max(np.abs(o64[k]).max(), 1e-6)
        err = np.abs(gk - o32[k]).max()
        spread = np.abs(o32[k] - o64[k]).max()
        assert err <= 4.0 * spread + tol * scale, (f"{k}: max |gpu - oracle32| {err:.3e}, oracle32-vs-64 "
                                                   f"{spread:.3e}, scale {scale:.3e}")
    return g, o64


def test_adjoint_box_stack(oracle_mod, gpu_device):
    model, state, control = scenes.box_stack(16, seed=0)
    cfg = config(K=16)
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=16)
    st = _advance(oracle_mod, model, cfg, state, control, col, 3)
    _compare(oracle_mod, model, cfg, st, control, col, gpu_device)


def test_adjoint_tall_stack_full_residual_differences(oracle_mod, gpu_device):
    # six-cube stacks: 66-72 rows per world (two row slots per lane), so the pose
    # pull-back takes the full-residual central differences instead of the row-local ones
    model, state, control = scenes.box_stack(8, n_cubes=6, seed=0)
    cfg = config(K=48)
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=48)
    st = _advance(oracle_mod, model, cfg, state, control, col, 2)
    _compare(oracle_mod, model, cfg, st, control, col, gpu_device, seed=3)


def test_adjoint_box_stack_block_jacobi(oracle_mod, gpu_device):
    model, state, control = scenes.box_stack(16, seed=0)
    cfg = config(K=16, precond="per_body_pair")
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=16)
    st = _advance(oracle_mod, model, cfg, state, control, col, 3)
    _compare(oracle_mod, model, cfg, st, control, col, gpu_device, seed=1)


def test_adjoint_skid_steer(oracle_mod, gpu_device):
    model, state, control = scenes.skid_steer(32, seed=3)
    cfg = config(K=32)
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=32)
    st = _advance(oracle_mod, model, cfg, state, control, col, 4)
    g, o64 = _compare(oracle_mod, model, cfg, st, control, col, gpu_device, seed=2)
    assert np.abs(o64["joint_target_vel"]).max() > 0


@pytest.mark.parametrize("soft", [False, True])
@pytest.mark.parametrize("scene", ["skid_steer", "box_stack"])
def test_adjoint_frozen_friction(oracle_mod, gpu_device, scene, soft):
    """adjoint.friction_linearization = "frozen" (the stick / slip compliance surrogate of
    adjoint/adjoint_friction.py, with the 0.9 threshold or the sigmoid blend; friction
    rows left out of w^T dR/dxi) against the oracle's restatement."""
    from ostrich_amd.config import AdjointConfig
    adj = AdjointConfig(friction_linearization="frozen", soft_blending=soft)
    if scene == "skid_steer":
        model, state, control = scenes.skid_steer(32, seed=3)
        K, steps = 32, 4
    else:
        model, state, control = scenes.box_stack(16, seed=0)
        K, steps = 16, 3
    cfg = config(K=K, adjoint=adj)
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=K)
    st = _advance(oracle_mod, model, cfg, state, control, col, steps)
    g, o64 = _compare(oracle_mod, model, cfg, st, control, col, gpu_device, seed=4)
    assert np.abs(g["shape_mu"]).max() == 0.0


def test_adjoint_free_body_closed_form(gpu_device):
    """No contacts: w_u = M^-1 f_drive, dL/du- = -M w_u = g_u + dt g_p (translation)."""
    model, state = single_box(mass=2.0, pose=(0.3, -0.2, 3.0, 0, 0, 0, 1))
    state = {k: np.asarray(v, np.float32) for k, v in state.items()}
    cfg = config(nr_iters=4, K=4)
    eng = _engine(model, cfg, gpu_device)
    s_in = State.from_numpy(state, gpu_device)
    s_out = s_in.clone()
    ctl = Control.from_numpy(zero_control(model), gpu_device)
    con = Contacts.from_dict(no_contacts(model), gpu_device)
    eng.step(s_in, s_out, ctl, con, DT)
    gq = torch.tensor([[1.0, -2.0, 0.5, 0, 0, 0, 0]])
    gu = torch.tensor([[0.3, 0.1, -0.4, 0, 0, 0]])
    g = eng.step_backward(s_in, s_out, ctl, con, DT, eng.constraint_impulses(), eng.constraint_impulses_prev(),
                          gq, gu)
    expect = gu[0, :3] + DT * gq[0, :3]
    np.testing.assert_allclose(g["body_qd"][0, :3].cpu().numpy(), expect.numpy(), rtol=1e-6)
    np.testing.assert_allclose(g["body_f"][0, :3].cpu().numpy(), (DT * expect / 2.0).numpy(), rtol=1e-6)
    np.testing.assert_allclose(g["body_q"][0, :3].cpu().numpy(), gq[0, :3].numpy(), rtol=1e-6)


def test_backward_rejected_on_plain_path(gpu_device):
    model, state = single_box()
    eng = _engine(model, config(K=4, path="plain"), gpu_device)
    s = State.from_numpy({k: np.asarray(v, np.float32) for k, v in state.items()}, gpu_device)
    with pytest.raises(NotImplementedError):
        eng.step_backward(s, s, Control.from_numpy(zero_control(model), gpu_device),
                          Contacts.from_dict(no_contacts(model), gpu_device), DT, None, None, None, None)


def test_adjoint_does_not_depend_on_the_wave_split(gpu_device):
    """With few worlds each world's pose central differences run on several waves (the
    split), with many on one; the first 64 worlds of a 2048-world batch (one wave each)
    must get bit-identical gradients from a 64-world solver (four waves each) fed the same
    step."""
    from ostrich_amd import OstrichEngine
    cfg = config(K=128)
    big_m, big_s, big_c = scenes.skid_steer(2048, seed=1)
    small_m, _, small_c = scenes.skid_steer(64, seed=1)
    B = big_m.bodies_per_world
    big = OstrichEngine(big_m, config=cfg, device=gpu_device, differentiable_simulation=True)
    s_in = State.from_numpy(big_s, gpu_device)
    s_out = s_in.clone()
    ctl = Control.from_numpy(big_c, gpu_device)
    con = collide.HipCollider(big_m, device=gpu_device, max_contacts_per_world=128).collide(s_in.body_q)
    big.step(s_in, s_out, ctl, con, DT)
    lam, lam_prev = big.constraint_impulses(), big.constraint_impulses_prev()
    rng = np.random.default_rng(4)
    gq = torch.as_tensor(rng.normal(size=(2048 * B, 7)).astype(np.float32), device=gpu_device)
    gu = torch.as_tensor(rng.normal(size=(2048 * B, 6)).astype(np.float32), device=gpu_device)
    g_big = big.step_backward(s_in, s_out, ctl, con, DT, lam, lam_prev, gq, gu, adjoint_multipliers=True)

    n = 64 * B
    small = OstrichEngine(small_m, config=cfg, device=gpu_device, differentiable_simulation=True)
    si = State(s_in.body_q[:n].clone(), s_in.body_qd[:n].clone(), s_in.body_f[:n].clone())
    so = State(s_out.body_q[:n].clone(), s_out.body_qd[:n].clone(), s_out.body_f[:n].clone())
    D = small_m.dofs_per_world
    ctl_s = Control(ctl.joint_target_pos[:64 * D].clone(), ctl.joint_target_vel[:64 * D].clone())
    con_s = collide.HipCollider(small_m, device=gpu_device, max_contacts_per_world=128).collide(si.body_q)
    g_small = small.step_backward(si, so, ctl_s, con_s, DT, lam[:64].clone(), lam_prev[:64].clone(), gq[:n].clone(),
                                  gu[:n].clone(), adjoint_multipliers=True)
    torch.cuda.synchronize()
    print(f"[adjoint split] 64 of 2048 worlds vs 64 worlds, B={B}", flush=True)
    for k, v in g_small.items():
        a = g_big[k].reshape(-1)[:v.numel()].reshape(v.shape)
        assert torch.equal(a, v), k


def test_forward_only_engine_skips_lambda_prev(gpu_device):
    """An engine built without differentiable_simulation does not write lambda_prev_iter
    (and says so when asked for it); its forward results are bit-identical to a
    differentiable engine's."""
    from ostrich_amd import OstrichEngine
    model, state, control = scenes.skid_steer(64, seed=1)
    cfg = config(K=128)
    s_in = State.from_numpy(state, gpu_device)
    ctl = Control.from_numpy(control, gpu_device)
    con = collide.HipCollider(model, device=gpu_device, max_contacts_per_world=128).collide(s_in.body_q)
    outs = []
    for diff in (False, True):
        eng = OstrichEngine(model, config=cfg, device=gpu_device, differentiable_simulation=diff)
        s_out = s_in.clone()
        eng.step(s_in, s_out, ctl, con, DT)
        outs.append((s_out.body_q.cpu(), s_out.body_qd.cpu(), eng.constraint_impulses().cpu()))
        if not diff:
            with pytest.raises(RuntimeError, match="differentiable_simulation"):
                eng.constraint_impulses_prev()
        else:
            assert torch.isfinite(eng.constraint_impulses_prev()).all()
    for a, b in zip(*outs):
        assert torch.equal(a, b)
