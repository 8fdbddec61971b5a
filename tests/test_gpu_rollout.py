"""Checkpointed rollouts and the torch bridge on the GPU.

* Checkpoint spacing does not change the gradient: k = 1, 3 and N give bit-identical
  results (the solver is deterministic, so re-simulated windows equal the forward).
* Each adjoint step of the chain matches the oracle's adjoint evaluated on the GPU's own
  recorded step (same tolerance rule as tests/test_gpu_adjoint.py).
* `OstrichStep` / `OstrichRollout` hand torch autograd the same gradients as the direct
  calls.
"""
import numpy as np
import pytest
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.containers import Contacts, Control, State
from ostrich_amd.differentiable import CheckpointedRollout, OstrichRollout, OstrichStep
from tests.helpers import config

pytestmark = pytest.mark.gpu

DT = 0.05
N = 6


def _setup(gpu_device, worlds=16):
    from ostrich_amd import OstrichEngine
    model, state, control = scenes.skid_steer(worlds, seed=5, max_obstacles=4, mu_range=(0.8, 0.8))
    cfg = config(K=32)
    eng = OstrichEngine(model, config=cfg, device=gpu_device, differentiable_simulation=True)
    col = collide.PrimitiveCollider(model, device=gpu_device, gap=0.02, max_contacts_per_world=32)
    collide_fn = lambda q: Contacts.from_dict(col.collide(q), gpu_device)
    rng = np.random.default_rng(11)
    tv = rng.uniform(-3, 3, size=(N, model.num_worlds * model.dofs_per_world)).astype(np.float32)
    controls = [Control(torch.as_tensor(control["joint_target_pos"], device=gpu_device),
                        torch.as_tensor(tv[t], device=gpu_device)) for t in range(N)]
    s0 = State.from_numpy(state, gpu_device)
    return model, cfg, eng, collide_fn, controls, s0


def _seed(model, gpu_device, seed=0):
    rng = np.random.default_rng(seed)
    WB = model.num_worlds * model.bodies_per_world
    return (torch.as_tensor(rng.normal(size=(WB, 7)), dtype=torch.float32, device=gpu_device),
            torch.as_tensor(rng.normal(size=(WB, 6)), dtype=torch.float32, device=gpu_device))


def test_checkpoint_spacing_is_exact(gpu_device):
    model, cfg, eng, collide_fn, controls, s0 = _setup(gpu_device)
    gq, gu = _seed(model, gpu_device)
    res = []
    for k in (1, 3, N):
        ro = CheckpointedRollout(eng, collide_fn, DT, checkpoint_every=k)
        fin = ro.forward(s0, controls)
        g = ro.backward(gq, gu)
        print(f"[rollout] k={k} recomputed {ro.recomputed_steps}", flush=True)
        assert ro.recomputed_steps == N
        res.append((fin.body_q.cpu(), {n: v.cpu() for n, v in g.items()}))
    for fin, g in res[1:]:
        assert torch.equal(fin, res[0][0])
        for n in g:
            assert torch.equal(g[n], res[0][1][n]), n
    assert res[0][1]["joint_target_vel"].abs().max() > 0


def test_chain_matches_oracle_per_step(gpu_device, oracle_mod):
    model, cfg, eng, collide_fn, controls, s0 = _setup(gpu_device, worlds=8)
    gq, gu = _seed(model, gpu_device, 1)
    ro = CheckpointedRollout(eng, collide_fn, DT, checkpoint_every=N)
    ro.forward(s0, controls)
    buf = ro._window(0)
    np_ = lambda t: t.detach().cpu().numpy()
    for t in reversed(range(N)):
        r = buf[t]
        print(f"[rollout] oracle chain step {t}", flush=True)
        g = eng.step_backward(r.state_in, r.state_out, r.control, r.contacts, DT, r.lam, r.lam_prev, gq, gu)
        c = {k: np_(v) if torch.is_tensor(v) else v for k, v in r.contacts.to_dict().items()}
        si = dict(body_q=np_(r.state_in.body_q), body_qd=np_(r.state_in.body_qd), body_f=np_(r.state_in.body_f))
        so = dict(body_q=np_(r.state_out.body_q), body_qd=np_(r.state_out.body_qd))
        ctl = dict(joint_target_pos=np_(r.control.joint_target_pos), joint_target_vel=np_(r.control.joint_target_vel))
        o32 = oracle_mod.adjoint(model, cfg, si, so, ctl, c, DT, np_(r.lam), np_(r.lam_prev), np_(gq), np_(gu),
                                 precision=32)
        o64 = oracle_mod.adjoint(model, cfg, si, so, ctl, c, DT, np_(r.lam), np_(r.lam_prev), np_(gq), np_(gu),
                                 precision=64, fd_eps=1e-6)
        for k in ("body_q", "body_qd", "joint_target_vel"):
            gk = np_(g[k]).reshape(o64[k].shape)
            scale = max(np.abs(o64[k]).max(), 1e-6)
            err = np.abs(gk - o32[k]).max()
            spread = np.abs(o32[k] - o64[k]).max()
            assert err <= 4 * spread + 1e-3 * scale, f"step {t} {k}: {err:.3e} (spread {spread:.3e}, scale {scale:.3e})"
        gq, gu = g["body_q"], g["body_qd"]


def test_autograd_functions_match_direct_calls(gpu_device):
    model, cfg, eng, collide_fn, controls, s0 = _setup(gpu_device, worlds=4)
    gq, gu = _seed(model, gpu_device, 2)
    # one step through torch autograd
    q = s0.body_q.clone().requires_grad_(True)
    qd = s0.body_qd.clone().requires_grad_(True)
    tv = controls[0].joint_target_vel.clone().requires_grad_(True)
    SC = model.shapes_per_world + model.num_global_shapes
    mu = torch.as_tensor(model.arrays["shape_mu"], dtype=torch.float32, device=gpu_device).clone().requires_grad_(True)
    con = collide_fn(s0.body_q)
    q1, qd1 = OstrichStep.apply(eng, con, DT, q, qd, s0.body_f, controls[0].joint_target_pos, tv, mu)
    ((q1 * gq).sum() + (qd1 * gu).sum()).backward()
    ref_out = s0.clone()
    eng.step(s0, ref_out, Control(controls[0].joint_target_pos, tv.detach()), con, DT)
    g = eng.step_backward(s0, ref_out, Control(controls[0].joint_target_pos, tv.detach()), con, DT,
                          eng.constraint_impulses(), eng.constraint_impulses_prev(), gq, gu)
    assert torch.equal(q.grad, g["body_q"])
    assert torch.equal(qd.grad, g["body_qd"])
    assert torch.equal(tv.grad, g["joint_target_vel"].reshape(-1))
    assert mu.grad.shape == (model.num_worlds * SC,)
    # whole rollout
    tp = torch.stack([c.joint_target_pos for c in controls])
    tvs = torch.stack([c.joint_target_vel for c in controls]).clone().requires_grad_(True)
    qN, qdN = OstrichRollout.apply(eng, collide_fn, DT, 2, s0.body_q, s0.body_qd, tp, tvs)
    ((qN * gq).sum() + (qdN * gu).sum()).backward()
    ro = CheckpointedRollout(eng, collide_fn, DT, checkpoint_every=2)
    ro.forward(s0, controls)
    gr = ro.backward(gq, gu)
    assert torch.equal(tvs.grad, gr["joint_target_vel"].reshape(N, -1))
    # the rollout with the friction coefficients as a parameter (0.1 instead of the
    # model's 0.8, so that wheels slip and dL/dmu is not zero): the same gradients as a
    # CheckpointedRollout under that friction, and dL/dmu as its per-step sum
    mu_r = torch.full((model.num_worlds * SC,), 0.1, device=gpu_device).requires_grad_(True)
    tvs2 = tvs.detach().clone().requires_grad_(True)
    qN2, qdN2 = OstrichRollout.apply(eng, collide_fn, DT, 2, s0.body_q, s0.body_qd, tp, tvs2, mu_r)
    ((qN2 * gq).sum() + (qdN2 * gu).sum()).backward()
    eng.set_shape_mu(mu_r.detach())
    ro2 = CheckpointedRollout(eng, collide_fn, DT, checkpoint_every=2)
    fin2 = ro2.forward(s0, controls)
    gr2 = ro2.backward(gq, gu)
    eng.set_shape_mu(eng.model_shape_mu())
    assert torch.equal(qN2, fin2.body_q) and not torch.equal(qN2, qN)
    assert torch.equal(tvs2.grad, gr2["joint_target_vel"].reshape(N, -1))
    assert torch.equal(mu_r.grad, gr2["shape_mu"].reshape(-1))
    assert mu_r.grad.abs().max() > 0


def _setup_seeded(gpu_device, worlds=16):
    """As _setup, with the cross-step warm starter seeding Newton (warm_start.seed_iterate)."""
    from ostrich_amd import OstrichEngine
    from ostrich_amd.config import WarmStartConfig
    model, cfg, _, collide_fn, controls, s0 = _setup(gpu_device, worlds)
    base = config(K=32)
    cfg = type(base)(nr=base.nr, linear=base.linear, contacts=base.contacts, execution=base.execution,
                     warm_start=WarmStartConfig(enabled=True, seed_iterate=True))
    return model, cfg, OstrichEngine(model, config=cfg, device=gpu_device, differentiable_simulation=True), collide_fn, controls, s0


def test_seeded_rollout_is_a_function_of_its_inputs(gpu_device):
    """With the warm starter seeding Newton, each checkpoint carries the warm starter's
    snapshot and a rollout starts from a reset one: checkpoint spacing 1, 3 and N give
    bit-identical final states and gradients, and running the rollout again (after other
    steps have changed the warm starter) gives the same bits."""
    model, cfg, eng, collide_fn, controls, s0 = _setup_seeded(gpu_device)
    gq, gu = _seed(model, gpu_device, 3)
    res = []
    for k in (1, 3, N, 1):
        ro = CheckpointedRollout(eng, collide_fn, DT, checkpoint_every=k)
        fin = ro.forward(s0, controls)
        g = ro.backward(gq, gu)
        res.append((fin.body_q.cpu(), {n: v.cpu() for n, v in g.items()}))
        # unrelated steps in between move the engine's warm-start snapshot
        junk = s0.clone()
        eng.step(s0, junk, controls[0], collide_fn(s0.body_q), DT)
    for fin, g in res[1:]:
        assert torch.equal(fin, res[0][0])
        for n in g:
            assert torch.equal(g[n], res[0][1][n]), n
    assert res[0][1]["joint_target_vel"].abs().max() > 0


def test_step_backward_without_arguments(gpu_device):
    """The reference's form (core/base_engine.py:528, simulation/differentiable_simulator.py:
    234-271): the forward loop records each step with TrajectoryBuffer.save_step(i,
    engine.data), the backward loop restores it with load_step(i, engine.data), calls
    engine.step_backward() with no arguments and carries the gradients with
    save_gradients / save_pose_gradients.  It gives the same bits as the explicit form
    (CheckpointedRollout.backward), and right after a step the no-argument call equals the
    explicit call on that step's data."""
    from ostrich_amd.differentiable import TrajectoryBuffer
    model, cfg, eng, collide_fn, controls, s0 = _setup(gpu_device, worlds=8)
    gq, gu = _seed(model, gpu_device, 4)
    # one step: no-argument call right after it
    s1 = s0.clone()
    con = collide_fn(s0.body_q)
    eng.step(s0, s1, controls[0], con, DT)
    eng.data.body_q_grad, eng.data.body_qd_grad = gq, gu
    g0 = {k: v.clone() for k, v in eng.step_backward().items()}
    g1 = eng.step_backward(s0, s1, controls[0], con, DT, eng.constraint_impulses(), eng.constraint_impulses_prev(),
                           gq, gu)
    for k in g1:
        assert torch.equal(g0[k], g1[k]), k
    # the reference's loop over a rollout
    traj = TrajectoryBuffer()
    states = [s0.clone()]
    for t in range(N):
        nxt = states[-1].clone()
        eng.step(states[-1], nxt, controls[t], collide_fn(states[-1].body_q), DT)
        traj.save_step(t, eng.data)
        states.append(nxt)
    traj.set_output_grad(N, gq, gu)
    for t in reversed(range(N)):
        traj.load_step(t, eng.data)
        eng.data.zero_gradients()
        eng.step_backward()
        traj.save_gradients(t, eng.data)
        traj.save_pose_gradients(t, eng.data)
    ro = CheckpointedRollout(eng, collide_fn, DT, checkpoint_every=N)
    ro.forward(s0, controls)
    gr = ro.backward(gq, gu)
    assert torch.equal(traj.body_q_grad[0], gr["body_q"])
    assert torch.equal(traj.body_qd_grad[0], gr["body_qd"])
    for t in range(N):
        assert torch.equal(traj.input_grads[t]["joint_target_vel"], gr["joint_target_vel"][t])
