"""The reference's wheeled-robot gradient checks
(tests/differentiable_simulator/test_wheeled_robot.py:46-53, 151-340) restated on the
oracle's FP64 adjoint: the four-wheeled Taros-4 robot (examples/taros_4/common.py)
started 5 cm above the ground as there, velocity motors k_p 1000, Newton <= 20,
PCR <= 200 to 1e-8; loss w . body_qd+ with w drawn from numpy's legacy generator with
the reference's seeds.  dL/d(wheel target velocity) of the adjoint against central
differences of the FP64 forward step (h 1e-4; 1e-3 through five steps), within the
reference's bands: 0.15 straight and multi-step, 0.5 for the turns.

Beyond the bands: from the airborne start every step converges (Newton stops at
||h|| <= 1e-5 in two iterations) and the adjoint matches the differences to 1e-5 relative (1e-4 through five steps,
where the differences use h 1e-3)."""
import numpy as np
import pytest
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.config import ContactsConfig, LinearSolverConfig, NewtonRaphsonConfig, OstrichEngineConfig

DT = 0.01
CASES = {  # name: (friction, wheel targets fl fr rl rr, seed, band)
    "straight": (0.8, (5.0, 5.0, 5.0, 5.0), 42, 0.15),
    "differential_turn": (0.5, (2.0, 10.0, 2.0, 10.0), 123, 0.5),
    "hard_turn": (0.3, (0.0, 20.0, 0.0, 20.0), 99, 0.5),
}


def reference_config(K=16):
    return OstrichEngineConfig(nr=NewtonRaphsonConfig(max_iters=20),
                               linear=LinearSolverConfig(max_iters=200, tol=1e-8, atol=1e-8),
                               contacts=ContactsConfig(max_per_world=K))


def taros_case(name):
    friction, targets, seed, band = CASES[name]
    model, state, control = scenes.taros4(1, friction=friction, z=0.8)
    control = dict(control, joint_target_vel=control["joint_target_vel"].astype(np.float64))
    control["joint_target_vel"][6:10] = targets
    np.random.seed(seed)
    w = np.random.randn(model.bodies_per_world * 6).astype(np.float32).astype(np.float64)
    state = {k: np.asarray(v, np.float64) for k, v in state.items()}
    return model, state, control, w, band


def contacts_of(model, body_q, K=16):
    col = collide.PrimitiveCollider(model, gap=0.05, max_contacts_per_world=K)
    c = col.collide(torch.as_tensor(np.asarray(body_q), dtype=torch.float32))
    return {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}


def rollout_loss(oracle_mod, model, cfg, state, control, w, steps):
    """w . body_qd after `steps` FP64 oracle steps (contacts from each step's poses)."""
    st = dict(state)
    for _ in range(steps):
        o = oracle_mod.step(model, cfg, st, control, contacts_of(model, st["body_q"]), DT, precision=64)
        st = dict(st, body_q=o["body_q"], body_qd=o["body_qd"])
    return float(w @ st["body_qd"].reshape(-1))


def fd_wheel_gradient(oracle_mod, model, cfg, state, control, w, steps=1, h=1e-4):
    """Central differences in the wheel targets.  Targets cross the ABI as float32, so
    the step taken is the float32 difference of the two perturbed targets (at 5 rad/s it
    is 1.4e-3 larger than 2h: all of the deviation the reference's test meets there)."""
    g = np.zeros(4)
    for d in range(4):
        vals, taken = [], []
        for sgn in (1.0, -1.0):
            c = dict(control, joint_target_vel=control["joint_target_vel"].copy())
            c["joint_target_vel"][6 + d] = np.float32(c["joint_target_vel"][6 + d] + sgn * h)
            taken.append(float(c["joint_target_vel"][6 + d]))
            vals.append(rollout_loss(oracle_mod, model, cfg, state, c, w, steps))
        g[d] = (vals[0] - vals[1]) / (taken[0] - taken[1])
    return g


def rel_err(a, f):
    return np.abs(a - f) / np.maximum(np.maximum(np.abs(a), np.abs(f)), 1e-10)


def oracle_adjoint_wheel_gradient(oracle_mod, model, cfg, state, control, w, steps=1):
    """Forward `steps` steps, then the adjoint chain back (the reference's trajectory
    buffer loop, test_wheeled_robot.py:257-266): dL/d(target_vel) summed over steps."""
    recs, st = [], dict(state)
    for _ in range(steps):
        c = contacts_of(model, st["body_q"])
        o = oracle_mod.step(model, cfg, st, control, c, DT, precision=64)
        recs.append((st, o, c))
        st = dict(st, body_q=o["body_q"], body_qd=o["body_qd"])
    B = model.bodies_per_world
    gq, gu = np.zeros((B, 7)), w.reshape(B, 6).copy()
    total = np.zeros(model.dofs_per_world)
    for st, o, c in reversed(recs):
        g = oracle_mod.adjoint(model, cfg, st, dict(body_q=o["body_q"], body_qd=o["body_qd"]), control, c, DT,
                               o["lam"], o["lam_prev"], gq, gu, precision=64, fd_eps=1e-6)
        total += g["joint_target_vel"].reshape(-1)
        gq, gu = g["body_q"].reshape(B, 7), g["body_qd"].reshape(B, 6)
    return total[6:10], recs


@pytest.mark.parametrize("name", list(CASES))
def test_wheel_velocity_gradient_matches_fd(oracle_mod, name):
    model, state, control, w, band = taros_case(name)
    cfg = reference_config()
    ga, recs = oracle_adjoint_wheel_gradient(oracle_mod, model, cfg, state, control, w)
    o = recs[0][1]
    assert o["merit"][0] <= cfg.nr.atol ** 2  # a converged step: the adjoint's premise
    fd = fd_wheel_gradient(oracle_mod, model, cfg, state, control, w)
    err = rel_err(ga, fd)
    print(f"[{name}] adjoint {np.round(ga, 5)} fd {np.round(fd, 5)} rel {np.round(err, 5)}")
    assert np.abs(fd).max() > 0.05
    assert err.max() < band
    assert err.max() < 1e-5


def test_multi_step_wheel_velocity_gradient_matches_fd(oracle_mod):
    """Five steps at 5 rad/s, seed 42, h 1e-3 (test_wheeled_robot.py:223-336)."""
    model, state, control, w, band = taros_case("straight")
    cfg = reference_config()
    ga, _ = oracle_adjoint_wheel_gradient(oracle_mod, model, cfg, state, control, w, steps=5)
    fd = fd_wheel_gradient(oracle_mod, model, cfg, state, control, w, steps=5, h=1e-3)
    err = rel_err(ga, fd)
    print(f"[multi-step] adjoint {np.round(ga, 5)} fd {np.round(fd, 5)} rel {np.round(err, 5)}")
    assert err.max() < 0.15
    assert err.max() < 1e-4  # h 1e-3: truncation of the differences
