"""Pins the CPU oracle (oracle/fb_oracle.cpp) before it is trusted as the checker.

Each test ties a piece of the restatement to something outside it: a closed form, an
invariant, a brute-force dense computation, an independent textbook algorithm, or a
known-answer band taken from the reference's own tests.  The reference ships no
numeric fixtures (its tests assert bands and finite-difference agreement only).
"""
import math

import numpy as np
import pytest
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.model import (DOF_TARGET_VELOCITY, GEO_BOX, GEO_PLANE, IDENTITY7, Model, Shape, WorldBuilder,
                               box_inertia, cylinder_inertia_y)
from tests.helpers import config, no_contacts, rollout, single_box, zero_control

G = 9.81


# ---------------------------------------------------------------- FB function -----
@pytest.mark.parametrize("a,b", [(0.3, 0.0), (0.0, 2.0), (0.0, 0.0), (1.5, 0.0)])
def test_fb_zero_exactly_on_complementarity_set(oracle_mod, a, b):
    """phi(a,b) = 0 iff a >= 0, b >= 0, a b = 0 (eps = 0): the defining NCP property."""
    v = oracle_mod.fb_diff(a, b, 1.0, 1.0, 0.0)[0]
    assert abs(v) < 1e-12


@pytest.mark.parametrize("a,b", [(0.4, 0.7), (-0.2, 0.5), (0.5, -0.1), (-0.3, -0.2)])
def test_fb_nonzero_off_complementarity_set(oracle_mod, a, b):
    v = oracle_mod.fb_diff(a, b, 1.0, 1.0, 0.0)[0]
    assert abs(v) > 1e-3
    # sign: positive inside the positive quadrant, negative if either argument is negative
    assert (v > 0) == (a > 0 and b > 0)


def test_fb_symmetry_and_derivatives(oracle_mod):
    rng = np.random.default_rng(0)
    for _ in range(50):
        a, b = rng.normal(size=2)
        al, be = rng.uniform(0.1, 3.0, size=2)
        eps = 10.0 ** rng.uniform(-8, -2)
        v, da, db = oracle_mod.fb_diff(a, b, al, be, eps)
        v2, da2, db2 = oracle_mod.fb_diff(b, a, be, al, eps)
        assert v == pytest.approx(v2, abs=1e-14)
        assert da == pytest.approx(db2, abs=1e-14)
        h = 1e-6
        fd_a = (oracle_mod.fb_diff(a + h, b, al, be, eps)[0] - oracle_mod.fb_diff(a - h, b, al, be, eps)[0]) / (2 * h)
        fd_b = (oracle_mod.fb_diff(a, b + h, al, be, eps)[0] - oracle_mod.fb_diff(a, b - h, al, be, eps)[0]) / (2 * h)
        assert da == pytest.approx(fd_a, rel=1e-5, abs=1e-7)
        assert db == pytest.approx(fd_b, rel=1e-5, abs=1e-7)


# ----------------------------------------------------------- pose integration -----
def _quat_from_axis_angle(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.concatenate([axis * math.sin(ang / 2), [math.cos(ang / 2)]])


def _qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + bw * ax + ay * bz - az * by, aw * by + bw * ay + az * bx - ax * bz,
                     aw * bz + bw * az + ax * by - ay * bx, aw * bw - ax * bx - ay * by - az * bz])


def _rot(q, v):
    qv = np.concatenate([v, [0.0]])
    qc = np.array([-q[0], -q[1], -q[2], q[3]])
    return _qmul(_qmul(q, qv), qc)[:3]


def test_integrate_pose_zero_velocity_is_identity(oracle_mod):
    q = _quat_from_axis_angle([1, 2, 3], 0.7)
    pose = np.concatenate([[0.3, -0.2, 1.0], q])
    out = oracle_mod.integrate_pose(np.zeros(6), pose, [0.1, 0.2, -0.05], 0.05)
    np.testing.assert_allclose(out, pose, atol=1e-14)


def test_integrate_pose_com_moves_with_linear_velocity(oracle_mod):
    """Implicit-Euler position update of the centre of mass: p_com += v dt, orientation kept."""
    q = _quat_from_axis_angle([0.3, -1, 0.2], 1.1)
    com = np.array([0.1, -0.2, 0.05])
    pose = np.concatenate([[0.5, 0.1, 2.0], q])
    v = np.array([1.5, -0.3, 0.7])
    dt = 0.05
    out = oracle_mod.integrate_pose(np.concatenate([v, [0, 0, 0]]), pose, com, dt)
    np.testing.assert_allclose(out[3:], q, atol=1e-14)
    com_before = pose[:3] + _rot(q, com)
    com_after = out[:3] + _rot(out[3:], com)
    np.testing.assert_allclose(com_after - com_before, v * dt, atol=1e-14)


def test_integrate_pose_rotation_matches_exponential_map(oracle_mod):
    """Angular update is first order in dt: error to exp(w dt) is O(dt^2), unit norm."""
    q0 = _quat_from_axis_angle([1, 0, 1], 0.4)
    w = np.array([0.4, -1.2, 2.0])
    for dt in (1e-2, 5e-3):
        out = oracle_mod.integrate_pose(np.concatenate([[0, 0, 0], w]), np.concatenate([[0, 0, 0], q0]),
                                        np.zeros(3), dt)
        ang = np.linalg.norm(w) * dt
        exact = _qmul(_quat_from_axis_angle(w, ang), q0)  # world-frame angular velocity
        err = min(np.linalg.norm(out[3:] - exact), np.linalg.norm(out[3:] + exact))
        assert abs(np.linalg.norm(out[3:]) - 1.0) < 1e-14
        assert err < 0.6 * (np.linalg.norm(w) * dt) ** 2


# ------------------------------------------------------------- free bodies -------
def test_free_fall_closed_form(oracle_mod):
    """No contacts: M(u - u_prev) = dt (f + m g)  =>  u = u_prev + dt (g + f/m)."""
    model, state = single_box(mass=2.0, pose=(0, 0, 3.0, 0, 0, 0, 1))
    state["body_qd"][0] = [0.5, -0.2, 0.1, 0, 0, 0]
    state["body_f"][0] = [4.0, 0.0, -2.0, 0, 0, 0]
    cfg = config(nr_iters=4, atol=1e-9, K=4)
    dt = 0.05
    o = oracle_mod.step(model, cfg, state, zero_control(model), no_contacts(model), dt, precision=64)
    u = np.array([0.5, -0.2, 0.1]) + dt * (np.array([0, 0, -G]) + np.array([4.0, 0.0, -2.0]) / 2.0)
    np.testing.assert_allclose(o["body_qd"][0, :3], u, atol=1e-12)
    np.testing.assert_allclose(o["body_q"][0, :3], np.array([0, 0, 3.0]) + dt * u, atol=1e-12)


def test_symmetric_spin_is_constant(oracle_mod):
    """Isotropic inertia: w x (I w) = 0, so a free cube keeps its angular velocity."""
    model, state = single_box(pose=(0, 0, 3.0, 0, 0, 0, 1))
    state["body_qd"][0] = [0, 0, 0, 0.3, 1.0, -0.5]
    o = oracle_mod.step(model, config(nr_iters=4, K=4), state, zero_control(model), no_contacts(model), 0.02,
                        precision=64)
    np.testing.assert_allclose(o["body_qd"][0, 3:], [0.3, 1.0, -0.5], atol=1e-12)


# ------------------------------------------------------------- resting contact ---
def _world_pairs(model, contacts, w):
    """(shape0, shape1) of the contacts of world w in packing (input) order."""
    sw = model.arrays["shape_world"]
    out = []
    n = int(np.asarray(contacts["count"]).reshape(-1)[0])
    s0 = np.asarray(contacts["shape0"]); s1 = np.asarray(contacts["shape1"])
    for i in range(n):
        a, b = int(s0[i]), int(s1[i])
        ww = sw[a] if sw[a] >= 0 else sw[b]
        if ww == w and a != b:
            out.append((a, b))
    return out


def test_resting_stack_supports_weight_per_interface(oracle_mod):
    """Converged normal impulses of a 3-cube stack: each interface carries the weight above it."""
    model, state, control = scenes.box_stack(2, 3, seed=0)
    cfg = config(nr_iters=30, atol=1e-7, pcr_iters=64, tol=1e-9, patol=1e-10, K=16)
    st, outs = rollout(oracle_mod, model, cfg, state, control, steps=3, dt=0.05, K=16, gap=0.05)
    col = collide.PrimitiveCollider(model, gap=0.05, max_contacts_per_world=16)
    c = col.collide(torch.as_tensor(st["body_q"], dtype=torch.float32))
    c = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}
    o = oracle_mod.step(model, cfg, st, control, c, 0.05, precision=64)
    nc, nj, nctrl = oracle_mod.dims(model, cfg)
    dt, m = 0.05, 1.0
    for w in range(2):
        pairs = _world_pairs(model, c, w)
        lam = o["lam"][w, nj + nctrl: nj + nctrl + 16]
        per = {}
        for k, p in enumerate(pairs):
            per[p] = per.get(p, 0.0) + lam[k]
        totals = sorted(per.values())
        np.testing.assert_allclose(totals, [1 * m * G, 2 * m * G, 3 * m * G], rtol=2e-3)
        assert np.abs(o["body_qd"][3 * w: 3 * w + 3]).max() < 1e-3


# ----------------------------------------------------- operator / PCR brute force -
def _c1_system(oracle_mod, W=3, seed=3, **kw):
    model, state, control = scenes.skid_steer(W, seed=seed)
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=32)
    c = col.collide(torch.as_tensor(state["body_q"], dtype=torch.float32))
    c = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}
    state = dict(state)
    rng = np.random.default_rng(seed)
    state["body_qd"] = rng.normal(scale=0.3, size=state["body_qd"].shape)
    cfg = config(K=32, **kw)
    return model, state, control, c, cfg


def test_operator_matches_dense_schur_matrix(oracle_mod):
    model, state, control, c, cfg = _c1_system(oracle_mod)
    nc, _, _ = oracle_mod.dims(model, cfg)
    x = np.random.default_rng(1).normal(size=(model.num_worlds, nc))
    ls = oracle_mod.linear_system(model, cfg, state, control, c, 0.05, x=x)
    for w in range(model.num_worlds):
        A, act = ls["A"][w], ls["active"][w] > 0
        assert act.sum() > 10
        np.testing.assert_allclose(A, A.T, atol=1e-9)
        xa = np.where(act, x[w], 0.0)
        np.testing.assert_allclose(ls["Ax"][w], A @ xa, rtol=1e-10, atol=1e-9)
        ev = np.linalg.eigvalsh(A[np.ix_(act, act)])
        assert ev.min() > 0.0
        np.testing.assert_allclose(ls["pinv"][w][act], 1.0 / (np.diag(A)[act] + 1e-6), rtol=1e-12)


def _textbook_pcr(A, b, Pdiag, iters):
    """Preconditioned conjugate residual (Saad, Iterative Methods, Alg. 6.20 with M^-1 = Pdiag)."""
    x = np.zeros_like(b)
    r = b - A @ x
    z = Pdiag * r
    p = z.copy()
    Az = A @ z
    Ap = Az.copy()
    for _ in range(iters):
        y = Pdiag * Ap
        alpha = (z @ Az) / (y @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z_new = z - alpha * y
        Az_new = A @ z_new
        beta = (z_new @ Az_new) / (z @ Az)
        z, Az = z_new, Az_new
        p = z + beta * p
        Ap = Az + beta * Ap
    return x


def test_pcr_matches_textbook_and_dense_solve(oracle_mod):
    model, state, control, c, _ = _c1_system(oracle_mod)
    cfg3 = config(K=32, pcr_iters=3, tol=0.0, patol=0.0, termination="per_world")
    ls3 = oracle_mod.linear_system(model, cfg3, state, control, c, 0.05)
    cfgn = config(K=32, pcr_iters=400, tol=1e-13, patol=0.0, termination="per_world")
    lsn = oracle_mod.linear_system(model, cfgn, state, control, c, 0.05)
    for w in range(model.num_worlds):
        act = ls3["active"][w] > 0
        A = ls3["A"][w][np.ix_(act, act)]
        b = ls3["rhs"][w][act]
        P = ls3["pinv"][w][act]
        ref3 = _textbook_pcr(A, b, P, 3)
        np.testing.assert_allclose(ls3["pcr_x"][w][act], ref3, rtol=1e-8, atol=1e-10 * np.abs(ref3).max())
        exact = np.linalg.solve(A, b)
        np.testing.assert_allclose(lsn["pcr_x"][w][act], exact, rtol=1e-6, atol=1e-8 * np.abs(exact).max())
        assert np.all(lsn["pcr_x"][w][~act] == 0.0)


def test_block_jacobi_pcr_solves_same_system(oracle_mod):
    """Per-body-pair block preconditioner changes the iterates, not the solution."""
    model, state, control, c, _ = _c1_system(oracle_mod)
    cfgb = config(K=32, pcr_iters=400, tol=1e-13, patol=0.0, precond="per_body_pair")
    lsb = oracle_mod.linear_system(model, cfgb, state, control, c, 0.05)
    for w in range(model.num_worlds):
        act = lsb["active"][w] > 0
        exact = np.linalg.solve(lsb["A"][w][np.ix_(act, act)], lsb["rhs"][w][act])
        np.testing.assert_allclose(lsb["pcr_x"][w][act], exact, rtol=1e-6, atol=1e-8 * np.abs(exact).max())


def test_block_jacobi_exact_for_single_pair(oracle_mod):
    """A world whose active rows all share one body pair: the block preconditioner is A^-1,
    so PCR converges in one iteration."""
    model, state = single_box(pose=(0, 0, 0.245, 0, 0, 0, 1))
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=8)
    c = col.collide(torch.as_tensor(state["body_q"], dtype=torch.float32))
    c = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}
    cfg1 = config(K=8, pcr_iters=1, tol=0.0, patol=0.0, precond="per_body_pair")
    ls = oracle_mod.linear_system(model, cfg1, state, zero_control(model), c, 0.05)
    act = ls["active"][0] > 0
    exact = np.linalg.solve(ls["A"][0][np.ix_(act, act)], ls["rhs"][0][act])
    np.testing.assert_allclose(ls["pcr_x"][0][act], exact, rtol=1e-8)


# ------------------------------------------------------- reference test bands ----
def _joint_world(jtype):
    """Cube (half 0.5, 1000 kg) hanging from the world by one joint at z=5
    (tests/test_joint_constraints.py:140-200)."""
    wb = WorldBuilder()
    b = wb.add_body(1000.0, box_inertia(1000.0, 0.5, 0.5, 0.5))
    Xp = (0.0, 0.0, 5.0, 0, 0, 0, 1)
    Xc = (0.0, 0.0, 1.0, 0, 0, 0, 1)
    if jtype == "revolute":
        wb.add_joint_revolute(-1, b, Xp, Xc, axis=(0.0, 1.0, 0.0))
    elif jtype == "prismatic":
        wb.add_joint_prismatic(-1, b, Xp, Xc, axis=(0.0, 1.0, 0.0))
    elif jtype == "fixed":
        wb.add_joint_fixed(-1, b, Xp, Xc)
    else:
        wb.add_joint_ball(-1, b, Xp, Xc)
    model = Model.from_worlds([wb], [])
    state = dict(body_q=np.array([[0, 0, 4.0, 0, 0, 0, 1.0]]), body_qd=np.zeros((1, 6)), body_f=np.zeros((1, 6)))
    return model, state


def _frames(q):
    p, quat = q[:3], q[3:]
    parent_pos, parent_q = np.array([0, 0, 5.0]), np.array([0, 0, 0, 1.0])
    child_pos = p + _rot(quat, np.array([0, 0, 1.0]))
    return parent_pos, parent_q, child_pos, quat


@pytest.mark.parametrize("jtype", ["revolute", "prismatic", "fixed", "ball"])
def test_joint_drift_under_heavy_load(oracle_mod, jtype):
    """tests/test_joint_constraints.py:140-267: 30 steps, dt 0.01, 1e4 N load, compliance
    1e-8: position and axis error < 1e-3; free DOFs move by > 1e-2."""
    from ostrich_amd.config import ComplianceConfig
    model, state = _joint_world(jtype)
    cfg = config(nr_iters=10, atol=1e-3, pcr_iters=10, K=1, compliance=ComplianceConfig(joint=1e-8))
    st = {k: v.copy() for k, v in state.items()}
    q0 = st["body_q"][0].copy()
    for _ in range(30):
        st["body_f"][0] = [1e4, 1e4, 1e4, 0, 0, 0]
        o = oracle_mod.step(model, cfg, st, zero_control(model), no_contacts(model), 0.01, precision=64)
        st["body_q"], st["body_qd"] = o["body_q"], o["body_qd"]
    pp, pq, cp, cq = _frames(st["body_q"][0])
    pos_err = np.linalg.norm(cp - pp) if jtype != "prismatic" else np.linalg.norm(np.delete(cp - pp, 1))
    ax = np.array([0, 1.0, 0])
    if jtype == "revolute":
        ang_err = np.linalg.norm(_rot(cq, ax) - ax)
        motion = abs(2 * math.atan2(np.dot(cq[:3], ax), cq[3]))
    elif jtype == "ball":
        ang_err = 0.0
        motion = np.linalg.norm(cq[:3])
    else:
        ang_err = 2 * np.linalg.norm(cq[:3])
        motion = abs((cp - pp)[1]) if jtype == "prismatic" else 1.0
    assert pos_err < 1e-3
    assert ang_err < 1e-3
    assert motion > 1e-2


def _aniso_box(mu_box, ground_mu, axis=(0.0, 0.0, 0.0), mu_perp=-1.0):
    model, state = single_box(mass=1.0, h=0.25, mu=mu_box, ground_mu=ground_mu,
                              friction_axis=axis, mu_perp=mu_perp)
    return model, state


def test_anisotropic_friction_slides_along_low_mu_axis(oracle_mod):
    """tests/test_anisotropic_friction.py:94-139: with a zero-friction ground, an
    anisotropic box pushed diagonally slides further along its low-mu axis."""
    model, state = _aniso_box(1.0, 0.0, axis=(1.0, 0.0, 0.0), mu_perp=0.05)
    cfg = config(nr_iters=20, atol=1e-5, pcr_iters=32, K=8)
    st = {k: v.copy() for k, v in state.items()}
    st["body_q"][0, 2] = 0.249
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=8)
    for _ in range(20):
        st["body_f"][0] = [6.0, 6.0, 0, 0, 0, 0]
        c = col.collide(torch.as_tensor(st["body_q"], dtype=torch.float32))
        o = oracle_mod.step(model, cfg, st, zero_control(model), c, 0.01, precision=64)
        st["body_q"], st["body_qd"] = o["body_q"], o["body_qd"]
    dx, dy = st["body_q"][0, 0], st["body_q"][0, 1]
    assert dy > 2.0 * abs(dx)


def test_subthreshold_force_sticks(oracle_mod):
    """tests/test_anisotropic_friction.py:141-159: a push below mu m g does not slide the box."""
    model, state = _aniso_box(0.8, 0.8)
    cfg = config(nr_iters=20, atol=1e-5, pcr_iters=32, K=8)
    st = {k: v.copy() for k, v in state.items()}
    st["body_q"][0, 2] = 0.249
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=8)
    for _ in range(20):
        st["body_f"][0] = [0.3 * 0.8 * G, 0, 0, 0, 0, 0]
        c = col.collide(torch.as_tensor(st["body_q"], dtype=torch.float32))
        o = oracle_mod.step(model, cfg, st, zero_control(model), c, 0.01, precision=64)
        st["body_q"], st["body_qd"] = o["body_q"], o["body_qd"]
    assert abs(st["body_q"][0, 0]) < 2e-3


def _stribeck_decel(oracle_mod, vx0, mss=2.0, vs=0.2, mu=0.5):
    """tests/test_stribeck_friction.py:113-146: 1 m cube of 100 kg settled 50 steps at
    dt 0.01, kicked to vx0, decel measured over one 0.005 s step."""
    from ostrich_amd.config import LinearSolverConfig, NewtonRaphsonConfig, OstrichEngineConfig, ContactsConfig
    model, state = single_box(mass=100.0, h=0.5, mu=mu, pose=(0, 0, 0.6, 0, 0, 0, 1),
                              mu_stiction_scale=mss, v_stribeck=vs)
    cfg = OstrichEngineConfig(nr=NewtonRaphsonConfig(max_iters=20),
                              linear=LinearSolverConfig(max_iters=200, tol=1e-8, atol=1e-8),
                              contacts=ContactsConfig(max_per_world=16))
    col = collide.PrimitiveCollider(model, gap=0.05, max_contacts_per_world=16)
    st = {k: v.copy() for k, v in state.items()}
    for _ in range(50):
        c = col.collide(torch.as_tensor(st["body_q"], dtype=torch.float32))
        o = oracle_mod.step(model, cfg, st, zero_control(model), c, 0.01, precision=64)
        st["body_q"], st["body_qd"] = o["body_q"], o["body_qd"]
    st["body_qd"] = st["body_qd"].copy()
    st["body_qd"][0, 0] = vx0
    c = col.collide(torch.as_tensor(st["body_q"], dtype=torch.float32))
    o = oracle_mod.step(model, cfg, st, zero_control(model), c, 0.005, precision=64)
    assert o["body_qd"][0, 0] > 0.0
    return (vx0 - o["body_qd"][0, 0]) / 0.005


def test_stribeck_deceleration_band_and_ratio(oracle_mod):
    """tests/test_stribeck_friction.py:173-178: 0.15 mu g < decel_fast < 1.2 mu g and
    decel_slow / decel_fast > 1.5."""
    mu = 0.5
    fast = _stribeck_decel(oracle_mod, 2.0)
    slow = _stribeck_decel(oracle_mod, 0.05)
    assert 0.15 * mu * G < fast < 1.2 * mu * G
    assert slow / fast > 1.5


# ------------------------------------------------------------- motor rows -------
def test_velocity_motor_closed_form(oracle_mod):
    """Unloaded wheel on a revolute joint to the world, velocity mode: with the joint rows
    holding the axle, I (qd - qd_prev) = lambda dt and qd - target + alpha lambda dt = 0,
    alpha = 1/(dt ke)  =>  qd = (target + alpha I qd_prev) / (1 + alpha I)."""
    from ostrich_amd.config import ComplianceConfig
    wb = WorldBuilder()
    b = wb.add_body(1.0, cylinder_inertia_y(1.0, 0.18, 0.05))
    wb.add_joint_revolute(-1, b, (0, 0, 1.0, 0, 0, 0, 1), axis=(0.0, 1.0, 0.0), mode=DOF_TARGET_VELOCITY,
                          ke=40.0)
    model = Model.from_worlds([wb], [], gravity=(0.0, 0.0, 0.0))
    state = dict(body_q=np.array([[0, 0, 1.0, 0, 0, 0, 1.0]]), body_qd=np.array([[0, 0, 0, 0, 0.5, 0.0]]),
                 body_f=np.zeros((1, 6)))
    control = dict(joint_target_pos=np.zeros(1, np.float32), joint_target_vel=np.array([3.0], np.float32))
    dt = 0.05
    cfg = config(nr_iters=10, atol=1e-9, pcr_iters=64, tol=1e-12, patol=1e-14, K=1,
                 compliance=ComplianceConfig(joint=1e-10))
    o = oracle_mod.step(model, cfg, state, control, no_contacts(model), dt, precision=64)
    I = 0.5 * 1.0 * 0.18 ** 2
    alpha = 1.0 / (dt * 40.0)
    qd = (3.0 + alpha * I * 0.5) / (1.0 + alpha * I)
    assert o["body_qd"][0, 4] == pytest.approx(qd, rel=1e-6)


# ------------------------------------------------------- loop semantics ---------
def test_backtracking_picks_min_merit_after_min_iter(oracle_mod):
    """core/backtracking.py + reductions.py:28-62: picked = argmin over [min_iter, n)."""
    model, state, control = scenes.skid_steer(4, seed=5)
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=16)
    c = col.collide(torch.as_tensor(state["body_q"], dtype=torch.float32))
    c = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}
    cfg = config(nr_iters=12, atol=1e-9, K=16, termination="global")
    o = oracle_mod.step(model, cfg, state, control, c, 0.05, precision=64)
    n = int(o["nr_iters"][0])
    for w in range(4):
        hist = o["merit_hist"][w, :n]
        k = 2 + int(np.argmin(hist[2:]))
        assert o["best_iter"][w] == k
        assert o["merit"][w] == hist[k]


def test_global_termination_runs_all_worlds_equally(oracle_mod):
    model, state, control = scenes.skid_steer(6, seed=2)
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=16)
    c = col.collide(torch.as_tensor(state["body_q"], dtype=torch.float32))
    c = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}
    og = oracle_mod.step(model, config(K=16, termination="global"), state, control, c, 0.05)
    op = oracle_mod.step(model, config(K=16, termination="per_world"), state, control, c, 0.05)
    assert len(set(og["nr_iters"].tolist())) == 1
    assert og["nr_iters"][0] == op["nr_iters"].max()


def test_contact_packing_order_and_overflow(oracle_mod):
    """core/contacts.py:25-99: per-world slots follow input order; overflow is flagged."""
    model, state, control = scenes.box_stack(3, 2, seed=1)
    col = collide.PrimitiveCollider(model, gap=0.05, max_contacts_per_world=16)
    c = col.collide(torch.as_tensor(state["body_q"], dtype=torch.float32))
    c = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}
    o = oracle_mod.step(model, config(K=16), state, control, c, 0.05)
    n = [len(_world_pairs(model, c, w)) for w in range(3)]
    assert o["contact_count"].tolist() == n
    o2 = oracle_mod.step(model, config(K=4), state, control, c, 0.05)
    assert all((f & 2) for f in o2["flags"])
