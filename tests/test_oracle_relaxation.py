"""Pins of the oracle's structural relaxation modes (engine_config.py:188-287,
RELAXATION_BRANCH.md), each against what the mode defines:

* friction "bilateral" (friction_constraint.py:284-295: w = 0, rows v_t + 0 lambda = 0):
  every loaded contact point stops slipping, so a sphere sliding without spin starts
  rolling with v = v0 / (1 + I / (m r^2)) = 5/7 v0 in one step (the textbook result);
* friction "bilateral_patch" (:39-76): only the first contact of each shape pair carries
  friction rows; that point stops slipping, the others carry no friction impulse;
* contact "equality" (contact_constraint.py:112-124: phi = signed distance): a box pulled
  off the ground stays on it, the normal forces summing to exactly the force that stops it,
  -m (v0 - g dt) / dt, where the complementarity mode lets it fly with v0 - g dt and normal
  forces that vanish with the Fischer-Burmeister smoothing (linearly in eps^2).
"""
import math

import numpy as np
import pytest
import torch

from ostrich_amd import collide
from ostrich_amd.config import ComplianceConfig, RelaxationConfig
from ostrich_amd.model import GEO_PLANE, GEO_SPHERE, IDENTITY7, Model, Shape, WorldBuilder, sphere_inertia
from tests.helpers import config, single_box, zero_control

G = 9.81
_TIGHT = dict(nr_iters=40, atol=1e-9, pcr_iters=200, tol=1e-12, patol=1e-14)


def _cfg(K, friction="cone", contact="complementarity", nr_iters=40, **kw):
    tight = dict(_TIGHT, nr_iters=nr_iters)
    return config(K=K, relaxation=RelaxationConfig(friction=friction, contact=contact), **tight, **kw)


def _collide(model, state, K, gap=0.02):
    col = collide.PrimitiveCollider(model, gap=gap, max_contacts_per_world=K)
    c = col.collide(torch.as_tensor(state["body_q"], dtype=torch.float32))
    return {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}


def _sphere(mu=0.8, r=0.1, m=1.0):
    wb = WorldBuilder()
    b = wb.add_body(m, sphere_inertia(m, r))
    wb.add_shape(Shape(GEO_SPHERE, b, IDENTITY7, (r, 0.0, 0.0), mu=mu))
    wb.add_joint_free(b)
    return Model.from_worlds([wb], [Shape(GEO_PLANE, -1, IDENTITY7, (0, 0, 0), mu=mu)])


@pytest.mark.parametrize("friction", ["bilateral", "bilateral_patch"])
def test_bilateral_friction_rolls_a_sliding_sphere(oracle_mod, friction):
    r, v0, dt = 0.1, 2.0, 0.01
    model = _sphere(r=r)
    st = dict(body_q=np.array([[0, 0, r, 0, 0, 0, 1.0]]), body_qd=np.array([[v0, 0, 0, 0, 0, 0.0]]),
              body_f=np.zeros((1, 6)))
    o = oracle_mod.step(model, _cfg(1, friction), st, zero_control(model), _collide(model, st, 1), dt, precision=64)
    v, w = o["body_qd"][0, 0], o["body_qd"][0, 4]
    assert v == pytest.approx(v0 / (1.0 + 0.4), rel=1e-5)  # 5/7 v0 (I = 2/5 m r^2)
    assert v - w * r == pytest.approx(0.0, abs=1e-7)       # contact point at rest
    # the cone mode instead lets it keep sliding (only mu g dt of speed lost)
    oc = oracle_mod.step(model, _cfg(1), st, zero_control(model), _collide(model, st, 1), dt, precision=64)
    assert oc["body_qd"][0, 0] > 0.95 * v0


def test_bilateral_patch_keeps_friction_on_the_first_point_of_a_pair(oracle_mod):
    """A box drifting at 1 cm/s on its four bottom corners (one shape pair): with
    bilateral_patch only the first contact carries friction; it stops slipping and the
    other three carry none.  (Pinning one corner of a box sliding fast asks Newton for a
    sudden yaw; from 0.2 m/s up the lagged load gate stalls it, so the drift is slow.)"""
    model, state = single_box(mass=1.0, h=0.25, mu=0.8)
    st = {k: v.copy() for k, v in state.items()}
    st["body_qd"][0] = [0.01, 0.0, 0.0, 0.0, 0.0, 0.0]
    c = _collide(model, st, 8)
    o = oracle_mod.step(model, _cfg(8, "bilateral_patch", nr_iters=60), st, zero_control(model), c, 0.01,
                        precision=64)
    n = int(o["contact_count"][0])
    assert n == 4 and len(set(zip(np.ravel(c["shape0"])[:n], np.ravel(c["shape1"])[:n]))) == 1
    assert o["merit"][0] < 1e-10
    _, nj, nctrl = oracle_mod.dims(model, _cfg(8))
    K = 8
    lam_t = o["lam"][0, nj + nctrl + K: nj + nctrl + K + 2 * n].reshape(n, 2)
    assert np.linalg.norm(lam_t[0]) > 0.1
    assert np.abs(lam_t[1:]).max() == 0.0
    # velocity of the first contact point after the step (step-start lever arm)
    u = o["body_qd"][0]
    q0 = st["body_q"][0]
    lever = _rotate(q0[3:], np.asarray(c["point0"][0], float))
    vp = u[:3] + np.cross(u[3:], lever)
    assert np.abs(vp[:2]).max() < 1e-5
    # in the bilateral mode every corner stops slipping
    ob = oracle_mod.step(model, _cfg(8, "bilateral"), st, zero_control(model), c, 0.01, precision=64)
    ub = ob["body_qd"][0]
    for k in range(n):
        vk = ub[:3] + np.cross(ub[3:], _rotate(q0[3:], np.asarray(c["point0"][k], float)))
        assert np.abs(vk[:2]).max() < 1e-5


def _rotate(q, v):
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return R @ v


def test_equality_contact_holds_a_box_that_is_pulled_up(oracle_mod):
    """A resting box given 1 m/s upward.  Equality: it stays on the ground and the four
    normal forces sum to -m (v0 - g dt) / dt (what stops it in one step).  Complementarity:
    it flies with v0 - g dt, and its normal forces are the smoothing's residue, linear in
    eps^2 (0 in the limit)."""
    m, v0, dt = 1.0, 1.0, 0.01
    model, state = single_box(mass=m, h=0.25, mu=0.5)
    st = {k: v.copy() for k, v in state.items()}
    st["body_qd"][0] = [0.0, 0.0, v0, 0.0, 0.0, 0.0]
    c = _collide(model, st, 8)
    _, nj, nctrl = oracle_mod.dims(model, _cfg(8))
    oe = oracle_mod.step(model, _cfg(8, contact="equality"), st, zero_control(model), c, dt, precision=64)
    n = int(oe["contact_count"][0])
    assert n == 4
    lam_n = oe["lam"][0, nj + nctrl: nj + nctrl + n]
    assert abs(oe["body_q"][0, 2] - 0.25) < 1e-8
    assert np.abs(oe["body_qd"][0]).max() < 1e-6
    assert lam_n.sum() == pytest.approx(-m * (v0 - G * dt) / dt, rel=1e-6)
    assert (lam_n < 0).all()
    res = []
    for eps_sq in (1e-8, 1e-10):
        oc = oracle_mod.step(model, _cfg(8, compliance=ComplianceConfig(contact_fb_smooth_eps_sq=eps_sq)), st,
                             zero_control(model), c, dt, precision=64)
        assert oc["body_qd"][0, 2] == pytest.approx(v0 - G * dt, abs=1e-6)
        assert oc["body_q"][0, 2] == pytest.approx(0.25 + (v0 - G * dt) * dt, abs=1e-8)
        res.append(oc["lam"][0, nj + nctrl: nj + nctrl + n])
    assert (res[0] > 0).all() and res[0].max() < 0.01 * m * G
    np.testing.assert_allclose(res[0] / res[1], 100.0, rtol=1e-3)
