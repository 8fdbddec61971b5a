"""The structural relaxation modes (engine_config.py:188-287) on the gfx950 solver:
friction "bilateral" / "bilateral_patch" (friction_constraint.py:39-76, :284-295) and
contact "equality" (contact_constraint.py:112-124).  Parity with the CPU oracle on box
stacks, to the same bar as the cone/complementarity parity (converged worlds within 1e-4
of scale of the FP32 oracle, capped worlds within four times their FP32-vs-FP64 spread),
and the closed forms the oracle is pinned to, now computed on the GPU."""
import numpy as np
import pytest
import torch

from ostrich_amd import scenes
from ostrich_amd.config import RelaxationConfig
from tests.helpers import config, single_box, zero_control
from tests.test_gpu_fused import G, _engine, _gpu_step, _run_compare
from tests.test_oracle_relaxation import _collide, _sphere

pytestmark = pytest.mark.gpu

MODES = [("bilateral", "complementarity"), ("bilateral_patch", "complementarity"), ("cone", "equality"),
         ("bilateral", "equality")]


@pytest.mark.parametrize("friction,contact", MODES)
def test_box_stack_parity_relaxation_modes(oracle_mod, gpu_device, friction, contact):
    model, state, control = scenes.box_stack(16, 3, seed=0)
    cfg = config(K=16, relaxation=RelaxationConfig(friction=friction, contact=contact))
    nr_diff, conv = _run_compare(oracle_mod, model, cfg, state, control, steps=4, dt=0.05, device=gpu_device, K=16)
    print(f"[relaxation {friction}/{contact}] converged fraction {conv.mean():.2f}, "
          f"|nr gpu - nr f32| <= 1 in {np.mean(nr_diff <= 1) if nr_diff.size else 1.0:.2f} of converged",
          flush=True)
    if nr_diff.size:
        assert np.mean(nr_diff <= 1) >= 0.9


def test_bilateral_friction_rolls_a_sliding_sphere_gpu(gpu_device):
    r, v0, dt = 0.1, 2.0, 0.01
    model = _sphere(r=r)
    st = dict(body_q=np.array([[0, 0, r, 0, 0, 0, 1.0]], np.float32),
              body_qd=np.array([[v0, 0, 0, 0, 0, 0]], np.float32), body_f=np.zeros((1, 6), np.float32))
    for friction in ("bilateral", "bilateral_patch"):
        cfg = config(K=1, nr_iters=40, atol=1e-6, pcr_iters=64,
                     relaxation=RelaxationConfig(friction=friction))
        g = _gpu_step(_engine(model, cfg, gpu_device), st, zero_control(model), _collide(model, st, 1), dt,
                      gpu_device)
        v, w = g["body_qd"][0, 0], g["body_qd"][0, 4]
        assert v == pytest.approx(v0 / 1.4, rel=1e-4)
        assert abs(v - w * r) < 1e-4


def test_equality_contact_holds_a_box_that_is_pulled_up_gpu(oracle_mod, gpu_device):
    m, v0, dt = 1.0, 1.0, 0.01
    model, state = single_box(mass=m, h=0.25, mu=0.5)
    st = {k: np.asarray(v, np.float32).copy() for k, v in state.items()}
    st["body_qd"][0] = [0.0, 0.0, v0, 0.0, 0.0, 0.0]
    c = _collide(model, st, 8)
    cfg = config(K=8, nr_iters=40, atol=1e-6, pcr_iters=64, relaxation=RelaxationConfig(contact="equality"))
    g = _gpu_step(_engine(model, cfg, gpu_device), st, zero_control(model), c, dt, gpu_device)
    _, nj, nctrl = oracle_mod.dims(model, cfg)
    n = int(oracle_mod.step(model, cfg, st, zero_control(model), c, dt, precision=32)["contact_count"][0])
    assert n == 4 and int(np.asarray(c["count"])[0]) == 4
    lam_n = g["lam"][0, nj + nctrl: nj + nctrl + n]
    assert abs(g["body_q"][0, 2] - 0.25) < 1e-5
    assert np.abs(g["body_qd"][0]).max() < 1e-3
    assert lam_n.sum() == pytest.approx(-m * (v0 - G * dt) / dt, rel=1e-4)
