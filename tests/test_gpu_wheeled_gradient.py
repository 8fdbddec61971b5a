"""The reference's wheeled-robot gradient checks
(tests/differentiable_simulator/test_wheeled_robot.py:151-340) on the gfx950 adjoint,
driven the reference's way: engine.step, TrajectoryBuffer.save_step / load_step,
engine.step_backward() with no arguments, save_gradients / save_pose_gradients.  The
FP32 GPU gradient dL/d(wheel target velocity) against central differences of the FP64
oracle forward (h 1e-4; 1e-3 through five steps), within the reference's bands (0.15
straight and multi-step, 0.5 for the turns) and within 1e-4 relative (the FP64 oracle
adjoint meets them to 1e-5)."""
import numpy as np
import pytest
import torch

from ostrich_amd.containers import Contacts, Control, State
from ostrich_amd.differentiable import TrajectoryBuffer
from tests.test_oracle_wheeled_gradient import (CASES, DT, contacts_of, fd_wheel_gradient,
                                                oracle_adjoint_wheel_gradient, reference_config, rel_err,
                                                taros_case)

pytestmark = pytest.mark.gpu


def gpu_wheel_gradient(model, cfg, state, control, w, device, steps=1):
    from ostrich_amd import OstrichEngine
    eng = OstrichEngine(model, config=cfg, device=device, differentiable_simulation=True)
    ctl = Control.from_numpy({k: np.asarray(v, np.float32) for k, v in control.items()}, device)
    states = [State.from_numpy({k: np.asarray(v, np.float32) for k, v in state.items()}, device)]
    buf = TrajectoryBuffer()
    for i in range(steps):
        s_out = states[i].clone()
        con = Contacts.from_dict(contacts_of(model, states[i].body_q.cpu().numpy()), device)
        eng.step(states[i], s_out, ctl, con, DT)
        buf.save_step(i, eng.data)
        states.append(s_out)
    B = model.bodies_per_world
    buf.set_output_grad(steps, torch.zeros(B, 7, device=device),
                        torch.as_tensor(w.reshape(B, 6), dtype=torch.float32, device=device))
    total = torch.zeros(model.dofs_per_world, device=device)
    for i in reversed(range(steps)):
        buf.load_step(i, eng.data)
        eng.data.zero_gradients()
        eng.step_backward()
        buf.save_gradients(i, eng.data)
        buf.save_pose_gradients(i, eng.data)
        total += buf.input_grads[i]["joint_target_vel"].reshape(-1)
    return total.cpu().numpy().astype(np.float64)[6:10]


@pytest.mark.parametrize("name", list(CASES))
def test_wheel_velocity_gradient_matches_fd_gpu(oracle_mod, gpu_device, name):
    model, state, control, w, band = taros_case(name)
    cfg = reference_config()
    ga = gpu_wheel_gradient(model, cfg, state, control, w, gpu_device)
    go, _ = oracle_adjoint_wheel_gradient(oracle_mod, model, cfg, state, control, w)
    fd = fd_wheel_gradient(oracle_mod, model, cfg, state, control, w)
    err, err_o = rel_err(ga, fd), rel_err(go, fd)
    print(f"[{name}] gpu {np.round(ga, 5)} fd {np.round(fd, 5)} rel {np.round(err, 5)} "
          f"(fp64 adjoint rel {np.round(err_o, 5)})", flush=True)
    assert err.max() < band
    assert err_o.max() < 1e-5 and err.max() < 1e-4


def test_multi_step_wheel_velocity_gradient_matches_fd_gpu(oracle_mod, gpu_device):
    model, state, control, w, band = taros_case("straight")
    cfg = reference_config()
    ga = gpu_wheel_gradient(model, cfg, state, control, w, gpu_device, steps=5)
    fd = fd_wheel_gradient(oracle_mod, model, cfg, state, control, w, steps=5, h=1e-3)
    err = rel_err(ga, fd)
    print(f"[multi-step] gpu {np.round(ga, 5)} fd {np.round(fd, 5)} rel {np.round(err, 5)}", flush=True)
    assert err.max() < 0.15
    assert err.max() < 1e-4
