"""Step logs held on the device and written to disk.

* `DeviceLog` - the history buffers of core/history_group.py:12-133: a set of device
  arrays copied into slot `step` of their [capacity, ...] buffers by `ofb_log_capture`
  (one gfx950 launch for all arrays), with the step counter on the device, so a capture
  can be recorded into the same graph as the step it follows.
* `SimulationLog` - logging/sim_logger.py:12-218 (SimulationHDF5Logger): per step the
  dynamics state, the constraint impulses and the solver history (Newton iterations,
  ||h||^2, picked candidate, flags, and per Newton iteration ||h||^2 and PCR iterations).
* `DatasetLog` - logging/dataset_logger.py:16-141 (DatasetHDF5Logger): the state-only
  training log (state, joint targets, impulses, the step's input contacts).
* `load_log` - the reader (logging/hdf5_reader.py).

The reference writes HDF5 (h5py, groups data/, model/, dims/ and attributes num_steps,
dt).  h5py is not part of this environment, so the same groups are written as keys
"data/<name>", "model/<name>", "dims/<name>", "attrs/<name>" of a compressed numpy .npz
archive; `load_log` returns them as nested dicts.
"""
import ctypes as C
from typing import Dict, Optional

import numpy as np
import torch

from . import _lib


class DeviceLog:
    """Buffers [capacity, *src.shape] for named 4-byte device tensors, filled by `capture`
    at the device-side step index (`step_index`, int32 [1]); captures at an index outside
    [0, capacity) write nothing (history_group.py:29-31)."""

    def __init__(self, capacity: int, sources: Dict[str, torch.Tensor], step_index: Optional[torch.Tensor] = None):
        if capacity < 1 or not sources:
            raise ValueError("a log needs capacity >= 1 and at least one array")
        self.lib = _lib.load()
        self.capacity = capacity
        dev = next(iter(sources.values())).device
        for n, t in sources.items():
            if t.device != dev or t.element_size() != 4 or not t.is_contiguous() or t.numel() == 0:
                raise ValueError(f"log source {n}: needs a non-empty contiguous 4-byte tensor on {dev}")
        self.device = dev
        self.sources = dict(sources)
        self.buffers = {n: torch.zeros((capacity,) + tuple(t.shape), dtype=t.dtype, device=dev)
                        for n, t in sources.items()}
        self.step_index = step_index if step_index is not None else torch.zeros(1, dtype=torch.int32, device=dev)
        names = list(sources)
        n = len(names)
        src = (C.c_void_p * n)(*[sources[k].data_ptr() for k in names])
        dst = (C.c_void_p * n)(*[self.buffers[k].data_ptr() for k in names])
        words = (C.c_int64 * n)(*[sources[k].numel() for k in names])
        self._h = C.c_void_p()
        _lib.check(self.lib.ofb_log_create(n, src, dst, words, capacity, C.byref(self._h)), "ofb_log_create")

    def capture(self, advance: bool = True):
        """Copies every source into slot step_index (then step_index += 1 with advance)."""
        stream = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check(self.lib.ofb_log_capture(self._h, C.c_void_p(self.step_index.data_ptr()), 1 if advance else 0,
                                            stream), "ofb_log_capture")

    def numpy(self):
        return {k: v.cpu().numpy() for k, v in self.buffers.items()}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self.lib.ofb_log_destroy(h)
            self._h = None


def _model_arrays(model):
    out = {k: np.asarray(v) for k, v in model.arrays.items()}
    for k in ("shape_geo", "shape_xform", "shape_size", "shape_body_2d"):
        if hasattr(model, k):
            out[k] = np.asarray(getattr(model, k))
    return out


def _dims(engine):
    m = engine.model
    return dict(num_worlds=m.num_worlds, bodies_per_world=m.bodies_per_world, joints_per_world=m.joints_per_world,
                dofs_per_world=m.dofs_per_world, shapes_per_world=m.shapes_per_world,
                num_global_shapes=m.num_global_shapes, num_constraints=engine.num_constraints,
                contacts_per_world=engine.config.contacts.max_per_world,
                newton_max_iters=engine.config.nr.max_iters)


def save_log(path, num_steps, dt, data, model_arrays, dims):
    """Writes the data/, model/, dims/ groups and the num_steps / dt attributes."""
    flat = {f"data/{k}": np.asarray(v) for k, v in data.items()}
    flat.update({f"model/{k}": np.asarray(v) for k, v in model_arrays.items()})
    flat.update({f"dims/{k}": np.asarray(v) for k, v in dims.items()})
    flat["attrs/num_steps"] = np.asarray(num_steps)
    flat["attrs/dt"] = np.asarray(dt, np.float64)
    with open(path, "wb") as f:
        np.savez_compressed(f, **flat)


def load_log(path):
    """{"data": {...}, "model": {...}, "dims": {...}, "attrs": {...}} of a saved log."""
    out = {"data": {}, "model": {}, "dims": {}, "attrs": {}}
    with np.load(path, allow_pickle=False) as z:
        for k in z.files:
            g, name = k.split("/", 1)
            v = z[k]
            out.setdefault(g, {})[name] = v.item() if v.ndim == 0 else v
    return out


class SimulationLog:
    """SimulationHDF5Logger (logging/sim_logger.py) for an OstrichEngine stepping
    state_in -> state_out with `control`: `capture_step()` after each step (eager or in the
    step's graph) logs, at the log's device step counter,

    * dynamics state: body_pose, body_vel (state_out), body_pose_prev, body_vel_prev,
      ext_force (state_in), joint_target_pos, joint_target_vel;
    * constraint data: _constr_force (lambda [W, N_c], reference row layout);
    * solver history: iter_count, res_norm_sq, candidates_best_idx, flags, pcr_iter_count
      [W], and per Newton iteration candidates_res_norm_sq and pcr_history_iter_count
      [newton_max_iters, W] (NaN / -1 past a world's last iteration and past the 32nd: the
      fused kernel keeps the first 32 iterations of a world on chip, so with
      nr.max_iters > 32 rows 32 and later read NaN / -1 even for iterations that ran).

    The engine's iteration history is attached for the log's lifetime."""

    def __init__(self, num_steps, engine, state_in, state_out, control, dt, log_dynamics_state=True,
                 log_constraint_data=True, log_linear_system_data=True):
        self.engine, self.dt, self.num_steps = engine, dt, num_steps
        W, dev = engine.num_worlds, engine.device
        self.stats = {k: torch.zeros(W, dtype=torch.int32, device=dev)
                      for k in ("nr_iters", "pcr_iters", "best_iter", "flags")}
        self.stats["merit"] = torch.zeros(W, dtype=torch.float32, device=dev)
        self.lam = torch.zeros(W, engine.num_constraints, dtype=torch.float32, device=dev)
        cap = engine.config.nr.max_iters
        self.hist_merit = torch.zeros(cap, W, dtype=torch.float32, device=dev)
        self.hist_pcr = torch.zeros(cap, W, dtype=torch.int32, device=dev)
        engine.set_iteration_history(self.hist_merit, self.hist_pcr)
        src = {}
        if log_dynamics_state:
            src.update(ext_force=state_in.body_f, body_pose=state_out.body_q, body_vel=state_out.body_qd,
                       body_pose_prev=state_in.body_q, body_vel_prev=state_in.body_qd)
            if control.joint_target_pos.numel():
                src.update(joint_target_pos=control.joint_target_pos, joint_target_vel=control.joint_target_vel)
        if log_constraint_data:
            src["_constr_force"] = self.lam
        if log_linear_system_data:
            src.update(iter_count=self.stats["nr_iters"], res_norm_sq=self.stats["merit"],
                       candidates_best_idx=self.stats["best_iter"], flags=self.stats["flags"],
                       pcr_iter_count=self.stats["pcr_iters"], candidates_res_norm_sq=self.hist_merit,
                       pcr_history_iter_count=self.hist_pcr)
        self.log = DeviceLog(num_steps, src)

    @property
    def step_index(self):
        return self.log.step_index

    def capture_step(self):
        """Logs the engine's last step (device copies only; records into a graph)."""
        self.engine.stats(out=self.stats)
        self.engine.constraint_impulses(out=self.lam)
        self.log.capture()

    def save(self, path):
        """Writes the log (the reference's save_to_hdf5 groups; see the module docstring)."""
        save_log(path, self.num_steps, self.dt, self.log.numpy(), _model_arrays(self.engine.model), _dims(self.engine))

    def close(self):
        self.engine.set_iteration_history(None)


class DatasetLog:
    """DatasetHDF5Logger (logging/dataset_logger.py): per step the state (ext_force,
    body_pose, body_vel, body_pose_prev, body_vel_prev), joint targets, the constraint
    impulses and the step's input contacts (contact_count, contact_point0/1,
    contact_normal, contact_shape0/1, contact_thickness0/1 from the Newton contact arrays;
    the margins stand for the thicknesses, core/contacts.py:85-86)."""

    def __init__(self, num_steps, engine, state_in, state_out, control, contacts, dt):
        self.engine, self.dt, self.num_steps = engine, dt, num_steps
        self.lam = torch.zeros(engine.num_worlds, engine.num_constraints, dtype=torch.float32, device=engine.device)
        src = dict(ext_force=state_in.body_f, body_pose=state_out.body_q, body_vel=state_out.body_qd,
                   body_pose_prev=state_in.body_q, body_vel_prev=state_in.body_qd)
        if control.joint_target_pos.numel():
            src.update(joint_target_pos=control.joint_target_pos, joint_target_vel=control.joint_target_vel)
        src["_constr_force"] = self.lam
        src.update(contact_count=contacts.rigid_contact_count, contact_point0=contacts.rigid_contact_point0,
                   contact_point1=contacts.rigid_contact_point1, contact_normal=contacts.rigid_contact_normal,
                   contact_shape0=contacts.rigid_contact_shape0, contact_shape1=contacts.rigid_contact_shape1,
                   contact_thickness0=contacts.rigid_contact_margin0,
                   contact_thickness1=contacts.rigid_contact_margin1)
        self.log = DeviceLog(num_steps, src)

    @property
    def step_index(self):
        return self.log.step_index

    def capture_step(self):
        self.engine.constraint_impulses(out=self.lam)
        self.log.capture()

    def save(self, path):
        save_log(path, self.num_steps, self.dt, self.log.numpy(), _model_arrays(self.engine.model), _dims(self.engine))


class AdjointLog:
    """AdjointHDF5Logger (logging/adjoint_logger.py:15-132): per differentiated step t, in
    slot t, the adjoint multipliers (_w = w_lambda [W, N_c], w_u [W*B, 6]), the incoming
    gradients (body_pose_grad, body_vel_grad), the gradients of the step's inputs
    (body_pose_prev_grad, body_vel_prev_grad, ext_force_grad, joint_target_pos_grad,
    joint_target_vel_grad, shape_mu_grad) and the adjoint PCR iterations per world
    (pcr_iter_count).  `capture_step(t, grads, grad_body_q, grad_body_qd)` takes the dict
    `step_backward(..., adjoint_multipliers=True)` returned; `CheckpointedRollout.backward`
    calls it with `adjoint_log=`."""

    def __init__(self, num_steps, engine, dt):
        self.engine, self.dt, self.num_steps = engine, dt, num_steps
        W, dev = engine.num_worlds, engine.device
        m = engine.model
        B, D, SC = m.bodies_per_world, m.dofs_per_world, m.shapes_per_world + m.num_global_shapes
        z = lambda *s, dt=torch.float32: torch.zeros(*s, dtype=dt, device=dev)
        self.stage = dict(_w=z(W, engine.num_constraints), w_u=z(W * B, 6), body_pose_grad=z(W * B, 7),
                          body_vel_grad=z(W * B, 6), body_pose_prev_grad=z(W * B, 7),
                          body_vel_prev_grad=z(W * B, 6), ext_force_grad=z(W * B, 6), shape_mu_grad=z(W, SC),
                          pcr_iter_count=z(W, dt=torch.int32))
        if D:
            self.stage.update(joint_target_pos_grad=z(W, D), joint_target_vel_grad=z(W, D))
        self._stats = {k: z(W, dt=torch.int32) for k in ("nr_iters", "pcr_iters", "best_iter", "flags")}
        self._stats["merit"] = z(W)
        self.log = DeviceLog(num_steps, self.stage)

    _FROM = dict(_w="w_lambda", w_u="w_u", body_pose_prev_grad="body_q", body_vel_prev_grad="body_qd",
                 ext_force_grad="body_f", shape_mu_grad="shape_mu", joint_target_pos_grad="joint_target_pos",
                 joint_target_vel_grad="joint_target_vel")

    def capture_step(self, t, grads, grad_body_q, grad_body_qd):
        """Logs the backward of step t (device copies; slot t, nothing past num_steps)."""
        for k, src in self._FROM.items():
            if k in self.stage:
                if src not in grads:
                    raise KeyError(f"{src} missing: call step_backward with adjoint_multipliers=True")
                self.stage[k].copy_(grads[src].reshape(self.stage[k].shape))
        self.stage["body_pose_grad"].copy_(grad_body_q.reshape(self.stage["body_pose_grad"].shape))
        self.stage["body_vel_grad"].copy_(grad_body_qd.reshape(self.stage["body_vel_grad"].shape))
        self.engine.stats(out=self._stats)
        self.stage["pcr_iter_count"].copy_(self._stats["pcr_iters"])
        self.log.step_index.fill_(int(t))
        self.log.capture(advance=False)

    def save(self, path):
        save_log(path, self.num_steps, self.dt, self.log.numpy(), _model_arrays(self.engine.model), _dims(self.engine))
