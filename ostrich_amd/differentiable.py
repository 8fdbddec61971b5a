"""Differentiable stepping: the torch bridge and checkpointed rollouts.

* `OstrichStep` - a `torch.autograd.Function` around one solver step whose backward is
  the gfx950 adjoint (`OstrichEngine.step_backward`, core/base_engine.py:528-743).  It
  plays the part of the reference's torch bridges (learning/differentiable_newton_step.py:
  33-250, learning/differentiable_linear_solve.py:16-103) for the whole step.
* `TrajectoryBuffer` - per-step records the adjoint needs (input/output state, control,
  contacts, lambda, lambda_prev_iter), as simulation/trajectory_buffer.py:106-351 keeps
  them, held on the device.
* `CheckpointedRollout` - forward over N steps keeping only the states at every k-th step
  (and, when the warm starter seeds Newton, its snapshot there); the backward
  re-simulates each window of k steps from its checkpoint into a TrajectoryBuffer and
  runs the adjoint backwards through it (simulation/differentiable_simulator.py:234-271
  with the windowed recompute of experiments/4_scalability_box/ostrich_sim_ckpt.py).
  Memory is O(N/k + k) step records instead of O(N).  The solver is deterministic (fixed
  accumulation order) and a rollout starts from a reset warm starter, so it is a function
  of (state0, controls) alone and a re-simulated window reproduces the forward bit for
  bit.
* `OstrichRollout` - a `torch.autograd.Function` over a whole checkpointed rollout.

Gradient flow between steps follows the reference: dL/dq- of step t is added to the
pose gradient of state t, and dL/du- of step t becomes the velocity gradient of state t
(trajectory_buffer.py:266-297, accumulated here with the loss's own gradient at t).
With `gradient_normalization` the carried (pose, velocity) gradient is rescaled to unit
norm per step, as TrajectoryBuffer.normalize_gradients (:299-351).
"""
from dataclasses import dataclass
from typing import Callable, List, Optional

import torch

from .containers import Contacts, Control, State


@dataclass
class StepRecord:
    state_in: State
    state_out: State
    control: Control
    contacts: Contacts
    lam: torch.Tensor
    lam_prev: torch.Tensor


class TrajectoryBuffer:
    """Step records (simulation/trajectory_buffer.py:106-351).

    Two ways in: `save_step(record)` appends a StepRecord (the checkpointed rollout), and
    the reference's `save_step(step_idx, data)` records the step an engine just took from
    its EngineData (`engine.data`).  `load_step(step_idx, data)` puts record step_idx back
    into an engine's data with dL/dq+, dL/du+ of the step's output state (state
    step_idx + 1) so that `engine.step_backward()` differentiates it; `save_gradients`
    and `save_pose_gradients` carry the step's input gradients to state step_idx
    (trajectory_buffer.py:266-297)."""

    def __init__(self):
        self.steps: List[StepRecord] = []
        self.body_q_grad = {}   # state index -> dL/dq
        self.body_qd_grad = {}  # state index -> dL/du
        self.input_grads = {}   # step index -> gradients of the step's inputs

    def save_step(self, rec, data=None, contacts=None):
        if isinstance(rec, StepRecord):
            self.steps.append(rec)
            return
        step_idx = int(rec)
        if step_idx != len(self.steps):
            raise ValueError(f"steps are recorded in order: expected step {len(self.steps)}, got {step_idx}")
        lam = data.lam if data.lam is not None else data.engine_lam()
        lam_prev = data.lam_prev if data.lam_prev is not None else data.engine_lam_prev()
        self.steps.append(StepRecord(data.state_in.clone(), data.state_out.clone(), data.control,
                                     contacts if contacts is not None else data.contacts, lam.clone(),
                                     lam_prev.clone()))

    def set_output_grad(self, state_idx, grad_q, grad_qd):
        """dL/dq, dL/du of state state_idx from the loss (the tape gradients the reference's
        buffer holds in body_pose.grad / body_vel.grad)."""
        self.body_q_grad[state_idx] = grad_q
        self.body_qd_grad[state_idx] = grad_qd

    def load_step(self, step_idx, data, contacts=None):
        r = self.steps[step_idx]
        data.state_in, data.state_out, data.control, data.contacts = r.state_in, r.state_out, r.control, r.contacts
        data.lam, data.lam_prev = r.lam, r.lam_prev
        data.body_q_grad = self.body_q_grad.get(step_idx + 1)
        data.body_qd_grad = self.body_qd_grad.get(step_idx + 1)

    def save_gradients(self, step_idx, data):
        """dL/du of state step_idx and the step's input gradients from data.grads."""
        self.body_qd_grad[step_idx] = data.grads["body_qd"]
        self.input_grads[step_idx] = {k: v for k, v in data.grads.items() if k not in ("body_q", "body_qd")}

    def save_pose_gradients(self, step_idx, data):
        """Adds dL/dq- of the step to the pose gradient of state step_idx."""
        g = data.grads["body_q"]
        prev = self.body_q_grad.get(step_idx)
        self.body_q_grad[step_idx] = g if prev is None else prev + g

    def __len__(self):
        return len(self.steps)

    def __getitem__(self, i) -> StepRecord:
        return self.steps[i]

    def clear(self):
        self.steps.clear()
        self.body_q_grad.clear()
        self.body_qd_grad.clear()
        self.input_grads.clear()


def _record_step(engine, state_in: State, control: Control, contacts: Contacts, dt: float) -> StepRecord:
    out = state_in.clone()
    engine.step(state_in, out, control, contacts, dt)
    return StepRecord(state_in, out, control, contacts, engine.constraint_impulses(),
                      engine.constraint_impulses_prev())


class OstrichStep(torch.autograd.Function):
    """(body_q, body_qd, body_f, target_pos, target_vel, shape_mu) -> (body_q+, body_qd+).

    `contacts` are treated as constant (collision is not differentiated, as in the
    reference).  `shape_mu` (W*(S+G)) may be None to keep the model's friction."""

    @staticmethod
    def forward(ctx, engine, contacts, dt, body_q, body_qd, body_f, target_pos, target_vel, shape_mu=None):
        # friction coefficients of this step: the given ones, else the model's; the engine
        # holds one set, so the step's set is kept and put back for its backward
        mu = shape_mu.detach().clone() if shape_mu is not None else engine.model_shape_mu()
        engine.set_shape_mu(mu)
        s_in = State(body_q.detach().clone().contiguous(), body_qd.detach().clone().contiguous(),
                     body_f.detach().clone().contiguous())
        ctl = Control(target_pos.detach().clone().contiguous(), target_vel.detach().clone().contiguous())
        rec = _record_step(engine, s_in, ctl, contacts, dt)
        ctx.engine, ctx.dt, ctx.rec, ctx.mu = engine, dt, rec, mu
        ctx.has_mu = shape_mu is not None
        return rec.state_out.body_q.clone(), rec.state_out.body_qd.clone()

    @staticmethod
    def backward(ctx, g_q, g_qd):
        r = ctx.rec
        if g_q is None:
            g_q = torch.zeros_like(r.state_out.body_q)
        if g_qd is None:
            g_qd = torch.zeros_like(r.state_out.body_qd)
        ctx.engine.set_shape_mu(ctx.mu)  # the friction of this step, whatever later steps set
        g = ctx.engine.step_backward_explicit(r.state_in, r.state_out, r.control, r.contacts, ctx.dt, r.lam,
                                              r.lam_prev, g_q, g_qd)
        gmu = g["shape_mu"].reshape(-1) if ctx.has_mu else None
        return (None, None, None, g["body_q"], g["body_qd"], g["body_f"],
                g["joint_target_pos"].reshape(r.control.joint_target_pos.shape),
                g["joint_target_vel"].reshape(r.control.joint_target_vel.shape), gmu)


class CheckpointedRollout:
    """N steps of `engine` with contacts from `collide(body_q) -> Contacts`, checkpoints
    every `checkpoint_every` steps."""

    def __init__(self, engine, collide: Callable[[torch.Tensor], Contacts], dt: float, checkpoint_every: int = 10,
                 gradient_normalization: Optional[bool] = None):
        if checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        self.engine, self.collide, self.dt, self.k = engine, collide, dt, checkpoint_every
        gn = engine.config.adjoint.gradient_normalization if gradient_normalization is None else gradient_normalization
        self.gradient_normalization = gn
        self.checkpoints: List[State] = []
        self.ws_checkpoints: List[Optional[torch.Tensor]] = []
        self.controls: List[Control] = []
        self.final: Optional[State] = None
        self.recomputed_steps = 0

    def _step(self, state: State, t: int) -> StepRecord:
        contacts = self.collide(state.body_q)
        return _record_step(self.engine, state, self.controls[t], contacts, self.dt)

    def _seeds_from_warm_start(self):
        ws = self.engine.config.warm_start
        return ws.enabled and ws.seed_iterate

    def forward(self, state0: State, controls: List[Control]) -> State:
        """Runs len(controls) steps from state0; keeps the states at steps 0, k, 2k, ...
        and, when the warm starter seeds Newton, its snapshot there (no per-step records:
        the backward re-simulates each window from its checkpoint with the same kernels,
        which reproduces these steps bit for bit).  The rollout starts from a reset warm
        starter, so it depends on (state0, controls) only."""
        self.controls = list(controls)
        self.checkpoints = []
        self.ws_checkpoints = []
        seeds = self._seeds_from_warm_start()
        if seeds:
            self.engine.reset_warm_start()
        s, nxt = state0.clone(), state0.clone()
        for t in range(len(self.controls)):
            if t % self.k == 0:
                self.checkpoints.append(s.clone())
                self.ws_checkpoints.append(self.engine.warm_start_state() if seeds else None)
            self.engine.step(s, nxt, self.controls[t], self.collide(s.body_q), self.dt)
            s, nxt = nxt, s
        self.final = s
        return s

    def _window(self, j: int) -> TrajectoryBuffer:
        buf = TrajectoryBuffer()
        t0 = j * self.k
        t1 = min(t0 + self.k, len(self.controls))
        s = self.checkpoints[j].clone()
        if self.ws_checkpoints and self.ws_checkpoints[j] is not None:
            self.engine.set_warm_start_state(self.ws_checkpoints[j])
        for t in range(t0, t1):
            rec = self._step(s, t)
            buf.save_step(rec)
            s = rec.state_out
        self.recomputed_steps += t1 - t0
        return buf

    def backward(self, grad_q: torch.Tensor, grad_qd: torch.Tensor,
                 step_loss_grads: Optional[Callable[[int], Optional[tuple]]] = None, adjoint_log=None):
        """Adjoint of the whole rollout from dL/dq_N, dL/du_N.

        `step_loss_grads(t)` may return (dL/dq_t, dL/du_t) of a loss on intermediate state
        t (1 <= t < N), added to the carried gradient.  Returns dict: body_q, body_qd (of
        the initial state), body_f [N, ...], joint_target_pos / joint_target_vel [N, W, D],
        joint_target_ke, joint_target_kd [W, D] and shape_mu [W, S+G] summed over steps.
        `adjoint_log` (sim_log.AdjointLog) logs every step's adjoint in slot t."""
        N = len(self.controls)
        gq = grad_q.reshape(self.final.body_q.shape).clone()
        gu = grad_qd.reshape(self.final.body_qd.shape).clone()
        out = dict(body_f=[None] * N, joint_target_pos=[None] * N, joint_target_vel=[None] * N)
        acc = {}
        for j in reversed(range(len(self.checkpoints))):
            buf = self._window(j)
            t0 = j * self.k
            for i in reversed(range(len(buf))):
                t = t0 + i
                r = buf[i]
                g = self.engine.step_backward_explicit(r.state_in, r.state_out, r.control, r.contacts, self.dt,
                                                       r.lam, r.lam_prev, gq, gu,
                                                       adjoint_multipliers=adjoint_log is not None)
                if adjoint_log is not None:
                    adjoint_log.capture_step(t, g, gq, gu)
                out["body_f"][t] = g["body_f"]
                out["joint_target_pos"][t] = g["joint_target_pos"]
                out["joint_target_vel"][t] = g["joint_target_vel"]
                for k in ("joint_target_ke", "joint_target_kd", "shape_mu"):
                    acc[k] = g[k] if k not in acc else acc[k] + g[k]
                gq, gu = g["body_q"], g["body_qd"]
                if step_loss_grads is not None and t > 0:
                    extra = step_loss_grads(t)
                    if extra is not None:
                        gq = gq + extra[0].reshape(gq.shape)
                        gu = gu + extra[1].reshape(gu.shape)
                if self.gradient_normalization and t > 0:
                    nrm = torch.sqrt((gq * gq).sum() + (gu * gu).sum())
                    scale = torch.where(nrm > 0, 1.0 / nrm, torch.zeros_like(nrm))
                    gq, gu = gq * scale, gu * scale
        out = {k: torch.stack(v) for k, v in out.items()}
        out.update(acc)
        out["body_q"], out["body_qd"] = gq, gu
        return out


class OstrichRollout(torch.autograd.Function):
    """(body_q0, body_qd0, target_pos [N, W*D], target_vel [N, W*D], shape_mu [W*(S+G)] or
    None) -> (body_q_N, body_qd_N) through a CheckpointedRollout (body_f is held at zero).
    `shape_mu` sets the friction of every step (None keeps the model's) and receives
    dL/dmu summed over the steps (the rollout of simulation/differentiable_simulator.py:
    234-271 with the friction coefficients as a parameter)."""

    @staticmethod
    def forward(ctx, engine, collide, dt, checkpoint_every, body_q0, body_qd0, target_pos, target_vel,
                shape_mu=None):
        N = target_vel.shape[0]
        s0 = State(body_q0.detach().contiguous(), body_qd0.detach().contiguous(), torch.zeros_like(body_qd0))
        controls = [Control(target_pos[t].detach().contiguous(), target_vel[t].detach().contiguous())
                    for t in range(N)]
        mu = shape_mu.detach().clone() if shape_mu is not None else engine.model_shape_mu()
        engine.set_shape_mu(mu)
        ro = CheckpointedRollout(engine, collide, dt, checkpoint_every)
        fin = ro.forward(s0, controls)
        ctx.ro, ctx.mu, ctx.has_mu = ro, mu, shape_mu is not None
        return fin.body_q.clone(), fin.body_qd.clone()

    @staticmethod
    def backward(ctx, g_q, g_qd):
        ro = ctx.ro
        if g_q is None:
            g_q = torch.zeros_like(ro.final.body_q)
        if g_qd is None:
            g_qd = torch.zeros_like(ro.final.body_qd)
        ro.engine.set_shape_mu(ctx.mu)  # the friction of this rollout, whatever ran since
        g = ro.backward(g_q, g_qd)
        N = len(ro.controls)
        gmu = g["shape_mu"].reshape(-1) if ctx.has_mu else None
        return (None, None, None, None, g["body_q"], g["body_qd"], g["joint_target_pos"].reshape(N, -1),
                g["joint_target_vel"].reshape(N, -1), gmu)


class TrajectoryAdam:
    """Adam over a control trajectory `params` [steps, worlds, dofs] (a contiguous float32
    device tensor), restating optim/trajectory_adam.py:56-137: the dof slice
    [dof_offset, dof_offset + num_dofs) is updated with per-element gradient clamping,
    biased moments and bias correction with the 0-based step counter.  The update runs in
    the gfx950 library (`ofb_trajectory_adam_step`)."""

    def __init__(self, params: torch.Tensor, lr: float = 0.05, betas=(0.9, 0.999), eps: float = 1e-7,
                 dof_offset: int = 0, num_dofs: Optional[int] = None, clip_grad: float = 100.0):
        if params.dim() != 3 or params.dtype != torch.float32 or not params.is_contiguous():
            raise ValueError("params must be a contiguous float32 tensor [steps, worlds, dofs]")
        self.params = params
        self.lr, (self.beta1, self.beta2), self.eps = lr, betas, eps
        self.dof_offset = dof_offset
        self.num_dofs = params.shape[2] - dof_offset if num_dofs is None else num_dofs
        self.clip_grad = clip_grad
        self.t = 0
        self.m = torch.zeros_like(params)
        self.v = torch.zeros_like(params)

    def step(self, grad: torch.Tensor):
        import ctypes as C

        from . import _lib
        if grad.shape != self.params.shape or grad.dtype != torch.float32:
            raise ValueError("grad must match params")
        grad = grad.contiguous()
        lib = _lib.load()
        S, W, D = self.params.shape
        stream = C.c_void_p(torch.cuda.current_stream(self.params.device).cuda_stream)
        _lib.check(lib.ofb_trajectory_adam_step(
            C.c_void_p(self.params.data_ptr()), C.c_void_p(grad.data_ptr()), C.c_void_p(self.m.data_ptr()),
            C.c_void_p(self.v.data_ptr()), S, W, D, self.dof_offset, self.num_dofs, C.c_float(self.lr),
            C.c_float(self.beta1), C.c_float(self.beta2), C.c_float(self.eps), C.c_float(self.clip_grad), self.t,
            stream), "ofb_trajectory_adam_step")
        self.t += 1

    def reset_internal_state(self):
        self.m.zero_()
        self.v.zero_()
        self.t = 0
