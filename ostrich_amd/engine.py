"""OstrichEngine: the solver plugin, with the reference's entry points.

Mirrors `OstrichEngine` (src/ostrich/core/engine.py:14-71) and the parts of
`OstrichEngineBase` (core/base_engine.py:111-751) a caller uses:
`step(state_in, state_out, control, contacts, dt)`, `reset_timestep_counter()` and the
per-step diagnostics.  All arithmetic runs in the gfx950 library behind the C ABI
(include/ostrich_fb.h); this class only moves pointers.
"""
import ctypes as C
from typing import Optional

import torch

from . import _abi, _lib
from .config import OstrichEngineConfig
from .containers import Contacts, Control, State


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class EngineData:
    """The step data `step_backward()` reads and the gradients it writes, the part of the
    reference's EngineData (core/engine_data.py:42-300) that its adjoint uses: the step's
    input state (q-, u-, f), output state (q+, u+), control, contacts, dt, lambda and
    lambda_prev_iter, the incoming dL/dq+ and dL/du+, and after the call the gradients
    of the step's inputs (`grads`: body_q, body_qd, body_f, joint_target_pos,
    joint_target_vel, joint_target_ke, joint_target_kd, shape_mu).

    `OstrichEngine.step` fills the step fields with the tensors it was called with (lambda
    and lambda_prev_iter are read from the solver when `step_backward()` runs, unless set);
    `TrajectoryBuffer.load_step` (ostrich_amd.differentiable) overwrites them with a
    recorded step, as simulation/trajectory_buffer.py:235-264 does."""

    def __init__(self, engine=None):
        self._engine = engine
        self.state_in = self.state_out = self.control = self.contacts = None
        self.dt = None
        self.lam = self.lam_prev = None
        self.body_q_grad = self.body_qd_grad = None
        self.grads = {}

    def engine_lam(self):
        """lambda of the engine's last step (its solver buffer)."""
        return self._engine.constraint_impulses()

    def engine_lam_prev(self):
        return self._engine.constraint_impulses_prev()

    def zero_gradients(self):
        """EngineData.zero_gradients: clears the outputs of the last step_backward()."""
        self.grads = {}


class OstrichEngine:
    def __init__(self, model, sim_steps=None, config: OstrichEngineConfig = OstrichEngineConfig(),
                 logging_config=None, differentiable_simulation=False, device="cuda:0",
                 rigid_contact_max=None):
        if not torch.cuda.is_available():
            raise RuntimeError("OstrichEngine needs a HIP device (torch.cuda.is_available() is False)")
        self.lib = _lib.load()
        self.model = model
        self.config = config
        self.sim_steps = sim_steps
        self.differentiable_simulation = differentiable_simulation
        self.device = torch.device(device)
        self.timestep = 0
        W = model.num_worlds
        # model arrays live on the device for the duration of ofb_create (the library copies them)
        dev = {}
        for name, dtype, _, _ in _abi.MODEL_ARRAYS:
            arr = model.arrays[name]
            td = torch.int32 if dtype == "i4" else torch.float32
            dev[name] = torch.as_tensor(arr, dtype=td, device=self.device).contiguous()
        ptrs = {k: v.data_ptr() for k, v in dev.items()}
        abi_model = model.to_abi(ptrs)
        cfg = config.to_abi()
        cap = rigid_contact_max if rigid_contact_max is not None else W * config.contacts.max_per_world
        cfg.rigid_contact_max = int(cap)
        self.rigid_contact_max = int(cap)
        h = C.c_void_p()
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.ofb_create(C.byref(abi_model), C.byref(cfg), config.path_id, C.byref(h)), "ofb_create")
        self._h = h
        # lambda_prev_iter, the adjoint's input, is written out only for a differentiable
        # engine (the reference allocates its adjoint arrays only then, base_engine.py:146)
        _lib.check(self.lib.ofb_set_keep_lambda_prev(h, 1 if differentiable_simulation else 0),
                   "ofb_set_keep_lambda_prev")
        del dev
        nc, oc, on, of = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(self.lib.ofb_dims(h, C.byref(nc), C.byref(oc), C.byref(on), C.byref(of)), "ofb_dims")
        self.num_constraints = nc.value
        self.offset_ctrl, self.offset_n, self.offset_f = oc.value, on.value, of.value
        self.num_worlds = W
        # keep the argument structs alive between calls
        self._in = _abi.OfbStateIn()
        self._out = _abi.OfbStateOut()
        self._ctl = _abi.OfbControl()
        self._con = _abi.OfbContacts()
        self.data = EngineData(self)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.device)
                self.lib.ofb_destroy(h)
            except Exception:
                pass
            self._h = None

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def step(self, state_in: State, state_out: State, control: Control, contacts: Contacts, dt: float):
        """core/engine.py:25-71."""
        for t in (state_in.body_q, state_in.body_qd, state_in.body_f, state_out.body_q, state_out.body_qd):
            if t.dtype != torch.float32 or not t.is_contiguous() or t.device != self.device:
                raise ValueError("state tensors must be contiguous float32 on the solver device")
        if contacts.rigid_contact_max > self.rigid_contact_max:
            raise ValueError(f"contacts hold {contacts.rigid_contact_max} slots, more than "
                             f"rigid_contact_max={self.rigid_contact_max} given at construction")
        self._in.body_q, self._in.body_qd, self._in.body_f = (
            state_in.body_q.data_ptr(), state_in.body_qd.data_ptr(), state_in.body_f.data_ptr())
        self._out.body_q, self._out.body_qd = state_out.body_q.data_ptr(), state_out.body_qd.data_ptr()
        self._ctl.joint_target_pos = control.joint_target_pos.data_ptr()
        self._ctl.joint_target_vel = control.joint_target_vel.data_ptr()
        c = self._fill_contacts(contacts)
        _lib.check(self.lib.ofb_step(self._h, C.byref(self._in), C.byref(self._out), C.byref(self._ctl),
                                     C.byref(c), C.c_float(dt), self._stream()), "ofb_step")
        self.timestep += 1
        d = self.data
        d.state_in, d.state_out, d.control, d.contacts, d.dt = state_in, state_out, control, contacts, dt
        d.lam = d.lam_prev = None  # the solver's own buffers hold this step's

    def reset_timestep_counter(self):
        """core/base_engine.py:278-279."""
        self.timestep = 0

    def stats(self, out=None):
        """Per-world diagnostics of the last step (tensors on the device).  `out` may be a
        dict of preallocated tensors (nr_iters, pcr_iters, best_iter, flags int32 [W];
        merit float32 [W]) to fill in place, e.g. inside a captured graph."""
        W = self.num_worlds
        if out is None:
            out = {k: torch.zeros(W, dtype=torch.int32, device=self.device)
                   for k in ("nr_iters", "pcr_iters", "best_iter", "flags")}
            out["merit"] = torch.zeros(W, dtype=torch.float32, device=self.device)
        _lib.check(self.lib.ofb_get_stats(self._h, _ptr(out["nr_iters"]), _ptr(out["pcr_iters"]),
                                          _ptr(out["merit"]), _ptr(out["best_iter"]), _ptr(out["flags"]),
                                          self._stream()), "ofb_get_stats")
        return out

    def constraint_impulses(self, out=None):
        """lambda [W, N_c] of the picked iterate (reference row layout), into `out` if given."""
        lam = out if out is not None else torch.zeros(self.num_worlds, self.num_constraints, dtype=torch.float32,
                                                      device=self.device)
        _lib.check(self.lib.ofb_get_lambda(self._h, _ptr(lam), self._stream()), "ofb_get_lambda")
        return lam

    def set_iteration_history(self, merit=None, pcr_iters=None):
        """Per-Newton-iteration history of every following step into merit (float32) and
        pcr_iters (int32) device tensors [capacity, W]: ||h||^2 after each iteration and
        its PCR iterations, NaN / -1 past a world's last iteration.  None detaches."""
        if merit is None:
            _lib.check(self.lib.ofb_set_iteration_history(self._h, None, None, 0), "ofb_set_iteration_history")
            self._hist_keepalive = None
            return
        if (merit.dtype != torch.float32 or pcr_iters.dtype != torch.int32 or merit.shape != pcr_iters.shape
                or merit.dim() != 2 or merit.shape[1] != self.num_worlds or not merit.is_contiguous()
                or not pcr_iters.is_contiguous()):
            raise ValueError("history buffers must be contiguous [capacity, W] float32 / int32 tensors")
        _lib.check(self.lib.ofb_set_iteration_history(self._h, _ptr(merit), _ptr(pcr_iters), merit.shape[0]),
                   "ofb_set_iteration_history")
        self._hist_keepalive = (merit, pcr_iters)

    def set_warm_start_net(self, net):
        """Attaches a learned warm start (`ostrich_amd.learned.WarmStartNet`) or detaches it
        (None).  Acts with warm_start.enabled and warm_start.seed_iterate."""
        _lib.check(self.lib.ofb_set_warm_start_net(self._h, net._h if net is not None else None),
                   "ofb_set_warm_start_net")
        self._net_keepalive = net

    def contact_features(self):
        """Learned-warm-start features [W, C_max, 12] of the last step and the impulse scale
        m_eff |g| [W, C_max]."""
        W, K = self.num_worlds, self.config.contacts.max_per_world
        f = torch.zeros(W, K, 12, device=self.device)
        sc = torch.zeros(W, K, device=self.device)
        _lib.check(self.lib.ofb_get_contact_features(self._h, _ptr(f), _ptr(sc), self._stream()),
                   "ofb_get_contact_features")
        return f, sc

    def linesearch_steps(self):
        """The merit line search's candidate step sizes (empty when it is off)."""
        import numpy as np
        n = C.c_int32()
        buf = (C.c_float * 512)()
        _lib.check(self.lib.ofb_get_linesearch_steps(self._h, buf, C.byref(n)), "ofb_get_linesearch_steps")
        return np.array(buf[:n.value], np.float32)

    def reset_warm_start(self):
        """Forgets the warm starter's snapshot (next step cold-starts every contact)."""
        _lib.check(self.lib.ofb_warm_start_reset(self._h, self._stream()), "ofb_warm_start_reset")

    def warm_start_state(self):
        """The warm starter's snapshot as an opaque uint8 device tensor (empty when the warm
        start is off); `set_warm_start_state` puts it back."""
        n = C.c_int64()
        _lib.check(self.lib.ofb_warm_start_state_size(self._h, C.byref(n)), "ofb_warm_start_state_size")
        buf = torch.empty(n.value, dtype=torch.uint8, device=self.device)
        if n.value:
            _lib.check(self.lib.ofb_warm_start_state_get(self._h, _ptr(buf), self._stream()),
                       "ofb_warm_start_state_get")
        return buf

    def set_warm_start_state(self, buf):
        n = C.c_int64()
        _lib.check(self.lib.ofb_warm_start_state_size(self._h, C.byref(n)), "ofb_warm_start_state_size")
        if buf.numel() != n.value or buf.dtype != torch.uint8 or buf.device != self.device:
            raise ValueError(f"warm-start state must be {n.value} uint8 bytes on {self.device}")
        if n.value:
            _lib.check(self.lib.ofb_warm_start_state_set(self._h, _ptr(buf.contiguous()), self._stream()),
                       "ofb_warm_start_state_set")

    def lambda_seed(self):
        """The Newton seed [W, N_c] the last step started from."""
        out = torch.zeros(self.num_worlds, self.num_constraints, dtype=torch.float32, device=self.device)
        _lib.check(self.lib.ofb_get_lambda_seed(self._h, _ptr(out), self._stream()), "ofb_get_lambda_seed")
        return out

    def set_lambda_seed(self, seed):
        seed = seed.to(self.device, torch.float32).contiguous()
        _lib.check(self.lib.ofb_set_lambda_seed(self._h, _ptr(seed), self._stream()), "ofb_set_lambda_seed")
        self._seed_keepalive = seed

    def packed_contacts(self):
        W, K = self.num_worlds, self.config.contacts.max_per_world
        d = self.device
        out = dict(count=torch.zeros(W, dtype=torch.int32, device=d),
                   shape0=torch.zeros(W * K, dtype=torch.int32, device=d),
                   shape1=torch.zeros(W * K, dtype=torch.int32, device=d),
                   point0=torch.zeros(W * K * 3, dtype=torch.float32, device=d),
                   point1=torch.zeros(W * K * 3, dtype=torch.float32, device=d),
                   normal=torch.zeros(W * K * 3, dtype=torch.float32, device=d),
                   thickness0=torch.zeros(W * K, dtype=torch.float32, device=d),
                   thickness1=torch.zeros(W * K, dtype=torch.float32, device=d))
        _lib.check(self.lib.ofb_get_packed_contacts(
            self._h, *(_ptr(out[k]) for k in ("count", "shape0", "shape1", "point0", "point1", "normal",
                                               "thickness0", "thickness1")), self._stream()),
            "ofb_get_packed_contacts")
        return out

    def last_solve_ms(self):
        """Duration of the last step's solve (HIP events on the step's stream)."""
        ms = C.c_float()
        _lib.check(self.lib.ofb_last_solve_ms(self._h, C.byref(ms)), "ofb_last_solve_ms")
        return ms.value

    def last_kernel_ms(self):
        """Duration of the last step's paired-class fused kernel (the launch holding the
        worlds of at most 32 rows), from HIP events on the step's stream around it."""
        ms = C.c_float()
        _lib.check(self.lib.ofb_last_kernel_ms(self._h, C.byref(ms)), "ofb_last_kernel_ms")
        return ms.value

    def set_class_timing(self, enabled=True):
        """Record HIP events around every row-class launch of each eager fused step."""
        _lib.check(self.lib.ofb_set_class_timing(self._h, 1 if enabled else 0), "ofb_set_class_timing")

    def class_times(self):
        """Per row class of the last step: dict(rows_cap, worlds, ms) lists (ms -1: the
        class was not launched or not timed).  Blocks until the step has finished."""
        import numpy as np
        rows, worlds, ms = (np.zeros(16, np.int32), np.zeros(16, np.int32), np.zeros(16, np.float32))
        n = C.c_int32()
        _lib.check(self.lib.ofb_get_class_times(self._h, 16, rows.ctypes.data, worlds.ctypes.data, ms.ctypes.data,
                                                C.byref(n)), "ofb_get_class_times")
        k = n.value
        return dict(rows_cap=rows[:k].tolist(), worlds=worlds[:k].tolist(), ms=[float(x) for x in ms[:k]])

    def class_shapes(self):
        """Per row class: dict(slots, lanes, max_rows) lists, the row slots per lane and the
        lanes per world of the kernel instance the class runs (32: two worlds per
        wavefront, 64: one, 128 / 256 / 512: one world over 2 / 4 / 8 wavefronts), and the
        largest row count of a world in the class (0: not launched).  A world of n rows is
        in the first class whose max_rows is >= n."""
        import numpy as np
        slots, lanes, max_rows = (np.zeros(16, np.int32) for _ in range(3))
        n = C.c_int32()
        _lib.check(self.lib.ofb_get_class_shapes(self._h, 16, slots.ctypes.data, lanes.ctypes.data,
                                                 max_rows.ctypes.data, C.byref(n)), "ofb_get_class_shapes")
        k = n.value
        return dict(slots=slots[:k].tolist(), lanes=lanes[:k].tolist(), max_rows=max_rows[:k].tolist())

    def fused_launch(self):
        """(waves per workgroup, pair_small_worlds) of the fused path."""
        w, p = C.c_int32(), C.c_int32()
        _lib.check(self.lib.ofb_get_fused_launch(self._h, C.byref(w), C.byref(p)), "ofb_get_fused_launch")
        return w.value, bool(p.value)

    def set_fused_launch(self, waves_per_block, pair_small_worlds):
        _lib.check(self.lib.ofb_set_fused_launch(self._h, int(waves_per_block), 1 if pair_small_worlds else 0),
                   "ofb_set_fused_launch")

    def fused_layout(self) -> int:
        """Row slots per lane of the fused path's worlds of more than 256 rows (2 or 4)."""
        v = C.c_int32()
        _lib.check(self.lib.ofb_get_fused_layout(self._h, C.byref(v)), "ofb_get_fused_layout")
        return v.value

    def set_fused_layout(self, large_world_slots: int):
        _lib.check(self.lib.ofb_set_fused_layout(self._h, int(large_world_slots)), "ofb_set_fused_layout")

    def autotune(self, state_in: State, control: Control, contacts: Contacts, dt: float, candidates=None,
                 repeats=3, pairing=True, layouts=None):
        """Chooses the fused launch settings for this batch from measurements.

        Each candidate (waves per workgroup in {1, 2, 4}; with `pairing` small worlds
        two per wavefront or one; with `layouts` the row slots per lane of the worlds of
        more than 256 rows, 2 or 4) runs one warm-up step and `repeats` timed steps of the
        given inputs on a scratch solver built from the same model and config, timed with
        HIP events on the current stream; the fastest is applied to this solver.  By
        default the layout is not timed but follows the batch: the wide one (4) on the
        Jacobi path when the batch holds at least 256 worlds of more than 256 rows (about
        one per CU), else the default (2).  With fewer such worlds they are all resident at
        once in either layout, and the default one solves each faster; with many, the wide
        one keeps more of them resident (on 1024 tracked robots 2.2 against 2.9 ms per
        step), while a step timed on a batch's first inputs, where the robots have not yet
        settled into their contacts, can rank the two the other way.  Every step starts
        from the same warm-start snapshot (this solver's), so the candidates solve the same
        systems from the same seeds.  This solver's state (warm-start snapshot, last
        results) is left as it was.
        Returns (best (waves_per_block, pair_small_worlds, large_world_slots),
        {candidate: best ms})."""
        if self.config.execution.path != "fused":
            raise ValueError("autotune applies to the fused path")
        scratch = OstrichEngine(self.model, config=self.config, device=self.device,
                                rigid_contact_max=self.rigid_contact_max)
        out = state_in.clone()
        stream = torch.cuda.current_stream(self.device)
        snapshot = self.warm_start_state()
        if layouts is None:
            layouts = (self.fused_layout(),)
            if self.config.linear.preconditioner_type == "jacobi":
                scratch.set_warm_start_state(snapshot)
                scratch.step(state_in, out, control, contacts, dt)  # packs this batch's contacts
                K = self.config.contacts.max_per_world
                rows = (self.num_constraints - 3 * K) + 3 * scratch.packed_contacts()["count"]
                layouts = (4,) if int((rows > 256).sum().item()) >= 256 else (2,)
        if candidates is None:
            pairs = (True, False) if pairing else (self.fused_launch()[1],)
            candidates = [(w, p, l) for l in layouts for p in pairs for w in (1, 2, 4)]
        times = {}
        for wpb, pair, slots in candidates:
            scratch.set_fused_launch(wpb, pair)
            scratch.set_fused_layout(slots)
            if snapshot is not None:
                scratch.set_warm_start_state(snapshot)
            scratch.step(state_in, out, control, contacts, dt)
            best = float("inf")
            for _ in range(repeats):
                if snapshot is not None:
                    scratch.set_warm_start_state(snapshot)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                scratch.step(state_in, out, control, contacts, dt)
                e1.record(stream)
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
            times[(wpb, bool(pair), int(slots))] = best
        torch.cuda.synchronize(self.device)
        del scratch
        choice = min(times, key=times.get)
        self.set_fused_launch(choice[0], choice[1])
        self.set_fused_layout(choice[2])
        return choice, times

    def capture(self, state_in: State, state_out: State, control: Control, contacts: Optional[Contacts], dt: float,
                collider=None, advance: bool = False):
        """Records one full step (contact packing, warm start, the fused launches, the
        warm-start snapshot) over these exact tensors into a HIP graph and returns it
        (a torch.cuda.CUDAGraph): `graph.replay()` runs a step that reads the tensors'
        current contents, so a caller refreshes them in place (State.copy_,
        Contacts.copy_) between replays.  The step allocates nothing and launches only
        kernels (no memset or host synchronisation), so it records as is; no timing events
        are recorded inside the graph.  Fused path only: the plain path reads its stop flag
        on the host every Newton iteration.

        With `collider` (a HipCollider) the graph first generates the contacts of
        state_in.body_q into the collider's own buffers and steps with those (`contacts`
        is then ignored); with `advance` it ends by copying state_out's poses and velocities into state_in, so
        consecutive replays run consecutive simulation steps with no host work between."""
        if self.config.execution.path != "fused":
            raise ValueError("graph capture applies to the fused path")
        if collider is None and contacts is None:
            raise ValueError("capture needs contacts or a collider")
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            if collider is not None:
                contacts = collider.collide(state_in.body_q)
            self.step(state_in, state_out, control, contacts, dt)
            if advance:
                state_in.body_q.copy_(state_out.body_q)
                state_in.body_qd.copy_(state_out.body_qd)
        torch.cuda.current_stream(self.device).wait_stream(s)
        self.timestep -= 1  # recording is not a step
        return g

    def constraint_impulses_prev(self):
        """lambda_prev_iter [W, N_c] as the last step left it (the snapshot of its last
        Newton iteration, base_engine.py:441); with `constraint_impulses()` it is what the
        reference's TrajectoryBuffer keeps per step for the adjoint.  Kept only by an
        engine built with differentiable_simulation=True."""
        if not self.differentiable_simulation:
            raise RuntimeError("lambda_prev_iter is kept only by an engine built with differentiable_simulation=True "
                               "(the adjoint's input; a forward-only engine skips writing it)")
        lam = torch.zeros(self.num_worlds, self.num_constraints, dtype=torch.float32, device=self.device)
        _lib.check(self.lib.ofb_get_lambda_prev(self._h, _ptr(lam), self._stream()), "ofb_get_lambda_prev")
        return lam

    def model_shape_mu(self):
        """The model's per-shape friction coefficients [W*(S+G)] on the device."""
        return torch.as_tensor(self.model.arrays["shape_mu"], dtype=torch.float32, device=self.device).contiguous()

    def set_shape_mu(self, shape_mu):
        """Replaces the per-shape friction coefficients [W*(S+G)] copied at construction."""
        mu = shape_mu.to(self.device, torch.float32).contiguous().reshape(-1)
        if mu.numel() != self.num_worlds * (self.model.shapes_per_world + self.model.num_global_shapes):
            raise ValueError("shape_mu must hold W*(S+G) values")
        _lib.check(self.lib.ofb_set_shape_mu(self._h, _ptr(mu), self._stream()), "ofb_set_shape_mu")
        self._mu_keepalive = mu

    def _fill_contacts(self, contacts):
        c = self._con
        c.count = contacts.rigid_contact_count.data_ptr()
        c.max_contacts = contacts.rigid_contact_max
        c.point0 = contacts.rigid_contact_point0.data_ptr()
        c.point1 = contacts.rigid_contact_point1.data_ptr()
        c.normal = contacts.rigid_contact_normal.data_ptr()
        c.shape0 = contacts.rigid_contact_shape0.data_ptr()
        c.shape1 = contacts.rigid_contact_shape1.data_ptr()
        c.margin0 = contacts.rigid_contact_margin0.data_ptr()
        c.margin1 = contacts.rigid_contact_margin1.data_ptr()
        return c

    def step_backward(self, *args, **kwargs):
        """Adjoint of one step (core/base_engine.py:528-743).

        `step_backward()` with no arguments is the reference's form: it reads the step
        from `self.data` (filled by the last `step` or by `TrajectoryBuffer.load_step`),
        with dL/dq+ and dL/du+ from data.body_q_grad / data.body_qd_grad (zero when unset),
        and leaves the gradients in data.grads (also returned).  The explicit form takes
        the same data as arguments: `step_backward(state_in, state_out, control, contacts,
        dt, lam, lam_prev, grad_body_q, grad_body_qd, adjoint_multipliers=False)`.  See
        `step_backward_explicit`."""
        if args or any(k != "adjoint_multipliers" for k in kwargs):
            return self.step_backward_explicit(*args, **kwargs)
        d = self.data
        if d.state_in is None:
            raise RuntimeError("step_backward() needs a step: call step() or TrajectoryBuffer.load_step first")
        lam = d.lam if d.lam is not None else self.constraint_impulses()
        lam_prev = d.lam_prev if d.lam_prev is not None else self.constraint_impulses_prev()
        gq = d.body_q_grad if d.body_q_grad is not None else torch.zeros_like(d.state_out.body_q)
        gu = d.body_qd_grad if d.body_qd_grad is not None else torch.zeros_like(d.state_out.body_qd)
        d.grads = self.step_backward_explicit(d.state_in, d.state_out, d.control, d.contacts, d.dt, lam, lam_prev,
                                              gq, gu, **kwargs)
        return d.grads

    def step_backward_explicit(self, state_in: State, state_out: State, control: Control, contacts: Contacts,
                               dt: float, lam, lam_prev, grad_body_q, grad_body_qd, adjoint_multipliers=False):
        """Adjoint of one step with its data as arguments: the step's input state (q-, u-,
        f), its output state (q+, u+), control, contacts, lambda and lambda_prev_iter
        (`constraint_impulses()` / `constraint_impulses_prev()` after that step) and dL/dq+,
        dL/du+.  Returns a dict of dL/d(input): body_q, body_qd, body_f, joint_target_pos,
        joint_target_vel, joint_target_ke, joint_target_kd, shape_mu (per world copy), plus
        w_lambda and w_u when `adjoint_multipliers`.  Per-world adjoint PCR counts and flags
        are then in `stats()`."""
        if self.config.execution.path != "fused":
            raise NotImplementedError("step_backward runs on the fused path")
        d, W = self.device, self.num_worlds
        B, D = self.model.bodies_per_world, self.model.dofs_per_world
        SC = self.model.shapes_per_world + self.model.num_global_shapes

        def f32(t, n):
            t = t.to(d, torch.float32).contiguous().reshape(-1)
            if t.numel() != n:
                raise ValueError(f"expected {n} values, got {t.numel()}")
            return t

        gq, gu = f32(grad_body_q, W * B * 7), f32(grad_body_qd, W * B * 6)
        la, lp = f32(lam, W * self.num_constraints), f32(lam_prev, W * self.num_constraints)
        out = dict(body_q=torch.empty(W * B, 7, device=d), body_qd=torch.empty(W * B, 6, device=d),
                   body_f=torch.empty(W * B, 6, device=d),
                   joint_target_pos=torch.empty(W, D, device=d), joint_target_vel=torch.empty(W, D, device=d),
                   joint_target_ke=torch.empty(W, D, device=d), joint_target_kd=torch.empty(W, D, device=d),
                   shape_mu=torch.empty(W, SC, device=d))
        if adjoint_multipliers:
            out["w_lambda"] = torch.empty(W, self.num_constraints, device=d)
            out["w_u"] = torch.empty(W * B, 6, device=d)
        si, so = _abi.OfbStateIn(), _abi.OfbStateOut()
        si.body_q, si.body_qd, si.body_f = (state_in.body_q.data_ptr(), state_in.body_qd.data_ptr(),
                                            state_in.body_f.data_ptr())
        so.body_q, so.body_qd = state_out.body_q.data_ptr(), state_out.body_qd.data_ptr()
        ctl = _abi.OfbControl()
        ctl.joint_target_pos = control.joint_target_pos.data_ptr()
        ctl.joint_target_vel = control.joint_target_vel.data_ptr()
        gin, gout = _abi.OfbAdjointIn(), _abi.OfbAdjointOut()
        gin.body_q_grad, gin.body_qd_grad = gq.data_ptr(), gu.data_ptr()
        for k, field in (("body_q", "body_q_grad"), ("body_qd", "body_qd_grad"), ("body_f", "body_f_grad"),
                         ("joint_target_pos", "joint_target_pos_grad"),
                         ("joint_target_vel", "joint_target_vel_grad"),
                         ("joint_target_ke", "joint_target_ke_grad"), ("joint_target_kd", "joint_target_kd_grad"),
                         ("shape_mu", "shape_mu_grad"), ("w_lambda", "adjoint_lambda"), ("w_u", "adjoint_body")):
            if k in out and out[k].numel():
                setattr(gout, field, out[k].data_ptr())
        con = self._fill_contacts(contacts)
        _lib.check(self.lib.ofb_step_backward(self._h, C.byref(si), C.byref(so), C.byref(ctl), C.byref(con),
                                              _ptr(la), _ptr(lp), C.byref(gin), C.byref(gout), C.c_float(dt),
                                              self._stream()), "ofb_step_backward")
        self._bwd_keepalive = (gq, gu, la, lp)
        return out
