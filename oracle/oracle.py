"""ctypes binding of the CPU oracle (oracle/fb_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The product
path (ostrich_amd) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from ostrich_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfb_oracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        _lib.oracle_step.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.c_float, C.c_int, vp,
                                     vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        _lib.oracle_adjoint.argtypes = ([vp] * 10 + [C.c_float] + [vp] * 4 + [C.c_int, C.c_double]
                                        + [vp] * 12)
        _lib.oracle_adjoint.restype = C.c_int
        _lib.oracle_weighted_residual.argtypes = [vp] * 10 + [C.c_float] + [vp] * 4
        _lib.oracle_weighted_residual.restype = C.c_double
        _lib.oracle_warm_start.argtypes = [vp] * 5 + [C.c_float, C.c_int] + [vp] * 9
        _lib.oracle_warm_start.restype = C.c_int
        _lib.oracle_linesearch_steps.argtypes = [vp, vp]
        _lib.oracle_linesearch_steps.restype = C.c_int
        _lib.oracle_step.restype = C.c_int
        _lib.oracle_dims.argtypes = [vp, vp, vp, vp, vp]
        _lib.oracle_set_threads.argtypes = [C.c_int]
        _lib.oracle_set_threads.restype = C.c_int
        _lib.oracle_fb_diff.argtypes = [C.c_double] * 5 + [vp]
        _lib.oracle_integrate_pose.argtypes = [vp, vp, vp, C.c_double, vp]
        _lib.oracle_linear_system.argtypes = [vp] * 8 + [C.c_float] + [vp] * 8
        _lib.oracle_pack.argtypes = [vp] * 4 + [C.c_int] + [vp] * 8
        _lib.oracle_pack.restype = C.c_int
    return _lib


def set_threads(n):
    """Threads of the oracle's per-world loops (OpenMP); n <= 0 keeps the current
    setting.  Returns the number in effect."""
    return lib().oracle_set_threads(int(n))


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class HostModel:
    """Model arrays pinned in host memory with an OfbModel pointing at them."""

    def __init__(self, model):
        self.model = model
        self.arrays = {k: np.ascontiguousarray(v) for k, v in model.arrays.items()}
        ptrs = {k: self.arrays[k].ctypes.data for k in self.arrays}
        self.abi = model.to_abi(ptrs)


class HostContacts:
    def __init__(self, contacts):
        c = {k: (v.detach().cpu().numpy() if hasattr(v, "detach") else v) for k, v in contacts.items()}
        self.arrays = dict(
            count=np.ascontiguousarray(np.asarray(c["count"], np.int32).reshape(1)),
            point0=np.ascontiguousarray(np.asarray(c["point0"], np.float32)),
            point1=np.ascontiguousarray(np.asarray(c["point1"], np.float32)),
            normal=np.ascontiguousarray(np.asarray(c["normal"], np.float32)),
            shape0=np.ascontiguousarray(np.asarray(c["shape0"], np.int32)),
            shape1=np.ascontiguousarray(np.asarray(c["shape1"], np.int32)),
            margin0=np.ascontiguousarray(np.asarray(c["margin0"], np.float32)),
            margin1=np.ascontiguousarray(np.asarray(c["margin1"], np.float32)),
        )
        a = self.arrays
        self.abi = _abi.OfbContacts()
        self.abi.count = a["count"].ctypes.data
        self.abi.max_contacts = int(c["max_contacts"])
        for k in ("point0", "point1", "normal", "shape0", "shape1", "margin0", "margin1"):
            setattr(self.abi, k, a[k].ctypes.data)


def dims(model, config):
    hm = HostModel(model)
    cfg = config.to_abi()
    nc, nj, nctrl = C.c_int32(), C.c_int32(), C.c_int32()
    lib().oracle_dims(C.byref(hm.abi), C.byref(cfg), C.byref(nc), C.byref(nj), C.byref(nctrl))
    return nc.value, nj.value, nctrl.value


def pack(model, config, state, contacts, precision=32):
    """Packed contacts after core/contacts.py:25-99 and the configured per-pair reduction
    (collision/top_k.py, fps.py, cluster.py), with state['body_q'] as the step-start
    poses.  Returns count [W] and the slot arrays [W*K(*3)] (point0/1, normal,
    thickness0/1, shape0/1), the layout of OstrichEngine.packed_contacts()."""
    hm = HostModel(model)
    hc = contacts if isinstance(contacts, HostContacts) else HostContacts(contacts)
    cfg = config.to_abi()
    W, K = model.num_worlds, config.contacts.max_per_world
    q = np.ascontiguousarray(np.asarray(state["body_q"], np.float64).reshape(-1, 7))
    out = dict(count=np.zeros(W, np.int32), shape0=np.zeros(W * K, np.int32), shape1=np.zeros(W * K, np.int32),
               point0=np.zeros(W * K * 3), point1=np.zeros(W * K * 3), normal=np.zeros(W * K * 3),
               thickness0=np.zeros(W * K), thickness1=np.zeros(W * K))
    rc = lib().oracle_pack(C.byref(hm.abi), C.byref(cfg), _p(q), C.byref(hc.abi), precision,
                           *(_p(out[k]) for k in ("count", "shape0", "shape1", "point0", "point1", "normal",
                                                  "thickness0", "thickness1")))
    if rc != 0:
        raise RuntimeError("oracle_pack failed")
    return out


def step(model, config, state, control, contacts, dt, precision=32, lam_seed=None, host_model=None):
    """One solver step on the CPU.  Returns a dict with body_q, body_qd (float64), lambda,
    nr_iters, pcr_iters, merit, best_iter, flags, contact_count."""
    hm = host_model or HostModel(model)
    hc = contacts if isinstance(contacts, HostContacts) else HostContacts(contacts)
    cfg = config.to_abi()
    W = model.num_worlds
    WB = W * model.bodies_per_world
    nc, _, _ = dims(model, config)
    q_in = np.ascontiguousarray(np.asarray(state["body_q"], np.float64).reshape(WB, 7))
    qd_in = np.ascontiguousarray(np.asarray(state["body_qd"], np.float64).reshape(WB, 6))
    f_in = np.ascontiguousarray(np.asarray(state["body_f"], np.float32).reshape(WB, 6))
    tp = np.ascontiguousarray(np.asarray(control["joint_target_pos"], np.float32))
    tv = np.ascontiguousarray(np.asarray(control["joint_target_vel"], np.float32))
    seed = None if lam_seed is None else np.ascontiguousarray(np.asarray(lam_seed, np.float64))
    out = dict(
        body_q=np.zeros((WB, 7)), body_qd=np.zeros((WB, 6)), lam=np.zeros((W, nc)),
        nr_iters=np.zeros(W, np.int32), pcr_iters=np.zeros(W, np.int32), merit=np.zeros(W),
        best_iter=np.zeros(W, np.int32), flags=np.zeros(W, np.int32), contact_count=np.zeros(W, np.int32),
        merit_hist=np.full((W, config.nr.max_iters), np.nan),
        residual=np.zeros((W, 6 * model.bodies_per_world + nc)),
        lam_prev=np.zeros((W, nc)),
    )
    rc = lib().oracle_step(C.byref(hm.abi), C.byref(cfg), _p(q_in), _p(qd_in), _p(f_in), _p(tp), _p(tv),
                           C.byref(hc.abi), C.c_float(dt), precision, _p(seed), _p(out["body_q"]),
                           _p(out["body_qd"]), _p(out["lam"]), _p(out["nr_iters"]), _p(out["pcr_iters"]),
                           _p(out["merit"]), _p(out["best_iter"]), _p(out["flags"]), _p(out["contact_count"]),
                           _p(out["merit_hist"]), _p(out["residual"]), _p(out["lam_prev"]))
    if rc != 0:
        raise RuntimeError(f"oracle_step failed ({rc})")
    return out


def adjoint(model, config, state_in, state_out, control, contacts, dt, lam, lam_prev, grad_q, grad_qd,
            precision=64, fd_eps=None, host_model=None):
    """Adjoint of one step at the state the forward step left (OstrichEngine.step_backward).

    state_in: body_q, body_qd, body_f of the step start; state_out: body_q, body_qd after
    the step; grad_q [W*B, 7] and grad_qd [W*B, 6] are dL/dq+ and dL/du+.  Returns dL/dq-,
    dL/du-, dL/df, dL/dtarget_pos, dL/dtarget_vel, dL/dke, dL/dkd, dL/dmu (per world
    shape), the adjoint multipliers w_lambda and w_u, PCR iterations and flags."""
    hm = host_model or HostModel(model)
    hc = contacts if isinstance(contacts, HostContacts) else HostContacts(contacts)
    cfg = config.to_abi()
    W, B, D = model.num_worlds, model.bodies_per_world, model.dofs_per_world
    SC = model.shapes_per_world + model.num_global_shapes
    WB = W * B
    nc, _, _ = dims(model, config)
    f64 = lambda a, shape: np.ascontiguousarray(np.asarray(a, np.float64).reshape(shape))
    f32 = lambda a, shape: np.ascontiguousarray(np.asarray(a, np.float32).reshape(shape))
    qp, qdp = f64(state_in["body_q"], (WB, 7)), f64(state_in["body_qd"], (WB, 6))
    fi = f32(state_in["body_f"], (WB, 6))
    q, qd = f64(state_out["body_q"], (WB, 7)), f64(state_out["body_qd"], (WB, 6))
    tp = f32(control["joint_target_pos"], (W * D,)) if D else np.zeros(1, np.float32)
    tv = f32(control["joint_target_vel"], (W * D,)) if D else np.zeros(1, np.float32)
    la, lp = f64(lam, (W, nc)), f64(lam_prev, (W, nc))
    gq, gu = f64(grad_q, (WB, 7)), f64(grad_qd, (WB, 6))
    eps = config.adjoint.pose_fd_eps if fd_eps is None else fd_eps
    out = dict(body_q=np.zeros((WB, 7)), body_qd=np.zeros((WB, 6)), body_f=np.zeros((WB, 6)),
               joint_target_pos=np.zeros((W, max(D, 1))), joint_target_vel=np.zeros((W, max(D, 1))),
               joint_target_ke=np.zeros((W, max(D, 1))), joint_target_kd=np.zeros((W, max(D, 1))),
               shape_mu=np.zeros((W, SC)), w_lambda=np.zeros((W, nc)), w_u=np.zeros((WB, 6)),
               pcr_iters=np.zeros(W, np.int32), flags=np.zeros(W, np.int32))
    rc = lib().oracle_adjoint(C.byref(hm.abi), C.byref(cfg), _p(qp), _p(qdp), _p(fi), _p(q), _p(qd), _p(tp), _p(tv),
                              C.byref(hc.abi), C.c_float(dt), _p(la), _p(lp), _p(gq), _p(gu), precision,
                              C.c_double(eps), _p(out["body_q"]), _p(out["body_qd"]), _p(out["body_f"]),
                              _p(out["joint_target_pos"]), _p(out["joint_target_vel"]), _p(out["joint_target_ke"]),
                              _p(out["joint_target_kd"]), _p(out["shape_mu"]), _p(out["w_lambda"]), _p(out["w_u"]),
                              _p(out["pcr_iters"]), _p(out["flags"]))
    if rc != 0:
        raise RuntimeError(f"oracle_adjoint failed ({rc})")
    return out


def linear_system(model, config, state, control, contacts, dt, x=None):
    """FP64 dense Schur system of the first Newton iteration (for brute-force checks)."""
    hm = HostModel(model)
    hc = HostContacts(contacts)
    cfg = config.to_abi()
    W = model.num_worlds
    WB = W * model.bodies_per_world
    nc, _, _ = dims(model, config)
    q_in = np.ascontiguousarray(np.asarray(state["body_q"], np.float64).reshape(WB, 7))
    qd_in = np.ascontiguousarray(np.asarray(state["body_qd"], np.float64).reshape(WB, 6))
    f_in = np.ascontiguousarray(np.asarray(state["body_f"], np.float32).reshape(WB, 6))
    tp = np.ascontiguousarray(np.asarray(control["joint_target_pos"], np.float32))
    tv = np.ascontiguousarray(np.asarray(control["joint_target_vel"], np.float32))
    xv = np.zeros((W, nc)) if x is None else np.ascontiguousarray(np.asarray(x, np.float64))
    A = np.zeros((W, nc, nc)); rhs = np.zeros((W, nc)); pinv = np.zeros((W, nc)); Ax = np.zeros((W, nc))
    act = np.zeros((W, nc))
    px = np.zeros((W, nc))
    pit = np.zeros(W, np.int32)
    lib().oracle_linear_system(C.byref(hm.abi), C.byref(cfg), _p(q_in), _p(qd_in), _p(f_in), _p(tp), _p(tv),
                               C.byref(hc.abi), C.c_float(dt), _p(xv), _p(A), _p(rhs), _p(pinv), _p(Ax), _p(act),
                               _p(px), _p(pit))
    return dict(A=A, rhs=rhs, pinv=pinv, Ax=Ax, active=act, pcr_x=px, pcr_iters=pit)


def fb_diff(a, b, alpha, beta, eps_sq):
    out = np.zeros(3)
    lib().oracle_fb_diff(a, b, alpha, beta, eps_sq, _p(out))
    return out


def integrate_pose(vel, pose, com, dt):
    v = np.ascontiguousarray(np.asarray(vel, np.float64))
    p = np.ascontiguousarray(np.asarray(pose, np.float64))
    c = np.ascontiguousarray(np.asarray(com, np.float64))
    out = np.zeros(7)
    lib().oracle_integrate_pose(_p(v), _p(p), _p(c), C.c_double(dt), _p(out))
    return out


def weighted_residual(model, config, state_in, state_out, control, contacts, dt, lam, lam_prev, w_lambda, w_u,
                      host_model=None):
    """FP64 sum_rows w_lambda h_c + sum_bodies w_u . h_d at the given step data."""
    hm = host_model or HostModel(model)
    hc = contacts if isinstance(contacts, HostContacts) else HostContacts(contacts)
    cfg = config.to_abi()
    W, B, D = model.num_worlds, model.bodies_per_world, model.dofs_per_world
    WB = W * B
    nc, _, _ = dims(model, config)
    f64 = lambda a, shape: np.ascontiguousarray(np.asarray(a, np.float64).reshape(shape))
    f32 = lambda a, shape: np.ascontiguousarray(np.asarray(a, np.float32).reshape(shape))
    tp = f32(control["joint_target_pos"], (W * D,)) if D else np.zeros(1, np.float32)
    tv = f32(control["joint_target_vel"], (W * D,)) if D else np.zeros(1, np.float32)
    args = [f64(state_in["body_q"], (WB, 7)), f64(state_in["body_qd"], (WB, 6)), f32(state_in["body_f"], (WB, 6)),
            f64(state_out["body_q"], (WB, 7)), f64(state_out["body_qd"], (WB, 6)), tp, tv]
    tail = [f64(lam, (W, nc)), f64(lam_prev, (W, nc)), f64(w_lambda, (W, nc)), f64(w_u, (WB, 6))]
    return lib().oracle_weighted_residual(C.byref(hm.abi), C.byref(cfg), *[_p(a) for a in args], C.byref(hc.abi),
                                          C.c_float(dt), *[_p(a) for a in tail])


class WarmStarter:
    """Cross-step contact warm start (collision/warm_start.py, position_match) with its
    snapshot held here.  `apply` returns the seed [W, N_c] of a step from the step-start
    state and contacts; `snapshot` stores the step's contacts and lambda."""

    def __init__(self, model, config):
        self.model, self.config = model, config
        self.hm = HostModel(model)
        W, K = model.num_worlds, config.contacts.max_per_world
        self.nc, _, _ = dims(model, config)
        self.count = np.zeros(W, np.int32)
        self.b0 = np.zeros(W * K, np.int32)
        self.b1 = np.zeros(W * K, np.int32)
        self.p = np.zeros(W * K * 3)
        self.n = np.zeros(W * K * 3)
        self.ln = np.zeros(W * K)
        self.lt = np.zeros(W * K * 2)

    def _call(self, state, contacts, dt, mode, lam=None):
        hc = contacts if isinstance(contacts, HostContacts) else HostContacts(contacts)
        cfg = self.config.to_abi()
        WB = self.model.num_worlds * self.model.bodies_per_world
        q = np.ascontiguousarray(np.asarray(state["body_q"], np.float64).reshape(WB, 7))
        qd = np.ascontiguousarray(np.asarray(state["body_qd"], np.float64).reshape(WB, 6))
        seed = np.zeros((self.model.num_worlds, self.nc))
        la = None if lam is None else np.ascontiguousarray(np.asarray(lam, np.float64).reshape(-1))
        rc = lib().oracle_warm_start(C.byref(self.hm.abi), C.byref(cfg), _p(q), _p(qd), C.byref(hc.abi),
                                     C.c_float(dt), mode, _p(la), _p(self.count), _p(self.b0), _p(self.b1),
                                     _p(self.p), _p(self.n), _p(self.ln), _p(self.lt), _p(seed))
        if rc != 0:
            raise RuntimeError("oracle_warm_start failed")
        return seed

    def apply(self, state, contacts, dt):
        return self._call(state, contacts, dt, 0)

    def snapshot(self, state, contacts, dt, lam):
        self._call(state, contacts, dt, 1, lam)


def linesearch_steps(config):
    cfg = config.to_abi()
    out = np.zeros(1024, np.float32)
    n = lib().oracle_linesearch_steps(C.byref(cfg), _p(out))
    return out[:n].copy()
