#!/usr/bin/env python
"""World-steps/s of the Fisher-Burmeister Newton contact solver on one MI355X.

Workload: the uneven-batch skid-steer scene - 4096 worlds, each a 10 kg chassis with three
velocity-driven wheels among 0-24 static boxes (robots start among and against them),
mu U(0.4,1.0), dt 0.05 s, Newton <= 20 iterations to ||h|| <= 1e-5, PCR <= 16 iterations
(Jacobi), contacts within 0.1 m.  The timed window is the scene's 200-step run (after a
few warm-up steps).  The headline variant seeds Newton with the cross-step contact warm
starter ("fused + warm start") and replays each step from a HIP graph; the other variants
(plain per-kernel path, cold start, block-Jacobi, learned warm start) are reported beside
it.  A step is one full solver step (contact packing + Newton loop + backtracking) over
all worlds; contacts are generated between steps by the primitive collider on the GPU and
are not timed, as the reference times collision separately
(profiling/engine_profiler.py:46-53).

Prints ONE JSON line.  `value` is whole-job world-steps/s of the default path over the
timed steps (every world-step, converged or stopped at the Newton cap, as the metric's
run-to-tolerance-or-cap definition counts them), computed from HIP events around each
step on the solver's stream; `converged_world_steps_per_s` counts only the world-steps
that reached ||h|| <= 1e-5.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ostrich_amd import collide, scenes  # noqa: E402
from ostrich_amd.config import (ContactsConfig, ExecutionConfig, LinearSolverConfig,  # noqa: E402
                                NewtonRaphsonConfig, OstrichEngineConfig, WarmStartConfig)
from ostrich_amd.containers import Contacts, Control, State  # noqa: E402

DT = 0.05
K_CONTACTS = 128
HBM_PEAK_GBS = 8000.0
# contact gap of the skid-steer scene: the reference's Helhest scenes generate contacts
# within rigid_gap = 0.1-1.0 m (examples/helhest/gradient/*.py 0.1, obstacle_benchmark.py
# 1.0), so worlds near obstacles carry many speculative contact rows
C1_GAP = 0.1


def solver_config(path, termination="per_world", precond="jacobi", seed="cold", paired=True, K=K_CONTACTS):
    """seed: "cold" (lambda = 0, the reference default), "prev" (the cross-step contact warm
    starter seeds Newton: warm_start.seed_iterate), "learned" (prev + the learned
    per-contact network, gated per world).  paired: worlds of at most 32 rows run two per
    wavefront on the assembled Schur matrix (fused path)."""
    return OstrichEngineConfig(
        nr=NewtonRaphsonConfig(max_iters=20, atol=1e-5),
        linear=LinearSolverConfig(max_iters=16, preconditioner_type=precond),
        contacts=ContactsConfig(max_per_world=K),
        warm_start=WarmStartConfig(enabled=True, seed_iterate=seed in ("prev", "learned")),
        execution=ExecutionConfig(path=path, termination=termination, pair_small_worlds=paired),
    )


def compulsory_bytes_per_world(model, n_c, n_rows):
    """Bytes the fused kernel must move per world-step at minimum: bodies read
    (pose 28, velocity 24, wrench 24, masses 8, inertias 72, com 12) and written (52),
    joints 80 each, dofs 32 each (axis, mode, gains, targets), shapes 36 each, packed
    contacts 52 each, lambda out 4 per row of the reference layout, stats 20."""
    B, J, D = model.bodies_per_world, model.joints_per_world, model.dofs_per_world
    S = model.shapes_per_world + model.num_global_shapes
    return 220 * B + 80 * J + 32 * D + 36 * S + 52 * n_c + 4 * n_rows + 20


def operand_bytes_per_world(n_rows_active, nr_iters, pcr_iters):
    """Operand bytes the algorithm touches per world-step (on-chip in the fused kernel):
    400 B per active row per Newton iteration (assembly, RHS, update and residual
    operands) and 304 B per active row per PCR iteration: the matvec (Jacobian pair read
    twice 96, body indices 8, x / C / Az 12, body scatter and gather 96), the Jacobi
    scaling 12, three dot products 20 and two vector updates 60."""
    return n_rows_active * (400.0 * nr_iters + 304.0 * pcr_iters)


NET_PATH = os.path.join(ROOT, "ostrich_amd", "data", "warm_start_net.npy")
GRAPH_PATH = "fused:jacobi:prev:graph"


def contacts_of(collider, body_q, device):
    """Contacts of body_q: the HIP collider writes its persistent buffers in place; the
    host-side PrimitiveCollider returns fresh tensors."""
    if isinstance(collider, collide.HipCollider):
        return collider.collide(body_q)
    return Contacts.from_dict(collider.collide(body_q), device)


def row_class(rows, max_rows):
    """Class index of worlds with `rows` active rows (tensor): the first class whose largest
    row count (OstrichEngine.class_shapes()["max_rows"], 0 for a class not launched) holds
    them, as the solver buckets them."""
    idx = torch.full_like(rows, -1)
    for c in reversed(range(len(max_rows))):
        if max_rows[c] > 0:
            idx = torch.where(rows <= max_rows[c], torch.full_like(rows, c), idx)
    return idx


def summarise(per_step, W, steps, step_ms, atol, nr_cap):
    """Distribution figures of a variant's timed steps from per-step per-world arrays."""
    nr = np.concatenate(per_step["nr"])
    pcr = np.concatenate(per_step["pcr"])
    merit = np.concatenate(per_step["merit"])
    cnt = np.concatenate(per_step["cnt"])
    conv = merit <= atol * atol
    capped = (nr >= nr_cap) & ~conv
    t = np.sum(step_ms) / 1e3
    h_cap = np.sqrt(merit[capped]) if capped.any() else np.zeros(1)
    hist_edges = [0, 3, 4, 5, 8, 16, 32, 64, 100, 128, 10 ** 9]
    hist = np.histogram(cnt, bins=hist_edges)[0]
    return dict(
        world_steps_per_s=W * steps / t, converged_world_steps_per_s=float(conv.sum()) / t,
        ms_per_step=float(np.mean(step_ms)),
        nr_mean=float(nr.mean()), nr_p95=float(np.percentile(nr, 95)),
        pcr_per_step_mean=float(pcr.mean()), pcr_per_step_p95=float(np.percentile(pcr, 95)),
        pcr_per_nr=float(pcr.sum() / max(nr.sum(), 1)),
        converged_fraction=float(conv.mean()), capped_fraction=float(capped.mean()),
        capped_h={"median": float(np.median(h_cap)), "p90": float(np.percentile(h_cap, 90)),
                  "max": float(h_cap.max())},
        contacts_mean=float(cnt.mean()), contacts_p50=float(np.median(cnt)),
        contacts_p95=float(np.percentile(cnt, 95)), contacts_max=int(cnt.max()), contacts_min=int(cnt.min()),
        contacts_hist={f"{a}-{b - 1}" if b < 10 ** 9 else f">={a}": int(h) for a, b, h in
                       zip(hist_edges[:-1], hist_edges[1:], hist)})


def run_path(variant, model, state, control, steps, warmup, device, collider, K=K_CONTACTS, tune=False,
             class_timing=False, large_world_slots=None):
    """One solver variant "path:preconditioner[:seed[:option...]]" over warmup + steps
    closed-loop steps.  Options: "unpaired" (worlds of at most 32 rows one per wavefront),
    "graph" (each step is a replay of a HIP graph recorded over static tensors; the
    collider's contacts are copied into them before each replay, untimed).  tune: run the
    launch-shape autotuner on the first step's inputs before timing.  Step times are HIP
    events on the stream around each step (or replay); with class_timing (eager fused
    steps) every row-class launch is also timed by its own events.  large_world_slots
    (without tune): the layout of the worlds of more than 256 rows, else the default."""
    from ostrich_amd import OstrichEngine
    parts = variant.split(":")
    path, precond = parts[0], parts[1]
    seed = parts[2] if len(parts) > 2 else "cold"
    opts = set(parts[3:])
    paired = "unpaired" not in opts
    graph = "graph" in opts
    cfg = solver_config(path, precond=precond, seed=seed, paired=paired, K=K)
    eng = OstrichEngine(model, config=cfg, device=device)
    if large_world_slots is not None and path == "fused":
        eng.set_fused_layout(large_world_slots)
    if seed == "learned":
        from ostrich_amd.learned import WarmStartNet
        if not os.path.exists(NET_PATH):
            raise RuntimeError(f"{NET_PATH} missing: train it with python -m ostrich_amd.train_warm_start")
        eng.set_warm_start_net(WarmStartNet(device, params=np.load(NET_PATH)))
    s_a = State.from_numpy(state, device)
    s_b = s_a.clone()
    ctl = Control.from_numpy(control, device)
    stream = torch.cuda.current_stream(device)
    W = model.num_worlds
    base_rows = model_rows(model)
    launch = None
    if tune and path == "fused":
        con0 = contacts_of(collider, s_a.body_q, device)
        launch, times = eng.autotune(s_a, ctl, con0, DT, pairing=paired)
        launch = dict(waves_per_block=launch[0], pair_small_worlds=launch[1], large_world_slots=launch[2],
                      candidates_ms={f"{w}x{'paired' if p else 'single'}/{l}": round(t, 4)
                                     for (w, p, l), t in times.items()})
        print(f"[bench] autotune {variant}: {launch}", file=sys.stderr, flush=True)
    class_timing = class_timing and path == "fused" and not graph
    if class_timing:
        eng.set_class_timing(True)
    g, con_static = None, None
    if graph:
        con_static = contacts_of(collider, s_a.body_q, device).clone()
        g = eng.capture(s_a, s_b, ctl, con_static, DT)
    step_ms = []
    per = dict(nr=[], pcr=[], merit=[], cnt=[])
    cls_ms, cls_worlds, cls_bytes = [], [], []
    for k in range(warmup + steps):
        con = contacts_of(collider, s_a.body_q, device)
        if graph:
            con_static.copy_(con)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        if graph:
            g.replay()
        else:
            eng.step(s_a, s_b, ctl, con, DT)
        e1.record(stream)
        e1.synchronize()
        if k >= warmup:
            step_ms.append(e0.elapsed_time(e1))
            st = eng.stats()
            cnt = eng.packed_contacts()["count"].clamp(max=K)
            per["nr"].append(st["nr_iters"].cpu().numpy())
            per["pcr"].append(st["pcr_iters"].cpu().numpy())
            per["merit"].append(st["merit"].cpu().numpy())
            per["cnt"].append(cnt.cpu().numpy())
            if class_timing:
                ct = eng.class_times()
                shapes = eng.class_shapes()
                rows = base_rows + 3 * cnt
                idx = row_class(rows, shapes["max_rows"])
                by = compulsory_bytes_per_world(model, cnt.double(), eng.num_constraints)
                cls_ms.append(ct["ms"])
                cls_worlds.append(ct["worlds"])
                cls_bytes.append([float(by[idx == c].sum()) for c in range(len(ct["ms"]))])
        if graph:
            s_a.copy_(s_b)  # the graph reads s_a and writes s_b
        else:
            s_a, s_b = s_b, s_a
    torch.cuda.synchronize()
    out = dict(steps=steps, warmup=warmup, variant=variant)
    out.update(summarise(per, W, steps, step_ms, cfg.nr.atol, cfg.nr.max_iters))
    out["step_ms_first20"] = float(np.mean(step_ms[:20]))
    if class_timing and cls_ms:
        ms = np.array(cls_ms)
        wk = np.array(cls_worlds)
        by = np.array(cls_bytes)
        classes = []
        shapes = eng.class_shapes()
        for c in range(ms.shape[1]):
            timed = ms[:, c] >= 0
            if not timed.any() and wk[:, c].max() == 0:
                continue
            classes.append(dict(rows_cap=int(shapes["slots"][c] * shapes["lanes"][c]),
                                max_rows=int(shapes["max_rows"][c]), slots_per_lane=int(shapes["slots"][c]),
                                lanes_per_world=int(shapes["lanes"][c]), worlds_mean=float(wk[:, c].mean()),
                                ms_mean=float(ms[timed, c].mean()) if timed.any() else None,
                                bytes_mean=float(by[:, c].mean())))
        out["classes"] = classes
    if launch is not None:
        out["launch"] = launch
    return out


def full_step_leg(model, state, control, device, steps, warmup, launch=None):
    """Whole simulation steps: the HIP collider, the solver step and the state advance
    recorded in one graph and replayed `steps` times back to back (no host work between
    steps).  Also times the collider alone (HIP events over 50 calls)."""
    from ostrich_amd import OstrichEngine
    cfg = solver_config("fused", seed="prev")
    eng = OstrichEngine(model, config=cfg, device=device)
    if launch is not None:
        eng.set_fused_launch(launch["waves_per_block"], launch["pair_small_worlds"])
        eng.set_fused_layout(launch["large_world_slots"])
    col = collide.HipCollider(model, device=device, gap=C1_GAP, max_contacts_per_world=K_CONTACTS)
    s_a = State.from_numpy(state, device)
    s_b = s_a.clone()
    ctl = Control.from_numpy(control, device)
    stream = torch.cuda.current_stream(device)
    g = eng.capture(s_a, s_b, ctl, None, DT, collider=col, advance=True)
    for _ in range(warmup):
        g.replay()
    e0, e1, e2, e3 = (torch.cuda.Event(enable_timing=True) for _ in range(4))
    e0.record(stream)
    for _ in range(steps):
        g.replay()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1)
    st = eng.stats()
    e2.record(stream)
    for _ in range(50):
        col.collide(s_a.body_q)
    e3.record(stream)
    e3.synchronize()
    W = model.num_worlds
    return dict(world_steps_per_s=W * steps / (ms / 1e3), ms_per_step=ms / steps, steps=steps, warmup=warmup,
                collide_ms=e2.elapsed_time(e3) / 50, contacts=int(col.contacts.rigid_contact_count.item()),
                nr_mean_last_step=float(st["nr_iters"].float().mean().item()),
                note="collide (ofb_collide) + solver step + state advance per graph replay, skid-steer scene")


def model_rows(model):
    """Joint and control rows of one world (every world shares the topology)."""
    from ostrich_amd.model import JOINT_BALL, JOINT_FIXED, JOINT_PRISMATIC, JOINT_REVOLUTE
    J, D = model.joints_per_world, model.dofs_per_world
    per = {JOINT_PRISMATIC: 5, JOINT_REVOLUTE: 5, JOINT_BALL: 3, JOINT_FIXED: 6}
    jt = model.arrays["joint_type"][:J]
    nj = sum(per.get(int(t), 0) for t in jt)
    qs = model.arrays["joint_qd_start"][:J]
    modes = model.arrays["joint_dof_mode"][:D] if D else []
    nctrl = sum(1 for t, q in zip(jt, qs) if int(t) in (JOINT_PRISMATIC, JOINT_REVOLUTE) and int(modes[q]) != 0)
    return nj + nctrl


def scene_leg(name, model_fn, worlds, K, steps, warmup, device, variants=("fused:jacobi:prev",), gap=0.02,
              hf_contacts="point", tune=True):
    """Another workload through the same solver paths (a parity-test configuration, not
    the headline): world-steps/s, Newton counts and contact counts per variant, each
    fused variant with the launch settings the autotuner picks for this workload."""
    model, state, control = model_fn(worlds)
    collider = collide.HipCollider(model, device=device, gap=gap, max_contacts_per_world=K, hf_contacts=hf_contacts)
    out = dict(worlds=worlds, contact_slots=K, gap=gap, heightfield_contacts=hf_contacts, steps=steps,
               warmup=warmup)
    for v in variants:
        print(f"[bench] {name} path={v} worlds={worlds}", file=sys.stderr, flush=True)
        out[v] = run_path(v, model, state, control, steps, warmup, device, collider, K=K, tune=tune)
        print(f"[bench] {name} {v}: {out[v]}", file=sys.stderr, flush=True)
    return out


def learned_warm_start_leg(device, skid_result=None, skid_prev=None, stack_worlds=64, max_cubes=32, steps=10,
                           warmup=3):
    """The learned warm start on its evaluation batch: 2048 box-stack worlds of 1-32 cubes
    (64 worlds per cube count, seeds 3000 + n; the network was trained on skid-steer and
    tracked-flipper rollouts only)
    and the 4096 skid-steer worlds of the headline (the fused:jacobi:learned and
    fused:jacobi:prev variants above).  Per cube count the same closed-loop steps run with
    the previous-step seed and with the network gated per world; reported: Newton
    iterations per world-step and world-steps/s both ways, over the stacks and over the
    whole batch."""
    rows = []
    for n in range(1, max_cubes + 1):
        model, state, control = scenes.box_stack(stack_worlds, n_cubes=n, seed=3000 + n)
        K = max(8, 4 * n + 4)
        col = collide.HipCollider(model, device=device, gap=0.02, max_contacts_per_world=K)
        row = dict(cubes=n)
        for v in ("fused:jacobi:prev", "fused:jacobi:learned"):
            r = run_path(v, model, state, control, steps, warmup, device, col, K=K)
            row[v.split(":")[2]] = dict(nr_mean=r["nr_mean"], ms_per_step=r["ms_per_step"],
                                        converged_fraction=r["converged_fraction"])
        rows.append(row)

    def total(rs, key, extra=None):
        ws = stack_worlds * steps * len(rs)
        nr = sum(stack_worlds * steps * r[key]["nr_mean"] for r in rs)
        ms = sum(steps * r[key]["ms_per_step"] for r in rs)
        if extra:
            ws += extra["worlds"] * extra["steps"]
            nr += extra["worlds"] * extra["steps"] * extra["nr_mean"]
            ms += extra["steps"] * extra["ms_per_step"]
        return dict(world_steps=ws, nr_mean=nr / ws, world_steps_per_s=ws / (ms / 1e3))
    out = dict(stack_worlds=stack_worlds, cube_counts=f"1-{max_cubes}", steps=steps, warmup=warmup,
               per_cube_count=rows, stacks=dict(prev=total(rows, "prev"), learned=total(rows, "learned")))
    if skid_result and skid_prev and "nr_mean" in skid_result and "nr_mean" in skid_prev:
        sk = lambda r: dict(worlds=4096, steps=r["steps"], nr_mean=r["nr_mean"], ms_per_step=r["ms_per_step"])
        out["whole_batch"] = dict(prev=total(rows, "prev", sk(skid_prev)),
                                  learned=total(rows, "learned", sk(skid_result)))
        out["whole_batch"]["nr_change"] = out["whole_batch"]["learned"]["nr_mean"] / \
            out["whole_batch"]["prev"]["nr_mean"] - 1.0
    out["stacks"]["nr_change"] = out["stacks"]["learned"]["nr_mean"] / out["stacks"]["prev"]["nr_mean"] - 1.0
    return out


def adjoint_leg(device, worlds=1024, steps=50, checkpoint_every=10, seed=3):
    """Differentiable 50-step rollout of skid-steer robots on a flat plane (mu 0.8, wheel
    velocity targets U(-3,3) per step), loss on the final chassis x, y and yaw, gradients
    w.r.t. every step's targets and the friction coefficient (grad_mu_norm: the norm over
    worlds of dL/dmu, all of a world's shapes moved together).  Times the
    forward rollout and the checkpointed backward (re-simulation of every window plus
    one adjoint per step) with HIP events on the solver stream."""
    from ostrich_amd import OstrichEngine
    from ostrich_amd.differentiable import CheckpointedRollout
    model, state, control = scenes.skid_steer(worlds, seed=seed, max_obstacles=0, mu_range=(0.8, 0.8))
    cfg = solver_config("fused")
    eng = OstrichEngine(model, config=cfg, device=device, differentiable_simulation=True)
    col = collide.HipCollider(model, device=device, gap=0.02, max_contacts_per_world=K_CONTACTS)
    collide_fn = lambda q: col.collide(q).clone()  # each step record keeps its own contacts
    rng = np.random.default_rng(seed)
    D, B = model.dofs_per_world, model.bodies_per_world
    controls = [Control(torch.as_tensor(control["joint_target_pos"], device=device),
                        torch.as_tensor(rng.uniform(-3, 3, worlds * D).astype(np.float32), device=device))
                for _ in range(steps)]
    s0 = State.from_numpy(state, device)
    target = torch.as_tensor(np.stack([rng.uniform(-1, 1, worlds), rng.uniform(-1, 1, worlds),
                                       rng.uniform(-np.pi / 4, np.pi / 4, worlds)], 1), dtype=torch.float32,
                             device=device)
    stream = torch.cuda.current_stream(device)
    def loss_grad(fin):
        """Loss on the final chassis x, y, yaw and its gradient w.r.t. body_q [W*B, 7]."""
        q = fin.body_q.reshape(worlds, B, 7)[:, 0].detach().clone().requires_grad_(True)
        yaw = torch.atan2(2 * (q[:, 6] * q[:, 5] + q[:, 3] * q[:, 4]), 1 - 2 * (q[:, 4] ** 2 + q[:, 5] ** 2))
        loss = ((q[:, 0] - target[:, 0]) ** 2 + (q[:, 1] - target[:, 1]) ** 2 + (yaw - target[:, 2]) ** 2).sum()
        loss.backward()
        gq = torch.zeros_like(fin.body_q).reshape(worlds, B, 7)
        gq[:, 0] = q.grad
        return loss, gq.reshape(-1, 7)

    ro = CheckpointedRollout(eng, collide_fn, DT, checkpoint_every=checkpoint_every)
    # warm-up over the whole rollout and the loss: the first launch of each kernel
    # instance loads its code object and torch's first autograd call starts its engine,
    # neither of which is rollout time
    ro.backward(loss_grad(ro.forward(s0, controls))[1], torch.zeros_like(s0.body_qd))
    torch.cuda.synchronize()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    ro = CheckpointedRollout(eng, collide_fn, DT, checkpoint_every=checkpoint_every)
    e0.record(stream)
    fin = ro.forward(s0, controls)
    e1.record(stream)
    loss, gq = loss_grad(fin)
    g = ro.backward(gq, torch.zeros_like(fin.body_qd))
    e2.record(stream)
    e2.synchronize()
    fwd, bwd = e0.elapsed_time(e1), e1.elapsed_time(e2)
    # the adjoint solves alone: a second backward with each step_backward's own HIP-event
    # time (packing + adjoint kernels on the step's stream) collected after the call
    adj = []
    inner = eng.step_backward_explicit

    def timed_backward(*a, **kw):
        out = inner(*a, **kw)
        adj.append(eng.last_solve_ms())
        return out
    eng.step_backward_explicit = timed_backward
    ro.backward(gq, torch.zeros_like(fin.body_qd))
    del eng.step_backward_explicit  # back to the class's method
    if len(adj) != steps:
        raise RuntimeError(f"adjoint timing saw {len(adj)} adjoint solves for {steps} steps")
    return dict(worlds=worlds, steps=steps, checkpoint_every=checkpoint_every, forward_ms=fwd, backward_ms=bwd,
                backward_ms_per_step=bwd / steps, adjoint_solve_ms_per_step=float(np.mean(adj)),
                loss=float(loss.item()), grad_target_vel_norm=float(g["joint_target_vel"].norm().item()),
                grad_mu_norm=float(g["shape_mu"].reshape(worlds, -1).sum(1).norm().item()),
                note="forward and backward include the collider and host work of every step; the backward "
                     "re-simulates each 10-step window from its checkpoint; adjoint_solve_ms_per_step is the "
                     "adjoint solve of one step (HIP events on its stream)")


def cpu_baseline(model_fn, sample_worlds, sample_steps, precond="jacobi", seed="prev", threads=1, gap=C1_GAP):
    """The oracle (FP32 CPU restatement of the solve; FP64 warm starter) on the first
    `sample_worlds` worlds of the same seeded scene, with the same Newton seeding as the
    headline, on `threads` host threads (OpenMP over worlds); warm start + solver calls
    are timed, collision is not."""
    from oracle import oracle
    used = oracle.set_threads(threads)
    model, state, control = model_fn(sample_worlds)
    cfg = solver_config("fused", precond=precond, seed=seed)
    col = collide.PrimitiveCollider(model, device="cpu", gap=gap, max_contacts_per_world=K_CONTACTS)
    hm = oracle.HostModel(model)
    ws = oracle.WarmStarter(model, cfg) if seed != "cold" else None
    st = {k: np.asarray(v, np.float64) for k, v in state.items()}
    t = 0.0
    for _ in range(sample_steps):
        c = col.collide(torch.as_tensor(st["body_q"], dtype=torch.float32))
        hc = oracle.HostContacts(c)
        t0 = time.perf_counter()
        lam_seed = ws.apply(st, hc, DT) if ws is not None else None
        o = oracle.step(model, cfg, st, control, hc, DT, precision=32, lam_seed=lam_seed, host_model=hm)
        if ws is not None:
            ws.snapshot(st, hc, DT, o["lam"])
        t += time.perf_counter() - t0
        st["body_q"], st["body_qd"] = o["body_q"], o["body_qd"]
    return dict(value=sample_worlds * sample_steps / t, unit="world-steps/s", cores=used, kind="port",
                sample=f"oracle FP32 (C++ CPU restatement, {used} thread{'s' if used > 1 else ''}), first "
                       f"{sample_worlds} worlds of the same seeded scene, {sample_steps} closed-loop steps, Newton "
                       f"seed '{seed}', solver and warm-start calls only ({t:.1f} s)")


def host_cpu():
    """Host CPU model and logical core count (lscpu / nproc where available)."""
    info = dict(nproc=os.cpu_count())
    try:
        import subprocess
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, ValueError):
        pass
    return info


def fp64_error_leg(model_fn, worlds, lead_steps, device, variant="fused:jacobi:prev"):
    """Error of one GPU step against the FP64 oracle on the same inputs.  The first
    `worlds` worlds of the scene are stepped on the GPU for `lead_steps` closed-loop steps;
    the next step's inputs (state, contacts, the warm starter's Newton seed) are then
    solved by the GPU, the FP32 oracle and the FP64 oracle.  Reports per-world maximum
    pose / velocity errors of the GPU and of the FP32 oracle against FP64 (the FP32
    rounding spread of this system), over all worlds and over the worlds that all three
    solves converge, and the Newton-count agreement."""
    from oracle import oracle
    from ostrich_amd import OstrichEngine
    model, state, control = model_fn(worlds)
    parts = variant.split(":")
    cfg = solver_config(parts[0], precond=parts[1], seed=parts[2])
    eng = OstrichEngine(model, config=cfg, device=device)
    col = collide.HipCollider(model, device=device, gap=C1_GAP, max_contacts_per_world=K_CONTACTS)
    s_a = State.from_numpy(state, device)
    s_b = s_a.clone()
    ctl = Control.from_numpy(control, device)
    for _ in range(lead_steps):
        eng.step(s_a, s_b, ctl, col.collide(s_a.body_q), DT)
        s_a, s_b = s_b, s_a
    con = col.collide(s_a.body_q).clone()
    st_in = {k: getattr(s_a, k).cpu().numpy() for k in ("body_q", "body_qd", "body_f")}
    eng.step(s_a, s_b, ctl, con, DT)
    seed = eng.lambda_seed().cpu().numpy()
    stats = {k: v.cpu().numpy() for k, v in eng.stats().items()}
    q_gpu, u_gpu = s_b.body_q.cpu().numpy(), s_b.body_qd.cpu().numpy()
    c = con.to_dict() if hasattr(con, "to_dict") else {
        "count": con.rigid_contact_count, "point0": con.rigid_contact_point0, "point1": con.rigid_contact_point1,
        "normal": con.rigid_contact_normal, "shape0": con.rigid_contact_shape0, "shape1": con.rigid_contact_shape1,
        "margin0": con.rigid_contact_margin0, "margin1": con.rigid_contact_margin1,
        "max_contacts": con.rigid_contact_point0.shape[0]}
    hc = oracle.HostContacts(c)
    o32 = oracle.step(model, cfg, st_in, control, hc, DT, precision=32, lam_seed=seed)
    o64 = oracle.step(model, cfg, st_in, control, hc, DT, precision=64, lam_seed=seed)
    B = model.bodies_per_world
    def per_world(a, b, cols):
        return np.abs(a.reshape(worlds, B, -1)[..., cols] - b.reshape(worlds, B, -1)[..., cols]).max(axis=(1, 2))
    gq, gu = per_world(q_gpu, o64["body_q"], slice(0, 3)), per_world(u_gpu, o64["body_qd"], slice(0, 6))
    sq, su = per_world(o32["body_q"], o64["body_q"], slice(0, 3)), per_world(o32["body_qd"], o64["body_qd"], slice(0, 6))
    atol_sq = cfg.nr.atol ** 2
    conv = (stats["merit"] <= atol_sq) & (o32["merit"] <= atol_sq) & (o64["merit"] <= atol_sq)
    def dist(x, m=None):
        x = x if m is None else x[m]
        if x.size == 0:
            return None
        return {"median": float(np.median(x)), "p95": float(np.percentile(x, 95)), "max": float(x.max())}
    return dict(worlds=worlds, step=lead_steps, variant=variant, converged_all_three=int(conv.sum()),
                gpu_vs_fp64_pos_m=dist(gq), gpu_vs_fp64_vel_mps=dist(gu),
                fp32_oracle_vs_fp64_pos_m=dist(sq), fp32_oracle_vs_fp64_vel_mps=dist(su),
                gpu_vs_fp64_pos_m_converged=dist(gq, conv), gpu_vs_fp64_vel_mps_converged=dist(gu, conv),
                fp32_oracle_vs_fp64_vel_mps_converged=dist(su, conv),
                nr_within_1_of_fp64=float(np.mean(np.abs(stats["nr_iters"] - o64["nr_iters"]) <= 1)),
                nr_within_1_of_fp64_converged=float(np.mean(np.abs(stats["nr_iters"] - o64["nr_iters"])[conv] <= 1))
                if conv.any() else None)


def _yaw_loss(q, target, W, B):
    """Loss of the differentiable rollout on the final chassis x, y and yaw, per world, and
    its gradient w.r.t. body_q [W*B, 7] (numpy, float64)."""
    qc = q.reshape(W, B, 7)[:, 0]
    x, y, qx, qy, qz, qw = qc[:, 0], qc[:, 1], qc[:, 3], qc[:, 4], qc[:, 5], qc[:, 6]
    a, b = 2 * (qw * qz + qx * qy), 1 - 2 * (qy ** 2 + qz ** 2)
    yaw = np.arctan2(a, b)
    L = (x - target[:, 0]) ** 2 + (y - target[:, 1]) ** 2 + (yaw - target[:, 2]) ** 2
    g = np.zeros((W, B, 7))
    g[:, 0, 0], g[:, 0, 1] = 2 * (x - target[:, 0]), 2 * (y - target[:, 1])
    dyaw, den = 2 * (yaw - target[:, 2]), a * a + b * b
    da, db = b / den * dyaw, -a / den * dyaw
    g[:, 0, 3], g[:, 0, 4] = da * 2 * qy, da * 2 * qx - db * 4 * qy
    g[:, 0, 5], g[:, 0, 6] = da * 2 * qw - db * 4 * qz, da * 2 * qz
    return L, g.reshape(W * B, 7)


def adjoint_fd_check(device, worlds=8, horizons=(1, 50), h=1e-4, seed=3):
    """The differentiable rollout's gradients on `worlds` skid-steer worlds (flat plane,
    mu 0.8, wheel targets U(-3,3) per step, loss on the final chassis x, y, yaw) against
    central differences (step h) of the FP64 oracle rollout: dL/d(targets) along two
    random directions and dL/dmu (every shape's mu moved together), per world.  Reported
    for the GPU adjoint (FP32) and for the FP64 oracle adjoint along the FP64 trajectory,
    whose deviation from the differences is the method's own (its Newton matrix keeps
    contact Jacobians at the step-start poses and contact compliance on its diagonal, and
    capped Newton loops make the rollout non-smooth)."""
    import copy
    from oracle import oracle
    from ostrich_amd import OstrichEngine
    from ostrich_amd.differentiable import CheckpointedRollout
    model, state, control = scenes.skid_steer(worlds, seed=seed, max_obstacles=0, mu_range=(0.8, 0.8))
    cfg = solver_config("fused")
    W, D, B = worlds, model.dofs_per_world, model.bodies_per_world
    N = max(horizons)
    rng = np.random.default_rng(seed)
    TV = np.stack([rng.uniform(-3, 3, W * D) for _ in range(N)]).astype(np.float32)
    target = np.stack([rng.uniform(-1, 1, W), rng.uniform(-1, 1, W), rng.uniform(-np.pi / 4, np.pi / 4, W)], 1)
    hcol = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=32)
    cfg_o = solver_config("fused", K=32)
    tp = control["joint_target_pos"]
    def contacts(q):
        c = hcol.collide(torch.as_tensor(np.asarray(q), dtype=torch.float32))
        return {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}
    def rollout64(tv, n, m=model, rec=False):
        st, recs = {k: np.asarray(v, np.float64) for k, v in state.items()}, []
        for t in range(n):
            ctl = dict(joint_target_pos=tp, joint_target_vel=tv[t])
            c = contacts(st["body_q"])
            o = oracle.step(m, cfg_o, st, ctl, c, DT, precision=64)
            if rec:
                recs.append((st, o, c, ctl))
            st = dict(st, body_q=o["body_q"], body_qd=o["body_qd"])
        return st, recs
    eng = OstrichEngine(model, config=solver_config("fused", K=32), device=device, differentiable_simulation=True)
    gcol = collide.HipCollider(model, device=device, gap=0.02, max_contacts_per_world=32)
    dirs = [np.random.default_rng(100 + k).normal(size=TV.shape) for k in range(2)]
    out = dict(worlds=W, h=h, directions="two random directions in the wheel targets, and mu")
    def rel(a, f):
        r = np.abs(a - f) / np.maximum(np.abs(f), 1e-12)
        return {"median": float(np.median(r)), "max": float(r.max())}
    for n in horizons:
        # GPU gradient (FP32 adjoint through the checkpointed rollout)
        controls = [Control(torch.as_tensor(tp, device=device), torch.as_tensor(TV[t], device=device))
                    for t in range(n)]
        ro = CheckpointedRollout(eng, lambda q: gcol.collide(q).clone(), DT, checkpoint_every=10)
        fin = ro.forward(State.from_numpy(state, device), controls)
        _, gq = _yaw_loss(fin.body_q.cpu().numpy().astype(np.float64), target, W, B)
        g = ro.backward(torch.as_tensor(gq, dtype=torch.float32, device=device), torch.zeros_like(fin.body_qd))
        g_tv = g["joint_target_vel"].cpu().numpy().reshape(n, W * D).astype(np.float64)
        g_mu = g["shape_mu"].cpu().numpy().reshape(W, -1).sum(1).astype(np.float64)
        # FP64 oracle adjoint along the FP64 trajectory
        st, recs = rollout64(TV, n, rec=True)
        L0, gq64 = _yaw_loss(st["body_q"], target, W, B)
        gu64, o_tv, o_mu = np.zeros((W * B, 6)), np.zeros((n, W * D)), np.zeros(W)
        for t in reversed(range(n)):
            s_in, o, c, ctl = recs[t]
            a = oracle.adjoint(model, cfg_o, s_in, dict(body_q=o["body_q"], body_qd=o["body_qd"]), ctl, c, DT,
                               o["lam"], o["lam_prev"], gq64, gu64, precision=64, fd_eps=1e-6)
            o_tv[t] = a["joint_target_vel"].reshape(-1)
            o_mu += a["shape_mu"].sum(1)
            gq64, gu64 = a["body_q"], a["body_qd"]
        res = dict(steps=n, converged_fraction_fp64=float(np.mean([r[1]["merit"] <= cfg_o.nr.atol ** 2
                                                                     for r in recs])))
        for k, v in enumerate(dirs):
            vt = v[:n]
            tp_, tm_ = (TV[:n] + h * vt).astype(np.float32), (TV[:n] - h * vt).astype(np.float32)
            fd = (_yaw_loss(rollout64(tp_, n)[0]["body_q"], target, W, B)[0]
                  - _yaw_loss(rollout64(tm_, n)[0]["body_q"], target, W, B)[0])
            step = (tp_.astype(np.float64) - tm_.astype(np.float64)).reshape(n, W, D)
            res[f"targets_dir{k}"] = dict(gpu_vs_fd=rel((g_tv.reshape(n, W, D) * step).sum((0, 2)), fd),
                                          fp64_adjoint_vs_fd=rel((o_tv.reshape(n, W, D) * step).sum((0, 2)), fd))
        mp, mm = copy.deepcopy(model), copy.deepcopy(model)
        mp.arrays["shape_mu"] = (model.arrays["shape_mu"] + h).astype(np.float32)
        mm.arrays["shape_mu"] = (model.arrays["shape_mu"] - h).astype(np.float32)
        dmu = float(mp.arrays["shape_mu"].astype(np.float64)[0] - mm.arrays["shape_mu"].astype(np.float64)[0])
        fd_mu = (_yaw_loss(rollout64(TV, n, m=mp)[0]["body_q"], target, W, B)[0]
                 - _yaw_loss(rollout64(TV, n, m=mm)[0]["body_q"], target, W, B)[0]) / dmu
        res["mu"] = dict(gpu=g_mu.tolist(), fp64_adjoint=o_mu.tolist(), fd=fd_mu.tolist())
        out[f"horizon_{n}"] = res
    return out


METRIC = "world-steps/s at FB residual\u22641e-5 (plain vs fused+warm-start); adjoint ms"


def kernel_name(c):
    """Kernel instance of a class entry of run_path's "classes"."""
    R, G = c["slots_per_lane"], c["lanes_per_world"]
    if G == 32:
        return "fused_step_kernel<1, *, 32, true> (worlds of at most 32 rows, two per wavefront)"
    if G == 64 and R == 1:
        return "fused_step_kernel<1, *, 64, true> (worlds of at most 64 rows, dense Schur matrix)"
    if G == 64:
        return f"fused_step_kernel<{R}, *, 64, false> (worlds of at most {R * G} rows, matrix-free operator)"
    return (f"fused_step_kernel<{R}, false, {G}, false> (worlds of {c['max_rows'] // 2 + 1}-{c['max_rows']} rows "
            f"over {G // 64} wavefronts, matrix-free operator)")


def dominant_class_roofline(res):
    """Roofline of the eager variant's dominant row-class launch, the one that keeps the most
    wavefronts busy for the longest (wavefronts of its worlds x its mean duration): the
    compulsory HBM bytes of the worlds it solved per launch over its HIP-event time."""
    classes = [c for c in res.get("classes", []) if c.get("ms_mean")]
    if not classes:
        return None
    def wave_ms(x):
        waves = x["worlds_mean"] * x["lanes_per_world"] / 64.0
        return x["ms_mean"] * waves
    c = max(classes, key=wave_ms)
    achieved = c["bytes_mean"] / (c["ms_mean"] / 1e3) / 1e9
    kind = kernel_name(c)
    return dict(achieved=achieved, frac=achieved / HBM_PEAK_GBS, kernel=kind, kernel_ms=c["ms_mean"],
                worlds_per_launch=c["worlds_mean"], bytes_per_launch=c["bytes_mean"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed steps (the scene's run is 200 steps)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--paths",
                    default=f"fused:jacobi:prev,{GRAPH_PATH},fused:jacobi:prev:unpaired,fused:jacobi:cold,"
                            "plain:jacobi:cold,fused:per_body_pair:prev,fused:jacobi:learned",
                    help="comma-separated path:preconditioner[:seed[:option]] variants; the headline is "
                         f"{GRAPH_PATH} when listed, else the first")
    ap.add_argument("--plain-steps", type=int, default=20,
                    help="timed steps of the plain per-kernel path (the first of the same run)")
    ap.add_argument("--cpu-sample-worlds", type=int, default=128)
    ap.add_argument("--cpu-sample-steps", type=int, default=40)
    ap.add_argument("--cpu-all-core-worlds", type=int, default=1024)
    ap.add_argument("--fp64-worlds", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-scenes", action="store_true", help="skip the tracked-robot and box-stack legs")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape (no autotuner run)")
    ap.add_argument("--scene-only", choices=("tracked_flipper", "box_stack"), default=None,
                    help="run only that workload's leg (previous-step warm start, class timing) and print its JSON "
                         "(for profiling runs)")
    ap.add_argument("--large-world-slots", type=int, choices=(2, 4), default=None,
                    help="with --no-tune and --scene-only: row slots per lane of the worlds of more than 256 rows "
                         "(the layout the autotuner picks for that workload, for profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device copy (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: worlds are independent, so the solver shards by world range "
              "with no exchange; multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    device = "cuda:0"
    if args.scene_only:
        if args.scene_only == "tracked_flipper":
            model, state, control = scenes.tracked_flipper(1024, seed=2)
            col = collide.HipCollider(model, device=device, gap=0.05, max_contacts_per_world=256, hf_contacts="cells")
            K = 256
        else:
            model, state, control = scenes.box_stack(16, 3, seed=0)
            col = collide.HipCollider(model, device=device, gap=0.02, max_contacts_per_world=16)
            K = 16
        res = run_path("fused:jacobi:prev", model, state, control, args.steps, args.warmup, device, col, K=K,
                       class_timing=True, tune=not args.no_tune, large_world_slots=args.large_world_slots)
        print(json.dumps(dict(workload=args.scene_only, worlds=model.num_worlds, **res)))
        return
    W = args.worlds
    model, state, control = scenes.skid_steer(W, seed=1)
    collider = collide.HipCollider(model, device=device, gap=C1_GAP, max_contacts_per_world=K_CONTACTS)
    paths = [p for p in args.paths.split(",") if p]
    results = {}
    for p in paths:
        steps = min(args.steps, args.plain_steps) if p.startswith("plain") else args.steps
        print(f"[bench] path={p} worlds={W} warmup={args.warmup} steps={steps}", file=sys.stderr, flush=True)
        try:
            results[p] = run_path(p, model, state, control, steps, args.warmup, device, collider,
                                  tune=not args.no_tune and p in (paths[0], GRAPH_PATH),
                                  class_timing=p == paths[0])
        except RuntimeError as e:
            results[p] = dict(error=str(e))
        print(f"[bench] {p}: {results[p]}", file=sys.stderr, flush=True)
    # headline: the autotuned fused step replayed from a HIP graph when that variant ran,
    # else the first variant; the roofline comes from the eager variant's class events
    # (a graph records no events)
    eager_path = paths[0]
    main_path = GRAPH_PATH if "world_steps_per_s" in results.get(GRAPH_PATH, {}) else eager_path
    r = results[main_path]
    re_ = results[eager_path]
    roof = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None}
    dom = dominant_class_roofline(re_)
    if dom:
        roof.update(dom)
        tpath = os.path.join(ROOT, "profiles", "fused_step_traffic_skid_steer_4096.json")
        if os.path.exists(tpath):
            with open(tpath) as f:
                t = json.load(f)
            name = dom["kernel"].replace("*", "false").split(" (")[0]
            k = t.get("kernels", {}).get(name)
            if k is not None:
                roof["traffic"] = k["fetch_bytes"] + k["write_bytes"]
                roof["traffic_source"] = ("not measured in this run: HBM bytes per launch of that kernel from "
                                          "rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a 256 MiB copy, "
                                          f"{os.path.relpath(tpath, ROOT)}")
        roof["note"] = ("achieved = compulsory HBM bytes of the class's worlds per launch over the launch's "
                        "HIP-event duration; the kernel is bound by instruction issue and LDS latency of one "
                        "wavefront per world (two per wavefront in the paired class), not by HBM")
    short = {k: r[k] for k in ("nr_mean", "nr_p95", "pcr_per_step_mean", "pcr_per_step_p95", "pcr_per_nr",
                               "converged_fraction", "capped_fraction", "capped_h", "contacts_mean", "contacts_p50",
                               "contacts_p95", "contacts_max", "contacts_min", "contacts_hist") if k in r}
    out = {
        "metric": METRIC,
        "value": r["world_steps_per_s"],
        "unit": "world-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded skid-steer scene, contacts from the HIP primitive collider)",
        "config": {"workload": f"uneven-batch skid-steer, {W} worlds among 0-24 static boxes (robots start among "
                               "and against them), contacts within 0.1 m, dt 0.05, 200-step run, NR<=20 to "
                               "||h||<=1e-5 or the cap, PCR<=16 (Jacobi, the library default), 128 contact slots, "
                               "Newton seeded by the previous step's contacts", "worlds": W,
                   "dt": DT, "variant": main_path, "termination": "per_world",
                   "launch": r.get("launch", re_.get("launch"))},
        "converged_world_steps_per_s": r["converged_world_steps_per_s"],
        "roofline": roof,
        "solver": short,
        "classes": re_.get("classes"),
        "paths": {p: results[p] for p in paths},
    }
    fj, pj = results.get("fused:jacobi:cold", {}), results.get("plain:jacobi:cold", {})
    if "step_ms_first20" in fj and "ms_per_step" in pj:
        # same window: the plain path's timed steps are the first of the run
        out["fused_over_plain_cold"] = pj["ms_per_step"] / fj["step_ms_first20"]
    fw = results.get(main_path, {})
    if "step_ms_first20" in fw and "ms_per_step" in pj:
        out["fused_warm_over_plain_cold"] = pj["ms_per_step"] / fw["step_ms_first20"]
    print("[bench] full simulation step (collide + solve, one graph)", file=sys.stderr, flush=True)
    out["full_step"] = full_step_leg(model, state, control, device, args.steps, args.warmup,
                                     launch=r.get("launch", re_.get("launch")))
    print(f"[bench] full_step: {out['full_step']}", file=sys.stderr, flush=True)
    if not args.no_scenes:
        # tracked flipper robots on a Gaussian-bump heightfield (1024 worlds, anisotropic
        # track friction, position-controlled flippers, four grid-cell contacts per track
        # sphere within 0.05 m) and 3-cube box stacks (16 worlds)
        sc_steps = min(args.steps, 40)
        out["tracked_flipper"] = scene_leg("tracked_flipper", lambda n: scenes.tracked_flipper(n, seed=2), 1024, 256,
                                           sc_steps, max(args.warmup, 6), device,
                                           ("fused:jacobi:prev", "fused:jacobi:cold"), gap=0.05,
                                           hf_contacts="cells")
        out["box_stack"] = scene_leg("box_stack", lambda n: scenes.box_stack(n, 3, seed=0), 16, 16, sc_steps,
                                     args.warmup, device, ("fused:jacobi:prev", "fused:jacobi:cold",
                                                           "plain:jacobi:cold"))
    if not args.no_scenes and "fused:jacobi:learned" in results:
        print("[bench] learned warm start on its evaluation batch", file=sys.stderr, flush=True)
        out["learned_warm_start"] = learned_warm_start_leg(device, results.get("fused:jacobi:learned"),
                                                           results.get("fused:jacobi:prev"))
        print(f"[bench] learned_warm_start: stacks {out['learned_warm_start']['stacks']}, whole batch "
              f"{out['learned_warm_start'].get('whole_batch')}", file=sys.stderr, flush=True)
    if not args.no_adjoint:
        print("[bench] adjoint rollout", file=sys.stderr, flush=True)
        out["adjoint"] = adjoint_leg(device)
        # the metric's second figure: the adjoint solve of one step (1024 worlds)
        out["adjoint_ms"] = out["adjoint"]["adjoint_solve_ms_per_step"]
        print(f"[bench] adjoint: {out['adjoint']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        mp = main_path.split(":")
        seed = mp[2] if len(mp) > 2 and mp[2] != "learned" else "cold"
        fn = lambda n: scenes.skid_steer(n, seed=1)
        out["cpu_baseline"] = cpu_baseline(fn, args.cpu_sample_worlds, args.cpu_sample_steps, precond=mp[1],
                                           seed=seed, threads=1)
        print(f"[bench] cpu_baseline: {out['cpu_baseline']}", file=sys.stderr, flush=True)
        host = host_cpu()
        out["cpu_baseline_all_cores"] = cpu_baseline(fn, args.cpu_all_core_worlds, args.cpu_sample_steps,
                                                     precond=mp[1], seed=seed, threads=min(16, host["nproc"] or 1))
        out["cpu_baseline_all_cores"]["host"] = host
        print(f"[bench] cpu_baseline_all_cores: {out['cpu_baseline_all_cores']}", file=sys.stderr, flush=True)
        # the oracle as the checker of the GPU results: FP64 error of one step, and the
        # differentiable rollout's gradients against FP64 central differences
        print("[bench] fp64 error", file=sys.stderr, flush=True)
        out["fp64_error"] = fp64_error_leg(lambda n: scenes.skid_steer(n, seed=1), args.fp64_worlds, 40, device)
        print(f"[bench] fp64_error: {out['fp64_error']}", file=sys.stderr, flush=True)
        if not args.no_adjoint:
            print("[bench] adjoint fd check", file=sys.stderr, flush=True)
            out.setdefault("adjoint", {})["fd_check"] = adjoint_fd_check(device)
            print(f"[bench] adjoint fd_check: {out['adjoint']['fd_check']}", file=sys.stderr, flush=True)
    if args.copy_calibration:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=device)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
