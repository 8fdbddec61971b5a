"""Trains and evaluates the learned per-contact warm start on the GPU.

    python -m ostrich_amd.train_warm_start --out ostrich_amd/data/warm_start_net.npy

Data: (feature, converged impulse) pairs logged by the solver itself from skid-steer
worlds (4 bodies, 0-24 obstacles) and box stacks of 1-8 cubes, with the cross-step warm
starter seeding Newton (so the "previous impulse" features are what the solver sees at
run time); only contacts of worlds whose Newton loop converged (||h|| <= atol) are kept.
Training: Adam by the library's MFMA forward/backward kernels.  Evaluation (disjoint
seeds): mean Newton iterations per step with the previous-step warm start and with the
learned one (per world the lower initial residual picks the seed), and the largest
difference of the converged states.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from . import collide, scenes
from .config import (ContactsConfig, ExecutionConfig, LinearSolverConfig, NewtonRaphsonConfig, OstrichEngineConfig,
                     WarmStartConfig)
from .containers import Control, State
from .engine import OstrichEngine
from .learned import WarmStartNet

DT = 0.05


def solver_config(K, seeded=True, nr_iters=20):
    return OstrichEngineConfig(nr=NewtonRaphsonConfig(max_iters=nr_iters, atol=1e-5),
                               linear=LinearSolverConfig(max_iters=16),
                               contacts=ContactsConfig(max_per_world=K),
                               warm_start=WarmStartConfig(enabled=True, seed_iterate=seeded),
                               execution=ExecutionConfig(path="fused"))


def _scene(kind, worlds, seed, n_cubes=3):
    if kind == "skid":
        model, state, control = scenes.skid_steer(worlds, seed=seed)
        return model, state, control, 128
    model, state, control = scenes.box_stack(worlds, n_cubes=n_cubes, seed=seed)
    return model, state, control, max(8, 8 * n_cubes)


def rollout(kind, worlds, seed, steps, device, n_cubes=3, net=None, seeded=True, log_data=False):
    """Runs `steps` steps; returns per-step mean Newton iterations, the final state and,
    with log_data, the (features, targets) of converged worlds' contacts."""
    model, state, control, K = _scene(kind, worlds, seed, n_cubes)
    cfg = solver_config(K, seeded=seeded)
    eng = OstrichEngine(model, config=cfg, device=device)
    if net is not None:
        eng.set_warm_start_net(net)
    col = collide.HipCollider(model, device=device, gap=0.02, max_contacts_per_world=K)
    s_a = State.from_numpy(state, device)
    s_b = s_a.clone()
    ctl = Control.from_numpy(control, device)
    nr, xs, ts, learned = [], [], [], []
    atol_sq = cfg.nr.atol ** 2
    for _ in range(steps):
        con = col.collide(s_a.body_q)
        eng.step(s_a, s_b, ctl, con, DT)
        st = eng.stats()
        nr.append(st["nr_iters"].float().mean().item())
        learned.append(((st["flags"] & 32) != 0).float().mean().item())
        if log_data:
            f, sc = eng.contact_features()
            lam = eng.constraint_impulses()
            on, of = eng.offset_n, eng.offset_f
            W = model.num_worlds
            tn = lam[:, on:on + K]
            tt = lam[:, of:of + 2 * K].reshape(W, K, 2)
            tgt = torch.cat([tn[:, :, None], tt], 2) / sc.clamp_min(1e-12)[:, :, None]
            mask = (sc > 0) & (st["merit"] <= atol_sq)[:, None]
            xs.append(f[mask])
            ts.append(tgt[mask])
        s_a, s_b = s_b, s_a
    data = (torch.cat(xs), torch.cat(ts)) if log_data else None
    return dict(nr=nr, learned_fraction=learned, final_q=s_a.body_q.clone(), data=data)


def collect(device, scale=1.0, log=print):
    xs, ts = [], []
    plan = [("skid", int(4096 * scale), 1001, 60, 3)] + [
        ("stack", int(512 * scale), 1100 + n, 40, n) for n in (1, 2, 3, 4, 6, 8)]
    for kind, W, seed, steps, n in plan:
        r = rollout(kind, W, seed, steps, device, n_cubes=n, log_data=True)
        x, t = r["data"]
        log(f"[collect] {kind} n={n} W={W}: {x.shape[0]} pairs, NR mean {np.mean(r['nr']):.2f}")
        xs.append(x)
        ts.append(t)
    return torch.cat(xs), torch.cat(ts)


def evaluate(net, device, scale=1.0, log=print):
    rows = []
    plan = [("skid", int(4096 * scale), 1, 40, 3)] + [("stack", int(256 * scale), 7 + n, 30, n)
                                                      for n in (1, 2, 3, 4, 6, 8)]
    for kind, W, seed, steps, n in plan:
        base = rollout(kind, W, seed, steps, device, n_cubes=n)
        lrn = rollout(kind, W, seed, steps, device, n_cubes=n, net=net)
        row = dict(scene=kind, cubes=n if kind == "stack" else None, worlds=W, steps=steps,
                   nr_prev_step=float(np.mean(base["nr"])), nr_learned=float(np.mean(lrn["nr"])),
                   learned_seed_fraction=float(np.mean(lrn["learned_fraction"])),
                   max_state_diff=float((base["final_q"] - lrn["final_q"]).abs().max().item()))
        log(f"[eval] {row}")
        rows.append(row)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    device = "cuda:0"
    log = lambda m: print(m, flush=True)
    t0 = time.time()
    X, T = collect(device, args.scale, log)
    log(f"[train] {X.shape[0]} pairs ({time.time() - t0:.1f} s)")
    net = WarmStartNet(device, seed=0)
    t1 = time.time()
    hist = net.fit(X, T, epochs=args.epochs, batch=4096, lr=2e-3, log=log)
    torch.cuda.synchronize()
    log(f"[train] {args.epochs} epochs in {time.time() - t1:.1f} s, loss {hist[0]:.4e} -> {hist[-1]:.4e}")
    rows = evaluate(net, device, args.scale, log)
    if args.out:
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        np.save(args.out, net.params().cpu().numpy())
    if args.report:
        with open(args.report, "w") as f:
            json.dump(dict(pairs=int(X.shape[0]), loss=hist, eval=rows), f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
