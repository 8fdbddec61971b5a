"""Trains and evaluates the learned per-contact warm start on the GPU.

    python -m ostrich_amd.train_warm_start --out ostrich_amd/data/warm_start_net.npy

Data: (feature, converged impulse) pairs logged by the solver itself from uneven-batch
skid-steer worlds (4 bodies among 0-24 obstacles, contacts within 0.1 m) and tracked
flipper robots on the bump heightfield, with the cross-step warm starter seeding Newton
(so the "previous impulse" features are what the solver sees at run time); only contacts
of worlds whose Newton loop converged (||h|| <= atol) are kept, and rollouts of further
seeds are added until `--pairs` pairs are logged.  Training: Adam by the library's MFMA
forward/backward kernels.  Evaluation (seeds disjoint from the training ones): 4096
skid-steer worlds and 2048 box stacks of 1-32 cubes (64 per count), mean Newton
iterations per step with the previous-step warm start and with the learned one (per
world the lower initial residual picks the seed), and, on each step of the learned run,
the largest difference between the two seeds' converged states from identical inputs.
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


SKID_GAP = 0.1


def _scene(kind, worlds, seed, n_cubes=3):
    """(model, state, control, contact slots, contact gap) of a scene kind."""
    if kind == "skid":
        model, state, control = scenes.skid_steer(worlds, seed=seed)
        return model, state, control, 128, SKID_GAP
    if kind == "flipper":
        model, state, control = scenes.tracked_flipper(worlds, seed=seed)
        return model, state, control, 64, 0.02
    model, state, control = scenes.box_stack(worlds, n_cubes=n_cubes, seed=seed)
    return model, state, control, max(8, 4 * n_cubes + 4), 0.02


def rollout(kind, worlds, seed, steps, device, n_cubes=3, net=None, seeded=True, log_data=False, check=False):
    """Runs `steps` steps; returns per-step mean Newton iterations and, with log_data, the
    (features, targets) of converged worlds' contacts.  With `check` (and a network) each
    step is also solved from the same inputs by a second engine that holds the same
    warm-start snapshot but no network: `conv_diff` is the largest pose difference of
    the worlds both converge, `nr_base` that engine's Newton counts."""
    model, state, control, K, gap = _scene(kind, worlds, seed, n_cubes)
    cfg = solver_config(K, seeded=seeded)
    eng = OstrichEngine(model, config=cfg, device=device)
    ref = OstrichEngine(model, config=cfg, device=device) if (check and net is not None) else None
    if net is not None:
        eng.set_warm_start_net(net)
    col = collide.HipCollider(model, device=device, gap=gap, max_contacts_per_world=K)
    s_a = State.from_numpy(state, device)
    s_b = s_a.clone()
    s_r = s_a.clone()
    ctl = Control.from_numpy(control, device)
    nr, nr_base, xs, ts, learned, conv_diff = [], [], [], [], [], 0.0
    atol_sq = cfg.nr.atol ** 2
    for _ in range(steps):
        con = col.collide(s_a.body_q)
        if ref is not None:
            ref.set_warm_start_state(eng.warm_start_state())
        eng.step(s_a, s_b, ctl, con, DT)
        st = eng.stats()
        nr.append(st["nr_iters"].float().mean().item())
        learned.append(((st["flags"] & 32) != 0).float().mean().item())
        if ref is not None:
            ref.step(s_a, s_r, ctl, con, DT)
            sr = ref.stats()
            nr_base.append(sr["nr_iters"].float().mean().item())
            both = (st["merit"] <= atol_sq) & (sr["merit"] <= atol_sq)
            d = (s_b.body_q - s_r.body_q).abs().reshape(model.num_worlds, -1).amax(1)
            if both.any():
                conv_diff = max(conv_diff, d[both].max().item())
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
    return dict(nr=nr, nr_base=nr_base, learned_fraction=learned, conv_diff=conv_diff, data=data)


def collect(device, pairs=2_000_000, log=print, max_rounds=8):
    """Rollouts of training seeds (2000+) of the skid-steer and tracked-flipper scenes,
    alternating, until `pairs` pairs are logged (or max_rounds rounds of both)."""
    xs, ts, total = [], [], 0
    for r in range(max_rounds):
        for kind, W, steps in (("skid", 4096, 60), ("flipper", 1024, 60)):
            seed = 2000 + 10 * r + (0 if kind == "skid" else 1)
            out = rollout(kind, W, seed, steps, device, log_data=True)
            x, t = out["data"]
            total += x.shape[0]
            log(f"[collect] {kind} seed {seed} W={W}: {x.shape[0]} pairs (total {total}), "
                f"NR mean {np.mean(out['nr']):.2f}")
            xs.append(x)
            ts.append(t)
        if total >= pairs:
            break
    X, T = torch.cat(xs), torch.cat(ts)
    return X[:pairs], T[:pairs]


def evaluate(net, device, log=print, skid_worlds=4096, stack_worlds=64, max_cubes=32):
    """Per scene: mean Newton iterations with the previous-step seed (a run without the
    network) and with the network, the fraction of worlds that took the network's seed,
    and the largest converged-state difference of the per-step check; plus the totals
    over all world-steps."""
    rows = []
    plan = [("skid", skid_worlds, 1, 40, 3)] + [("stack", stack_worlds, 7 + n, 30, n)
                                                 for n in range(1, max_cubes + 1)]
    for kind, W, seed, steps, n in plan:
        base = rollout(kind, W, seed, steps, device, n_cubes=n)
        lrn = rollout(kind, W, seed, steps, device, n_cubes=n, net=net, check=True)
        row = dict(scene=kind, cubes=n if kind == "stack" else None, worlds=W, steps=steps,
                   nr_prev_step=float(np.mean(base["nr"])), nr_learned=float(np.mean(lrn["nr"])),
                   nr_same_inputs_without_net=float(np.mean(lrn["nr_base"])),
                   learned_seed_fraction=float(np.mean(lrn["learned_fraction"])),
                   max_converged_pose_diff=lrn["conv_diff"])
        log(f"[eval] {row}")
        rows.append(row)
    def total(pred):
        sel = [r for r in rows if pred(r)]
        ws = np.array([r["worlds"] * r["steps"] for r in sel], float)
        prev = float(np.sum(ws * [r["nr_prev_step"] for r in sel]) / ws.sum())
        lrn = float(np.sum(ws * [r["nr_learned"] for r in sel]) / ws.sum())
        return dict(world_steps=int(ws.sum()), nr_prev_step=prev, nr_learned=lrn, change=lrn / prev - 1.0)
    summary = dict(all=total(lambda r: True), skid=total(lambda r: r["scene"] == "skid"),
                   stacks=total(lambda r: r["scene"] == "stack"),
                   max_converged_pose_diff=max(r["max_converged_pose_diff"] for r in rows))
    log(f"[eval] summary {summary}")
    return rows, summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--pairs", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--report", default=None)
    ap.add_argument("--eval-only", default=None, help="evaluate the network stored in this .npy file")
    args = ap.parse_args()
    device = "cuda:0"
    log = lambda m: print(m, flush=True)
    report = {}
    if args.eval_only:
        net = WarmStartNet(device, params=torch.as_tensor(np.load(args.eval_only)))
    else:
        t0 = time.time()
        X, T = collect(device, args.pairs, log)
        log(f"[train] {X.shape[0]} pairs ({time.time() - t0:.1f} s)")
        net = WarmStartNet(device, seed=0)
        t1 = time.time()
        hist = net.fit(X, T, epochs=args.epochs, batch=4096, lr=2e-3, log=log)
        torch.cuda.synchronize()
        log(f"[train] {args.epochs} epochs in {time.time() - t1:.1f} s, loss {hist[0]:.4e} -> {hist[-1]:.4e}")
        report.update(pairs=int(X.shape[0]), loss=hist, train_s=time.time() - t1)
        if args.out:
            os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
            np.save(args.out, net.params().cpu().numpy())
    rows, summary = evaluate(net, device, log)
    report.update(eval=rows, summary=summary)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
