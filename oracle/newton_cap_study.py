"""Why worlds stop at the Newton cap: a CPU study on the FP64 oracle (test infrastructure).

Runs the oracle in FP64 on the first worlds of the uneven skid-steer scene (128 worlds, 40
closed-loop steps) and of the tracked-flipper scene (32 worlds, 30 steps), with the
cross-step warm start seeding Newton, under solver settings that separate the linear
solve's truncation from the Newton iteration itself:

  bench        NR <= 20 to ||h|| <= 1e-5, PCR <= 16 at tol 1e-5 (the headline settings)
  pcr200       the same with PCR <= 200 at tol 1e-8 (tests/differentiable_simulator/
               test_wheeled_robot.py:46-53 in the reference): near-exact linear solves
  pcr64        PCR <= 64 at tol 1e-6
  pcr200_nr100 PCR <= 200 at 1e-8 and NR <= 100: does Newton converge given time?
  refyaml      examples/conf/engine/ostrich.yaml:4-27 (NR <= 16, backtrack_min_iter 4,
               PCR <= 26 at 1e-3, compliances 6e-10 / 1e-10 / 1e-8)

Prints one markdown table row per (scene, setting): Newton and PCR iterations, converged
fraction (||h|| <= 1e-5 at the picked iterate), the fraction whose picked ||h|| is not
finite (the Newton iteration diverged) and ||h|| of the other capped world-steps.

    python -m oracle.newton_cap_study [--threads 8]
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def settings():
    from ostrich_amd.config import ComplianceConfig, LinearSolverConfig, NewtonRaphsonConfig
    return {
        "bench": (NewtonRaphsonConfig(max_iters=20, atol=1e-5), LinearSolverConfig(max_iters=16), None),
        "pcr200": (NewtonRaphsonConfig(max_iters=20, atol=1e-5),
                   LinearSolverConfig(max_iters=200, tol=1e-8, atol=1e-8), None),
        "pcr64": (NewtonRaphsonConfig(max_iters=20, atol=1e-5),
                  LinearSolverConfig(max_iters=64, tol=1e-6, atol=1e-6), None),
        "pcr200_nr100": (NewtonRaphsonConfig(max_iters=100, atol=1e-5),
                         LinearSolverConfig(max_iters=200, tol=1e-8, atol=1e-8), None),
        "refyaml": (NewtonRaphsonConfig(max_iters=16, atol=1e-5, backtrack_min_iter=4),
                    LinearSolverConfig(max_iters=26, tol=1e-3, atol=1e-3),
                    ComplianceConfig(joint=6e-10, contact=1e-10, friction=1e-8)),
    }


def run(scene, worlds, steps, name, skip=3):
    import torch
    from ostrich_amd import collide, scenes
    from ostrich_amd.config import ContactsConfig, OstrichEngineConfig, WarmStartConfig
    from oracle import oracle
    if scene == "skid_steer":
        model, state, control = scenes.skid_steer(worlds, seed=1)
        gap, K, hf = 0.1, 128, "point"
    else:
        model, state, control = scenes.tracked_flipper(worlds, seed=2)
        gap, K, hf = 0.05, 256, "cells"
    nr_cfg, lin_cfg, comp = settings()[name]
    kw = dict(nr=nr_cfg, linear=lin_cfg, contacts=ContactsConfig(max_per_world=K),
              warm_start=WarmStartConfig(enabled=True, seed_iterate=True))
    if comp is not None:
        kw["compliance"] = comp
    cfg = OstrichEngineConfig(**kw)
    col = collide.PrimitiveCollider(model, device="cpu", gap=gap, max_contacts_per_world=K, hf_contacts=hf)
    hm = oracle.HostModel(model)
    ws = oracle.WarmStarter(model, cfg)
    st = {k: np.asarray(v, np.float64) for k, v in state.items()}
    nr, pcr, merit = [], [], []
    t0 = time.time()
    for s in range(steps):
        c = col.collide(torch.as_tensor(st["body_q"], dtype=torch.float32))
        hc = oracle.HostContacts(c)
        seed = ws.apply(st, hc, 0.05)
        o = oracle.step(model, cfg, st, control, hc, 0.05, precision=64, lam_seed=seed, host_model=hm)
        ws.snapshot(st, hc, 0.05, o["lam"])
        st["body_q"], st["body_qd"] = o["body_q"], o["body_qd"]
        if s >= skip:  # the first steps settle the robots onto their contacts
            nr.append(o["nr_iters"])
            pcr.append(o["pcr_iters"])
            merit.append(o["merit"])
    nr, pcr, merit = (np.concatenate(a) for a in (nr, pcr, merit))
    conv = merit <= 1e-10
    nonfinite = ~np.isfinite(merit)
    capped = ~conv & ~nonfinite
    h = np.sqrt(merit[capped]) if capped.any() else np.array([np.nan])
    return dict(scene=scene, setting=name, nr=nr.mean(), nr_p95=np.percentile(nr, 95), pcr_per_nr=pcr.sum() / nr.sum(),
                converged=conv.mean(), nonfinite=nonfinite.mean(), h_med=np.median(h), h_p90=np.percentile(h, 90),
                secs=time.time() - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    from oracle import oracle
    oracle.set_threads(args.threads)
    print("| scene | setting | NR/step | NR p95 | PCR/NR | converged | non-finite ||h|| | capped ||h|| median | "
          "capped ||h|| p90 |")
    print("|---|---|---|---|---|---|---|---|---|")
    plan = [("skid_steer", 128, 40, n) for n in settings()] + \
           [("tracked_flipper", 32, 30, n) for n in ("bench", "pcr200", "pcr64", "refyaml")]
    for scene, W, steps, name in plan:
        r = run(scene, W, steps, name)
        print(f"| {scene} ({W} worlds, {steps} steps) | {name} | {r['nr']:.2f} | {r['nr_p95']:.0f} | "
              f"{r['pcr_per_nr']:.1f} | {r['converged']:.3f} | {r['nonfinite']:.3f} | {r['h_med']:.1e} | "
              f"{r['h_p90']:.1e} |", flush=True)


if __name__ == "__main__":
    main()
