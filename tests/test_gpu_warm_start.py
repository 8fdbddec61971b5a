"""The gfx950 cross-step warm starter against the oracle's, and its effect.

The oracle's warm starter is driven with the GPU run's own step data (start-of-step
state, contacts, lambda), so every step's seed can be compared: matched impulses are
copies (equal to FP32 rounding of the projection), cold starts follow the same formulas.
"""
import numpy as np
import pytest
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.config import WarmStartConfig
from ostrich_amd.containers import Contacts, Control, State
from tests.helpers import config

pytestmark = pytest.mark.gpu

DT = 0.05


def _run(gpu_device, model, state, control, cfg, steps, oracle_ws=None, K=16):
    from ostrich_amd import OstrichEngine
    eng = OstrichEngine(model, config=cfg, device=gpu_device)
    col = collide.PrimitiveCollider(model, device=gpu_device, gap=0.02, max_contacts_per_world=K)
    s_a = State.from_numpy(state, gpu_device)
    s_b = s_a.clone()
    ctl = Control.from_numpy(control, gpu_device)
    nr, seeds = [], []
    for t in range(steps):
        c = col.collide(s_a.body_q)
        con = Contacts.from_dict(c, gpu_device)
        eng.step(s_a, s_b, ctl, con, DT)
        nr.append(eng.stats()["nr_iters"].float().mean().item())
        if oracle_ws is not None:
            cn = {k: (v.cpu().numpy() if torch.is_tensor(v) else v) for k, v in c.items()}
            st = dict(body_q=s_a.body_q.cpu().numpy(), body_qd=s_a.body_qd.cpu().numpy())
            seeds.append((eng.lambda_seed().cpu().numpy(), oracle_ws.apply(st, cn, DT)))
            oracle_ws.snapshot(st, cn, DT, eng.constraint_impulses().cpu().numpy())
        s_a, s_b = s_b, s_a
    return nr, seeds


@pytest.mark.parametrize("method", ["position_match", "pair_aggregate"])
def test_warm_start_seed_matches_oracle(gpu_device, oracle_mod, method):
    """Both warm-start methods (collision/warm_start.py: position_match, pair_aggregate)."""
    model, state, control = scenes.box_stack(16, seed=2)
    cfg = config(K=16, warm_start=WarmStartConfig(seed_iterate=True, method=method))
    ws = oracle_mod.WarmStarter(model, cfg)
    print(f"[warm start] parity, 16 box stacks, {method}", flush=True)
    _, seeds = _run(gpu_device, model, state, control, cfg, 8, oracle_ws=ws)
    for t, (g, o) in enumerate(seeds):
        scale = max(np.abs(o).max(), 1.0)
        assert np.abs(g - o).max() <= 1e-5 * scale, f"step {t}: {np.abs(g - o).max():.3e}"
    assert np.abs(seeds[-1][1]).max() > 0


def test_warm_start_cuts_newton_iterations(gpu_device):
    model, state, control = scenes.box_stack(64, seed=4)
    res = {}
    for seeded in (False, True):
        cfg = config(K=16, warm_start=WarmStartConfig(seed_iterate=seeded))
        print(f"[warm start] seeded={seeded}", flush=True)
        nr, _ = _run(gpu_device, model, state, control, cfg, 12)
        res[seeded] = np.mean(nr[3:])
    assert res[True] < 0.7 * res[False], res
