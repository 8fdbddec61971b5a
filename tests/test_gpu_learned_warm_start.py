"""The learned warm start inside the solver.

A network whose outputs are zero proposes lambda = 0 for every contact; the solver must
take it only for worlds where it starts from a lower residual than the previous-step seed
(flag bit 5), and a world that converges from either seed must land on the same state
(to the Newton tolerance) - the network moves only the iteration count."""
import numpy as np
import pytest
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.config import WarmStartConfig
from ostrich_amd.containers import Contacts, Control, State
from tests.helpers import config

pytestmark = pytest.mark.gpu

DT = 0.05


def test_gate_and_converged_results(gpu_device):
    from ostrich_amd import OstrichEngine
    from ostrich_amd.learned import WarmStartNet
    model, state, control = scenes.box_stack(64, seed=9)
    cfg = config(K=16, warm_start=WarmStartConfig(seed_iterate=True))
    col = collide.PrimitiveCollider(model, device=gpu_device, gap=0.02, max_contacts_per_world=16)
    eng = OstrichEngine(model, config=cfg, device=gpu_device)
    s_a = State.from_numpy(state, gpu_device)
    s_b = s_a.clone()
    ctl = Control.from_numpy(control, gpu_device)
    for _ in range(4):  # build a warm-start history
        eng.step(s_a, s_b, ctl, Contacts.from_dict(col.collide(s_a.body_q), gpu_device), DT)
        s_a, s_b = s_b, s_a
    con = Contacts.from_dict(col.collide(s_a.body_q), gpu_device)
    # two engines with the same history: replay the snapshot by re-running the last step
    net = WarmStartNet(gpu_device, params=torch.zeros(WarmStartNet(gpu_device).num_params))
    outs = {}
    for use_net in (False, True):
        e2 = OstrichEngine(model, config=cfg, device=gpu_device)
        s1 = State.from_numpy(state, gpu_device)
        s2 = s1.clone()
        for _ in range(4):
            e2.step(s1, s2, ctl, Contacts.from_dict(col.collide(s1.body_q), gpu_device), DT)
            s1, s2 = s2, s1
        if use_net:
            e2.set_warm_start_net(net)
        print(f"[learned] use_net={use_net}", flush=True)
        e2.step(s1, s2, ctl, con, DT)
        st = e2.stats()
        outs[use_net] = dict(q=s2.body_q.cpu().numpy(), merit=st["merit"].cpu().numpy(),
                             flags=st["flags"].cpu().numpy(), nr=st["nr_iters"].cpu().numpy())
    a, b = outs[False], outs[True]
    assert (a["flags"] & 32).sum() == 0
    conv = (a["merit"] <= 1e-10) & (b["merit"] <= 1e-10)
    assert conv.sum() > 10
    B = model.bodies_per_world
    dq = np.abs(a["q"] - b["q"]).reshape(-1, B * 7).max(1)
    assert dq[conv].max() <= 1e-4, dq[conv].max()
    # worlds that kept the previous-step seed reproduce it exactly
    kept = (b["flags"] & 32) == 0
    np.testing.assert_array_equal(a["nr"][kept], b["nr"][kept])
