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
G = 9.81


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


def _one_step_with_and_without(model, state, control, cfg, K, net, device, warm=4, gap=0.02):
    """Two engines with the same warm-start history (`warm` steps); the next step from
    identical inputs without and with the network."""
    from ostrich_amd import OstrichEngine
    col = collide.HipCollider(model, device=device, gap=gap, max_contacts_per_world=K)
    outs = {}
    for use_net in (False, True):
        eng = OstrichEngine(model, config=cfg, device=device)
        s1 = State.from_numpy(state, device)
        s2 = s1.clone()
        ctl = Control.from_numpy(control, device)
        for _ in range(warm):
            eng.step(s1, s2, ctl, col.collide(s1.body_q), DT)
            s1, s2 = s2, s1
        con = col.collide(s1.body_q)
        if use_net:
            eng.set_warm_start_net(net)
        eng.step(s1, s2, ctl, con, DT)
        st = {k: v.cpu().numpy() for k, v in eng.stats().items()}
        outs[use_net] = dict(q=s2.body_q.cpu().numpy(), qd=s2.body_qd.cpu().numpy(), **st)
    return outs[False], outs[True]


@pytest.mark.parametrize("scene", ["skid_steer", "stack_1", "stack_3", "stack_8", "stack_32"])
def test_trained_network_moves_only_the_iteration_count(gpu_device, scene):
    """The shipped network (ostrich_amd/data/warm_start_net.npy, trained on other seeds)
    on uneven skid-steer worlds and on stacks of 1, 3, 8 and 32 cubes, one step from
    identical inputs without and with it.  Worlds that keep the previous-step seed
    reproduce it bit for bit; worlds that took the network's seed and converge are
    solutions of the same system (||h|| <= atol).  On the box stacks a converged world
    lands on the same state with either seed (1e-4 of scale).  On the skid-steer worlds
    it need not: the lagged friction weight makes every |lambda_f| >= mu lambda_n of a
    sticking contact a fixed point, so two seeds can converge to different solutions of
    the residual; the pose differences are reported."""
    import os
    from ostrich_amd.learned import WarmStartNet
    path = os.path.join(os.path.dirname(__file__), "..", "ostrich_amd", "data", "warm_start_net.npy")
    net = WarmStartNet(gpu_device, params=torch.as_tensor(np.load(path)))
    if scene == "skid_steer":
        model, state, control = scenes.skid_steer(512, seed=77)
        K, gap = 128, 0.1
    else:
        n = int(scene.split("_")[1])
        model, state, control = scenes.box_stack(64, n_cubes=n, seed=300 + n)
        K, gap = max(8, 4 * n + 4), 0.02
    cfg = config(K=K, warm_start=WarmStartConfig(seed_iterate=True))
    a, b = _one_step_with_and_without(model, state, control, cfg, K, net, gpu_device, gap=gap)
    took = (b["flags"] & 32) != 0
    atol_sq = cfg.nr.atol ** 2
    conv = (a["merit"] <= atol_sq) & (b["merit"] <= atol_sq)
    W = model.num_worlds
    dq = np.abs(a["q"] - b["q"]).reshape(W, -1).max(1)
    du = np.abs(a["qd"] - b["qd"]).reshape(W, -1).max(1)
    moved = conv & took
    print(f"[learned {scene}] network seed taken in {took.mean():.2f} of worlds; NR mean {a['nr_iters'].mean():.2f} "
          f"-> {b['nr_iters'].mean():.2f} (worlds that took it: {a['nr_iters'][took].mean() if took.any() else 0:.2f} "
          f"-> {b['nr_iters'][took].mean() if took.any() else 0:.2f}); converged in both {conv.mean():.2f}; "
          f"pose difference of those that took it, median {np.median(dq[moved]) if moved.any() else 0:.2e} "
          f"max {dq[moved].max() if moved.any() else 0:.2e}", flush=True)
    kept = ~took
    np.testing.assert_array_equal(a["nr_iters"][kept], b["nr_iters"][kept])
    np.testing.assert_array_equal(a["q"].reshape(W, -1)[kept], b["q"].reshape(W, -1)[kept])
    assert (b["merit"][took & (b["nr_iters"] < cfg.nr.max_iters)] <= atol_sq).all()
    if scene != "skid_steer" and conv.any():
        assert dq[conv].max() <= 1e-4 and du[conv].max() <= 1e-4 * G * DT + 1e-4, (dq[conv].max(), du[conv].max())
