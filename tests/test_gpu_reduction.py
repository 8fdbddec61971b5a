"""Per-pair contact reduction on the GPU (fb_reduce.hip) against the oracle: the packed
contact set after top_k / fps / cluster must equal the FP64 oracle's bit for bit (the
reducer decides in FP64 on both sides, so the near-ties of flat contact patches fall the
same way; the survivors are copies of input contacts, in packed order), and a full step
with the reduction on must match the oracle's step."""
import numpy as np
import pytest
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.config import ContactReductionConfig, ContactsConfig, OstrichEngineConfig
from ostrich_amd.containers import Contacts, Control, State
from tests.helpers import config

pytestmark = pytest.mark.gpu


def _tilted_stacks(worlds, seed):
    """3-cube stacks with roll/pitch of up to 0.06 rad, so the corner contacts of a face
    sit at depths that differ by far more than rounding (no ties for the ranking)."""
    model, state, control = scenes.box_stack(worlds, 3, seed=seed)
    rng = np.random.default_rng(seed)
    q = state["body_q"].reshape(-1, 7).astype(np.float64)
    for i in range(q.shape[0]):
        r, p = rng.uniform(-0.06, 0.06, size=2)
        qa = np.array([np.sin(r / 2), 0, 0, np.cos(r / 2)])
        qb = np.array([0, np.sin(p / 2), 0, np.cos(p / 2)])
        x1, y1, z1, w1 = q[i, 3:]
        for qq in (qa, qb):
            x2, y2, z2, w2 = qq
            x1, y1, z1, w1 = (w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                              w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2)
        q[i, 3:] = (x1, y1, z1, w1)
        q[i, 2] -= 0.01 * (i % 3 + 1)  # every level penetrates a little
    state["body_q"] = q.astype(np.float32)
    return model, state, control


def _cfg(policy, k):
    base = config(K=16)
    return OstrichEngineConfig(nr=base.nr, linear=base.linear, execution=base.execution,
                               contacts=ContactsConfig(max_per_world=16, reduction=ContactReductionConfig(
                                   policy=policy, max_per_pair=k)))


@pytest.mark.parametrize("policy,k", [("top_k", 2), ("top_k", 1), ("fps", 3), ("fps", 2), ("cluster", 2),
                                      ("cluster", 3)])
def test_reduced_contacts_match_oracle_bit_exact(oracle_mod, gpu_device, policy, k):
    from ostrich_amd import OstrichEngine
    model, state, control = _tilted_stacks(24, seed=11)
    col = collide.PrimitiveCollider(model, gap=0.05, max_contacts_per_world=16)
    c = col.collide(torch.as_tensor(state["body_q"]))
    c = {kk: (v.numpy() if hasattr(v, "numpy") else v) for kk, v in c.items()}
    cfg = _cfg(policy, k)
    eng = OstrichEngine(model, config=cfg, device=gpu_device)
    s_in = State.from_numpy(state, gpu_device)
    s_out = s_in.clone()
    eng.step(s_in, s_out, Control.from_numpy(control, gpu_device), Contacts.from_dict(c, gpu_device), 0.05)
    g = {kk: v.cpu().numpy() for kk, v in eng.packed_contacts().items()}
    o = oracle_mod.pack(model, cfg, state, c, precision=64)
    full = oracle_mod.pack(model, config(K=16), state, c, precision=64)["count"]
    print(f"[reduce] {policy} K={k}: contacts {full.sum()} -> {o['count'].sum()}", flush=True)
    assert (o["count"] < full).any()
    np.testing.assert_array_equal(g["count"], o["count"])
    W, K = model.num_worlds, 16
    for w in range(W):
        n = int(o["count"][w])
        sl = slice(w * K, w * K + n)
        np.testing.assert_array_equal(g["shape0"][sl], o["shape0"][sl])
        np.testing.assert_array_equal(g["shape1"][sl], o["shape1"][sl])
        for key in ("point0", "point1", "normal"):
            np.testing.assert_array_equal(g[key][3 * w * K: 3 * (w * K + n)],
                                          o[key][3 * w * K: 3 * (w * K + n)].astype(np.float32))


def test_step_with_reduction_matches_oracle(oracle_mod, gpu_device):
    from ostrich_amd import OstrichEngine
    model, state, control = _tilted_stacks(16, seed=12)
    col = collide.PrimitiveCollider(model, gap=0.05, max_contacts_per_world=16)
    c = col.collide(torch.as_tensor(state["body_q"]))
    c = {kk: (v.numpy() if hasattr(v, "numpy") else v) for kk, v in c.items()}
    cfg = _cfg("top_k", 3)
    eng = OstrichEngine(model, config=cfg, device=gpu_device)
    s_in = State.from_numpy(state, gpu_device)
    s_out = s_in.clone()
    eng.step(s_in, s_out, Control.from_numpy(control, gpu_device), Contacts.from_dict(c, gpu_device), 0.05)
    o32 = oracle_mod.step(model, cfg, state, control, c, 0.05, precision=32)
    o64 = oracle_mod.step(model, cfg, state, control, c, 0.05, precision=64)
    W = model.num_worlds
    for key, scale in (("body_q", 1.0), ("body_qd", 9.81 * 0.05)):
        g = getattr(s_out, key).cpu().numpy().astype(np.float64)
        eg = np.abs(g - o32[key]).reshape(W, -1).max(1)
        ep = np.abs(o32[key] - o64[key]).reshape(W, -1).max(1)
        assert (eg <= 4.0 * ep + 1e-4 * scale).all(), (key, eg.max(), ep.max())
    np.testing.assert_array_equal(eng.stats()["nr_iters"].cpu().numpy() >= 1, True)
