"""Parity of the plain per-kernel path with the oracle, in both termination modes."""
import numpy as np
import pytest
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.config import NewtonRaphsonConfig
from tests.helpers import config
from tests.test_gpu_fused import G, _engine, _gpu_step

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("termination", ["global", "per_world"])
def test_plain_box_stack_matches_oracle(oracle_mod, gpu_device, termination):
    """Resting stacks (sticking friction, unique solution): strict agreement, and in the
    global mode every world runs the same number of Newton iterations as the oracle's
    batch-global loop (base_engine.py:42-80)."""
    model, state, control = scenes.box_stack(16, 3, seed=0)
    cfg = config(K=16, termination=termination, path="plain")
    eng = _engine(model, cfg, gpu_device)
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=16)
    st = {k: np.asarray(v, np.float32) for k, v in state.items()}
    for _ in range(3):
        c = col.collide(torch.as_tensor(st["body_q"], dtype=torch.float32))
        c = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}
        g = _gpu_step(eng, st, control, c, 0.05, gpu_device)
        o = oracle_mod.step(model, cfg, st, control, c, 0.05, precision=32)
        np.testing.assert_allclose(g["body_q"], o["body_q"], atol=1e-4)
        np.testing.assert_allclose(g["body_qd"], o["body_qd"], atol=1e-4 * G * 0.05)
        if termination == "global":
            assert len(set(g["nr_iters"].tolist())) == 1
            assert abs(int(g["nr_iters"][0]) - int(o["nr_iters"][0])) <= 1
        else:
            assert np.mean(np.abs(g["nr_iters"] - o["nr_iters"]) <= 1) >= 0.9
        st = {k: np.asarray(v, np.float32) for k, v in dict(st, body_q=o["body_q"], body_qd=o["body_qd"]).items()}


@pytest.mark.parametrize("path", ["plain", "fused"])
def test_fixed_iterations_plain_and_fused_agree(oracle_mod, gpu_device, path):
    """Two Newton iterations on skid-steer worlds: both GPU paths match the FP32 oracle
    within 1e-4 of scale."""
    model, state, control = scenes.skid_steer(64, seed=3, layout="clear")
    base = config(K=128, path=path)
    cfg = type(base)(nr=NewtonRaphsonConfig(max_iters=2, backtrack_min_iter=0, atol=1e-5), linear=base.linear,
                     contacts=base.contacts, execution=base.execution)
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=128)
    st = {k: np.asarray(v, np.float32).copy() for k, v in state.items()}
    st["body_qd"] = np.random.default_rng(5).normal(scale=0.3, size=st["body_qd"].shape).astype(np.float32)
    c = col.collide(torch.as_tensor(st["body_q"], dtype=torch.float32))
    c = {k: (v.numpy() if hasattr(v, "numpy") else v) for k, v in c.items()}
    g = _gpu_step(_engine(model, cfg, gpu_device), st, control, c, 0.05, gpu_device)
    o = oracle_mod.step(model, cfg, st, control, c, 0.05, precision=32)
    np.testing.assert_allclose(g["body_q"], o["body_q"], atol=1e-4)
    np.testing.assert_allclose(g["body_qd"], o["body_qd"], atol=1e-4 * G * 0.05)
    lscale = np.maximum(np.abs(o["lam"]).max(1, keepdims=True), 1.0)
    assert np.all(np.abs(g["lam"] - o["lam"]) <= 1e-4 * lscale)
