"""Pins the oracle's merit line search (core/linesearch.py, core/engine_data.py:15-39):
the candidate set equals the reference's numpy construction, and with the line search
one Newton iteration never ends above the full Newton step (alpha = 1 is a candidate
and the argmin keeps the smallest merit)."""
import numpy as np
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.config import LinesearchConfig
from tests.helpers import config

DT = 0.05


def _reference_steps(ls):
    cons = np.logspace(np.log10(ls.min_step), np.log10(ls.conservative_upper_bound), ls.conservative_step_count)
    half = ls.optimistic_window / 2.0
    opt = np.linspace(1.0 - half, 1.0 + half, ls.optimistic_step_count)
    s = np.concatenate([cons, opt])
    s.sort()
    s[np.argmin(np.abs(s - 1.0))] = 1.0
    return s.astype(np.float32)


def test_candidate_steps_match_reference_construction(oracle_mod):
    for ls in (LinesearchConfig(enabled=True), LinesearchConfig(enabled=True, conservative_step_count=5,
                                                                 optimistic_step_count=3, optimistic_window=0.5)):
        got = oracle_mod.linesearch_steps(config(linesearch=ls))
        np.testing.assert_array_equal(got, _reference_steps(ls))
        assert 1.0 in got


def test_linesearch_never_worse_than_full_step(oracle_mod):
    for make in (lambda: scenes.box_stack(6, seed=3), lambda: scenes.skid_steer(6, seed=4)):
        model, state, control = make()
        col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=32)
        c = col.collide(torch.as_tensor(np.asarray(state["body_q"], np.float32)))
        full = oracle_mod.step(model, config(K=32, nr_iters=3, atol=0.0), state, control, c, DT, precision=64)
        ls = oracle_mod.step(model, config(K=32, nr_iters=3, atol=0.0, linesearch=LinesearchConfig(enabled=True)),
                             state, control, c, DT, precision=64)
        # the first Newton iteration starts from the same point in both runs
        assert (ls["merit_hist"][:, 0] <= full["merit_hist"][:, 0] * (1 + 1e-12)).all()
        assert np.isfinite(ls["merit"]).all()
