"""Pins the oracle's cross-step contact warm start (collision/warm_start.py, position
match) on hand-built two-step scenes with known answers: the cold-start gravity split,
the impact impulse, the sliding-friction seed, exact copies of matched impulses, the
v*dt prediction of moving contacts, and the effect on Newton iterations of a resting
stack when the iterate is seeded.

The reference's cold-start kernel takes the force on body 0 as -n lambda
(warm_start.py:246-253, 383-441), the convention from before its contact packing
flipped Newton's normals (core/contacts.py:80-83); its contact rows push body 0 along
+n.  So for physical contacts (normal from shape1 toward shape0 after the flip) the
gravity and impact terms never fire, which the oracle restates as is
(`test_cold_start_is_silent_for_flipped_normals`).  The formulas themselves are pinned
on contacts whose normals follow the cold-start kernel's own convention.
"""
import numpy as np
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.config import WarmStartConfig
from tests.helpers import config, single_box, zero_control

DT = 0.05
G = 9.81


def _setup(oracle_mod, vel=(0, 0, 0, 0, 0, 0), z=0.25, mu=0.5):
    model, state = single_box(mass=1.0, mu=mu, pose=(0, 0, z, 0, 0, 0, 1))
    state["body_qd"][0] = vel
    cfg = config(K=8, nr_iters=30, warm_start=WarmStartConfig(seed_iterate=True))
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=8)
    ws = oracle_mod.WarmStarter(model, cfg)
    nc, nj, nctrl = oracle_mod.dims(model, cfg)
    return model, state, cfg, col, ws, nj + nctrl


def _contacts(col, state):
    return col.collide(torch.as_tensor(np.asarray(state["body_q"], np.float32)))


def _reversed(c):
    """The same contacts with Newton normals reversed (the cold-start kernel's convention)."""
    d = {k: (np.asarray(v).copy() if hasattr(v, "__len__") else v) for k, v in c.items()}
    d["normal"] = -d["normal"]
    return d


def test_cold_start_is_silent_for_flipped_normals(oracle_mod):
    """Physical contacts of a box resting / falling on the ground: no gravity, no impact
    seed (see the module docstring)."""
    for vel in ((0, 0, 0, 0, 0, 0), (0, 0, -1.0, 0, 0, 0)):
        model, state, cfg, col, ws, on = _setup(oracle_mod, vel=vel)
        seed = ws.apply(state, _contacts(col, state), DT)
        assert np.abs(seed).max() == 0.0


def test_cold_start_splits_weight_over_supporting_contacts(oracle_mod):
    model, state, cfg, col, ws, on = _setup(oracle_mod)
    c = _reversed(_contacts(col, state))
    n = int(np.asarray(c["count"])[0])
    assert n == 4
    seed = ws.apply(state, c, DT)
    np.testing.assert_allclose(seed[0, on:on + n], G / 4, rtol=1e-6)
    assert np.abs(seed[0, on + 8:]).max() == 0.0  # at rest: no friction seed


def test_cold_start_impact_dominates_when_falling(oracle_mod):
    """Box falling at 1 m/s: v_n = (v1 - v0).n = -1 with the reversed normal, so the seed
    is m_eff (-v_n) / dt with m_eff = 1 kg against the static ground (20 N), above the
    gravity share (2.45 N)."""
    model, state, cfg, col, ws, on = _setup(oracle_mod, vel=(0, 0, -1.0, 0, 0, 0))
    c = _reversed(_contacts(col, state))
    seed = ws.apply(state, c, DT)
    np.testing.assert_allclose(seed[0, on:on + 4], 1.0 / DT, rtol=1e-6)


def test_cold_start_sliding_friction_opposes_motion(oracle_mod):
    model, state, cfg, col, ws, on = _setup(oracle_mod, vel=(0.5, 0.0, 0, 0, 0, 0), mu=0.5)
    c = _reversed(_contacts(col, state))
    seed = ws.apply(state, c, DT)
    ln = seed[0, on:on + 4]
    assert (ln > 0).all()
    ft = seed[0, on + 8:on + 16].reshape(4, 2)
    np.testing.assert_allclose(np.linalg.norm(ft, axis=1), 0.5 * ln, rtol=1e-6)
    # rebuild the world-frame seed from the tangent basis of the packed normal (-Newton)
    n = -np.asarray(c["normal"], np.float64)[0]
    a = np.array([1.0, 0, 0]) if abs(n[0]) <= min(abs(n[1]), abs(n[2])) else (
        np.array([0, 1.0, 0]) if abs(n[1]) <= abs(n[2]) else np.array([0, 0, 1.0]))
    t1 = np.cross(n, a); t1 /= np.linalg.norm(t1); t2 = np.cross(n, t1)
    f = ft[0, 0] * t1 + ft[0, 1] * t2
    # v_rel = v1 - v0 = -v_box (ground is body 1): the seed points along -v_rel = +x
    assert f[0] > 0 and abs(f[1]) < 1e-9


def test_matched_contacts_copy_previous_impulses(oracle_mod):
    model, state, cfg, col, ws, on = _setup(oracle_mod)
    st = {k: np.asarray(v, np.float64) for k, v in state.items()}
    c = _contacts(col, st)
    o = oracle_mod.step(model, cfg, st, zero_control(model), c, DT, precision=64, lam_seed=ws.apply(st, c, DT))
    ws.snapshot(st, c, DT, o["lam"])
    st2 = dict(st, body_q=o["body_q"], body_qd=o["body_qd"])
    c2 = _contacts(col, st2)
    seed = ws.apply(st2, c2, DT)
    np.testing.assert_allclose(seed[0, on:on + 4], o["lam"][0, on:on + 4], rtol=1e-12)
    np.testing.assert_allclose(seed[0, on + 8:on + 16], o["lam"][0, on + 8:on + 16], atol=1e-12)


def test_moving_contacts_match_through_prediction(oracle_mod):
    """Box translating at 1 m/s: contacts move 5 cm per step, ten times the 5 mm floor;
    the v*dt prediction brings them back within it."""
    model, state, cfg, col, ws, on = _setup(oracle_mod, z=0.25)
    st = {k: np.asarray(v, np.float64) for k, v in state.items()}
    c = _contacts(col, st)
    lam = np.zeros((1, oracle_mod.dims(model, cfg)[0]))
    lam[0, on:on + 4] = [1.0, 2.0, 3.0, 4.0]
    ws.snapshot(st, c, DT, lam)
    st2 = dict(st)
    st2["body_q"] = st["body_q"].copy()
    st2["body_q"][0, 0] += 1.0 * DT
    st2["body_qd"] = st["body_qd"].copy()
    st2["body_qd"][0, 0] = 1.0
    c2 = _contacts(col, st2)
    seed = ws.apply(st2, c2, DT)
    np.testing.assert_allclose(seed[0, on:on + 4], [1.0, 2.0, 3.0, 4.0])
    # without the velocity the same displacement is a mismatch (cold start instead)
    st3 = dict(st2, body_qd=st["body_qd"])
    seed3 = ws.apply(st3, c2, DT)
    np.testing.assert_allclose(seed3[0, on:on + 4], 0.0)


def test_seeded_iterate_cuts_newton_iterations_on_resting_stacks(oracle_mod):
    """Three-cube stacks: from the warm start the Newton loop needs fewer iterations than
    from lambda = 0 (the reference keeps seed_iterate off by default)."""
    model, state, control = scenes.box_stack(4, seed=0)
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=16)
    nr = {}
    for seeded in (False, True):
        cfg = config(K=16, warm_start=WarmStartConfig(seed_iterate=seeded))
        ws = oracle_mod.WarmStarter(model, cfg)
        st = {k: np.asarray(v, np.float64) for k, v in state.items()}
        its = []
        for t in range(12):
            c = col.collide(torch.as_tensor(st["body_q"], dtype=torch.float32))
            seed = ws.apply(st, c, DT) if seeded else None
            o = oracle_mod.step(model, cfg, st, control, c, DT, precision=64, lam_seed=seed)
            if seeded:
                ws.snapshot(st, c, DT, o["lam"])
            its.append(o["nr_iters"].mean())
            st = dict(st, body_q=o["body_q"], body_qd=o["body_qd"])
        nr[seeded] = np.mean(its[3:])
    assert nr[True] < 0.7 * nr[False], nr


def _drop_contact(c, i):
    """The flat contact list without entry i (later entries move up one slot)."""
    d = {k: (v.clone() if torch.is_tensor(v) else np.array(v, copy=True)) for k, v in c.items()}
    n = int(d["count"][0])
    for k, v in d.items():
        if k in ("count", "max_contacts"):
            continue
        v[i:n - 1] = v[i + 1:n].clone() if torch.is_tensor(v) else v[i + 1:n].copy()
    d["count"][0] = n - 1
    return d


def test_pair_aggregate_spreads_each_pairs_total(oracle_mod):
    """warm_start.method = pair_aggregate (collision/warm_start.py:896-1010): each body
    pair's previous total normal impulse is spread evenly over the pair's current
    contacts, friction seeds stay zero.  Two stacked cubes: four cube-cube contacts and
    three cube-ground contacts with impulses 1..7 in the snapshot; the same contacts give
    10/4 and 18/3 per contact, and with one cube-cube contact gone the first pair's 10 is
    spread over the three that remain.  The position-match method on the same inputs
    copies the contacts' own impulses instead."""
    model, state, control = scenes.box_stack(1, n_cubes=2, seed=0)
    K = 16
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=K)
    st = {k: np.asarray(v, np.float64) for k, v in state.items()}
    c = col.collide(torch.as_tensor(np.asarray(st["body_q"], np.float32)))
    s0, s1 = np.asarray(c["shape0"][:7]), np.asarray(c["shape1"][:7])
    assert int(c["count"][0]) == 7 and (s0[:4] == s0[0]).all() and (s1[:4] == s1[0]).all()
    assert (s0[4:] == s0[4]).all() and (s1[4:] == s1[4]).all() and (s0[0], s1[0]) != (s0[4], s1[4])
    lam_n = np.arange(1.0, 8.0)
    res = {}
    for method in ("pair_aggregate", "position_match"):
        cfg = config(K=K, warm_start=WarmStartConfig(seed_iterate=True, method=method))
        ws = oracle_mod.WarmStarter(model, cfg)
        nc, nj, nctrl = oracle_mod.dims(model, cfg)
        on, of = nj + nctrl, nj + nctrl + K
        lam = np.zeros((1, nc))
        lam[0, on:on + 7] = lam_n
        lam[0, of:of + 14] = 0.3  # friction impulses: not carried by pair_aggregate
        ws.snapshot(st, c, DT, lam)
        res[method] = (ws.apply(st, c, DT), ws.apply(st, _drop_contact(c, 3), DT), on, of)
    seed, seed_drop, on, of = res["pair_aggregate"]
    np.testing.assert_allclose(seed[0, on:on + 7], [2.5] * 4 + [6.0] * 3, rtol=1e-12)
    np.testing.assert_allclose(seed[0, of:of + 14], 0.0)
    np.testing.assert_allclose(seed_drop[0, on:on + 6], [10.0 / 3] * 3 + [6.0] * 3, rtol=1e-12)
    np.testing.assert_allclose(seed_drop[0, on + 6:on + K], 0.0)
    pm, _, on, of = res["position_match"]
    np.testing.assert_allclose(pm[0, on:on + 7], lam_n, rtol=1e-12)
