"""Contact reduction (collision/top_k.py, fps.py, cluster.py) in the oracle, pinned by
contact sets whose depths and positions are chosen by construction: body poses are the
identity, each contact's body point sits at depth d below the ground point along the
normal, so the reducer's depth is d and its midpoint is (x, y, -d/2).
"""
import numpy as np
import pytest

from oracle import oracle
from ostrich_amd.config import ContactReductionConfig, ContactsConfig, OstrichEngineConfig
from ostrich_amd.model import GEO_BOX, GEO_PLANE, IDENTITY7, Model, Shape, WorldBuilder, box_inertia

K_SLOTS = 16


def _world(n_bodies=3):
    wb = WorldBuilder()
    for _ in range(n_bodies):
        b = wb.add_body(1.0, box_inertia(1.0, 0.1, 0.1, 0.1))
        wb.add_shape(Shape(GEO_BOX, b, IDENTITY7, (0.1, 0.1, 0.1)))
        wb.add_joint_free(b)
    model = Model.from_worlds([wb], [Shape(GEO_PLANE, -1, IDENTITY7, (0, 0, 0))])
    q = np.tile(np.array([0, 0, 0, 0, 0, 0, 1.0]), (n_bodies, 1))
    return model, dict(body_q=q)


def _contacts(items):
    """items: (body, x, y, depth); shape0 = the body's box (flat id 1 + body), shape1 =
    the ground (flat id 0), Newton normal from the body down to the ground."""
    n = len(items)
    c = dict(count=np.array([n], np.int32), max_contacts=n, point0=np.zeros((n, 3), np.float32),
             point1=np.zeros((n, 3), np.float32), normal=np.tile(np.float32([0, 0, -1]), (n, 1)),
             shape0=np.zeros(n, np.int32), shape1=np.zeros(n, np.int32), margin0=np.zeros(n, np.float32),
             margin1=np.zeros(n, np.float32))
    for i, (b, x, y, d) in enumerate(items):
        c["point0"][i] = (x, y, -d)
        c["point1"][i] = (x, y, 0.0)
        c["shape0"][i] = 1 + b
    return c


def _cfg(policy, k, **kw):
    return OstrichEngineConfig(contacts=ContactsConfig(max_per_world=K_SLOTS, reduction=ContactReductionConfig(
        policy=policy, max_per_pair=k, **kw)))


def _kept(model, state, items, cfg):
    out = oracle.pack(model, cfg, state, _contacts(items), precision=64)
    n = int(out["count"][0])
    p0 = out["point0"][: 3 * n].reshape(n, 3)
    # identify each survivor by its (x, y, depth) in the input list
    keys = [(round(float(p[0]), 6), round(float(p[1]), 6), round(-float(p[2]), 6)) for p in p0]
    ref = [(round(x, 6), round(y, 6), round(d, 6)) for (_, x, y, d) in items]
    return [ref.index(k) for k in keys]


def test_none_keeps_everything():
    model, st = _world()
    items = [(b, 0.01 * i, 0.0, 0.001 * (i + 1)) for i, b in enumerate([0, 1, 0, 2, 1])]
    assert _kept(model, st, items, _cfg("none", 1)) == [0, 1, 2, 3, 4]


def test_top_k_keeps_the_k_deepest_of_each_pair_in_input_order():
    model, st = _world()
    rng = np.random.default_rng(0)
    bodies = rng.integers(0, 3, size=14)
    depths = rng.permutation(14) * 1e-3 + 1e-3
    items = [(int(b), 0.02 * i, 0.01, float(d)) for i, (b, d) in enumerate(zip(bodies, depths))]
    for k in (1, 2, 3):
        kept = _kept(model, st, items, _cfg("top_k", k))
        assert kept == sorted(kept)  # packed order survives
        for b in range(3):
            idx = [i for i in range(14) if bodies[i] == b]
            deepest = sorted(idx, key=lambda i: -depths[i])[:k]
            assert sorted(i for i in kept if bodies[i] == b) == sorted(deepest), (k, b)


def test_top_k_ties_break_by_index():
    model, st = _world(1)
    items = [(0, 0.0, 0.0, 0.002), (0, 0.1, 0.0, 0.003), (0, 0.2, 0.0, 0.002), (0, 0.3, 0.0, 0.002)]
    assert _kept(model, st, items, _cfg("top_k", 2)) == [0, 1]


def test_fps_seeds_with_the_deepest_then_spreads():
    model, st = _world(1)
    # x positions 0, 1, 2, 3, 10; the deepest is x = 1, then the farthest from it (x = 10),
    # then the one whose nearest pick is farthest (x = 3: 2 m from x = 1)
    items = [(0, 0.0, 0.0, 0.001), (0, 1.0, 0.0, 0.005), (0, 2.0, 0.0, 0.002), (0, 3.0, 0.0, 0.003),
             (0, 10.0, 0.0, 0.0005)]
    assert _kept(model, st, items, _cfg("fps", 1)) == [1]
    assert _kept(model, st, items, _cfg("fps", 2)) == [1, 4]
    assert _kept(model, st, items, _cfg("fps", 3)) == [1, 3, 4]
    assert _kept(model, st, items, _cfg("fps", 9)) == [0, 1, 2, 3, 4]


def test_cluster_collapses_duplicates_and_keeps_the_deepest_representatives():
    model, st = _world(1)
    # two near-duplicate groups (1 mm apart) and one lone contact
    items = [(0, 0.0, 0.0, 0.002), (0, 0.001, 0.0, 0.004), (0, 0.5, 0.0, 0.001), (0, 0.501, 0.0, 0.0015),
             (0, 1.0, 0.0, 0.003)]
    assert _kept(model, st, items, _cfg("cluster", 4)) == [1, 3, 4]
    assert _kept(model, st, items, _cfg("cluster", 2)) == [1, 4]
    # farther than the position threshold: no cluster
    spread = [(0, 0.0, 0.0, 0.002), (0, 0.01, 0.0, 0.004)]
    assert _kept(model, st, spread, _cfg("cluster", 4)) == [0, 1]


def test_cluster_second_pass_also_ranks_earlier_pairs():
    """cluster.py:205-238 ranks every marked contact of the world against the current
    leader's pair, so a later pair with K deeper representatives drops an earlier pair's
    representative (restated as the reference's serial order gives it)."""
    model, st = _world(2)
    items = [(0, 0.0, 0.0, 0.001), (1, 1.0, 0.0, 0.005)]
    assert _kept(model, st, items, _cfg("cluster", 1)) == [1]
    # the other order: the deeper pair leads, the later shallow pair drops nothing
    items = [(1, 1.0, 0.0, 0.005), (0, 0.0, 0.0, 0.001)]
    assert _kept(model, st, items, _cfg("cluster", 1)) == [0, 1]
    # top_k and fps treat pairs independently
    assert _kept(model, st, [(0, 0.0, 0.0, 0.001), (1, 1.0, 0.0, 0.005)], _cfg("top_k", 1)) == [0, 1]
    assert _kept(model, st, [(0, 0.0, 0.0, 0.001), (1, 1.0, 0.0, 0.005)], _cfg("fps", 1)) == [0, 1]


def test_reduction_config_validation():
    with pytest.raises(NotImplementedError):
        ContactReductionConfig(policy="hull")
    with pytest.raises(ValueError):
        ContactReductionConfig(policy="top_k", max_per_pair=0)
