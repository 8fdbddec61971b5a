"""The gfx950 collider (`ofb_collide`) against PrimitiveCollider on the host.

Both generate the same candidates and tests in FP32; the kernels' rounding (sin/cos,
fused multiply-adds) differs from torch's host arithmetic in the last bits, so a
candidate whose distance lies within rounding of the gap (or whose tangent excess lies
within rounding of the slack) may flip.  Validity is monotone in both thresholds, so for
each world the host collider runs at (gap +- DELTA, slack +- DELTA): a world whose two
contact sets agree has no candidate that close to a threshold, and its HIP contacts must
equal the host's exactly in number, order and shape ids, with geometry within 2e-5 (m,
unit normals); any other world's HIP count must lie between the two host counts.  The
cap (W * max_contacts_per_world, kept in flat order) is checked against the HIP
collider's own uncapped output, which the comparison above pins.
"""
import numpy as np
import pytest
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.model import (GEO_BOX, GEO_CYLINDER, GEO_PLANE, GEO_SPHERE, IDENTITY7, Model, Shape, WorldBuilder,
                               box_inertia, cylinder_inertia_y)

pytestmark = pytest.mark.gpu

GAP, SLACK, DELTA = 0.02, 0.06, 2e-5


def _mixed_scene(W=64, seed=3):
    """Dynamic box, cylinder and sphere over a ground plane with a static box obstacle
    and a dynamic box below the sphere (every shape-pair kind the collider handles)."""
    rng = np.random.default_rng(seed)
    worlds, poses = [], []
    for _ in range(W):
        wb = WorldBuilder()
        b0 = wb.add_body(1.0, box_inertia(1.0, 0.2, 0.1, 0.15))
        wb.add_shape(Shape(GEO_BOX, b0, IDENTITY7, (0.2, 0.1, 0.15), mu=0.5))
        wb.add_joint_free(b0)
        b1 = wb.add_body(1.0, cylinder_inertia_y(1.0, 0.12, 0.05))
        wb.add_shape(Shape(GEO_CYLINDER, b1, IDENTITY7, (0.12, 0.05, 0.0), mu=0.5))
        wb.add_joint_free(b1)
        b2 = wb.add_body(0.5, box_inertia(0.5, 0.05, 0.05, 0.05))
        wb.add_shape(Shape(GEO_SPHERE, b2, IDENTITY7, (0.05, 0.0, 0.0), mu=0.5))
        wb.add_joint_free(b2)
        b3 = wb.add_body(2.0, box_inertia(2.0, 0.3, 0.3, 0.1))
        wb.add_shape(Shape(GEO_BOX, b3, IDENTITY7, (0.3, 0.3, 0.1), mu=0.5))
        wb.add_joint_free(b3)
        wb.add_shape(Shape(GEO_BOX, -1, (1.0, 0.0, 0.1, 0.0, 0.0, 0.0, 1.0), (0.3, 0.3, 0.1), mu=0.5))
        worlds.append(wb)
        z = rng.uniform(-0.02, 0.04, 4)
        poses.append([(0.0, 0.0, 0.15 + z[0], 0, 0, 0, 1), (1.0, 0.0, 0.32 + z[1], 0, 0, 0, 1),
                      (-1.0, 0.0, 0.25 + z[2], 0, 0, 0, 1), (-1.0, 0.0, 0.1 + z[3], 0, 0, 0, 1)])
    model = Model.from_worlds(worlds, [Shape(GEO_PLANE, -1, IDENTITY7, (0.0, 0.0, 0.0), mu=0.5)])
    return model, np.asarray(poses, np.float64).reshape(-1, 7)


def _perturb(q, rng, pos=0.03, rot=0.15):
    q = np.array(q, np.float64).reshape(-1, 7).copy()
    q[:, :3] += rng.uniform(-pos, pos, (q.shape[0], 3))
    half = 0.5 * rng.uniform(-rot, rot, (q.shape[0], 3))
    dq = np.concatenate([np.sin(half), np.cos(np.linalg.norm(half, axis=1, keepdims=True))], 1)
    x1, y1, z1, w1 = dq.T
    x2, y2, z2, w2 = q[:, 3:].T
    qq = np.stack([w1 * x2 + w2 * x1 + y1 * z2 - z1 * y2, w1 * y2 + w2 * y1 + z1 * x2 - x1 * z2,
                   w1 * z2 + w2 * z1 + x1 * y2 - y1 * x2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2], 1)
    q[:, 3:] = qq / np.linalg.norm(qq, axis=1, keepdims=True)
    return q.astype(np.float32)


def _host(model, q, K, gap=GAP, slack=SLACK, hf="point"):
    return collide.PrimitiveCollider(model, gap=gap, lateral_slack=slack, max_contacts_per_world=K,
                                     hf_contacts=hf).collide(torch.as_tensor(q))


def _world_of(model, shape0):
    return (shape0.long() - model.num_global_shapes) // model.shapes_per_world


def _by_world(model, c):
    n = int(c["count"][0])
    w = _world_of(model, c["shape0"][:n])
    return n, w, torch.bincount(w, minlength=model.num_worlds)


def _compare(model, hip, q, gap=GAP, hf="point"):
    """Per-world comparison of the uncapped HIP contacts with the host's; returns the
    number of worlds compared exactly."""
    big = 4096
    ref = _host(model, q, big, gap, hf=hf)
    lo = _by_world(model, _host(model, q, big, gap - DELTA, SLACK - DELTA, hf))[2]
    hi = _by_world(model, _host(model, q, big, gap + DELTA, SLACK + DELTA, hf))[2]
    out = hip.collide(torch.as_tensor(q, device="cuda:0").contiguous())
    torch.cuda.synchronize()
    n = int(out.rigid_contact_count.item())
    mine = dict(count=torch.tensor([n]), shape0=out.rigid_contact_shape0[:n].cpu(),
                shape1=out.rigid_contact_shape1[:n].cpu(), point0=out.rigid_contact_point0[:n].cpu(),
                point1=out.rigid_contact_point1[:n].cpu(), normal=out.rigid_contact_normal[:n].cpu())
    _, w_mine, cnt_mine = _by_world(model, mine)
    n_ref, w_ref, cnt_ref = _by_world(model, ref)
    assert torch.all(w_mine[1:] >= w_mine[:-1]), "contacts not in world order"
    assert torch.all(lo <= cnt_mine) and torch.all(cnt_mine <= hi)
    clear = torch.nonzero(lo == hi).squeeze(-1)
    assert clear.numel() >= model.num_worlds // 2, f"only {clear.numel()} worlds clear of the thresholds"
    off_m = torch.cumsum(cnt_mine, 0) - cnt_mine
    off_r = torch.cumsum(cnt_ref, 0) - cnt_ref
    for w in clear.tolist():
        k = int(cnt_ref[w])
        assert int(cnt_mine[w]) == k, w
        if k == 0:
            continue
        sm, sr = slice(int(off_m[w]), int(off_m[w]) + k), slice(int(off_r[w]), int(off_r[w]) + k)
        assert torch.equal(mine["shape0"][sm], ref["shape0"][sr]), w
        assert torch.equal(mine["shape1"][sm], ref["shape1"][sr]), w
        for f in ("point0", "point1", "normal"):
            torch.testing.assert_close(mine[f][sm], ref[f][sr], atol=2e-5, rtol=1e-5, msg=f"world {w} {f}")
    assert torch.count_nonzero(out.rigid_contact_margin0[:n]) == 0
    assert torch.count_nonzero(out.rigid_contact_margin1[:n]) == 0
    return clear.numel(), n


def _scene(name):
    if name == "skid_steer":
        model, state, _ = scenes.skid_steer(128, seed=1)
    elif name == "box_stack":
        model, state, _ = scenes.box_stack(64, n_cubes=3, seed=0)
    elif name in ("tracked_flipper", "tracked_flipper_cells"):
        model, state, _ = scenes.tracked_flipper(64, seed=2)
    else:
        model, q = _mixed_scene()
        return model, q
    return model, np.asarray(state["body_q"], np.float64)


@pytest.mark.parametrize("name", ["skid_steer", "box_stack", "tracked_flipper", "tracked_flipper_cells", "mixed"])
def test_hip_collider_matches_host_collider(name, gpu_device):
    """tracked_flipper_cells: four contacts per track sphere (the grid cells around it) at
    a 0.05 m gap, the tracked-robot workload's contact density (tens to ~240 a world)."""
    model, q0 = _scene(name)
    hf, gap = ("cells", 0.05) if name.endswith("_cells") else ("point", GAP)
    hip = collide.HipCollider(model, device=gpu_device, gap=gap, lateral_slack=SLACK, max_contacts_per_world=4096,
                              hf_contacts=hf)
    rng = np.random.default_rng(11)
    for k in range(4):
        q = _perturb(q0, rng) if k else np.asarray(q0, np.float32)
        worlds, n = _compare(model, hip, q, gap, hf)
        print(f"[collide] {name} pose set {k}: {n} contacts, {worlds}/{model.num_worlds} worlds exact", flush=True)


@pytest.mark.parametrize("K", [1, 3])
def test_hip_collider_cap_keeps_the_first_contacts_in_flat_order(K, gpu_device):
    model, q0 = _scene("mixed")
    full = collide.HipCollider(model, device=gpu_device, max_contacts_per_world=4096)
    capped = collide.HipCollider(model, device=gpu_device, max_contacts_per_world=K)
    q = torch.as_tensor(_perturb(q0, np.random.default_rng(7)), device=gpu_device).contiguous()
    a, b = full.collide(q), capped.collide(q)
    torch.cuda.synchronize()
    n_all, n = int(a.rigid_contact_count.item()), int(b.rigid_contact_count.item())
    cap = model.num_worlds * K
    assert n_all > cap and n == cap
    for f in ("rigid_contact_point0", "rigid_contact_point1", "rigid_contact_normal", "rigid_contact_shape0",
              "rigid_contact_shape1"):
        assert torch.equal(getattr(a, f)[:n], getattr(b, f)[:n]), f


def test_hip_collider_in_a_captured_graph(gpu_device):
    model, q0 = _scene("skid_steer")
    eager = collide.HipCollider(model, device=gpu_device, max_contacts_per_world=128)
    graphed = collide.HipCollider(model, device=gpu_device, max_contacts_per_world=128)
    q_static = torch.as_tensor(np.asarray(q0, np.float32), device=gpu_device).contiguous()
    graphed.collide(q_static)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        graphed.collide(q_static)
    rng = np.random.default_rng(5)
    for k in range(4):
        q = torch.as_tensor(_perturb(q0, rng), device=gpu_device).contiguous()
        q_static.copy_(q)
        g.replay()
        a = eager.collide(q)
        torch.cuda.synchronize()
        b = graphed.contacts
        n = int(a.rigid_contact_count.item())
        print(f"[collide] graph replay {k}: {n} contacts", flush=True)
        assert int(b.rigid_contact_count.item()) == n
        for f in ("rigid_contact_point0", "rigid_contact_point1", "rigid_contact_normal", "rigid_contact_shape0",
                  "rigid_contact_shape1"):
            assert torch.equal(getattr(a, f)[:n], getattr(b, f)[:n]), (k, f)
        junk = torch.randn(1 << 16, device=gpu_device).sum()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)


def test_hip_collider_rejects_bad_inputs(gpu_device):
    model, q0 = _scene("box_stack")
    hip = collide.HipCollider(model, device=gpu_device, max_contacts_per_world=8)
    with pytest.raises(ValueError):
        hip.collide(torch.zeros(3, 7, device=gpu_device))
    with pytest.raises(ValueError):
        hip.collide(torch.as_tensor(np.asarray(q0, np.float32)))
