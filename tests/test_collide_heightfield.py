"""Sphere-heightfield contacts of the primitive collider (the tracked-robot workload).

Newton's own heightfield collision is not available here, so these tests pin the
collider by geometry: bilinear interpolation is exact on planes, so a heightfield sampled
from a plane must give the plane's contact (normal, distance, points) for every sphere.
"""
import numpy as np
import torch

from ostrich_amd import collide, scenes
from ostrich_amd.model import GEO_HEIGHTFIELD, GEO_SPHERE, IDENTITY7, Model, Shape, WorldBuilder, sphere_inertia


def _sphere_world(points, r=0.05, hf=None):
    worlds = []
    for _ in range(len(points)):
        wb = WorldBuilder()
        b = wb.add_body(1.0, sphere_inertia(1.0, r))
        wb.add_shape(Shape(GEO_SPHERE, b, IDENTITY7, (r, 0.0, 0.0), mu=0.7))
        wb.add_joint_free(b)
        worlds.append(wb)
    model = Model.from_worlds(worlds, [Shape(GEO_HEIGHTFIELD, -1, IDENTITY7, (0, 0, 0), mu=0.9)], heightfield=hf)
    q = np.zeros((len(points), 7), np.float32)
    q[:, :3] = points
    q[:, 6] = 1.0
    return model, q


def _plane_field(a, b, c, n=64, spacing=0.05):
    origin = (-0.5 * (n - 1) * spacing, -0.5 * (n - 1) * spacing)
    xs = origin[0] + spacing * np.arange(n)
    X, Y = np.meshgrid(xs, xs)
    return dict(heights=(a * X + b * Y + c).astype(np.float32), origin=origin, spacing=(spacing, spacing))


def test_planar_heightfield_gives_plane_contacts():
    a, b, c, r = 0.2, -0.1, 0.3, 0.05
    rng = np.random.default_rng(0)
    xy = rng.uniform(-1.0, 1.0, size=(16, 2))
    dist = rng.uniform(-0.01, 0.015, size=16)  # signed sphere-surface distance, all within the gap
    nrm = np.array([-a, -b, 1.0]) / np.sqrt(a * a + b * b + 1.0)
    foot = np.column_stack([xy, a * xy[:, 0] + b * xy[:, 1] + c])
    centres = foot + nrm * (r + dist)[:, None]
    model, q = _sphere_world(centres, r, _plane_field(a, b, c))
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=4)
    out = col.collide(torch.as_tensor(q))
    assert int(out["count"][0]) == 16
    # Newton's convention: the normal points from shape0 (the sphere) to shape1 (terrain)
    np.testing.assert_allclose(out["normal"][:16].numpy(), np.tile(-nrm, (16, 1)), atol=1e-5)
    p0 = out["point0"][:16].numpy()          # sphere body frame (identity rotation)
    p1 = out["point1"][:16].numpy()          # world frame (static terrain)
    np.testing.assert_allclose(p0, np.tile(-nrm * r, (16, 1)), atol=1e-5)
    gap = ((centres - nrm * r) - p1) @ nrm   # signed distance along the normal
    np.testing.assert_allclose(gap, dist, atol=2e-5)
    # p1 lies on the plane
    np.testing.assert_allclose(p1[:, 2], a * p1[:, 0] + b * p1[:, 1] + c, atol=2e-5)
    # shape ids: the sphere of world w (flat id G + w) against the global terrain (0)
    np.testing.assert_array_equal(out["shape0"][:16].numpy(), 1 + np.arange(16))
    np.testing.assert_array_equal(out["shape1"][:16].numpy(), np.zeros(16))


def test_far_spheres_make_no_contact():
    model, q = _sphere_world(np.array([[0.0, 0.0, 0.3 + 0.05 + 0.05]]), 0.05, _plane_field(0.0, 0.0, 0.3))
    out = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=4).collide(torch.as_tensor(q))
    assert int(out["count"][0]) == 0


def test_surface_interpolation_matches_grid():
    hf = scenes.gaussian_bump_heightfield(7, n=32)
    model, _ = _sphere_world(np.zeros((1, 3)), 0.05, hf)
    col = collide.PrimitiveCollider(model)
    ox, oy = hf["origin"]
    dx = hf["spacing"][0]
    j, i = np.meshgrid(np.arange(31), np.arange(31), indexing="ij")
    z, gx, gy = col.heightfield_surface(torch.as_tensor(ox + dx * i, dtype=torch.float32),
                                        torch.as_tensor(oy + dx * j, dtype=torch.float32))
    np.testing.assert_allclose(z.numpy(), hf["heights"][:31, :31], atol=1e-6)
    rng = np.random.default_rng(1)
    x, y = rng.uniform(ox, ox + 30 * dx, size=(2, 200))
    z, gx, gy = col.heightfield_surface(torch.as_tensor(x, dtype=torch.float32), torch.as_tensor(y, dtype=torch.float32))
    np.testing.assert_allclose(z.numpy(), scenes.heightfield_height(hf, x, y), atol=1e-5)
    e = 1e-4
    fd = (scenes.heightfield_height(hf, x + e, y) - scenes.heightfield_height(hf, x - e, y)) / (2 * e)
    inside = np.abs(((x - ox) / dx) - np.round((x - ox) / dx)) > 0.01  # away from cell edges
    np.testing.assert_allclose(gx.numpy()[inside], fd[inside], atol=2e-3)


def test_tracked_flipper_scene_shapes():
    model, state, control = scenes.tracked_flipper(4, seed=2)
    assert model.bodies_per_world == 5 and model.joints_per_world == 5 and model.dofs_per_world == 10
    assert model.shapes_per_world == 60 and model.num_global_shapes == 1
    t = control["joint_target_pos"].reshape(4, 10)
    assert np.all(t[:, :6] == 0) and np.all(np.abs(t[:, 6:]) <= 0.8)
    col = collide.PrimitiveCollider(model, gap=0.02, max_contacts_per_world=64)
    out = col.collide(torch.as_tensor(state["body_q"]))
    assert int(out["count"][0]) >= 4  # every robot starts touching the terrain


def _brute_cell(hf, c, ci, cj, n=300):
    """Closest point of grid cell (ci, cj)'s two triangles to c by dense barycentric
    sampling, and its distance."""
    h = hf["heights"]
    (ox, oy), (sx, sy) = hf["origin"], hf["spacing"]
    P = lambda i, j: np.array([ox + i * sx, oy + j * sy, h[j, i]], np.float64)
    P00, P10, P01, P11 = P(ci, cj), P(ci + 1, cj), P(ci, cj + 1), P(ci + 1, cj + 1)
    s, t = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))
    m = s + t <= 1.0
    s, t = s[m], t[m]
    best, bd = None, np.inf
    for a, b, cc in ((P00, P10, P11), (P00, P11, P01)):
        pts = a + s[:, None] * (b - a) + t[:, None] * (cc - a)
        d = np.linalg.norm(pts - c, axis=1)
        k = int(d.argmin())
        if d[k] < bd:
            bd, best = d[k], pts[k]
    return best, bd


def test_cell_contacts_are_closest_points_of_the_cells():
    """Mode "cells": a sphere's contact with each of the four grid cells around its
    nearest grid vertex lies at the closest point of the cell (checked against dense
    sampling of the cell's two triangles), with the normal from that point to the centre
    and distance |c - p| - r; a cell gives none only when the centre is behind the planes
    of both its triangles (a neighbouring triangle holds the contact)."""
    hf = scenes.gaussian_bump_heightfield(7, n=32, spacing=0.05)
    (ox, oy), (sx, sy) = hf["origin"], hf["spacing"]
    rng = np.random.default_rng(3)
    r = 0.05
    xy = rng.uniform(ox + 3 * sx, ox + 27 * sx, size=(12, 2))
    z = scenes.heightfield_height(hf, xy[:, 0], xy[:, 1]) + r + rng.uniform(-0.01, 0.015, 12)
    centres = np.column_stack([xy, z])
    model, q = _sphere_world(centres, r, hf)
    col = collide.PrimitiveCollider(model, gap=0.5, max_contacts_per_world=4, hf_contacts="cells")
    c = torch.as_tensor(centres, dtype=torch.float32)[:, None, :]
    foot, nrm, dist, ok = (t.numpy() for t in col._heightfield_cell_contacts(c, torch.full((12, 1), r)))
    h = hf["heights"]
    P = lambda i, j: np.array([ox + i * sx, oy + j * sy, h[j, i]], np.float64)
    used = 0
    for w in range(12):
        cw = centres[w]
        I = int(np.floor((cw[0] - ox) / sx + 0.5))
        J = int(np.floor((cw[1] - oy) / sy + 0.5))
        for qd in range(4):
            ci, cj = I - 1 + (qd & 1), J - 1 + (qd >> 1)
            if not ok[w, 0, qd]:
                P00, P10, P01, P11 = P(ci, cj), P(ci + 1, cj), P(ci, cj + 1), P(ci + 1, cj + 1)
                for A, B_, C_ in ((P00, P10, P11), (P00, P11, P01)):
                    n = np.cross(B_ - A, C_ - A)
                    assert (cw - A) @ n <= 1e-6 * np.linalg.norm(n)
                continue
            used += 1
            pb, db = _brute_cell(hf, cw, ci, cj)
            assert np.linalg.norm(foot[w, 0, qd] - pb) < 4e-3 * sx
            d = np.linalg.norm(cw - foot[w, 0, qd])
            assert abs(d - db) < 1e-4 * sx + 1e-6
            assert abs(dist[w, 0, qd] - (d - r)) < 1e-5
            np.testing.assert_allclose(nrm[w, 0, qd], (cw - foot[w, 0, qd]) / d, atol=2e-4)
    assert used >= 36
    # the contacts the collider emits are those cells' (normal from the sphere to the terrain)
    out = col.collide(torch.as_tensor(q))
    n = int(out["count"][0])
    assert n == used
    np.testing.assert_allclose(out["normal"][:n].numpy(), -nrm[ok], atol=1e-6)
    np.testing.assert_allclose(out["point1"][:n].numpy(), foot[ok], atol=1e-6)


def test_cell_contacts_on_a_flat_field():
    """Flat terrain at height 0.3, sphere centre at (x, y, 0.3 + r + g): the cell under the
    centre gives distance g with a vertical normal; a cell beside it, at lateral offset l
    from the centre, gives sqrt((r + g)^2 + l^2) - r; cells off the grid give none."""
    r, g = 0.05, 0.004
    hf = _plane_field(0.0, 0.0, 0.3, n=16, spacing=0.05)
    (ox, oy), (sx, sy) = hf["origin"], hf["spacing"]
    x, y = ox + 7.3 * sx, oy + 7.6 * sy    # nearest vertex (7, 8): cells i 6..7, j 7..8
    model, q = _sphere_world(np.array([[x, y, 0.3 + r + g]]), r, hf)
    out = collide.PrimitiveCollider(model, gap=0.5, max_contacts_per_world=4, hf_contacts="cells").collide(
        torch.as_tensor(q))
    assert int(out["count"][0]) == 4
    p1 = out["point1"][:4].numpy()
    cells = [(6, 7), (7, 7), (6, 8), (7, 8)]
    for k, (ci, cj) in enumerate(cells):
        px = np.clip(x, ox + ci * sx, ox + (ci + 1) * sx)
        py = np.clip(y, oy + cj * sy, oy + (cj + 1) * sy)
        np.testing.assert_allclose(p1[k], [px, py, 0.3], atol=1e-5)
    d = np.linalg.norm(np.array([x, y, 0.3 + r + g]) - p1, axis=1) - r
    assert abs(d[1] - g) < 1e-5                                            # the cell under the centre
    # at the corner of the grid only the cell inside it remains
    model, q = _sphere_world(np.array([[ox + 0.2 * sx, oy + 0.3 * sy, 0.3 + r]]), r, hf)
    out = collide.PrimitiveCollider(model, gap=0.5, max_contacts_per_world=4, hf_contacts="cells").collide(
        torch.as_tensor(q))
    assert int(out["count"][0]) == 1
