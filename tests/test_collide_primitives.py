"""Plane and box contacts of the primitive collider, pinned by closed forms.

The host collider is the reference the GPU collider is checked against; these cases fix
its geometry where the answer is known exactly: a box resting on a plane touches it with
its four lower corners at distance 0, a box tilted about one axis with the two corners of
its lowest edge, a sphere on a plane or on a box face with the point below its centre, a
cylinder lying on a plane with the lowest point of its middle circle, and the Newton
conventions (normal from shape0 to shape1, point0 in the body frame, point1 in the frame
of shape1's body or the world).
"""
import math

import numpy as np
import torch

from ostrich_amd import collide
from ostrich_amd.model import (GEO_BOX, GEO_CYLINDER, GEO_PLANE, GEO_SPHERE, IDENTITY7, Model, Shape, WorldBuilder,
                               box_inertia, cylinder_inertia_y, sphere_inertia)


def _one_body(geo, size, pose, inertia, obstacle=None):
    wb = WorldBuilder()
    b = wb.add_body(1.0, inertia)
    wb.add_shape(Shape(geo, b, IDENTITY7, size, mu=0.5))
    wb.add_joint_free(b)
    if obstacle is not None:
        wb.add_shape(obstacle)
    model = Model.from_worlds([wb], [Shape(GEO_PLANE, -1, IDENTITY7, (0.0, 0.0, 0.0), mu=0.5)])
    return model, np.asarray([pose], np.float32)


def _contacts(model, q, gap=0.02):
    out = collide.PrimitiveCollider(model, gap=gap, max_contacts_per_world=64).collide(torch.as_tensor(q))
    n = int(out["count"][0])
    return {k: (v[:n].numpy() if hasattr(v, "numpy") and k != "count" else v) for k, v in out.items()}, n


def test_resting_box_touches_with_its_four_lower_corners():
    h = (0.3, 0.2, 0.1)
    model, q = _one_body(GEO_BOX, h, (1.0, -2.0, 0.1, 0, 0, 0, 1), box_inertia(1.0, *h))
    c, n = _contacts(model, q)
    assert n == 4
    np.testing.assert_allclose(c["point0"][:, 2], -0.1, atol=1e-7)  # body frame: the lower face
    np.testing.assert_allclose(np.abs(c["point0"][:, :2]), np.tile([0.3, 0.2], (4, 1)), atol=1e-7)
    np.testing.assert_allclose(c["point1"][:, 2], 0.0, atol=1e-6)  # world frame: on the plane
    np.testing.assert_allclose(c["normal"], np.tile([0.0, 0.0, -1.0], (4, 1)), atol=1e-7)  # box -> plane
    assert set(c["shape1"].tolist()) == {0}  # the global plane is flat shape 0


def test_tilted_box_touches_with_its_lowest_edge():
    h = 0.25
    ang = math.radians(30.0)
    qy = (0.0, math.sin(ang / 2), 0.0, math.cos(ang / 2))
    # rotated about y: the lowest corners (x, z) = (+h, -h) and (-h, -h) drop to
    # -h (sin + cos); lift the centre so that exactly those two corners sit on the plane
    zc = h * (math.sin(ang) + math.cos(ang))
    model, q = _one_body(GEO_BOX, (h, h, h), (0.0, 0.0, zc) + qy, box_inertia(1.0, h, h, h))
    c, n = _contacts(model, q, gap=1e-3)
    assert n == 2
    np.testing.assert_allclose(c["point1"][:, 2], 0.0, atol=1e-6)
    np.testing.assert_allclose(np.sort(c["point0"][:, 1]), [-h, h], atol=1e-7)


def test_sphere_on_plane_and_on_a_box_face():
    r = 0.1
    model, q = _one_body(GEO_SPHERE, (r, 0, 0), (0.3, 0.4, r - 0.005, 0, 0, 0, 1), sphere_inertia(1.0, r))
    c, n = _contacts(model, q)
    assert n == 1
    np.testing.assert_allclose(c["point0"][0], [0.0, 0.0, -r], atol=1e-7)
    np.testing.assert_allclose(c["point1"][0], [0.3, 0.4, 0.0], atol=1e-6)
    np.testing.assert_allclose(c["normal"][0], [0.0, 0.0, -1.0], atol=1e-7)
    # a static box obstacle whose top face (z = 0.5) is 0.01 below the sphere
    box = Shape(GEO_BOX, -1, (2.0, 0.0, 0.3, 0, 0, 0, 1), (0.4, 0.4, 0.2), mu=0.5)
    model, q = _one_body(GEO_SPHERE, (r, 0, 0), (2.1, 0.05, 0.5 + r + 0.01, 0, 0, 0, 1), sphere_inertia(1.0, r), box)
    c, n = _contacts(model, q)
    assert n == 1
    np.testing.assert_allclose(c["point1"][0], [2.1, 0.05, 0.5], atol=1e-6)
    np.testing.assert_allclose(c["point0"][0], [0.0, 0.0, -r], atol=1e-6)
    np.testing.assert_allclose(c["normal"][0], [0.0, 0.0, -1.0], atol=1e-6)


def test_lying_cylinder_touches_with_the_lowest_point_of_its_middle_circle():
    r, hh = 0.18, 0.05
    # axis (local y) kept horizontal, rotated about the world z by 40 degrees
    a = math.radians(40.0)
    qz = (0.0, 0.0, math.sin(a / 2), math.cos(a / 2))
    model, q = _one_body(GEO_CYLINDER, (r, hh, 0.0), (0.5, 0.5, r - 0.002) + qz, cylinder_inertia_y(1.0, r, hh))
    c, n = _contacts(model, q)
    assert n == 1
    np.testing.assert_allclose(c["point1"][0], [0.5, 0.5, 0.0], atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(c["point0"][0]), r, atol=1e-6)  # on the rim, body frame
    np.testing.assert_allclose(c["point0"][0][1], 0.0, atol=1e-7)  # middle circle


def test_candidates_beyond_the_gap_are_dropped():
    h = (0.3, 0.2, 0.1)
    model, q = _one_body(GEO_BOX, h, (0.0, 0.0, 0.1 + 0.03, 0, 0, 0, 1), box_inertia(1.0, *h))
    assert _contacts(model, q, gap=0.02)[1] == 0
    assert _contacts(model, q, gap=0.04)[1] == 4
