"""Model description consumed by the solver: the arrays of OstrichModel
(src/ostrich/core/model.py:171-342) built directly, without Newton.

A `WorldBuilder` collects the bodies, shapes and joints of one world; `Model.from_worlds`
stacks worlds that share one topology (model.py:200-203 requires uniform worlds) and
appends the global shapes (ground plane, heightfield) the way
`OstrichModel._broadcast_shape_array` does (model.py:305-319).  Shape geometry is kept
alongside for the primitive collider in `ostrich_amd.collide`.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _abi

JOINT_PRISMATIC, JOINT_REVOLUTE, JOINT_BALL, JOINT_FIXED, JOINT_FREE = 0, 1, 2, 3, 4
DOF_NONE, DOF_TARGET_POSITION, DOF_TARGET_VELOCITY = 0, 1, 2
GEO_PLANE, GEO_BOX, GEO_CYLINDER, GEO_SPHERE, GEO_HEIGHTFIELD = 0, 1, 2, 3, 4

# joint type -> number of DOFs (Newton joint_dof_dim)
_JOINT_DOFS = {JOINT_PRISMATIC: 1, JOINT_REVOLUTE: 1, JOINT_BALL: 3, JOINT_FIXED: 0, JOINT_FREE: 6}

IDENTITY7 = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)


def box_inertia(mass, hx, hy, hz):
    return np.diag([mass / 3.0 * (hy * hy + hz * hz), mass / 3.0 * (hx * hx + hz * hz),
                    mass / 3.0 * (hx * hx + hy * hy)])


def cylinder_inertia_y(mass, radius, half_height):
    """Solid cylinder whose axis is the body y axis."""
    h = 2.0 * half_height
    ixx = mass * (3.0 * radius * radius + h * h) / 12.0
    return np.diag([ixx, 0.5 * mass * radius * radius, ixx])


def sphere_inertia(mass, radius):
    return np.eye(3) * (0.4 * mass * radius * radius)


@dataclass
class Shape:
    geo: int
    body: int                      # world-local body, -1 = static
    xform: tuple = IDENTITY7       # shape frame in the body frame (world frame if static)
    size: tuple = (0.0, 0.0, 0.0)  # box half extents / (radius, half_height, 0) / (radius,0,0)
    mu: float = 0.5
    friction_axis: tuple = (0.0, 0.0, 0.0)
    mu_perp: float = -1.0
    mu_stiction_scale: float = -1.0
    v_stribeck: float = -1.0
    stribeck_lateral_only: float = -1.0


@dataclass
class WorldBuilder:
    bodies: list = field(default_factory=list)   # (mass, inertia 3x3, com 3)
    shapes: list = field(default_factory=list)
    joints: list = field(default_factory=list)
    dofs: list = field(default_factory=list)     # (axis, mode, ke, kd)

    def add_body(self, mass, inertia, com=(0.0, 0.0, 0.0)):
        self.bodies.append((float(mass), np.asarray(inertia, dtype=np.float64), tuple(com)))
        return len(self.bodies) - 1

    def add_shape(self, shape: Shape):
        self.shapes.append(shape)
        return len(self.shapes) - 1

    def _add_joint(self, jtype, parent, child, X_p, X_c, axis=(0.0, 0.0, 0.0), mode=DOF_NONE,
                   ke=0.0, kd=0.0, compliance=-1.0, enabled=True):
        qd_start = len(self.dofs)
        ndof = _JOINT_DOFS[jtype]
        for k in range(ndof):
            if jtype in (JOINT_PRISMATIC, JOINT_REVOLUTE):
                self.dofs.append((tuple(axis), mode, float(ke), float(kd)))
            else:
                self.dofs.append(((0.0, 0.0, 0.0), DOF_NONE, 0.0, 0.0))
        self.joints.append(dict(type=jtype, parent=parent, child=child, X_p=tuple(X_p), X_c=tuple(X_c),
                                qd_start=qd_start, enabled=1 if enabled else 0, compliance=float(compliance)))
        return len(self.joints) - 1

    def add_joint_free(self, child):
        return self._add_joint(JOINT_FREE, -1, child, IDENTITY7, IDENTITY7)

    def add_joint_revolute(self, parent, child, X_p, X_c=IDENTITY7, axis=(0.0, 1.0, 0.0),
                           mode=DOF_NONE, ke=0.0, kd=0.0, compliance=-1.0):
        return self._add_joint(JOINT_REVOLUTE, parent, child, X_p, X_c, axis, mode, ke, kd, compliance)

    def add_joint_prismatic(self, parent, child, X_p, X_c=IDENTITY7, axis=(1.0, 0.0, 0.0),
                            mode=DOF_NONE, ke=0.0, kd=0.0, compliance=-1.0):
        return self._add_joint(JOINT_PRISMATIC, parent, child, X_p, X_c, axis, mode, ke, kd, compliance)

    def add_joint_ball(self, parent, child, X_p, X_c=IDENTITY7, compliance=-1.0):
        return self._add_joint(JOINT_BALL, parent, child, X_p, X_c, compliance=compliance)

    def add_joint_fixed(self, parent, child, X_p, X_c=IDENTITY7, compliance=-1.0):
        return self._add_joint(JOINT_FIXED, parent, child, X_p, X_c, compliance=compliance)


class Model:
    """Batched model arrays (world-major) plus collider geometry."""

    def __init__(self):
        self.arrays = {}

    @staticmethod
    def from_worlds(worlds, global_shapes=(), gravity=(0.0, 0.0, -9.81), heightfield=None):
        """heightfield: geometry of the GEO_HEIGHTFIELD global shape, if there is one: a dict
        with `heights` [ny, nx] (metres, row = y index), `origin` (x, y) of node (0, 0) and
        `spacing` (dx, dy).  Only the collider reads it; the solver sees the shape's
        material (mu) like any other static shape."""
        m = Model()
        m.heightfield = heightfield
        W = len(worlds)
        w0 = worlds[0]
        B, S, J, D = len(w0.bodies), len(w0.shapes), len(w0.joints), len(w0.dofs)
        for w in worlds:
            if (len(w.bodies), len(w.shapes), len(w.joints), len(w.dofs)) != (B, S, J, D):
                raise ValueError("all worlds must share one topology (core/model.py:200-203)")
            if [j["type"] for j in w.joints] != [j["type"] for j in w0.joints]:
                raise ValueError("all worlds must share one joint layout")
        G = len(global_shapes)
        m.num_worlds, m.bodies_per_world, m.joints_per_world = W, B, J
        m.dofs_per_world, m.shapes_per_world, m.num_global_shapes = D, S, G
        m.gravity = tuple(float(g) for g in gravity)
        a = m.arrays
        a["body_mass"] = np.array([[b[0] for b in w.bodies] for w in worlds], np.float32).reshape(-1)
        a["body_inv_mass"] = np.where(a["body_mass"] > 0, 1.0 / np.maximum(a["body_mass"], 1e-30), 0.0).astype(np.float32)
        inertia = np.array([[b[1] for b in w.bodies] for w in worlds], np.float64).reshape(W * B, 3, 3)
        inv_inertia = np.linalg.inv(inertia)
        a["body_inertia"] = inertia.astype(np.float32).reshape(-1)
        a["body_inv_inertia"] = inv_inertia.astype(np.float32).reshape(-1)
        a["body_com"] = np.array([[b[2] for b in w.bodies] for w in worlds], np.float32).reshape(-1)
        jl = [j for w in worlds for j in w.joints]
        a["joint_type"] = np.array([j["type"] for j in jl], np.int32)
        a["joint_parent"] = np.array([j["parent"] for j in jl], np.int32)
        a["joint_child"] = np.array([j["child"] for j in jl], np.int32)
        a["joint_qd_start"] = np.array([j["qd_start"] for j in jl], np.int32)
        a["joint_enabled"] = np.array([j["enabled"] for j in jl], np.int32)
        a["joint_X_p"] = np.array([j["X_p"] for j in jl], np.float32).reshape(-1)
        a["joint_X_c"] = np.array([j["X_c"] for j in jl], np.float32).reshape(-1)
        a["joint_compliance"] = np.array([j["compliance"] for j in jl], np.float32)
        dl = [d for w in worlds for d in w.dofs]
        a["joint_axis"] = np.array([d[0] for d in dl], np.float32).reshape(-1) if dl else np.zeros(0, np.float32)
        a["joint_dof_mode"] = np.array([d[1] for d in dl], np.int32)
        a["joint_target_ke"] = np.array([d[2] for d in dl], np.float32)
        a["joint_target_kd"] = np.array([d[3] for d in dl], np.float32)
        # shape columns: per-world [0,S) then globals [S,S+G) (model.py:305-319)
        cols = [list(w.shapes) + list(global_shapes) for w in worlds]
        a["shape_body"] = np.array([[s.body for s in c] for c in cols], np.int32).reshape(-1)
        a["shape_mu"] = np.array([[s.mu for s in c] for c in cols], np.float32).reshape(-1)
        a["shape_friction_axis"] = np.array([[s.friction_axis for s in c] for c in cols], np.float32).reshape(-1)
        a["shape_mu_perp"] = np.array([[s.mu_perp for s in c] for c in cols], np.float32).reshape(-1)
        a["shape_mu_stiction_scale"] = np.array([[s.mu_stiction_scale for s in c] for c in cols], np.float32).reshape(-1)
        a["shape_v_stribeck"] = np.array([[s.v_stribeck for s in c] for c in cols], np.float32).reshape(-1)
        a["shape_stribeck_lateral_only"] = np.array([[s.stribeck_lateral_only for s in c] for c in cols], np.float32).reshape(-1)
        # Newton flat shape layout [globals | world_0 | world_1 | ...] -> world
        a["shape_world"] = np.concatenate([np.full(G, -1, np.int32),
                                           np.repeat(np.arange(W, dtype=np.int32), S)])
        for k in list(a):
            a[k] = np.ascontiguousarray(a[k])
        # collider geometry (per world columns, same order as above)
        m.shape_geo = np.array([[s.geo for s in c] for c in cols], np.int32)
        m.shape_xform = np.array([[s.xform for s in c] for c in cols], np.float64)
        m.shape_size = np.array([[s.size for s in c] for c in cols], np.float64)
        m.shape_body_2d = np.array([[s.body for s in c] for c in cols], np.int32)
        return m

    # Newton flat shape id of world-local column `col`
    def flat_shape_id(self, world, col):
        S, G = self.shapes_per_world, self.num_global_shapes
        col = np.asarray(col)
        return np.where(col < S, G + world * S + col, col - S)

    def to_abi(self, ptrs):
        """Fill an OfbModel from a dict name -> address (host or device)."""
        m = _abi.OfbModel()
        m.num_worlds = self.num_worlds
        m.bodies_per_world = self.bodies_per_world
        m.joints_per_world = self.joints_per_world
        m.dofs_per_world = self.dofs_per_world
        m.shapes_per_world = self.shapes_per_world
        m.num_global_shapes = self.num_global_shapes
        for i in range(3):
            m.gravity[i] = self.gravity[i]
        for name, _, _, _ in _abi.MODEL_ARRAYS:
            setattr(m, name, ptrs[name])
        return m
