"""Seeded synthetic scenes for the solver workloads.

box_stack   - independent worlds, each a ground plane plus a vertical stack of cubes
              (edge 0.5 m, 1 kg) with seeded x/y offsets U(-0.02,0.02) m and yaw
              U(-0.1,0.1) rad, mu 0.6, 4 contacts per face contact.
skid_steer  - a 10 kg chassis box 0.6x0.4x0.2 m with three 1 kg cylinder wheels
              (r 0.18 m, width 0.1 m) on velocity-controlled revolute joints (the
              topology of the reference's Helhest robot, examples/helhest/common.py:206-335:
              free base + three revolute wheels, axis y), wheel targets U(-5,5) rad/s,
              0-24 static boxes with edge U(0.05,0.3) m placed uniformly in a 4x4 m area,
              mu U(0.4,1.0).

Each builder returns (model, state, control) with numpy arrays; `state` holds body_q,
body_qd, body_f in Newton's layout and `control` the joint targets.
"""
import math

import numpy as np

from .model import (DOF_TARGET_VELOCITY, GEO_BOX, GEO_CYLINDER, GEO_PLANE, IDENTITY7, Model,
                    Shape, WorldBuilder, box_inertia, cylinder_inertia_y)


def _yaw_quat(yaw):
    return (0.0, 0.0, math.sin(0.5 * yaw), math.cos(0.5 * yaw))


def ground_plane(mu=0.6):
    return Shape(GEO_PLANE, -1, IDENTITY7, (0.0, 0.0, 0.0), mu=mu)


def box_stack(num_worlds=16, n_cubes=3, seed=0, edge=0.5, mass=1.0, mu=0.6, drop=0.0):
    rng = np.random.default_rng(seed)
    h = 0.5 * edge
    worlds, qs = [], []
    for _ in range(num_worlds):
        wb = WorldBuilder()
        poses = []
        for i in range(n_cubes):
            b = wb.add_body(mass, box_inertia(mass, h, h, h))
            wb.add_shape(Shape(GEO_BOX, b, IDENTITY7, (h, h, h), mu=mu))
            wb.add_joint_free(b)
            dx, dy = rng.uniform(-0.02, 0.02, size=2)
            yaw = rng.uniform(-0.1, 0.1)
            z = h + i * edge + drop * (i + 1)
            poses.append((dx, dy, z) + _yaw_quat(yaw))
        worlds.append(wb)
        qs.append(poses)
    model = Model.from_worlds(worlds, [ground_plane(mu)])
    B = n_cubes
    body_q = np.array(qs, np.float32).reshape(num_worlds * B, 7)
    state = dict(body_q=body_q, body_qd=np.zeros((num_worlds * B, 6), np.float32),
                 body_f=np.zeros((num_worlds * B, 6), np.float32))
    control = dict(joint_target_pos=np.zeros(num_worlds * model.dofs_per_world, np.float32),
                   joint_target_vel=np.zeros(num_worlds * model.dofs_per_world, np.float32))
    return model, state, control


SKID_CHASSIS_HALF = (0.3, 0.2, 0.1)
SKID_WHEEL_R = 0.18
SKID_WHEEL_HH = 0.05
SKID_WHEEL_POS = ((0.2, 0.26, -0.05), (0.2, -0.26, -0.05), (-0.45, 0.0, -0.05))
SKID_MAX_OBSTACLES = 24


def skid_steer(num_worlds=4096, seed=1, max_obstacles=SKID_MAX_OBSTACLES, ke=150.0, kd=0.0,
               target_range=5.0, mu_range=(0.4, 1.0), area=4.0, clear_radius=0.8):
    """Uneven-batch skid-steer worlds (4 bodies, 15 joint rows, 3 velocity motor rows)."""
    rng = np.random.default_rng(seed)
    worlds, qs, targets = [], [], []
    z0 = SKID_WHEEL_R - SKID_WHEEL_POS[0][2] + 0.002
    for _ in range(num_worlds):
        mu = float(rng.uniform(*mu_range))
        wb = WorldBuilder()
        chassis = wb.add_body(10.0, box_inertia(10.0, *SKID_CHASSIS_HALF))
        wb.add_shape(Shape(GEO_BOX, chassis, IDENTITY7, SKID_CHASSIS_HALF, mu=mu))
        wb.add_joint_free(chassis)
        poses = [(0.0, 0.0, z0, 0.0, 0.0, 0.0, 1.0)]
        for pos in SKID_WHEEL_POS:
            wheel = wb.add_body(1.0, cylinder_inertia_y(1.0, SKID_WHEEL_R, SKID_WHEEL_HH))
            wb.add_shape(Shape(GEO_CYLINDER, wheel, IDENTITY7, (SKID_WHEEL_R, SKID_WHEEL_HH, 0.0), mu=mu))
            wb.add_joint_revolute(chassis, wheel, X_p=tuple(pos) + (0.0, 0.0, 0.0, 1.0),
                                  axis=(0.0, 1.0, 0.0), mode=DOF_TARGET_VELOCITY, ke=ke, kd=kd)
            poses.append((pos[0], pos[1], z0 + pos[2], 0.0, 0.0, 0.0, 1.0))
        # per-world ground plane so each world carries its own mu
        wb.add_shape(Shape(GEO_PLANE, -1, IDENTITY7, (0.0, 0.0, 0.0), mu=mu))
        n_obs = int(rng.integers(0, max_obstacles + 1))
        for k in range(max_obstacles):
            if k < n_obs:
                e = float(rng.uniform(0.05, 0.3))
                while True:
                    x, y = rng.uniform(-0.5 * area, 0.5 * area, size=2)
                    if x * x + y * y > clear_radius * clear_radius:
                        break
                wb.add_shape(Shape(GEO_BOX, -1, (float(x), float(y), 0.5 * e, 0.0, 0.0, 0.0, 1.0),
                                   (0.5 * e, 0.5 * e, 0.5 * e), mu=mu))
            else:  # unused slot: far below the ground, never in contact
                wb.add_shape(Shape(GEO_BOX, -1, (0.0, 0.0, -100.0 - k, 0.0, 0.0, 0.0, 1.0),
                                   (0.01, 0.01, 0.01), mu=mu))
        worlds.append(wb)
        qs.append(poses)
        t = np.zeros(len(wb.dofs), np.float32)
        t[6:9] = rng.uniform(-target_range, target_range, size=3)
        targets.append(t)
    model = Model.from_worlds(worlds, [])
    B = model.bodies_per_world
    body_q = np.array(qs, np.float32).reshape(num_worlds * B, 7)
    state = dict(body_q=body_q, body_qd=np.zeros((num_worlds * B, 6), np.float32),
                 body_f=np.zeros((num_worlds * B, 6), np.float32))
    control = dict(joint_target_pos=np.zeros(num_worlds * model.dofs_per_world, np.float32),
                   joint_target_vel=np.concatenate(targets).astype(np.float32))
    return model, state, control

