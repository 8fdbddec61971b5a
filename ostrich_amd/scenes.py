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
taros4      - the reference's four-wheeled Taros-4 gradient-test robot
              (examples/taros_4/common.py:15-50): 1200 kg chassis box 2.74x1.5x1.0 m with
              its centre of mass 0.2 m up, four 50 kg cylinder wheels (r 0.45 m, width
              0.3 m) at x +-1.0, y +-0.9, z -0.3 on revolute joints about y.
tracked_flipper - a 40 kg body box 0.9x0.5x0.3 m with four 3 kg flippers on
              position-controlled revolute joints (targets U(-0.8,0.8) rad), six tracks of ten
              r = 0.05 m sphere contact points with anisotropic friction (mu 1.0 along the
              track, 0.5 across), on a shared 256x256 heightfield at 2 cm spacing made of 32
              Gaussian bumps (amplitude U(0,0.15) m, sigma U(0.1,0.5) m).

Each builder returns (model, state, control) with numpy arrays; `state` holds body_q,
body_qd, body_f in Newton's layout and `control` the joint targets.
"""
import math

import numpy as np

from .model import (DOF_TARGET_POSITION, DOF_TARGET_VELOCITY, GEO_BOX, GEO_CYLINDER, GEO_HEIGHTFIELD,
                    GEO_PLANE, GEO_SPHERE, IDENTITY7, Model, Shape, WorldBuilder, box_inertia,
                    cylinder_inertia_y)


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


def _skid_parts():
    """Axis-aligned bounds (lo, hi) of the robot's parts at its spawn pose (origin, yaw 0):
    the chassis box and the bounding box of each wheel cylinder (axis y)."""
    z0 = SKID_WHEEL_R - SKID_WHEEL_POS[0][2] + 0.002
    hx, hy, hz = SKID_CHASSIS_HALF
    parts = [((-hx, -hy, z0 - hz), (hx, hy, z0 + hz))]
    for x, y, z in SKID_WHEEL_POS:
        c = (x, y, z0 + z)
        parts.append(((c[0] - SKID_WHEEL_R, c[1] - SKID_WHEEL_HH, c[2] - SKID_WHEEL_R),
                      (c[0] + SKID_WHEEL_R, c[1] + SKID_WHEEL_HH, c[2] + SKID_WHEEL_R)))
    return z0, parts


def _place_against(x, y, half, parts, clearance=1e-3):
    """Moves a ground box of half extent `half` centred at (x, y) out of every robot part it
    overlaps, along the horizontal axis of least travel, so it ends `clearance` from that
    part's face; returns the new centre.  A box that clears every part stays where it is."""
    for _ in range(8):
        moved = False
        for lo, hi in parts:
            if 2.0 * half <= lo[2]:  # below the part: no overlap in z
                continue
            ox = min(x + half, hi[0]) - max(x - half, lo[0])
            oy = min(y + half, hi[1]) - max(y - half, lo[1])
            if ox <= -clearance or oy <= -clearance:
                continue
            cx, cy = 0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1])
            # candidate moves to each side of the part; the shortest wins
            moves = [(hi[0] + half + clearance - x, 0.0), (lo[0] - half - clearance - x, 0.0),
                     (0.0, hi[1] + half + clearance - y), (0.0, lo[1] - half - clearance - y)]
            # prefer the side the box centre is already on when travels tie
            moves.sort(key=lambda m: (abs(m[0]) + abs(m[1]),
                                      -((x - cx) * m[0] + (y - cy) * m[1])))
            x, y = x + moves[0][0], y + moves[0][1]
            moved = True
        if not moved:
            break
    return x, y


def skid_steer(num_worlds=4096, seed=1, max_obstacles=SKID_MAX_OBSTACLES, ke=150.0, kd=0.0,
               target_range=5.0, mu_range=(0.4, 1.0), area=4.0, layout="cluttered", clear_radius=0.8):
    """Uneven-batch skid-steer worlds (4 bodies, 15 joint rows, 3 velocity motor rows).

    The robot starts at the origin; 0-24 static boxes are placed uniformly in the
    `area` x `area` square around it.  layout "cluttered": a box that would overlap the
    robot is moved out along its shortest horizontal path to rest 1 mm from the part it
    hit, so robots start among and against obstacles (as the reference's obstacle
    benchmark drives its robot into boxes, examples/helhest/obstacle_benchmark.py:3-8).
    layout "clear": boxes keep out of a disc of `clear_radius` around the robot."""
    if layout not in ("cluttered", "clear"):
        raise ValueError(f"layout must be 'cluttered' or 'clear', got {layout!r}")
    rng = np.random.default_rng(seed)
    worlds, qs, targets = [], [], []
    z0, parts = _skid_parts()
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
                if layout == "clear":
                    while True:
                        x, y = rng.uniform(-0.5 * area, 0.5 * area, size=2)
                        if x * x + y * y > clear_radius * clear_radius:
                            break
                else:
                    x, y = rng.uniform(-0.5 * area, 0.5 * area, size=2)
                    x, y = _place_against(float(x), float(y), 0.5 * e, parts)
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


# the reference's Taros-4 robot (examples/taros_4/common.py:15-50, 183-290)
TAROS_CHASSIS_HALF = (1.37, 0.75, 0.5)
TAROS_CHASSIS_MASS = 1200.0
TAROS_CHASSIS_I = (600.0, 1000.0, 1000.0)
TAROS_CHASSIS_COM = (0.0, 0.0, 0.2)
TAROS_WHEEL_R = 0.45
TAROS_WHEEL_HH = 0.15
TAROS_WHEEL_MASS = 50.0
TAROS_WHEEL_I = (3.0, 3.0, 5.0)
TAROS_WHEEL_POS = ((1.0, 0.9, -0.3), (1.0, -0.9, -0.3), (-1.0, 0.9, -0.3), (-1.0, -0.9, -0.3))


def taros4(num_worlds=1, friction=0.8, control_mode="velocity", k_p=1000.0, k_d=0.0, z=0.75):
    """Taros-4 worlds: free chassis plus four wheels (front-left, front-right, rear-left,
    rear-right; dofs 6-9) on revolute joints about y, velocity or position motors with
    gains k_p / k_d, wheel and ground friction `friction`.  The chassis origin starts at
    height `z`: 0.75 sets the wheels on the ground (the reference's tests start it at 0.8,
    5 cm up, inside its 5 cm contact gap, tests/differentiable_simulator/
    test_wheeled_robot.py:29-41)."""
    if control_mode not in ("velocity", "position"):
        raise ValueError(f"control_mode must be 'velocity' or 'position', got {control_mode!r}")
    mode = DOF_TARGET_VELOCITY if control_mode == "velocity" else DOF_TARGET_POSITION
    worlds, qs = [], []
    for _ in range(num_worlds):
        wb = WorldBuilder()
        chassis = wb.add_body(TAROS_CHASSIS_MASS, np.diag(TAROS_CHASSIS_I), com=TAROS_CHASSIS_COM)
        wb.add_shape(Shape(GEO_BOX, chassis, TAROS_CHASSIS_COM + (0.0, 0.0, 0.0, 1.0), TAROS_CHASSIS_HALF,
                           mu=friction))
        wb.add_joint_free(chassis)
        poses = [(0.0, 0.0, z, 0.0, 0.0, 0.0, 1.0)]
        for pos in TAROS_WHEEL_POS:
            wheel = wb.add_body(TAROS_WHEEL_MASS, np.diag(TAROS_WHEEL_I))
            wb.add_shape(Shape(GEO_CYLINDER, wheel, IDENTITY7, (TAROS_WHEEL_R, TAROS_WHEEL_HH, 0.0), mu=friction))
            wb.add_joint_revolute(chassis, wheel, X_p=tuple(pos) + (0.0, 0.0, 0.0, 1.0), axis=(0.0, 1.0, 0.0),
                                  mode=mode, ke=k_p, kd=k_d)
            poses.append((pos[0], pos[1], z + pos[2], 0.0, 0.0, 0.0, 1.0))
        wb.add_shape(Shape(GEO_PLANE, -1, IDENTITY7, (0.0, 0.0, 0.0), mu=friction))
        worlds.append(wb)
        qs.append(poses)
    model = Model.from_worlds(worlds, [])
    B = model.bodies_per_world
    body_q = np.array(qs, np.float32).reshape(num_worlds * B, 7)
    state = dict(body_q=body_q, body_qd=np.zeros((num_worlds * B, 6), np.float32),
                 body_f=np.zeros((num_worlds * B, 6), np.float32))
    control = dict(joint_target_pos=np.zeros(num_worlds * model.dofs_per_world, np.float32),
                   joint_target_vel=np.zeros(num_worlds * model.dofs_per_world, np.float32))
    return model, state, control


# tracked flipper robot on rough terrain
TF_BODY_HALF = (0.45, 0.25, 0.15)
TF_BODY_MASS = 40.0
TF_FLIPPER_MASS = 3.0
TF_FLIPPER_LEN = 0.4
TF_FLIPPER_HALF = (0.2, 0.04, 0.06)
TF_SPHERE_R = 0.05
TF_TRACK_N = 10
TF_MAIN_TRACK_Y = 0.2
TF_HINGES = ((0.45, 0.33, -0.05, 1.0), (0.45, -0.33, -0.05, 1.0),   # (x, y, z, direction along x)
             (-0.45, 0.33, -0.05, -1.0), (-0.45, -0.33, -0.05, -1.0))
TF_KE, TF_KD = 2000.0, 50.0   # flipper position gains (N m / rad, N m s / rad)


def gaussian_bump_heightfield(seed, n=256, spacing=0.02, bumps=32, amp=(0.0, 0.15), sigma=(0.1, 0.5)):
    """Heights [n, n] of a sum of Gaussian bumps on a square grid centred on the origin."""
    rng = np.random.default_rng(seed)
    ext = (n - 1) * spacing
    origin = (-0.5 * ext, -0.5 * ext)
    xs = origin[0] + spacing * np.arange(n)
    X, Y = np.meshgrid(xs, xs)  # row = y index
    h = np.zeros((n, n))
    for _ in range(bumps):
        a = rng.uniform(*amp)
        sg = rng.uniform(*sigma)
        cx, cy = rng.uniform(origin[0], origin[0] + ext, size=2)
        h += a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * sg * sg))
    return dict(heights=h.astype(np.float32), origin=origin, spacing=(spacing, spacing))


def heightfield_height(hf, x, y):
    """Bilinear height of the heightfield at world (x, y) (numpy; clamped to the grid)."""
    h = np.asarray(hf["heights"], np.float64)
    ny, nx = h.shape
    u = np.clip((np.asarray(x) - hf["origin"][0]) / hf["spacing"][0], 0.0, nx - 1.0)
    v = np.clip((np.asarray(y) - hf["origin"][1]) / hf["spacing"][1], 0.0, ny - 1.0)
    i = np.minimum(np.floor(u).astype(int), nx - 2)
    j = np.minimum(np.floor(v).astype(int), ny - 2)
    fu, fv = u - i, v - j
    return (h[j, i] * (1 - fu) * (1 - fv) + h[j, i + 1] * fu * (1 - fv) + h[j + 1, i] * (1 - fu) * fv
            + h[j + 1, i + 1] * fu * fv)


def _quat_apply(q, v):
    x, y, z, w = q
    u = np.array([x, y, z])
    v = np.asarray(v, np.float64)
    return v * (2 * w * w - 1) + u * (2 * u.dot(v)) + np.cross(u, v) * (2 * w)


def tracked_flipper(num_worlds=1024, seed=2, target_range=0.8, place=1.2, terrain_seed=None, ke=TF_KE, kd=TF_KD):
    """Tracked flipper robots (5 bodies, 20 joint rows, 4 position motor rows) on one shared
    heightfield; each world places its robot at a seeded position and yaw, just above the
    terrain.  Every track sphere carries friction_axis = its body's x axis with mu 1.0 and
    mu_perp 0.0 against a terrain of mu 1.0: the reference averages the two shapes'
    coefficients (friction_constraint.py:143-156), so the effective coefficients are 1.0
    along the track and 0.5 across it."""
    rng = np.random.default_rng(seed)
    hf = gaussian_bump_heightfield(seed if terrain_seed is None else terrain_seed)
    terrain = Shape(GEO_HEIGHTFIELD, -1, IDENTITY7, (0.0, 0.0, 0.0), mu=1.0)
    sphere_kw = dict(mu=1.0, friction_axis=(1.0, 0.0, 0.0), mu_perp=0.0)
    bh = TF_BODY_HALF
    worlds, qs, targets = [], [], []
    for _ in range(num_worlds):
        wb = WorldBuilder()
        body = wb.add_body(TF_BODY_MASS, box_inertia(TF_BODY_MASS, *bh))
        wb.add_joint_free(body)
        pts = []  # (body, local point) of every track sphere, for the placement
        for side in (1.0, -1.0):
            for k in range(TF_TRACK_N):
                x = -0.4 + 0.8 * k / (TF_TRACK_N - 1)
                p = (x, side * TF_MAIN_TRACK_Y, -bh[2])
                wb.add_shape(Shape(GEO_SPHERE, body, p + (0.0, 0.0, 0.0, 1.0), (TF_SPHERE_R, 0.0, 0.0), **sphere_kw))
                pts.append(p)
        hinges = []
        for hx, hy, hz, d in TF_HINGES:
            fl = wb.add_body(TF_FLIPPER_MASS, box_inertia(TF_FLIPPER_MASS, *TF_FLIPPER_HALF))
            c_off = (d * 0.5 * TF_FLIPPER_LEN, 0.0, 0.0)  # flipper centre from its hinge
            wb.add_joint_revolute(body, fl, X_p=(hx, hy, hz, 0.0, 0.0, 0.0, 1.0),
                                  X_c=(-c_off[0], 0.0, 0.0, 0.0, 0.0, 0.0, 1.0), axis=(0.0, 1.0, 0.0),
                                  mode=DOF_TARGET_POSITION, ke=ke, kd=kd)
            for k in range(TF_TRACK_N):
                s = 0.04 + (TF_FLIPPER_LEN - 0.04) * k / (TF_TRACK_N - 1)
                p = (d * s - c_off[0], 0.0, -TF_FLIPPER_HALF[2])
                wb.add_shape(Shape(GEO_SPHERE, fl, p + (0.0, 0.0, 0.0, 1.0), (TF_SPHERE_R, 0.0, 0.0), **sphere_kw))
                pts.append((hx + d * s, hy, hz - TF_FLIPPER_HALF[2]))
            hinges.append((hx + c_off[0], hy, hz))
        x0, y0 = rng.uniform(-place, place, size=2)
        yaw = rng.uniform(-np.pi, np.pi)
        q = _yaw_quat(yaw)
        pw = np.array([_quat_apply(q, p) for p in pts]) + np.array([x0, y0, 0.0])
        ground = heightfield_height(hf, pw[:, 0], pw[:, 1])
        z0 = float(np.max(ground - pw[:, 2])) + TF_SPHERE_R + 0.002
        poses = [(x0, y0, z0) + q]
        for c in hinges:
            cw = _quat_apply(q, c)
            poses.append((x0 + cw[0], y0 + cw[1], z0 + cw[2]) + q)
        worlds.append(wb)
        qs.append(poses)
        t = np.zeros(len(wb.dofs), np.float32)
        t[6:10] = rng.uniform(-target_range, target_range, size=4)
        targets.append(t)
    model = Model.from_worlds(worlds, [terrain], heightfield=hf)
    B = model.bodies_per_world
    body_q = np.array(qs, np.float32).reshape(num_worlds * B, 7)
    state = dict(body_q=body_q, body_qd=np.zeros((num_worlds * B, 6), np.float32),
                 body_f=np.zeros((num_worlds * B, 6), np.float32))
    control = dict(joint_target_pos=np.concatenate(targets).astype(np.float32),
                   joint_target_vel=np.zeros(num_worlds * model.dofs_per_world, np.float32))
    return model, state, control
