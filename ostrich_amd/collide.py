"""Primitive collider for the benchmark scenes (plane / box / cylinder / sphere /
heightfield).

It produces Newton-format rigid contacts (flat arrays; normal from shape0 to shape1;
point0/point1 in the frame of the owning body, world frame for static shapes; margins),
the input format of the solver boundary (core/contacts.py:25-99).  Newton's own
`collide` is not available here, so contact geometry is this module's approximation:
sample points of the moving shape are tested against planes and boxes.  The routine is
written with torch tensor ops so the same code runs on host tensors (tests) and on the
GPU (benchmark loop); the solver never depends on it.

Candidate pairs are fixed per model; each pair contributes a fixed block of candidate
points, and valid candidates are emitted in (world, pair, point) order, which makes the
flat contact order deterministic.
"""
import math

import numpy as np
import torch

from .model import GEO_BOX, GEO_CYLINDER, GEO_HEIGHTFIELD, GEO_PLANE, GEO_SPHERE


def _quat_rotate(q, v):
    """q [...,4] xyzw (unit), v [...,3]."""
    u = q[..., :3]
    w = q[..., 3:4]
    c = 2.0 * w * w - 1.0
    d = 2.0 * (u * v).sum(-1, keepdim=True)
    return v * c + u * d + torch.cross(u, v, dim=-1) * (2.0 * w)


def _quat_rotate_inv(q, v):
    qc = torch.cat([-q[..., :3], q[..., 3:4]], dim=-1)
    return _quat_rotate(qc, v)


def _quat_mul(a, b):
    ax, ay, az, aw = a.unbind(-1)
    bx, by, bz, bw = b.unbind(-1)
    return torch.stack([aw * bx + bw * ax + ay * bz - az * by,
                        aw * by + bw * ay + az * bx - ax * bz,
                        aw * bz + bw * az + ax * by - ay * bx,
                        aw * bw - ax * bx - ay * by - az * bz], dim=-1)


def _box_corners(h):
    s = torch.tensor([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)],
                     dtype=h.dtype, device=h.device)
    return s * h[..., None, :]  # [..., 8, 3]


def _cylinder_samples(r, hh, k):
    ang = torch.arange(k, dtype=r.dtype, device=r.device) * (2.0 * math.pi / k)
    pts = []
    for y in (-1.0, 0.0, 1.0):
        pts.append(torch.stack([torch.cos(ang) * r[..., None], torch.full_like(ang.expand(r.shape + (k,)), 0.0) + y * hh[..., None],
                                torch.sin(ang) * r[..., None]], dim=-1))
    return torch.cat(pts, dim=-2)  # [..., 3k, 3], axis = local y


HF_CONTACT_MODES = {"point": 1, "cells": 4}


def _closest_on_triangle(p, a, b, c):
    """Closest point of triangle (a, b, c) to p, broadcast over leading dimensions, and
    whether it lies in the face's interior region (Ericson, Real-Time Collision Detection,
    5.1.5: the Voronoi region of p among the triangle's vertices, edges and face, tested
    in that order)."""
    dot = lambda x, y: (x * y).sum(-1, keepdim=True)
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = dot(ab, ap), dot(ac, ap)
    bp = p - b
    d3, d4 = dot(ab, bp), dot(ac, bp)
    cp = p - c
    d5, d6 = dot(ab, cp), dot(ac, cp)
    vc = d1 * d4 - d3 * d2
    vb = d5 * d2 - d1 * d6
    va = d3 * d6 - d5 * d4
    r_a = (d1 <= 0) & (d2 <= 0)
    r_b = (d3 >= 0) & (d4 <= d3)
    r_ab = (vc <= 0) & (d1 >= 0) & (d3 <= 0)
    r_c = (d6 >= 0) & (d5 <= d6)
    r_ac = (vb <= 0) & (d2 >= 0) & (d6 <= 0)
    r_bc = (va <= 0) & (d4 - d3 >= 0) & (d5 - d6 >= 0)
    denom = 1.0 / (va + vb + vc)
    res = a + ab * (vb * denom) + ac * (vc * denom)                           # face
    res = torch.where(r_bc, b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6))), res)
    res = torch.where(r_ac, a + ac * (d2 / (d2 - d6)), res)
    res = torch.where(r_c, c, res)
    res = torch.where(r_ab, a + ab * (d1 / (d1 - d3)), res)
    res = torch.where(r_b, b, res)
    res = torch.where(r_a, a, res)
    face = ~(r_a | r_b | r_ab | r_c | r_ac | r_bc)
    return res, face[..., 0]


class PrimitiveCollider:
    """hf_contacts: "point" (one contact per body sphere: the bilinear surface below its
    centre) or "cells" (up to four per sphere: the four grid cells around the grid vertex
    nearest its centre, each cell two triangles, the contact at the closest point of the
    nearer triangle, as a triangle-mesh narrow phase generates them)."""

    def __init__(self, model, device="cpu", gap=0.02, cylinder_samples=16, lateral_slack=0.06,
                 max_contacts_per_world=128, hf_contacts="point"):
        if hf_contacts not in HF_CONTACT_MODES:
            raise ValueError(f"hf_contacts must be one of {tuple(HF_CONTACT_MODES)}, got {hf_contacts!r}")
        self.hf_contacts = hf_contacts
        self.model = model
        self.device = torch.device(device)
        self.gap = gap
        self.k = cylinder_samples
        self.slack = lateral_slack
        self.max_per_world = max_contacts_per_world
        dt = torch.float32
        self.geo = torch.as_tensor(model.shape_geo, device=self.device)           # [W, SC]
        self.xform = torch.as_tensor(model.shape_xform, dtype=dt, device=self.device)  # [W, SC, 7]
        self.size = torch.as_tensor(model.shape_size, dtype=dt, device=self.device)    # [W, SC, 3]
        self.sbody = torch.as_tensor(model.shape_body_2d, device=self.device)    # [W, SC]
        # candidate pairs from world 0 (uniform topology): (col_a, col_b) with a dynamic
        g0 = model.shape_geo[0]
        b0 = model.shape_body_2d[0]
        SC = g0.shape[0]
        # bodies joined by a joint do not collide with each other (Newton's default
        # parent/child collision filter)
        J = model.joints_per_world
        jp = model.arrays["joint_parent"][:J]
        jc = model.arrays["joint_child"][:J]
        joined = {(int(min(p, c)), int(max(p, c))) for p, c in zip(jp, jc) if p >= 0 and c >= 0}
        pairs = []
        for a in range(SC):
            if b0[a] < 0 or g0[a] == GEO_PLANE:
                continue
            for b in range(SC):
                if b == a or g0[b] not in (GEO_PLANE, GEO_BOX):
                    continue
                if b0[b] >= 0 and b0[b] == b0[a]:
                    continue
                if b0[b] >= 0 and (int(min(b0[a], b0[b])), int(max(b0[a], b0[b]))) in joined:
                    continue
                if b0[b] >= 0:
                    # dynamic-dynamic: test only the shape with the larger column against
                    # the lower one (one direction per pair, as in a stack)
                    if not (a > b):
                        continue
                pairs.append((a, b))
        self.pairs = pairs
        # spheres on bodies against the heightfield (one contact per sphere: the point of
        # the bilinear surface below its centre, with the surface normal there)
        hcols = [c for c in range(SC) if g0[c] == GEO_HEIGHTFIELD]
        self.hf_col = hcols[0] if hcols else -1
        self.hf_spheres = [a for a in range(SC) if b0[a] >= 0 and g0[a] == GEO_SPHERE] if hcols else []
        if self.hf_spheres:
            hf = model.heightfield
            self.hf_h = torch.as_tensor(np.asarray(hf["heights"], np.float32), device=self.device)
            self.hf_origin = tuple(float(v) for v in hf["origin"])
            self.hf_spacing = tuple(float(v) for v in hf["spacing"])

    def _shape_pose(self, body_q, col):
        """World pose [W,7] of shape column `col` for every world."""
        W = self.model.num_worlds
        B = self.model.bodies_per_world
        X = self.xform[:, col]
        bidx = self.sbody[:, col]
        if int(bidx[0]) < 0:
            return X
        bq = body_q.view(W, B, 7)[torch.arange(W, device=self.device), bidx.long()]
        p = bq[:, :3] + _quat_rotate(bq[:, 3:], X[:, :3])
        q = _quat_mul(bq[:, 3:], X[:, 3:])
        return torch.cat([p, q], dim=-1)

    def _body_pose(self, body_q, col):
        W, B = self.model.num_worlds, self.model.bodies_per_world
        bidx = self.sbody[:, col]
        if int(bidx[0]) < 0:
            out = torch.zeros(W, 7, dtype=body_q.dtype, device=self.device)
            out[:, 6] = 1.0
            return out
        return body_q.view(W, B, 7)[torch.arange(W, device=self.device), bidx.long()]

    def _samples(self, col):
        """Sample points of shape `col` in its shape frame, [W, P, 3]."""
        geo = int(self.model.shape_geo[0, col])
        size = self.size[:, col]
        if geo == GEO_BOX:
            return _box_corners(size)
        if geo == GEO_CYLINDER:
            return _cylinder_samples(size[:, 0], size[:, 1], self.k)
        if geo == GEO_SPHERE:
            return torch.zeros(size.shape[0], 1, 3, dtype=size.dtype, device=self.device)
        raise ValueError(f"no samples for geometry {geo}")

    def heightfield_surface(self, x, y):
        """Bilinear height and its gradient at world (x, y) (clamped to the grid)."""
        h = self.hf_h
        ny, nx = h.shape
        u = ((x - self.hf_origin[0]) / self.hf_spacing[0]).clamp(0.0, nx - 1.0)
        v = ((y - self.hf_origin[1]) / self.hf_spacing[1]).clamp(0.0, ny - 1.0)
        i = u.floor().long().clamp(max=nx - 2)
        j = v.floor().long().clamp(max=ny - 2)
        fu, fv = u - i, v - j
        h00, h10, h01, h11 = h[j, i], h[j, i + 1], h[j + 1, i], h[j + 1, i + 1]
        z = h00 * (1 - fu) * (1 - fv) + h10 * fu * (1 - fv) + h01 * (1 - fu) * fv + h11 * fu * fv
        dzdx = ((h10 - h00) * (1 - fv) + (h11 - h01) * fv) / self.hf_spacing[0]
        dzdy = ((h01 - h00) * (1 - fu) + (h11 - h10) * fu) / self.hf_spacing[1]
        return z, dzdx, dzdy

    def _heightfield_cell_contacts(self, c, r):
        """Contacts of sphere centres c [W,Ns,3] (radii r [W,Ns]) with the four grid cells
        around each centre's nearest grid vertex: [W,Ns,4] point on the terrain, unit normal
        from the terrain to the sphere, signed distance, cell-in-grid mask."""
        h = self.hf_h
        ny, nx = h.shape
        sx, sy = self.hf_spacing
        ox, oy = self.hf_origin
        u = (c[..., 0] - ox) / sx
        v = (c[..., 1] - oy) / sy
        I = torch.floor(u + 0.5).long()
        J = torch.floor(v + 0.5).long()
        di = torch.tensor([0, 1, 0, 1], device=self.device)
        dj = torch.tensor([0, 0, 1, 1], device=self.device)
        ci = I[..., None] - 1 + di                                                 # [W,Ns,4]
        cj = J[..., None] - 1 + dj
        inside = (ci >= 0) & (ci <= nx - 2) & (cj >= 0) & (cj <= ny - 2)
        i0, j0 = ci.clamp(0, nx - 2), cj.clamp(0, ny - 2)
        def vert(i, j):
            return torch.stack([ox + i.to(torch.float32) * sx, oy + j.to(torch.float32) * sy, h[j, i]], dim=-1)
        P00, P10, P01, P11 = vert(i0, j0), vert(i0 + 1, j0), vert(i0, j0 + 1), vert(i0 + 1, j0 + 1)
        cc = c[..., None, :].expand_as(P00)
        rr = r[..., None]

        def tri(A, B_, C_):
            """contact with one triangle: foot, normal, distance, usable"""
            nt = torch.cross(B_ - A, C_ - A, dim=-1)
            nt = nt / nt.norm(dim=-1, keepdim=True)
            q, face = _closest_on_triangle(cc, A, B_, C_)
            e = cc - q
            side = (e * nt).sum(-1)
            d = e.norm(dim=-1)
            # inside the face region the contact is the face's (signed: the centre may be
            # below it); beside it, the nearest edge or vertex, for a centre in front of
            # the face only (behind it, a neighbouring triangle holds the contact), along
            # the face normal when the centre is within r/4 of that edge or vertex (the
            # direction to it is then dominated by rounding)
            edge_ok = (side > 0) & (d > 1e-6)
            flat = face | (d <= 0.25 * rr)
            nrm = torch.where(flat[..., None], nt, e / d.clamp_min(1e-6)[..., None])
            dist = torch.where(flat, side, d) - rr
            return q, nrm, dist, face | edge_ok, face

        qa, na, da, ua, fa = tri(P00, P10, P11)
        qb, nb, db, ub, fb = tri(P00, P11, P01)
        # a face contact wins over an edge or vertex one; between two of a kind the nearer,
        # near-ties going to the first triangle so that rounding does not pick between two
        # distinct feet
        use_a = ua & (~ub | (fa & ~fb) | ((fa == fb) & (da <= db + 1e-6)))
        foot = torch.where(use_a[..., None], qa, qb)
        nrm = torch.where(use_a[..., None], na, nb)
        dist = torch.where(use_a, da, db)
        return foot, nrm, dist, inside & (ua | ub)

    def _heightfield_contacts(self, body_q):
        W, B = self.model.num_worlds, self.model.bodies_per_world
        cols = torch.as_tensor(self.hf_spheres, device=self.device)
        bidx = self.sbody[:, cols].long()                                          # [W,Ns]
        bq = body_q.view(W, B, 7)[torch.arange(W, device=self.device)[:, None], bidx]  # [W,Ns,7]
        c = bq[..., :3] + _quat_rotate(bq[..., 3:], self.xform[:, cols, :3])
        r = self.size[:, cols, 0]                                                   # [W,Ns]
        fa = torch.as_tensor(self.model.flat_shape_id(np.arange(W)[:, None], np.array(self.hf_spheres)[None, :]),
                             device=self.device)
        fb = torch.as_tensor(self.model.flat_shape_id(np.arange(W), np.full(W, self.hf_col)), device=self.device)
        if self.hf_contacts == "cells":
            p1w, nrm, dist, inside = self._heightfield_cell_contacts(c, r)        # [W,Ns,4]
            cq = c[..., None, :].expand_as(p1w)
            pw = cq - nrm * r[..., None, None]
            p0 = _quat_rotate_inv(bq[..., None, 3:].expand(*p1w.shape[:-1], 4), pw - bq[..., None, :3])
            Ns = len(self.hf_spheres)
            flat = lambda t: t.reshape(W, Ns * 4, *t.shape[3:])
            s0 = fa[..., None].expand(W, Ns, 4)
            return (flat(p0), flat(p1w), flat(-nrm), flat(s0), fb[:, None].expand(W, Ns * 4),
                    flat(inside & (dist < self.gap)))
        z, gx, gy = self.heightfield_surface(c[..., 0], c[..., 1])
        nrm = torch.stack([-gx, -gy, torch.ones_like(z)], dim=-1)
        nrm = nrm / nrm.norm(dim=-1, keepdim=True)
        ground = torch.stack([c[..., 0], c[..., 1], z], dim=-1)
        dist = ((c - ground) * nrm).sum(-1) - r
        pw = c - nrm * r[..., None]
        p1w = pw - nrm * dist[..., None]
        p0 = _quat_rotate_inv(bq[..., 3:], pw - bq[..., :3])
        return p0, p1w, -nrm, fa, fb[:, None].expand_as(fa), dist < self.gap

    def collide(self, body_q):
        """body_q: [W*B, 7] tensor on the collider device.  Returns a dict of flat
        Newton-format contact tensors (count[1], point0, point1, normal, shape0, shape1,
        margin0, margin1) with max_contacts = W * max_contacts_per_world."""
        body_q = body_q.to(self.device, torch.float32)
        W = self.model.num_worlds
        out_p0, out_p1, out_n, out_s0, out_s1, out_valid = [], [], [], [], [], []
        for (a, b) in self.pairs:
            geo_a = int(self.model.shape_geo[0, a])
            geo_b = int(self.model.shape_geo[0, b])
            Xa = self._shape_pose(body_q, a)
            Ba = self._body_pose(body_q, a)
            Xb = self._shape_pose(body_q, b)
            Bb = self._body_pose(body_q, b)
            if geo_b == GEO_PLANE:
                nrm = _quat_rotate(Xb[:, 3:], torch.tensor([0.0, 0.0, 1.0], device=self.device).expand(W, 3))
                if geo_a == GEO_CYLINDER:
                    # lowest point of the middle circle
                    ax = _quat_rotate(Xa[:, 3:], torch.tensor([0.0, 1.0, 0.0], device=self.device).expand(W, 3))
                    d = -(nrm - (nrm * ax).sum(-1, keepdim=True) * ax)
                    d = d / d.norm(dim=-1, keepdim=True).clamp_min(1e-9)
                    pw = (Xa[:, :3] + d * self.size[:, a, 0:1])[:, None, :]
                elif geo_a == GEO_SPHERE:
                    pw = (Xa[:, :3] - nrm * self.size[:, a, 0:1])[:, None, :]
                else:
                    loc = self._samples(a)
                    pw = Xa[:, None, :3] + _quat_rotate(Xa[:, None, 3:], loc)
                dist = ((pw - Xb[:, None, :3]) * nrm[:, None, :]).sum(-1)          # [W,P]
                p1w = pw - dist[..., None] * nrm[:, None, :]
                n_out = nrm[:, None, :].expand_as(pw)
                valid = dist < self.gap
            else:  # box
                hb = self.size[:, b]                                                  # [W,3]
                if geo_a == GEO_SPHERE:
                    loc = self._samples(a)
                else:
                    loc = self._samples(a)
                pw = Xa[:, None, :3] + _quat_rotate(Xa[:, None, 3:], loc)
                q = _quat_rotate_inv(Xb[:, None, 3:], pw - Xb[:, None, :3])         # [W,P,3]
                excess = q.abs() - hb[:, None, :]
                if int(self.model.shape_body_2d[0, b]) >= 0:
                    # dynamic box below: face chosen from the relative centre position
                    dc = _quat_rotate_inv(Xb[:, 3:], Xa[:, :3] - Xb[:, :3])
                    ha = self.size[:, a]
                    score = dc.abs() / (hb + ha.norm(dim=-1, keepdim=True))
                    axis = score.argmax(-1)[:, None].expand(q.shape[:-1])
                else:
                    axis = excess.argmax(-1)
                onehot = torch.nn.functional.one_hot(axis, 3).to(q.dtype)
                sgn = torch.where(q >= 0, 1.0, -1.0)
                ex_n = (excess * onehot).sum(-1)
                lat_ok = ((excess <= self.slack) | (onehot > 0)).all(-1)
                qp = torch.maximum(torch.minimum(q, hb[:, None, :]), -hb[:, None, :])
                qp = qp * (1 - onehot) + onehot * sgn * hb[:, None, :]
                n_loc = onehot * sgn
                if geo_a == GEO_SPHERE:
                    r = self.size[:, a, 0][:, None]
                    ex_n = ex_n - r
                    pw = pw - _quat_rotate(Xb[:, None, 3:], n_loc) * r[..., None]
                dist = ex_n
                n_out = _quat_rotate(Xb[:, None, 3:], n_loc)
                p1w = Xb[:, None, :3] + _quat_rotate(Xb[:, None, 3:], qp)
                valid = (dist < self.gap) & lat_ok
            # frames of the owning bodies
            p0 = _quat_rotate_inv(Ba[:, None, 3:], pw - Ba[:, None, :3])
            p1 = _quat_rotate_inv(Bb[:, None, 3:], p1w - Bb[:, None, :3])
            P = pw.shape[1]
            out_p0.append(p0)
            out_p1.append(p1)
            out_n.append(-n_out)  # Newton: from shape0 (a) to shape1 (b)
            fa = torch.as_tensor(self.model.flat_shape_id(np.arange(W), np.full(W, a)), device=self.device)
            fb = torch.as_tensor(self.model.flat_shape_id(np.arange(W), np.full(W, b)), device=self.device)
            out_s0.append(fa[:, None].expand(W, P))
            out_s1.append(fb[:, None].expand(W, P))
            out_valid.append(valid)
        if self.hf_spheres:
            p0, p1, n_neg, s0, s1, valid = self._heightfield_contacts(body_q)
            out_p0.append(p0)
            out_p1.append(p1)
            out_n.append(n_neg)
            out_s0.append(s0)
            out_s1.append(s1)
            out_valid.append(valid)
        p0 = torch.cat(out_p0, 1)
        p1 = torch.cat(out_p1, 1)
        nn = torch.cat(out_n, 1)
        s0 = torch.cat(out_s0, 1)
        s1 = torch.cat(out_s1, 1)
        valid = torch.cat(out_valid, 1)
        idx = valid.reshape(-1).nonzero().squeeze(-1)
        n = int(idx.numel())
        cap = W * self.max_per_world
        n_keep = min(n, cap)
        idx = idx[:n_keep]
        def take(t, width):
            flat = t.reshape(-1, width) if width > 1 else t.reshape(-1)
            buf = torch.zeros((cap, width) if width > 1 else (cap,), dtype=flat.dtype, device=self.device)
            buf[:n_keep] = flat[idx]
            return buf
        return dict(
            count=torch.tensor([n_keep], dtype=torch.int32, device=self.device),
            max_contacts=cap,
            point0=take(p0, 3).float().contiguous(),
            point1=take(p1, 3).float().contiguous(),
            normal=take(nn, 3).float().contiguous(),
            shape0=take(s0.to(torch.int32), 1).contiguous(),
            shape1=take(s1.to(torch.int32), 1).contiguous(),
            margin0=torch.zeros(cap, dtype=torch.float32, device=self.device),
            margin1=torch.zeros(cap, dtype=torch.float32, device=self.device),
        )


class HipCollider:
    """PrimitiveCollider's contact generation as gfx950 kernels (`ofb_collide`).

    Same candidate pairs, sample points, tests and (world, pair, point) order as
    PrimitiveCollider; the contacts are written into persistent device tensors (one
    `Contacts` object reused by every call), so a graph captured over `collide` and the
    solver step replays both.  Requires the HIP library and a GPU device."""

    def __init__(self, model, device="cuda:0", gap=0.02, cylinder_samples=16, lateral_slack=0.06,
                 max_contacts_per_world=128, hf_contacts="point"):
        import ctypes as C

        from . import _abi, _lib
        from .containers import Contacts
        self._lib = _lib.load()
        self.model = model
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("HipCollider runs on a GPU device")
        plan = PrimitiveCollider(model, device="cpu", gap=gap, cylinder_samples=cylinder_samples,
                                 lateral_slack=lateral_slack, max_contacts_per_world=max_contacts_per_world,
                                 hf_contacts=hf_contacts)
        self.pairs = plan.pairs
        self.hf_spheres = plan.hf_spheres
        W, SC = model.shape_geo.shape
        # host arrays the description points to (copied by ofb_collider_create)
        keep = dict(
            pairs=np.ascontiguousarray(np.array(self.pairs, np.int32).reshape(-1, 2)),
            geo=np.ascontiguousarray(model.shape_geo, np.int32),
            body=np.ascontiguousarray(model.shape_body_2d, np.int32),
            xform=np.ascontiguousarray(model.shape_xform, np.float32),
            size=np.ascontiguousarray(model.shape_size, np.float32),
            flat=np.ascontiguousarray(model.flat_shape_id(np.arange(W)[:, None], np.arange(SC)[None, :]), np.int32),
            hfs=np.ascontiguousarray(np.array(self.hf_spheres, np.int32)),
        )
        ptr = lambda a: a.ctypes.data if a.size else None
        d = _abi.OfbColliderDesc()
        d.num_worlds, d.bodies_per_world, d.shape_cols = W, model.bodies_per_world, SC
        d.num_pairs = len(self.pairs)
        d.pairs, d.shape_geo, d.shape_body = ptr(keep["pairs"]), ptr(keep["geo"]), ptr(keep["body"])
        d.shape_xform, d.shape_size, d.flat_shape_id = ptr(keep["xform"]), ptr(keep["size"]), ptr(keep["flat"])
        d.cylinder_samples, d.gap, d.lateral_slack = cylinder_samples, gap, lateral_slack
        d.max_contacts_per_world = max_contacts_per_world
        d.num_hf_spheres = len(self.hf_spheres)
        d.hf_spheres = ptr(keep["hfs"])
        d.hf_col = plan.hf_col
        if self.hf_spheres:
            h = np.ascontiguousarray(np.asarray(model.heightfield["heights"], np.float32))
            keep["hf"] = h
            d.hf_ny, d.hf_nx = h.shape
            d.hf_origin[0], d.hf_origin[1] = plan.hf_origin[0], plan.hf_origin[1]
            d.hf_spacing[0], d.hf_spacing[1] = plan.hf_spacing[0], plan.hf_spacing[1]
            d.hf_heights = h.ctypes.data
            d.hf_points = HF_CONTACT_MODES[hf_contacts]
        self._handle = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self._lib.ofb_collider_create(C.byref(d), C.byref(self._handle)), "ofb_collider_create")
        cap = W * max_contacts_per_world
        self.max_contacts = cap
        z = lambda *s, dt=torch.float32: torch.zeros(*s, dtype=dt, device=self.device)
        self.contacts = Contacts(z(1, dt=torch.int32), z(cap, 3), z(cap, 3), z(cap, 3), z(cap, dt=torch.int32),
                                 z(cap, dt=torch.int32), z(cap), z(cap))
        c = self.contacts
        self._out = _abi.OfbContactBuffers(
            c.rigid_contact_count.data_ptr(), cap, c.rigid_contact_point0.data_ptr(),
            c.rigid_contact_point1.data_ptr(), c.rigid_contact_normal.data_ptr(), c.rigid_contact_shape0.data_ptr(),
            c.rigid_contact_shape1.data_ptr(), c.rigid_contact_margin0.data_ptr(), c.rigid_contact_margin1.data_ptr())

    def collide(self, body_q):
        """Contacts of body_q [W*B, 7] (float32, on the collider's device), written into
        `self.contacts` on the current stream; returns that Contacts object."""
        import ctypes as C

        from . import _lib
        W, B = self.model.num_worlds, self.model.bodies_per_world
        if (body_q.device != self.device or body_q.dtype != torch.float32 or not body_q.is_contiguous()
                or body_q.numel() != W * B * 7):
            raise ValueError(f"body_q must be a contiguous float32 [{W * B}, 7] tensor on {self.device}")
        stream = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check(self._lib.ofb_collide(self._handle, C.c_void_p(body_q.data_ptr()), C.byref(self._out), stream),
                   "ofb_collide")
        return self.contacts

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self._lib.ofb_collider_destroy(h)
            self._handle = None
