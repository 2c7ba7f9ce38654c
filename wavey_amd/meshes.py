"""Rooms and cabinets as triangle meshes, for ``FdtdGrid.build_room``.

A mesh room is a dict (the fields of ``wv_room_mesh``):

    {"positions": (V, 3), "triangles": (T, 3), "tri_surface": (T,),
     "fits": [(beta0, [(r, m, k), ...]), ...],   # one per surface
     "prism": bool, "height": h, "regions": [[outer (n, 2), hole, ...], ...]}

Triangles are wound so their normals point into the room, as
``room_from_mesh`` leaves them (core/wavey-geom/src/room.rs:425-431); the
grid builder only looks at the winding where two crossings of one column ray
fall together (fdtd.rs:399-403).
"""
from __future__ import annotations

import numpy as np

RIGID = (0.0, [])


class MeshBuilder:
    """Collects triangles, their surfaces and one fit per surface."""

    def __init__(self):
        self.pos, self.tri, self.surf, self.fits = [], [], [], []

    def surface(self, fit):
        self.fits.append(fit)
        return len(self.fits) - 1

    def polygon(self, s, pts):
        """Planar simple polygon of surface s, in winding order."""
        base = len(self.pos)
        self.pos += [tuple(float(c) for c in p) for p in pts]
        for t in triangulate(pts):
            self.tri.append(tuple(base + i for i in t))
            self.surf.append(s)

    def patch_grid(self, fits, patch, u0, u1, v0, v1, point, flip):
        """One surface per patch of a gu x gv grid over [u0,u1] x [v0,v1];
        point(u, v) is the 3D point, the winding runs u then v (reversed when
        flip), patch (iu, iv) takes fits[ids[iv*gu + iu]]."""
        gu, gv, ids = patch
        us = [u0 + (u1 - u0) * i / gu for i in range(gu + 1)]
        vs = [v0 + (v1 - v0) * i / gv for i in range(gv + 1)]
        us[-1], vs[-1] = u1, v1
        for iv in range(gv):
            for iu in range(gu):
                s = self.surface(fits[int(ids[iv * gu + iu])])
                q = [point(us[iu], vs[iv]), point(us[iu + 1], vs[iv]), point(us[iu + 1], vs[iv + 1]),
                     point(us[iu], vs[iv + 1])]
                self.polygon(s, q[::-1] if flip else q)

    def room(self, **extra):
        out = {"positions": np.array(self.pos, dtype=np.float64).reshape(-1, 3),
               "triangles": np.array(self.tri, dtype=np.uint32).reshape(-1, 3),
               "tri_surface": np.array(self.surf, dtype=np.uint32), "fits": list(self.fits), "prism": False}
        out.update(extra)
        return out


def triangulate(pts):
    """Ear clipping of a planar simple polygon; index triples that keep the
    polygon's winding."""
    P = np.asarray(pts, dtype=np.float64)
    n = len(P)
    if n < 3:
        return []
    normal = np.zeros(3)
    for i in range(n):  # Newell's normal
        a, b = P[i], P[(i + 1) % n]
        normal += np.array([(a[1] - b[1]) * (a[2] + b[2]), (a[2] - b[2]) * (a[0] + b[0]),
                            (a[0] - b[0]) * (a[1] + b[1])])
    drop = int(np.argmax(np.abs(normal)))
    q = P[:, [k for k in range(3) if k != drop]]
    area2 = sum(q[i, 0] * q[(i + 1) % n, 1] - q[(i + 1) % n, 0] * q[i, 1] for i in range(n))
    sign = 1.0 if area2 > 0 else -1.0

    def turn(o, a, b):
        return ((a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])) * sign

    idx = list(range(n))
    out = []
    while len(idx) > 3:
        m = len(idx)
        for k in range(m):
            i0, i1, i2 = idx[(k - 1) % m], idx[k], idx[(k + 1) % m]
            if turn(q[i0], q[i1], q[i2]) <= 0:
                continue  # reflex or straight corner
            if any(turn(q[i0], q[i1], q[j]) >= 0 and turn(q[i1], q[i2], q[j]) >= 0 and turn(q[i2], q[i0], q[j]) >= 0
                   for j in idx if j not in (i0, i1, i2)):
                continue  # another vertex inside the ear
            out.append((i0, i1, i2))
            idx.pop(k)
            break
        else:
            raise ValueError("polygon is not simple")
    out.append(tuple(idx))
    return out


def box_room(room, prism=False):
    """The box room with pillars and surface patches that
    ``FdtdGrid.build_box`` takes, as a mesh room: one surface per patch with
    that patch's fit, patches laid out as the box builder splits each
    surface. Each pillar adds four sides and two rigid caps on the floor and
    ceiling, wound against them, so a column ray through the pillar crosses
    four times. ``prism`` gives the extruded form: the plan rectangle with
    the pillars as holes, height L2."""
    L = [float(v) for v in room["L"]]
    fits = room["fits"]
    pat = room["patches"]
    pillars = [tuple(float(v) for v in p) for p in room.get("pillars", [])]
    if len(pat) != 6 + 4 * len(pillars):
        raise ValueError("one patch entry per surface is needed")
    M = MeshBuilder()
    # Surface: (u range, v range, point(u, v), flip) with the normal (u then
    # v, reversed by flip) pointing into the room.
    M.patch_grid(fits, pat[0], 0.0, L[1], 0.0, L[2], lambda u, v: (0.0, u, v), False)   # x = 0, +x
    M.patch_grid(fits, pat[1], 0.0, L[1], 0.0, L[2], lambda u, v: (L[0], u, v), True)   # x = L0, -x
    M.patch_grid(fits, pat[2], 0.0, L[0], 0.0, L[2], lambda u, v: (u, 0.0, v), True)    # y = 0, +y
    M.patch_grid(fits, pat[3], 0.0, L[0], 0.0, L[2], lambda u, v: (u, L[1], v), False)  # y = L1, -y
    M.patch_grid(fits, pat[4], 0.0, L[0], 0.0, L[1], lambda u, v: (u, v, 0.0), False)   # floor, +z
    M.patch_grid(fits, pat[5], 0.0, L[0], 0.0, L[1], lambda u, v: (u, v, L[2]), True)   # ceiling, -z
    for p, (x0, x1, y0, y1) in enumerate(pillars):
        s = 6 + 4 * p
        M.patch_grid(fits, pat[s], y0, y1, 0.0, L[2], lambda u, v: (x0, u, v), True)       # facing -x
        M.patch_grid(fits, pat[s + 1], y0, y1, 0.0, L[2], lambda u, v: (x1, u, v), False)  # facing +x
        M.patch_grid(fits, pat[s + 2], x0, x1, 0.0, L[2], lambda u, v: (u, y0, v), False)  # facing -y
        M.patch_grid(fits, pat[s + 3], x0, x1, 0.0, L[2], lambda u, v: (u, y1, v), True)   # facing +y
        cap = M.surface(RIGID)
        M.polygon(cap, [(x0, y0, 0.0), (x0, y1, 0.0), (x1, y1, 0.0), (x1, y0, 0.0)])  # -z, against the floor
        M.polygon(cap, [(x0, y0, L[2]), (x1, y0, L[2]), (x1, y1, L[2]), (x0, y1, L[2])])  # +z, against the ceiling
    extra = {}
    if prism:
        outer = [(0.0, 0.0), (L[0], 0.0), (L[0], L[1]), (0.0, L[1])]
        holes = [[(x0, y0), (x0, y1), (x1, y1), (x1, y0)] for x0, x1, y0, y1 in pillars]
        extra = {"prism": True, "height": L[2], "regions": [[outer] + holes]}
    return M.room(**extra)


def swept(profile, width, fits):
    """A profile in (x, z), counter-clockwise, swept along y from 0 to
    ``width`` (the rooms of core/wavey-solver/tests/mesh_rooms.rs:17-31):
    one surface per profile edge (fits[i] for edge i), then the end caps at
    y = 0 and y = width (fits[n], fits[n+1])."""
    n = len(profile)
    M = MeshBuilder()
    for i in range(n):
        (xa, za), (xb, zb) = profile[i], profile[(i + 1) % n]
        s = M.surface(fits[i])
        # Normal (-w dz, 0, w dx): the left side of a counter-clockwise profile.
        M.polygon(s, [(xa, 0.0, za), (xb, 0.0, zb), (xb, width, zb), (xa, width, za)])
    s0, s1 = M.surface(fits[n]), M.surface(fits[n + 1])
    M.polygon(s0, [(x, 0.0, z) for x, z in profile][::-1])  # +y
    M.polygon(s1, [(x, width, z) for x, z in profile])       # -y
    return M.room()


CUBOID_TRIS = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
               (1, 2, 6), (1, 6, 5), (2, 3, 7), (2, 7, 6), (3, 0, 4), (3, 4, 7)]


def cuboid(size):
    """A cabinet: the box [-s/2, s/2] per axis, 8 vertices and 12 triangles
    wound outward (the local frame Solid::from_mesh takes, fdtd.rs:433)."""
    hx, hy, hz = (0.5 * float(s) for s in size)
    verts = [(-hx, -hy, -hz), (hx, -hy, -hz), (hx, hy, -hz), (-hx, hy, -hz),
             (-hx, -hy, hz), (hx, -hy, hz), (hx, hy, hz), (-hx, hy, hz)]
    return np.array(verts), np.array(CUBOID_TRIS, dtype=np.uint32)


def rotation(axis, angle):
    """Rotation matrix about ``axis`` by ``angle`` radians (Rodrigues)."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)
