"""Rooms for the grid-builder tests: a patched box with pillars, the swept
rooms of core/wavey-solver/tests/mesh_rooms.rs, a sloped ceiling, an
extruded L-shaped plan with a hole, and turned cuboid cabinets."""
import numpy as np

from wavey_amd import fdtd, meshes


def fit(i):
    """Distinct real and reactive fits, so every surface's slots differ."""
    return (0.01 * (i + 1), [(2.0 + i, 1e-3 * (i + 1), 50.0 + 10 * i)])


def patched_box():
    """Patch edges and pillar sides lie off every cell centre at the spacings
    used here: where a ray meets an edge between two patches exactly, the
    box builder and a triangle mesh settle the tie differently."""
    rng = np.random.default_rng(5)
    pillars = [(1.131, 1.712, 0.873, 1.291), (2.461, 2.834, 1.627, 2.371)]
    fits = [fit(i) for i in range(5)] + [meshes.RIGID]
    grids = [(2, 1), (3, 2), (1, 2), (2, 2), (3, 3), (2, 3)]
    patches = [(gu, gv, rng.integers(0, 6, gu * gv).tolist()) for gu, gv in grids]
    patches += [(1, 2, rng.integers(0, 6, 2).tolist()) for _ in range(4 * len(pillars))]
    return {"L": (3.713, 2.917, 2.339), "pillars": pillars, "fits": fits, "patches": patches}


def stepped_balcony():
    """mesh_rooms.rs:76-78: stepped floor and a balcony slab, swept 5 m."""
    profile = [(0.0, 0.0), (6.0, 0.0), (6.0, 1.0), (10.0, 1.0), (10.0, 4.0), (0.0, 4.0), (0.0, 2.6), (3.0, 2.6),
               (3.0, 2.2), (0.0, 2.2)]
    return meshes.swept(profile, 5.0, [fit(i) for i in range(len(profile) + 2)]), profile


def sloped(A=3.3, B=2.1, h0=2.6, h1=1.7):
    """Floor z=0, walls x=0, x=A, y=0, y=B, a ceiling from z=h0 at x=0 to h1 at x=A."""
    profile = [(0.0, 0.0), (A, 0.0), (A, h1), (0.0, h0)]
    return meshes.swept(profile, B, [fit(i) for i in range(6)]), (A, B, h0, h1)


def cabinet(pos=(1.6, 1.2, 0.9), size=(0.55, 0.4, 0.7), angle=0.47):
    v, t = meshes.cuboid(size)
    R = meshes.rotation((0.3, 0.2, 1.0), angle)
    return fdtd.solid_from_mesh(v, t, R, pos), (R, np.asarray(pos), np.asarray(size))


def l_prism():
    """An extruded L-shaped plan with a square hole (a pillar), height 2.4."""
    outer = [(0.0, 0.0), (3.1, 0.0), (3.1, 1.3), (1.4, 1.3), (1.4, 2.7), (0.0, 2.7)]
    hole = [(0.5, 0.4), (0.5, 0.9), (0.9, 0.9), (0.9, 0.4)]
    H = 2.4
    M = meshes.MeshBuilder()
    for i in range(len(outer)):  # walls, inward for a counter-clockwise plan
        (xa, ya), (xb, yb) = outer[i], outer[(i + 1) % len(outer)]
        M.polygon(M.surface(fit(i)), [(xa, ya, 0.0), (xa, ya, H), (xb, yb, H), (xb, yb, 0.0)][::-1])
    for i in range(len(hole)):  # the hole's sides face out of the hole
        (xa, ya), (xb, yb) = hole[i], hole[(i + 1) % len(hole)]
        M.polygon(M.surface(fit(6)), [(xa, ya, 0.0), (xa, ya, H), (xb, yb, H), (xb, yb, 0.0)][::-1])
    M.polygon(M.surface(fit(7)), [(x, y, 0.0) for x, y in outer])
    M.polygon(M.surface(fit(8)), [(x, y, H) for x, y in outer][::-1])
    return M.room(prism=True, height=H, regions=[[outer, hole]])
