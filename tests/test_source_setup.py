"""CPU: the wave job's source set-up. FdtdGrid::air_along (fdtd.rs:516-533)
and the choice of source node, driver point and component taps of
run.rs:601-624, in the oracle and the library.

Pins of the oracle: air_along's closed form on a grid with a box cabinet
(from inside the cabinet along an axis, the first node past the cabinet's
face on the row of the rounded start; every node of the row is visited,
as the march steps half a cell), its fallback to the nearest air node when
no air lies within 2 m along the axis, and a start in the air returning its
own rounded node; the set-up's cases against the pieces it is made of
(taps_at of the centre or of the baffle point, source_components of the
nearest node, the single-node fallback next to a wall). The library matches
the oracle bit for bit on every case and on random points and axes.
"""
import numpy as np

import oracle
from wavey_amd import fdtd

from test_cabinet import C, cabinet_grid, solid_has


def rnd(v):
    """Rust's f64::round (half away from zero) for v >= 0; Python's round
    takes halves to even."""
    return int(np.floor(v + 0.5))


def grids():
    mesh, solid = cabinet_grid()
    return oracle.build_room(mesh, C, 0.1, [solid]), fdtd.FdtdGrid.build_room(mesh, C, 0.1, [solid]), solid


def test_air_along_closed_form():
    og, g, solid = grids()
    dx = og.dx
    for p in [(1.32, 1.155, 0.85), (1.05, 1.0, 0.6), (1.2, 1.3, 1.1), (1.61, 0.97, 1.16)]:
        for axis in [(1.0, 0.0, 0.0), (-1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0),
                     (0.0, 0.0, -1.0)]:
            k = int(np.argmax(np.abs(axis)))
            s = int(axis[k])
            idx = [rnd((p[i] - og.origin[i]) / dx) for i in range(3)]
            while solid_has(solid, og, idx):
                idx[k] += s
            expect = og.index(*idx)
            assert og.flags[expect] & 64
            got = oracle.air_along(og, p, axis)
            assert got == expect, (p, axis)
            assert g.air_along(p, axis) == got
    # From the air: the rounded node itself.
    p = (0.52, 0.47, 0.55)
    idx = [rnd((p[i] - og.origin[i]) / dx) for i in range(3)]
    assert oracle.air_along(og, p, (0.0, 1.0, 0.0)) == og.index(*idx)


def test_air_along_falls_back_to_the_nearest_air_node():
    og, g, _ = grids()
    # Below the floor, pointing down: no air within 2 m along the axis.
    p = (1.5, 1.1, -0.4)
    assert oracle.air_along(og, p, (0.0, 0.0, -1.0)) == oracle.nearest_air(og, p)
    assert g.air_along(p, (0.0, 0.0, -1.0)) == oracle.nearest_air(og, p)
    # Far outside the grid: none at all.
    assert oracle.air_along(og, (40.0, 40.0, 40.0), (1.0, 0.0, 0.0)) is None
    assert g.air_along((40.0, 40.0, 40.0), (1.0, 0.0, 0.0)) is None


def test_air_along_library_equals_oracle_random():
    og, g, _ = grids()
    rng = np.random.default_rng(5)
    for _ in range(300):
        p = rng.uniform(-0.3, 3.4, 3)
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        assert g.air_along(p, a) == oracle.air_along(og, p, a)


def test_wave_source_taps_cases():
    og, g, solid = grids()
    center = (1.32 + 0.2, 1.155, 0.85)  # inside the cabinet, near its +x face
    axis = (1.0, 0.0, 0.0)
    # Calibrated: the baffle point's taps (some nodes are in the cabinet and
    # drop out).
    comps, fb = oracle.wave_source_taps(og, center, axis, 1, 1)
    assert comps == [oracle.taps_at(og, oracle.baffle_point(og, center, axis))] and not fb
    # Uncalibrated, order 0, in the air: the centre's taps.
    air_pt = (0.61, 0.72, 1.43)
    comps, fb = oracle.wave_source_taps(og, air_pt, axis, 0, 0)
    assert comps == [oracle.taps_at(og, air_pt)] and not fb
    # Order 1 in the air: the nearest node's monopole and dipoles.
    comps, fb = oracle.wave_source_taps(og, air_pt, axis, 0, 1)
    assert comps == oracle.source_components(og, oracle.nearest_air(og, air_pt), 1) and len(comps) == 4 and not fb
    # Order 1 against a wall: the nearest node alone, with the note.
    wall_pt = (0.05, 1.0, 1.0)
    node = oracle.nearest_air(og, wall_pt)
    comps, fb = oracle.wave_source_taps(og, wall_pt, axis, 0, 1)
    assert comps == [[(node, 1.0)]] and fb
    # Uncalibrated order 1 from inside the cabinet: the nearest air node.
    comps, fb = oracle.wave_source_taps(og, center, axis, 0, 1)
    assert comps[0] == [(oracle.nearest_air(og, center), 1.0)]
    # Outside the grid: None.
    assert oracle.wave_source_taps(og, (40.0, 40.0, 40.0), axis, 0, 0) is None
    assert fdtd.wave_source_taps(g, (40.0, 40.0, 40.0), axis, 0, 0) is None
    # The library, bit for bit.
    rng = np.random.default_rng(9)
    pts = [center, air_pt, wall_pt] + [tuple(rng.uniform(0.0, 3.1, 3)) for _ in range(40)]
    for p in pts:
        for cal in (0, 1):
            for order in (0, 1):
                a = rng.normal(size=3)
                a /= np.linalg.norm(a)
                for ax in (axis, tuple(a)):
                    assert fdtd.wave_source_taps(g, p, ax, cal, order) == oracle.wave_source_taps(og, p, ax, cal,
                                                                                                   order), (p, cal)
