"""Turns a seeded room config into solver inputs with the library's host side.

The wave job of the reference does the same per source
(core/wavey-solver/src/run.rs:593-661): receivers snap to the nearest air
node, an omni source spreads over trilinear taps (taps_at, falling back to
the nearest node), a directional source runs as its monopole and three
dipole components (source_components) combined in frequency; the library's
wv_wave_source_taps makes that choice (run.rs:601-624).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import fdtd


@dataclass
class Case:
    cfg: dict
    grid: fdtd.FdtdGrid
    signal: np.ndarray
    receivers: list
    runs: list      # [(source index, component index, taps)]


def build_case(cfg, threads=0, steps=None, builder="box"):
    """Grid, source signal, receivers and source runs of a room config.
    ``builder``: "box" (the box-room builder), "mesh" or "prism" (the same
    room as triangles through the general-room builder: same grid)."""
    if builder == "box":
        g = fdtd.FdtdGrid.build_box(cfg["room"], cfg["c"], cfg["dx"], threads)
    else:
        from . import meshes

        mesh = meshes.box_room(cfg["room"], prism=(builder == "prism"))
        g = fdtd.FdtdGrid.build_room(mesh, cfg["c"], cfg["dx"], threads=threads)
    T = cfg["steps"] if steps is None else steps
    signal = fdtd.source_signal(cfg["f_src"], g.dt, T)
    receivers = []
    for p in cfg["receivers"]:
        r = g.nearest_air(p)
        if r is None:
            raise fdtd.FdtdError(-1, f"receiver {p} has no air node nearby")
        receivers.append(r)
    runs = []
    for si, s in enumerate(cfg["sources"]):
        # run.rs:601-624: an omni source is its trilinear taps (order 0), a
        # cardioid the nearest node's monopole and dipoles (order 1); no
        # cabinet calibrations here.
        order = 0 if s["kind"] == "omni" else 1
        got = fdtd.wave_source_taps(g, s["pos"], s.get("axis", (1.0, 0.0, 0.0)), 0, order)
        if got is None:
            raise fdtd.FdtdError(-1, f"source {si} is not inside the grid")
        for ci, taps in enumerate(got[0]):
            runs.append((si, ci, taps))
    return Case(cfg=cfg, grid=g, signal=signal, receivers=receivers, runs=runs)


def wave_sources(case, freqs):
    """The wave job's sources of a case (run.rs:611-643): an omni source is
    one run of its taps; a cardioid D = 1/2 + 1/2 d.axis is its four
    component runs with per-frequency weights (fdtd.rs:812-815)."""
    from . import configs

    out = []
    for si, s in enumerate(case.cfg["sources"]):
        comps = [taps for (i, _, taps) in case.runs if i == si]
        src = {"comps": comps, "weights": None, "h0": None, "r_ref": 1.0}
        if s["kind"] != "omni" and len(comps) == 4:
            src["weights"] = configs.cardioid_weights(s["axis"], freqs, case.grid.dx, case.cfg["c"],
                                                      fdtd.component_weights)
        out.append(src)
    return out
