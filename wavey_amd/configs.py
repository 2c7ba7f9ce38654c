"""Seeded synthetic workloads for the wave solver (box rooms, admittances, sources).

Room specs are plain dicts (see ``FdtdGrid.build_box``):
``{"L", "pillars", "fits", "patches"}``, with one patch entry per surface
(walls x=0, x=L0, y=0, y=L1, floor, ceiling, then each pillar's sides facing
-x, +x, -y, +y).

Workloads (sizes use the reference's padding rule n = ceil(L/dx) + 2,
core/wavey-solver/src/fdtd.rs:295-299, and dt = dx/(sqrt(3) c), :137-138,290):

* ``shoebox``: 4 x 3 x 2.5 m, dx 0.1 m (42 x 32 x 27), 2000 steps, six walls
  with one real admittance of Paris absorption 0.2, an omni Gaussian-derivative
  pulse (250 Hz) at (1.1, 0.9, 1.2) m, 4 receivers >= 0.5 m from the walls.
* ``pillared_room``: 12 x 9 x 4 m, dx 0.04 m (302 x 227 x 102), two 0.6 m
  floor-to-ceiling pillars, 6 surface groups each with a 3-term sum of
  first-order admittance sections, one cardioid (4 component runs), an 8 x 8
  receiver grid at z = 1.2 m, 6000 steps.
* ``hall``: 40 x 25 x 15 m, dx 0.04 m (1002 x 627 x 377), eight 1 m pillars,
  10 materials on surface patches, 2 cardioids, 256 receivers, 12000 steps.
* ``room_mix``: 48 rooms of uneven size at dx 0.05 m with 0-3 pillars, 1-4
  omni speakers each (one solve per speaker), 16 receivers, 0.5 s.
* ``serving_trace``: Poisson arrivals of rooms drawn as in ``room_mix`` with
  cell counts log-uniform in 0.2M-60M and lengths of 0.3-1.0 s.

A first-order section rho/(s + a) of the admittance is the reference's branch
1/(r + s m + k/s) with m = 1/rho, r = a/rho, k = 0
(core/wavey-materials/src/impedance_fit.rs:25-55).
"""
from __future__ import annotations

import math

import numpy as np

C_AIR = 343.0
COURANT = 0.5773502691896257
MASTER_SEED = 20261203


def dt_of(dx, c=C_AIR):
    return COURANT * dx / c


def grid_dims(L, dx):
    return tuple(int(math.ceil(L[k] / dx)) + 2 for k in range(3))


def paris_np(beta, n=256):
    """Statistical absorption of admittance beta (impedance_fit.rs:73-84), vectorised."""
    beta = np.atleast_1d(np.asarray(beta, dtype=np.complex128))
    h = math.pi / 2.0 / n
    th = (np.arange(n) + 0.5) * h
    c = np.cos(th)[None, :]
    r = (c - beta[:, None]) / (c + beta[:, None])
    return ((1.0 - np.abs(r) ** 2) * np.sin(2.0 * th)[None, :] * h).sum(axis=1)


def section_admittance(rho, a, f):
    s = 2j * math.pi * np.asarray(f, dtype=np.float64)
    return sum(r / (s + aa) for r, aa in zip(rho, a))


def first_order_fit(rng, n_sections=3, f_lo=20.0, f_hi=500.0, alpha_range=(0.05, 0.6)):
    """A passive admittance sum_i rho_i/(s + a_i), poles log-uniform in 2 pi [f_lo, f_hi],
    residues positive and scaled by one factor so the mean Paris absorption over
    32 log-spaced frequencies in [f_lo, f_hi] equals a target drawn from alpha_range.
    Returned as a fit: (beta0 = 0, [(r, m, k)])."""
    a = 2.0 * math.pi * np.exp(rng.uniform(math.log(f_lo), math.log(f_hi), n_sections))
    w = rng.uniform(0.2, 1.0, n_sections)
    rho0 = w * a  # each section's low-frequency admittance rho/a ~ w
    target = rng.uniform(*alpha_range)
    fq = f_lo * (f_hi / f_lo) ** (np.arange(32) / 31.0)

    def mean_alpha(scale):
        return float(paris_np(section_admittance(rho0 * scale, a, fq)).mean())

    lo, hi = 1e-6, 1.0
    while mean_alpha(hi) < target and hi < 1e6:
        hi *= 2.0
    for _ in range(100):
        mid = math.sqrt(lo * hi)
        if mean_alpha(mid) < target:
            lo = mid
        else:
            hi = mid
    rho = rho0 * math.sqrt(lo * hi)
    branches = [(float(aa / r), float(1.0 / r), 0.0) for r, aa in zip(rho, a)]
    return (0.0, branches)


def real_fit(alpha, fit_real_impedance):
    """Frequency-independent real admittance of Paris absorption alpha (material.rs:96-111)."""
    return (1.0 / fit_real_impedance(alpha), [])


def uniform_patches(n_pillars, fit_of_surface):
    return [(1, 1, [fit_of_surface(s)]) for s in range(6 + 4 * n_pillars)]


def _place_pillars(rng, L, dx, count, side, margin=1.0, gap=0.5, tries=2000):
    """Up to `count` square pillars of side `side` (or uniform in a (lo, hi)
    tuple), corners on multiples of dx, `margin` from the walls and `gap`
    apart. Corners on cell faces keep every face between two cell centres."""
    out = []
    for _ in range(tries):
        if len(out) == count:
            break
        w = float(rng.uniform(*side)) if isinstance(side, tuple) else float(side)
        n_w = max(1, int(round(w / dx)))
        if L[0] - 2 * margin - n_w * dx <= 0 or L[1] - 2 * margin - n_w * dx <= 0:
            break
        i0 = int(round(rng.uniform(margin, L[0] - margin - n_w * dx) / dx))
        j0 = int(round(rng.uniform(margin, L[1] - margin - n_w * dx) / dx))
        p = (i0 * dx, (i0 + n_w) * dx, j0 * dx, (j0 + n_w) * dx)
        if all(p[1] + gap <= q[0] or q[1] + gap <= p[0] or p[3] + gap <= q[2] or q[3] + gap <= p[2] for q in out):
            out.append(p)
    return out


def _in_pillars(p, pillars, pad=0.0):
    return any(q[0] - pad < p[0] < q[1] + pad and q[2] - pad < p[1] < q[3] + pad for q in pillars)


def _free_point(rng, L, pillars, clear=0.5, pad=0.3):
    for _ in range(10000):
        p = np.array([rng.uniform(clear, L[k] - clear) for k in range(3)])
        if not _in_pillars(p, pillars, pad):
            return p
    raise RuntimeError("no free point")


def _unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def shoebox(fit_real_impedance, seed=MASTER_SEED):
    rng = np.random.default_rng([seed, 0])
    L = (4.0, 3.0, 2.5)
    fit = real_fit(0.2, fit_real_impedance)
    room = {"L": L, "pillars": [], "fits": [fit], "patches": uniform_patches(0, lambda s: 0)}
    rx = [np.array([rng.uniform(0.5, L[k] - 0.5) for k in range(3)]) for _ in range(4)]
    return {"name": "shoebox", "room": room, "dx": 0.1, "c": C_AIR, "steps": 2000, "f_src": 250.0,
            "sources": [{"pos": np.array([1.1, 0.9, 1.2]), "kind": "omni"}], "receivers": rx, "seed": seed}


def pillared_room(seed=MASTER_SEED):
    rng = np.random.default_rng([seed, 1])
    L = (12.0, 9.0, 4.0)
    dx = 0.04
    pillars = _place_pillars(rng, L, dx, 2, 0.6)
    fits = [first_order_fit(rng) for _ in range(6)]
    # Groups: floor, ceiling, x-walls, y-walls, pillar 1, pillar 2.
    group = {0: 2, 1: 2, 2: 3, 3: 3, 4: 0, 5: 1}

    def fit_of(s):
        return group[s] if s < 6 else 4 + (s - 6) // 4

    room = {"L": L, "pillars": pillars, "fits": fits, "patches": uniform_patches(len(pillars), fit_of)}
    src = {"pos": _free_point(rng, L, pillars, clear=1.0, pad=0.5), "kind": "cardioid", "axis": _unit(rng)}
    xs = np.linspace(1.0, L[0] - 1.0, 8)
    ys = np.linspace(1.0, L[1] - 1.0, 8)
    rx = [np.array([x, y, 1.2]) for y in ys for x in xs]
    return {"name": "pillared_room", "room": room, "dx": dx, "c": C_AIR, "steps": 6000, "f_src": 250.0,
            "sources": [src], "receivers": rx, "seed": seed}


def hall(seed=MASTER_SEED, dx=0.04, steps=12000):
    rng = np.random.default_rng([seed, 2])
    L = (40.0, 25.0, 15.0)
    pillars = _place_pillars(rng, L, dx, 8, 1.0, margin=2.0, gap=1.0)
    fits = [first_order_fit(rng) for _ in range(10)]
    grids = [(5, 3), (5, 3), (8, 3), (8, 3), (8, 5), (8, 5)]
    patches = [(gu, gv, rng.integers(0, 10, gu * gv).tolist()) for gu, gv in grids]
    patches += [(1, 1, [int(rng.integers(0, 10))]) for _ in range(4 * len(pillars))]
    room = {"L": L, "pillars": pillars, "fits": fits, "patches": patches}
    srcs = [{"pos": _free_point(rng, L, pillars, clear=2.0, pad=1.0), "kind": "cardioid", "axis": _unit(rng)}
            for _ in range(2)]
    rx = [_free_point(rng, L, pillars, clear=0.5, pad=0.1) for _ in range(256)]
    return {"name": "hall", "room": room, "dx": dx, "c": C_AIR, "steps": steps, "f_src": 250.0,
            "sources": srcs, "receivers": rx, "seed": seed}


def material_library(seed=MASTER_SEED, n=8):
    rng = np.random.default_rng([seed, 3])
    return [first_order_fit(rng) for _ in range(n)]


def mix_room(rng, lib, dx=0.05, dims=None, duration=0.5, n_rx=16):
    L = tuple(float(v) for v in (dims if dims is not None else
                                 (rng.uniform(3, 15), rng.uniform(3, 12), rng.uniform(2.4, 6))))
    n_p = int(rng.integers(0, 4))
    pillars = _place_pillars(rng, L, dx, n_p, (0.3, 0.8), margin=0.8, gap=0.4) if min(L[:2]) > 2.5 else []
    fit_idx = rng.integers(0, len(lib), 6 + 4 * len(pillars))
    used = sorted(set(int(i) for i in fit_idx))
    remap = {u: k for k, u in enumerate(used)}
    room = {"L": L, "pillars": pillars, "fits": [lib[u] for u in used],
            "patches": [(1, 1, [remap[int(i)]]) for i in fit_idx]}
    n_spk = int(rng.integers(1, 5))
    clear = min(0.5, 0.2 * min(L))
    srcs = [{"pos": _free_point(rng, L, pillars, clear=clear, pad=0.2), "kind": "omni"} for _ in range(n_spk)]
    rx = [_free_point(rng, L, pillars, clear=clear, pad=0.1) for _ in range(n_rx)]
    steps = int(math.ceil(duration / dt_of(dx)))
    return {"name": "mix_room", "room": room, "dx": dx, "c": C_AIR, "steps": steps, "f_src": 250.0,
            "sources": srcs, "receivers": rx}


def room_mix(seed=MASTER_SEED, n_rooms=48):
    rng = np.random.default_rng([seed, 4])
    lib = material_library(seed)
    return [mix_room(rng, lib) for _ in range(n_rooms)]


def serving_trace(rate_hz, n_requests, seed=MASTER_SEED, cells_range=(0.2e6, 60e6), dx=0.05):
    """Requests: (arrival time s, room config). Rooms as room_mix, rescaled so
    their padded cell count is log-uniform in cells_range."""
    rng = np.random.default_rng([seed, 5])
    lib = material_library(seed)
    t = 0.0
    out = []
    for _ in range(n_requests):
        t += float(rng.exponential(1.0 / rate_hz))
        want = math.exp(rng.uniform(math.log(cells_range[0]), math.log(cells_range[1])))
        shape = np.array([rng.uniform(3, 15), rng.uniform(3, 12), rng.uniform(2.4, 6)])
        vol_cells = want * dx ** 3
        scale = (vol_cells / np.prod(shape)) ** (1.0 / 3.0)
        dims = np.maximum(shape * scale, 1.0)
        dur = float(rng.uniform(0.3, 1.0))
        out.append((t, mix_room(rng, lib, dx=dx, dims=dims, duration=dur)))
    return out


def cardioid_weights(axis, freqs, dx, c, component_weights):
    """Per-frequency weights of the 4 components of D = 1/2 + 1/2 d.axis
    (a0 = 1/2, a = axis/2; fdtd.rs:812-815, run.rs:625-643)."""
    a0 = complex(0.5, 0.0)
    a = [complex(0.5 * v, 0.0) for v in axis]
    return np.array([component_weights(a0, a, 2.0 * math.pi * f / c, dx) for f in freqs])
