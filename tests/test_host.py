"""Host side of the library against the oracle: grids, taps, receivers,
signals and transfer functions must be identical (bit-exact for the integer
layout and for the reference-order float arithmetic)."""
import math

import numpy as np
import pytest

import oracle as O
from wavey_amd import configs, fdtd


def same_grid(cfg_room, c, dx):
    gp = fdtd.FdtdGrid.build_box(cfg_room, c, dx)
    go = O.build_box(cfg_room, c, dx)
    assert gp.n == go.n
    assert gp.origin == go.origin and gp.dx == go.dx and gp.dt == go.dt
    assert np.array_equal(gp.flags, go.flags)
    assert np.array_equal(gp.b_cell, go.b_cell)
    assert np.array_equal(gp.b_start, go.b_start)
    assert np.array_equal(gp.slots, go.slots)
    return gp, go


def test_shoebox_grid():
    cfg = configs.shoebox(fdtd.fit_real_impedance)
    gp, _ = same_grid(cfg["room"], cfg["c"], cfg["dx"])
    assert gp.n == (42, 32, 27) and len(gp.b_cell) == 5528


def test_pillared_room_grid():
    cfg = configs.pillared_room()
    gp, _ = same_grid(cfg["room"], cfg["c"], cfg["dx"])
    assert gp.n == (302, 227, 102)
    assert len(cfg["room"]["pillars"]) == 2
    # 6 surface groups of 3 sections: corners carry up to 3 groups = 9 slots.
    assert np.diff(gp.b_start).max() == 9


def test_hall_grid_coarse():
    cfg = configs.hall(dx=0.2)
    gp, _ = same_grid(cfg["room"], cfg["c"], 0.2)
    assert len(cfg["room"]["pillars"]) == 8 and len(cfg["room"]["fits"]) == 10


def test_hall_dimensions():
    cfg = configs.hall()
    assert configs.grid_dims(cfg["room"]["L"], cfg["dx"]) == (1002, 627, 377)


def test_mix_room_grids():
    for cfg in configs.room_mix(n_rooms=6):
        same_grid(cfg["room"], cfg["c"], 0.1)


def test_random_rooms_with_patches():
    rng = np.random.default_rng(42)
    for trial in range(6):
        L = tuple(float(v) for v in rng.uniform(1.0, 3.5, 3))
        dx = float(rng.uniform(0.06, 0.15))
        pil = configs._place_pillars(rng, L, dx, int(rng.integers(0, 3)), (0.2, 0.5), margin=0.3, gap=0.2)
        fits = [(float(rng.uniform(0.01, 0.3)), [])] + [configs.first_order_fit(rng) for _ in range(3)]
        fits.append(fits[1])  # an equal fit under another index merges
        patches = []
        for s in range(6 + 4 * len(pil)):
            gu, gv = int(rng.integers(1, 4)), int(rng.integers(1, 4))
            patches.append((gu, gv, rng.integers(0, len(fits), gu * gv).tolist()))
        same_grid({"L": L, "pillars": pil, "fits": fits, "patches": patches}, 343.0, dx)


@pytest.fixture(scope="module")
def grids():
    cfg = configs.pillared_room()
    room = cfg["room"]
    return fdtd.FdtdGrid.build_box(room, 343.0, 0.08), O.build_box(room, 343.0, 0.08)


def test_nearest_air_and_taps(grids):
    gp, go = grids
    rng = np.random.default_rng(9)
    for _ in range(200):
        p = rng.uniform([-0.5, -0.5, -0.5], [12.5, 9.5, 4.5])
        assert gp.nearest_air(p) == O.nearest_air(go, p)
        tp, to = gp.taps_at(p), O.taps_at(go, p)
        assert tp == to


def test_source_components(grids):
    gp, go = grids
    rng = np.random.default_rng(10)
    air = np.nonzero(gp.flags & 64)[0]
    for c in rng.choice(air, 300):
        for order in (0, 1):
            assert fdtd.source_components(gp, int(c), order) == O.source_components(go, int(c), order)


def test_source_signal_bitwise():
    for f_max, dt, n in ((250.0, configs.dt_of(0.1), 2000), (600.0, 9.6e-5, 500), (1600.0, 3.3e-5, 3000)):
        assert np.array_equal(fdtd.source_signal(f_max, dt, n), O.source_signal(f_max, dt, n))


def test_to_transfer_and_weights():
    rng = np.random.default_rng(12)
    cfg = configs.shoebox(fdtd.fit_real_impedance)
    gp = fdtd.FdtdGrid.build_box(cfg["room"], 343.0, 0.1)
    go = O.build_box(cfg["room"], 343.0, 0.1)
    sig = O.source_signal(250.0, gp.dt, 700)
    freqs = [40.0, 95.5, 250.0]
    spec = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    h0 = rng.normal(size=3) + 1j * rng.normal(size=3)
    a = fdtd.to_transfer(spec, sig, gp, freqs, 343.0, 1.3, h0)
    b = O.to_transfer(spec, sig, go, freqs, 343.0, 1.3, h0)
    assert np.allclose(a, b, rtol=1e-13, atol=0)
    a0, av = 0.5 + 0.1j, [0.2 - 0.3j, 0.1j, -0.4 + 0j]
    assert np.allclose(fdtd.component_weights(a0, av, 1.7, 0.04), O.component_weights(a0, av, 1.7, 0.04),
                       rtol=1e-15, atol=0)


def test_materials():
    rng = np.random.default_rng(14)
    for alpha in (0.01, 0.2, 0.5, 0.94, 0.99):
        assert fdtd.fit_real_impedance(alpha) == O.fit_real_impedance(alpha)
    for _ in range(10):
        fit = configs.first_order_fit(rng)
        fit = (float(rng.uniform(0, 0.2)), fit[1] + [(1.0, 0.0, 2000.0)])
        for f in (5.0, 120.0, 700.0):
            assert abs(fdtd.admittance(fit, f) - O.admittance(fit, f)) < 1e-14 * abs(O.admittance(fit, f))
        beta = complex(rng.uniform(0, 1), rng.uniform(-1, 1))
        assert abs(fdtd.paris(beta) - O.paris(beta)) < 1e-15


def test_first_order_fit_meets_target():
    # Config convention: the residue scale puts the mean Paris absorption over
    # 32 log-spaced points in 20-500 Hz on the drawn target.
    rng = np.random.default_rng(15)
    for _ in range(5):
        state = rng.bit_generator.state
        fit = configs.first_order_fit(rng)
        rng2 = np.random.default_rng()
        rng2.bit_generator.state = state
        rng2.uniform(size=3)
        rng2.uniform(size=3)
        target = rng2.uniform(0.05, 0.6)
        fq = 20.0 * (500.0 / 20.0) ** (np.arange(32) / 31.0)
        alpha = np.mean([O.paris(O.admittance(fit, f)) for f in fq])
        assert abs(alpha - target) < 1e-9
        for r, m, k in fit[1]:
            assert r > 0 and m > 0 and k == 0.0
            assert 2 * math.pi * 20 * 0.999 < r / m < 2 * math.pi * 500 * 1.001


def test_settings_functions():
    """fdtd.rs:44-82, 141-176: crossover weight, f_max, grid_fmax, headroom,
    phase error against their closed forms and the oracle."""
    L = fdtd.lib()
    assert fdtd.crossover_weight(250.0, 100.0) == 1.0 and fdtd.crossover_weight(250.0, 176.0) == 1.0
    assert fdtd.crossover_weight(250.0, 354.0) == 0.0 and fdtd.crossover_weight(250.0, 1000.0) == 0.0
    assert abs(fdtd.crossover_weight(250.0, 250.0) - 0.5) < 1e-15
    for f in [190.0, 220.0, 280.0, 330.0]:
        x = math.log2(f / 250.0)
        assert abs(fdtd.crossover_weight(250.0, f) - 0.5 * (1 + math.cos(math.pi * (x + 0.5)))) < 1e-15
        # Power complementary with the geometric side mirrored about the crossover.
        assert abs(fdtd.crossover_weight(250.0, f) + fdtd.crossover_weight(250.0, 250.0 ** 2 / f) - 1.0) < 1e-12
    assert L.wv_settings_f_max(250.0) == 250.0 * math.sqrt(2.0)
    assert [L.wv_accuracy_headroom(a) for a in range(3)] == [1.0, 1.5, 2.0]
    assert [L.wv_accuracy_cabinet_cells(a) for a in range(3)] == [4.0, 6.0, 8.0]
    freqs = [50.0, 100.0, 150.0, 300.0, 400.0]
    # Highest frequency with weight > 0 is 300 Hz; times 1.5, capped at 250 sqrt 2.
    assert fdtd.grid_fmax(250.0, freqs, "standard") == min(300.0 * 1.5, 250.0 * math.sqrt(2.0))
    assert fdtd.grid_fmax(250.0, [50.0, 100.0], "draft") == 100.0
    assert fdtd.grid_fmax(250.0, [500.0], "high") == 250.0 * math.sqrt(2.0)
    for f, dx in [(63.0, 0.04), (250.0, 0.04), (125.0, 0.1)]:
        e = fdtd.phase_error_deg_per_m(f, 343.0, dx)
        assert e == O.phase_error_deg_per_m(f, 343.0, dx)
    # The scheme's axial dispersion is second order in dx.
    e1, e2 = fdtd.phase_error_deg_per_m(100.0, 343.0, 0.04), fdtd.phase_error_deg_per_m(100.0, 343.0, 0.02)
    assert 3.9 < e1 / e2 < 4.1


def _corner_slots(fits, builder):
    """Slots of the node at cell (1, 3, 1) of a 1.0 x 0.8 x 0.6 m room at dx
    0.1 m, whose missing faces are the wall x = 0 (fit 0) and the floor (fit
    1); the other surfaces take fit 2."""
    from wavey_amd import meshes

    room = {"L": (1.0, 0.8, 0.6), "pillars": [], "fits": fits,
            "patches": [(1, 1, [0]), (1, 1, [2]), (1, 1, [2]), (1, 1, [2]), (1, 1, [1]), (1, 1, [2])]}
    if builder == "box":
        g = fdtd.FdtdGrid.build_box(room, 343.0, 0.1)
        og = O.build_box(room, 343.0, 0.1)
    else:
        mesh = meshes.box_room(room)
        g = fdtd.FdtdGrid.build_room(mesh, 343.0, 0.1)
        og = O.build_room(mesh, 343.0, 0.1)
    for name in ("flags", "b_cell", "b_start", "slots"):
        assert np.array_equal(getattr(g, name), getattr(og, name)), name
    b = int(np.flatnonzero(g.b_cell == g.index(1, 3, 1))[0])
    return g.slots[g.b_start[b]:g.b_start[b + 1]]


@pytest.mark.parametrize("builder", ["box", "mesh"])
def test_faces_merge_only_where_whole_fits_are_equal(builder):
    """fdtd.rs:542 groups a node's faces by AdmittanceFit's derived PartialEq
    (impedance_fit.rs:34-44), which compares the band and both errors too:
    equal fits merge into one conductance slot of admittance x 2 (inv_d = 2
    beta0, fdtd.rs:552-553, 561-562), value-equal fits with different
    metadata stay two slots of inv_d = beta0, and a fit holding a NaN equals
    no fit, itself included."""
    beta = 0.3
    other = (0.05, [])
    meta = {"band": (20.0, 300.0), "rms_error": 0.01, "max_error": 0.02}
    merged = _corner_slots([(beta, [], meta), (beta, [], dict(meta)), other], builder)
    assert merged.shape == (1, 4) and merged[0, 0] == 2 * beta
    apart = _corner_slots([(beta, [], meta), (beta, [], dict(meta, max_error=0.03)), other], builder)
    assert apart.shape == (2, 4) and np.all(apart[:, 0] == beta)
    band = _corner_slots([(beta, [], meta), (beta, [], dict(meta, band=(20.0, 250.0))), other], builder)
    assert band.shape == (2, 4)
    # No metadata on either: equal (zeros), merged.
    assert _corner_slots([(beta, []), (beta, []), other], builder).shape == (1, 4)
    # A NaN in the metadata: never equal, even the same fit on both faces.
    nan = dict(meta, rms_error=float("nan"))
    assert _corner_slots([(beta, [], nan), (beta, [], nan), other], builder).shape == (2, 4)
    fits = [(beta, [], nan), (0.7, []), other]
    room_one_fit = _corner_slots([fits[0], fits[0], other], builder)
    assert room_one_fit.shape == (2, 4) and np.all(room_one_fit[:, 0] == beta)
