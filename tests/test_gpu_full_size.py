"""GPU: the production pass shapes at full size, with data that reaches the
receivers, against the FP64 CPU oracle.

* The hall (1002 x 627 x 377, its default FP32 three-step passes with
  64-plane z-chunks over six chunks, four tiles across, 53 row blocks):
  taps next to the floor, beside a pillar on the last plane of the first
  z-chunk, against a wall on the first plane of the third chunk, on the
  last plane of the fourth and under the ceiling, plus the cardioid's
  monopole node; receivers on both sides of the chunk boundaries 64|65,
  128|129, 192|193, 256|257 and 320|321, next to walls, floor, ceiling and
  the pillar, each within about 70 cells of a tap, so the pulse reaches it
  within the 300 steps. One, two and three steps per pass bit-identical;
  FP64 on the device against the oracle (1e-9), FP32 against the oracle
  (1e-4), overall and per receiver; most samples nonzero.
* A tall room of 8.7 Mi cells and 152 planes (302 x 190 x 152, two pillars,
  ten three-section materials): three steps per pass with z-chunks 64, 32
  and the default, two steps and single steps, bit-identical, against the
  oracle. Its 150 air planes span three 64-plane chunks, so the second and
  third chunk and the second bitmap word of a lane are compared with the
  oracle, not only with the device's own single steps.
* The pillared room at its full configuration (302 x 227 x 102, 6000 steps,
  the cardioid's four component runs, 64 receivers): every component in FP64
  (1e-9) and FP32 (1e-4) against the oracle.
* The largest grid the library takes (589.8 M cells, 597.7 M pitched): taps
  and receivers at the far corner, default passes against single steps and
  both precisions against the oracle.

Reference: core/wavey-solver/tests/fdtd.rs:171-204 (the room with a pillar,
GPU against CPU).
"""
import numpy as np
import pytest

import oracle
from wavey_amd import configs, fdtd, workload

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def oracle_grid(g):
    """The library's grid as the oracle's (the same fields; the two builders
    give identical grids, tested on the CPU)."""
    return oracle.Grid(n=tuple(g.n), origin=tuple(g.origin), dx=g.dx, dt=g.dt, flags=g.flags,
                       b_cell=np.ascontiguousarray(g.b_cell, dtype=np.uint32),
                       b_start=np.ascontiguousarray(g.b_start, dtype=np.uint32),
                       slots=np.ascontiguousarray(g.slots, dtype=np.float64).reshape(-1, 4))


def air_cell(g, x, y, z):
    i = g.index(x, y, z)
    assert g.flags[i] & fdtd.AIR, f"cell ({x}, {y}, {z}) is not air"
    return i


def per_receiver(a, b):
    return [rel_l2(a[r], b[r]) for r in range(len(b)) if np.linalg.norm(b[r]) > 0]


HALL_STEPS = 300


def hall_taps_receivers(g, case):
    """Taps and receivers of the hall run. The pulse peaks about 122 steps in
    and travels 0.577 cells per step, so every receiver is within about 70
    cells of a tap."""
    nx, ny, nz = g.n
    assert (nx, ny, nz) == (1002, 627, 377)
    # The -x face of pillar 0 on row y = 283 (the pillar spans y 10.8-11.8 m).
    row = g.flags.reshape(nz, ny, nx)[64, 283, :] & fdtd.AIR
    px0 = int(np.flatnonzero(row[200:] == 0)[0]) + 200 - 1  # last air cell before the pillar
    src = case.runs[0][2][0][0]
    taps = [(src, 1.0),
            (air_cell(g, 500, 300, 1), 0.7),        # next to the floor
            (air_cell(g, px0, 283, 64), -0.9),      # beside the pillar, last plane of chunk 0
            (air_cell(g, 1, 100, 129), 0.8),        # against the wall x = 0, first plane of chunk 2
            (air_cell(g, 700, 450, 256), 0.6),      # last plane of chunk 3
            (air_cell(g, 700, 450, nz - 2), 0.5)]   # under the ceiling, last chunk
    sx, sy, sz = (src % nx, (src // nx) % ny, src // (nx * ny))
    pts = [(sx + 3, sy, sz), (sx, sy - 7, sz + 2), (sx + 10, sy + 10, sz + 10), (sx, sy, 128), (sx, sy, 129),
           (500, 300, 2), (520, 300, 1), (500, 330, 40), (500, 300, 64), (500, 300, 65),
           (px0 - 1, 283, 64), (px0, 283, 65), (px0 - 14, 270, 128), (px0, 300, 129), (px0 + 27, 283, 64),
           (1, 100, 128), (2, 100, 129), (1, 120, 192), (1, 100, 193), (30, 100, 129), (1, 30, 129),
           (720, 470, 256), (700, 450, 257), (700, 430, 192), (700, 450, 193),
           (700, 450, nz - 3), (700, 450, 320), (700, 450, 321), (650, 450, nz - 2)]
    rx = [air_cell(g, *p) for p in pts]
    return taps, rx


@pytest.fixture(scope="module")
def hall():
    cfg = configs.hall()
    case = workload.build_case(cfg, steps=HALL_STEPS)
    taps, rx = hall_taps_receivers(case.grid, case)
    return case, taps, rx


def test_hall_full_size_pass_modes_and_oracle(hall):
    case, taps, rx = hall
    g = case.grid
    sig = case.signal
    print(f"case: hall {g.n}, {HALL_STEPS} steps, {len(taps)} taps, {len(rx)} receivers, fp32 passes "
          f"of 3 (default), 2 and 1, fp64 default", flush=True)
    outs = {}
    with fdtd.Plan(g, "fp32") as plan:
        info = plan.pass_info()
        assert info["steps_per_pass"] == 3 and info["pair_zchunk"] == 64
        outs[3] = plan.run(taps, sig, rx, mode=fdtd.MODE_IR)[0]
        plan.set_pass(2)
        outs[2] = plan.run(taps, sig, rx, mode=fdtd.MODE_IR)[0]
        plan.set_pass(1)
        outs[1] = plan.run(taps, sig, rx, mode=fdtd.MODE_IR)[0]
    with fdtd.Plan(g, "fp64") as plan:
        assert plan.pass_info()["steps_per_pass"] == 3
        d64 = plan.run(taps, sig, rx, mode=fdtd.MODE_IR)[0]
    one = outs[1]
    assert np.all(np.isfinite(one))
    nonzero = np.count_nonzero(one, axis=1)
    print(f"hall: nonzero samples per receiver min {nonzero.min()} of {HALL_STEPS}, "
          f"{np.count_nonzero(one) / one.size:.2f} of all", flush=True)
    assert nonzero.min() >= 20 and np.count_nonzero(one) > 0.4 * one.size
    assert np.array_equal(one, outs[2]) and np.array_equal(one, outs[3])
    assert np.count_nonzero(d64, axis=1).min() >= 20

    print("hall: FP64 oracle run", flush=True)
    ref = oracle.run(oracle_grid(g), taps, sig, rx, threads=0)[0]
    e64, e32 = rel_l2(d64, ref), rel_l2(outs[3], ref)
    w64, w32 = max(per_receiver(d64, ref)), max(per_receiver(outs[3], ref))
    print(f"hall: device FP64 vs oracle {e64:.3e} (worst receiver {w64:.3e}); "
          f"FP32 {e32:.3e} (worst receiver {w32:.3e})", flush=True)
    assert e64 < 1e-9 and w64 < 1e-9
    assert e32 < 1e-4 and w32 < 1e-4


def test_hall_every_component_run(hall):
    """All eight component runs of the hall's wave job (two cardioid sources,
    a monopole and three dipoles each; run.rs:601-661), not only the first:
    each run's series at cells near its source, the plain FP32 run (three
    steps per pass) against the FP64 run of the same scheme (1e-4 relative
    L2), and the same eight runs as one batch, bit-identical to the plain
    FP32 runs."""
    case, _, _ = hall
    g = case.grid
    sig = case.signal
    nx, ny, nz = g.n
    assert len(case.runs) == 8 and {si for si, _, _ in case.runs} == {0, 1}
    rxs = []
    for _, _, taps in case.runs:
        c0 = taps[0][0]
        x, y, z = c0 % nx, (c0 // nx) % ny, c0 // (nx * ny)
        pts = [(x + 3, y, z), (x, y - 5, z + 2), (x + 8, y + 8, z + 8), (x - 12, y + 4, z - 6)]
        rxs.append([air_cell(g, *p) for p in pts])
    plain32, plain64 = [], []
    for k, (si, ci, taps) in enumerate(case.runs):
        print(f"case: hall source {si} component {ci} ({len(taps)} taps), {HALL_STEPS} steps, fp32 and fp64",
              flush=True)
        for prec, out in (("fp32", plain32), ("fp64", plain64)):
            with fdtd.Plan(g, prec) as plan:
                out.append(plan.run(taps, sig, rxs[k], mode=fdtd.MODE_IR)[0])
        err = rel_l2(plain32[k], plain64[k])
        print(f"  FP32 vs FP64 relative L2 {err:.3e}", flush=True)
        assert np.count_nonzero(plain64[k], axis=1).min() >= 100
        assert err < 1e-4
    print("case: hall, the eight runs as one batch, fp32", flush=True)
    with fdtd.Batch("fp32") as b:
        ids = [b.add(g, taps, sig, rxs[k]) for k, (_, _, taps) in enumerate(case.runs)]
        assert b.run() == fdtd.WV_OK
        for k, i in enumerate(ids):
            assert np.array_equal(b.result(i)[0], plain32[k]), f"run {k}"


def tall_room(seed=configs.MASTER_SEED):
    """12 x 7.5 x 6 m at dx 0.04 m: 302 x 190 x 152 cells (8.7 Mi, three steps
    per pass by default), two 0.6 m pillars, ten three-section materials on
    the walls' patches and the pillars."""
    rng = np.random.default_rng([seed, 6])
    L = (12.0, 7.5, 6.0)
    dx = 0.04
    pillars = configs._place_pillars(rng, L, dx, 2, 0.6, margin=1.5, gap=1.0)
    fits = [configs.first_order_fit(rng) for _ in range(10)]
    grids = [(2, 3), (2, 3), (3, 3), (3, 3), (3, 2), (3, 2)]
    patches = [(gu, gv, rng.integers(0, 10, gu * gv).tolist()) for gu, gv in grids]
    patches += [(1, 1, [int(rng.integers(0, 10))]) for _ in range(4 * len(pillars))]
    room = {"L": L, "pillars": pillars, "fits": fits, "patches": patches}
    return {"name": "tall_room", "room": room, "dx": dx, "c": configs.C_AIR, "steps": 600, "f_src": 250.0,
            "sources": [{"pos": np.array([3.1, 2.3, 2.5]), "kind": "omni"}],
            "receivers": [np.array([rng.uniform(0.3, L[0] - 0.3), rng.uniform(0.3, L[1] - 0.3),
                                    rng.uniform(0.3, L[2] - 0.3)]) for _ in range(12)]}


def test_tall_room_multi_chunk_passes_against_oracle():
    cfg = tall_room()
    case = workload.build_case(cfg)
    g = case.grid
    nx, ny, nz = g.n
    assert (nx, ny, nz) == (302, 190, 152) and g.cells >= 8 << 20
    taps = list(case.runs[0][2]) + [(air_cell(g, 40, 30, 64), 0.6), (air_cell(g, 250, 150, 129), -0.4),
                                    (air_cell(g, 150, 1, 100), 0.5)]
    rx = list(case.receivers) + [air_cell(g, x, y, z) for (x, y, z) in
                                 ((40, 30, 65), (41, 30, 64), (40, 31, 128), (250, 150, 128), (250, 149, 129),
                                  (150, 2, 100), (150, 1, 65), (300, 187, 150), (1, 1, 1), (150, 95, 129),
                                  (150, 95, 64), (150, 95, 65))]
    sig = case.signal
    ref = oracle.run(oracle_grid(g), taps, sig, rx, threads=0)[0]
    for prec, tol in (("fp64", 1e-9), ("fp32", 1e-4)):
        with fdtd.Plan(g, prec) as plan:
            assert plan.pass_info()["steps_per_pass"] == 3
            outs = {"default": plan.run(taps, sig, rx, mode=fdtd.MODE_IR)[0]}
            modes = [(3, 64), (3, 32), (2, 32), (2, 16), (1, 0)] if prec != "fp64" else [(3, 32), (3, 16), (2, 16),
                                                                                        (1, 0)]
            for spp, zc in modes:
                plan.set_pass(spp, zc)
                outs[(spp, zc)] = plan.run(taps, sig, rx, mode=fdtd.MODE_IR)[0]
        one = outs[(1, 0)]
        assert np.count_nonzero(one, axis=1).min() >= 20
        for k, o in outs.items():
            assert np.array_equal(one, o), f"{prec} {k} differs from single steps"
        err = rel_l2(one, ref)
        worst = max(per_receiver(one, ref))
        print(f"tall room {prec}: {len(outs)} pass modes identical; vs oracle {err:.3e} "
              f"(worst receiver {worst:.3e})", flush=True)
        assert err < tol and worst < tol


@pytest.fixture(scope="module")
def pillared_full():
    cfg = configs.pillared_room()
    case = workload.build_case(cfg)
    assert case.grid.n == (302, 227, 102) and len(case.signal) == 6000 and len(case.receivers) == 64
    assert len(case.runs) == 4
    return case, oracle_grid(case.grid)


@pytest.mark.parametrize("comp", [0, 1, 2, 3])
def test_pillared_room_full_config_against_oracle(pillared_full, comp):
    case, og = pillared_full
    taps = case.runs[comp][2]
    print(f"case: pillared room 302x227x102, 6000 steps, component {comp} ({len(taps)} taps), "
          f"64 receivers, FP64 oracle", flush=True)
    ref = oracle.run(og, taps, case.signal, case.receivers, threads=0)[0]
    for prec, tol in (("fp64", 1e-9), ("fp32", 1e-4)):
        ir = fdtd.Gpu(prec).run_ir(case.grid, taps, case.signal, case.receivers)
        err = rel_l2(ir, ref)
        worst = max(per_receiver(ir, ref))
        print(f"pillared room component {comp} {prec}: vs oracle {err:.3e} (worst receiver {worst:.3e})",
              flush=True)
        assert err < tol and worst < tol


def test_largest_grid_the_library_takes():
    """The hall's box at dx 0.0295 m: 1358 x 850 x 511 = 589.8 M cells, just
    under MAX_CELLS (600 M), 597.7 M pitched cells (32-bit cell offsets with
    the least headroom any accepted grid leaves). Taps and receivers next to
    the far corner (the last air cell, the largest offsets) and the near one;
    a 20-step pulse over 60 steps. The default three-step passes equal single
    steps bit for bit; FP32 within 1e-4 and FP64 within 1e-9 of the oracle."""
    cfg = configs.hall(dx=0.0295, steps=60)
    g = fdtd.FdtdGrid.build_box(cfg["room"], cfg["c"], cfg["dx"], 0)
    nx, ny, nz = g.n
    assert (nx, ny, nz) == (1358, 850, 511) and g.cells <= 600_000_000
    ex, ey, ez = nx - 3, ny - 3, nz - 3  # the last air cell on each axis
    assert not g.flags[g.index(ex + 1, ey + 1, ez + 1)] & fdtd.AIR
    taps = [(air_cell(g, ex, ey, ez), 1.0), (air_cell(g, ex - 4, ey - 3, ez - 2), -0.6),
            (air_cell(g, 3, 4, 5), 0.8)]
    pts = [(ex, ey, ez), (ex - 1, ey, ez), (ex - 12, ey - 3, ez), (ex, ey - 15, ez - 2), (ex - 6, ey - 6, ez - 20),
           (ex - 20, ey - 10, ez - 5), (1, 1, 1), (12, 4, 5), (3, 20, 9), (5, 5, 25)]
    receivers = [air_cell(g, *p) for p in pts]
    signal = np.zeros(60)
    signal[:20] = np.sin(np.pi * np.arange(1, 21) / 21.0)
    print(f"case: largest grid {g.n} ({g.cells} cells), 60 steps, FP64 oracle", flush=True)
    ref = oracle.run(oracle_grid(g), taps, signal, receivers, threads=0)[0]
    assert np.count_nonzero(ref) > 0.5 * ref.size
    for prec, tol in (("fp32", 1e-4), ("fp64", 1e-9)):
        with fdtd.Plan(g, prec) as plan:
            print(f"case: largest grid, {prec}, pass mode {plan.pass_info()['steps_per_pass']}", flush=True)
            ir, _ = plan.run(taps, signal, receivers, mode=fdtd.MODE_IR)
            plan.set_pass(1)
            print(f"case: largest grid, {prec}, single steps", flush=True)
            one, _ = plan.run(taps, signal, receivers, mode=fdtd.MODE_IR)
        assert np.array_equal(ir, one)
        err = rel_l2(ir, ref)
        worst = max(per_receiver(ir, ref))
        print(f"largest grid {prec}: pass modes identical; vs oracle {err:.3e} (worst receiver {worst:.3e})",
              flush=True)
        assert err < tol and worst < tol
