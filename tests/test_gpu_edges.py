"""GPU: edge cases of the run against the FP64 oracle and between pass modes.

* no receivers; runs of 1, 2, 3 and 130 steps with one, two and three steps
  per pass;
* cancellation between step blocks and the progress callback
  (gpu_fdtd.rs:202-205, 259);
* a tap on a solid cell and the same cell given twice (the reference injects
  into any cell and adds duplicate taps, fdtd.rs:718-720, gpu_fdtd.rs:67-75),
  with a receiver on the solid cell;
* air up to the grid's x ends (a caller's grid without the builder's padding
  columns), so that cells at x = 1 and x = nx - 2 are interior: every pass
  mode bit-identical and FP64 against the oracle;
* an empty signal, and bad run arguments, through every entry point;
* every single-step block shape and z-chunk bit-identical to the default.
"""
import ctypes as C
import dataclasses

import numpy as np
import pytest

import oracle
from wavey_amd import configs, fdtd, workload

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def shoebox():
    cfg = configs.shoebox(fdtd.fit_real_impedance)
    case = workload.build_case(cfg)
    og = oracle.build_box(cfg["room"], cfg["c"], cfg["dx"])
    return case, og


def test_no_receivers(shoebox):
    case, _ = shoebox
    print("case: shoebox, no receivers", flush=True)
    with fdtd.Plan(case.grid, "fp32") as plan:
        ir, sp = plan.run(case.runs[0][2], case.signal[:300], [], [62.5], mode=fdtd.MODE_BOTH)
    assert ir.shape == (0, 300) and sp.shape == (0, 1)


def test_zero_steps_every_entry_point(shoebox):
    """An empty signal runs no step (gpu_fdtd.rs:200-203): all-zero spectra,
    an empty receiver series and no progress call, through the plan, both
    one-call entry points, a batch beside a run of 200 steps and the server;
    the longer run beside it is unchanged."""
    case, _ = shoebox
    taps, rx = case.runs[0][2], case.receivers
    freqs = [62.5, 125.0]
    empty, sig = case.signal[:0], case.signal[:200]
    R = len(rx)
    print("case: shoebox, zero steps, plan", flush=True)
    calls = []
    with fdtd.Plan(case.grid, "fp32") as plan:
        ir, sp = plan.run(taps, empty, rx, freqs, mode=fdtd.MODE_BOTH, progress=calls.append)
        assert ir.shape == (R, 0) and sp.shape == (R, 2) and np.all(sp == 0)
        ref, _ = plan.run(taps, sig, rx, mode=fdtd.MODE_IR)
    assert calls == []
    assert np.count_nonzero(ref) > 0
    for fn in (fdtd.fdtd_run, fdtd.fdtd_run_noalloc):
        print(f"case: shoebox, zero steps, {fn.__name__}", flush=True)
        ir, sp = fn(case.grid, taps, empty, rx, freqs, fdtd.MODE_BOTH, "fp64", progress=calls.append)
        assert ir.shape == (R, 0) and sp.shape == (R, 2) and np.all(sp == 0)
    assert calls == []
    print("case: shoebox, zero steps, batch beside 200 steps", flush=True)
    with fdtd.Batch("fp32") as b:
        i0 = b.add(case.grid, taps, empty, rx)
        i1 = b.add(case.grid, taps, sig, rx)
        assert b.run() == fdtd.WV_OK
        ir0, sp0 = b.result(i0, freqs, mode=fdtd.MODE_BOTH)
        ir1, _ = b.result(i1)
    assert ir0.shape == (R, 0) and np.all(sp0 == 0)
    assert np.array_equal(ir1, ref)
    print("case: shoebox, zero steps, server beside 200 steps", flush=True)
    with fdtd.Server("fp32") as sv:
        r0 = sv.submit(case.grid, taps, empty, rx)
        r1 = sv.submit(case.grid, taps, sig, rx)
        ir0, _, _ = sv.wait(r0)
        ir1, _, _ = sv.wait(r1)
        st = sv.stats()
    assert ir0.shape == (R, 0)
    assert np.array_equal(ir1, ref)
    assert st["completed"] == 2 and st["failed"] == 0


def test_run_argument_errors(shoebox):
    """Bad run arguments come back as status codes with a message (where the
    reference would panic or return Err), from the plan, the one-call entry
    point, the batch and the server, and the plan runs normally afterwards: a
    tap on the grid's first plane or beyond its last cell, a receiver beyond
    the last cell (WV_ERR_INVALID), and a receiver record of 2^31 steps
    (WV_ERR_LIMIT, refused before the signal is read)."""
    case, _ = shoebox
    g = case.grid
    taps, rx, sig = case.runs[0][2], case.receivers, case.signal[:50]
    cells = g.cells
    bad = [("tap on plane 0", [(g.n[0] + 1, 1.0)], rx, "source tap"),
           ("tap beyond the grid", [(cells, 1.0)], rx, "source tap"),
           ("receiver beyond the grid", taps, [cells], "receiver outside")]
    with fdtd.Plan(g, "fp32") as plan:
        ref, _ = plan.run(taps, sig, rx, mode=fdtd.MODE_IR)
        for name, t, r, msg in bad:
            print(f"case: shoebox, {name}", flush=True)
            for call in (lambda: plan.run(t, sig, r, mode=fdtd.MODE_IR),
                         lambda: fdtd.fdtd_run(g, t, sig, r, (), fdtd.MODE_IR)):
                with pytest.raises(fdtd.FdtdError) as e:
                    call()
                assert e.value.status == fdtd.WV_ERR_INVALID and msg in str(e.value)
            with fdtd.Batch("fp32") as b:
                with pytest.raises(fdtd.FdtdError) as e:
                    b.add(g, t, sig, r)
                assert e.value.status == fdtd.WV_ERR_INVALID
            with fdtd.Server("fp32") as sv:
                with pytest.raises(fdtd.FdtdError) as e:
                    sv.submit(g, t, sig, r)
                assert e.value.status == fdtd.WV_ERR_INVALID
        again, _ = plan.run(taps, sig, rx, mode=fdtd.MODE_IR)
    assert np.array_equal(again, ref) and np.count_nonzero(ref) > 0
    print("case: shoebox, 2^31 steps", flush=True)
    one = np.zeros(1)
    rc = np.ascontiguousarray(rx, dtype=np.uint64)
    err = C.create_string_buffer(256)
    st = fdtd.lib().wv_fdtd_run(C.byref(g.view()), fdtd._taps_array(taps), len(taps),
                                one.ctypes.data_as(C.POINTER(C.c_double)), 1 << 31,
                                rc.ctypes.data_as(C.POINTER(C.c_uint64)), len(rc), None, 0, fdtd.MODE_IR, 0,
                                None, fdtd.PROGRESS_FN(), None, None, None, err, 256)
    assert st == fdtd.WV_ERR_LIMIT and b"too large" in err.value


def test_every_single_step_shape_bit_identical(shoebox):
    """The single-step kernel's block shapes (7, 3, 15 and 1 tiles per block,
    one or four special-cell waves) and z-chunks (1-32 planes) split the
    same cell updates differently over blocks and waves: every shape gives
    the default shape's series bit for bit, FP32 and FP64, on the shoebox
    (the small-grid default: one-tile blocks) and on a room of 0.4 M cells.
    The default is checked against the oracle in the other tests."""
    case, _ = shoebox
    rng = np.random.default_rng(3)
    cfg = configs.mix_room(rng, configs.material_library(), dx=0.05, dims=(3.4, 3.8, 3.3))
    cfg["steps"] = 160
    room = workload.build_case(cfg)
    for name, c, steps in (("shoebox", case, 200), ("room", room, 160)):
        taps, sig, rx = c.runs[0][2], c.signal[:steps], c.receivers
        for prec in ("fp32", "fp64"):
            with fdtd.Plan(c.grid, prec) as plan:
                plan.set_pass(1)
                ref, _ = plan.run(taps, sig, rx, mode=fdtd.MODE_IR)
                assert np.count_nonzero(ref) > 0.3 * ref.size
                for variant in range(4):
                    for zc in (1, 2, 4, 8, 16, 32):
                        print(f"case: {name} {c.grid.n}, {prec}, variant {variant}, z-chunk {zc}", flush=True)
                        plan.set_tile(zc, variant)
                        ir, _ = plan.run(taps, sig, rx, mode=fdtd.MODE_IR)
                        assert np.array_equal(ir, ref), (name, prec, variant, zc)


@pytest.mark.parametrize("steps", [1, 2, 3, 130])
def test_short_runs_both_pass_modes(shoebox, steps):
    case, og = shoebox
    print(f"case: shoebox, {steps} steps, fp64, one, two and three steps per pass", flush=True)
    sig = case.signal[:steps]
    ref = oracle.run(og, case.runs[0][2], sig, case.receivers, threads=4)[0]
    with fdtd.Plan(case.grid, "fp64") as plan:
        plan.set_pass(1)
        one, _ = plan.run(case.runs[0][2], sig, case.receivers, mode=fdtd.MODE_IR)
        plan.set_pass(2, 8)
        two, _ = plan.run(case.runs[0][2], sig, case.receivers, mode=fdtd.MODE_IR)
        plan.set_pass(3, 8)
        three, _ = plan.run(case.runs[0][2], sig, case.receivers, mode=fdtd.MODE_IR)
    assert np.array_equal(one, two) and np.array_equal(one, three)
    if np.linalg.norm(ref) > 0:
        assert rel_l2(one, ref) < 1e-9
    else:
        assert np.all(one == 0)


def test_cancel_from_progress(shoebox):
    case, _ = shoebox
    print("case: shoebox, cancel from the progress callback", flush=True)
    flag = np.zeros(1, dtype=np.int32)
    seen = []

    def progress(f):
        seen.append(f)
        flag[0] = 1

    with fdtd.Plan(case.grid, "fp32") as plan:
        out = plan.run(case.runs[0][2], case.signal, case.receivers, cancel=flag, progress=progress)
    assert out is None
    assert seen and seen[0] < 1.0


def test_progress_reaches_one(shoebox):
    case, _ = shoebox
    seen = []
    with fdtd.Plan(case.grid, "fp32") as plan:
        out = plan.run(case.runs[0][2], case.signal, case.receivers, progress=seen.append)
    assert out is not None
    assert seen == sorted(seen) and seen[-1] == 1.0


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-4)])
@pytest.mark.parametrize("spp", [1, 2, 3])
def test_solid_and_duplicate_taps(shoebox, prec, tol, spp):
    case, og = shoebox
    g = case.grid
    nx, ny, nz = g.n
    solid = g.index(0, ny // 2, nz // 2)  # the padding layer at x = 0
    assert not (g.flags[solid] & fdtd.AIR)
    air = case.runs[0][2][0][0]
    taps = [(solid, 0.5), (air, 0.7), (air, 0.3)]
    rx = list(case.receivers) + [solid]
    print(f"case: shoebox, tap on a solid cell and a duplicate tap, {prec}, steps/pass {spp}", flush=True)
    ref = oracle.run(og, taps, case.signal[:800], rx, threads=4)[0]
    with fdtd.Plan(g, prec) as plan:
        plan.set_pass(spp, 8)
        ir, _ = plan.run(taps, case.signal[:800], rx, mode=fdtd.MODE_IR)
    assert rel_l2(ir, ref) < tol
    assert rel_l2(ir[-1], ref[-1]) < tol  # the solid cell: its injections accumulate


def without_x_padding(g):
    """The grid less its columns x = 0 and x = nx - 1: the cells next to the
    walls become the grid's end columns and the cells beside them interior.
    Works for the library's and the oracle's grid (same fields)."""
    nx, ny, nz = g.n
    keep = np.zeros(nx, bool)
    keep[1:-1] = True
    flags = np.ascontiguousarray(g.flags.reshape(nz, ny, nx)[:, :, keep].ravel())
    z, r = np.divmod(g.b_cell.astype(np.int64), nx * ny)
    y, x = np.divmod(r, nx)
    assert np.all((x >= 1) & (x <= nx - 2)), "a boundary cell in a padding column"
    b_cell = ((z * ny + y) * (nx - 2) + (x - 1)).astype(np.uint32)
    origin = (g.origin[0] + g.dx, g.origin[1], g.origin[2])
    return dataclasses.replace(g, n=(nx - 2, ny, nz), origin=origin, flags=flags, b_cell=b_cell)


def test_air_at_the_grid_x_ends_all_pass_modes():
    cfg = dict(configs.pillared_room(), dx=0.08, steps=600)
    case = workload.build_case(cfg)
    g = without_x_padding(case.grid)
    og = without_x_padding(oracle.build_box(cfg["room"], cfg["c"], cfg["dx"]))
    nx, ny, nz = g.n
    f = g.flags.reshape(nz, ny, nx)
    assert np.count_nonzero((f[:, :, 1] & 127) == 127) > 0 and np.count_nonzero((f[:, :, nx - 2] & 127) == 127) > 0
    src = g.nearest_air(g.center(case.runs[0][2][0][0]) - np.array([g.dx, 0.0, 0.0]))
    taps = [(src, 1.0)]
    rx = [g.index(1, y, z) for y, z in ((ny // 2, nz // 2), (ny // 3, nz // 4))] + \
         [g.index(nx - 2, ny // 2, nz // 3), g.index(2, ny // 4, nz // 2)]
    sig = case.signal[:600]
    print(f"case: pillared room without x padding, grid {g.n}, 600 steps, passes of 1, 2, 3", flush=True)
    ref = oracle.run(og, taps, sig, rx, threads=4)[0]
    for prec in ("fp64", "fp32"):
        with fdtd.Plan(g, prec) as plan:
            outs = []
            for spp, zc in ((1, 0), (2, 8), (3, 8), (3, 0)):
                plan.set_pass(spp, zc)
                outs.append(plan.run(taps, sig, rx, mode=fdtd.MODE_IR)[0])
        for o in outs[1:]:
            assert np.array_equal(outs[0], o), prec
        assert np.all(np.isfinite(outs[0]))
        assert rel_l2(outs[0], ref) < (1e-9 if prec == "fp64" else 1e-4), prec
