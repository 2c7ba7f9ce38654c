"""GPU: fused three-step passes (boundary cells inside the z-march) in the
batch, against plain runs.

A fused pass forms every cell with the single-step kernel's arithmetic (the
interior update and the boundary solve, same operands), so each instance's
receiver series must be bit-identical to its plain run (``Plan.run`` with one
step per launch; the plain FP64 path is checked against the oracle in
test_gpu_parity.py). The cases cover step counts that end inside a 128-step
block and are not multiples of three (passes of one and two steps), several
span blocks and z-chunks, rows wider than a wave, rooms thinner than a span
block, slots with mass-spring branches (state g), FP64 slot state, and a
room too wide for the fused form (it keeps single steps in the same batch).
"""
import numpy as np
import pytest

from wavey_amd import configs, fdtd, workload

pytestmark = pytest.mark.gpu


def mix_rooms(n, seed=7, steps=(260, 700), dx=0.1, dims=None):
    rng = np.random.default_rng(seed)
    lib = configs.material_library(seed)
    out = []
    for i in range(n):
        d = dims[i] if dims else (rng.uniform(2.5, 6.0), rng.uniform(2.5, 5.0), rng.uniform(2.4, 3.2))
        cfg = configs.mix_room(rng, lib, dx=dx, dims=d, n_rx=6)
        cfg["steps"] = int(rng.integers(steps[0], steps[1]))
        out.append(workload.build_case(cfg))
    return out


def plain_ir(case, taps, prec="fp32"):
    with fdtd.Plan(case.grid, prec) as p:
        p.set_pass(1, 0)
        return p.run(taps, case.signal, case.receivers, mode=fdtd.MODE_IR)[0]


def batch_irs(items, prec, form, graphs=True, zchunk=0):
    with fdtd.Batch(prec) as b:
        ids = [b.add(c.grid, taps, c.signal, c.receivers) for c, taps in items]
        if zchunk:
            b.set_tile(-1, zchunk)
        b.set_launch(form)
        assert b.run(graphs=graphs) == fdtd.WV_OK
        t = b.timing()
        info = [b.info(i) for i in ids]
        return [b.result(i)[0] for i in ids], t, info


@pytest.fixture(scope="module")
def rooms():
    cases = mix_rooms(5)
    items = [(c, taps) for c in cases for (_, _, taps) in c.runs]
    return cases, items


@pytest.mark.parametrize("prec", ["fp32", "fp32_state64"])
@pytest.mark.parametrize("graphs", [True, False])
def test_fused_equals_single_steps(rooms, prec, graphs):
    cases, items = rooms
    print(f"case: fused batch of {len(items)} instances, {prec}, graphs={graphs}", flush=True)
    ref, _, _ = batch_irs(items, prec, fdtd.BATCH_STEPS, graphs)
    out, t, info = batch_irs(items, prec, fdtd.BATCH_FUSED, graphs)
    assert t["launch"] == fdtd.BATCH_FUSED
    assert {d["zchunk"] for d in info} == {32}
    nonzero = 0
    for k in range(len(items)):
        assert np.array_equal(out[k], ref[k]), f"instance {k}"
        nonzero += int(np.count_nonzero(out[k]))
    assert nonzero > 1000
    for ci, c in enumerate(cases[:2]):
        k = next(k for k, (cc, _) in enumerate(items) if cc is c)
        assert np.array_equal(out[k], plain_ir(c, c.runs[0][2], prec)), f"room {ci}"


@pytest.mark.parametrize("zchunk", [4, 16])
def test_fused_z_chunks(rooms, zchunk):
    """Blocks of 4 and 16 planes (several z-chunks per room, special cells
    and receivers near chunk ends) give the same series."""
    cases, items = rooms
    print(f"case: fused batch, z-chunk {zchunk}", flush=True)
    ref, _, _ = batch_irs(items, "fp32", fdtd.BATCH_STEPS)
    out, _, info = batch_irs(items, "fp32", fdtd.BATCH_FUSED, zchunk=zchunk)
    assert {d["zchunk"] for d in info} == {zchunk}
    for k in range(len(items)):
        assert np.array_equal(out[k], ref[k]), f"instance {k}"


def test_fused_wide_thin_and_too_wide_rooms():
    """Rows of 292 cells (pitch 320: 2816-cell span steps, pillars, several
    spans per plane), rooms thinner than a span in x, y and z, and a 40 m
    room (pitch 832 > 768) that keeps single steps in the same batch."""
    dims = [(14.5, 3.4, 2.6), (0.45, 6.0, 2.5), (6.0, 0.45, 2.5), (5.0, 4.0, 0.35), (40.0, 2.4, 2.4)]
    cases = mix_rooms(len(dims), seed=21, steps=(300, 420), dx=0.05, dims=dims)
    items = [(c, c.runs[0][2]) for c in cases]
    print(f"case: fused batch, grids {[c.grid.n for c in cases]}", flush=True)
    out, _, info = batch_irs(items, "fp32", fdtd.BATCH_FUSED)
    assert info[0]["zchunk"] == 32 and info[-1]["zchunk"] != 32  # the 40 m room is not fused
    for k, (c, taps) in enumerate(items):
        ref = plain_ir(c, taps)
        assert np.all(np.isfinite(ref))
        assert np.array_equal(out[k], ref), f"room {dims[k]}"


def test_fused_spring_branches():
    """Library materials with mass-spring branches (slot state g live), the
    room of tests/fdtd.rs:352-377, 700 steps: fused equals single steps in
    FP32 and with FP64 slot state."""
    import json
    import os
    lib = {m["id"]: m for m in json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                                          "library_materials.json")))}
    ids = ["audience-standing", "mineral-wool-50", "plaster-rough", "carpet-heavy", "wood-panel",
           "ceiling-industrial"]
    fits = fdtd.fit_surfaces([lib[i] for i in ids], (20.0, 300.0))
    assert any(k > 0 for f in fits for (_, _, k) in f[1])
    c = 343.0
    dx = c / (300.0 * 12.0)
    room = {"L": (5.0, 4.0, 3.0), "pillars": [], "fits": fits, "patches": [(1, 1, [i]) for i in range(6)]}
    g = fdtd.FdtdGrid.build_box(room, c, dx)
    src = g.nearest_air((1.2, 1.1, 1.3))
    rx = [g.nearest_air(p) for p in ((4.1, 3.2, 1.5), (2.5, 0.5, 2.6))]
    sig = fdtd.source_signal(300.0, g.dt, 700)
    for prec in ("fp32", "fp32_state64"):
        print(f"case: fused spring room {g.n}, {prec}", flush=True)
        with fdtd.Plan(g, prec) as p:
            p.set_pass(1, 0)
            ref = p.run([(src, 1.0)], sig, rx, mode=fdtd.MODE_IR)[0]
        with fdtd.Batch(prec) as b:
            i = b.add(g, [(src, 1.0)], sig, rx)
            b.set_launch(fdtd.BATCH_FUSED)
            b.run()
            out = b.result(i)[0]
            assert b.info(i)["zchunk"] == 32
        assert np.count_nonzero(ref) > 100
        assert np.array_equal(out, ref), prec


def test_fused_autotune_keeps_results(rooms):
    cases, items = rooms
    print("case: autotune with the fused form among the candidates", flush=True)
    ref, _, _ = batch_irs(items, "fp32", fdtd.BATCH_STEPS)
    with fdtd.Batch("fp32") as b:
        ids = [b.add(c.grid, taps, c.signal, c.receivers) for c, taps in items]
        b.autotune(trial_steps=6)
        b.run()
        form = b.timing()["launch"]
        zc = {b.info(i)["zchunk"] for i in ids}
        out = [b.result(i)[0] for i in ids]
    print(f"  tuned form {form}, z-chunks {zc}", flush=True)
    assert form in (fdtd.BATCH_STEPS, fdtd.BATCH_PERSISTENT, fdtd.BATCH_FUSED)
    if form == fdtd.BATCH_FUSED:
        assert zc <= {16, 32, 64}
    for k in range(len(items)):
        assert np.array_equal(out[k], ref[k]), f"instance {k}"
