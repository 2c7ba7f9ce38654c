"""The committed FP64 golden series (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle).

CPU: the oracle reproduces them, and the library's host side (grid builder,
source signal, receiver nodes, taps) reproduces their inputs exactly.
GPU: the device runs match them, FP64 at 1e-9 and FP32 at 1e-4 relative L2.
"""
import os

import numpy as np
import pytest

import oracle
from wavey_amd import configs, fdtd

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def runs_of(g):
    out, k = [], 0
    for n in g["tap_counts"]:
        out.append([(int(c), float(w)) for c, w in zip(g["tap_cells"][k:k + n], g["tap_weights"][k:k + n])])
        k += n
    return out


def cfg_of(name):
    if name == "shoebox":
        return configs.shoebox(fdtd.fit_real_impedance)
    return dict(configs.pillared_room(), dx=0.08, steps=1500)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("name", ["shoebox", "pillared_mini"])
def test_oracle_reproduces_golden(name):
    g = load(name)
    cfg = cfg_of(name)
    og = oracle.build_box(cfg["room"], cfg["c"], cfg["dx"])
    assert tuple(og.n) == tuple(g["grid_n"])
    for taps, ref in zip(runs_of(g), g["ir"]):
        ir = oracle.run(og, taps, g["signal"], [int(r) for r in g["receivers"]], threads=8)[0]
        assert rel_l2(ir, ref) < 1e-12


@pytest.mark.parametrize("name", ["shoebox", "pillared_mini"])
def test_library_host_side_reproduces_golden_inputs(name):
    g = load(name)
    cfg = cfg_of(name)
    grid = fdtd.FdtdGrid.build_box(cfg["room"], cfg["c"], cfg["dx"])
    assert tuple(grid.n) == tuple(g["grid_n"]) and grid.dt == float(g["dt"])
    assert np.array_equal(fdtd.source_signal(cfg["f_src"], grid.dt, len(g["signal"])), g["signal"])
    pts = cfg["receivers"] if name == "shoebox" else cfg["receivers"][::8]
    assert [grid.nearest_air(p) for p in pts] == [int(r) for r in g["receivers"]]
    if name == "shoebox":
        assert grid.taps_at(cfg["sources"][0]["pos"]) == runs_of(g)[0]
    else:
        node = grid.nearest_air(cfg["sources"][0]["pos"])
        assert fdtd.source_components(grid, node, 1)[:2] == runs_of(g)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["shoebox", "pillared_mini"])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-4)])
def test_device_matches_golden(name, prec, tol):
    g = load(name)
    cfg = cfg_of(name)
    print(f"case: golden {name} {prec}", flush=True)
    grid = fdtd.FdtdGrid.build_box(cfg["room"], cfg["c"], cfg["dx"])
    with fdtd.Plan(grid, prec) as plan:
        for taps, ref in zip(runs_of(g), g["ir"]):
            ir, _ = plan.run(taps, g["signal"], [int(r) for r in g["receivers"]], mode=fdtd.MODE_IR)
            assert rel_l2(ir, ref) < tol
