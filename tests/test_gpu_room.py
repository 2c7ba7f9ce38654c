"""GPU: runs on grids of general rooms (fdtd.rs:289-461) against the FP64
oracle on the oracle's own grid of the same mesh.

The builder's grid is the oracle's bit for bit (checked here again, and on
the CPU in test_room_grid.py); the run then has to match as on box rooms:
FP64 within 1e-9 relative L2 over all receiver samples, FP32 within 1e-4.
Rooms: a sloped ceiling with a turned cabinet, an extruded L-shaped plan
with a hole and a cabinet, the stepped floor with a balcony slab.
"""
import numpy as np
import pytest

import oracle
from room_cases import cabinet, l_prism, sloped, stepped_balcony
from wavey_amd import fdtd

pytestmark = pytest.mark.gpu

C = 343.0
TOL = {"fp64": 1e-9, "fp32": 1e-4}


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def case(name):
    if name == "sloped_cabinet":
        mesh, _ = sloped()
        solid, _ = cabinet(pos=(1.6, 1.0, 0.9))
        return mesh, [solid], 0.05, (0.7, 0.6, 1.1), [(2.6, 1.5, 0.9), (1.6, 0.35, 1.5), (0.4, 1.7, 2.0)]
    if name == "l_prism_cabinet":
        solid, _ = cabinet(pos=(2.2, 0.7, 1.0), size=(0.4, 0.35, 0.5), angle=0.9)
        return l_prism(), [solid], 0.05, (0.6, 2.1, 1.2), [(2.8, 0.4, 1.1), (1.1, 1.1, 0.5), (0.3, 0.2, 2.0)]
    mesh, _ = stepped_balcony()
    return mesh, [], 0.1, (2.0, 2.5, 1.5), [(8.0, 2.5, 2.2), (1.5, 4.0, 3.2), (5.0, 1.0, 0.8)]


@pytest.fixture(scope="module", params=["sloped_cabinet", "l_prism_cabinet", "stepped_balcony"])
def room_run(request):
    mesh, solids, dx, src, rxp = case(request.param)
    g = fdtd.FdtdGrid.build_room(mesh, C, dx, solids)
    og = oracle.build_room(mesh, C, dx, solids)
    assert np.array_equal(g.flags, og.flags) and np.array_equal(g.slots, og.slots)
    steps = 700
    sig = oracle.source_signal(250.0, og.dt, steps)
    taps = g.taps_at(src)
    assert taps == oracle.taps_at(og, src)
    rx = [g.nearest_air(p) for p in rxp]
    ref = oracle.run(og, taps, sig, rx, threads=8)[0]
    return request.param, g, taps, sig, rx, ref


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("spp", [1, 2, 3])
def test_room_ir_matches_oracle(room_run, prec, spp):
    name, g, taps, sig, rx, ref = room_run
    print(f"case: {name} {g.n} {len(sig)} steps {prec} steps/pass {spp}", flush=True)
    with fdtd.Plan(g, prec) as plan:
        plan.set_pass(spp, 8)
        ir, _ = plan.run(taps, sig, rx, mode=fdtd.MODE_IR)
    err = rel_l2(ir, ref)
    print(f"relative L2 {err:.3e}")
    assert np.all(np.isfinite(ir))
    assert err < TOL[prec]
