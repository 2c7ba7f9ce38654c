"""GPU: the reference's analytic wave tests, run on the HIP path.

* A directional source built from the four first-order components
  (core/wavey-solver/tests/fdtd.rs:449-525): in a large room, before the
  first reflection, the measured directivity on a 1.6 m circle and above the
  source equals the first-order fit including the dipoles' near-field term,
  within 0.05. The cardioid case (the reference's second model is a sample
  speaker file, which is not part of this path).
* Oblique reflection off a frequency-dependent floor (fdtd.rs:527-595): the
  floor-reflected part, isolated with a second run over a rigid floor and the
  analytic direct sound (with the scheme's axial phase lag), equals the
  spherical-wave reflection factor of the fitted impedance within 0.05 at
  incidence 30-75 degrees.
"""
import math

import numpy as np
import pytest

from analytic import spherical
from wavey_amd import fdtd
from wavey_amd.meshes import RIGID

pytestmark = pytest.mark.gpu

C = 343.0


def spectra_of(plan, taps, signal, rx, freqs, grid, h0):
    _, sp = plan.run(taps, signal, rx, freqs, mode=fdtd.MODE_SPECTRA)
    return fdtd.to_transfer(sp, signal, grid, freqs, C, 1.0, h0)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_directional_source_reproduces_directivity(prec):
    f_max, ppw = 500.0, 12.0
    dx = C / (f_max * ppw)
    room = {"L": (14.0, 12.0, 10.0), "pillars": [], "fits": [RIGID], "patches": [(1, 1, [0])] * 6}
    g = fdtd.FdtdGrid.build_box(room, C, dx)
    centre = np.array([7.0, 6.0, 5.0])
    src = g.nearest_air(centre)
    comps = fdtd.source_components(g, src, 1)
    assert len(comps) == 4
    r = 1.6
    angles = [0.0, 60.0, 90.0, 135.0, 180.0, 270.0]
    rx = [g.nearest_air(centre + r * np.array([math.cos(math.radians(a)), math.sin(math.radians(a)), 0.0]))
          for a in angles]
    rx.append(g.nearest_air(centre + np.array([0.0, 0.0, r])))
    steps = int(0.018 / g.dt)  # before the first reflection (>= 5 m away)
    signal = fdtd.source_signal(f_max, g.dt, steps)
    freqs = [120.0, 200.0, 300.0]
    nf = len(freqs)
    print(f"case: directivity, grid {g.n}, {steps} steps, {prec}", flush=True)
    with fdtd.Plan(g, prec) as plan:
        spectra = [spectra_of(plan, taps, signal, rx, freqs, g, [1.0] * nf) for taps in comps]
    # A cardioid turned 90 degrees about z: local +x is the room's +y.
    yaw = math.radians(90.0)
    to_room = np.array([[math.cos(yaw), -math.sin(yaw), 0.0], [math.sin(yaw), math.cos(yaw), 0.0], [0.0, 0.0, 1.0]])
    dirs = fdtd.first_order_directions()
    d = 0.5 + 0.5 * dirs[:, 0]
    ps = g.center(src)
    worst = 0.0
    for fi, f in enumerate(freqs):
        a0, a, res = fdtd.fit_first_order(d, 1.0, to_room)
        assert res < 1e-9  # a cardioid is first order
        k = 2.0 * math.pi * f / C
        w = fdtd.component_weights(a0, a, k, g.dx)
        for ri, cell in enumerate(rx):
            total = sum(w[ci] * spectra[ci][ri * nf + fi] for ci in range(len(comps)))
            measured = total / spectra[0][ri * nf + fi]
            v = g.center(cell) - ps
            dist = float(np.linalg.norm(v))
            u = v / dist
            near = 1.0 + 1.0 / complex(0.0, k * dist)
            fit = a0 + (a[0] * u[0] + a[1] * u[1] + a[2] * u[2]) * near
            worst = max(worst, abs(measured - fit))
    print(f"worst |measured - first-order fit| {worst:.4f}")
    assert worst < 0.05


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_oblique_reflection_matches_spherical_wave_theory(prec):
    f_max, ppw = 600.0, 10.0
    dx = C / (f_max * ppw)
    alpha = 0.3
    floor = 4  # surfaces: x = 0, x = L, y = 0, y = L, floor, ceiling
    fit, info = fdtd.fit_admittance({"id": "floor", "absorption": alpha}, (20.0, f_max))

    def grid(floor_fit):
        patches = [(1, 1, [0])] * 6
        patches[floor] = (1, 1, [1])
        return fdtd.FdtdGrid.build_box({"L": (20.0, 14.0, 8.0), "pillars": [], "fits": [RIGID, floor_fit],
                                        "patches": patches}, C, dx)

    g_imp, g_rig = grid(fit), grid(RIGID)
    assert g_imp.n == g_rig.n
    src = g_imp.nearest_air(np.array([8.0, 7.0, 1.2]))
    sp = g_imp.center(src)
    h = float(sp[2])  # the floor is at z = 0
    angles = [30.0, 45.0, 60.0, 70.0, 75.0]
    rx = [g_imp.nearest_air(sp + np.array([2.0 * h * math.tan(math.radians(a)), 0.0, 0.0])) for a in angles]
    steps = int(12.0 / C / g_imp.dt)
    signal = fdtd.source_signal(f_max, g_imp.dt, steps)
    freqs = [200.0, 300.0, 450.0]
    nf = len(freqs)
    h0 = [1.0] * nf
    print(f"case: oblique reflection, grid {g_imp.n}, {steps} steps, {prec}, fit error {info['max_error']:.3f}",
          flush=True)
    with fdtd.Plan(g_imp, prec) as p:
        p_imp = spectra_of(p, [(src, 1.0)], signal, rx, freqs, g_imp, h0)
    with fdtd.Plan(g_rig, prec) as p:
        p_rig = spectra_of(p, [(src, 1.0)], signal, rx, freqs, g_rig, h0)
    worst = 0.0
    for ri, cell in enumerate(rx):
        xp = g_imp.center(cell)
        d = float(np.linalg.norm(xp - sp))
        r2 = math.sqrt(d * d + 4.0 * h * h)
        cos = 2.0 * h / r2
        for fi, f in enumerate(freqs):
            k = 2.0 * math.pi * f / C
            lag = math.radians(fdtd.phase_error_deg_per_m(f, C, g_imp.dx)) * d
            direct = (1.0 / d) * complex(math.cos(-k * d - lag), math.sin(-k * d - lag))
            q = (p_imp[ri * nf + fi] - direct) / (p_rig[ri * nf + fi] - direct)
            zeta = 1.0 / fdtd.admittance(fit, f)
            err = abs(q - spherical(zeta, cos, k, r2))
            worst = max(worst, err)
            print(f"theta {math.degrees(math.acos(cos)):4.1f} {f:3.0f} Hz: |Q| {abs(q):.3f} "
                  f"spherical {abs(spherical(zeta, cos, k, r2)):.3f} |dQ| {err:.3f}")
    assert worst < 0.05, worst
