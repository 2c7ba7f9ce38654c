"""GPU: the reference's analytic FDTD tests (core/wavey-solver/tests/fdtd.rs)
run on the HIP path in FP32, the product's arithmetic: the same checks that
pin the FP64 oracle in test_oracle.py, with the device computing the run.

* direct sound before the first reflection: 1/r within 0.1 dB and the
  scheme's axial phase lag within 0.3 degrees (fdtd.rs:13-47);
* rigid-box and 1 m cube eigenfrequencies within 0.5 % and 0.6 %
  (fdtd.rs:49-145);
* plane-wave reflection off a frequency-dependent end wall of a rigid tube:
  (zeta - 1)/(zeta + 1) within 0.03 (fdtd.rs:262-321, with this build's
  first-order-section admittances);
* multi-section walls: each third of a long run contributes a tenth of the
  previous third (fdtd.rs:325-350), in FP64 on the device (the FP32 rounding
  floor of a 9000-step run lies above the last third's contribution).
"""
import math

import numpy as np
import pytest

from wavey_amd import configs, fdtd
from wavey_amd.meshes import RIGID

pytestmark = pytest.mark.gpu

C = 343.0
COURANT = 1.0 / math.sqrt(3.0)


def box_room(L, fit):
    return {"L": L, "pillars": [], "fits": [fit], "patches": [(1, 1, [0])] * 6}


def run_spectra(g, taps, sig, rx, freqs, prec="fp32"):
    with fdtd.Plan(g, prec) as plan:
        _, sp = plan.run(taps, sig, rx, freqs, mode=fdtd.MODE_SPECTRA)
    return sp


def test_direct_sound_is_calibrated():
    f_max = 600.0
    g = fdtd.FdtdGrid.build_box(box_room((12.0, 10.0, 8.0), (1.0 / fdtd.fit_real_impedance(0.2), [])), C,
                                C / (f_max * 10.0))
    src = g.nearest_air((6.0, 5.0, 4.0))
    rcv = g.nearest_air((7.5, 5.0, 4.0))
    r = float(np.linalg.norm(g.center(rcv) - g.center(src)))
    steps = int(0.016 / g.dt)
    sig = fdtd.source_signal(f_max, g.dt, steps)
    freqs = [150.0, 250.0, 350.0, 450.0]
    print(f"case: direct sound, grid {g.n}, {steps} steps, fp32", flush=True)
    t = fdtd.to_transfer(run_spectra(g, [(src, 1.0)], sig, [rcv], freqs), sig, g, freqs, C, 1.0, np.ones(4))
    for fi, f in enumerate(freqs):
        k = 2 * math.pi * f / C
        want = np.exp(-1j * k * r) / r
        db = 20 * math.log10(abs(t[fi]) / abs(want))
        deg = math.degrees(np.angle(t[fi] / want))
        predicted = -fdtd.phase_error_deg_per_m(f, C, g.dx) * r
        assert abs(db) < 0.1, (f, db)
        assert abs(deg - predicted) < 0.3, (f, deg, predicted)


def peak_errors(mag, freqs, modes):
    worst, checked = 0.0, 0
    for m in modes:
        if any(1e-9 < abs(o / m - 1) < 0.03 for o in modes):
            continue  # not well separated
        idx = np.nonzero(np.abs(freqs / m - 1) < 0.015)[0]
        k = idx[np.argmax(mag[idx])]
        worst = max(worst, abs(freqs[k] / m - 1))
        checked += 1
    return worst, checked


def test_rigid_box_eigenfrequencies():
    L = (3.0, 2.5, 2.0)
    g = fdtd.FdtdGrid.build_box(box_room(L, RIGID), C, 0.1)
    src = g.nearest_air((0.05, 0.05, 0.05))
    rcv = g.nearest_air((L[0] - 0.05, L[1] - 0.05, L[2] - 0.05))
    steps = int(2.0 / g.dt)
    sig = fdtd.source_signal(400.0, g.dt, steps)
    freqs = 40.0 + np.arange(1000) * 0.08
    print(f"case: rigid box modes, grid {g.n}, {steps} steps, fp32", flush=True)
    mag = np.abs(run_spectra(g, [(src, 1.0)], sig, [rcv], freqs)[0])
    modes = [C / 2 * math.sqrt((a / L[0]) ** 2 + (b / L[1]) ** 2 + (c / L[2]) ** 2)
             for a in range(4) for b in range(4) for c in range(4)]
    worst, checked = peak_errors(mag, freqs, [m for m in modes if 45.0 < m < 115.0])
    assert checked >= 4
    assert worst < 0.005, worst


def test_rigid_cube_modes_to_800_hz():
    n = math.ceil(1.0 / (C / (800.0 * 15.0)))
    f_max = C * n / 15.0
    g = fdtd.FdtdGrid.build_box(box_room((1.0, 1.0, 1.0), RIGID), C, C / (f_max * 15.0))
    src = g.nearest_air((0.2, 0.2, 0.2))
    rcv = g.nearest_air((0.9, 0.9, 0.9))
    steps = int(4.0 / g.dt)
    sig = fdtd.source_signal(1600.0, g.dt, steps)
    freqs = 90.0 + np.arange(7400) * 0.1
    print(f"case: rigid 1 m cube modes, grid {g.n}, {steps} steps, fp32", flush=True)
    mag = np.abs(run_spectra(g, [(src, 1.0)], sig, [rcv], freqs)[0])
    modes = [C / 2 * math.sqrt(a * a + b * b + c * c) for a in range(6) for b in range(a + 1) for c in range(b + 1)]
    worst, checked = peak_errors(mag, freqs, [m for m in modes if 100.0 <= m <= 800.0])
    assert checked >= 5
    assert worst < 0.006, worst


def test_tube_reflection_matches_admittance():
    f_max, ppw = 500.0, 15.0
    dx = C / (f_max * ppw)
    cells = lambda v: round(v / dx) * dx  # noqa: E731
    ln, w = cells(30.0), cells(0.3)
    rng = np.random.default_rng(11)
    for trial in range(2):
        fit = configs.first_order_fit(rng, alpha_range=(0.3, 0.6))
        # Surface 0 (x = 0) carries the fit, the rest is rigid.
        room = {"L": (ln, w, w), "pillars": [], "fits": [RIGID, fit], "patches": [(1, 1, [1])] + [(1, 1, [0])] * 5}
        g = fdtd.FdtdGrid.build_box(room, C, dx)
        src = g.nearest_air((15.0, w / 2, w / 2))
        rcv = g.nearest_air((2.0, w / 2, w / 2))
        xr = g.center(rcv)[0]
        steps = int(0.075 / g.dt)
        sig = fdtd.source_signal(f_max, g.dt, steps)
        freqs = 40.0 + np.arange(15) * 22.0
        print(f"case: tube reflection, trial {trial}, grid {g.n}, {steps} steps, fp32", flush=True)
        spec_all = run_spectra(g, [(src, 1.0)], sig, [rcv], freqs)[0]
        long = {"L": (ln * 2 + cells(15.0), w, w), "pillars": [], "fits": [RIGID], "patches": [(1, 1, [0])] * 6}
        gl = fdtd.FdtdGrid.build_box(long, C, dx)
        shift = cells(30.0)
        src_l = gl.nearest_air((15.0 + shift, w / 2, w / 2))
        rcv_l = gl.nearest_air((xr + shift, w / 2, w / 2))
        spec_inc = run_spectra(gl, [(src_l, 1.0)], sig, [rcv_l], freqs)[0]
        worst = 0.0
        for fi, f in enumerate(freqs):
            om = 2 * math.pi * f
            k_num = 2.0 / g.dx * math.asin(math.sin(om * g.dt / 2) / COURANT)
            r_meas = (spec_all[fi] - spec_inc[fi]) / spec_inc[fi] * np.exp(1j * k_num * 2 * xr)
            zeta = 1.0 / fdtd.admittance(fit, f)
            worst = max(worst, abs(r_meas - (zeta - 1) / (zeta + 1)))
        assert worst < 0.03, worst


def test_branch_boundaries_decay():
    rng = np.random.default_rng(13)
    fits = [configs.first_order_fit(rng, alpha_range=(0.1, 0.3)) for _ in range(6)]
    room = {"L": (3.0, 2.4, 2.1), "pillars": [], "fits": fits, "patches": [(1, 1, [s]) for s in range(6)]}
    g = fdtd.FdtdGrid.build_box(room, C, C / (300.0 * 12.0))
    src = g.nearest_air((0.7, 0.6, 0.5))
    rcv = g.nearest_air((2.3, 1.7, 1.6))
    steps = 9000
    sig = fdtd.source_signal(300.0, g.dt, steps)
    print(f"case: multi-section walls decay, grid {g.n}, {steps} steps, fp64", flush=True)
    with fdtd.Plan(g, "fp64") as plan:
        ir, _ = plan.run([(src, 1.0)], sig, [rcv], mode=fdtd.MODE_IR)
    x = ir[0]
    assert np.all(np.isfinite(x))
    n = np.arange(steps)
    for f in (80.0, 200.0):
        e = x * np.exp(-2j * math.pi * f * g.dt * n)
        first, mid, last = (abs(e[: steps // 3].sum()), abs(e[steps // 3: 2 * steps // 3].sum()),
                            abs(e[2 * steps // 3:].sum()))
        assert mid < 0.1 * first and last < 0.1 * mid, (f, first, mid, last)
