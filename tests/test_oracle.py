"""Pins the FP64 oracle to what the mathematics fixes.

Each test names the reference test or closed form it restates
(core/wavey-solver/tests/fdtd.rs, core/wavey-materials/src/*.rs). Checks are
chosen so a plausible slip in the restatement (a dropped term, a wrong sign,
a transposed index) fails one of them; none retypes the oracle's formula.
"""
import math

import numpy as np
import pytest

import oracle as O
from wavey_amd import configs

C = 343.0


def box_room(L, fit, pillars=()):
    return {"L": L, "pillars": list(pillars), "fits": [fit],
            "patches": [(1, 1, [0])] * (6 + 4 * len(pillars))}


RIGID = (0.0, [])


# ------------------------------------------------------------ materials

@pytest.mark.parametrize("alpha", [0.02, 0.1, 0.3, 0.6, 0.9])
def test_fit_real_impedance_reproduces_absorption(alpha):
    # material.rs:146-157: the fitted real impedance has Paris absorption alpha.
    z = O.fit_real_impedance(alpha)
    assert z >= 1.5
    assert abs(O.paris(complex(1.0 / z, 0.0)) - alpha) < 1e-6


def test_paris_limits():
    # Rigid: no absorption. Matched (beta = 1) normal incidence is the
    # textbook maximum region; the integral of (1-|R|^2) sin 2t over [0, pi/2]
    # for beta = 1 has the closed form 8 ln2 - 5 = 0.54518 ... times ... checked
    # by adaptive quadrature instead.
    assert O.paris(0j) == 0.0
    from scipy.integrate import quad

    for beta in (0.3 + 0.1j, 1.0, 0.05 - 0.02j):
        def f(t):
            c = math.cos(t)
            r = (c - beta) / (c + beta)
            return (1 - abs(r) ** 2) * math.sin(2 * t)

        want = quad(f, 0, math.pi / 2, limit=200)[0]
        assert abs(O.paris(complex(beta), 256) - want) < 2e-5


def test_admittance_of_sections_is_passive_and_matches_closed_form():
    # impedance_fit.rs:48-55 for branches 1/(r + jwm + k/jw); a first-order
    # section rho/(s+a) is m = 1/rho, r = a/rho, k = 0.
    rng = np.random.default_rng(1)
    fit = configs.first_order_fit(rng)
    for f in (10.0, 63.0, 250.0, 900.0):
        rho = [1.0 / m for (_, m, _) in fit[1]]
        a = [r / m for (r, m, _) in fit[1]]
        want = sum(ro / (2j * math.pi * f + aa) for ro, aa in zip(rho, a))
        got = O.admittance(fit, f)
        assert abs(got - want) < 1e-12 * abs(want)
        assert got.real >= 0.0
    # A spring branch: 1/(k/jw) = jw/k.
    k = 3000.0
    assert abs(O.admittance((0.0, [(0.0, 0.0, k)]), 100.0) - 2j * math.pi * 100.0 / k) < 1e-12


# -------------------------------------------------------------- boundary

def branch_system(raw, counts, p, pp, lap, lam, dt, u0, g0):
    """Brute force: the branch equations of fdtd.rs:18-30 as one linear
    system in (p+, u+_j), solved with numpy.

    mass branch (m > 0): m (u+ - u)/dt + r (u+ + u)/2 + k (g+ + g)/2 = (p+ + 2p + p-)/4
    no mass:             r u+ + k g+ = (p+ + p)/2
    with g+ = g + dt (u+ + u)/2 and p+ = 2p - p- + lap - lam sum_j (u+_j - u_j)."""
    S = len(raw)
    A = np.zeros((S + 1, S + 1))
    b = np.zeros(S + 1)
    A[0, 0] = 1.0
    b[0] = 2 * p - pp + lap + lam * sum(u0)
    for j in range(S):
        A[0, 1 + j] = lam
    for j, ((r, m, k), n) in enumerate(zip(raw, counts)):
        r, m, k = r / n, m / n, k / n
        row = 1 + j
        if m > 0:
            # m/dt u+ + r/2 u+ + k/2 (g + dt/2 (u+ + u)) + k/2 g - p+/4 = m/dt u - r/2 u + (2p + p-)/4
            A[row, 1 + j] = m / dt + r / 2 + k * dt / 4
            A[row, 0] = -0.25
            b[row] = m / dt * u0[j] - r / 2 * u0[j] - k * dt / 4 * u0[j] - k * g0[j] + (2 * p + pp) / 4
        else:
            # r u+ + k (g + dt/2 (u+ + u)) = (p+ + p)/2
            A[row, 1 + j] = r + k * dt / 2
            A[row, 0] = -0.5
            b[row] = p / 2 - k * g0[j] - k * dt / 2 * u0[j]
    x = np.linalg.solve(A, b)
    return x[0], x[1:]


def test_boundary_update_solves_branch_equations():
    # fdtd.rs:581-598 closed form against the brute-force solve of the
    # coupled system, for mixed resistive / mass / spring branches and
    # merged faces (impedance / count, fdtd.rs:552-553).
    rng = np.random.default_rng(7)
    dt = configs.dt_of(0.05)
    lam = O.COURANT
    for trial in range(20):
        fits = []
        for _ in range(3):
            beta0 = rng.uniform(0.0, 0.3) if rng.random() < 0.6 else 0.0
            br = [(rng.uniform(0.5, 5.0), rng.uniform(0.0, 0.02) if rng.random() < 0.7 else 0.0,
                   rng.uniform(0.0, 2000.0) if rng.random() < 0.5 else 0.0) for _ in range(rng.integers(0, 3))]
            fits.append((beta0, br))
        faces = list(rng.integers(0, 3, rng.integers(1, 4)))
        slots = O.boundary_slots(faces, fits, dt)
        # The raw branches behind the slots, in slot order, with their face counts.
        raw, counts = [], []
        seen = []
        for f in faces:
            if f not in seen:
                seen.append(f)
        for f in seen:
            n = faces.count(f)
            b0, br = fits[f]
            if b0 > 0:
                raw.append((1.0 / b0, 0.0, 0.0))
                counts.append(n)
            for x in br:
                raw.append(x)
                counts.append(n)
        assert len(raw) == len(slots)
        if not raw:
            continue
        p, pp, lap = rng.normal(), rng.normal(), rng.normal() * 0.1
        u0 = rng.normal(size=len(raw)) * 0.05
        g0 = rng.normal(size=len(raw)) * 1e-4
        st = np.stack([u0, g0], axis=1).copy()
        nxt = O.boundary_update(slots, st, p, pp, lap, lam, dt)
        want_p, want_u = branch_system(raw, counts, p, pp, lap, lam, dt, u0, g0)
        assert abs(nxt - want_p) < 1e-10 * max(1.0, abs(want_p))
        assert np.allclose(st[:, 0], want_u, rtol=1e-9, atol=1e-12)
        assert np.allclose(st[:, 1], g0 + 0.5 * dt * (want_u + u0), rtol=1e-9, atol=1e-14)


def test_pure_conductance_is_the_classic_update():
    # fdtd.rs:28-30: (1 + lam b/2) p+ = 2p + lap - (1 - lam b/2) p- for m = k = 0.
    dt = configs.dt_of(0.1)
    lam = O.COURANT
    beta = 0.137
    slots = O.boundary_slots([0], [(beta, [])], dt)
    st = np.zeros((1, 2))
    rng = np.random.default_rng(3)
    p_prev, p = 0.0, 0.0
    for n in range(50):
        lap = rng.normal() * 0.1
        # Keep the state consistent with the scheme it came from.
        nxt = O.boundary_update(slots, st, p, p_prev, lap, lam, dt)
        classic = (2 * p + lap - (1 - lam * beta / 2) * p_prev) / (1 + lam * beta / 2)
        assert abs(nxt - classic) < 1e-12 * max(1.0, abs(classic))
        p_prev, p = p, nxt


def test_slots_merge_equal_fits_and_scale_by_count():
    dt = configs.dt_of(0.04)
    fa = (0.0, [(2.0, 0.01, 0.0)])
    fb = (0.25, [])
    s_a = O.boundary_slots([0], [fa, fb], dt)
    s_aa = O.boundary_slots([0, 0], [fa, fb], dt)
    s_ab = O.boundary_slots([0, 1], [fa, fb], dt)
    assert len(s_a) == 1 and len(s_aa) == 1 and len(s_ab) == 2
    # Two faces in parallel: r, m halved -> d halved -> inv_d doubled.
    assert abs(s_aa[0, 0] - 2 * s_a[0, 0]) < 1e-12 * s_a[0, 0]
    assert s_a[0, 3] == 0.25 and s_ab[1, 3] == 0.0
    # Resistive slot of beta0: inv_d = beta0.
    assert abs(s_ab[1, 0] - 0.25) < 1e-15


# -------------------------------------------------------------- grid

def test_shoebox_grid_matches_reference_sizes():
    # fdtd.rs:293-299: n = ceil(L/dx) + 2; one solid padding layer; the
    # survey's shoebox has 42 x 32 x 27 cells, 30000 air, 5528 boundary nodes.
    g = O.build_box(box_room((4.0, 3.0, 2.5), (0.03, [])), C, 0.1)
    assert g.n == (42, 32, 27)
    air = (g.flags & 64) != 0
    assert air.sum() == 40 * 30 * 25
    assert len(g.b_cell) == 40 * 30 * 25 - 38 * 28 * 23
    assert np.all(np.diff(g.b_cell.astype(np.int64)) > 0)
    # Flags bit d says the neighbour in direction d is air.
    f = g.flags.reshape(27, 32, 42)
    a = air.reshape(27, 32, 42)
    zz, yy, xx = np.nonzero(a)
    nbrs = [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]
    for d, (dz, dy, dx) in enumerate(nbrs):
        assert np.array_equal((f[zz, yy, xx] >> d) & 1, a[zz + dz, yy + dy, xx + dx].astype(np.uint8))
    # Corner node: three faces of one fit merge into one slot of 3x admittance.
    i = g.index(1, 1, 1)
    b = int(np.searchsorted(g.b_cell, i))
    assert g.b_cell[b] == i and g.b_start[b + 1] - g.b_start[b] == 1
    assert abs(g.slots[g.b_start[b], 0] - 3 * 0.03) < 1e-15


def test_pillar_cells_are_solid_and_faces_use_pillar_fits():
    fits = [(0.05, []), (0.2, [])]
    pil = [(1.0, 1.5, 1.0, 1.5)]
    patches = [(1, 1, [0])] * 6 + [(1, 1, [1])] * 4
    g = O.build_box({"L": (3.0, 3.0, 2.0), "pillars": pil, "fits": fits, "patches": patches}, C, 0.1)
    # Cells with centres inside the pillar are not air.
    for x in range(11, 16):
        for y in range(11, 16):
            assert not g.flags[g.index(x, y, 5)] & 64
    # The node just -x of the pillar has its +x face on the pillar: beta 0.2.
    i = g.index(10, 13, 5)
    b = int(np.searchsorted(g.b_cell, i))
    assert g.b_cell[b] == i
    assert abs(g.slots[g.b_start[b], 0] - 0.2) < 1e-15


# -------------------------------------------------------------- sources

def test_source_signal_shape():
    # fdtd.rs:601-610: s(t) = -(t - t0)/sigma^2 exp(-(t - t0)^2 / 2 sigma^2) dt,
    # sigma = 0.342/f_max, t0 = 6 sigma: odd about t0, extrema +-dt/sigma e^-1/2 at
    # t0 -+ sigma, starts at zero.
    f_max = 250.0
    sigma = 0.342 / f_max
    dt = sigma / 100.0
    s = O.source_signal(f_max, dt, 4000)
    n0 = 6 * sigma / dt
    k = int(round(n0))
    assert abs(k - n0) < 1e-9
    assert abs(s[k]) < 1e-12 * s.max()
    for j in (1, 17, 300, 590):
        assert abs(s[k + j] + s[k - j]) < 1e-9 * s.max()
    j = 100
    assert abs(s[k - j] - dt / sigma * math.exp(-0.5)) < 1e-9 * s.max()
    assert np.argmax(s) == k - j
    assert abs(s[0]) < 1e-6 * s.max()


def test_taps_are_trilinear_and_fall_back_to_nearest_air():
    g = O.build_box(box_room((6.0, 4.0, 3.0), (0.03, [])), C, 343.0 / (200.0 * 12.0))
    p = np.array([1.53, 1.27, 1.41])
    taps = O.taps_at(g, p)
    assert len(taps) == 8  # fdtd.rs:158-159 of the tests
    w = np.array([t[1] for t in taps])
    assert abs(w.sum() - 1.0) < 1e-12 and np.all(w > 0)
    centroid = sum(wi * g.center(c) for c, wi in taps)
    assert np.allclose(centroid, p, atol=1e-12)
    # Next to a wall most nodes are solid: renormalised over the air ones.
    q = np.array([0.01, 2.0, 1.5])
    t2 = O.taps_at(g, q)
    assert all(g.flags[c] & 64 for c, _ in t2)
    assert abs(sum(w for _, w in t2) - 1.0) < 1e-12


def test_nearest_air_is_the_nearest():
    g = O.build_box(box_room((2.0, 1.5, 1.2), (0.03, []), pillars=[(0.8, 1.2, 0.5, 0.9)]), C, 0.1)
    air = np.nonzero(g.flags & 64)[0]
    centres = np.array([g.center(i) for i in air])
    rng = np.random.default_rng(5)
    for _ in range(40):
        p = rng.uniform([-0.3, -0.3, -0.3], [2.3, 1.8, 1.5])
        i = O.nearest_air(g, p)
        d = ((centres - p) ** 2).sum(axis=1)
        assert abs(((g.center(i) - p) ** 2).sum() - d.min()) < 1e-12


def test_source_components_are_axis_dipoles():
    g = O.build_box(box_room((3.0, 3.0, 3.0), (0.03, [])), C, 0.1)
    c = g.index(10, 12, 14)
    comps = O.source_components(g, c, 1)
    assert comps[0] == [(c, 1.0)]
    nx, nxy = g.n[0], g.n[0] * g.n[1]
    for k, s in enumerate((1, nx, nxy)):
        assert comps[k + 1] == [(c + s, 1.0), (c - s, -1.0)]
    # Against a wall the dipoles do not fit: None (fdtd.rs:736-738).
    assert O.source_components(g, g.index(1, 12, 14), 1) is None
    assert O.source_components(g, g.index(1, 12, 14), 0) == [[(g.index(1, 12, 14), 1.0)]]


def test_component_weights_reproduce_first_order_pattern():
    # fdtd.rs:810-815: monopole + dipoles weighted by a/(2j sin k dx) give
    # a0 + a.d in the far field of the two-point dipole for small k dx.
    a0 = 0.5 + 0j
    a = [0.3 + 0.1j, -0.2 + 0j, 0.4j]
    k, dx = 2 * math.pi * 100 / C, 0.04
    w = O.component_weights(a0, a, k, dx)
    assert w[0] == a0
    for j in range(3):
        # weight * (e^{jk dx cos} - e^{-jk dx cos}) at cos = 1 is a_j exactly.
        assert abs(w[1 + j] * 2j * math.sin(k * dx) - a[j]) < 1e-14


# ----------------------------------------------------------- time loop

def run_spectra(g, taps, sig, rx, freqs):
    _, sp, st = O.run(g, taps, sig, rx, freqs=freqs, ir=False, threads=8)
    assert st == 0
    return sp


def test_direct_sound_is_calibrated():
    # tests/fdtd.rs:13-47: large room, receiver 1.5 m from the source on a grid
    # axis, run stopped before the first reflection: 1/r amplitude within
    # 0.1 dB and the phase lag of the analytic axial dispersion within 0.3 deg.
    f_max = 600.0
    dx = C / (f_max * 10.0)
    g = O.build_box(box_room((12.0, 10.0, 8.0), (1.0 / O.fit_real_impedance(0.2), [])), C, dx)
    src = O.nearest_air(g, (6.0, 5.0, 4.0))
    rcv = O.nearest_air(g, (7.5, 5.0, 4.0))
    r = float(np.linalg.norm(g.center(rcv) - g.center(src)))
    steps = int(0.016 / g.dt)
    sig = O.source_signal(f_max, g.dt, steps)
    freqs = [150.0, 250.0, 350.0, 450.0]
    sp = run_spectra(g, [(src, 1.0)], sig, [rcv], freqs)
    t = O.to_transfer(sp, sig, g, freqs, C, 1.0, np.ones(len(freqs)))
    for fi, f in enumerate(freqs):
        k = 2 * math.pi * f / C
        want = np.exp(-1j * k * r) / r
        db = 20 * math.log10(abs(t[fi]) / abs(want))
        deg = math.degrees(np.angle(t[fi] / want))
        predicted = -O.phase_error_deg_per_m(f, C, g.dx) * r
        assert abs(db) < 0.1, (f, db)
        assert abs(deg - predicted) < 0.3, (f, deg, predicted)


def box_modes(L, fmin, fmax, nmax):
    out = []
    for a in range(nmax):
        for b in range(nmax):
            for c in range(nmax):
                f = C / 2 * math.sqrt((a / L[0]) ** 2 + (b / L[1]) ** 2 + (c / L[2]) ** 2)
                if fmin < f < fmax:
                    out.append(f)
    return out


def test_rigid_box_eigenfrequencies():
    # tests/fdtd.rs:49-90: rigid 3 x 2.5 x 2 m box at dx = 0.1 m: the strongest
    # response within +-1.5 % of each well separated mode is within 0.5 %.
    L = (3.0, 2.5, 2.0)
    g = O.build_box(box_room(L, RIGID), C, 0.1)
    assert len(g.b_cell) == 0
    src = O.nearest_air(g, (0.05, 0.05, 0.05))
    rcv = O.nearest_air(g, (L[0] - 0.05, L[1] - 0.05, L[2] - 0.05))
    steps = int(2.0 / g.dt)
    sig = O.source_signal(400.0, g.dt, steps)
    freqs = 40.0 + np.arange(1000) * 0.08
    mag = np.abs(run_spectra(g, [(src, 1.0)], sig, [rcv], freqs)[0])
    modes = box_modes(L, 45.0, 115.0, 4)
    checked = 0
    for m in modes:
        if any(1e-9 < abs(o / m - 1) < 0.03 for o in modes):
            continue
        idx = np.nonzero(np.abs(freqs / m - 1) < 0.015)[0]
        k = idx[np.argmax(mag[idx])]
        assert abs(freqs[k] / m - 1) < 0.005, (m, freqs[k])
        checked += 1
    assert checked >= 4


@pytest.mark.slow
def test_rigid_cube_modes_to_800_hz():
    # tests/fdtd.rs:96-145 (Li et al. 2022 cube): 35 cells across 1 m, modes to
    # 800 Hz within 0.6 %.
    n = math.ceil(1.0 / (C / (800.0 * 15.0)))
    f_max = C * n / 15.0
    g = O.build_box(box_room((1.0, 1.0, 1.0), RIGID), C, C / (f_max * 15.0))
    src = O.nearest_air(g, (0.2, 0.2, 0.2))
    rcv = O.nearest_air(g, (0.9, 0.9, 0.9))
    steps = int(4.0 / g.dt)
    sig = O.source_signal(1600.0, g.dt, steps)
    freqs = 90.0 + np.arange(7400) * 0.1
    mag = np.abs(run_spectra(g, [(src, 1.0)], sig, [rcv], freqs)[0])
    modes = []
    for a in range(6):
        for b in range(a + 1):
            for c in range(b + 1):
                m = C / 2 * math.sqrt(a * a + b * b + c * c)
                if 100.0 <= m <= 800.0:
                    modes.append(m)
    worst, checked = 0.0, 0
    for m in modes:
        if any(1e-9 < abs(o / m - 1) < 0.03 for o in modes):
            continue
        idx = np.nonzero(np.abs(freqs / m - 1) < 0.015)[0]
        k = idx[np.argmax(mag[idx])]
        worst = max(worst, abs(freqs[k] / m - 1))
        checked += 1
    assert checked >= 5
    assert worst < 0.006


def test_tube_reflection_matches_admittance():
    # tests/fdtd.rs:262-321 with this build's first-order-section admittances
    # in place of fitted library materials: plane wave in a rigid tube hits
    # the end wall x = 0; |R_measured - (zeta - 1)/(zeta + 1)| < 0.03.
    f_max, ppw = 500.0, 15.0
    dx = C / (f_max * ppw)
    cells = lambda v: round(v / dx) * dx  # noqa: E731
    ln, w = cells(30.0), cells(0.3)
    rng = np.random.default_rng(11)
    for trial in range(2):
        fit = configs.first_order_fit(rng, alpha_range=(0.3, 0.6))
        patches = [(1, 1, [0]), (1, 1, [1]), (1, 1, [0]), (1, 1, [0]), (1, 1, [0]), (1, 1, [0])]
        room = {"L": (ln, w, w), "pillars": [], "fits": [RIGID, fit], "patches": patches}
        # Surface 0 is x = 0 in this layout: give the fit to it.
        room["patches"] = [(1, 1, [1])] + [(1, 1, [0])] * 5
        g = O.build_box(room, C, dx)
        src = O.nearest_air(g, (15.0, w / 2, w / 2))
        rcv = O.nearest_air(g, (2.0, w / 2, w / 2))
        xr = g.center(rcv)[0]
        steps = int(0.075 / g.dt)
        sig = O.source_signal(f_max, g.dt, steps)
        freqs = 40.0 + np.arange(15) * 22.0
        spec_all = run_spectra(g, [(src, 1.0)], sig, [rcv], freqs)[0]
        long = {"L": (ln * 2 + cells(15.0), w, w), "pillars": [], "fits": [RIGID], "patches": [(1, 1, [0])] * 6}
        gl = O.build_box(long, C, dx)
        shift = cells(30.0)
        src_l = O.nearest_air(gl, (15.0 + shift, w / 2, w / 2))
        rcv_l = O.nearest_air(gl, (xr + shift, w / 2, w / 2))
        spec_inc = run_spectra(gl, [(src_l, 1.0)], sig, [rcv_l], freqs)[0]
        worst = 0.0
        for fi, f in enumerate(freqs):
            om = 2 * math.pi * f
            k_num = 2.0 / g.dx * math.asin(math.sin(om * g.dt / 2) / O.COURANT)
            r_meas = (spec_all[fi] - spec_inc[fi]) / spec_inc[fi] * np.exp(1j * k_num * 2 * xr)
            zeta = 1.0 / O.admittance(fit, f)
            r_true = (zeta - 1) / (zeta + 1)
            worst = max(worst, abs(r_meas - r_true))
        assert worst < 0.03, worst


def test_branch_boundaries_decay():
    # tests/fdtd.rs:325-350 with this build's multi-section walls: the
    # spectrum's contribution from each successive third of the run shrinks
    # tenfold and nothing grows. The run length keeps the last third above
    # the FP64 noise floor for these walls.
    rng = np.random.default_rng(13)
    fits = [configs.first_order_fit(rng, alpha_range=(0.1, 0.3)) for _ in range(6)]
    room = {"L": (3.0, 2.4, 2.1), "pillars": [], "fits": fits, "patches": [(1, 1, [s]) for s in range(6)]}
    g = O.build_box(room, C, C / (300.0 * 12.0))
    assert np.any(np.diff(g.b_start) >= 3)
    src = O.nearest_air(g, (0.7, 0.6, 0.5))
    rcv = O.nearest_air(g, (2.3, 1.7, 1.6))
    steps = 9000
    sig = O.source_signal(300.0, g.dt, steps)
    ir, _, _ = O.run(g, [(src, 1.0)], sig, [rcv], threads=8)
    x = ir[0]
    assert np.all(np.isfinite(x))
    n = np.arange(steps)
    for f in (80.0, 200.0):
        e = x * np.exp(-2j * math.pi * f * g.dt * n)
        first, mid, last = (abs(e[: steps // 3].sum()), abs(e[steps // 3: 2 * steps // 3].sum()),
                            abs(e[2 * steps // 3:].sum()))
        assert mid < 0.1 * first and last < 0.1 * mid, (f, first, mid, last)


def test_ir_and_spectra_agree():
    # The DFT of the recorded series at the receivers is the run's spectra
    # (fdtd.rs:673-683): a direct numpy DFT of the IR in FP64.
    cfg = configs.shoebox(O.fit_real_impedance)
    g = O.build_box(cfg["room"], C, 0.1)
    sig = O.source_signal(250.0, g.dt, 600)
    taps = O.taps_at(g, cfg["sources"][0]["pos"])
    rx = [O.nearest_air(g, p) for p in cfg["receivers"]]
    freqs = [20.0, 77.7, 250.0]
    ir, sp, _ = O.run(g, taps, sig, rx, freqs=freqs, threads=4)
    n = np.arange(600)
    for fi, f in enumerate(freqs):
        x = f * g.dt * n
        ph = -2 * math.pi * (x - np.trunc(x))
        want = (ir * np.exp(1j * ph)[None, :]).sum(axis=1)
        assert np.allclose(sp[:, fi], want, rtol=1e-12, atol=1e-18)
    assert np.all(ir[:, 0] == 0.0)  # p^0 is recorded before any injection


def test_threads_do_not_change_results():
    cfg = configs.pillared_room()
    g = O.build_box(cfg["room"], C, 0.16)
    sig = O.source_signal(250.0, g.dt, 200)
    node = O.nearest_air(g, cfg["sources"][0]["pos"])
    rx = [O.nearest_air(g, p) for p in cfg["receivers"][:8]]
    a, _, _ = O.run(g, [(node, 1.0)], sig, rx, threads=1)
    b, _, _ = O.run(g, [(node, 1.0)], sig, rx, threads=8)
    assert np.array_equal(a, b)


def test_cancel_returns_cancelled():
    cfg = configs.shoebox(O.fit_real_impedance)
    g = O.build_box(cfg["room"], C, 0.1)
    sig = O.source_signal(250.0, g.dt, 300)
    _, _, st = O.run(g, [(g.index(10, 10, 10), 1.0)], sig, [g.index(5, 5, 5)], threads=1, cancel=True)
    assert st == 1
