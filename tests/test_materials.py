"""Material absorption and admittance fits on the CPU: Spectrum::eval
(spectrum.rs:24-44), Material::absorption_at (material.rs:72-93) and
fit_admittance (impedance_fit.rs:158-235).

The oracle's restatement is pinned by the reference's own tests on its
built-in material library (tests/golden/library_materials.json is the
reference's core/wavey-materials/src/library.json, data only):
every material fits over 20-424 Hz with max error < 0.05 and RMS < 0.02
(impedance_fit.rs:241-252), every branch has Q <= 2 over 20-707 Hz
(:254-261), every fit is passive (:263-272), the standing audience is
nearly transparent at 40 Hz and absorbing at 300 Hz (:274-282); and by
closed forms: a constant absorption is fitted by a conductance alone, the
one whose 48-angle Paris absorption is alpha; no absorption gives the rigid fit; the
spectrum's log interpolation (spectrum.rs:76-82); a matched impedance
absorbs what the Paris integral of R = (cos - 1)/(cos + 1) gives. The
library's fits are the oracle's bit for bit.
"""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from wavey_amd import fdtd

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = json.load(open(os.path.join(HERE, "golden", "library_materials.json")))


def admittance(fit, f):
    beta0, branches = fit
    jw = 2j * np.pi * max(f, 1e-6)
    return beta0 + sum(1.0 / (r + jw * m + k / jw) for r, m, k in branches)


@pytest.fixture(scope="module")
def low_band_fits():
    with ThreadPoolExecutor(max_workers=8) as ex:
        mine = list(ex.map(lambda m: fdtd.fit_admittance(m, (20.0, 424.0), 2), LIBRARY))
        ref = list(ex.map(lambda m: oracle.fit_admittance(m, (20.0, 424.0), 2), LIBRARY))
    return mine, ref


def test_library_fits_over_low_band(low_band_fits):
    """impedance_fit.rs:241-252, and the library's fits are the oracle's."""
    mine, ref = low_band_fits
    assert len(LIBRARY) == 26
    for m, a, b in zip(LIBRARY, mine, ref):
        assert a == b, m["id"]
        assert b[1]["max_error"] < 0.05, m["id"]
        assert b[1]["rms_error"] < 0.02, m["id"]
        assert b[1]["band"] == (20.0, 424.0)


def test_library_fits_are_passive(low_band_fits):
    """impedance_fit.rs:263-272"""
    _, ref = low_band_fits
    for m, (fit, _) in zip(LIBRARY, ref):
        for i in range(200):
            assert admittance(fit, 1.0 + i * 25.0).real >= 0.0, m["id"]


def test_audience_is_nearly_transparent_at_low_frequency(low_band_fits):
    """impedance_fit.rs:274-282"""
    _, ref = low_band_fits
    k = [m["id"] for m in LIBRARY].index("audience-standing")
    fit = ref[k][0]
    assert oracle.paris(admittance(fit, 40.0)) < 0.08
    assert oracle.paris(admittance(fit, 300.0)) > 0.3


def test_branches_are_broad():
    """impedance_fit.rs:254-261 over 20-707 Hz."""
    with ThreadPoolExecutor(max_workers=8) as ex:
        fits = list(ex.map(lambda m: oracle.fit_admittance(m, (20.0, 707.0), 2)[0], LIBRARY))
    for m, (_, branches) in zip(LIBRARY, fits):
        for r, mm, k in branches:
            assert np.sqrt(mm * k) / r <= 2.0 + 1e-9, m["id"]


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5])
def test_constant_absorption_is_a_conductance(alpha):
    """The fit's cost integrates Paris with 48 angles (impedance_fit.rs:192),
    so the conductance solves paris48(beta) = alpha; the reported error
    integrates with 96 (:183)."""
    mat = {"id": "flat", "absorption": alpha}
    fit, err = oracle.fit_admittance(mat, (20.0, 500.0), 2)
    assert fit[1] == []
    assert abs(oracle.paris(fit[0], 48) - alpha) < 1e-7  # the simplex stops at a cost spread of 1e-14
    beta = 1.0 / oracle.fit_real_impedance(alpha)  # the 256-angle solution
    assert abs(fit[0] - beta) < 0.01 * beta
    e96 = abs(oracle.paris(fit[0], 96) - alpha)
    assert abs(err["rms_error"] - e96) < 1e-15 and abs(err["max_error"] - e96) < 1e-15
    assert fdtd.fit_admittance(mat, (20.0, 500.0), 2) == (fit, err)


def test_no_absorption_is_rigid():
    fit, err = oracle.fit_admittance({"id": "rigid", "absorption": 0.0}, (20.0, 500.0), 2)
    assert fit == (0.0, []) and err["rms_error"] == 0.0 and err["max_error"] == 0.0
    assert fdtd.fit_admittance({"id": "rigid", "absorption": 0.0}, (20.0, 500.0), 2) == (fit, err)


def test_spectrum_and_absorption():
    """spectrum.rs:76-82; material.rs:72-93, 137-143."""
    s = {"absorption": {"freqs_hz": [100.0, 1000.0], "values": [0.0, 1.0]}}
    assert oracle.material_absorption(s, 50.0) == 0.0
    assert oracle.material_absorption(s, 5000.0) == 1.0
    assert abs(oracle.material_absorption(s, 316.227766) - 0.5) < 1e-6
    assert abs(oracle.material_absorption(s, 10 ** 2.25) - 0.25) < 1e-12
    assert oracle.material_absorption({"absorption": 1.7}, 100.0) == 1.0  # clamped
    # Impedance-defined: matched (zeta = 1) and near rigid.
    th = (np.arange(256) + 0.5) * (np.pi / 2 / 256)
    r = (np.cos(th) - 1.0) / (np.cos(th) + 1.0)
    matched = float(np.sum((1 - r * r) * np.sin(2 * th) * (np.pi / 2 / 256)))
    got = oracle.material_absorption({"absorption": 0.0, "impedance": (1.0, 0.0)}, 100.0)
    assert abs(got - matched) < 1e-12 and 0.9 < got < 0.96
    assert oracle.material_absorption({"absorption": 0.0, "impedance": (1e9, 0.0)}, 100.0) < 1e-8
    for mat in [s, {"absorption": 0.3}, {"absorption": 0.0, "impedance": (3.0, {"freqs_hz": [50, 500], "values": [-1, 2]})}]:
        for f in [30.0, 100.0, 250.0, 700.0]:
            assert fdtd.material_absorption(mat, f) == oracle.material_absorption(mat, f)


def test_fit_surfaces_per_material():
    """fdtd.rs:213-222: one fit per distinct id, returned per surface."""
    mats = [LIBRARY[0], LIBRARY[2], LIBRARY[0]]
    fits = fdtd.fit_surfaces(mats, (20.0, 424.0))
    for k, m in ((0, LIBRARY[0]), (1, LIBRARY[2]), (2, LIBRARY[0])):
        fit, meta = oracle.fit_admittance(m, (20.0, 424.0))
        assert fits[k][:2] == fit and fits[k][2] == meta
