"""CPU: the closed forms of tests/analytic.py against the reference's own
checks of them (core/wavey-solver/src/reflection.rs:94-125)."""
import numpy as np
from scipy.special import wofz

from analytic import plane, spherical


def test_faddeeva_reference_values():
    # Abramowitz & Stegun 7.1 (reflection.rs:95-111): w(0) = 1, w(i) = e erfc(1),
    # w(1), w(2 + i), and the lower half-plane.
    cases = [(0j, 1 + 0j), (1j, 0.427583576155807 + 0j), (1 + 0j, 0.367879441171442 + 0.607157705841394j),
             (2 + 1j, 0.140239581366278 + 0.222213440179899j)]
    for z, want in cases:
        assert abs(wofz(z) - want) < 1e-12
    got = wofz(-1j)
    assert abs(got.real - (2 * np.e - 0.427583576155807)) < 1e-11 and abs(got.imag) < 1e-11


def test_spherical_limits():
    # reflection.rs:113-125: a rigid surface reflects fully; far from the
    # source Q approaches the plane-wave factor; near grazing on a hard
    # surface Q stays near +1 while the plane-wave factor goes to -1.
    zeta = 12 - 8j
    assert abs(spherical(1e12 + 0j, 0.3, 5.0, 10.0) - 1.0) < 1e-6
    assert abs(spherical(zeta, 0.6, 50.0, 1e4) - plane(zeta, 0.6)) < 1e-3
    assert spherical(80 - 20j, 0.01, 5.0, 10.0).real > 0.5
    assert plane(80 - 20j, 0.01).real < 0.0


def test_spherical_normal_incidence_far_field():
    # At normal incidence and large kr the factor is (zeta - 1)/(zeta + 1).
    for zeta in (3.0 + 1j, 1.5 - 0.5j, 40.0 + 0j):
        assert abs(spherical(zeta, 1.0, 200.0, 1e4) - (zeta - 1) / (zeta + 1)) < 2e-3
