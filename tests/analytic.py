"""Closed forms for the analytic wave tests (test infrastructure).

The spherical-wave reflection factor of a locally reacting plane
(Weyl-van der Pol, in the form of Chien and Soroka), as the reference states
it in core/wavey-solver/src/reflection.rs:71-88, in the solver's e^{jwt}
convention; the Faddeeva function is scipy.special.wofz.
"""
import numpy as np
from scipy.special import wofz


def spherical(zeta, cos, k, r):
    """Reflection factor for normalized impedance zeta, incidence cosine cos
    (to the normal), wavenumber k and unfolded path length r
    (reflection.rs:74-82)."""
    cos = min(max(cos, 0.0), 1.0)
    beta = 1.0 / np.conj(zeta)  # literature convention: conjugate impedance
    rp = (cos - beta) / (cos + beta)
    w = complex(0.5, 0.5) * np.sqrt(k * r) * (cos + beta)
    f = 1.0 + 1j * np.sqrt(np.pi) * w * wofz(w)
    return np.conj(rp + (1.0 - rp) * f)


def plane(zeta, cos):
    """Plane-wave reflection factor (reflection.rs:85-88)."""
    zc = zeta * max(cos, 1e-6)
    return (zc - 1.0) / (zc + 1.0)
