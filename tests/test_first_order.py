"""fit_first_order (fdtd.rs:749-807) on the CPU.

Pins: the sampling directions are unit vectors spread evenly in z (the
golden-angle spiral); a cardioid D = 1/2 + 1/2 u.axis is first order, so
under any rotation its fit is exact (a0 = 1/2, a = R x / 2, residual 0);
a pattern with second-order terms gets the least-squares fit constrained to
the on-axis value, which is checked against the same problem solved as a
KKT system with numpy (an independent solver); the fit meets the on-axis
value exactly. The library's fit is the oracle's bit for bit.
"""
import numpy as np

import oracle
from wavey_amd import fdtd, meshes


def test_directions():
    u = oracle.first_order_directions()
    assert u.shape == (400, 3)
    np.testing.assert_allclose(np.linalg.norm(u, axis=1), 1.0, atol=1e-15)
    np.testing.assert_allclose(u[:, 2], 1.0 - 2.0 * (np.arange(400) + 0.5) / 400, rtol=0, atol=1e-15)
    assert abs(u.mean(axis=0)).max() < 0.01
    assert np.array_equal(fdtd.first_order_directions(), u)


def test_cardioid_is_exact():
    u = oracle.first_order_directions()
    for seed in range(5):
        R = meshes.rotation(np.random.default_rng(seed).normal(size=3), 0.3 + seed)
        d = 0.5 + 0.5 * u[:, 0]
        a0, a, res = oracle.fit_first_order(d, 1.0, R)
        assert abs(a0 - 0.5) < 1e-13
        np.testing.assert_allclose(np.array(a), 0.5 * R[:, 0], atol=1e-13)
        assert res < 1e-13
        assert fdtd.fit_first_order(d, 1.0, R) == (a0, a, res)


def test_second_order_pattern_is_the_constrained_least_squares():
    u = oracle.first_order_directions()
    R = meshes.rotation((0.2, -0.4, 1.0), 0.9)
    pat = lambda v: (0.2 + 0.1j) + 0.4 * v[..., 0] + 0.25 * v[..., 0] ** 2 - 0.1j * v[..., 1] * v[..., 2]  # noqa: E731
    d = pat(u)
    want = pat(np.array([1.0, 0.0, 0.0]))
    a0, a, res = oracle.fit_first_order(d, want, R)
    # KKT: minimise |B c - d|^2 subject to e.c = want, B rows (1, R u).
    B = np.hstack([np.ones((400, 1)), u @ R.T])
    e = np.concatenate([[1.0], R[:, 0]])
    K = np.zeros((5, 5))
    K[:4, :4] = 2 * B.T @ B
    K[:4, 4] = e
    K[4, :4] = e
    sol = np.linalg.solve(K.astype(complex), np.concatenate([2 * B.T @ d, [want]]))
    c = np.array([a0] + a)
    np.testing.assert_allclose(c, sol[:4], rtol=0, atol=1e-12)
    assert abs(c @ e - want) < 1e-13
    fit = B @ c
    assert abs(res - np.sqrt(np.sum(abs(fit - d) ** 2) / np.sum(abs(d) ** 2))) < 1e-13
    assert 0.01 < res < 0.5
    assert fdtd.fit_first_order(d, want, R) == (a0, a, res)
