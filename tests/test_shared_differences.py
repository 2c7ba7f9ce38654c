"""CPU: the FP32 interior update formed from shared neighbour differences is
the same, bit for bit, as the update formed from six differences per cell.

The device tiles (``row_update`` in wavey_amd/csrc/step_kernels.h) form
``(xm - c) + (xp - c)`` as ``dx[e+1] - dx[e]`` from one difference per pair of
neighbouring cells, and ``(zm - c) + (zp - c)`` as ``dzp - dzm`` with ``dzm``
carried from the plane below; the ring and special cells (``interior_update``)
form every difference. The two agree because round-to-nearest is symmetric:
RN(a - b) = -RN(b - a), and x + (-y) is the operation x - y. The check runs
both orders in float32 (numpy rounds each operation to nearest) over fields
chosen where rounding matters: near-uniform values (the room's DC offset),
values across a binade boundary, wide dynamic range, and exact zeros.
"""
import numpy as np
import pytest

F = np.float32


def six_differences(c, xm, xp, ym, yp, zm, zp):
    # interior_update's order: ((xm - c) + (xp - c)) + ((ym - c) + (yp - c)) + ((zm - c) + (zp - c))
    return ((xm - c) + (xp - c)) + ((ym - c) + (yp - c)) + ((zm - c) + (zp - c))


def shared_differences(row, ym, yp, zm_plane, zp_plane):
    """A row of cells c[0..n-1] with x neighbours row[-1] and row[n] at the ends."""
    c = row[1:-1]
    dx = row[1:] - row[:-1]          # dx[e] = c[e] - (the cell before it)
    dzm = c - zm_plane               # carried from the plane below: zp - c there
    dzp = zp_plane - c
    return (dx[1:] - dx[:-1]) + ((ym - c) + (yp - c)) + (dzp - dzm)


FIELDS = {
    # near-uniform: a DC offset with small waves on it
    "dc": lambda r, n: F(7.3e-4) + (r.standard_normal(n) * 1e-9).astype(F),
    # across the binade boundary at 2^-10
    "binade": lambda r, n: (F(2.0 ** -10) * (1 + (r.standard_normal(n) * 1e-3))).astype(F),
    # wide dynamic range, both signs
    "wide": lambda r, n: (r.standard_normal(n) * np.exp(r.uniform(-30, 10, n))).astype(F),
    # many exact ties and zeros
    "zeros": lambda r, n: (r.integers(-2, 3, n) * F(0.25)).astype(F),
}


@pytest.mark.parametrize("kind", sorted(FIELDS))
def test_shared_differences_are_bit_identical(kind):
    r = np.random.default_rng(17 + sorted(FIELDS).index(kind))
    n = 4096
    gen = FIELDS[kind]
    for _ in range(25):
        row = gen(r, n + 2)
        ym, yp, zm, zp = (gen(r, n) for _ in range(4))
        c = row[1:-1]
        full = six_differences(c, row[:-2], row[2:], ym, yp, zm, zp)
        shared = shared_differences(row, ym, yp, zm, zp)
        assert full.dtype == F and shared.dtype == F
        # bit for bit, except that a zero may differ in sign (both compare equal
        # and a zero sum is never divided by)
        same = (full.view(np.uint32) == shared.view(np.uint32)) | ((full == 0) & (shared == 0))
        assert same.all(), (kind, np.flatnonzero(~same)[:5])


def test_an_unshared_order_would_differ():
    """The identity needs the pairing: summing x differences in another order
    (xm - c, then + (ym - c), ...) rounds differently on the same fields, so
    the check above would catch a reordered kernel."""
    r = np.random.default_rng(5)
    n = 4096
    row = FIELDS["wide"](r, n + 2)
    ym, yp, zm, zp = (FIELDS["wide"](r, n) for _ in range(4))
    c = row[1:-1]
    full = six_differences(c, row[:-2], row[2:], ym, yp, zm, zp)
    other = (((((row[:-2] - c) + (ym - c)) + (row[2:] - c)) + (yp - c)) + (zm - c)) + (zp - c)
    assert (full.view(np.uint32) != other.view(np.uint32)).any()
