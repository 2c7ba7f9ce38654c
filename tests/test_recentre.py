"""CPU: the FP32 re-centring of the device kernels (wv::recentre), restated
in numpy on a small rigid room.

The interior update c + ((c - q) + l2 sum(p_nb - c)) and the slot-free
boundary update of the kernels are formed from differences, so a uniform
shift of p^n and p^{n-1} leaves them unchanged in exact arithmetic: checked
in FP64 (shifted and unshifted runs agree to rounding). In FP32 a field that
holds a mean pressure D resolves the waves on top of it only in steps of
ulp(D): with D = 1e-3 and a wave of 1e-7 the plain FP32 series is far from
FP64, while the same field stored relative to its mean (the run's offset,
the records getting it back in FP64 at read-out) is within 1e-4, and stays
there through re-centring at steps 128, 256, 512, ... (while the mean
persists, the stored mean of p^n over the air cells moves from both fields
into the offset). The full
schedule on a room whose mean builds up from the source is checked on the
device (the spring-branch room of test_gpu_parity.py).
"""
import numpy as np

L2 = 1.0 / 3.0


def run(steps, dtype, recentre=False, shift0=0.0, mean=1e-3, bump=1e-7, dims=(20, 18, 16)):
    """Receiver series of a rigid room that starts at rest at pressure
    `mean` with a Gaussian bump of height `bump`; fields of `dtype` holding
    the pressure less the offset (initially shift0), the update in the
    kernels' difference form, records p^n + offset before the update."""
    nx, ny, nz = dims
    air = np.zeros((nz, ny, nx), bool)
    air[1:-1, 1:-1, 1:-1] = True
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    g = bump * np.exp(-((x - 8.3) ** 2 + (y - 7.1) ** 2 + (z - 6.4) ** 2) / 4.0)
    l2 = dtype(L2)
    if dtype == np.float32 and np.float64(l2) > L2:
        l2 = np.nextafter(l2, dtype(0))
    p = np.where(air, (mean - shift0) + g, 0.0).astype(dtype)
    q = p.copy()
    d0 = float(shift0)
    nb = [np.roll(air, s, axis=a) & air for a, s in ((2, 1), (2, -1), (1, 1), (1, -1), (0, 1), (0, -1))]
    rx = [(5, 5, 5), (10, 12, 15), (14, 16, 18)]
    out = np.zeros((len(rx), steps))
    n_air = int(air.sum())
    prev_mean = 0.0
    for n in range(steps):
        if recentre and n >= 128 and n & (n - 1) == 0:
            # The check of wv::recentre: the first records the pressure's
            # mean; later ones move the stored mean into the offset while the
            # pressure's mean persists (at least half of the last check's),
            # else the offset back into the field.
            stored = float(np.sum(p[air], dtype=np.float64) / n_air)
            pmean = stored + d0
            m = dtype(0)
            if n > 128:
                m = dtype(stored) if abs(pmean) >= 0.5 * abs(prev_mean) else dtype(-d0)
            prev_mean = pmean
            p = (p - m).astype(dtype)
            q = (q - m).astype(dtype)
            d0 += float(m)
        for k, r in enumerate(rx):
            out[k, n] = float(p[r]) + d0  # the stored value, its offset added in FP64 at read-out
        d = np.zeros_like(p)
        for k, (a, s) in enumerate(((2, 1), (2, -1), (1, 1), (1, -1), (0, 1), (0, -1))):
            d = (d + np.where(nb[k], np.roll(p, s, axis=a) - p, dtype(0))).astype(dtype)
        new = (p + ((p - q) + l2 * d)).astype(dtype)
        q = np.where(air, new, q)
        p, q = q, p
    return out - mean


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_uniform_shift_changes_nothing_in_exact_arithmetic():
    a = run(400, np.float64)
    b = run(400, np.float64, shift0=3.7e-4)
    c = run(400, np.float64, recentre=True)
    assert rel(b, a) < 1e-6 and rel(c, a) < 1e-6  # the wave is 1e-4 of the pressure: FP64 keeps 1e-12 of it


def test_offset_restores_fp32_resolution():
    """The same room stored relative to its mean (the state re-centring
    brings a run to): FP32 within 1e-4 of FP64, where the field that holds
    the mean is not; re-centring at steps 128, 256, ... keeps it there."""
    ref = run(1200, np.float64)
    plain = rel(run(1200, np.float32), ref)
    offset = rel(run(1200, np.float32, shift0=1e-3), ref)
    kept = rel(run(1200, np.float32, shift0=1e-3, recentre=True), ref)
    assert plain > 1e-2, plain  # ulp(1e-3) is about 1e-10, the wave about 1e-9 at the receivers
    assert offset < 1e-4 and kept < 1e-4, (plain, offset, kept)
