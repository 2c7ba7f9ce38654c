"""GPU: the library's cabinet calibration (wv_calibrate_cabinet,
fdtd.rs:873-955) against the FP64 oracle's (ob_calibrate_cabinet, run_cpu_with
as the runner) on the same cabinet.

FP64 runs: within 1e-9 relative L2 over every on-axis and directivity value.
FP32 runs: within 1e-4. The batched and the one-after-another forms give the
same values bit for bit.
"""
import time

import numpy as np
import pytest

import oracle
from wavey_amd import fdtd, meshes

pytestmark = pytest.mark.gpu

C = 343.0
FREQS = [10.0, 25.0, 40.0, 60.0, 80.0, 100.0]


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def reference():
    v, t = meshes.cuboid((0.4, 0.3, 0.5))
    ac = (0.15, 0.0, 0.05)
    ref = oracle.calibrate_cabinet(v, t, ac, C, 100.0, 10.0, 3.0, FREQS, threads=8)
    return v, t, ac, ref


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-4)])
def test_calibration_matches_oracle(reference, prec, tol):
    v, t, ac, ref = reference
    print(f"case: cabinet 0.4x0.3x0.5 m, dx 0.1 m, {prec}, batched and plan", flush=True)
    t0 = time.perf_counter()
    got = fdtd.calibrate_cabinet(v, t, ac, C, 100.0, 10.0, 3.0, FREQS, precision=prec, batched=True)
    print(f"calibration (two runs, one batch): {time.perf_counter() - t0:.3f} s")
    e_on = rel_l2(got["on_axis"], ref["on_axis"])
    e_dir = rel_l2(got["directivity"], ref["directivity"])
    print(f"relative L2: on axis {e_on:.3e}, directivity {e_dir:.3e}")
    assert np.array_equal(got["angles_deg"], ref["angles_deg"])
    assert e_on < tol and e_dir < tol
    assert np.all(got["directivity"][0] == 1.0)
    plain = fdtd.calibrate_cabinet(v, t, ac, C, 100.0, 10.0, 3.0, FREQS, precision=prec, batched=False)
    assert np.array_equal(plain["on_axis"], got["on_axis"])
    assert np.array_equal(plain["directivity"], got["directivity"])


def test_calibration_cancel(reference):
    v, t, ac, _ = reference
    flag = np.ones(1, dtype=np.int32)
    print("case: cabinet calibration cancelled before it starts", flush=True)
    assert fdtd.calibrate_cabinet(v, t, ac, C, 100.0, 10.0, 3.0, FREQS, batched=True, cancel=flag) is None
    assert fdtd.calibrate_cabinet(v, t, ac, C, 100.0, 10.0, 3.0, FREQS, batched=False, cancel=flag) is None
