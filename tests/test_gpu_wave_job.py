"""GPU: the wave job's FDTD part (wv_wave_job, run.rs:593-661) against the
same composition done with the oracle: every component's FP64 spectra
(fdtd.rs:644-725), to_transfer (fdtd.rs:819-840) and the cardioid's
component weights (fdtd.rs:812-815), summed per source.

FP64 device against the oracle: 1e-9 relative L2 over the transfer
functions; FP32: 1e-4. The batched and one-plan forms are bit-identical.
"""
import numpy as np
import pytest

import oracle
from wavey_amd import configs, fdtd, workload

pytestmark = pytest.mark.gpu

FREQS = [31.5, 50.0, 80.0, 125.0, 200.0]


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def job():
    # The pillared room's cardioid plus an omni source, coarse grid, 1200 steps.
    cfg = dict(configs.pillared_room(), dx=0.08, steps=1200)
    cfg["sources"] = list(cfg["sources"]) + [{"pos": (3.0, 2.5, 1.5), "kind": "omni"}]
    case = workload.build_case(cfg)
    sources = workload.wave_sources(case, FREQS)
    for s in sources:
        s["h0"] = np.exp(1j * np.linspace(0.1, 0.5, len(FREQS)))
        s["r_ref"] = 1.0
    og = oracle.build_box(cfg["room"], cfg["c"], cfg["dx"])
    ref = []
    for s in sources:
        t = np.zeros((len(case.receivers), len(FREQS)), dtype=np.complex128)
        for ci, taps in enumerate(s["comps"]):
            _, spec, _ = oracle.run(og, taps, case.signal, case.receivers, freqs=FREQS, ir=False, threads=8)
            tc = oracle.to_transfer(spec, case.signal, og, FREQS, cfg["c"], s["r_ref"], s["h0"])
            tc = np.asarray(tc).reshape(len(case.receivers), len(FREQS))
            w = np.ones((len(FREQS), 1)) if s["weights"] is None else np.asarray(s["weights"])
            t += w[:, ci][None, :] * tc
        ref.append(t)
    return case, sources, np.array(ref)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-4)])
def test_wave_job_matches_oracle_composition(job, prec, tol):
    case, sources, ref = job
    print(f"case: wave job, pillared room dx 0.08, cardioid + omni, {prec}", flush=True)
    one = fdtd.wave_job(case.grid, sources, case.signal, case.receivers, FREQS, case.cfg["c"], prec, batched=False)
    bat = fdtd.wave_job(case.grid, sources, case.signal, case.receivers, FREQS, case.cfg["c"], prec, batched=True)
    assert one.shape == ref.shape
    assert np.array_equal(one, bat)
    for s in range(len(sources)):
        assert rel_l2(one[s], ref[s]) < tol
