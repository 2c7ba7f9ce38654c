"""GPU: the batched and served forms at their full configurations.

* The uneven room mix: 48 seeded rooms at dx 0.05 m, 1-4 omni speakers each
  solved as its own instance (103 instances), 0.5 s = 5941 steps, 16
  receivers, one autotuned batch. Every instance's receiver series is
  bit-identical to its plain run (the same kernels and arithmetic per cell);
  a sample is within 1e-4 relative L2 of an FP64 run of the same inputs.
* The serving trace: 40 requests with padded cells log-uniform in
  0.2M-60M and 0.3-1.0 s of simulated time, submitted from a client thread
  at Poisson arrival times that offer 80 % of 200 Gcell-updates/s. Requests
  that arrive while others run are admitted and batched (the server reports
  more than one running at once); each result equals the request's plain
  run alone, bit for bit, and a sample is within 1e-4 of FP64.

The reference runs these solves one after another (run.rs:645-661).
"""
import threading
import time

import numpy as np
import pytest

from wavey_amd import configs, fdtd, workload

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_room_mix_batch_equals_plain_runs():
    cases = [workload.build_case(cfg) for cfg in configs.room_mix()]
    inst = [(c, taps) for c in cases for (_, _, taps) in c.runs]
    assert len(cases) == 48 and len(inst) == 103 and len(cases[0].signal) == 5941
    print(f"case: room mix, {len(inst)} instances, plain runs", flush=True)
    plain = []
    t0 = time.perf_counter()
    for c in cases:
        with fdtd.Plan(c.grid, "fp32") as p:
            for (_, _, taps) in c.runs:
                plain.append(p.run(taps, c.signal, c.receivers, mode=fdtd.MODE_IR)[0])
    t_plain = time.perf_counter() - t0
    print(f"case: room mix, one batch (autotuned); plain runs took {t_plain:.2f} s", flush=True)
    t0 = time.perf_counter()
    with fdtd.Batch("fp32") as b:
        ids = b.add_many([(c.grid, taps, c.signal, c.receivers) for c, taps in inst], threads=8)
        b.autotune(trial_steps=8)
        assert b.run() == fdtd.WV_OK
        batched = [b.result(i)[0] for i in ids]
    t_batch = time.perf_counter() - t0
    same = [k for k in range(len(inst)) if np.array_equal(batched[k], plain[k])]
    print(f"room mix: {len(same)}/{len(inst)} instances bit-identical to plain; plain {t_plain:.2f} s, "
          f"batch {t_batch:.2f} s ({t_plain / t_batch:.2f}x)", flush=True)
    assert len(same) == len(inst)
    for k in range(0, len(inst), 17):
        c, taps = inst[k]
        with fdtd.Plan(c.grid, "fp64") as p:
            ref = p.run(taps, c.signal, c.receivers, mode=fdtd.MODE_IR)[0]
        err = rel_l2(batched[k], ref)
        print(f"room mix: instance {k} ({c.grid.n}) FP32 vs FP64 {err:.2e}", flush=True)
        assert err < 1e-4


def test_serving_trace_each_result_equals_its_plain_run():
    trace = configs.serving_trace(1.0, 40)
    cases = [workload.build_case(cfg) for _, cfg in trace]
    works = [c.grid.air_cells() * len(c.signal) for c in cases]
    rate = 0.8 * 200e9 / float(np.mean(works))
    arrivals = [t / rate for t, _ in trace]
    sizes = sorted(c.grid.cells for c in cases)
    assert sizes[0] < 0.3e6 and sizes[-1] > 40e6
    print(f"case: serving trace, {len(cases)} requests, rate {rate:.2f}/s, plain runs", flush=True)
    plain = []
    for c in cases:
        with fdtd.Plan(c.grid, "fp32") as p:
            plain.append(p.run(c.runs[0][2], c.signal, c.receivers, mode=fdtd.MODE_IR)[0])
    print("case: serving trace through the server (lanes policy)", flush=True)
    ids = [None] * len(cases)
    with fdtd.Server("fp32") as server:
        start = time.perf_counter()

        def client():
            for k, (a, c) in enumerate(zip(arrivals, cases)):
                dt = start + a - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids[k] = server.submit(c.grid, c.runs[0][2], c.signal, c.receivers)

        th = threading.Thread(target=client)
        th.start()
        th.join()
        served, lat = [], []
        for k in range(len(cases)):
            ir, _, latency = server.wait(ids[k])
            served.append(ir)
            lat.append(latency)
        st = server.stats()
    same = sum(int(np.array_equal(x, y)) for x, y in zip(served, plain))
    print(f"serving: {same}/{len(cases)} bit-identical to plain runs alone; at most {int(st['max_running'])} "
          f"running at once; p50 {np.percentile(lat, 50):.3f} s, p99 {np.percentile(lat, 99):.3f} s", flush=True)
    assert same == len(cases)
    assert st["failed"] == 0 and st["max_running"] >= 2
    for k in range(0, len(cases), 10):
        c = cases[k]
        with fdtd.Plan(c.grid, "fp64") as p:
            ref = p.run(c.runs[0][2], c.signal, c.receivers, mode=fdtd.MODE_IR)[0]
        err = rel_l2(served[k], ref)
        print(f"serving: request {k} ({c.grid.n}) FP32 vs FP64 {err:.2e}", flush=True)
        assert err < 1e-4
