"""GPU: the batched launch and the request server against plain runs.

Each instance of a batch, and each served request, runs the same kernels and
arithmetic per cell as its plain run (``Plan.run``), whatever the tile shape
and whichever instances share the launch, so the receiver series must be
bit-identical. The plain FP64 path is itself checked against the FP64 oracle
in test_gpu_parity.py, and one batched FP64 instance is checked against the
oracle here (1e-9 relative L2).
"""
import time

import numpy as np
import pytest

import oracle
from wavey_amd import configs, fdtd, workload

pytestmark = pytest.mark.gpu


def small_rooms(n, seed=7, steps=(260, 700)):
    """Mix-recipe rooms at dx 0.1 m with different sizes and step counts (not
    multiples of 128, so instances finish inside a graph block)."""
    rng = np.random.default_rng(seed)
    lib = configs.material_library(seed)
    out = []
    for i in range(n):
        dims = (rng.uniform(2.5, 6.0), rng.uniform(2.5, 5.0), rng.uniform(2.4, 3.2))
        cfg = configs.mix_room(rng, lib, dx=0.1, dims=dims, n_rx=6)
        cfg["steps"] = int(rng.integers(steps[0], steps[1]))
        case = workload.build_case(cfg)
        out.append(case)
    return out


@pytest.fixture(scope="module")
def rooms():
    cases = small_rooms(5)
    plain = {}
    for prec in ("fp32", "fp64"):
        for i, c in enumerate(cases):
            with fdtd.Plan(c.grid, prec) as p:
                plain[(prec, i)] = p.run(c.runs[0][2], c.signal, c.receivers, mode=fdtd.MODE_IR)[0]
    return cases, plain


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_equals_plain_runs(rooms, graphs):
    cases, plain = rooms
    print(f"case: batch of {len(cases)} rooms, fp32, graphs={graphs}", flush=True)
    with fdtd.Batch("fp32") as b:
        ids = [b.add(c.grid, c.runs[0][2], c.signal, c.receivers) for c in cases]
        assert b.run(graphs=graphs) == fdtd.WV_OK
        t = b.timing()
        for i, iid in enumerate(ids):
            ir, _ = b.result(iid)
            assert ir.shape == plain[("fp32", i)].shape
            assert np.array_equal(ir, plain[("fp32", i)]), f"instance {i}"
    if graphs:
        assert t["graph_launches"] > 0 and t["graph_captures"] > 0
    else:
        assert t["graph_launches"] == 0


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("graphs", [True, False])
def test_persistent_launch_equals_plain_runs(rooms, prec, graphs):
    """One persistent launch per 128-step block: each of the 8 XCDs runs its
    share of the instances, an instance's step waiting only on its own
    previous step. All speaker runs of the five rooms (every XCD gets a
    share), step counts that end inside a block (the table and the shares
    rebuilt as instances finish): each instance bit-identical to the same
    instance in a launch per step, and the first run of each room to its
    plain run."""
    cases, plain = rooms
    inst = [(ci, c, taps) for ci, c in enumerate(cases) for (_, _, taps) in c.runs]
    assert len(inst) >= 8
    print(f"case: persistent batch of {len(inst)} instances, {prec}, graphs={graphs}", flush=True)
    out = {}
    for form in (fdtd.BATCH_STEPS, fdtd.BATCH_PERSISTENT):
        with fdtd.Batch(prec) as b:
            ids = [b.add(c.grid, taps, c.signal, c.receivers) for _, c, taps in inst]
            b.set_launch(form)
            assert b.run(graphs=graphs) == fdtd.WV_OK
            assert b.timing()["launch"] == form
            out[form] = [b.result(i)[0] for i in ids]
    first = {}
    for k, (ci, _, _) in enumerate(inst):
        first.setdefault(ci, k)
        assert np.array_equal(out[fdtd.BATCH_PERSISTENT][k], out[fdtd.BATCH_STEPS][k]), f"instance {k}"
    for ci, k in first.items():
        assert np.array_equal(out[fdtd.BATCH_PERSISTENT][k], plain[(prec, ci)]), f"room {ci}"


def test_persistent_launch_after_autotune(rooms):
    """The autotuner times both launch forms; the form it keeps, and the
    persistent form set afterwards, keep every result."""
    cases, plain = rooms
    with fdtd.Batch("fp32") as b:
        ids = [b.add(c.grid, c.runs[0][2], c.signal, c.receivers) for c in cases]
        b.autotune(trial_steps=4)
        tuned = b.timing()["launch"]
        assert tuned in (fdtd.BATCH_STEPS, fdtd.BATCH_PERSISTENT)
        b.set_launch(fdtd.BATCH_PERSISTENT)
        assert b.run() == fdtd.WV_OK
        for i, iid in enumerate(ids):
            assert np.array_equal(b.result(iid)[0], plain[("fp32", i)]), f"instance {i}"
        with pytest.raises(fdtd.FdtdError):
            b.set_launch(7)


def test_batch_autotuned_shapes_keep_results(rooms):
    cases, plain = rooms
    print("case: batch autotune, fp32", flush=True)
    with fdtd.Batch("fp32") as b:
        ids = [b.add(c.grid, c.runs[0][2], c.signal, c.receivers) for c in cases]
        b.autotune(trial_steps=4)
        zchunks = {b.info(i)["zchunk"] for i in ids}
        assert zchunks <= {2, 4, 8}
        b.run()
        for i, iid in enumerate(ids):
            assert np.array_equal(b.result(iid)[0], plain[("fp32", i)]), f"instance {i}"


def test_batch_fp64_matches_plain_and_oracle(rooms):
    cases, plain = rooms
    print("case: batch fp64 vs plain fp64 and the oracle", flush=True)
    with fdtd.Batch("fp64") as b:
        ids = [b.add(c.grid, c.runs[0][2], c.signal, c.receivers) for c in cases[:3]]
        b.set_tile(variant=1, zchunk=2)
        b.run()
        irs = [b.result(i)[0] for i in ids]
    for i in range(3):
        assert np.array_equal(irs[i], plain[("fp64", i)])
    c = cases[0]
    og = oracle.build_box(c.cfg["room"], c.cfg["c"], c.cfg["dx"])
    ref = oracle.run(og, c.runs[0][2], c.signal, c.receivers, threads=8)[0]
    err = np.linalg.norm(irs[0] - ref) / np.linalg.norm(ref)
    assert err < 1e-9


def test_batch_spectra_equal_plain(rooms):
    cases, _ = rooms
    freqs = [40.0, 80.0, 160.0]
    c = cases[1]
    with fdtd.Plan(c.grid, "fp32") as p:
        _, ref = p.run(c.runs[0][2], c.signal, c.receivers, freqs, mode=fdtd.MODE_SPECTRA)
    with fdtd.Batch("fp32") as b:
        iid = b.add(c.grid, c.runs[0][2], c.signal, c.receivers)
        b.run()
        _, sp = b.result(iid, freqs, mode=fdtd.MODE_SPECTRA)
    assert np.array_equal(sp, ref)


def test_server_results_same_alone_and_in_mix(rooms):
    """Requests submitted while others run join the running batch; each
    result equals the request run alone."""
    cases, plain = rooms
    print("case: server, 5 requests in two waves, fp32", flush=True)
    with fdtd.Server("fp32") as sv:
        first = [sv.submit(c.grid, c.runs[0][2], c.signal, c.receivers) for c in cases[:2]]
        time.sleep(0.01)
        second = [sv.submit(c.grid, c.runs[0][2], c.signal, c.receivers) for c in cases[2:]]
        out = [sv.wait(r) for r in first + second]
        st = sv.stats()
    for i, (ir, _, lat) in enumerate(out):
        assert np.array_equal(ir, plain[("fp32", i)]), f"request {i}"
        assert lat > 0
    assert st["completed"] == 5 and st["failed"] == 0
    assert st["max_running"] >= 2
    assert 0 < st["p50_s"] <= st["p95_s"] <= st["p99_s"]


@pytest.mark.parametrize("prec", ["fp64", "fp32_state64"])
def test_server_and_batch_in_other_precisions(rooms, prec):
    """The server (lanes policy) and the batch in FP64 and in FP32 with FP64
    boundary state: every result equals the plain run of the same precision
    bit for bit (the server mixes the requests as they arrive; the batch's
    small table takes one-tile blocks)."""
    cases, plain = rooms
    ref = []
    for i, c in enumerate(cases):
        if (prec, i) in plain:
            ref.append(plain[(prec, i)])
        else:
            with fdtd.Plan(c.grid, prec) as p:
                ref.append(p.run(c.runs[0][2], c.signal, c.receivers, mode=fdtd.MODE_IR)[0])
    print(f"case: server, {len(cases)} requests, {prec}", flush=True)
    with fdtd.Server(prec) as sv:
        ids = [sv.submit(c.grid, c.runs[0][2], c.signal, c.receivers) for c in cases]
        out = [sv.wait(r) for r in ids]
        st = sv.stats()
    for i, (ir, _, _) in enumerate(out):
        assert np.array_equal(ir, ref[i]), f"request {i}"
    assert st["completed"] == len(cases) and st["failed"] == 0
    print(f"case: batch, {len(cases)} rooms, {prec}", flush=True)
    with fdtd.Batch(prec) as b:
        bids = [b.add(c.grid, c.runs[0][2], c.signal, c.receivers) for c in cases]
        assert b.run() == fdtd.WV_OK
        for i, iid in enumerate(bids):
            assert np.array_equal(b.result(iid)[0], ref[i]), f"instance {i}"


def test_batch_cancel_and_resume(rooms):
    """A batch cancelled from its progress callback returns WV_CANCELLED
    between step blocks, as the reference's run does (gpu_fdtd.rs:202-205);
    run again, it resumes where it stopped, its progress rises to 1, and
    every instance still equals its plain run."""
    cases, plain = rooms
    flag = np.zeros(1, dtype=np.int32)
    seen = []

    def stop_at_first(f):
        seen.append(f)
        flag[0] = 1

    print(f"case: batch of {len(cases)} rooms cancelled at its first progress report, then resumed", flush=True)
    with fdtd.Batch("fp32") as b:
        ids = [b.add(c.grid, c.runs[0][2], c.signal, c.receivers) for c in cases]
        assert b.run(cancel=flag, progress=stop_at_first) == fdtd.WV_CANCELLED
        assert len(seen) == 1 and 0.0 < seen[0] < 1.0
        flag[0] = 0
        again = []
        assert b.run(cancel=flag, progress=again.append) == fdtd.WV_OK
        assert again and again[-1] == 1.0 and all(x <= y for x, y in zip(again, again[1:]))
        for i, iid in enumerate(ids):
            assert np.array_equal(b.result(iid)[0], plain[("fp32", i)]), f"instance {i}"


def test_server_admission_cap_serializes(rooms):
    """A cap that holds one request at a time: requests run one after another
    and still give the same results."""
    cases, plain = rooms
    need = max(fdtd.footprint_estimate(c.grid, "fp32", len(c.signal), len(c.receivers), len(c.runs[0][2]))
               for c in cases[:3])
    print(f"case: server with cap {need} bytes", flush=True)
    with fdtd.Server("fp32", hbm_cap_bytes=need + 1) as sv:
        ids = [sv.submit(c.grid, c.runs[0][2], c.signal, c.receivers) for c in cases[:3]]
        out = [sv.wait(r) for r in ids]
        st = sv.stats()
    for i, (ir, _, _) in enumerate(out):
        assert np.array_equal(ir, plain[("fp32", i)])
    assert st["max_running"] == 1


def test_server_rejects_oversized_request(rooms):
    cases, _ = rooms
    c = cases[0]
    with fdtd.Server("fp32", hbm_cap_bytes=1 << 20) as sv:
        rid = sv.submit(c.grid, c.runs[0][2], c.signal, c.receivers)
        with pytest.raises(fdtd.FdtdError):
            sv.wait(rid)


def large_room(steps=601, seed=11):
    """A mix-recipe room on a grid of more than 4 Mi cells, so its plain run
    takes two steps per pass; an odd step count that is not a multiple of 128."""
    rng = np.random.default_rng(seed)
    lib = configs.material_library(seed)
    cfg = configs.mix_room(rng, lib, dx=0.05, dims=(14.0, 10.0, 4.0), n_rx=6)
    cfg["steps"] = steps
    return workload.build_case(cfg)


@pytest.fixture(scope="module")
def big():
    c = large_room()
    assert c.grid.cells >= 4 << 20
    with fdtd.Plan(c.grid, "fp32") as p:
        ir = p.run(c.runs[0][2], c.signal, c.receivers, mode=fdtd.MODE_IR)[0]
        assert p.pass_info()["steps_per_pass"] >= 2
    return c, ir


@pytest.mark.parametrize("policy,fill", [("all", 0), ("shortest", 0), ("shortest", 1), ("lanes", 0), ("lanes", 1)])
def test_server_policies_keep_results(rooms, policy, fill):
    """Whichever requests each block advances (all, or least work first up
    to a fill; fill 1 runs one request per block), every result equals its
    plain run."""
    cases, plain = rooms
    print(f"case: server policy {policy}, fill {fill}, 5 requests, fp32", flush=True)
    with fdtd.Server("fp32", policy=policy, fill_cells=fill) as sv:
        ids = [sv.submit(c.grid, c.runs[0][2], c.signal, c.receivers) for c in cases]
        out = [sv.wait(r) for r in ids]
        st = sv.stats()
    for i, (ir, _, _) in enumerate(out):
        assert np.array_equal(ir, plain[("fp32", i)]), f"request {i}"
    assert st["completed"] == 5 and st["failed"] == 0
    assert st["solo_blocks"] == 0  # no grid here has two-step passes


@pytest.mark.parametrize("policy", ["all", "shortest", "lanes"])
def test_server_large_request_alone_runs_pairs(rooms, big, policy):
    """A large request (601 steps) and a small one (274 steps): under "all"
    and "shortest" they share blocks, then the large one runs alone on the
    plain run's two-step kernels; under "lanes" its two-step plan takes the
    large-request lane (its own stream, the same kernels) beside the small
    one. Both results equal their plain runs."""
    cases, plain = rooms
    c, ref = big
    print(f"case: server policy {policy}, a {c.grid.cells}-cell request and a small one, fp32", flush=True)
    with fdtd.Server("fp32", policy=policy) as sv:
        ids = [sv.submit(c.grid, c.runs[0][2], c.signal, c.receivers),
               sv.submit(cases[3].grid, cases[3].runs[0][2], cases[3].signal, cases[3].receivers)]
        (ir_big, _, _), (ir_small, _, _) = [sv.wait(r) for r in ids]
        st = sv.stats()
    assert np.array_equal(ir_big, ref)
    assert np.array_equal(ir_small, plain[("fp32", 3)])
    if policy == "lanes":
        assert st["lane_blocks"] > 0
    else:
        assert st["solo_blocks"] > 0


def test_batch_last_large_instance_runs_pairs(rooms, big):
    """In a batch, the large instance outlives the small ones and finishes
    alone on the two-step kernels, bit-identical to its plain run."""
    cases, plain = rooms
    c, ref = big
    print(f"case: batch of a {c.grid.cells}-cell instance and two small ones, fp32", flush=True)
    with fdtd.Batch("fp32") as b:
        ids = [b.add(c.grid, c.runs[0][2], c.signal, c.receivers)]
        ids += [b.add(x.grid, x.runs[0][2], x.signal[:150], x.receivers) for x in cases[:2]]
        b.run()
        t = b.timing()
        assert np.array_equal(b.result(ids[0])[0], ref)
        for i in range(2):
            assert np.array_equal(b.result(ids[1 + i])[0], plain[("fp32", i)][:, :150])
    assert t["solo_blocks"] > 0


def test_server_large_lane_beside_the_batch(rooms, big):
    """Policy "lanes" with a 1 Mi-cell lane threshold: two large requests take
    the second stream one after the other while the small ones run as a batch
    beside them; every result equals its plain run."""
    cases, plain = rooms
    c, ref = big
    print("case: server policy lanes, lane_cells 1Mi, two large and three small requests, fp32", flush=True)
    with fdtd.Server("fp32", policy="lanes", lane_cells=1 << 20) as sv:
        ids = [sv.submit(c.grid, c.runs[0][2], c.signal, c.receivers),
               sv.submit(c.grid, c.runs[0][2], c.signal[:400], c.receivers)]
        ids += [sv.submit(x.grid, x.runs[0][2], x.signal, x.receivers) for x in cases[:3]]
        out = [sv.wait(r)[0] for r in ids]
        st = sv.stats()
    assert np.array_equal(out[0], ref)
    assert np.array_equal(out[1], ref[:, :400])
    for i in range(3):
        assert np.array_equal(out[2 + i], plain[("fp32", i)]), f"small request {i}"
    assert st["completed"] == 5 and st["failed"] == 0
    assert st["lane_blocks"] >= 8  # 601 and 400 steps in 128-step blocks


def test_footprint_estimate_bounds_device_bytes(rooms, big):
    """The admission estimate is at least what an instance holds on the
    device, in each pass mode: small rooms (one step per pass), a 4.5M-cell
    room (two) and the hall (three, five field buffers)."""
    cases, _ = rooms
    c_big, _ = big
    hall = workload.build_case(configs.hall(), steps=12)
    print("case: footprint estimates of small rooms, a 4.5M-cell room and the hall, fp32", flush=True)
    with fdtd.Batch("fp32") as b:
        for c in [cases[0], cases[1], c_big, hall]:
            iid = b.add(c.grid, c.runs[0][2], c.signal, c.receivers)
            held = b.info(iid)["device_bytes"]
            est = fdtd.footprint_estimate(c.grid, "fp32", len(c.signal), len(c.receivers), len(c.runs[0][2]))
            assert held <= est, (c.grid.n, held, est)
            assert est < 1.5 * held + (64 << 20), (c.grid.n, held, est)


def test_hall_takes_the_large_lane_and_runs_alone_in_a_batch(rooms):
    """A request on the full hall grid (three steps per pass in its plain
    run) takes the server's large lane, and in a batch it finishes alone on
    the three-step kernels; both give its plain run's series."""
    cases, plain = rooms
    hall = workload.build_case(configs.hall(), steps=300)
    print("case: hall request (300 steps) in the server and in a batch beside a small room, fp32", flush=True)
    with fdtd.Plan(hall.grid, "fp32") as p:
        ref = p.run(hall.runs[0][2], hall.signal, hall.receivers, mode=fdtd.MODE_IR)[0]
        assert p.pass_info()["steps_per_pass"] == 3
    with fdtd.Server("fp32") as sv:
        ids = [sv.submit(hall.grid, hall.runs[0][2], hall.signal, hall.receivers),
               sv.submit(cases[3].grid, cases[3].runs[0][2], cases[3].signal, cases[3].receivers)]
        out = [sv.wait(r)[0] for r in ids]
        st = sv.stats()
    assert np.array_equal(out[0], ref)
    assert np.array_equal(out[1], plain[("fp32", 3)])
    assert st["lane_blocks"] >= 3
    with fdtd.Batch("fp32") as b:
        ids = [b.add(hall.grid, hall.runs[0][2], hall.signal, hall.receivers),
               b.add(cases[0].grid, cases[0].runs[0][2], cases[0].signal[:150], cases[0].receivers)]
        b.run()
        t = b.timing()
        assert np.array_equal(b.result(ids[0])[0], ref)
        assert np.array_equal(b.result(ids[1])[0], plain[("fp32", 0)][:, :150])
    assert t["solo_blocks"] >= 1
