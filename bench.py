"""Benchmark of the wave-solver step on one MI355X.

A step is one time step of the FDTD run over the whole grid. FP32 grids of
8 Mi cells and more (the hall; FP64 from 7 Mi) take three steps per pass:
k_tri1 (cells three or more from a wall three times in one z-march, special
cells' first step, receivers) and two launches of k_pair2 (special cells and the cells near the
walls, second and third steps, receivers), so per-step figures are a third
of a pass; grids of 2.25 Mi cells and more (FP64: 0.75 Mi) take two steps per
pass (k_pair1 and k_pair2), smaller grids one launch of k_step per step.
Workload (default): the 40 x 25 x 15 m hall at dx = 0.04 m (1002 x 627 x 377
cells, 232.5M air cells, 8 pillars, 10 three-section admittances on surface
patches, 256 receivers), FP32 fields, one cardioid source component. Inputs
are built and resident in HBM before the timed region; W warm-up steps run
first, then exactly K steps are timed (the call returns after a stream
synchronize on both sides).

Prints one JSON line with the metric, the roofline of the step kernel
(algorithmic bytes per launch / measured launch time, HIP events on the
kernel's stream; HBM traffic per launch from two rocprofv3 counter passes of
a short run, when rocprofv3 is installed), a CPU baseline (the FP64 oracle,
threaded, on a bounded sample of the same grid), the FP32 receiver error
against an FP64 run of the same inputs, and the second half of the metric:
the batched form on the 48-room uneven mix against sequential plain runs
(`room_mix`) and the served Poisson trace, loaded at 80 % of the mix's
measured throughput, with its p50/p95/p99 latency against a sequential
plain queue (`serving`).
"""
import argparse
import glob
import json
import os
import shutil
import sqlite3
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(st, n_recv, n_taps, fp64=False):
    """Bytes per step: p read once, prev read and written (12 B per FP32 air
    cell), 4 B boundary index per node, per slot 12 B constants + u read and
    written when no slot has a spring (k = 0; else 16 B constants + (u, g)
    read and written = 32 B), 4 B per receiver and 8 B per tap."""
    f = 2 if fp64 else 1
    per_slot = 32 if st["slots_with_k"] else 20
    return (12 * f * st["air_cells"] + 4 * st["boundary_nodes"] + per_slot * f * st["slots"]
            + 4 * f * n_recv + 8 * n_taps)


def host_cpu():
    """CPU model, logical CPUs of the host and of this process's affinity."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return model, os.cpu_count() or 1, affinity


def cpu_threads():
    """Every CPU of this process's affinity, within the host's CPU share for
    one GPU (OMP_NUM_THREADS where it is set, else 16)."""
    _, _, affinity = host_cpu()
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(affinity, share))


def cpu_baseline(cfg, case, n1=20, n2=170, threads=None):
    """The FP64 oracle (restatement of fdtd.rs:644-725, z-slab threaded) on the
    same grid, taps and receivers, from rest, for n2 steps and for n1: both
    pay the same set-up (field allocation and first touch), so the step rate
    is their difference quotient over n2 - n1 steps. The signal is the
    workload's source pulse at that length, whatever the GPU run's length."""
    import oracle
    from wavey_amd import fdtd

    threads = threads or cpu_threads()
    og = oracle.build_box(cfg["room"], cfg["c"], cfg["dx"])
    taps = case.runs[0][2]
    rx = case.receivers
    sig = fdtd.source_signal(cfg["f_src"], og.dt, n2)

    def timed(n):
        t0 = time.perf_counter()
        oracle.run(og, taps, sig[:n], rx, threads=threads)
        return time.perf_counter() - t0

    t1 = timed(n1)
    t2 = timed(n2)
    dt = max(t2 - t1, 1e-9)
    air = int(np.count_nonzero(og.flags & 64))
    model, nproc, affinity = host_cpu()
    return {"value": round(air * (n2 - n1) / dt / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{cfg['name']} grid ({air} air cells), FP64 oracle, {threads} threads: runs of {n2} and "
                      f"{n1} steps from rest, {n2 - n1} steps in {dt:.1f} s (difference of the two runs), "
                      f"extrapolated per step",
            "cpu_model": model, "nproc": nproc, "affinity_cpus": affinity}


def _counter_kib(db_dir, counter, kernel):
    """Mean value per dispatch of one counter for kernels whose (demangled)
    name contains `kernel`, from a rocprofv3 SQLite output."""
    dbs = glob.glob(os.path.join(db_dir, "**", "*.db"), recursive=True)
    if not dbs:
        raise RuntimeError(f"no rocprofv3 database under {db_dir}")
    vals = []
    for db in dbs:
        con = sqlite3.connect(db)
        vals += [r[0] for r in con.execute(
            "select value from counters_collection where counter_name = ? and kernel_name like ?",
            (counter, "%" + kernel + "%"))]
        con.close()
    if not vals:
        raise RuntimeError(f"{counter} not recorded for {kernel}")
    return sum(vals) / len(vals), len(vals)


STEP_KERNELS = ["k_tri1", "k_pair1", "k_pair2", "k_step"]
TRAFFIC_STEPS = 12  # a whole number of passes of one, two or three steps


def measure_traffic(args, copy_bytes):
    """HBM bytes per step from two rocprofv3 counter passes (FETCH_SIZE, then
    WRITE_SIZE) over a 12-step run of this workload: the counter summed over
    every dispatch of the step kernels, divided by the 12 steps. Each counter
    is scaled by what it reads for the plain copy kernel of the same run,
    which moves copy_bytes in each direction."""
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    out = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix=f"wv_{ctr.lower()}_", dir="/tmp")
        cmd = [exe, "--pmc", ctr, "-d", d, "-o", "run", "--", sys.executable, os.path.abspath(__file__),
               "--no-cpu", "--no-traffic", "--no-mix", "--precision-steps", "0", "--steps", str(TRAFFIC_STEPS), "--warmup", "0", "--workload", args.workload,
               "--precision", args.precision, "--zchunk", str(args.zchunk), "--variant", str(args.variant),
               "--steps-per-pass", str(args.steps_per_pass), "--pair-zchunk", str(args.pair_zchunk)]
        print(f"bench: counter pass {ctr}", file=sys.stderr, flush=True)
        env = dict(os.environ, TMPDIR="/tmp")
        try:
            r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                               timeout=180)
            if r.returncode != 0:
                return None, f"rocprofv3 {ctr} pass exited with {r.returncode}"
            step_kib, n = 0.0, 0
            for k in STEP_KERNELS:
                try:
                    kib, cnt = _counter_kib(d, ctr, k)
                except RuntimeError:
                    continue  # not launched in this mode
                step_kib += kib * cnt
                n += cnt
            if n == 0:
                raise RuntimeError(f"{ctr} not recorded for any step kernel")
            step_kib /= TRAFFIC_STEPS
            copy_kib, _ = _counter_kib(d, ctr, "k_copy")
        except (subprocess.TimeoutExpired, RuntimeError, sqlite3.Error) as e:
            return None, f"{ctr} pass: {e}"
        finally:
            shutil.rmtree(d, ignore_errors=True)
        scale = copy_bytes / (copy_kib * 1024.0)
        out[ctr] = {"kib_per_step": round(step_kib, 1), "dispatches": n, "scale": round(scale, 4),
                    "bytes": int(step_kib * 1024.0 * scale)}
    return out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"], out


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def _instances(cases):
    """One instance per (room, source run): the per-speaker solves."""
    return [(ci, c, taps) for ci, c in enumerate(cases) for (_, _, taps) in c.runs]


def bench_room_mix(args):
    """The 48-room uneven mix: sequential plain runs (one plan per room, its
    speakers one after another) against one batch (all instances in one
    launch per step, autotuned shapes, 128-step blocks as HIP graphs). Every
    batched IR is compared with its plain IR (bit-identical expected); a
    sample of instances is also run in FP64 for the FP32 error. With
    args.oneshot also timed: one wv_fdtd_run call per solve (the device plan
    built and freed every call, as the reference calls Gpu::run_fdtd once
    per component run). Returns the result line."""
    from wavey_amd import configs, fdtd, workload

    rooms = configs.room_mix(n_rooms=args.rooms)
    t0 = time.perf_counter()
    cases = [workload.build_case(cfg, steps=args.mix_steps or None) for cfg in rooms]
    t_build = time.perf_counter() - t0
    inst = _instances(cases)
    work = sum(c.grid.air_cells() * len(c.signal) for _, c, _ in inst)
    print(f"bench: room mix {len(cases)} rooms, {len(inst)} instances, {work / 1e12:.2f} T cell-updates, "
          f"built in {t_build:.1f} s", file=sys.stderr, flush=True)

    # A plain copy of known size in the same process (the calibration of
    # HBM counter profiles of short runs, --mix-steps).
    copy = None
    if args.mix_steps:
        big = max(cases, key=lambda c: c.grid.air_cells())
        with fdtd.Plan(big.grid, args.precision) as p:
            nbytes = (big.grid.air_cells() * 4) // 4096 * 4096
            copy = {"bytes": int(nbytes), "ms": round(p.copy_probe(nbytes, reps=5), 4)}

    # Sequential plain runs.
    plain = []
    t0 = time.perf_counter()
    dev_ms = 0.0
    for ci, c in enumerate(cases):
        with fdtd.Plan(c.grid, args.precision) as p:
            for (_, _, taps) in c.runs:
                ir, _ = p.run(taps, c.signal, c.receivers, mode=fdtd.MODE_IR)
                dev_ms += p.last_timing()[0]
                plain.append(ir)
    t_plain = time.perf_counter() - t0
    print(f"bench: sequential plain {t_plain:.2f} s", file=sys.stderr, flush=True)

    # One wv_fdtd_run call per solve.
    t_oneshot, oneshot_same = None, 0
    if args.oneshot:
        t0 = time.perf_counter()
        for k, (_, c, taps) in enumerate(inst):
            ir, _ = fdtd.fdtd_run(c.grid, taps, c.signal, c.receivers, mode=fdtd.MODE_IR, precision=args.precision)
            oneshot_same += int(np.array_equal(ir, plain[k]))
        t_oneshot = time.perf_counter() - t0
        print(f"bench: sequential wv_fdtd_run calls {t_oneshot:.2f} s", file=sys.stderr, flush=True)

    # One batch.
    t0 = time.perf_counter()
    b = fdtd.Batch(args.precision)
    ids = b.add_many([(c.grid, taps, c.signal, c.receivers) for _, c, taps in inst], threads=8)
    t_add = time.perf_counter() - t0
    t0 = time.perf_counter()
    if not args.no_autotune:
        b.autotune(trial_steps=8)
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.run(graphs=not args.no_graphs)
    t_run = time.perf_counter() - t0
    bt = b.timing()
    t0 = time.perf_counter()
    batched = [b.result(i)[0] for i in ids]
    t_read = time.perf_counter() - t0
    zch = sorted({b.info(i)["zchunk"] for i in ids})
    b.close()
    t_batch = t_add + t_tune + t_run + t_read
    identical = sum(int(np.array_equal(x, y)) for x, y in zip(batched, plain))
    max_diff = max(float(np.max(np.abs(x - y))) for x, y in zip(batched, plain))

    # The other launch form, with the tuned shapes: one persistent launch per
    # 128-step block (or a launch per step, if the autotuner kept that one).
    other = fdtd.BATCH_STEPS if bt["launch"] == fdtd.BATCH_PERSISTENT else fdtd.BATCH_PERSISTENT
    with fdtd.Batch(args.precision) as b2:
        ids2 = b2.add_many([(c.grid, taps, c.signal, c.receivers) for _, c, taps in inst], threads=8)
        b2.set_launch(other)
        t0 = time.perf_counter()
        b2.run(graphs=not args.no_graphs)
        t_other = time.perf_counter() - t0
        same_other = sum(int(np.array_equal(b2.result(i)[0], plain[k])) for k, i in enumerate(ids2))
    forms = {fdtd.BATCH_STEPS: "launch_per_step", fdtd.BATCH_PERSISTENT: "persistent"}
    print(f"bench: batch run {t_run:.2f} s ({forms[bt['launch']]}), {forms[other]} {t_other:.2f} s",
          file=sys.stderr, flush=True)

    # FP32 against an FP64 run of the same scheme, on a sample of instances.
    errs = []
    if args.precision != "fp64":
        for k in range(0, len(inst), max(1, len(inst) // args.fp64_sample)):
            _, c, taps = inst[k]
            with fdtd.Plan(c.grid, "fp64") as p:
                ref, _ = p.run(taps, c.signal, c.receivers, mode=fdtd.MODE_IR)
            errs.append(rel_l2(batched[k], ref))
    n = len(inst)
    line = {
        "metric": "batched solves/s (uneven room mix)",
        "value": round(n / t_batch, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": {"fp32": "f32", "fp64": "f64", "fp32_state64": "f32+f64state"}[args.precision],
        "data": "synthetic (seeded rooms, materials, speakers, receivers)",
        "config": {"workload": f"room mix: {len(cases)} rooms, dx 0.05 m, {n} per-speaker instances, "
                               f"{len(cases[0].signal)} steps{' (0.5 s)' if not args.mix_steps else ''}, "
                               f"16 receivers each",
                   "instances": n, "cell_updates": int(work), "precision": args.precision,
                   "graphs": not args.no_graphs, "autotune": not args.no_autotune},
        "sequential_plain": {"solves_per_s": round(n / t_plain, 3), "seconds": round(t_plain, 3),
                             "device_step_loop_s": round(dev_ms / 1e3, 3)},
        "sequential_fdtd_run_calls": None if t_oneshot is None else {
            "solves_per_s": round(n / t_oneshot, 3), "seconds": round(t_oneshot, 3),
            "identical_to_plain": f"{oneshot_same}/{n}", "speedup_of_batch": round(t_oneshot / t_batch, 3)},
        "batched": {"seconds": round(t_batch, 3), "add_s": round(t_add, 3), "autotune_s": round(t_tune, 3),
                    "run_s": round(t_run, 3), "readback_s": round(t_read, 3),
                    "device_run_s": round(bt["device_ms"] / 1e3, 3),
                    "run_gcell_per_s": round(work / max(t_run, 1e-9) / 1e9, 1),
                    "graph_launches": bt["graph_launches"], "graph_captures": bt["graph_captures"],
                    "variant": bt["variant"], "zchunks": zch, "launch": forms[bt["launch"]]},
        "other_launch_form": {"launch": forms[other], "run_s": round(t_other, 3),
                              "run_gcell_per_s": round(work / max(t_other, 1e-9) / 1e9, 1),
                              "bit_identical_to_plain": f"{same_other}/{n}"},
        "speedup_vs_sequential_plain": round(t_plain / t_batch, 3),
        "speedup_step_loop_only": round(t_plain / max(t_run, 1e-9), 3),
        "parity": {"bit_identical_to_plain": f"{identical}/{n}", "max_abs_diff": max_diff,
                   "fp32_vs_fp64_rel_l2_max": max(errs) if errs else None, "fp64_sample": len(errs)},
        "end_to_end_gcell_per_s": round(work / t_batch / 1e9, 2),
        "setup_s": {"grid_build": round(t_build, 2)},
    }
    if copy:
        line["copy_probe"] = copy
    return line


def bench_serving(args, throughput=None):
    """Poisson arrivals of room-mix requests with padded cells log-uniform in
    0.2M-60M and 0.3-1.0 s of simulated time, at the rate that offers
    args.load of the given batched throughput (cell-updates/s). The server
    (admission by footprint, continuous batching) against sequential plain
    runs of the same trace: the plain latencies are those of a FIFO single
    queue with the measured per-request times and the same arrivals. With
    args.burst every request arrives at once: solves/s is then the capacity
    of each form (all 40 requests from first submission to last result).
    `throughput` (Gcell-updates/s) overrides args.throughput. Returns the
    result line."""
    import threading

    from wavey_amd import configs, fdtd, workload

    base = configs.serving_trace(1.0, args.requests)
    t0 = time.perf_counter()
    cases = [workload.build_case(cfg) for _, cfg in base]
    t_build = time.perf_counter() - t0
    works = [c.grid.air_cells() * len(c.signal) for c in cases]
    throughput = throughput or args.throughput
    rate = args.load * throughput * 1e9 / float(np.mean(works))
    arrivals = [0.0] * len(base) if args.burst else [t / rate for t, _ in base]
    print(f"bench: serving {len(cases)} requests, {sum(works) / 1e12:.2f} T cell-updates, rate {rate:.2f}/s, "
          f"built in {t_build:.1f} s", file=sys.stderr, flush=True)

    # Sequential plain: each request alone, in arrival order.
    plain, service = [], []
    for c in cases:
        t0 = time.perf_counter()
        with fdtd.Plan(c.grid, args.precision) as p:
            ir, _ = p.run(c.runs[0][2], c.signal, c.receivers, mode=fdtd.MODE_IR)
        service.append(time.perf_counter() - t0)
        plain.append(ir)
    fin, lat_plain = 0.0, []
    for a, sv in zip(arrivals, service):
        fin = max(fin, a) + sv
        lat_plain.append(fin - a)
    span_plain = fin - arrivals[0]
    print(f"bench: sequential plain {sum(service):.2f} s of service", file=sys.stderr, flush=True)

    # The server, fed by a client thread at the arrival times.
    server = fdtd.Server(args.precision, hbm_cap_bytes=int(args.cap_gb * 1e9), policy=args.policy,
                         fill_cells=args.fill, lane_cells=args.lane_cells)
    ids = [None] * len(cases)
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
    served = []
    lat = []
    for k in range(len(cases)):
        ir, _, l = server.wait(ids[k])
        served.append(ir)
        lat.append(l)
    st = server.stats()
    server.close()
    span = st["span_s"]
    for k, c in enumerate(cases):
        print(f"bench: request {k}: {c.grid.air_cells() / 1e6:.2f}M air cells, {len(c.signal)} steps, "
              f"arrival {arrivals[k]:.3f} s, service alone {service[k]:.3f} s, latency {lat[k]:.3f} s "
              f"(sequential plain {lat_plain[k]:.3f} s)", file=sys.stderr, flush=True)
    identical = sum(int(np.array_equal(x, y)) for x, y in zip(served, plain))
    errs = []
    if args.precision != "fp64":
        for k in range(0, len(cases), max(1, len(cases) // args.fp64_sample)):
            c = cases[k]
            with fdtd.Plan(c.grid, "fp64") as p:
                ref, _ = p.run(c.runs[0][2], c.signal, c.receivers, mode=fdtd.MODE_IR)
            errs.append(rel_l2(served[k], ref))

    def pct(v):
        return {"p50": round(float(np.percentile(v, 50)), 4), "p95": round(float(np.percentile(v, 95)), 4),
                "p99": round(float(np.percentile(v, 99)), 4)}

    n = len(cases)
    line = {
        "metric": "served solves/s and latency (%s)" % ("all requests at once" if args.burst else "Poisson trace"),
        "value": round(n / span, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": {"fp32": "f32", "fp64": "f64", "fp32_state64": "f32+f64state"}[args.precision],
        "data": "synthetic (seeded rooms, sizes, durations, arrivals)",
        "config": {"workload": f"serving trace: {n} requests, padded cells log-uniform 0.2M-60M, 0.3-1.0 s, "
                               + ("dx 0.05 m, all submitted at once" if args.burst else
                                  f"dx 0.05 m, Poisson rate {rate:.3f}/s = {args.load:.0%} of "
                                  f"{throughput:.1f} Gcell-updates/s"),
                   "policy": args.policy, "fill_cells": args.fill or 16 << 20,
                   "lane_cells": args.lane_cells or "multi-step plans from 2.25 Mi cells",
                   "requests": n, "cell_updates": int(sum(works)), "cap_bytes": int(args.cap_gb * 1e9)},
        "latency_s": pct(lat),
        "sequential_plain": {"solves_per_s": round(n / span_plain, 3), "latency_s": pct(lat_plain),
                             "service_s_total": round(sum(service), 3)},
        "speedup_vs_sequential_plain": round(span_plain / span, 3),
        "server": {"span_s": round(span, 3), "max_running": int(st["max_running"]),
                   "peak_device_bytes": int(st["peak_device_bytes"]), "graph_captures": int(st["graph_captures"]),
                   "graph_launches": int(st["graph_launches"]), "blocks_alone": int(st["solo_blocks"]),
                   "large_lane_blocks": int(st["lane_blocks"]),
                   "failed": int(st["failed"])},
        "parity": {"bit_identical_to_plain": f"{identical}/{n}",
                   "fp32_vs_fp64_rel_l2_max": max(errs) if errs else None, "fp64_sample": len(errs)},
        "setup_s": {"grid_build": round(t_build, 2)},
    }
    return line


def bench_error(args, cfg):
    """FP32 (and FP32 with FP64 slot state) against an FP64 run of the same
    scheme and inputs, on the device, over args.error_steps steps of the
    workload's first source component: relative L2 of the receiver series,
    per receiver (max) and over all."""
    from wavey_amd import fdtd, workload

    case = workload.build_case(cfg, steps=args.error_steps)
    taps = case.runs[0][2]
    out = {}
    irs = {}
    for prec in ("fp64", "fp32", "fp32_state64"):
        t0 = time.perf_counter()
        with fdtd.Plan(case.grid, prec) as plan:
            irs[prec], _ = plan.run(taps, case.signal, case.receivers, mode=fdtd.MODE_IR)
        out[prec + "_s"] = round(time.perf_counter() - t0, 2)
        print(f"bench: {prec} run {out[prec + '_s']} s", file=sys.stderr, flush=True)
    ref = irs["fp64"]
    for prec in ("fp32", "fp32_state64"):
        per = [rel_l2(irs[prec][r], ref[r]) for r in range(len(ref)) if np.linalg.norm(ref[r]) > 0]
        out[prec] = {"rel_l2_all": rel_l2(irs[prec], ref), "rel_l2_max_receiver": max(per),
                     "rel_l2_median_receiver": float(np.median(per))}
    line = {"metric": "FP32 receiver error against FP64 (same scheme, same inputs)",
            "config": {"workload": f"{cfg['name']} {'x'.join(str(v) for v in case.grid.n)} cells, "
                                   f"{args.error_steps} steps, {len(case.receivers)} receivers, one source component"},
            "tolerance": 1e-4, **out}
    return line


def precision_check(case, taps, ir32, steps, precision):
    """The timed run's receiver series against an FP64 device run of the same
    steps, and a longer FP32/FP64 pair over `steps` steps of the same grid,
    taps and receivers (the timed run is too short for the pulse to reach
    most receivers). Relative L2 over all receivers and the worst receiver."""
    from wavey_amd import fdtd

    def errs(a, ref):
        # Per receiver, those the pulse has reached: series above 1e-3 of the
        # largest one (the others hold only the stencil's precursor).
        norms = np.linalg.norm(ref, axis=1)
        per = [rel_l2(a[r], ref[r]) for r in range(len(ref)) if norms[r] > 1e-3 * norms.max()]
        return {"rel_l2_all": rel_l2(a, ref), "rel_l2_max_receiver": max(per) if per else None,
                "receivers_reached": len(per)}

    out = {"tolerance": 1e-4, "reference": "FP64 device run of the same scheme and inputs (itself within "
                                           "1e-9 of the FP64 CPU oracle in the parity tests)"}
    sig_long = fdtd.source_signal(case.cfg["f_src"], case.grid.dt, steps)
    irs = {}
    for prec in ("fp64", precision):
        with fdtd.Plan(case.grid, prec) as plan:
            if prec == "fp64":
                short, _ = plan.run(taps, case.signal, case.receivers, mode=fdtd.MODE_IR)
            irs[prec], _ = plan.run(taps, sig_long, case.receivers, mode=fdtd.MODE_IR)
    out["timed_run"] = {"steps": len(case.signal), **errs(ir32, short)}
    out["long_run"] = {"steps": steps, **errs(irs[precision], irs["fp64"])}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="hall",
                    choices=["hall", "pillared_room", "shoebox", "room_mix", "serving"])
    ap.add_argument("--rooms", type=int, default=48, help="room_mix: rooms")
    ap.add_argument("--mix-steps", type=int, default=0,
                    help="room_mix: steps per instance (0 = the workload's 0.5 s; fewer for counter profiles)")
    ap.add_argument("--requests", type=int, default=40, help="serving: requests")
    ap.add_argument("--load", type=float, default=0.8, help="serving: offered fraction of --throughput")
    ap.add_argument("--burst", action="store_true", help="serving: every request arrives at once")
    ap.add_argument("--throughput", type=float, default=200.0,
                    help="serving: batched room-mix throughput to load against, Gcell-updates/s "
                         "(end to end, as the room_mix mode measures it)")
    ap.add_argument("--cap-gb", type=float, default=288.0, help="serving: admission cap, GB")
    ap.add_argument("--policy", choices=["lanes", "shortest", "all"], default="lanes",
                    help="serving: requests each block advances (least work first up to --fill, large "
                         "requests on their own stream; least work first only; or all)")
    ap.add_argument("--lane-cells", type=int, default=0, help="serving: air cells from which a multi-step request "
                                                              "takes the large lane (0 = plans with two or "
                                                              "three steps per pass from 2.25 Mi cells)")
    ap.add_argument("--fill", type=int, default=0, help="serving: air cells per block for --policy shortest "
                                                        "(0 = the library's 16 Mi)")
    ap.add_argument("--fp64-sample", type=int, default=6, help="instances rerun in FP64 for the error")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--error-steps", type=int, default=0,
                    help="instead of timing: FP32 against FP64 over this many steps of the workload")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--grid-builder", default="box", choices=["box", "mesh", "prism"],
                    help="build the room with the box builder, or as triangles (a closed mesh, or an extruded plan)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp32_state64"])
    ap.add_argument("--zchunk", type=int, default=0)
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--steps-per-pass", type=int, default=0, help="1, 2 or 3; 0 = the plan's default for the grid")
    ap.add_argument("--pair-zchunk", type=int, default=0, help="z-chunk of multi-step passes; 0 = the default")
    ap.add_argument("--oneshot", action="store_true", help="room_mix: also time one wv_fdtd_run call per solve")
    ap.add_argument("--no-mix", action="store_true",
                    help="hall line: leave out the room-mix batch and the serving trace")
    ap.add_argument("--precision-steps", type=int, default=1500,
                    help="hall line: steps of the FP32-against-FP64 run (0 = none)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-traffic", action="store_true", help="skip the HBM counter passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU z-slab decomposition is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "room_mix":
        print(json.dumps(bench_room_mix(args)), flush=True)
        return
    if args.workload == "serving":
        print(json.dumps(bench_serving(args)), flush=True)
        return

    from wavey_amd import configs, fdtd, workload

    if args.workload == "hall":
        cfg = configs.hall()
    elif args.workload == "pillared_room":
        cfg = configs.pillared_room()
    else:
        cfg = configs.shoebox(fdtd.fit_real_impedance)
    if args.error_steps:
        print(json.dumps(bench_error(args, cfg)), flush=True)
        return
    total = args.warmup + args.steps
    t0 = time.perf_counter()
    case = workload.build_case(cfg, steps=max(total, 1), builder=args.grid_builder)
    t_build = time.perf_counter() - t0
    print(f"bench: {cfg['name']} grid {case.grid.n}, built ({args.grid_builder}) in {t_build:.1f} s",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    plan = fdtd.Plan(case.grid, args.precision)
    if args.zchunk or args.variant >= 0:
        plan.set_tile(args.zchunk, args.variant)
    if args.steps_per_pass or args.pair_zchunk:
        plan.set_pass(args.steps_per_pass, args.pair_zchunk)
    taps = case.runs[0][2]
    # A plain copy of one field buffer: the attainable-bandwidth reference and
    # the calibration dispatch for HBM traffic counters.
    copy_bytes = case.grid.cells * (8 if args.precision == "fp64" else 4)
    copy_ms = plan.copy_probe(copy_bytes, reps=5)
    plan.begin(taps, case.signal, case.receivers)
    t_setup = time.perf_counter() - t0
    st = plan.stats()
    pinfo = plan.pass_info()
    spp = pinfo["steps_per_pass"]
    if args.warmup:
        plan.advance(args.warmup)
    t0 = time.perf_counter()
    plan.advance(args.steps)
    wall = time.perf_counter() - t0
    dev_ms, steps_timed, _ = plan.last_timing()
    ir, _ = plan.finish((), fdtd.MODE_IR)
    if not np.all(np.isfinite(ir)):
        raise RuntimeError("non-finite receiver samples")
    plan.close()

    air = st["air_cells"]
    value = air * args.steps / wall / 1e9
    step_s = dev_ms * 1e-3 / max(steps_timed, 1)
    nbytes = algorithmic_bytes(st, len(case.receivers), len(taps), args.precision == "fp64")
    achieved = nbytes / step_s / 1e9
    line = {
        "metric": "FDTD Gcell-updates/s",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp64": "f64", "fp32_state64": "f32+f64state"}[args.precision],
        "data": "synthetic (seeded box room, admittances, source, receivers)",
        "config": {"workload": f"{cfg['name']} {'x'.join(str(v) for v in case.grid.n)} cells, dx {cfg['dx']} m, "
                               f"{air} air cells, {st['boundary_nodes']} boundary nodes, {st['slots']} slots, "
                               f"{len(case.receivers)} receivers, one source component",
                   "grid": list(case.grid.n), "air_cells": air, "precision": args.precision,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": {3: "k_tri1 + 2 x k_pair2 (three steps per pass, z-chunk %d)" % pinfo["pair_zchunk"],
                                2: "k_pair1 + k_pair2 (two steps per pass, z-chunk %d)" % pinfo["pair_zchunk"],
                                1: "k_step"}[spp],
                     "per": "step", "bytes_per_step": int(nbytes),
                     "step_us": round(step_s * 1e6, 2),
                     "copy_GBs": round(2 * copy_bytes / (copy_ms * 1e-3) / 1e9, 1), "copy_bytes": 2 * copy_bytes},
        "setup_s": {"grid_build": round(t_build, 2), "grid_builder": args.grid_builder,
                    "upload_and_layout": round(t_setup, 2)},
    }
    if args.precision_steps and args.precision != "fp64":
        t0 = time.perf_counter()
        line["fp32_vs_fp64"] = precision_check(case, taps, ir, args.precision_steps, args.precision)
        print(f"bench: precision check {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    if not args.no_traffic:
        traffic, detail = measure_traffic(args, copy_bytes)
        line["roofline"]["traffic"] = traffic
        line["roofline"]["traffic_detail"] = detail
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cfg, case)
    if not args.no_mix and args.workload == "hall":
        # The batched form on the uneven room mix, then the serving trace loaded
        # at args.load of the batch's measured end-to-end throughput.
        mix = bench_room_mix(args)
        line["room_mix"] = mix
        line["serving"] = bench_serving(args, throughput=mix["end_to_end_gcell_per_s"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
