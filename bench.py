"""Benchmark of the wave-solver step on one MI355X.

A step is one time step of the FDTD run over the whole grid: one launch of
the fused step kernel (interior tiles, boundary/special cells, receivers).
Workload (default): the 40 x 25 x 15 m hall at dx = 0.04 m (1002 x 627 x 377
cells, 232.5M air cells, 8 pillars, 10 three-section admittances on surface
patches, 256 receivers), FP32 fields, one cardioid source component. Inputs
are built and resident in HBM before the timed region; W warm-up steps run
first, then exactly K steps are timed (the call returns after a stream
synchronize on both sides).

Prints one JSON line with the metric, the roofline of the step kernel
(algorithmic bytes per launch / measured launch time, HIP events on the
kernel's stream; HBM traffic per launch from two rocprofv3 counter passes of
a short run, when rocprofv3 is installed) and a CPU baseline (the FP64
oracle, threaded, on a bounded sample of the same grid).
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


def cpu_baseline(cfg, case, budget_s=15.0, threads=None):
    """The FP64 oracle (restatement of fdtd.rs:644-725, z-slab threaded) on the
    same grid, from rest. Runs of n1 and n2 steps pay the same set-up (field
    allocation and first touch), so the step rate is the difference quotient;
    n2 is sized to about budget_s of stepping."""
    import oracle

    threads = threads or min(16, os.cpu_count() or 1)
    og = oracle.build_box(cfg["room"], cfg["c"], cfg["dx"])
    taps = case.runs[0][2]
    rx = case.receivers
    sig = case.signal

    def timed(n):
        t0 = time.perf_counter()
        oracle.run(og, taps, sig[:n], rx, threads=threads)
        return time.perf_counter() - t0

    n1 = 2
    t1 = timed(n1)
    t_probe = max(timed(2 * n1) - t1, 1e-6) / n1
    n2 = int(min(len(sig), max(2 * n1 + 1, n1 + budget_s / t_probe)))
    t2 = timed(n2)
    dt = max(t2 - t1, 1e-9)
    air = int(np.count_nonzero(og.flags & 64))
    return {"value": round(air * (n2 - n1) / dt / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{cfg['name']} grid, FP64 oracle, {threads} threads: runs of {n2} and {n1} steps from rest, "
                      f"{n2 - n1} steps in {dt:.1f} s (difference of the two runs)"}


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


def measure_traffic(args, copy_bytes):
    """HBM bytes per step-kernel launch from two rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE) over a short run of this workload. Each
    counter is scaled by what it reads for the plain copy kernel of the same
    run, which moves copy_bytes in each direction."""
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    out = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix=f"wv_{ctr.lower()}_", dir="/tmp")
        cmd = [exe, "--pmc", ctr, "-d", d, "-o", "run", "--", sys.executable, os.path.abspath(__file__),
               "--no-cpu", "--no-traffic", "--steps", "10", "--warmup", "2", "--workload", args.workload,
               "--precision", args.precision, "--zchunk", str(args.zchunk), "--variant", str(args.variant)]
        print(f"bench: counter pass {ctr}", file=sys.stderr, flush=True)
        env = dict(os.environ, TMPDIR="/tmp")
        try:
            r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                               timeout=180)
            if r.returncode != 0:
                return None, f"rocprofv3 {ctr} pass exited with {r.returncode}"
            step_kib, n = _counter_kib(d, ctr, "k_step")
            copy_kib, _ = _counter_kib(d, ctr, "k_copy")
        except (subprocess.TimeoutExpired, RuntimeError, sqlite3.Error) as e:
            return None, f"{ctr} pass: {e}"
        finally:
            shutil.rmtree(d, ignore_errors=True)
        scale = copy_bytes / (copy_kib * 1024.0)
        out[ctr] = {"kib_per_dispatch": round(step_kib, 1), "dispatches": n, "scale": round(scale, 4),
                    "bytes": int(step_kib * 1024.0 * scale)}
    return out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"], out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="hall", choices=["hall", "pillared_room", "shoebox"])
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp32_state64"])
    ap.add_argument("--zchunk", type=int, default=0)
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-traffic", action="store_true", help="skip the HBM counter passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU z-slab decomposition is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    from wavey_amd import configs, fdtd, workload

    if args.workload == "hall":
        cfg = configs.hall()
    elif args.workload == "pillared_room":
        cfg = configs.pillared_room()
    else:
        cfg = configs.shoebox(fdtd.fit_real_impedance)
    total = args.warmup + args.steps
    t0 = time.perf_counter()
    case = workload.build_case(cfg, steps=max(total, 1))
    t_build = time.perf_counter() - t0
    print(f"bench: {cfg['name']} grid {case.grid.n}, built in {t_build:.1f} s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    plan = fdtd.Plan(case.grid, args.precision)
    if args.zchunk or args.variant >= 0:
        plan.set_tile(args.zchunk, args.variant)
    taps = case.runs[0][2]
    # A plain copy of one field buffer: the attainable-bandwidth reference and
    # the calibration dispatch for HBM traffic counters.
    copy_bytes = case.grid.cells * (8 if args.precision == "fp64" else 4)
    copy_ms = plan.copy_probe(copy_bytes, reps=5)
    plan.begin(taps, case.signal, case.receivers)
    t_setup = time.perf_counter() - t0
    st = plan.stats()
    if args.warmup:
        plan.advance(args.warmup)
    t0 = time.perf_counter()
    plan.advance(args.steps)
    wall = time.perf_counter() - t0
    dev_ms, launches, _ = plan.last_timing()
    ir, _ = plan.finish((), fdtd.MODE_IR)
    if not np.all(np.isfinite(ir)):
        raise RuntimeError("non-finite receiver samples")
    plan.close()

    air = st["air_cells"]
    value = air * args.steps / wall / 1e9
    launch_s = dev_ms * 1e-3 / max(launches, 1)
    nbytes = algorithmic_bytes(st, len(case.receivers), len(taps), args.precision == "fp64")
    achieved = nbytes / launch_s / 1e9
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
                     "kernel": "k_step", "bytes_per_launch": int(nbytes),
                     "launch_us": round(launch_s * 1e6, 2),
                     "copy_GBs": round(2 * copy_bytes / (copy_ms * 1e-3) / 1e9, 1), "copy_bytes": 2 * copy_bytes},
        "setup_s": {"grid_build": round(t_build, 2), "upload_and_layout": round(t_setup, 2)},
    }
    if not args.no_traffic:
        traffic, detail = measure_traffic(args, copy_bytes)
        line["roofline"]["traffic"] = traffic
        line["roofline"]["traffic_detail"] = detail
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cfg, case)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
