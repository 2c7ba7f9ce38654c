"""Writes the FP64 golden receiver series of the shoebox (2000 steps, 4
receivers) and of the pillared room on a coarse grid (dx 0.08 m, 1500 steps,
its cardioid's monopole and +x dipole components, 8 of its receivers), from
the oracle (the FP64 restatement of fdtd.rs:644-725), with the inputs that
produced them. Run from the repository root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from wavey_amd import configs  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def shoebox_case():
    cfg = configs.shoebox(oracle.fit_real_impedance)
    og = oracle.build_box(cfg["room"], cfg["c"], cfg["dx"])
    sig = oracle.source_signal(cfg["f_src"], og.dt, cfg["steps"])
    rx = [oracle.nearest_air(og, p) for p in cfg["receivers"]]
    taps = oracle.taps_at(og, cfg["sources"][0]["pos"])
    return og, sig, rx, [taps]


def pillared_case():
    cfg = dict(configs.pillared_room(), dx=0.08, steps=1500)
    og = oracle.build_box(cfg["room"], cfg["c"], cfg["dx"])
    sig = oracle.source_signal(cfg["f_src"], og.dt, cfg["steps"])
    rx = [oracle.nearest_air(og, p) for p in cfg["receivers"][::8]]
    node = oracle.nearest_air(og, cfg["sources"][0]["pos"])
    comps = oracle.source_components(og, node, 1)
    return og, sig, rx, comps[:2]


CASES = {"shoebox": shoebox_case, "pillared_mini": pillared_case}


def main():
    for name, make in CASES.items():
        og, sig, rx, runs = make()
        irs = [oracle.run(og, taps, sig, rx, threads=8)[0] for taps in runs]
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), ir=np.array(irs), signal=np.asarray(sig),
                            receivers=np.asarray(rx, dtype=np.int64),
                            tap_cells=np.array([t[0] for taps in runs for t in taps], dtype=np.int64),
                            tap_weights=np.array([t[1] for taps in runs for t in taps]),
                            tap_counts=np.array([len(taps) for taps in runs], dtype=np.int64),
                            grid_n=np.asarray(og.n, dtype=np.int64), dt=og.dt)
        print(name, np.array(irs).shape)


if __name__ == "__main__":
    main()
