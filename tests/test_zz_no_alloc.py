"""GPU: the allocation-free entry point wv_fdtd_run_noalloc against
wv_fdtd_run, bit for bit.

Two calls in a row on the same device array, the second on a smaller grid
(the pillared room at dx 0.08 m, then the shoebox), in FP32 and FP64, with
receiver series and spectra; then a run too large for the array is refused
with WV_ERR_LIMIT and the array still serves a run afterwards; and a run
whose default pass mode does not fit the array takes single steps.
"""
import numpy as np
import pytest

from wavey_amd import configs, fdtd, workload

pytestmark = pytest.mark.gpu

FREQS = [31.25, 62.5, 125.0, 250.0]


def cases():
    big = dict(configs.pillared_room(), dx=0.08, steps=700)
    small = dict(configs.shoebox(fdtd.fit_real_impedance), steps=500)
    out = []
    for cfg, comp in ((big, 1), (small, 0)):
        c = workload.build_case(cfg)
        out.append((cfg["name"], c.grid, c.runs[comp][2], c.signal, c.receivers))
    assert out[0][1].cells > 4 * out[1][1].cells
    return out


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_noalloc_matches_fdtd_run_on_two_calls(prec):
    runs = cases()
    got = []
    for name, g, taps, sig, rx in runs:
        print(f"case: wv_fdtd_run_noalloc {name} {g.n}, {len(sig)} steps, {prec}", flush=True)
        got.append(fdtd.fdtd_run_noalloc(g, taps, sig, rx, FREQS, fdtd.MODE_BOTH, prec))
    for (name, g, taps, sig, rx), (ir, sp) in zip(runs, got):
        print(f"case: wv_fdtd_run {name} {g.n}, {prec}", flush=True)
        ref_ir, ref_sp = fdtd.fdtd_run(g, taps, sig, rx, FREQS, fdtd.MODE_BOTH, prec)
        assert np.count_nonzero(ref_ir) > 0.3 * ref_ir.size
        assert np.array_equal(ir, ref_ir), name
        assert np.array_equal(sp, ref_sp), name


def test_noalloc_refuses_a_run_beyond_its_array():
    room = {"L": (20.0, 16.0, 4.0), "pillars": [], "fits": [(0.5, [])], "patches": configs.uniform_patches(0, lambda s: 0)}
    g = fdtd.FdtdGrid.build_box(room, configs.C_AIR, 0.04)
    sig = fdtd.source_signal(250.0, g.dt, 10)
    src = g.nearest_air((10.0, 8.0, 2.0))
    with pytest.raises(fdtd.FdtdError) as e:
        fdtd.fdtd_run_noalloc(g, [(src, 1.0)], sig, [src], (), fdtd.MODE_IR, "fp32")
    assert e.value.status == -3 and "64 MiB" in str(e.value)
    # The array still serves a run after the refusal.
    name, g2, taps, sig2, rx = cases()[1]
    print(f"case: wv_fdtd_run_noalloc {name} after a refused run", flush=True)
    ir, _ = fdtd.fdtd_run_noalloc(g2, taps, sig2, rx, (), fdtd.MODE_IR, "fp32")
    ref, _ = fdtd.fdtd_run(g2, taps, sig2, rx, (), fdtd.MODE_IR, "fp32")
    assert np.array_equal(ir, ref)


def test_noalloc_takes_single_steps_where_passes_do_not_fit():
    # An FP64 grid of 1.37 M cells defaults to two steps per pass (four field
    # buffers), whose footprint exceeds the array; the allocation-free call
    # runs it with single steps (two buffers) and gives the same results.
    case = workload.build_case(dict(configs.pillared_room(), dx=0.07, steps=400))
    g, taps = case.grid, case.runs[1][2]
    with fdtd.Plan(g, "fp64") as plan:
        assert plan.pass_info()["steps_per_pass"] == 2
    need = fdtd.footprint_estimate(g, "fp64", len(case.signal), len(case.receivers), len(taps))
    assert need > 64 << 20
    print(f"case: wv_fdtd_run_noalloc {g.n} fp64, default pass mode beyond the array", flush=True)
    ir, sp = fdtd.fdtd_run_noalloc(g, taps, case.signal, case.receivers, FREQS, fdtd.MODE_BOTH, "fp64")
    ref_ir, ref_sp = fdtd.fdtd_run(g, taps, case.signal, case.receivers, FREQS, fdtd.MODE_BOTH, "fp64")
    assert np.count_nonzero(ref_ir) > 0.3 * ref_ir.size
    assert np.array_equal(ir, ref_ir)
    assert np.array_equal(sp, ref_sp)
