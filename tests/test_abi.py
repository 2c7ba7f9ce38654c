"""The C ABI: every function include/*.h declares is exported, and argument
errors come back as status codes with messages (no GPU needed)."""
import ctypes as C
import glob
import os
import re
import subprocess

import numpy as np

from wavey_amd import configs, fdtd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(wv_\w+)\s*\(", text):
            names.add(m.group(1))
    names.discard("wv_progress_fn")
    return names


def test_library_exports_every_declared_symbol():
    L = fdtd.lib()
    names = declared_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", fdtd.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = names - exported
    assert not missing, missing
    for n in names:
        getattr(L, n)
    assert set(fdtd.EXPORTS) <= exported


def test_library_is_gfx950():
    # The library built here carries a gfx950 code object.
    with open(fdtd.LIB_PATH, "rb") as f:
        assert b"gfx950" in f.read()


def test_invalid_grid_is_an_error_not_a_crash():
    L = fdtd.lib()
    g = fdtd.WvGrid()  # null arrays
    h = C.c_void_p()
    err = C.create_string_buffer(256)
    st = L.wv_plan_create(C.byref(g), 0, C.byref(h), err, 256)
    assert st == -1 and b"invalid grid" in err.value
    cfg = configs.shoebox(fdtd.fit_real_impedance)
    grid = fdtd.FdtdGrid.build_box(cfg["room"], 343.0, 0.1)
    st = L.wv_plan_create(C.byref(grid.view()), 7, C.byref(h), err, 256)
    assert st == -1
    # Boundary list pointing outside the grid.
    bad = fdtd.FdtdGrid(n=grid.n, origin=grid.origin, dx=grid.dx, dt=grid.dt, flags=grid.flags,
                        b_cell=np.array([grid.cells + 5], np.uint32), b_start=np.array([0, 1], np.uint32),
                        slots=np.zeros((1, 4)))
    st = L.wv_plan_create(C.byref(bad.view()), 0, C.byref(h), err, 256)
    assert st == -1 and b"out of range" in err.value


def test_grid_limit_is_reported():
    room = {"L": (60.0, 60.0, 30.0), "pillars": [], "fits": [(0.1, [])], "patches": [(1, 1, [0])] * 6}
    try:
        fdtd.FdtdGrid.build_box(room, 343.0, 0.04)
    except fdtd.FdtdError as e:
        assert e.status == -3 and "M cells" in str(e)
    else:
        raise AssertionError("a 1.7G-cell grid must be refused")


def test_no_device_is_a_device_error():
    # Without a GPU (this container) a valid grid fails at the first HIP call
    # with WV_ERR_DEVICE: the caller's CPU fallback path (run.rs:186-191).
    from conftest import gpu_available

    if gpu_available():
        return
    cfg = configs.shoebox(fdtd.fit_real_impedance)
    grid = fdtd.FdtdGrid.build_box(cfg["room"], 343.0, 0.1)
    try:
        fdtd.Plan(grid)
    except fdtd.FdtdError as e:
        assert e.status in (-2, -4) and str(e)
    else:
        raise AssertionError("expected a device error without a GPU")


def test_batch_and_server_without_device_are_errors():
    from conftest import gpu_available

    if gpu_available():
        return
    for make in (lambda: fdtd.Batch("fp32"), lambda: fdtd.Server("fp32")):
        try:
            make()
        except fdtd.FdtdError as e:
            assert e.status == -2 and str(e)
        else:
            raise AssertionError("expected a device error without a GPU")


def test_footprint_estimate_covers_fields():
    cfg = configs.shoebox(fdtd.fit_real_impedance)
    grid = fdtd.FdtdGrid.build_box(cfg["room"], 343.0, 0.1)
    n = grid.n
    pitched = ((n[0] + 63) // 64 * 64) * n[1] * n[2]
    e32 = fdtd.footprint_estimate(grid, "fp32", 2000, 4, 8)
    e64 = fdtd.footprint_estimate(grid, "fp64", 2000, 4, 8)
    assert e32 > 2 * 4 * pitched + 2000 * 4 * 4
    assert e64 > 2 * 8 * pitched and e64 > e32


C_PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "wavey_fdtd.h"

int main(void) {
    double sig[16];
    wv_source_signal(250.0, 1.6832e-4, 16, sig);
    /* An invalid grid must come back as an error with a message, not a crash. */
    wv_grid g;
    memset(&g, 0, sizeof g);
    wv_plan* plan = NULL;
    char err[256] = {0};
    wv_status st = wv_plan_create(&g, WV_FP32, &plan, err, sizeof err);
    printf("%.17g %d %s | %s\n", sig[5], (int)st, err, wv_version());
    return st == WV_ERR_INVALID && plan == NULL && err[0] ? 0 : 1;
}
"""


def test_header_builds_and_links_as_c(tmp_path):
    """A C99 program includes the header, links the library and calls it
    (no GPU needed for these calls)."""
    import subprocess

    src = tmp_path / "abi.c"
    src.write_text(C_PROGRAM)
    exe = tmp_path / "abi"
    libdir = os.path.dirname(fdtd.LIB_PATH)
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, str(src), "-o", str(exe), "-L", libdir,
                    "-lwavey_fdtd", f"-Wl,-rpath,{libdir}"], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = fdtd.source_signal(250.0, 1.6832e-4, 16)
    assert float(r.stdout.split()[0]) == ref[5]


def test_batch_launch_form_argument_errors():
    # A null batch or an unknown launch form is refused before any HIP call.
    L = fdtd.lib()
    err = C.create_string_buffer(256)
    for form in (fdtd.BATCH_STEPS, fdtd.BATCH_PERSISTENT, 5):
        assert L.wv_batch_set_launch(None, form, err, 256) == -1
        assert b"invalid argument" in err.value
