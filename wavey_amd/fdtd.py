"""Python view of the C ABI in include/wavey_fdtd.h (libwavey_fdtd.so).

Names and argument meanings follow the reference's FDTD interface in
core/wavey-solver/src/fdtd.rs and gpu_fdtd.rs:

* ``FdtdGrid``: the voxel grid (fdtd.rs:196-209) with ``build_box`` (the
  box-room case of ``build_dx``, fdtd.rs:289-377), ``index``, ``center``,
  ``nearest_air`` (:245-272) and ``taps_at`` (:467-492).
* ``source_signal`` (:601-610), ``source_components`` (:730-743),
  ``component_weights`` (:812-815), ``to_transfer`` (:819-840).
* ``Gpu.run_fdtd`` (gpu_fdtd.rs:45-54): ``Ok(Some(spectra))`` is the returned
  array, ``Ok(None)`` (cancelled) is ``None`` and ``Err(String)`` raises
  ``FdtdError`` with the message. ``Gpu.run_ir`` is ``run_cpu_ir``
  (fdtd.rs:630-641) on the device.

There is no CPU fallback here: every run goes through the HIP library, and a
missing library raises at import of the first call.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# WAVEY_FDTD_LIB names another build of the same library (for example the
# host-sanitizer build of wavey_amd/csrc/Makefile's `asan` target).
LIB_PATH = os.environ.get("WAVEY_FDTD_LIB") or os.path.join(_HERE, "libwavey_fdtd.so")

WV_OK, WV_CANCELLED = 0, 1
WV_ERR_INVALID, WV_ERR_DEVICE, WV_ERR_LIMIT, WV_ERR_NOMEM = -1, -2, -3, -4
MODE_SPECTRA, MODE_IR, MODE_BOTH = 1, 2, 3
FP32, FP64, FP32_STATE64 = 0, 1, 2
PRECISIONS = {"fp32": FP32, "fp64": FP64, "fp32_state64": FP32_STATE64}
SERVE_ALL, SERVE_SHORTEST, SERVE_LANES = 0, 1, 2
BATCH_STEPS, BATCH_PERSISTENT = 0, 1
COURANT = 0.5773502691896257
AIR = 1 << 6


class FdtdError(RuntimeError):
    """Err(String) of the reference run (gpu_fdtd.rs:54); carries the status."""

    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


class WvSlot(C.Structure):
    _fields_ = [("inv_d", C.c_double), ("cv", C.c_double), ("k", C.c_double), ("w", C.c_double)]


class WvGrid(C.Structure):
    _fields_ = [("n", C.c_uint64 * 3), ("dx", C.c_double), ("dt", C.c_double), ("origin", C.c_double * 3),
                ("flags", C.POINTER(C.c_uint8)), ("n_boundary", C.c_uint64), ("b_cell", C.POINTER(C.c_uint32)),
                ("b_start", C.POINTER(C.c_uint32)), ("slots", C.POINTER(WvSlot))]


class WvTap(C.Structure):
    _fields_ = [("cell", C.c_uint64), ("w", C.c_double)]


class WvBranch(C.Structure):
    _fields_ = [("r", C.c_double), ("m", C.c_double), ("k", C.c_double)]


class WvAdmittance(C.Structure):
    _fields_ = [("beta0", C.c_double), ("n_branches", C.c_uint32), ("branches", C.POINTER(WvBranch)),
                ("band", C.c_double * 2), ("rms_error", C.c_double), ("max_error", C.c_double)]


class WvBoxRoom(C.Structure):
    _fields_ = [("size", C.c_double * 3), ("n_pillars", C.c_uint32), ("pillars", C.POINTER(C.c_double)),
                ("n_fits", C.c_uint32), ("fits", C.POINTER(WvAdmittance)), ("patch_gu", C.POINTER(C.c_uint32)),
                ("patch_gv", C.POINTER(C.c_uint32)), ("patch_off", C.POINTER(C.c_uint32)),
                ("patch_fit", C.POINTER(C.c_uint32))]


class WvSpectrum(C.Structure):
    _fields_ = [("n", C.c_uint32), ("freqs_hz", C.POINTER(C.c_double)), ("values", C.POINTER(C.c_double))]


class WvMaterial(C.Structure):
    _fields_ = [("absorption", WvSpectrum), ("impedance_re", C.POINTER(WvSpectrum)),
                ("impedance_im", C.POINTER(WvSpectrum))]


class WvFitResult(C.Structure):
    _fields_ = [("beta0", C.c_double), ("n_branches", C.c_uint32), ("branches", WvBranch * 2),
                ("band", C.c_double * 2), ("rms_error", C.c_double), ("max_error", C.c_double)]


class WvRoomMesh(C.Structure):
    _fields_ = [("prism", C.c_int32), ("height", C.c_double), ("n_positions", C.c_uint32),
                ("positions", C.POINTER(C.c_double)), ("n_triangles", C.c_uint32),
                ("triangles", C.POINTER(C.c_uint32)), ("tri_surface", C.POINTER(C.c_uint32)),
                ("n_fits", C.c_uint32), ("fits", C.POINTER(WvAdmittance)), ("n_regions", C.c_uint32),
                ("region_cycles", C.POINTER(C.c_uint32)), ("cycle_off", C.POINTER(C.c_uint32)),
                ("cycle_xy", C.POINTER(C.c_double))]


class WvSolid(C.Structure):
    _fields_ = [("n_planes", C.c_uint32), ("planes", C.POINTER(C.c_double)), ("lo", C.c_double * 3),
                ("hi", C.c_double * 3)]


PROGRESS_FN = C.CFUNCTYPE(None, C.c_double, C.c_void_p)

class WvMicrophone(C.Structure):
    _fields_ = [("id", C.c_uint64), ("name", C.c_char_p), ("position", C.c_double * 3)]


class WvFdtdSettings(C.Structure):
    _fields_ = [("crossover_hz", C.c_double), ("duration_s", C.c_double), ("points_per_wavelength", C.c_double),
                ("source_order", C.c_uint32), ("cabinets", C.c_int32), ("accuracy", C.c_int32)]


class WvResults(C.Structure):
    _fields_ = [("nf", C.c_size_t), ("freqs", C.POINTER(C.c_double)), ("n_speakers", C.c_size_t),
                ("speaker_ids", C.POINTER(C.c_uint64)), ("n_mics", C.c_size_t), ("mics", C.POINTER(WvMicrophone)),
                ("transfer", C.POINTER(C.c_double)), ("backend", C.c_char_p), ("seconds", C.c_double),
                ("images", C.c_uint64), ("paths", C.c_uint64), ("diffracted_paths", C.c_uint64),
                ("receivers", C.c_uint64), ("n_notes", C.c_size_t), ("notes", C.POINTER(C.c_char_p)),
                ("fingerprint", C.c_uint64), ("rt_sabine", C.POINTER(C.c_double)),
                ("rt_eyring", C.POINTER(C.c_double)), ("fdtd", C.POINTER(WvFdtdSettings))]


class WvWaveSource(C.Structure):
    _fields_ = [("taps", C.POINTER(WvTap)), ("comp_off", C.POINTER(C.c_uint32)), ("n_comps", C.c_uint32),
                ("weights", C.POINTER(C.c_double)), ("h0", C.POINTER(C.c_double)), ("r_ref", C.c_double),
                ("on_axis", C.POINTER(C.c_double))]


EXPORTS = [
    "wv_fdtd_run", "wv_fdtd_run_noalloc", "wv_plan_create", "wv_plan_run", "wv_run_begin", "wv_run_advance", "wv_run_finish",
    "wv_plan_last_timing", "wv_plan_stats", "wv_plan_set_tile", "wv_plan_set_pass", "wv_plan_pass_info", "wv_plan_copy_probe", "wv_plan_free", "wv_grid_build_box",
    "wv_built_grid_view", "wv_grid_free", "wv_source_signal", "wv_nearest_air", "wv_air_along",
    "wv_wave_source_taps", "wv_taps_at",
    "wv_source_components", "wv_component_weights", "wv_to_transfer", "wv_admittance_eval", "wv_paris",
    "wv_fit_real_impedance", "wv_version", "wv_footprint_estimate", "wv_batch_create", "wv_batch_add",
    "wv_batch_autotune", "wv_batch_set_tile", "wv_batch_set_launch", "wv_batch_run", "wv_batch_result", "wv_batch_info",
    "wv_batch_timing", "wv_batch_free", "wv_serve_create", "wv_serve_submit", "wv_serve_wait", "wv_serve_stats", "wv_serve_set_policy",
    "wv_serve_free", "wv_wave_job", "wv_results_save", "wv_results_read_transfer", "wv_grid_build_room", "wv_solid_from_mesh", "wv_grid_spacing",
    "wv_baffle_point", "wv_blend", "wv_calibrate_cabinet", "wv_material_absorption", "wv_fit_admittance",
    "wv_first_order_directions", "wv_fit_first_order", "wv_accuracy_headroom", "wv_accuracy_cabinet_cells",
    "wv_settings_f_max", "wv_grid_fmax", "wv_crossover_weight", "wv_phase_error_deg_per_m",
]

_lib = None
_lock = threading.Lock()


def lib():
    """The loaded HIP library; raises when it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise FdtdError(WV_ERR_DEVICE, f"libwavey_fdtd.so is not built ({LIB_PATH}); run __graft_entry__.build()")
            L = C.CDLL(LIB_PATH)
            P = C.POINTER
            st, sz, dp = C.c_int, C.c_size_t, P(C.c_double)
            errbuf = [C.c_char_p, sz]
            L.wv_fdtd_run.argtypes = [P(WvGrid), P(WvTap), sz, dp, sz, P(C.c_uint64), sz, dp, sz, C.c_int, C.c_int,
                                      P(C.c_int32), PROGRESS_FN, C.c_void_p, dp, dp] + errbuf
            L.wv_fdtd_run.restype = st
            L.wv_fdtd_run_noalloc.argtypes = L.wv_fdtd_run.argtypes
            L.wv_fdtd_run_noalloc.restype = st
            L.wv_plan_create.argtypes = [P(WvGrid), C.c_int, P(C.c_void_p)] + errbuf
            L.wv_plan_create.restype = st
            L.wv_plan_run.argtypes = [C.c_void_p, P(WvTap), sz, dp, sz, P(C.c_uint64), sz, dp, sz, C.c_int,
                                      P(C.c_int32), PROGRESS_FN, C.c_void_p, dp, dp] + errbuf
            L.wv_plan_run.restype = st
            L.wv_run_begin.argtypes = [C.c_void_p, P(WvTap), sz, dp, sz, P(C.c_uint64), sz] + errbuf
            L.wv_run_begin.restype = st
            L.wv_run_advance.argtypes = [C.c_void_p, sz, P(C.c_int32), PROGRESS_FN, C.c_void_p] + errbuf
            L.wv_run_advance.restype = st
            L.wv_run_finish.argtypes = [C.c_void_p, dp, sz, C.c_int, dp, dp] + errbuf
            L.wv_run_finish.restype = st
            L.wv_plan_last_timing.argtypes = [C.c_void_p, dp, P(C.c_uint64), P(C.c_uint64)]
            L.wv_plan_last_timing.restype = st
            L.wv_plan_stats.argtypes = [C.c_void_p, P(C.c_uint64)]
            L.wv_plan_stats.restype = st
            L.wv_plan_set_tile.argtypes = [C.c_void_p, C.c_uint32, C.c_int32]
            L.wv_plan_set_tile.restype = st
            L.wv_plan_set_pass.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
            L.wv_plan_set_pass.restype = st
            L.wv_plan_pass_info.argtypes = [C.c_void_p, P(C.c_uint32)]
            L.wv_plan_pass_info.restype = st
            L.wv_plan_copy_probe.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, dp]
            L.wv_plan_copy_probe.restype = st
            L.wv_plan_free.argtypes = [C.c_void_p]
            L.wv_plan_free.restype = None
            L.wv_grid_build_box.argtypes = [P(WvBoxRoom), C.c_double, C.c_double, C.c_int, P(C.c_void_p)] + errbuf
            L.wv_grid_build_box.restype = st
            L.wv_grid_build_room.argtypes = [P(WvRoomMesh), P(WvSolid), sz, C.c_double, C.c_double, C.c_int,
                                             P(C.c_void_p)] + errbuf
            L.wv_grid_build_room.restype = st
            L.wv_solid_from_mesh.argtypes = [dp, C.c_uint32, P(C.c_uint32), C.c_uint32, dp, dp, dp, P(C.c_uint32),
                                             dp, dp]
            L.wv_solid_from_mesh.restype = st
            L.wv_grid_spacing.argtypes = [C.c_double, C.c_double, C.c_double, P(WvSolid), sz]
            L.wv_grid_spacing.restype = C.c_double
            L.wv_built_grid_view.argtypes = [C.c_void_p]
            L.wv_built_grid_view.restype = P(WvGrid)
            L.wv_grid_free.argtypes = [C.c_void_p]
            L.wv_grid_free.restype = None
            L.wv_source_signal.argtypes = [C.c_double, C.c_double, sz, dp]
            L.wv_source_signal.restype = None
            L.wv_nearest_air.argtypes = [P(WvGrid), dp, P(C.c_uint64)]
            L.wv_nearest_air.restype = st
            L.wv_air_along.argtypes = [P(WvGrid), dp, dp, P(C.c_uint64)]
            L.wv_air_along.restype = st
            L.wv_wave_source_taps.argtypes = [P(WvGrid), dp, dp, C.c_int, C.c_uint32, P(WvTap), P(C.c_uint32),
                                              P(C.c_uint32), P(C.c_int32)]
            L.wv_wave_source_taps.restype = st
            L.wv_taps_at.argtypes = [P(WvGrid), dp, P(WvTap), P(C.c_size_t)]
            L.wv_taps_at.restype = st
            L.wv_source_components.argtypes = [P(WvGrid), C.c_uint64, C.c_uint32, P(WvTap), P(C.c_uint32),
                                               P(C.c_uint32)]
            L.wv_source_components.restype = st
            L.wv_component_weights.argtypes = [dp, dp, C.c_double, C.c_double, dp]
            L.wv_component_weights.restype = None
            L.wv_to_transfer.argtypes = [dp, sz, dp, sz, C.c_double, C.c_double, dp, sz, C.c_double, C.c_double,
                                         dp, dp]
            L.wv_to_transfer.restype = None
            L.wv_admittance_eval.argtypes = [P(WvAdmittance), C.c_double, dp]
            L.wv_admittance_eval.restype = None
            L.wv_paris.argtypes = [C.c_double, C.c_double, C.c_uint32]
            L.wv_paris.restype = C.c_double
            L.wv_fit_real_impedance.argtypes = [C.c_double]
            L.wv_fit_real_impedance.restype = C.c_double
            L.wv_version.argtypes = []
            L.wv_version.restype = C.c_char_p
            L.wv_footprint_estimate.argtypes = [P(WvGrid), C.c_int, sz, sz, sz, P(C.c_uint64)]
            L.wv_footprint_estimate.restype = st
            L.wv_batch_create.argtypes = [C.c_int, P(C.c_void_p)] + errbuf
            L.wv_batch_create.restype = st
            L.wv_batch_add.argtypes = [C.c_void_p, P(WvGrid), P(WvTap), sz, dp, sz, P(C.c_uint64), sz,
                                       P(C.c_int32)] + errbuf
            L.wv_batch_add.restype = st
            L.wv_batch_autotune.argtypes = [C.c_void_p, C.c_uint32] + errbuf
            L.wv_batch_autotune.restype = st
            L.wv_batch_set_tile.argtypes = [C.c_void_p, C.c_int32, C.c_uint32] + errbuf
            L.wv_batch_set_tile.restype = st
            L.wv_batch_set_launch.argtypes = [C.c_void_p, C.c_int] + errbuf
            L.wv_batch_set_launch.restype = st
            L.wv_batch_run.argtypes = [C.c_void_p, C.c_int, P(C.c_int32), PROGRESS_FN, C.c_void_p] + errbuf
            L.wv_batch_run.restype = st
            L.wv_batch_result.argtypes = [C.c_void_p, C.c_int32, dp, sz, C.c_int, dp, dp] + errbuf
            L.wv_batch_result.restype = st
            L.wv_batch_info.argtypes = [C.c_void_p, C.c_int32, P(C.c_uint64)]
            L.wv_batch_info.restype = st
            L.wv_batch_timing.argtypes = [C.c_void_p, dp, P(C.c_uint64)]
            L.wv_batch_timing.restype = st
            L.wv_batch_free.argtypes = [C.c_void_p]
            L.wv_batch_free.restype = None
            L.wv_serve_create.argtypes = [C.c_int, C.c_uint64, P(C.c_void_p)] + errbuf
            L.wv_serve_create.restype = st
            L.wv_serve_submit.argtypes = [C.c_void_p, P(WvGrid), P(WvTap), sz, dp, sz, P(C.c_uint64), sz, dp, sz,
                                          C.c_int, P(C.c_uint64)] + errbuf
            L.wv_serve_submit.restype = st
            L.wv_serve_wait.argtypes = [C.c_void_p, C.c_uint64, dp, dp, dp] + errbuf
            L.wv_serve_wait.restype = st
            L.wv_serve_set_policy.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64]
            L.wv_serve_set_policy.restype = st
            L.wv_serve_stats.argtypes = [C.c_void_p, dp]
            L.wv_serve_stats.restype = st
            L.wv_serve_free.argtypes = [C.c_void_p]
            L.wv_serve_free.restype = None
            L.wv_results_save.argtypes = [C.c_char_p, P(WvResults)] + errbuf
            L.wv_results_save.restype = st
            L.wv_results_read_transfer.argtypes = [C.c_char_p, sz, sz, sz, P(C.c_float)] + errbuf
            L.wv_results_read_transfer.restype = st
            L.wv_wave_job.argtypes = [P(WvGrid), P(WvWaveSource), sz, dp, sz, P(C.c_uint64), sz, dp, sz, C.c_double,
                                      C.c_int, C.c_int, P(C.c_int32), dp] + errbuf
            L.wv_wave_job.restype = st
            L.wv_baffle_point.argtypes = [P(WvGrid), dp, dp, dp]
            L.wv_baffle_point.restype = None
            L.wv_blend.argtypes = [dp, dp, C.c_double, dp]
            L.wv_blend.restype = None
            L.wv_calibrate_cabinet.argtypes = [dp, C.c_uint32, P(C.c_uint32), C.c_uint32, dp, C.c_double, C.c_double,
                                               C.c_double, C.c_double, dp, sz, C.c_int, C.c_int, P(C.c_int32), dp,
                                               dp, dp] + errbuf
            L.wv_calibrate_cabinet.restype = st
            L.wv_material_absorption.argtypes = [P(WvMaterial), C.c_double]
            L.wv_material_absorption.restype = C.c_double
            L.wv_fit_admittance.argtypes = [P(WvMaterial), C.c_double, C.c_double, C.c_uint32, P(WvFitResult)]
            L.wv_fit_admittance.restype = st
            L.wv_first_order_directions.argtypes = [dp]
            L.wv_first_order_directions.restype = None
            L.wv_fit_first_order.argtypes = [dp, dp, dp, dp, dp]
            L.wv_fit_first_order.restype = C.c_double
            for name in ("wv_accuracy_headroom", "wv_accuracy_cabinet_cells"):
                getattr(L, name).argtypes = [C.c_int]
                getattr(L, name).restype = C.c_double
            L.wv_settings_f_max.argtypes = [C.c_double]
            L.wv_settings_f_max.restype = C.c_double
            L.wv_grid_fmax.argtypes = [C.c_double, C.c_int, dp, sz]
            L.wv_grid_fmax.restype = C.c_double
            L.wv_crossover_weight.argtypes = [C.c_double, C.c_double]
            L.wv_crossover_weight.restype = C.c_double
            L.wv_phase_error_deg_per_m.argtypes = [C.c_double, C.c_double, C.c_double]
            L.wv_phase_error_deg_per_m.restype = C.c_double
            _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(status, err):
    if status < 0:
        raise FdtdError(status, err.value.decode(errors="replace"))
    return status


# ------------------------------------------------------------------- grid

def _admittances(fits):
    """Fits as (beta0, [(r, m, k), ...]) or (beta0, branches, meta) with meta
    = {"band": (lo, hi), "rms_error", "max_error"} (zeros when absent)."""
    keep = []
    arr = (WvAdmittance * len(fits))()
    for i, fit in enumerate(fits):
        beta0, branches = fit[0], fit[1]
        meta = fit[2] if len(fit) > 2 and fit[2] else {}
        br = (WvBranch * max(1, len(branches)))()
        for j, (r, m, k) in enumerate(branches):
            br[j].r, br[j].m, br[j].k = float(r), float(m), float(k)
        keep.append(br)
        arr[i].beta0 = float(beta0)
        arr[i].n_branches = len(branches)
        arr[i].branches = C.cast(br, C.POINTER(WvBranch))
        arr[i].band[0], arr[i].band[1] = (float(v) for v in meta.get("band", (0.0, 0.0)))
        arr[i].rms_error = float(meta.get("rms_error", 0.0))
        arr[i].max_error = float(meta.get("max_error", 0.0))
    return arr, keep


def _room_mesh(mesh):
    u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))  # noqa: E731
    r = WvRoomMesh()
    pos = np.ascontiguousarray(np.asarray(mesh["positions"], dtype=np.float64).reshape(-1, 3))
    tri = np.ascontiguousarray(np.asarray(mesh["triangles"], dtype=np.uint32).reshape(-1, 3))
    ts = np.ascontiguousarray(np.asarray(mesh["tri_surface"], dtype=np.uint32).reshape(-1))
    fits, keep = _admittances(mesh["fits"])
    prism = bool(mesh.get("prism", False))
    regions = mesh.get("regions", []) if prism else []
    rc, off, xy = [0], [0], []
    for reg in regions:
        for cyc in reg:
            cyc = np.asarray(cyc, dtype=np.float64).reshape(-1, 2)
            xy.append(cyc)
            off.append(off[-1] + len(cyc))
        rc.append(len(off) - 1)
    rc = np.array(rc, dtype=np.uint32)
    off = np.array(off, dtype=np.uint32)
    xy = np.ascontiguousarray(np.concatenate(xy) if xy else np.zeros((1, 2)))
    r.prism = int(prism)
    r.height = float(mesh.get("height", 0.0))
    r.n_positions, r.positions = len(pos), _dp(pos)
    r.n_triangles, r.triangles, r.tri_surface = len(tri), u32(tri), u32(ts)
    r.n_fits, r.fits = len(mesh["fits"]), C.cast(fits, C.POINTER(WvAdmittance))
    r.n_regions = len(regions)
    r.region_cycles, r.cycle_off, r.cycle_xy = u32(rc), u32(off), _dp(xy)
    return r, keep + [pos, tri, ts, fits, rc, off, xy]


def _solids(solids):
    arr = (WvSolid * max(1, len(solids)))()
    keep = []
    for i, s in enumerate(solids):
        pl = np.ascontiguousarray(np.asarray(s["planes"], dtype=np.float64).reshape(-1, 4))
        keep.append(pl)
        arr[i].n_planes = len(pl)
        arr[i].planes = _dp(pl)
        for k in range(3):
            arr[i].lo[k], arr[i].hi[k] = float(s["lo"][k]), float(s["hi"][k])
    return arr, keep


def solid_from_mesh(verts, tris, rot=None, pos=(0.0, 0.0, 0.0)):
    """Solid::from_mesh (fdtd.rs:433-456): a convex mesh (outward winding)
    turned by ``rot`` (3x3, applied as rot @ v) and moved by ``pos``."""
    v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.uint32).reshape(-1, 3))
    R = np.ascontiguousarray(np.eye(3) if rot is None else np.asarray(rot, dtype=np.float64).reshape(3, 3))
    P = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3))
    planes = np.zeros((max(1, len(t)), 4))
    lo, hi = np.zeros(3), np.zeros(3)
    n = C.c_uint32()
    st = lib().wv_solid_from_mesh(_dp(v), len(v), t.ctypes.data_as(C.POINTER(C.c_uint32)), len(t), _dp(R), _dp(P),
                                  _dp(planes), C.byref(n), _dp(lo), _dp(hi))
    if st != 0:
        return None
    return {"planes": planes[:n.value].copy(), "lo": lo, "hi": hi}


def grid_spacing(c, f_max, ppw, solids=()):
    """FdtdGrid::build_with's spacing (fdtd.rs:282-286)."""
    sa, keep = _solids(solids)
    return float(lib().wv_grid_spacing(c, f_max, ppw, sa, len(solids)))


@dataclass
class FdtdGrid:
    """Voxelized room (fdtd.rs:196-209). Arrays are numpy; ``view()`` is the C struct."""
    n: tuple
    origin: tuple
    dx: float
    dt: float
    flags: np.ndarray
    b_cell: np.ndarray
    b_start: np.ndarray
    slots: np.ndarray  # (S, 4): inv_d, cv, k, w
    _keep: list = field(default_factory=list, repr=False)

    @classmethod
    def build_box(cls, room, c, dx, threads=0):
        """Box room with pillars (see ``wv_box_room``): fdtd.rs:289-377."""
        L = lib()
        r = WvBoxRoom()
        for k in range(3):
            r.size[k] = float(room["L"][k])
        pil = np.ascontiguousarray(np.array(room.get("pillars", []), dtype=np.float64).reshape(-1, 4))
        fits, keep = _admittances(room["fits"])
        patches = room["patches"]
        if len(patches) != 6 + 4 * len(pil):
            raise FdtdError(WV_ERR_INVALID, "one patch entry per surface is needed")
        gu = np.array([p[0] for p in patches], dtype=np.uint32)
        gv = np.array([p[1] for p in patches], dtype=np.uint32)
        off = np.zeros(len(patches), dtype=np.uint32)
        off[1:] = np.cumsum(gu * gv)[:-1]
        pf = np.concatenate([np.asarray(p[2], dtype=np.uint32).ravel() for p in patches]).astype(np.uint32)
        r.n_pillars = len(pil)
        r.pillars = _dp(pil)
        r.n_fits = len(fits)
        r.fits = C.cast(fits, C.POINTER(WvAdmittance))
        u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))  # noqa: E731
        r.patch_gu, r.patch_gv, r.patch_off, r.patch_fit = u32(gu), u32(gv), u32(off), u32(pf)
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        _check(L.wv_grid_build_box(C.byref(r), c, dx, threads, C.byref(h), err, 512), err)
        return cls._from_built(h)

    @classmethod
    def build_room(cls, mesh, c, dx, solids=(), threads=0):
        """General room with rigid convex solids (``wv_room_mesh``,
        ``wv_solid``): fdtd.rs:289-461.

        ``mesh`` = {"positions": (V, 3), "triangles": (T, 3), "tri_surface":
        (T,), "fits": [(beta0, [(r, m, k), ...]), ...] one per surface,
        "prism": bool, "height": h, "regions": [[outer (n, 2), hole, ...],
        ...]}; ``solids`` = [{"planes": (P, 4), "lo": (3,), "hi": (3,)}].
        """
        L = lib()
        r, keep = _room_mesh(mesh)
        sa, keep2 = _solids(solids)
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        _check(L.wv_grid_build_room(C.byref(r), sa, len(solids), c, dx, threads, C.byref(h), err, 512), err)
        return cls._from_built(h)

    @classmethod
    def _from_built(cls, h):
        L = lib()
        try:
            v = L.wv_built_grid_view(h).contents
            n = tuple(int(v.n[k]) for k in range(3))
            N = n[0] * n[1] * n[2]
            nb = int(v.n_boundary)
            flags = np.ctypeslib.as_array(v.flags, shape=(N,)).copy()
            b_cell = np.ctypeslib.as_array(v.b_cell, shape=(nb,)).copy() if nb else np.zeros(0, np.uint32)
            b_start = np.ctypeslib.as_array(v.b_start, shape=(nb + 1,)).copy()
            ns = int(b_start[-1])
            if ns:
                raw = np.ctypeslib.as_array(C.cast(v.slots, C.POINTER(C.c_double)), shape=(4 * ns,)).copy()
                slots = raw.reshape(ns, 4)
            else:
                slots = np.zeros((0, 4))
            return cls(n=n, origin=tuple(v.origin[k] for k in range(3)), dx=v.dx, dt=v.dt, flags=flags,
                       b_cell=b_cell, b_start=b_start, slots=np.ascontiguousarray(slots))
        finally:
            L.wv_grid_free(h)

    @property
    def cells(self):
        return self.n[0] * self.n[1] * self.n[2]

    def air_cells(self):
        return int(np.count_nonzero(self.flags & AIR))

    def index(self, x, y, z):
        return (z * self.n[1] + y) * self.n[0] + x

    def center(self, i):
        nx, ny = self.n[0], self.n[1]
        z, r = divmod(int(i), nx * ny)
        y, x = divmod(r, nx)
        return np.array([self.origin[0] + x * self.dx, self.origin[1] + y * self.dx, self.origin[2] + z * self.dx])

    def view(self):
        g = WvGrid()
        for k in range(3):
            g.n[k] = self.n[k]
            g.origin[k] = self.origin[k]
        g.dx, g.dt = self.dx, self.dt
        self.flags = np.ascontiguousarray(self.flags, dtype=np.uint8)
        self.b_cell = np.ascontiguousarray(self.b_cell, dtype=np.uint32)
        self.b_start = np.ascontiguousarray(self.b_start, dtype=np.uint32)
        self.slots = np.ascontiguousarray(self.slots, dtype=np.float64).reshape(-1, 4)
        g.flags = self.flags.ctypes.data_as(C.POINTER(C.c_uint8))
        g.n_boundary = len(self.b_cell)
        g.b_cell = self.b_cell.ctypes.data_as(C.POINTER(C.c_uint32))
        g.b_start = self.b_start.ctypes.data_as(C.POINTER(C.c_uint32))
        g.slots = self.slots.ctypes.data_as(C.POINTER(WvSlot))
        return g

    def nearest_air(self, p):
        out = C.c_uint64()
        pp = np.ascontiguousarray(p, dtype=np.float64)
        st = lib().wv_nearest_air(C.byref(self.view()), _dp(pp), C.byref(out))
        return None if st != WV_OK else int(out.value)

    def baffle_point(self, p, axis):
        """fdtd.rs:497-512"""
        pp = np.ascontiguousarray(p, dtype=np.float64)
        ax = np.ascontiguousarray(axis, dtype=np.float64)
        out = np.zeros(3)
        lib().wv_baffle_point(C.byref(self.view()), _dp(pp), _dp(ax), _dp(out))
        return out

    def air_along(self, p, axis):
        """fdtd.rs:516-533: the first air node along axis from p, else the
        nearest; None when there is none."""
        out = C.c_uint64()
        pp = np.ascontiguousarray(p, dtype=np.float64)
        ax = np.ascontiguousarray(axis, dtype=np.float64)
        st = lib().wv_air_along(C.byref(self.view()), _dp(pp), _dp(ax), C.byref(out))
        return None if st != WV_OK else int(out.value)

    def taps_at(self, p):
        taps = (WvTap * 8)()
        n = C.c_size_t()
        pp = np.ascontiguousarray(p, dtype=np.float64)
        st = lib().wv_taps_at(C.byref(self.view()), _dp(pp), taps, C.byref(n))
        if st != WV_OK:
            return None
        return [(int(taps[t].cell), float(taps[t].w)) for t in range(n.value)]


# ------------------------------------------------------------ sources etc.

def source_signal(f_max, dt, steps):
    out = np.zeros(int(steps))
    lib().wv_source_signal(f_max, dt, int(steps), _dp(out))
    return out


def source_components(grid, c, order):
    comps = (WvTap * 8)()
    ln = (C.c_uint32 * 4)()
    nc = C.c_uint32()
    st = lib().wv_source_components(C.byref(grid.view()), int(c), int(order), comps, ln, C.byref(nc))
    if st != WV_OK:
        return None
    return [[(int(comps[2 * k + t].cell), float(comps[2 * k + t].w)) for t in range(ln[k])] for k in range(nc.value)]


def wave_source_taps(grid, center, axis, calibrated, order):
    """run.rs:601-624 (wv_wave_source_taps): a wave-job source's component
    taps and whether a directional source fell back to one node; None when
    the source is not inside the grid."""
    comps = (WvTap * 8)()
    ln = (C.c_uint32 * 4)()
    nc = C.c_uint32()
    fb = C.c_int32()
    pp = np.ascontiguousarray(center, dtype=np.float64)
    ax = np.ascontiguousarray(axis, dtype=np.float64)
    st = lib().wv_wave_source_taps(C.byref(grid.view()), _dp(pp), _dp(ax), int(calibrated), int(order), comps, ln,
                                   C.byref(nc), C.byref(fb))
    if st != WV_OK:
        return None
    if nc.value == 1:
        return [[(int(comps[t].cell), float(comps[t].w)) for t in range(ln[0])]], bool(fb.value)
    return [[(int(comps[2 * c + t].cell), float(comps[2 * c + t].w)) for t in range(ln[c])]
            for c in range(nc.value)], bool(fb.value)


def component_weights(a0, a, k, dx):
    a0v = np.array([a0.real, a0.imag])
    av = np.array([[v.real, v.imag] for v in a]).ravel()
    out = np.zeros(8)
    lib().wv_component_weights(_dp(a0v), _dp(av), k, dx, _dp(out))
    return out[0::2] + 1j * out[1::2]


def to_transfer(spectra, signal, grid, freqs, c, r_ref, h0):
    spec = np.asarray(spectra, dtype=np.complex128).ravel()
    sr = np.ascontiguousarray(np.stack([spec.real, spec.imag], axis=1).ravel())
    sig = np.ascontiguousarray(signal, dtype=np.float64)
    fr = np.ascontiguousarray(freqs, dtype=np.float64)
    h = np.asarray(h0, dtype=np.complex128)
    hr = np.ascontiguousarray(np.stack([h.real, h.imag], axis=1).ravel())
    out = np.zeros(2 * len(spec))
    lib().wv_to_transfer(_dp(sr), len(spec), _dp(sig), len(sig), grid.dt, grid.dx, _dp(fr), len(fr), c, r_ref,
                         _dp(hr), _dp(out))
    return out[0::2] + 1j * out[1::2]


def admittance(fit, f):
    arr, keep = _admittances([fit])
    out = np.zeros(2)
    lib().wv_admittance_eval(C.cast(arr, C.POINTER(WvAdmittance)), f, _dp(out))
    return complex(out[0], out[1])


def paris(beta, n=256):
    return lib().wv_paris(beta.real, beta.imag, n)


def fit_real_impedance(alpha):
    return lib().wv_fit_real_impedance(alpha)


# -------------------------------------------------------------------- runs

def _taps_array(taps):
    arr = (WvTap * max(1, len(taps)))()
    for t, (cell, w) in enumerate(taps):
        arr[t].cell = int(cell)
        arr[t].w = float(w)
    return arr


class Plan:
    """A grid resident on the device (wv_plan): many runs, one upload."""

    def __init__(self, grid: FdtdGrid, precision="fp32"):
        self.grid = grid
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self._h = C.c_void_p()
        err = C.create_string_buffer(1024)
        _check(lib().wv_plan_create(C.byref(grid.view()), self.precision, C.byref(self._h), err, 1024), err)
        self._steps = 0
        self._nrecv = 0

    def close(self):
        if self._h:
            lib().wv_plan_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_tile(self, zchunk=0, variant=-1):
        st = lib().wv_plan_set_tile(self._h, int(zchunk), int(variant))
        if st != WV_OK:
            raise FdtdError(st, "invalid tile settings")

    def set_pass(self, steps_per_pass=0, pair_zchunk=0):
        """Steps per pass (1, 2 or 3; 0 = default); takes effect at the next
        begin/run."""
        st = lib().wv_plan_set_pass(self._h, int(steps_per_pass), int(pair_zchunk))
        if st != WV_OK:
            raise FdtdError(st, "invalid pass settings")

    def pass_info(self):
        out = (C.c_uint32 * 4)()
        lib().wv_plan_pass_info(self._h, out)
        return {"steps_per_pass": int(out[0]), "pair_zchunk": int(out[1]), "ring_cells": int(out[2]),
                "ring2_cells": int(out[3])}

    def copy_probe(self, nbytes, reps=10):
        """Mean ms of a plain copy of nbytes between the field buffers (clobbers them)."""
        ms = C.c_double()
        st = lib().wv_plan_copy_probe(self._h, int(nbytes), int(reps), C.byref(ms))
        if st != WV_OK:
            raise FdtdError(st, "copy probe failed")
        return ms.value

    def stats(self):
        out = (C.c_uint64 * 7)()
        lib().wv_plan_stats(self._h, out)
        keys = ["cells", "air_cells", "boundary_nodes", "slots", "special_cells", "pitched_cells", "slots_with_k"]
        return dict(zip(keys, [int(v) for v in out]))

    def begin(self, taps, signal, receivers):
        sig = np.ascontiguousarray(signal, dtype=np.float64)
        rc = np.ascontiguousarray(receivers, dtype=np.uint64)
        tp = _taps_array(taps)
        err = C.create_string_buffer(1024)
        _check(lib().wv_run_begin(self._h, tp, len(taps), _dp(sig), len(sig),
                                  rc.ctypes.data_as(C.POINTER(C.c_uint64)), len(rc), err, 1024), err)
        self._steps = len(sig)
        self._nrecv = len(rc)

    def advance(self, k, cancel=None, progress=None):
        err = C.create_string_buffer(1024)
        cflag = None if cancel is None else cancel
        cb = PROGRESS_FN(lambda f, _ctx: progress(f)) if progress else PROGRESS_FN()
        st = lib().wv_run_advance(self._h, int(k), cflag.ctypes.data_as(C.POINTER(C.c_int32)) if cflag is not None
                                  else None, cb, None, err, 1024)
        return _check(st, err)

    def finish(self, freqs=(), mode=MODE_BOTH):
        fr = np.ascontiguousarray(freqs, dtype=np.float64)
        R, T, F = self._nrecv, self._steps, len(fr)
        ir = np.zeros((R, T)) if mode & MODE_IR else None
        spec = np.zeros(2 * R * F) if (mode & MODE_SPECTRA and F) else None
        err = C.create_string_buffer(1024)
        _check(lib().wv_run_finish(self._h, _dp(fr) if F else None, F, mode,
                                   _dp(spec) if spec is not None else None, _dp(ir) if ir is not None else None,
                                   err, 1024), err)
        sp = (spec[0::2] + 1j * spec[1::2]).reshape(R, F) if spec is not None else None
        return ir, sp

    def last_timing(self):
        ms = C.c_double()
        steps = C.c_uint64()
        launches = C.c_uint64()
        lib().wv_plan_last_timing(self._h, C.byref(ms), C.byref(steps), C.byref(launches))
        return ms.value, steps.value, launches.value

    def run(self, taps, signal, receivers, freqs=(), mode=MODE_BOTH, cancel=None, progress=None):
        """Returns (ir [R][T], spectra [R][F]) or None when cancelled."""
        self.begin(taps, signal, receivers)
        st = self.advance(len(signal), cancel=cancel, progress=progress)
        if st == WV_CANCELLED:
            return None
        return self.finish(freqs, mode)


class Gpu:
    """The device context, as the reference's ``Gpu`` (gpu.rs:104-107)."""

    def __init__(self, precision="fp32"):
        self.precision = precision
        lib()

    def run_fdtd(self, g, source, signal, receivers, freqs, cancel=None, progress=None):
        """gpu_fdtd.rs:45-54: spectra [R][F] complex, None when cancelled, FdtdError on Err."""
        with Plan(g, self.precision) as p:
            out = p.run(source, signal, receivers, freqs, MODE_SPECTRA, cancel, progress)
        return None if out is None else out[1]

    def run_ir(self, g, source, signal, receivers, cancel=None, progress=None):
        """run_cpu_ir (fdtd.rs:630-641) on the device: [R][T] pressure series."""
        with Plan(g, self.precision) as p:
            out = p.run(source, signal, receivers, (), MODE_IR, cancel, progress)
        return None if out is None else out[0]


def _one_call(fn, g, taps, signal, receivers, freqs, mode, precision, cancel, progress):
    sig = np.ascontiguousarray(signal, dtype=np.float64)
    rc = np.ascontiguousarray(receivers, dtype=np.uint64)
    fr = np.ascontiguousarray(freqs, dtype=np.float64)
    R, T, F = len(rc), len(sig), len(fr)
    ir = np.zeros((R, T))
    spec = np.zeros(2 * R * max(F, 1))
    err = C.create_string_buffer(1024)
    cb = PROGRESS_FN(lambda f, _ctx: progress(f)) if progress else PROGRESS_FN()
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    st = fn(C.byref(g.view()), _taps_array(taps), len(taps), _dp(sig), T,
            rc.ctypes.data_as(C.POINTER(C.c_uint64)), R, _dp(fr) if F else None, F, mode, prec,
            cancel.ctypes.data_as(C.POINTER(C.c_int32)) if cancel is not None else None, cb, None,
            _dp(spec), _dp(ir), err, 1024)
    _check(st, err)
    if st == WV_CANCELLED:
        return None
    return ir, (spec[0::2] + 1j * spec[1::2])[: R * F].reshape(R, F)


def fdtd_run(g, taps, signal, receivers, freqs=(), mode=MODE_BOTH, precision="fp32", cancel=None, progress=None):
    """The one-call C entry point wv_fdtd_run."""
    return _one_call(lib().wv_fdtd_run, g, taps, signal, receivers, freqs, mode, precision, cancel, progress)


def fdtd_run_noalloc(g, taps, signal, receivers, freqs=(), mode=MODE_BOTH, precision="fp32", cancel=None,
                     progress=None):
    """wv_fdtd_run_noalloc: the same run with no device allocation (every
    device buffer in the library's 64 MiB static array; larger runs raise
    FdtdError with WV_ERR_LIMIT)."""
    return _one_call(lib().wv_fdtd_run_noalloc, g, taps, signal, receivers, freqs, mode, precision, cancel,
                     progress)


def footprint_estimate(grid, precision, steps, n_recv, n_taps):
    """Device bytes a run of this grid will hold (wv_footprint_estimate)."""
    out = C.c_uint64()
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    st = lib().wv_footprint_estimate(C.byref(grid.view()), prec, int(steps), int(n_recv), int(n_taps),
                                     C.byref(out))
    if st != WV_OK:
        raise FdtdError(st, "invalid grid")
    return int(out.value)


class Batch:
    """Many runs of different grids advanced together (wv_batch_*): one launch
    per step over every instance, 128-step blocks replayed as HIP graphs.
    Each instance's results equal its plain run's (``Plan.run``)."""

    def __init__(self, precision="fp32"):
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self._h = C.c_void_p()
        err = C.create_string_buffer(1024)
        _check(lib().wv_batch_create(self.precision, C.byref(self._h), err, 1024), err)
        self._shape = {}

    def close(self):
        if self._h:
            lib().wv_batch_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, grid: FdtdGrid, taps, signal, receivers):
        """Uploads the grid and starts its run; returns the instance id."""
        sig = np.ascontiguousarray(signal, dtype=np.float64)
        rc = np.ascontiguousarray(receivers, dtype=np.uint64)
        iid = C.c_int32()
        err = C.create_string_buffer(1024)
        _check(lib().wv_batch_add(self._h, C.byref(grid.view()), _taps_array(taps), len(taps), _dp(sig), len(sig),
                                  rc.ctypes.data_as(C.POINTER(C.c_uint64)), len(rc), C.byref(iid), err, 1024), err)
        self._shape[iid.value] = (len(rc), len(sig))
        return iid.value

    def add_many(self, items, threads=8):
        """add() for [(grid, taps, signal, receivers)], set up on several host
        threads at once; returns the ids in the order of items."""
        from concurrent.futures import ThreadPoolExecutor

        with ThreadPoolExecutor(max_workers=max(1, int(threads))) as ex:
            return list(ex.map(lambda it: self.add(*it), items))

    def autotune(self, trial_steps=16):
        err = C.create_string_buffer(1024)
        _check(lib().wv_batch_autotune(self._h, int(trial_steps), err, 1024), err)

    def set_tile(self, variant=-1, zchunk=0):
        err = C.create_string_buffer(1024)
        _check(lib().wv_batch_set_tile(self._h, int(variant), int(zchunk), err, 1024), err)

    def run(self, graphs=True, cancel=None, progress=None):
        """Runs every instance to its end; WV_OK or WV_CANCELLED."""
        err = C.create_string_buffer(1024)
        cb = PROGRESS_FN(lambda f, _ctx: progress(f)) if progress else PROGRESS_FN()
        st = lib().wv_batch_run(self._h, 1 if graphs else 0,
                                cancel.ctypes.data_as(C.POINTER(C.c_int32)) if cancel is not None else None, cb, None,
                                err, 1024)
        return _check(st, err)

    def result(self, iid, freqs=(), mode=MODE_IR):
        R, T = self._shape[iid]
        fr = np.ascontiguousarray(freqs, dtype=np.float64)
        F = len(fr)
        ir = np.zeros((R, T)) if mode & MODE_IR else None
        spec = np.zeros(2 * R * F) if (mode & MODE_SPECTRA and F) else None
        err = C.create_string_buffer(1024)
        _check(lib().wv_batch_result(self._h, int(iid), _dp(fr) if F else None, F, mode,
                                     _dp(spec) if spec is not None else None, _dp(ir) if ir is not None else None,
                                     err, 1024), err)
        sp = (spec[0::2] + 1j * spec[1::2]).reshape(R, F) if spec is not None else None
        return ir, sp

    def info(self, iid):
        out = (C.c_uint64 * 6)()
        if lib().wv_batch_info(self._h, int(iid), out) != WV_OK:
            raise FdtdError(WV_ERR_INVALID, "unknown instance")
        keys = ["cells", "air_cells", "steps", "zchunk", "size_class", "device_bytes"]
        return dict(zip(keys, [int(v) for v in out]))

    def set_launch(self, form):
        """BATCH_STEPS (a launch per step) or BATCH_PERSISTENT (one launch per
        128-step block, each XCD running its share of the instances)."""
        err = C.create_string_buffer(1024)
        _check(lib().wv_batch_set_launch(self._h, int(form), err, 1024), err)

    def timing(self):
        ms = C.c_double()
        out = (C.c_uint64 * 7)()
        lib().wv_batch_timing(self._h, C.byref(ms), out)
        keys = ["steps", "cell_updates", "graph_launches", "graph_captures", "variant", "solo_blocks", "launch"]
        d = dict(zip(keys, [int(v) for v in out]))
        d["device_ms"] = ms.value
        return d


class Server:
    """The request loop (wv_serve_*): admission by device footprint under a
    cap, continuous batching at 128-step boundaries, results on host memory.
    Grids passed to ``submit`` must stay alive until ``wait`` returns; this
    class keeps a reference."""

    def __init__(self, precision="fp32", hbm_cap_bytes=288 * 10**9, policy="lanes", fill_cells=0, lane_cells=0):
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self._h = C.c_void_p()
        err = C.create_string_buffer(1024)
        _check(lib().wv_serve_create(self.precision, int(hbm_cap_bytes), C.byref(self._h), err, 1024), err)
        self._req = {}
        self.set_policy(policy, fill_cells, lane_cells)

    def set_policy(self, policy="lanes", fill_cells=0, lane_cells=0):
        """Which admitted requests each 128-step block advances: "all";
        "shortest" remaining work first up to fill_cells air cells; "lanes":
        as "shortest", except the large requests (lane_cells 0: plans with
        two or three steps per pass on grids of 2.25 Mi cells and more; else
        multi-step plans of lane_cells air cells and more), one at a time in admission order on a second stream
        (fill_cells 0 = the library's 16 Mi)."""
        code = {"all": SERVE_ALL, "shortest": SERVE_SHORTEST, "lanes": SERVE_LANES}[policy]
        if lib().wv_serve_set_policy(self._h, code, int(fill_cells), int(lane_cells)) != 0:
            raise FdtdError(WV_ERR_INVALID, "wv_serve_set_policy: invalid argument")

    def close(self):
        if self._h:
            lib().wv_serve_free(self._h)
            self._h = C.c_void_p()
            self._req.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, grid: FdtdGrid, taps, signal, receivers, freqs=(), mode=MODE_IR):
        sig = np.ascontiguousarray(signal, dtype=np.float64)
        rc = np.ascontiguousarray(receivers, dtype=np.uint64)
        fr = np.ascontiguousarray(freqs, dtype=np.float64)
        rid = C.c_uint64()
        err = C.create_string_buffer(1024)
        view = grid.view()
        _check(lib().wv_serve_submit(self._h, C.byref(view), _taps_array(taps), len(taps), _dp(sig), len(sig),
                                     rc.ctypes.data_as(C.POINTER(C.c_uint64)), len(rc), _dp(fr) if len(fr) else None,
                                     len(fr), mode, C.byref(rid), err, 1024), err)
        # The library borrows the grid view (and the arrays behind it) until wait.
        self._req[rid.value] = (grid, view, len(rc), len(sig), len(fr), mode)
        return rid.value

    def wait(self, rid):
        """(ir [R][T] or None, spectra [R][F] or None, latency s)."""
        grid, view, R, T, F, mode = self._req[rid]
        ir = np.zeros((R, T)) if mode & MODE_IR else None
        spec = np.zeros(2 * R * F) if (mode & MODE_SPECTRA and F) else None
        lat = C.c_double()
        err = C.create_string_buffer(1024)
        try:
            _check(lib().wv_serve_wait(self._h, int(rid), _dp(spec) if spec is not None else None,
                                       _dp(ir) if ir is not None else None, C.byref(lat), err, 1024), err)
        finally:
            del self._req[rid]
        sp = (spec[0::2] + 1j * spec[1::2]).reshape(R, F) if spec is not None else None
        return ir, sp, lat.value

    def stats(self):
        out = (C.c_double * 13)()
        lib().wv_serve_stats(self._h, out)
        keys = ["completed", "p50_s", "p95_s", "p99_s", "span_s", "cell_updates", "max_running",
                "peak_device_bytes", "failed", "graph_captures", "solo_blocks", "graph_launches", "lane_blocks"]
        return dict(zip(keys, [float(v) for v in out]))


def wave_job(grid: FdtdGrid, sources, signal, receivers, freqs, c, precision="fp32", batched=True, cancel=None):
    """The wave solver's part of a wave job (run.rs:593-661, wv_wave_job).

    sources: [{"comps": [taps, ...], "weights": [nf][n_comps] complex or None
    (one component), "h0": [nf] complex or None, "r_ref": m, "on_axis": [nf]
    complex calibration or None}]. Returns the
    transfer functions [n_sources][n_recv][nf] complex, or None when cancelled."""
    sig = np.ascontiguousarray(signal, dtype=np.float64)
    rc = np.ascontiguousarray(receivers, dtype=np.uint64)
    fr = np.ascontiguousarray(freqs, dtype=np.float64)
    R, F = len(rc), len(fr)
    keep = []
    arr = (WvWaveSource * max(1, len(sources)))()
    for k, src in enumerate(sources):
        comps = src["comps"]
        taps = [t for comp in comps for t in comp]
        off = np.zeros(len(comps) + 1, dtype=np.uint32)
        off[1:] = np.cumsum([len(comp) for comp in comps])
        tp = _taps_array(taps)
        keep += [tp, off]
        e = arr[k]
        e.taps = tp
        e.comp_off = off.ctypes.data_as(C.POINTER(C.c_uint32))
        e.n_comps = len(comps)
        if src.get("weights") is not None:
            w = np.ascontiguousarray(np.asarray(src["weights"], dtype=np.complex128).reshape(F, len(comps)))
            wv = np.ascontiguousarray(w.view(np.float64))
            keep.append(wv)
            e.weights = _dp(wv)
        if src.get("h0") is not None:
            h = np.ascontiguousarray(np.asarray(src["h0"], dtype=np.complex128).view(np.float64))
            keep.append(h)
            e.h0 = _dp(h)
        e.r_ref = float(src.get("r_ref", 1.0))
        if src.get("on_axis") is not None:
            oa = np.ascontiguousarray(np.asarray(src["on_axis"], dtype=np.complex128).view(np.float64))
            keep.append(oa)
            e.on_axis = _dp(oa)
    out = np.zeros(2 * len(sources) * R * F)
    err = C.create_string_buffer(1024)
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    st = lib().wv_wave_job(C.byref(grid.view()), arr, len(sources), _dp(sig), len(sig),
                           rc.ctypes.data_as(C.POINTER(C.c_uint64)), R, _dp(fr) if F else None, F, float(c), prec,
                           1 if batched else 0,
                           cancel.ctypes.data_as(C.POINTER(C.c_int32)) if cancel is not None else None, _dp(out),
                           err, 1024)
    _check(st, err)
    if st == WV_CANCELLED:
        return None
    return (out[0::2] + 1j * out[1::2]).reshape(len(sources), R, F)


ACCURACIES = {"draft": 0, "standard": 1, "high": 2}


def save_results(directory, freqs, speaker_ids, mics, transfer, stats=None, fingerprint=0, rt_sabine=None,
                 rt_eyring=None, fdtd_settings=None):
    """SimulationResult::save (result.rs:444-461) of a wave-solver result:
    mics = [{"id", "name", "position"}], transfer [mic][speaker][freq]
    complex, stats = {"backend", "seconds", "images", "paths",
    "diffracted_paths", "receivers", "notes"}, fdtd_settings =
    {"crossover_hz", "duration_s", "points_per_wavelength", "source_order",
    "cabinets", "accuracy"} or None."""
    fr = np.ascontiguousarray(freqs, dtype=np.float64)
    ids = np.ascontiguousarray(speaker_ids, dtype=np.uint64)
    tr = np.ascontiguousarray(np.asarray(transfer, dtype=np.complex128).reshape(len(mics), len(ids), len(fr)))
    trv = np.ascontiguousarray(tr.view(np.float64))
    st = dict(stats or {})
    r = WvResults()
    r.nf, r.freqs = len(fr), _dp(fr)
    r.n_speakers, r.speaker_ids = len(ids), ids.ctypes.data_as(C.POINTER(C.c_uint64))
    marr = (WvMicrophone * max(1, len(mics)))()
    names = [str(m["name"]).encode() for m in mics]
    for k, m in enumerate(mics):
        marr[k].id = int(m["id"])
        marr[k].name = names[k]
        for i in range(3):
            marr[k].position[i] = float(m["position"][i])
    r.n_mics, r.mics = len(mics), C.cast(marr, C.POINTER(WvMicrophone))
    r.transfer = _dp(trv)
    r.backend = str(st.get("backend", "")).encode()
    r.seconds = float(st.get("seconds", 0.0))
    for k in ("images", "paths", "diffracted_paths", "receivers"):
        setattr(r, k, int(st.get(k, 0)))
    notes = [str(n).encode() for n in st.get("notes", [])]
    narr = (C.c_char_p * max(1, len(notes)))(*notes)
    r.n_notes, r.notes = len(notes), C.cast(narr, C.POINTER(C.c_char_p))
    r.fingerprint = int(fingerprint)
    keep = []
    for name, v in (("rt_sabine", rt_sabine), ("rt_eyring", rt_eyring)):
        if v is not None:
            a = np.ascontiguousarray(v, dtype=np.float64)
            keep.append(a)
            setattr(r, name, _dp(a))
    if fdtd_settings is not None:
        fs = WvFdtdSettings(float(fdtd_settings.get("crossover_hz", 250.0)),
                            float(fdtd_settings.get("duration_s", 0.0)),
                            float(fdtd_settings.get("points_per_wavelength", 15.0)),
                            int(fdtd_settings.get("source_order", 1)), int(bool(fdtd_settings.get("cabinets", True))),
                            ACCURACIES[fdtd_settings.get("accuracy", "high")])
        keep.append(fs)
        r.fdtd = C.pointer(fs)
    err = C.create_string_buffer(1024)
    _check(lib().wv_results_save(str(directory).encode(), C.byref(r), err, 1024), err)


def read_results_transfer(directory, n_mics, n_speakers, nf):
    """result.rs:463-476: [mic][speaker][freq] complex64 from results.bin."""
    out = np.zeros((n_mics, n_speakers, nf, 2), dtype=np.float32)
    err = C.create_string_buffer(1024)
    _check(lib().wv_results_read_transfer(str(directory).encode(), n_mics, n_speakers, nf,
                                          out.ctypes.data_as(C.POINTER(C.c_float)), err, 1024), err)
    return out.view(np.complex64).reshape(n_mics, n_speakers, nf)


def blend(ism, wave, w):
    """fdtd.rs:847-862: power-complementary crossfade (complex scalars)."""
    a = np.array([complex(ism).real, complex(ism).imag])
    b = np.array([complex(wave).real, complex(wave).imag])
    out = np.zeros(2)
    lib().wv_blend(_dp(a), _dp(b), float(w), _dp(out))
    return complex(out[0], out[1])


def calibrate_cabinet(verts, tris, ac, c, f_max, ppw, cells_across, freqs, precision="fp32", batched=True,
                      cancel=None):
    """CabinetCalibration of a cabinet mesh (fdtd.rs:873-955, wv_calibrate_cabinet):
    {"on_axis": [nf] complex, "angles_deg": [12], "directivity": [12][nf] complex},
    or None when cancelled."""
    v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.uint32).reshape(-1, 3))
    a = np.ascontiguousarray(np.asarray(ac, dtype=np.float64).reshape(3))
    fr = np.ascontiguousarray(freqs, dtype=np.float64)
    F = len(fr)
    on = np.zeros(2 * max(F, 1))
    ang = np.zeros(12)
    dr = np.zeros(2 * 12 * max(F, 1))
    err = C.create_string_buffer(1024)
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    st = lib().wv_calibrate_cabinet(_dp(v), len(v), t.ctypes.data_as(C.POINTER(C.c_uint32)), len(t), _dp(a),
                                    float(c), float(f_max), float(ppw), float(cells_across), _dp(fr) if F else None,
                                    F, prec, 1 if batched else 0,
                                    cancel.ctypes.data_as(C.POINTER(C.c_int32)) if cancel is not None else None,
                                    _dp(on), _dp(ang), _dp(dr), err, 1024)
    if st == WV_CANCELLED:
        return None
    _check(st, err)
    return {"on_axis": on[:2 * F].view(np.complex128).copy(), "angles_deg": ang,
            "directivity": dr[:2 * 12 * F].view(np.complex128).reshape(12, F).copy()}


def _spectrum(spec, keep):
    """A Spectrum (spectrum.rs:6-44): a number, or {"freqs_hz": [...], "values": [...]}."""
    sp = WvSpectrum()
    if isinstance(spec, dict):
        f = np.ascontiguousarray(spec["freqs_hz"], dtype=np.float64)
        v = np.ascontiguousarray(spec["values"], dtype=np.float64)
        sp.n = len(f)
    else:
        f = np.zeros(1)
        v = np.array([float(spec)])
        sp.n = 0
    keep += [f, v]
    sp.freqs_hz, sp.values = _dp(f), _dp(v)
    return sp


def _material(mat, keep):
    m = WvMaterial()
    m.absorption = _spectrum(mat.get("absorption", 0.0), keep)
    if mat.get("impedance") is not None:
        re, im = (_spectrum(x, keep) for x in mat["impedance"])
        keep += [re, im]
        m.impedance_re, m.impedance_im = C.pointer(re), C.pointer(im)
    return m


def material_absorption(mat, f):
    """Material::absorption_at (material.rs:72-78) for a material dict with the
    library's fields ("absorption", optional "impedance": (re, im) spectra)."""
    keep = []
    m = _material(mat, keep)
    return float(lib().wv_material_absorption(C.byref(m), float(f)))


def fit_admittance(mat, band, max_branches=2):
    """fit_admittance (impedance_fit.rs:158-235): ((beta0, [(r, m, k), ...]),
    {"band": (lo, hi), "rms_error": e, "max_error": e})."""
    keep = []
    m = _material(mat, keep)
    out = WvFitResult()
    st = lib().wv_fit_admittance(C.byref(m), float(band[0]), float(band[1]), int(max_branches), C.byref(out))
    if st != WV_OK:
        raise FdtdError(st, "wv_fit_admittance: invalid argument")
    fit = (out.beta0, [(out.branches[b].r, out.branches[b].m, out.branches[b].k) for b in range(out.n_branches)])
    return fit, {"band": (out.band[0], out.band[1]), "rms_error": out.rms_error, "max_error": out.max_error}


def fit_surfaces(materials, band, threads=8):
    """fit_surfaces (fdtd.rs:213-222): one fit per distinct material id, fitted
    in parallel, returned per surface as (beta0, branches, {"band",
    "rms_error", "max_error"}), the whole AdmittanceFit."""
    from concurrent.futures import ThreadPoolExecutor

    unique = {}
    for m in materials:
        unique.setdefault(m["id"], m)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        def whole(m):
            fit, meta = fit_admittance(m, band, 2)
            return (fit[0], fit[1], meta)

        fits = dict(zip(unique, ex.map(whole, unique.values())))
    return [fits[m["id"]] for m in materials]


def first_order_directions():
    """fdtd.rs:753-760: the 400 local directions fit_first_order samples."""
    out = np.zeros(1200)
    lib().wv_first_order_directions(_dp(out))
    return out.reshape(400, 3)


def fit_first_order(d, d_axis, to_room):
    """fit_first_order (fdtd.rs:749-807): d = the directivity at
    first_order_directions() (complex, [400]), d_axis on axis; returns
    (a0, [a_x, a_y, a_z], relative RMS residual)."""
    dv = np.ascontiguousarray(np.asarray(d, dtype=np.complex128).reshape(400).view(np.float64))
    ax = np.array([complex(d_axis).real, complex(d_axis).imag])
    R = np.ascontiguousarray(np.asarray(to_room, dtype=np.float64).reshape(3, 3))
    a0, a = np.zeros(2), np.zeros(6)
    res = lib().wv_fit_first_order(_dp(dv), _dp(ax), _dp(R), _dp(a0), _dp(a))
    return complex(a0[0], a0[1]), [complex(a[2 * i], a[2 * i + 1]) for i in range(3)], float(res)


ACCURACY = {"draft": 0, "standard": 1, "high": 2}


def crossover_weight(crossover_hz, f):
    """FdtdSettings::weight (fdtd.rs:166-175)."""
    return float(lib().wv_crossover_weight(float(crossover_hz), float(f)))


def grid_fmax(crossover_hz, freqs, accuracy="high"):
    """FdtdSettings::grid_fmax (fdtd.rs:148-154)."""
    fr = np.ascontiguousarray(freqs, dtype=np.float64)
    return float(lib().wv_grid_fmax(float(crossover_hz), ACCURACY[accuracy], _dp(fr), len(fr)))


def phase_error_deg_per_m(f, c, dx):
    """FdtdSettings::phase_error_deg_per_m (fdtd.rs:157-162)."""
    return float(lib().wv_phase_error_deg_per_m(float(f), float(c), float(dx)))
