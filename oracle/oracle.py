"""ctypes binding of the FP64 C oracle (fdtd_oracle.c) -- test infrastructure.

Room specs are plain data, the same dicts the product's grid builder takes:

    {"L": (lx, ly, lz),
     "pillars": [(x0, x1, y0, y1), ...],          # floor-to-ceiling boxes
     "fits": [(beta0, [(r, m, k), ...]), ...],    # AdmittanceFit values
     "patches": [(gu, gv, [fit index per patch, row-major iv*gu+iu]), ...]}

with one ``patches`` entry per surface: walls x=0, x=L0, y=0, y=L1, floor,
ceiling, then the four sides (facing -x, +x, -y, +y) of each pillar.

Pinning (the C file's header has the detail): every function is checked by
the tests; three parts are partly "parity unpinned": the exact-tie triangle
choice of the general-room builder, the exact Nelder-Mead trajectory of the
admittance fit and the composed outputs of the cabinet calibration.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_fdtd.so")
MAX_BRANCHES = 8
COURANT = 0.5773502691896257


class ObBranch(C.Structure):
    _fields_ = [("r", C.c_double), ("m", C.c_double), ("k", C.c_double)]


class ObFit(C.Structure):
    _fields_ = [("beta0", C.c_double), ("nb", C.c_int32), ("br", ObBranch * MAX_BRANCHES),
                ("band", C.c_double * 2), ("rms_error", C.c_double), ("max_error", C.c_double)]


class ObSlot(C.Structure):
    _fields_ = [("inv_d", C.c_double), ("cv", C.c_double), ("k", C.c_double), ("w", C.c_double)]


class ObRoom(C.Structure):
    _fields_ = [("L", C.c_double * 3), ("n_pillars", C.c_int32), ("pillars", C.POINTER(C.c_double)),
                ("n_fits", C.c_int32), ("fits", C.POINTER(ObFit)), ("gu", C.POINTER(C.c_int32)),
                ("gv", C.POINTER(C.c_int32)), ("patch_off", C.POINTER(C.c_int32)),
                ("patch_fit", C.POINTER(C.c_int32))]


class ObGrid(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("dx", C.c_double), ("dt", C.c_double), ("n", C.c_int64 * 3),
                ("flags", C.POINTER(C.c_uint8)), ("bidx", C.POINTER(C.c_uint32)), ("nb", C.c_int64),
                ("b_cell", C.POINTER(C.c_uint32)), ("b_start", C.POINTER(C.c_uint32)), ("ns", C.c_int64),
                ("slots", C.POINTER(ObSlot))]


class ObMeshRoom(C.Structure):
    _fields_ = [("prism", C.c_int32), ("height", C.c_double), ("n_pos", C.c_int32), ("pos", C.POINTER(C.c_double)),
                ("n_tri", C.c_int32), ("tri", C.POINTER(C.c_int32)), ("tri_surface", C.POINTER(C.c_int32)),
                ("n_fits", C.c_int32), ("fits", C.POINTER(ObFit)), ("n_regions", C.c_int32),
                ("region_cycles", C.POINTER(C.c_int32)), ("cycle_off", C.POINTER(C.c_int32)),
                ("cycle_xy", C.POINTER(C.c_double))]


class ObSolid(C.Structure):
    _fields_ = [("n_planes", C.c_int32), ("planes", C.POINTER(C.c_double)), ("lo", C.c_double * 3),
                ("hi", C.c_double * 3)]


class ObSpectrum(C.Structure):
    _fields_ = [("n", C.c_int32), ("f", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double))]


class ObMaterial(C.Structure):
    _fields_ = [("absorption", ObSpectrum), ("has_impedance", C.c_int32), ("imp_re", ObSpectrum),
                ("imp_im", ObSpectrum)]


class ObGridView(C.Structure):
    _fields_ = [("n", C.c_int64 * 3), ("dx", C.c_double), ("dt", C.c_double), ("flags", C.POINTER(C.c_uint8)),
                ("nb", C.c_int64), ("b_cell", C.POINTER(C.c_uint32)), ("b_start", C.POINTER(C.c_uint32)),
                ("slots", C.POINTER(ObSlot))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        L.ob_build_box.argtypes = [P(ObRoom), C.c_double, C.c_double, P(ObGrid)]
        L.ob_build_box.restype = C.c_int
        L.ob_grid_free.argtypes = [P(ObGrid)]
        L.ob_boundary_update.argtypes = [P(ObSlot), C.c_int, P(C.c_double), C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double]
        L.ob_boundary_update.restype = C.c_double
        L.ob_boundary_slots.argtypes = [P(C.c_int32), C.c_int, P(ObFit), C.c_double, P(ObSlot), C.c_int]
        L.ob_boundary_slots.restype = C.c_int
        L.ob_source_signal.argtypes = [C.c_double, C.c_double, C.c_int64, P(C.c_double)]
        L.ob_nearest_air.argtypes = [P(ObGrid), P(C.c_double)]
        L.ob_nearest_air.restype = C.c_int64
        L.ob_taps_at.argtypes = [P(ObGrid), P(C.c_double), P(C.c_int64), P(C.c_double)]
        L.ob_taps_at.restype = C.c_int
        L.ob_source_components.argtypes = [P(ObGrid), C.c_int64, C.c_int, P(C.c_int64), P(C.c_double), P(C.c_int)]
        L.ob_source_components.restype = C.c_int
        L.ob_component_weights.argtypes = [P(C.c_double), P(C.c_double), C.c_double, C.c_double, P(C.c_double)]
        L.ob_to_transfer.argtypes = [P(C.c_double), C.c_int64, P(C.c_double), C.c_int64, C.c_double, C.c_double,
                                     P(C.c_double), C.c_int64, C.c_double, C.c_double, P(C.c_double), P(C.c_double)]
        L.ob_phase_error_deg_per_m.argtypes = [C.c_double, C.c_double, C.c_double]
        L.ob_phase_error_deg_per_m.restype = C.c_double
        L.ob_admittance.argtypes = [P(ObFit), C.c_double, P(C.c_double)]
        L.ob_paris.argtypes = [C.c_double, C.c_double, C.c_int]
        L.ob_paris.restype = C.c_double
        L.ob_fit_real_impedance.argtypes = [C.c_double]
        L.ob_fit_real_impedance.restype = C.c_double
        L.ob_run.argtypes = [P(ObGridView), P(C.c_int64), P(C.c_double), C.c_int64, P(C.c_double), C.c_int64,
                             P(C.c_int64), C.c_int64, P(C.c_double), C.c_int64, P(C.c_double), P(C.c_double),
                             C.c_int, P(C.c_int32)]
        L.ob_run.restype = C.c_int
        L.ob_run_fields.argtypes = [P(ObGridView), P(C.c_int64), P(C.c_double), C.c_int64, P(C.c_double), C.c_int64,
                                    C.c_int, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.ob_run_fields.restype = C.c_int
        L.ob_ray_triangle.argtypes = [P(C.c_double)] * 5 + [P(C.c_double)]
        L.ob_ray_triangle.restype = C.c_int
        L.ob_parity_inside.argtypes = [P(ObMeshRoom), P(C.c_double)]
        L.ob_parity_inside.restype = C.c_int
        L.ob_solid_from_mesh.argtypes = [P(C.c_double), C.c_int32, P(C.c_int32), C.c_int32, P(C.c_double),
                                         P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double)]
        L.ob_solid_from_mesh.restype = C.c_int
        L.ob_solid_contains.argtypes = [P(ObSolid), P(C.c_double)]
        L.ob_solid_contains.restype = C.c_int
        L.ob_build_room.argtypes = [P(ObMeshRoom), P(ObSolid), C.c_int32, C.c_double, C.c_double, P(ObGrid)]
        L.ob_build_room.restype = C.c_int
        L.ob_baffle_point.argtypes = [P(ObGrid), P(C.c_double), P(C.c_double), P(C.c_double)]
        L.ob_baffle_point.restype = None
        L.ob_air_along.argtypes = [P(ObGrid), P(C.c_double), P(C.c_double)]
        L.ob_air_along.restype = C.c_int64
        L.ob_wave_source_taps.argtypes = [P(ObGrid), P(C.c_double), P(C.c_double), C.c_int, C.c_int, P(C.c_int64),
                                          P(C.c_double), P(C.c_int), P(C.c_int)]
        L.ob_wave_source_taps.restype = C.c_int
        L.ob_blend.argtypes = [P(C.c_double), P(C.c_double), C.c_double, P(C.c_double)]
        L.ob_blend.restype = None
        L.ob_calibrate_cabinet.argtypes = [P(C.c_double), C.c_int32, P(C.c_int32), C.c_int32, P(C.c_double), C.c_double,
                                           C.c_double, C.c_double, C.c_double, P(C.c_double), C.c_int32, C.c_int,
                                           P(C.c_double), P(C.c_double), P(C.c_double)]
        L.ob_calibrate_cabinet.restype = C.c_int
        L.ob_first_order_directions.argtypes = [P(C.c_double)]
        L.ob_first_order_directions.restype = None
        L.ob_fit_first_order.argtypes = [P(C.c_double)] * 5
        L.ob_fit_first_order.restype = C.c_double
        L.ob_spectrum_eval.argtypes = [P(C.c_double), P(C.c_double), C.c_int32, C.c_double]
        L.ob_spectrum_eval.restype = C.c_double
        L.ob_material_absorption.argtypes = [P(ObMaterial), C.c_double]
        L.ob_material_absorption.restype = C.c_double
        L.ob_fit_admittance.argtypes = [P(ObMaterial), C.c_double, C.c_double, C.c_int, P(ObFit), P(C.c_double),
                                        P(C.c_double)]
        L.ob_fit_admittance.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def make_fit(beta0, branches, meta=None):
    """An AdmittanceFit; meta = {"band": (lo, hi), "rms_error", "max_error"}
    (zeros when absent) takes part in the fit's equality."""
    f = ObFit()
    f.beta0 = float(beta0)
    f.nb = len(branches)
    if f.nb > MAX_BRANCHES:
        raise ValueError("too many branches")
    for i, (r, m, k) in enumerate(branches):
        f.br[i].r, f.br[i].m, f.br[i].k = float(r), float(m), float(k)
    if meta:
        f.band[0], f.band[1] = (float(v) for v in meta.get("band", (0.0, 0.0)))
        f.rms_error = float(meta.get("rms_error", 0.0))
        f.max_error = float(meta.get("max_error", 0.0))
    return f


@dataclass
class Grid:
    """FdtdGrid (fdtd.rs:196-209) as numpy arrays, plus the ctypes owner."""
    n: tuple
    origin: tuple
    dx: float
    dt: float
    flags: np.ndarray
    b_cell: np.ndarray
    b_start: np.ndarray
    slots: np.ndarray  # (S, 4) inv_d, cv, k, w
    _owner: object = None

    @property
    def cells(self):
        return self.n[0] * self.n[1] * self.n[2]

    def index(self, x, y, z):
        return (z * self.n[1] + y) * self.n[0] + x

    def center(self, i):
        nx, ny = self.n[0], self.n[1]
        z, r = divmod(int(i), nx * ny)
        y, x = divmod(r, nx)
        return np.array([self.origin[0] + x * self.dx, self.origin[1] + y * self.dx, self.origin[2] + z * self.dx])

    def _obgrid(self):
        g = ObGrid()
        for k in range(3):
            g.origin[k] = self.origin[k]
            g.n[k] = self.n[k]
        g.dx, g.dt = self.dx, self.dt
        g.flags = _ptr(self.flags, C.c_uint8)
        return g

    def view(self):
        v = ObGridView()
        for k in range(3):
            v.n[k] = self.n[k]
        v.dx, v.dt = self.dx, self.dt
        v.flags = _ptr(self.flags, C.c_uint8)
        v.nb = len(self.b_cell)
        v.b_cell = _ptr(self.b_cell, C.c_uint32)
        v.b_start = _ptr(self.b_start, C.c_uint32)
        v.slots = self.slots.ctypes.data_as(C.POINTER(ObSlot))
        return v


def _room_struct(room):
    keep = []
    r = ObRoom()
    for k in range(3):
        r.L[k] = float(room["L"][k])
    pil = np.ascontiguousarray(np.array(room.get("pillars", []), dtype=np.float64).reshape(-1, 4))
    fits = (ObFit * len(room["fits"]))(*[make_fit(*f) for f in room["fits"]])
    patches = room["patches"]
    n_surf = 6 + 4 * len(pil)
    if len(patches) != n_surf:
        raise ValueError(f"room needs {n_surf} surface patch entries, got {len(patches)}")
    gu = np.array([p[0] for p in patches], dtype=np.int32)
    gv = np.array([p[1] for p in patches], dtype=np.int32)
    off = np.zeros(n_surf, dtype=np.int32)
    off[1:] = np.cumsum(gu * gv)[:-1]
    pf = np.concatenate([np.asarray(p[2], dtype=np.int32).ravel() for p in patches]).astype(np.int32)
    keep += [pil, fits, gu, gv, off, pf]
    r.n_pillars = len(pil)
    r.pillars = _ptr(pil, C.c_double)
    r.n_fits = len(room["fits"])
    r.fits = fits
    r.gu, r.gv, r.patch_off, r.patch_fit = (_ptr(gu, C.c_int32), _ptr(gv, C.c_int32), _ptr(off, C.c_int32),
                                            _ptr(pf, C.c_int32))
    return r, keep


def _grid_from(g):
    n = tuple(int(g.n[k]) for k in range(3))
    N = n[0] * n[1] * n[2]
    flags = np.ctypeslib.as_array(g.flags, shape=(N,)).copy()
    nb = int(g.nb)
    b_cell = np.ctypeslib.as_array(g.b_cell, shape=(nb,)).copy() if nb else np.zeros(0, np.uint32)
    b_start = np.ctypeslib.as_array(g.b_start, shape=(nb + 1,)).copy()
    ns = int(g.ns)
    if ns:
        raw = np.ctypeslib.as_array(C.cast(g.slots, C.POINTER(C.c_double)), shape=(ns * 4,)).copy()
        slots = raw.reshape(ns, 4)
    else:
        slots = np.zeros((0, 4))
    return Grid(n=n, origin=tuple(g.origin[k] for k in range(3)), dx=g.dx, dt=g.dt, flags=flags,
                b_cell=b_cell.astype(np.uint32), b_start=b_start.astype(np.uint32),
                slots=np.ascontiguousarray(slots))


def build_box(room, c, dx):
    """fdtd.rs:289-377 (build_dx) for a box room with pillars."""
    L = lib()
    r, keep = _room_struct(room)
    g = ObGrid()
    st = L.ob_build_box(C.byref(r), c, dx, C.byref(g))
    if st != 0:
        raise RuntimeError(f"ob_build_box failed ({st})")
    try:
        return _grid_from(g)
    finally:
        L.ob_grid_free(C.byref(g))


# ------------------------------------------------------------ general rooms
#
# A mesh room is a dict
#     {"positions": (V, 3), "triangles": (T, 3), "tri_surface": (T,),
#      "fits": [(beta0, [(r, m, k), ...]), ...],       # one per surface
#      "prism": bool, "height": h,                     # extruded rooms
#      "regions": [[outer (n, 2), hole (m, 2), ...], ...]}
# and a solid {"planes": (P, 4), "lo": (3,), "hi": (3,)}.

def _mesh_struct(mesh):
    r = ObMeshRoom()
    pos = np.ascontiguousarray(np.asarray(mesh["positions"], dtype=np.float64).reshape(-1, 3))
    tri = np.ascontiguousarray(np.asarray(mesh["triangles"], dtype=np.int32).reshape(-1, 3))
    ts = np.ascontiguousarray(np.asarray(mesh["tri_surface"], dtype=np.int32))
    fits = (ObFit * max(1, len(mesh["fits"])))(*[make_fit(*f) for f in mesh["fits"]])
    regions = mesh.get("regions", []) if mesh.get("prism") else []
    rc, off, xy = [0], [0], []
    for reg in regions:
        for cyc in reg:
            cyc = np.asarray(cyc, dtype=np.float64).reshape(-1, 2)
            xy.append(cyc)
            off.append(off[-1] + len(cyc))
        rc.append(len(off) - 1)
    rc = np.array(rc, dtype=np.int32)
    off = np.array(off, dtype=np.int32)
    xy = np.ascontiguousarray(np.concatenate(xy) if xy else np.zeros((1, 2)))
    r.prism = int(bool(mesh.get("prism")))
    r.height = float(mesh.get("height", 0.0))
    r.n_pos, r.pos = len(pos), _ptr(pos, C.c_double)
    r.n_tri, r.tri, r.tri_surface = len(tri), _ptr(tri, C.c_int32), _ptr(ts, C.c_int32)
    r.n_fits, r.fits = len(mesh["fits"]), fits
    r.n_regions = len(regions)
    r.region_cycles, r.cycle_off, r.cycle_xy = _ptr(rc, C.c_int32), _ptr(off, C.c_int32), _ptr(xy, C.c_double)
    return r, [pos, tri, ts, fits, rc, off, xy]


def _solids_struct(solids):
    arr = (ObSolid * max(1, len(solids)))()
    keep = []
    for i, s in enumerate(solids):
        pl = np.ascontiguousarray(np.asarray(s["planes"], dtype=np.float64).reshape(-1, 4))
        keep.append(pl)
        arr[i].n_planes = len(pl)
        arr[i].planes = _ptr(pl, C.c_double)
        for k in range(3):
            arr[i].lo[k], arr[i].hi[k] = float(s["lo"][k]), float(s["hi"][k])
    return arr, keep


def build_room(mesh, c, dx, solids=()):
    """fdtd.rs:289-377 (build_dx) for a general room, column_fill (:385-422)
    for meshes, rigid convex solids (:424-461)."""
    L = lib()
    r, keep = _mesh_struct(mesh)
    sa, keep2 = _solids_struct(solids)
    g = ObGrid()
    st = L.ob_build_room(C.byref(r), sa, len(solids), c, dx, C.byref(g))
    if st != 0:
        L.ob_grid_free(C.byref(g))
        raise RuntimeError(f"ob_build_room failed ({st})")
    try:
        return _grid_from(g)
    finally:
        L.ob_grid_free(C.byref(g))


def parity_inside(mesh, points):
    """room.rs:355-367 (Room::contains for meshes) at each of ``points`` (N, 3)."""
    r, keep = _mesh_struct(mesh)
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    L = lib()
    base = pts.ctypes.data
    return np.array([bool(L.ob_parity_inside(C.byref(r), C.cast(base + 24 * i, C.POINTER(C.c_double))))
                     for i in range(len(pts))], dtype=bool)


def ray_triangle(o, d, tri):
    """bvh.rs:209-226; the hit's t or None."""
    a = [np.ascontiguousarray(np.asarray(v, dtype=np.float64)) for v in (o, d, tri[0], tri[1], tri[2])]
    t = np.zeros(1)
    ok = lib().ob_ray_triangle(*[_ptr(v, C.c_double) for v in a], _ptr(t, C.c_double))
    return float(t[0]) if ok else None


def solid_from_mesh(verts, tris, rot=None, pos=(0.0, 0.0, 0.0)):
    """fdtd.rs:433-456 (Solid::from_mesh); rot row-major, applied as rot @ v."""
    v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.int32).reshape(-1, 3))
    R = np.ascontiguousarray(np.eye(3) if rot is None else np.asarray(rot, dtype=np.float64).reshape(3, 3))
    P = np.ascontiguousarray(np.asarray(pos, dtype=np.float64))
    planes = np.zeros((max(1, len(t)), 4))
    lo, hi = np.zeros(3), np.zeros(3)
    n = lib().ob_solid_from_mesh(_ptr(v, C.c_double), len(v), _ptr(t, C.c_int32), len(t), _ptr(R, C.c_double),
                                 _ptr(P, C.c_double), _ptr(planes, C.c_double), _ptr(lo, C.c_double),
                                 _ptr(hi, C.c_double))
    if n < 0:
        return None
    return {"planes": planes[:n].copy(), "lo": lo, "hi": hi}


def solid_contains(solid, p):
    """fdtd.rs:458-460"""
    sa, keep = _solids_struct([solid])
    q = np.ascontiguousarray(np.asarray(p, dtype=np.float64))
    return bool(lib().ob_solid_contains(sa, _ptr(q, C.c_double)))


def source_signal(f_max, dt, steps):
    out = np.zeros(steps)
    lib().ob_source_signal(f_max, dt, steps, _ptr(out, C.c_double))
    return out


def nearest_air(grid, p):
    g = grid._obgrid()
    pp = np.asarray(p, dtype=np.float64)
    i = lib().ob_nearest_air(C.byref(g), _ptr(pp, C.c_double))
    return None if i < 0 else int(i)


def taps_at(grid, p):
    g = grid._obgrid()
    pp = np.asarray(p, dtype=np.float64)
    cells = np.zeros(8, dtype=np.int64)
    w = np.zeros(8)
    nt = lib().ob_taps_at(C.byref(g), _ptr(pp, C.c_double), _ptr(cells, C.c_int64), _ptr(w, C.c_double))
    if nt == 0:
        return None
    return [(int(cells[t]), float(w[t])) for t in range(nt)]


def source_components(grid, c, order):
    g = grid._obgrid()
    cc = np.zeros(8, dtype=np.int64)
    cw = np.zeros(8)
    ln = np.zeros(4, dtype=np.int32)
    nc = lib().ob_source_components(C.byref(g), int(c), int(order), _ptr(cc, C.c_int64), _ptr(cw, C.c_double),
                                    _ptr(ln, C.c_int))
    if nc == 0:
        return None
    return [[(int(cc[2 * k + t]), float(cw[2 * k + t])) for t in range(ln[k])] for k in range(nc)]


def air_along(grid, p, axis):
    """fdtd.rs:516-533; None when no air node is found."""
    g = grid._obgrid()
    pp = np.ascontiguousarray(p, dtype=np.float64)
    ax = np.ascontiguousarray(axis, dtype=np.float64)
    i = lib().ob_air_along(C.byref(g), _ptr(pp, C.c_double), _ptr(ax, C.c_double))
    return None if i < 0 else int(i)


def wave_source_taps(grid, center, axis, calibrated, order):
    """run.rs:601-624: (component taps, omni fallback) or None."""
    g = grid._obgrid()
    pp = np.ascontiguousarray(center, dtype=np.float64)
    ax = np.ascontiguousarray(axis, dtype=np.float64)
    cc = np.zeros(8, dtype=np.int64)
    cw = np.zeros(8)
    ln = np.zeros(4, dtype=np.int32)
    fb = np.zeros(1, dtype=np.int32)
    nc = lib().ob_wave_source_taps(C.byref(g), _ptr(pp, C.c_double), _ptr(ax, C.c_double), int(calibrated),
                                   int(order), _ptr(cc, C.c_int64), _ptr(cw, C.c_double), _ptr(ln, C.c_int),
                                   _ptr(fb, C.c_int))
    if nc == 0:
        return None
    if nc == 1:
        return [[(int(cc[t]), float(cw[t])) for t in range(ln[0])]], bool(fb[0])
    return [[(int(cc[2 * k + t]), float(cw[2 * k + t])) for t in range(ln[k])] for k in range(nc)], bool(fb[0])


def component_weights(a0, a, k, dx):
    a0v = np.array([a0.real, a0.imag])
    av = np.array([[v.real, v.imag] for v in a]).ravel()
    out = np.zeros(8)
    lib().ob_component_weights(_ptr(a0v, C.c_double), _ptr(av, C.c_double), k, dx, _ptr(out, C.c_double))
    return out[0::2] + 1j * out[1::2]


def to_transfer(spectra, signal, grid, freqs, c, r_ref, h0):
    spec = np.ascontiguousarray(np.asarray(spectra, dtype=np.complex128).ravel())
    sr = np.ascontiguousarray(np.stack([spec.real, spec.imag], axis=1).ravel())
    sig = np.ascontiguousarray(signal, dtype=np.float64)
    fr = np.ascontiguousarray(freqs, dtype=np.float64)
    h = np.asarray(h0, dtype=np.complex128)
    hr = np.ascontiguousarray(np.stack([h.real, h.imag], axis=1).ravel())
    out = np.zeros(2 * len(spec))
    lib().ob_to_transfer(_ptr(sr, C.c_double), len(spec), _ptr(sig, C.c_double), len(sig), grid.dt, grid.dx,
                         _ptr(fr, C.c_double), len(fr), c, r_ref, _ptr(hr, C.c_double), _ptr(out, C.c_double))
    return out[0::2] + 1j * out[1::2]


def phase_error_deg_per_m(f, c, dx):
    return lib().ob_phase_error_deg_per_m(f, c, dx)


def admittance(fit, f):
    ft = make_fit(*fit)
    out = np.zeros(2)
    lib().ob_admittance(C.byref(ft), f, _ptr(out, C.c_double))
    return complex(out[0], out[1])


def paris(beta, n=256):
    return lib().ob_paris(beta.real, beta.imag, n)


def fit_real_impedance(alpha):
    return lib().ob_fit_real_impedance(alpha)


def boundary_slots(face_fits, fits, dt):
    ff = np.asarray(face_fits, dtype=np.int32)
    fa = (ObFit * len(fits))(*[make_fit(*f) for f in fits])
    out = (ObSlot * 64)()
    ns = lib().ob_boundary_slots(_ptr(ff, C.c_int32), len(ff), fa, dt, out, 64)
    return np.array([[out[j].inv_d, out[j].cv, out[j].k, out[j].w] for j in range(ns)]).reshape(-1, 4)


def boundary_update(slots, states, p, p_prev, lap, lam, dt):
    """fdtd.rs:581-598; ``states`` (S, 2) is updated in place."""
    sl = np.ascontiguousarray(slots, dtype=np.float64)
    st = np.ascontiguousarray(states, dtype=np.float64)
    nxt = lib().ob_boundary_update(sl.ctypes.data_as(C.POINTER(ObSlot)), len(sl), _ptr(st, C.c_double), p, p_prev,
                                   lap, lam, dt)
    states[...] = st
    return nxt


def run(grid, taps, signal, receivers, freqs=(), ir=True, threads=0, cancel=None):
    """fdtd.rs:644-725 (run_cpu_with). Returns (ir [R][T] or None, spectra [R][F] or None, status)."""
    tc = np.ascontiguousarray([t[0] for t in taps], dtype=np.int64)
    tw = np.ascontiguousarray([t[1] for t in taps], dtype=np.float64)
    sig = np.ascontiguousarray(signal, dtype=np.float64)
    rc = np.ascontiguousarray(receivers, dtype=np.int64)
    fr = np.ascontiguousarray(freqs, dtype=np.float64)
    T, R, F = len(sig), len(rc), len(fr)
    ir_out = np.zeros((R, T)) if ir else None
    spec = np.zeros(2 * R * F) if F else None
    cflag = None
    if cancel is not None:
        cflag = np.array([1 if cancel else 0], dtype=np.int32)
    v = grid.view()
    st = lib().ob_run(C.byref(v), _ptr(tc, C.c_int64), _ptr(tw, C.c_double), len(tc), _ptr(sig, C.c_double), T,
                      _ptr(rc, C.c_int64), R, _ptr(fr, C.c_double), F,
                      _ptr(ir_out, C.c_double) if ir else None, _ptr(spec, C.c_double) if F else None, threads,
                      _ptr(cflag, C.c_int32) if cflag is not None else None)
    sp = None
    if F and st == 0:
        sp = (spec[0::2] + 1j * spec[1::2]).reshape(R, F)
    return ir_out, sp, st


def run_fields(grid, taps, signal, threads=0):
    """Fields p^T, p^(T-1) and the slot states after len(signal) steps."""
    tc = np.ascontiguousarray([t[0] for t in taps], dtype=np.int64)
    tw = np.ascontiguousarray([t[1] for t in taps], dtype=np.float64)
    sig = np.ascontiguousarray(signal, dtype=np.float64)
    N = grid.cells
    p = np.zeros(N)
    prev = np.zeros(N)
    st = np.zeros(2 * max(1, len(grid.slots)))
    v = grid.view()
    lib().ob_run_fields(C.byref(v), _ptr(tc, C.c_int64), _ptr(tw, C.c_double), len(tc), _ptr(sig, C.c_double),
                        len(sig), threads, _ptr(p, C.c_double), _ptr(prev, C.c_double), _ptr(st, C.c_double))
    return p, prev, st[: 2 * len(grid.slots)].reshape(-1, 2)


def baffle_point(grid, p, axis):
    """fdtd.rs:497-512"""
    pp = np.ascontiguousarray(p, dtype=np.float64)
    ax = np.ascontiguousarray(axis, dtype=np.float64)
    out = np.zeros(3)
    g = grid._obgrid()
    lib().ob_baffle_point(C.byref(g), _ptr(pp, C.c_double), _ptr(ax, C.c_double), _ptr(out, C.c_double))
    return out


def blend(ism, wave, w):
    """fdtd.rs:847-862"""
    a = np.array([complex(ism).real, complex(ism).imag])
    b = np.array([complex(wave).real, complex(wave).imag])
    out = np.zeros(2)
    lib().ob_blend(_ptr(a, C.c_double), _ptr(b, C.c_double), float(w), _ptr(out, C.c_double))
    return complex(out[0], out[1])


def calibrate_cabinet(verts, tris, ac, c, f_max, ppw, cells_across, freqs, threads=0):
    """fdtd.rs:890-955 with run_cpu_with as the runner; None where the reference returns None."""
    v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.int32).reshape(-1, 3))
    a = np.ascontiguousarray(np.asarray(ac, dtype=np.float64).reshape(3))
    fr = np.ascontiguousarray(freqs, dtype=np.float64)
    F = len(fr)
    on = np.zeros(2 * max(F, 1))
    ang = np.zeros(12)
    dr = np.zeros(2 * 12 * max(F, 1))
    st = lib().ob_calibrate_cabinet(_ptr(v, C.c_double), len(v), _ptr(t, C.c_int32), len(t), _ptr(a, C.c_double),
                                    float(c), float(f_max), float(ppw), float(cells_across), _ptr(fr, C.c_double), F,
                                    threads, _ptr(on, C.c_double), _ptr(ang, C.c_double), _ptr(dr, C.c_double))
    if st != 0:
        return None
    return {"on_axis": on[:2 * F].view(np.complex128).copy(), "angles_deg": ang,
            "directivity": dr[:2 * 12 * F].view(np.complex128).reshape(12, F).copy()}


def _spectrum(spec, keep):
    sp = ObSpectrum()
    if isinstance(spec, dict):
        f = np.ascontiguousarray(spec["freqs_hz"], dtype=np.float64)
        v = np.ascontiguousarray(spec["values"], dtype=np.float64)
        sp.n = len(f)
    else:
        f, v = np.zeros(1), np.array([float(spec)])
        sp.n = 0
    keep += [f, v]
    sp.f, sp.v = _ptr(f, C.c_double), _ptr(v, C.c_double)
    return sp


def _material(mat, keep):
    m = ObMaterial()
    m.absorption = _spectrum(mat.get("absorption", 0.0), keep)
    if mat.get("impedance") is not None:
        m.has_impedance = 1
        m.imp_re = _spectrum(mat["impedance"][0], keep)
        m.imp_im = _spectrum(mat["impedance"][1], keep)
    return m


def material_absorption(mat, f):
    """material.rs:72-78"""
    keep = []
    m = _material(mat, keep)
    return float(lib().ob_material_absorption(C.byref(m), float(f)))


def fit_admittance(mat, band, max_branches=2):
    """impedance_fit.rs:158-235: ((beta0, [(r, m, k), ...]), {"band", "rms_error", "max_error"})."""
    keep = []
    m = _material(mat, keep)
    out = ObFit()
    b = np.zeros(2)
    e = np.zeros(2)
    nb = lib().ob_fit_admittance(C.byref(m), float(band[0]), float(band[1]), int(max_branches), C.byref(out),
                                 _ptr(b, C.c_double), _ptr(e, C.c_double))
    fit = (out.beta0, [(out.br[k].r, out.br[k].m, out.br[k].k) for k in range(nb)])
    return fit, {"band": (b[0], b[1]), "rms_error": e[0], "max_error": e[1]}


def first_order_directions():
    """fdtd.rs:753-760"""
    out = np.zeros(1200)
    lib().ob_first_order_directions(_ptr(out, C.c_double))
    return out.reshape(400, 3)


def fit_first_order(d, d_axis, to_room):
    """fdtd.rs:749-807: (a0, [a_x, a_y, a_z], relative RMS residual)."""
    dv = np.ascontiguousarray(np.asarray(d, dtype=np.complex128).reshape(400).view(np.float64))
    ax = np.array([complex(d_axis).real, complex(d_axis).imag])
    R = np.ascontiguousarray(np.asarray(to_room, dtype=np.float64).reshape(3, 3))
    a0, a = np.zeros(2), np.zeros(6)
    res = lib().ob_fit_first_order(_ptr(dv, C.c_double), _ptr(ax, C.c_double), _ptr(R, C.c_double),
                                   _ptr(a0, C.c_double), _ptr(a, C.c_double))
    return complex(a0[0], a0[1]), [complex(a[2 * i], a[2 * i + 1]) for i in range(3)], float(res)
