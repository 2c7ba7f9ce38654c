/*
 * FP64 CPU oracle for the Wavey FDTD wave-solver path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * scheme (core/wavey-solver/src/fdtd.rs). It is used by tests/, by the smoke
 * check and as the CPU baseline leg of bench.py, always as the checker and
 * never as the thing measured or shipped. The product (wavey_amd/) never
 * links or calls it.
 *
 * Each function cites the reference file:line it restates. Ray casts test
 * every triangle instead of walking the reference's BVH (see the general
 * rooms below).
 */
#ifndef WAVEY_FDTD_ORACLE_H
#define WAVEY_FDTD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OB_MAX_BRANCHES 8

/* core/wavey-materials/src/impedance_fit.rs:27-44 (Branch, AdmittanceFit). */
typedef struct { double r, m, k; } ob_branch;
/* impedance_fit.rs:34-44 (AdmittanceFit) with its band and errors: fits
 * compare as the derived PartialEq does, every field with ==. */
typedef struct {
    double beta0;
    int32_t nb;
    ob_branch br[OB_MAX_BRANCHES];
    double band[2];
    double rms_error, max_error;
} ob_fit;

/* core/wavey-solver/src/fdtd.rs:178-187 (Slot). */
typedef struct { double inv_d, cv, k, w; } ob_slot;

/*
 * Box room [0,L0]x[0,L1]x[0,L2] with rectangular floor-to-ceiling pillars.
 * Surfaces: 0..5 are the walls x=0, x=L0, y=0, y=L1, floor z=0, ceiling
 * z=L2; pillar p owns surfaces 6+4p+{0,1,2,3} = its sides facing -x, +x,
 * -y, +y. Surface s is cut into a gu[s] x gv[s] patch grid over its (u,v)
 * extent (walls x: u=y v=z; walls y: u=x v=z; floor/ceiling: u=x v=y;
 * pillar sides: the same as the wall they are parallel to, over the pillar
 * extent), and patch (iu,iv) uses fit patch_fit[patch_off[s] + iv*gu[s] + iu].
 */
typedef struct {
    double L[3];
    int32_t n_pillars;
    const double* pillars; /* [n_pillars][4] = x0, x1, y0, y1 */
    int32_t n_fits;
    const ob_fit* fits;
    const int32_t* gu;
    const int32_t* gv;
    const int32_t* patch_off;
    const int32_t* patch_fit;
} ob_room;

/* core/wavey-solver/src/fdtd.rs:196-209 (FdtdGrid), boundary in CSR form. */
typedef struct {
    double origin[3], dx, dt;
    int64_t n[3];
    uint8_t* flags;   /* [N] */
    uint32_t* bidx;   /* [N] */
    int64_t nb;
    uint32_t* b_cell; /* [nb] */
    uint32_t* b_start;/* [nb+1] */
    int64_t ns;
    ob_slot* slots;   /* [ns] */
} ob_grid;

int ob_build_box(const ob_room* room, double c, double dx, ob_grid* out);
void ob_grid_free(ob_grid* g);

/* fdtd.rs:539-571 */
int ob_boundary_slots(const int32_t* face_fits, int n_faces, const ob_fit* fits, double dt,
                      ob_slot* out, int max_out);
/* fdtd.rs:581-598; states are interleaved (u, g) pairs. */
double ob_boundary_update(const ob_slot* slots, int n, double* states, double p, double p_prev,
                          double lap, double lam, double dt);
/* fdtd.rs:601-610 */
void ob_source_signal(double f_max, double dt, int64_t steps, double* out);
/* fdtd.rs:245-272; returns -1 when no air node is within 16 cells. */
int64_t ob_nearest_air(const ob_grid* g, const double p[3]);
/* fdtd.rs:497-512 (FdtdGrid::baffle_point). */
void ob_baffle_point(const ob_grid* g, const double p[3], const double axis[3], double out[3]);
/* fdtd.rs:516-533 (FdtdGrid::air_along); -1 for None. */
int64_t ob_air_along(const ob_grid* g, const double p[3], const double axis[3]);
/* run.rs:601-624: a wave-job source's component taps (cells/w [8], len[4]);
 * returns the number of components, 0 when the source is outside the grid. */
int ob_wave_source_taps(const ob_grid* g, const double center[3], const double axis[3], int calibrated, int order,
                        int64_t* cells, double* w, int* len, int* fallback);
/* fdtd.rs:847-862 (blend); complex values as (re, im). */
void ob_blend(const double ism[2], const double wave[2], double w, double out[2]);
/* fdtd.rs:467-492; returns the number of taps (0 = None). */
int ob_taps_at(const ob_grid* g, const double p[3], int64_t* cells, double* weights);
/* fdtd.rs:730-743; returns the number of components (0 = None).
 * comp_cells/comp_w are [4][2]; comp_len[4]. */
int ob_source_components(const ob_grid* g, int64_t c, int order, int64_t* comp_cells,
                         double* comp_w, int* comp_len);
/* fdtd.rs:753-760: the 400 golden-spiral directions [400][3], local frame. */
void ob_first_order_directions(double* out);
/* fdtd.rs:749-807 (fit_first_order): d [400] complex directivity at those
 * directions, d_axis the on-axis value, to_room row-major (applied as
 * to_room * v). Writes a0 (re, im) and a [3] (re, im); returns the relative
 * RMS residual. */
double ob_fit_first_order(const double* d, const double d_axis[2], const double to_room[9], double a0_out[2],
                          double a_out[6]);
/* fdtd.rs:812-815; a0, a[3] and out[4] are complex (re, im). */
void ob_component_weights(const double a0[2], const double a[6], double k, double dx, double out[8]);
/* fdtd.rs:819-840; spectra/h0/out complex interleaved. */
void ob_to_transfer(const double* spectra, int64_t n_spec, const double* signal, int64_t steps,
                    double dt, double dx, const double* freqs, int64_t nf, double c, double r_ref,
                    const double* h0, double* out);
/* fdtd.rs:157-162 */
double ob_phase_error_deg_per_m(double f, double c, double dx);
/* spectrum.rs:24-44 (Spectrum::eval); n = 0 is the scalar values[0]. */
double ob_spectrum_eval(const double* freqs_hz, const double* values, int32_t n, double f);
/* material.rs:7-31 (Material), the parts absorption_at reads. */
typedef struct { int32_t n; const double* f; const double* v; } ob_spectrum;
typedef struct {
    ob_spectrum absorption;
    int32_t has_impedance;
    ob_spectrum imp_re, imp_im;
} ob_material;
/* material.rs:72-78 */
double ob_material_absorption(const ob_material* m, double f);
/* impedance_fit.rs:158-235 (fit_admittance, with nelder_mead :87-129, decode
 * and encode_branch :139-155), max_branches <= 2. Returns the number of
 * branches kept; band_out = (lo, hi), err_out = (rms, max) over the band. */
int ob_fit_admittance(const ob_material* mat, double band_lo, double band_hi, int max_branches, ob_fit* out,
                      double band_out[2], double err_out[2]);
/* impedance_fit.rs:48-55, :73-84; material.rs:96-111 */
void ob_admittance(const ob_fit* fit, double f, double out[2]);
double ob_paris(double beta_re, double beta_im, int n);
double ob_fit_real_impedance(double alpha);

/*
 * fdtd.rs:644-725 (run_cpu_with). Records the IR [n_recv][steps] when ir_out
 * is non-null and the DFT spectra [n_recv][nf] (complex) when spec_out is
 * non-null. The grid is passed as plain arrays (the same view the GPU library
 * takes) so both sides can run on one grid. Returns 0, or 1 when the cancel
 * flag was seen set at a 64-step check.
 */
typedef struct {
    int64_t n[3];
    double dx, dt;
    const uint8_t* flags;
    int64_t nb;
    const uint32_t* b_cell;
    const uint32_t* b_start;
    const ob_slot* slots;
} ob_grid_view;

int ob_run(const ob_grid_view* g, const int64_t* tap_cells, const double* tap_w, int64_t n_taps,
           const double* signal, int64_t steps, const int64_t* recv, int64_t n_recv,
           const double* freqs, int64_t nf, double* ir_out, double* spec_out, int threads,
           const volatile int32_t* cancel);

/* Same loop, stopped after `steps` steps; returns both field arrays for tests. */
int ob_run_fields(const ob_grid_view* g, const int64_t* tap_cells, const double* tap_w,
                  int64_t n_taps, const double* signal, int64_t steps, int threads, double* p_out,
                  double* prev_out, double* state_out);

/*
 * General rooms (fdtd.rs:289-461, core/wavey-geom/src/room.rs:133-157,
 * core/wavey-geom/src/bvh.rs:63-226). Every ray cast here tests every
 * triangle (no BVH): the closest hit, and for all-hits queries the hits
 * sorted by t, equal t in triangle order. The reference's BVH visits
 * triangles in tree order, so the two differ only where two triangles are
 * hit at exactly the same t.
 *
 * prism = 1: inside when 0 < z < height and (x, y) is in a plan region
 * (region r = cycle region_cycles[r] as its outer boundary, the cycles up
 * to region_cycles[r+1] as holes; cycle k = points cycle_off[k] ..
 * cycle_off[k+1] of cycle_xy). prism = 0: column ray parity.
 */
typedef struct {
    int32_t prism;
    double height;
    int32_t n_pos;
    const double* pos;        /* [n_pos][3] */
    int32_t n_tri;
    const int32_t* tri;       /* [n_tri][3] */
    const int32_t* tri_surface;
    int32_t n_fits;           /* one per surface */
    const ob_fit* fits;
    int32_t n_regions;
    const int32_t* region_cycles;
    const int32_t* cycle_off;
    const double* cycle_xy;
} ob_mesh_room;

/* fdtd.rs:424-461 (Solid): n . p <= d for every plane, inside [lo, hi]. */
typedef struct {
    int32_t n_planes;
    const double* planes; /* [n_planes][4] = nx, ny, nz, d */
    double lo[3], hi[3];
} ob_solid;

/* bvh.rs:209-226: Moller-Trumbore, two-sided; returns 1 and *t on a hit. */
int ob_ray_triangle(const double o[3], const double d[3], const double a[3], const double b[3],
                    const double c[3], double* t);
/* room.rs:355-367: inside when at least two of three fixed rays cross an odd
 * number of triangles. */
int ob_parity_inside(const ob_mesh_room* room, const double p[3]);
/* fdtd.rs:433-456; rot row-major, applied as rot * v. Returns the number of
 * distinct planes written to planes_out ([n_tris][4] room), -1 for no
 * triangles. */
int ob_solid_from_mesh(const double* verts, int32_t n_verts, const int32_t* tris, int32_t n_tris,
                       const double rot[9], const double pos[3], double* planes_out, double lo[3],
                       double hi[3]);
int ob_solid_contains(const ob_solid* s, const double p[3]);
/* fdtd.rs:289-377 (build_dx) with column_fill (:385-422). */
int ob_build_room(const ob_mesh_room* room, const ob_solid* solids, int32_t n_solids, double c,
                  double dx, ob_grid* out);

/* fdtd.rs:890-955 (calibrate_cabinet) with run_cpu_with as the runner:
 * on_axis_out [nf] and dir_out [12][nf] complex (re, im), angles_out [12]
 * degrees. 0, or -1 where the reference returns None. */
int ob_calibrate_cabinet(const double* verts, int32_t n_verts, const int32_t* tris, int32_t n_tris,
                         const double ac[3], double c, double f_max, double ppw, double cells_across,
                         const double* freqs, int32_t nf, int threads, double* on_axis_out,
                         double* angles_out, double* dir_out);

#ifdef __cplusplus
}
#endif
#endif
