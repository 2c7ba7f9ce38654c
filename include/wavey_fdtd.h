/*
 * wavey_fdtd.h -- C ABI of the MI355X-native FDTD wave solver.
 *
 * Drop-in for the wave-solver run behind `wavey simulate --fdtd <Hz>`:
 * core/wavey-solver/src/run.rs:170-192 (run_fdtd), whose GPU implementation
 * is Gpu::run_fdtd (core/wavey-solver/src/gpu_fdtd.rs:45-277) and whose CPU
 * one is fdtd::run_cpu / run_cpu_ir (core/wavey-solver/src/fdtd.rs:616-641).
 *
 * Conventions (mirroring the reference):
 *  - Cell indices are the reference's: (z*ny + y)*nx + x, x fastest
 *    (fdtd.rs:227-229).
 *  - Inputs are borrowed for the call; the caller owns every buffer. The
 *    library copies inputs to the device and writes outputs into caller
 *    buffers.
 *  - Status: WV_OK = Ok(Some), WV_CANCELLED = Ok(None), negative = Err(String)
 *    with the message in `err` (run.rs:184-187 falls back to the CPU on Err).
 *  - `cancel` is polled between 128-step blocks (gpu_fdtd.rs:202-205);
 *    `progress` gets done/total after each block, on the calling thread.
 *  - One HIP device (the current one), one stream per plan, default priority.
 *    Entry points are reentrant; calls on different plans may overlap.
 */
#ifndef WAVEY_FDTD_H
#define WAVEY_FDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    WV_OK = 0,
    WV_CANCELLED = 1,
    WV_ERR_INVALID = -1, /* bad argument (the reference would panic or Err) */
    WV_ERR_DEVICE = -2,  /* HIP error: the caller falls back to the CPU */
    WV_ERR_LIMIT = -3,   /* grid/steps/boundary beyond what this build takes */
    WV_ERR_NOMEM = -4    /* device memory exhausted */
} wv_status;

/* Output forms: the reference GPU path returns spectra only (gpu_fdtd.rs:262-276),
 * run_cpu_ir returns the receiver time series (fdtd.rs:630-641). */
enum { WV_MODE_SPECTRA = 1, WV_MODE_IR = 2, WV_MODE_BOTH = 3 };

/* Arithmetic: FP32 fields and state (the reference GPU path), FP64 throughout
 * (the reference CPU path), or FP32 fields with FP64 boundary-filter state. */
enum { WV_FP32 = 0, WV_FP64 = 1, WV_FP32_STATE64 = 2 };

/* fdtd.rs:178-187 (Slot): u+ = (1/2 - w)*inv_d*p+ + (cv*u- - k*g- + p/2 + w*p-)*inv_d. */
typedef struct {
    double inv_d, cv, k, w;
} wv_slot;

/* fdtd.rs:196-209 (FdtdGrid) with the boundary list in CSR form:
 * node b sits at cell b_cell[b] and owns slots b_start[b] .. b_start[b+1]-1.
 * flags: bit 6 = air, bits 0-5 = neighbour (-x,+x,-y,+y,-z,+z) is air. */
typedef struct {
    uint64_t n[3];
    double dx, dt;
    double origin[3];
    const uint8_t* flags;     /* [n0*n1*n2] */
    uint64_t n_boundary;
    const uint32_t* b_cell;   /* [n_boundary] */
    const uint32_t* b_start;  /* [n_boundary + 1] */
    const wv_slot* slots;     /* [b_start[n_boundary]] */
} wv_grid;

/* One source tap (cell, weight): fdtd.rs:467-492, gpu_fdtd.rs:67-75. */
typedef struct {
    uint64_t cell;
    double w;
} wv_tap;

typedef void (*wv_progress_fn)(double fraction, void* ctx);

/* --------------------------------------------------------------- settings */

/* fdtd.rs:44-82 (WaveAccuracy) and fdtd.rs:141-176 (FdtdSettings). */
enum { WV_ACCURACY_DRAFT = 0, WV_ACCURACY_STANDARD = 1, WV_ACCURACY_HIGH = 2 };
/* fdtd.rs:61-67: grid headroom above the highest frequency needed. */
double wv_accuracy_headroom(int accuracy);
/* fdtd.rs:72-78: cells across a cabinet's smallest side. */
double wv_accuracy_cabinet_cells(int accuracy);
/* fdtd.rs:144-146: the top of the crossover band, crossover * sqrt(2). */
double wv_settings_f_max(double crossover_hz);
/* fdtd.rs:148-154 (grid_fmax). */
double wv_grid_fmax(double crossover_hz, int accuracy, const double* freqs, size_t nf);
/* fdtd.rs:166-175 (weight): the wave result's weight at f. */
double wv_crossover_weight(double crossover_hz, double f);
/* fdtd.rs:157-162 */
double wv_phase_error_deg_per_m(double f, double c, double dx);

/* ------------------------------------------------------------------ run */

/*
 * One solve: replaces Gpu::run_fdtd (gpu_fdtd.rs:45-54) and, with
 * WV_MODE_IR, run_cpu_ir (fdtd.rs:630-641). Steps = signal length. Records
 * p^n before the update of step n (fdtd.rs:673-683): ir_out [n_recv][steps]
 * (doubles), spectra_out [n_recv][nf] complex interleaved (re, im) with the
 * DFT sum_n p^n e^{-j 2 pi fract(f dt n)} (fdtd.rs:679-682).
 */
wv_status wv_fdtd_run(const wv_grid* grid, const wv_tap* taps, size_t n_taps, const double* signal,
                      size_t steps, const uint64_t* receivers, size_t n_recv, const double* freqs,
                      size_t nf, int mode, int precision, const volatile int32_t* cancel,
                      wv_progress_fn progress, void* ctx, double* spectra_out, double* ir_out,
                      char* err, size_t err_len);

/*
 * wv_fdtd_run without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * beside HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Every device buffer of the run is carved from one static
 * device array of 64 MiB per GPU, held from the library's load until the
 * process ends: FP32 grids up to about 3.5 M cells (FP64 about 1.8 M), less
 * the receiver record (steps x receivers x 4 or 8 bytes). A run whose
 * estimated footprint is larger returns WV_ERR_LIMIT with a message.
 * Arguments, results and status otherwise as wv_fdtd_run, with the same
 * kernels, so results equal wv_fdtd_run's bit for bit. The first call on a
 * device looks up the array's device address and creates that device's
 * stream and events. Calls on one GPU share its array and are run one at a
 * time (a per-device lock); the array keeps the last call's contents, and
 * every run sets up all it reads.
 */
wv_status wv_fdtd_run_noalloc(const wv_grid* grid, const wv_tap* taps, size_t n_taps, const double* signal,
                              size_t steps, const uint64_t* receivers, size_t n_recv, const double* freqs,
                              size_t nf, int mode, int precision, const volatile int32_t* cancel,
                              wv_progress_fn progress, void* ctx, double* spectra_out, double* ir_out,
                              char* err, size_t err_len);

/*
 * Plan form of the same run: the grid is uploaded and its device layout built
 * once, then many runs (source components, calibrations: run.rs:645-661,
 * fdtd.rs:941-942) reuse it. A plan is used by one thread at a time.
 */
typedef struct wv_plan wv_plan;
wv_status wv_plan_create(const wv_grid* grid, int precision, wv_plan** out, char* err, size_t err_len);
wv_status wv_plan_run(wv_plan* plan, const wv_tap* taps, size_t n_taps, const double* signal,
                      size_t steps, const uint64_t* receivers, size_t n_recv, const double* freqs,
                      size_t nf, int mode, const volatile int32_t* cancel, wv_progress_fn progress,
                      void* ctx, double* spectra_out, double* ir_out, char* err, size_t err_len);
/*
 * wv_plan_run in three parts, so a caller can advance a run in pieces
 * (progress, cancellation, timing a window of steps). wv_run_begin zeroes
 * the fields and state and builds the per-run device layout (taps join the
 * special-cell list); wv_run_advance enqueues k more steps (k <= remaining)
 * and returns when they are done; wv_run_finish reads the receivers back.
 * The reference loop is one piece (gpu_fdtd.rs:200-276).
 */
wv_status wv_run_begin(wv_plan* plan, const wv_tap* taps, size_t n_taps, const double* signal,
                       size_t steps, const uint64_t* receivers, size_t n_recv, char* err,
                       size_t err_len);
wv_status wv_run_advance(wv_plan* plan, size_t k, const volatile int32_t* cancel,
                         wv_progress_fn progress, void* ctx, char* err, size_t err_len);
wv_status wv_run_finish(wv_plan* plan, const double* freqs, size_t nf, int mode,
                        double* spectra_out, double* ir_out, char* err, size_t err_len);
/* Device time of the last wv_run_advance's step loop (HIP events on the
 * plan's stream), its step count and kernel launches. No reference
 * counterpart: measurement hook. */
wv_status wv_plan_last_timing(const wv_plan* plan, double* step_loop_ms, uint64_t* steps,
                              uint64_t* launches);
/* Sizes of the plan's device layout (for roofline accounting). No reference
 * counterpart. out: [cells, air_cells, boundary_nodes, slots, special_cells,
 * pitched_cells, slots_with_k]. */
wv_status wv_plan_stats(const wv_plan* plan, uint64_t out[7]);
/* Single-step kernel shape: z planes per block (0 = default, else a power of
 * two up to 32) and the block variant (-1 = default; 0..3 = blocks of 7, 3,
 * 15 and 1 plane tiles of 256 FP32 or 128 FP64 cells, plus one special-cell
 * wave, four with one tile). Defaults by grid size: FP32 variant 3 with
 * 1-plane chunks below 128 Ki cells, 2-plane below 1 Mi and 4-plane below
 * 2.5 Mi; FP64 variant 3 with 2-plane chunks below 128 Ki cells and 4-plane
 * below 1 Mi; else variant 0 with 4-plane chunks. Takes effect at the next
 * wv_run_begin; results do not depend on the shape. No reference
 * counterpart: the autotuner's knobs. */
wv_status wv_plan_set_tile(wv_plan* plan, uint32_t zchunk, int32_t variant);
/* Steps per pass over the fields: 1; 2 (two steps per z-march: deep
 * interior cells advance twice in one pass, the cells next to boundaries in
 * a second, sparse launch; four field buffers); or 3 (the cells three or more
 * cells from a boundary advance three times in one pass, the others in two
 * sparse launches; five field buffers); 0 = default. pair_zchunk: z planes
 * per block of a multi-step pass (0 = default, else a power of two up to 64;
 * two-step and FP64 passes take at most 32). Results are bit-identical in
 * every mode. Takes effect at the next
 * wv_run_begin. No reference counterpart: a tuning knob. */
wv_status wv_plan_set_pass(wv_plan* plan, uint32_t steps_per_pass, uint32_t pair_zchunk);
/* The pass mode in use (the current run's, else the next run's): out =
 * [steps per pass, pass z-chunk, ring cells (interior, next to a
 * non-interior cell) and cells within two of a non-interior cell (three
 * steps per pass only) of the current run]. By default FP32 grids of 8 Mi
 * cells and more (FP64: 7 Mi) take three steps per pass, from 2.25 Mi cells
 * (FP64: 0.75 Mi) two, smaller ones one. Default z-chunk: FP32 three-step
 * passes the longest of 64, 32, 16 that gives 640 tile blocks; FP64
 * three-step and every two-step pass 32 from 64 Mi cells; else 16, or 8
 * where 16 give fewer than 150 tile blocks. No reference counterpart. */
wv_status wv_plan_pass_info(const wv_plan* plan, uint32_t out[4]);
/* Copies `bytes` (at most one field buffer) between the plan's field
 * buffers `reps` times with a plain streaming kernel and returns the mean
 * time per copy. It overwrites the fields: call it outside a run. No
 * reference counterpart: the bandwidth and HBM-counter calibration. */
wv_status wv_plan_copy_probe(wv_plan* plan, uint64_t bytes, uint32_t reps, double* ms);
void wv_plan_free(wv_plan* plan);

/* Device bytes a run of this grid will hold (fields, layout, slot state,
 * records), estimated before any allocation; the server admits requests by
 * it. No reference counterpart. */
wv_status wv_footprint_estimate(const wv_grid* grid, int precision, size_t steps, size_t n_recv,
                                size_t n_taps, uint64_t* bytes);

/* ----------------------------------------------------------------- batch
 *
 * Many runs (rooms, the per-source and per-component runs of a wave job,
 * run.rs:645-661) of different grid sizes advanced together, in 128-step
 * blocks replayed as HIP graphs. A block is one launch per step over every
 * instance's tiles (WV_BATCH_STEPS) or one persistent launch (WV_BATCH_PERSISTENT:
 * each XCD runs a share of the instances through the whole block, an
 * instance's step waiting only on its own previous step). Each instance runs
 * the same arithmetic as wv_plan_run, so its results equal that run's. The reference runs them one after another.
 * wv_batch_add may be called from several threads at once (each instance is
 * set up on its own stream); every other call on a batch is made by one
 * thread at a time, with no wv_batch_add in progress. Inputs are copied at
 * wv_batch_add.
 */
typedef struct wv_batch wv_batch;
wv_status wv_batch_create(int precision, wv_batch** out, char* err, size_t err_len);
/* Uploads a grid and starts its run from rest; *id numbers instances 0, 1, ... */
wv_status wv_batch_add(wv_batch* batch, const wv_grid* grid, const wv_tap* taps, size_t n_taps,
                       const double* signal, size_t steps, const uint64_t* receivers, size_t n_recv,
                       int32_t* id, char* err, size_t err_len);
/* Times the whole batch over trial_steps steps per candidate: the block variant
 * for the launch, then the z-chunk of each size class (floor(log2(air
 * cells)), largest first), then the launch form. The choices are kept for
 * later batches and servers of the process (servers use WV_BATCH_STEPS);
 * every instance restarts from rest. */
wv_status wv_batch_autotune(wv_batch* batch, uint32_t trial_steps, char* err, size_t err_len);
/* Sets one shape for the whole batch (variant as wv_plan_set_tile; zchunk 0 =
 * default); instances restart from rest. */
wv_status wv_batch_set_tile(wv_batch* batch, int32_t variant, uint32_t zchunk, char* err,
                            size_t err_len);
enum { WV_BATCH_STEPS = 0, WV_BATCH_PERSISTENT = 1 };
/* Sets the launch form of the batch's blocks; instances keep their state.
 * WV_BATCH_PERSISTENT where the device's XCC_ID register does not name 8
 * XCDs returns WV_ERR_LIMIT. */
wv_status wv_batch_set_launch(wv_batch* batch, int form, char* err, size_t err_len);
/* Runs every instance to its last step. use_graphs = 0 issues the same
 * launches one by one. Cancel is polled between step blocks. An instance
 * left alone on a grid of at least 2.25 Mi cells (FP64: 0.75 Mi) runs its
 * remaining steps with the plain run's multi-step kernels (two or three steps per pass, as its
 * plain run; same results). */
wv_status wv_batch_run(wv_batch* batch, int use_graphs, const volatile int32_t* cancel,
                       wv_progress_fn progress, void* ctx, char* err, size_t err_len);
/* An instance's results, as wv_run_finish (ir_out [R][T], spectra [R][F][2]). */
wv_status wv_batch_result(wv_batch* batch, int32_t id, const double* freqs, size_t nf, int mode,
                          double* spectra_out, double* ir_out, char* err, size_t err_len);
/* out: [cells, air_cells, steps, zchunk, size_class, device_bytes]. */
wv_status wv_batch_info(const wv_batch* batch, int32_t id, uint64_t out[6]);
/* Last wv_batch_run: device ms (HIP events on the batch stream); out: [step
 * launches per instance block sequence, cell updates, graph launches, graph
 * captures, block variant, blocks run alone, launch form]. */
wv_status wv_batch_timing(const wv_batch* batch, double* ms, uint64_t out[7]);
void wv_batch_free(wv_batch* batch);

/* ----------------------------------------------------------------- serving
 *
 * A request loop on a worker thread. Submitted runs wait in a queue; each is
 * admitted when its estimated device footprint fits under hbm_cap_bytes
 * beside the running ones, joins the running batch at the next 128-step
 * boundary (continuous batching) and leaves it when done, its results copied
 * to host memory and its device memory freed. The grid passed to
 * wv_serve_submit is borrowed until wv_serve_wait returns for that request;
 * taps, signal, receivers and frequencies are copied. No reference
 * counterpart: the reference runs one solve at a time.
 */
typedef struct wv_server wv_server;
wv_status wv_serve_create(int precision, uint64_t hbm_cap_bytes, wv_server** out, char* err,
                          size_t err_len);
wv_status wv_serve_submit(wv_server* server, const wv_grid* grid, const wv_tap* taps,
                          size_t n_taps, const double* signal, size_t steps,
                          const uint64_t* receivers, size_t n_recv, const double* freqs, size_t nf,
                          int mode, uint64_t* id, char* err, size_t err_len);
/* Blocks until the request is done; results as wv_run_finish; *latency_s =
 * completion - submission. Each id can be waited for once. */
wv_status wv_serve_wait(wv_server* server, uint64_t id, double* spectra_out, double* ir_out,
                        double* latency_s, char* err, size_t err_len);
/* Which admitted requests each 128-step block advances.
 * WV_SERVE_ALL: every one, in one batched launch per step.
 * WV_SERVE_SHORTEST: least remaining work (air cells x steps left) first,
 * while the chosen air cells stay within fill_cells (0 = 16 Mi); the first
 * always runs, the others wait on the device.
 * WV_SERVE_LANES (the default): as WV_SERVE_SHORTEST for most requests; the
 * large ones (lane_cells = 0: those whose plan takes two or three steps per
 * pass on a grid of at least 2.25 Mi cells; else multi-step plans of at
 * least lane_cells air cells) run one at a time, in admission order and each
 * to its end, on a second stream beside the batch, which the device runs at
 * the same time.
 * A request that runs alone on a grid of at least 2.25 Mi cells (FP64:
 * 0.75 Mi) runs the plain run's multi-step kernels. Results do not depend on the policy. */
enum { WV_SERVE_ALL = 0, WV_SERVE_SHORTEST = 1, WV_SERVE_LANES = 2 };
wv_status wv_serve_set_policy(wv_server* server, int policy, uint64_t fill_cells, uint64_t lane_cells);
/* out: [completed, p50, p95, p99 latency s, first submission to last
 * completion s, cell updates, most requests running at once, peak device
 * bytes, failed, graph captures, blocks run alone on the batch stream, graph
 * launches, blocks on the large-request lane]. */
wv_status wv_serve_stats(wv_server* server, double out[13]);
/* Finishes every submitted request, then stops the worker. */
void wv_serve_free(wv_server* server);

/* -------------------------------------------------------------- wave job
 *
 * The wave solver's part of a wave job (run.rs:593-661): for each source, its
 * runs on one grid (one component for an omni or calibrated source's taps,
 * four for a first-order directional source: fdtd.rs:730-743; see
 * wv_wave_source_taps), each run's spectra turned into a transfer function
 * (to_transfer, fdtd.rs:819-840), divided by the cabinet calibration's
 * on-axis response where the source has one, and summed with the source's
 * per-frequency complex component weights (fdtd.rs:812-815).
 */
typedef struct {
    const wv_tap* taps;        /* every component's taps, back to back */
    const uint32_t* comp_off;  /* [n_comps + 1]: component c = taps[comp_off[c] .. comp_off[c+1]) */
    uint32_t n_comps;
    const double* weights;     /* [nf][n_comps] complex (re, im); NULL for a single component (weight 1) */
    const double* h0;          /* [nf] complex source spectrum; NULL for 1 */
    double r_ref;              /* reference distance of the source's sensitivity, m */
    const double* on_axis;     /* [nf] complex: the cabinet calibration's on-axis response, which every
                                  run's transfer function is divided by (run.rs:651-656); NULL for none */
} wv_wave_source;
/* transfer_out: [n_sources][n_recv][nf] complex (re, im). batched = 0 runs
 * the components one after another on one plan (the reference's order),
 * 1 hands every run to one batch; the results are the same bit for bit. */
wv_status wv_wave_job(const wv_grid* grid, const wv_wave_source* sources, size_t n_sources,
                      const double* signal, size_t steps, const uint64_t* receivers, size_t n_recv,
                      const double* freqs, size_t nf, double c, int precision, int batched,
                      const volatile int32_t* cancel, double* transfer_out, char* err,
                      size_t err_len);

/*
 * A cabinet's free-field calibration (CabinetCalibration, calibrate_cabinet:
 * fdtd.rs:873-955): verts/tris relative to the enclosure centre (outward
 * winding), acoustic centre ac in that frame, +x on axis. Builds a cubic box
 * with matched walls around it, runs the wave solver once with the cabinet
 * as a rigid solid and once without, and writes on_axis_out [nf] (gain of
 * cabinet + driver over the bare driver, complex (re, im)), angles_out [12]
 * (0, 30, ... 330 degrees, horizontal, +90 = left) and directivity_out
 * [12][nf] (complex, relative to on axis), faded toward 1 below the
 * frequency where the cabinet is half a radian across. batched = 1 runs the
 * two grids as one batch, 0 one after another; the results are the same.
 * WV_ERR_INVALID where the reference returns None.
 */
wv_status wv_calibrate_cabinet(const double* verts, uint32_t n_verts, const uint32_t* tris, uint32_t n_tris,
                               const double ac[3], double c, double f_max, double ppw, double cells_across,
                               const double* freqs, size_t nf, int precision, int batched,
                               const volatile int32_t* cancel, double* on_axis_out, double* angles_out,
                               double* directivity_out, char* err, size_t err_len);

/* ---------------------------------------------------------- result files
 *
 * SimulationResult::save (core/wavey-solver/src/result.rs:444-461), as
 * `wavey simulate` writes it (core/wavey-cli/src/main.rs:513-516), for the
 * wave solver's part of a result: results.json (serde's metadata, field
 * order of result.rs:16-90 and fdtd.rs:85-114, no cut planes, has_late
 * false) and results.bin (every microphone's transfer functions
 * [speaker][freq] as complex f32, microphones in order).
 */
enum { WV_RESULT_BANDS = 8 }; /* raytrace.rs:29-30: octave bands 63 Hz .. 8 kHz */
typedef struct {
    uint64_t id;
    const char* name;      /* UTF-8 */
    double position[3];
} wv_microphone;           /* receivers.rs:5-10 */
typedef struct {
    double crossover_hz, duration_s, points_per_wavelength;
    uint32_t source_order;
    int32_t cabinets;
    int32_t accuracy;      /* WV_ACCURACY_*; High is not written (fdtd.rs:112-113) */
} wv_fdtd_settings;        /* fdtd.rs:85-114 */
typedef struct {
    size_t nf;
    const double* freqs;
    size_t n_speakers;
    const uint64_t* speaker_ids;
    size_t n_mics;
    const wv_microphone* mics;
    const double* transfer;  /* [n_mics][n_speakers][nf] complex (re, im), stored as f32 */
    const char* backend;     /* RunStats (result.rs:80-90) */
    double seconds;
    uint64_t images, paths, diffracted_paths, receivers;
    size_t n_notes;
    const char* const* notes;
    uint64_t fingerprint;
    const double* rt_sabine; /* [WV_RESULT_BANDS] or NULL for zeros */
    const double* rt_eyring;
    const wv_fdtd_settings* fdtd; /* NULL: no wave stage (null) */
} wv_results;
/* Writes dir/results.json and dir/results.bin, creating dir. */
wv_status wv_results_save(const char* dir, const wv_results* r, char* err, size_t err_len);
/* result.rs:463-476: the transfers back from dir/results.bin, out
 * [n_mics][n_speakers][nf] complex f32 (re, im). */
wv_status wv_results_read_transfer(const char* dir, size_t n_mics, size_t n_speakers, size_t nf, float* out,
                                   char* err, size_t err_len);

/* ------------------------------------------------------------ grid build */

/* impedance_fit.rs:27-44 (Branch, AdmittanceFit). Faces of a node share
 * slots only where their whole fits are equal, as the reference's derived
 * PartialEq compares them (fdtd.rs:542): beta0, the branches, band,
 * rms_error and max_error, each with ==, so a fit holding a NaN equals no
 * fit, itself included. Callers without a fit's metadata pass zeros. */
typedef struct {
    double r, m, k;
} wv_branch;
typedef struct {
    double beta0;
    uint32_t n_branches;
    const wv_branch* branches;
    double band[2];
    double rms_error, max_error;
} wv_admittance;

/*
 * A box room [0,size] with rectangular floor-to-ceiling pillars: the rooms
 * FdtdGrid::build_dx (fdtd.rs:289-377) voxelizes for shoeboxes and sketches
 * with rectangular holes. Surfaces: walls x=0, x=size0, y=0, y=size1, floor,
 * ceiling, then the sides facing -x, +x, -y, +y of each pillar. Surface s is
 * split into patch_gu[s] x patch_gv[s] patches over its (u, v) extent (x-walls
 * u=y v=z, y-walls u=x v=z, floor/ceiling u=x v=y, pillar sides as the walls
 * they parallel, over the pillar); patch (iu, iv) uses
 * fits[patch_fit[patch_off[s] + iv*gu + iu]].
 */
typedef struct {
    double size[3];
    uint32_t n_pillars;
    const double* pillars; /* [n_pillars][4] = x0, x1, y0, y1 */
    uint32_t n_fits;
    const wv_admittance* fits;
    const uint32_t* patch_gu;
    const uint32_t* patch_gv;
    const uint32_t* patch_off;
    const uint32_t* patch_fit;
} wv_box_room;

typedef struct wv_built_grid wv_built_grid;
/* fdtd.rs:289-377 for box rooms: same flags, boundary order and slots. */
wv_status wv_grid_build_box(const wv_box_room* room, double c, double dx, int threads,
                            wv_built_grid** out, char* err, size_t err_len);
const wv_grid* wv_built_grid_view(const wv_built_grid* g);
void wv_grid_free(wv_built_grid* g);

/*
 * A general room: the triangles of Room (core/wavey-geom/src/room.rs:133-148)
 * with the surface of each triangle and one admittance fit per surface.
 * prism = 1 is a room extruded from a plan sketch: a cell is inside when
 * 0 < z < height and its (x, y) lies in one of the plan regions
 * (Room::contains, room.rs:151-157, with Region::contains and
 * point_in_polygon, regions.rs:44-47, 237-249); a region is cycle
 * region_cycles[r] (its outer boundary) followed by cycles up to
 * region_cycles[r+1] (its holes), cycle k being the points
 * cycle_xy[2*cycle_off[k] .. 2*cycle_off[k+1]). prism = 0 is a closed mesh:
 * inside by vertical ray parity (column_fill, fdtd.rs:385-422).
 */
typedef struct {
    int32_t prism;
    double height;
    uint32_t n_positions;
    const double* positions;     /* [n_positions][3] */
    uint32_t n_triangles;
    const uint32_t* triangles;   /* [n_triangles][3] vertex indices */
    const uint32_t* tri_surface; /* [n_triangles] surface of each triangle */
    uint32_t n_fits;             /* one fit per surface: > every tri_surface entry */
    const wv_admittance* fits;
    uint32_t n_regions;          /* prism rooms only */
    const uint32_t* region_cycles; /* [n_regions + 1] */
    const uint32_t* cycle_off;     /* [region_cycles[n_regions] + 1] */
    const double* cycle_xy;
} wv_room_mesh;

/* A rigid convex obstacle (a speaker cabinet): the points with
 * n . p <= d for every plane (nx, ny, nz, d) inside [lo, hi]
 * (Solid, fdtd.rs:424-461). */
typedef struct {
    uint32_t n_planes;
    const double* planes; /* [n_planes][4] */
    double lo[3], hi[3];
} wv_solid;

/* Solid::from_mesh (fdtd.rs:433-456): a convex triangle mesh with outward
 * winding, vertices turned by rot (row-major 3x3, applied as rot * v) and
 * moved by pos. planes_out holds up to n_tris planes; the distinct planes
 * are written and counted in *n_planes. WV_ERR_INVALID for no triangles. */
wv_status wv_solid_from_mesh(const double* verts, uint32_t n_verts, const uint32_t* tris,
                             uint32_t n_tris, const double rot[9], const double pos[3],
                             double* planes_out, uint32_t* n_planes, double lo[3], double hi[3]);
/* FdtdGrid::build_with's spacing (fdtd.rs:282-286): c / (f_max * ppw), made
 * finer where a solid is less than 8 cells across its smallest side. */
double wv_grid_spacing(double c, double f_max, double ppw, const wv_solid* solids, size_t n_solids);
/* FdtdGrid::build_dx (fdtd.rs:289-377) for a general room with cabinets:
 * same flags, boundary order and slots. The surface behind a missing face
 * is the nearest triangle hit within 1.5 dx along the face's axis; a miss
 * or a solid neighbour leaves the face rigid. */
wv_status wv_grid_build_room(const wv_room_mesh* room, const wv_solid* solids, size_t n_solids,
                             double c, double dx, int threads, wv_built_grid** out, char* err,
                             size_t err_len);

/* ---------------------------------------------------- source / receivers */

/* fdtd.rs:601-610 */
void wv_source_signal(double f_max, double dt, size_t steps, double* out);
/* fdtd.rs:245-272; returns 0 and *cell, or WV_ERR_INVALID when none (None). */
wv_status wv_nearest_air(const wv_grid* g, const double p[3], uint64_t* cell);
/* fdtd.rs:497-512: where a driver on a cabinet's baffle meets the air, from
 * p (inside the cabinet) along axis; p itself when no air is met within 2 m. */
void wv_baffle_point(const wv_grid* g, const double p[3], const double axis[3], double out[3]);
/* fdtd.rs:516-533 (FdtdGrid::air_along): the first air node along axis from
 * p (steps of dx/2, up to 2 m), else the nearest air node; WV_ERR_INVALID
 * when neither exists (None). */
wv_status wv_air_along(const wv_grid* g, const double p[3], const double axis[3], uint64_t* cell);
/* run.rs:601-624: a wave-job source's runs. The source node: with a cabinet
 * calibration (calibrated = 1) the first air node along the axis, else the
 * nearest air node, else the first along the axis. A calibrated source, or
 * one of order 0, is one component: the trilinear taps (taps_at) of the
 * driver point (calibrated: baffle_point in front of the acoustic centre;
 * else the centre), or the node alone. Order 1: source_components of the
 * node, or the node alone when it is too close to a surface
 * (*omni_fallback = 1, the reference's note). *n_comps = 1: comp_len[0]
 * (<= 8) taps from comps[0]; *n_comps = 4: component c at comps[2c] ..
 * comps[2c + comp_len[c]) as wv_source_components. WV_ERR_INVALID when the
 * source is not inside the grid (SolverError::Invalid). */
wv_status wv_wave_source_taps(const wv_grid* g, const double center[3], const double axis[3], int calibrated,
                              uint32_t order, wv_tap comps[8], uint32_t comp_len[4], uint32_t* n_comps,
                              int32_t* omni_fallback);
/* fdtd.rs:467-492; *n_taps <= 8; WV_ERR_INVALID when None. */
wv_status wv_taps_at(const wv_grid* g, const double p[3], wv_tap taps[8], size_t* n_taps);
/* fdtd.rs:730-743: comps[4][2], comp_len[4], *n_comps = 1 or 4. None -> WV_ERR_INVALID. */
wv_status wv_source_components(const wv_grid* g, uint64_t c, uint32_t order, wv_tap comps[8],
                               uint32_t comp_len[4], uint32_t* n_comps);
/* fdtd.rs:753-760: the 400 directions (unit vectors, the speaker's local
 * frame) at which fit_first_order samples the directivity: out [400][3]. */
void wv_first_order_directions(double out[1200]);
/* fdtd.rs:749-807 (fit_first_order): D(u) ~ a0 + a.u with u in room
 * coordinates, from the complex directivity d [400] (re, im) at those
 * directions and d_axis on axis (+x local); to_room row-major, applied as
 * to_room * v. Writes a0 [2], a [3][2]; returns the relative RMS residual. */
double wv_fit_first_order(const double* d, const double d_axis[2], const double to_room[9], double a0[2],
                          double a[6]);
/* fdtd.rs:812-815; complex values interleaved (re, im): a0[2], a[6], out[8]. */
void wv_component_weights(const double a0[2], const double a[6], double k, double dx, double out[8]);
/* fdtd.rs:819-840; spectra/h0/out complex interleaved. */
void wv_to_transfer(const double* spectra, size_t n_spec, const double* signal, size_t steps,
                    double dt, double dx, const double* freqs, size_t nf, double c, double r_ref,
                    const double* h0, double* out);
/* fdtd.rs:847-862: power-complementary crossfade of the image-source and
 * wave fields with wave weight w; complex values as (re, im). */
void wv_blend(const double ism[2], const double wave[2], double w, double out[2]);

/* ------------------------------------------------------------- materials */

/* impedance_fit.rs:48-55; out complex (re, im). */
void wv_admittance_eval(const wv_admittance* fit, double f, double out[2]);
/* impedance_fit.rs:73-84 */
double wv_paris(double beta_re, double beta_im, uint32_t n);
/* material.rs:96-111 */
double wv_fit_real_impedance(double alpha);

/* spectrum.rs:6-44 (Spectrum): n = 0 is the single value values[0], else a
 * table over freqs_hz interpolated linearly in ln f and held outside it. */
typedef struct {
    uint32_t n;
    const double* freqs_hz;
    const double* values;
} wv_spectrum;
/* material.rs:7-31 (Material), the fields the absorption reads: random-
 * incidence absorption, or a normalized impedance (both parts non-NULL). */
typedef struct {
    wv_spectrum absorption;
    const wv_spectrum* impedance_re;
    const wv_spectrum* impedance_im;
} wv_material;
/* material.rs:72-78 (Material::absorption_at). */
double wv_material_absorption(const wv_material* m, double f);
/* impedance_fit.rs:34-44 (AdmittanceFit) with at most two branches. */
typedef struct {
    double beta0;
    uint32_t n_branches;
    wv_branch branches[2];
    double band[2];
    double rms_error, max_error;
} wv_fit_result;
/* impedance_fit.rs:158-235 (fit_admittance): the admittance whose Paris
 * absorption follows the material's over band_lo..band_hi, with up to
 * max_branches (<= 2) resonant branches; fdtd.rs:213-222 (fit_surfaces)
 * calls it per distinct material with max_branches = 2. */
wv_status wv_fit_admittance(const wv_material* material, double band_lo, double band_hi,
                            uint32_t max_branches, wv_fit_result* out);

/* Library version string. */
const char* wv_version(void);

#ifdef __cplusplus
}
#endif
#endif
