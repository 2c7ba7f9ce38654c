// Device plan of one FDTD run (the grid's device layout, a run's state and
// its stream) and the helpers the plain and batched entry points share.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "tri_kernels.h"
#include "step_kernels.h"
#include "wavey_fdtd.h"
#include "wv_internal.h"

namespace wv {

struct DevErr {
    std::string msg;
    wv_status code;
};

#define WV_HIP(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw DevErr{std::string(#expr) + ": " + hipGetErrorString(e_),                      \
                         e_ == hipErrorOutOfMemory ? WV_ERR_NOMEM : WV_ERR_DEVICE};               \
    } while (0)

// A bump arena over device memory the library already holds (the
// allocation-free entry point's static array). While one is installed on the
// calling thread, DBuf::alloc carves 256-byte-aligned pieces from it instead
// of calling hipMalloc, and DBuf::release gives nothing back.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
};
inline thread_local Arena* t_arena = nullptr;

// Device allocation: owned (hipMalloc / hipFree), or a piece of the arena
// installed when it was allocated.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool owned = false;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        owned = false;
    }
    void alloc(size_t n) {
        if (n <= bytes && p) return;
        release();
        size_t b = std::max<size_t>(n, 256);
        if (Arena* a = t_arena) {
            const size_t off = (a->used + 255) & ~(size_t)255;
            if (off > a->cap || b > a->cap - off)
                throw DevErr{"the run needs more than the " + std::to_string(a->cap >> 20) +
                                 " MiB of device scratch the allocation-free entry point holds",
                             WV_ERR_LIMIT};
            p = a->base + off;
            a->used = off + b;
            bytes = b;
            return;
        }
        WV_HIP(hipMalloc(&p, b));
        bytes = b;
        owned = true;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

}  // namespace wv

struct wv_plan {
    using DBuf = wv::DBuf;
    int device = 0;
    hipStream_t stream = nullptr;
    int prec = WV_FP32;
    uint32_t nx = 0, ny = 0, nz = 0, pitch = 0, sz = 0;
    uint64_t npitched = 0, cells = 0, air = 0, nb = 0, ns = 0, ns_k = 0;
    double dx = 0, dt = 0;
    uint32_t zchunk = 0;      // requested (0 = default)
    uint32_t run_zchunk = 0;  // the current run's layout
    uint32_t nzb = 0;         // 32-plane blocks of the z-packed bitmaps
    int variant = -1;                          // requested (-1 = default by size)
    int run_variant = wv::kDefaultVariant;     // the current run's
    uint32_t max_slots = 0, n_sets = 0;

    DBuf flags, force_base, force, zbits, dbits, counts, offs;
    DBuf b_cell_sorted, b_set_sorted;
    DBuf set_n, set_off, set_c, st_u, st_g;  // set_c: {inv_d, cv, k, w} per slot
    DBuf fa, fb, fc, fd;  // fields; fc, fd only for two-step passes
    DBuf sp_rec, sp_tapw;
    DBuf tap_cell, tap_w, recv, ir, freqs, spec;
    uint32_t n_special = 0, sp_pad = 0, seg_stride = 0;
    // Multi-step passes: steps per pass (1, 2 or 3; 0 = default), their
    // z-chunk, the special cells' pitched indices, the ring (interior, not
    // deep) and, for three steps, the deep2 bitmap and the interior cells
    // that are not deep2.
    uint32_t steps_per_pass = 0, run_steps_per_pass = 1;
    uint32_t pair_zchunk = 0;
    DBuf sp_cell, ring_cell, ring_counts, ring_offs;
    uint32_t n_ring = 0;
    DBuf fe, d2bits, ring2_cell, ring2_counts, ring2_offs;  // fe: the fifth field buffer
    DBuf scan_tmp;  // chunk sums and bases of the compaction scans
    // FP32 re-centring: the run's offset (double), the pressure's mean at
    // the last check, the offsets in force from each check on (d0h[j],
    // j = offset_index(n), d0h[0] = 0), the mean's block partial sums and the
    // last shift (float), and the step last checked at.
    DBuf d0, prev_mean, d0h, mean_part, shift;
    size_t recentred_at = 0;
    uint32_t n_ring2 = 0;
    // Which field buffer (0..4: fa, fb, fc, fd, fe) holds p and which prev.
    int buf_p = 0, buf_q = 1;

    // Current run.
    bool active = false;
    size_t steps = 0, done = 0;
    uint32_t n_recv = 0;
    std::vector<double> signal;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    double last_ms = 0.0;
    uint64_t last_steps = 0, last_launches = 0;

    // The stream and events belong to someone else (the allocation-free
    // entry point's per-device set): plan_setup creates none, and they
    // outlive the plan.
    bool borrowed = false;

    size_t tsize() const { return prec == WV_FP64 ? 8 : 4; }
    size_t ssize() const { return prec == WV_FP32 ? 4 : 8; }
    uint32_t lane_cells() const { return (uint32_t)(16 / tsize()); }
    ~wv_plan() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (borrowed) return;
        for (auto e : {ev[0], ev[1], t0, t1})
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace wv {

struct Layout {
    uint32_t zchunk, bx, n_tiles, nzc, seg;
    uint64_t nwaves, positions;
};

inline Layout layout_of(const wv_plan* pl) {
    Layout L;
    L.zchunk = pl->run_zchunk;
    L.bx = kTileX * pl->lane_cells();
    L.n_tiles = div_up(pl->sz, L.bx);
    L.nzc = div_up(pl->nz - 2, L.zchunk);
    L.seg = (L.zchunk * L.bx + kBlock - 1) / kBlock * kBlock;
    L.nwaves = (uint64_t)L.nzc * L.n_tiles;
    L.positions = L.nwaves * L.seg;
    return L;
}

template <class T, class S>
inline StepParams<T, S> make_params(wv_plan* pl, uint32_t tiles_per_block) {
    StepParams<T, S> a{};
    const Layout L = layout_of(pl);
    a.zbits = pl->zbits.as<uint32_t>();
    a.nx = pl->nx;
    a.ny = pl->ny;
    a.nz = pl->nz;
    a.pitch = pl->pitch;
    a.sz = pl->sz;
    a.n_tiles = L.n_tiles;
    a.tile_blocks = div_up(L.n_tiles, tiles_per_block);
    a.zchunk = L.zchunk;
    a.main_blocks = a.tile_blocks * L.nzc;
    a.seg_base = pl->offs.as<uint32_t>();
    a.seg_stride = pl->seg_stride;
    a.sp_rec = pl->sp_rec.as<uint2>();
    a.sp_tapw = pl->sp_tapw.as<S>();
    a.sp_pad = pl->sp_pad;
    a.set_c = pl->set_c.as<S>();
    a.st_u = pl->st_u.as<S>();
    a.st_g = pl->st_g.as<S>();
    a.has_k = pl->ns_k > 0;
    a.max_slots = pl->max_slots;
    a.n_recv = pl->n_recv;
    a.recv = pl->recv.as<uint32_t>();
    a.d0 = pl->d0.as<double>();
    const double l2 = kCourant * kCourant;
    if (sizeof(T) == 4) {
        // FP32 lambda^2: 1/3 rounded toward zero, so 6 l2 <= 2 (the scheme stays
        // within its stability limit), and c2 = 2 - 6 l2 exactly, so a constant
        // field is still preserved.
        float l2f = (float)l2;
        if ((double)l2f > l2) l2f = std::nextafter(l2f, 0.0f);
        a.l2 = (T)l2f;
        a.c2 = (T)(2.0 - 6.0 * (double)l2f);
        a.l2s = (S)(sizeof(S) == 4 ? (double)l2f : l2);
    } else {
        a.l2 = (T)l2;
        a.c2 = (T)(2.0 - 6.0 * l2);
        a.l2s = (S)l2;
    }
    a.lam = (S)kCourant;
    a.half_dt = (S)(0.5 * pl->dt);
    return a;
}

// The z-chunk of a multi-step pass: the requested one, else the default by
// grid size and mode. Chunks must lie within the 32-plane blocks of the
// bitmaps, except FP32 three-step passes, which hold two bitmap words per
// lane and take up to 64 planes: by default the longest of 64, 32 and 16
// that still gives 640 tile blocks (2.5 per CU; the hall takes 64, a 58 Mi
// cell grid 32, a 29 Mi one 16). Other passes take 32 planes from 64 Mi
// cells. Below those, 16 planes, or 8 where 16 leave fewer than 150 tile
// blocks (on rooms of 0.8-7.5 M cells: below 150 blocks 8 planes were
// faster on all but one, by up to 31 %; from 150 blocks 16 were faster on
// all but one, by up to 11 %; 4 planes lost to 8 even at 270 blocks
// against 135).
constexpr uint64_t kPassMinBlocks = 150;
inline uint32_t pass_zchunk(const wv_plan* pl, uint32_t steps_per_pass) {
    const bool wide = steps_per_pass == 3 && pl->prec != WV_FP64;
    if (pl->pair_zchunk) return std::min(pl->pair_zchunk, wide ? 64u : 32u);
    if (!wide && pl->cells >= kPairLongChunkCells) return 32u;
    const uint64_t stride = pl->prec == WV_FP64 ? kTileStep<double> : kTileStep<float>;
    const uint64_t cols = std::max<uint64_t>(1, pl->nx > 4 ? (pl->nx - 4 + stride - 1) / stride : 1);
    const uint64_t rows = (pl->ny + (steps_per_pass == 3 ? kTriRows : kPairRows) - 1) /
                          (steps_per_pass == 3 ? kTriRows : kPairRows);
    auto blocks = [&](uint64_t zc) { return cols * rows * ((pl->nz - 2 + zc - 1) / zc); };
    if (wide) {
        for (uint32_t zc : {64u, 32u}) {
            if (blocks(zc) >= 640) return zc;
        }
    }
    return blocks(16) >= kPassMinBlocks ? 16u : 8u;
}

template <class F>
inline wv_status guarded(char* err, size_t err_len, F f) {
    try {
        return f();
    } catch (const DevErr& e) {
        set_err(err, err_len, e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_err(err, err_len, "host allocation failed");
        return WV_ERR_NOMEM;
    } catch (const std::exception& e) {
        // A host-side failure (a thread that cannot start): the caller falls back.
        set_err(err, err_len, std::string("internal error: ") + e.what());
        return WV_ERR_DEVICE;
    }
}

// Builds the plan's device copy of a grid (fields, flags, boundary marks,
// slot-set table). Throws DevErr.
void plan_setup(wv_plan* pl, const wv_grid* g);
// Starts a run on a plan: taps, signal, receivers; classification, special
// cell records, zeroed fields and state. Throws DevErr.
void run_begin(wv_plan* pl, const wv_tap* taps, size_t n_taps, const double* signal, size_t steps,
               const uint64_t* receivers, size_t n_recv);
// Returns the plan's begun run to step 0 (fields, records, slot state and
// re-centring offsets zeroed) without rebuilding its layout. Throws DevErr.
void run_rewind(wv_plan* pl);
// Enqueues steps [n, stop) of the plan's current run on `stream` (two-step
// passes where the run has them) without waiting; the plan's buf_p / buf_q
// then name the buffers of p^stop and p^(stop-1). Returns the launches.
// Throws DevErr.
uint64_t enqueue_steps(wv_plan* pl, size_t n, size_t stop, hipStream_t stream);
// FP32 re-centring. A room that keeps a mean pressure after the source
// pulse (walls that hardly absorb at DC) holds every FP32 cell near that
// mean, so the stored field resolves the waves on top of it, and the
// velocity p^n - p^{n-1}, only in steps of the mean's ulp, and the rounding
// that accumulates in the room's mean-pressure mode grows with the run.
// At steps 128, 256, 512, ... (powers of two) the mean of p^n over the air
// cells is taken (a fixed-order FP64 sum, so every pass mode and batch gets
// the same value). The first check records the pressure's mean; from the
// second on, while that mean keeps at least half of its value at the last
// check (a room that holds it), the stored mean moves from both fields
// (p^n and p^{n-1} less it, in FP32, every cell) into the run's offset d0,
// and once it decays (a room that absorbs it) the offset moves back into
// the fields. The interior update does not depend on a uniform shift (it
// is formed from differences); the boundary slots and the records add d0
// back. FP64 runs are never re-centred.
inline bool recentres_at(const wv_plan* pl, size_t n) {
    return pl->prec != WV_FP64 && n >= 128 && (n & (n - 1)) == 0;
}
// The first re-centring step after n (SIZE_MAX for FP64): multi-step passes
// end there.
inline size_t next_recentre(const wv_plan* pl, size_t n) {
    if (pl->prec == WV_FP64) return ~(size_t)0;
    size_t m = 128;
    while (m <= n) m <<= 1;
    return m;
}
// Enqueues the re-centring of step n on `stream` (p^n and p^{n-1} in the
// plan's buffers buf_p, buf_q), once per step of a run. Throws DevErr.
void recentre(wv_plan* pl, size_t n, hipStream_t stream);

// Spectra (FP64 DFT of the record) and/or the receiver series of the run's
// completed steps into caller buffers. Throws DevErr.
void run_finish(wv_plan* pl, const double* freqs, size_t nf, int mode, double* spectra_out, double* ir_out);
wv_status check_grid(const wv_grid* g, char* err, size_t err_len);
// Device bytes a plan holds (every buffer), and an estimate, before any
// allocation, of what a plan of this grid and run will hold (special cells
// counted as boundary nodes plus taps).
uint64_t plan_bytes(const wv_plan* pl);
uint64_t estimate_bytes(const wv_grid* g, int prec, size_t steps, size_t n_recv, size_t n_taps,
                        uint32_t steps_per_pass = 0);
// Checks of wv_run_begin's arguments against a plan (tap planes, receivers,
// record size); WV_OK or an error with its message.
wv_status check_run_args(const wv_plan* pl, const wv_tap* taps, size_t n_taps, const double* signal, size_t steps,
                         const uint64_t* receivers, size_t n_recv, char* err, size_t err_len);

}  // namespace wv
