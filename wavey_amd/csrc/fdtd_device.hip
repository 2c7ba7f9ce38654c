// MI355X (gfx950) FDTD wave-solver run: device layout, kernels, run loop.
//
// The scheme is Wavey's (core/wavey-solver/src/fdtd.rs:1-30): the 3D
// leapfrog p+ = (2 - K l2) p + l2 sum_air p_nb - p-, with locally reacting
// frequency-dependent boundaries solved per node in closed form
// (fdtd.rs:581-598). The layout is this build's own:
//
//  * Fields p, prev live in a pitched grid: row pitch = nx rounded up to 64
//    cells, so a wave's 16-byte-per-lane accesses cover aligned row segments.
//  * One launch per time step. A work-group owns a tile of the xy plane
//    (256 FP32 or 128 FP64 cells by TY rows, one 64-lane wave per row, 16
//    bytes per lane) and marches a z-chunk; the column's p at z-1 and z stay
//    in registers. Per plane a lane reads p at z+1, prev and the two
//    y-neighbour rows, and writes prev: 12 bytes per FP32 cell.
//  * Two bitmaps, packed along z (one 32-bit word per column per 32 planes,
//    read once per chunk), say which cells are plain interior air (six air
//    neighbours, no boundary slots, no source tap) and which are "special"
//    (every other air cell, and taps). Special cells are updated by the same
//    lane from the neighbour values it already holds: their metadata sit in
//    a list compacted (wave ballot + prefix sums) in the order the waves
//    consume them, so a lane finds its cell's entry from the bitmap alone.
//  * Boundary slot constants are deduplicated into a small table of slot
//    sets; slot state is stored slot-major, state[j][rank], so the reads of a
//    wave's special cells are contiguous.
//  * Receivers are recorded by extra blocks of the same launch. Within a step
//    p is only read and each prev cell is written by one thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "wavey_fdtd.h"
#include "wv_internal.h"

using namespace wv;

namespace {

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

// Owned device allocation.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        if (n <= bytes && p) return;
        release();
        size_t b = std::max<size_t>(n, 256);
        WV_HIP(hipMalloc(&p, b));
        bytes = b;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

constexpr uint32_t kBlock = 256;          // setup kernels and compaction blocks
constexpr uint32_t kTileX = 64;           // one 64-lane wave per tile row
constexpr uint32_t kStepsPerBlock = 128;  // cancel/progress granularity, as gpu_fdtd.rs:35
constexpr uint32_t kSubPlanes = 4;        // planes per special-cell pass of a z-chunk

// Special-cell record: {meta, first constant of its slot set}. Meta bits:
// 0-5 air neighbours (-x, +x, -y, +y, -z, +z), 6 solid, 7 tap, 8-13 slot
// count, 14-26 position in its wave's chunk (dz * tile width + x in tile).
constexpr uint32_t kMetaSolid = 1u << 6;  // a tap on a non-air cell: injection only
constexpr uint32_t kMetaTap = 1u << 7;    // has a tap weight
constexpr uint32_t kMetaSlotShift = 8;
constexpr uint32_t kMetaPosShift = 14;
constexpr uint32_t kMaxSlots = 63;

// ------------------------------------------------------------------ kernels

template <typename T, typename S>
struct StepParams {
    const T* __restrict__ p;
    T* __restrict__ prev;
    const uint32_t* __restrict__ zbits;  // interior bits [nzb][ny][pitch]
    uint32_t nx, ny, nz, pitch, sz;
    uint32_t tiles_x, tiles_y, zchunk, main_blocks;
    // Special cells, in wave order: wave w owns [seg_base[w * seg_stride],
    // seg_base[(w + 1) * seg_stride]).
    const uint32_t* __restrict__ seg_base;
    uint32_t seg_stride;
    const uint2* __restrict__ sp_rec;
    const S* __restrict__ sp_tapw;
    uint32_t sp_pad;  // row length of the slot-major state arrays
    const S* __restrict__ set_inv_d;
    const S* __restrict__ set_cv;
    const S* __restrict__ set_k;
    const S* __restrict__ set_w;
    S* __restrict__ st_u;  // [slot][rank]
    S* __restrict__ st_g;
    uint32_t has_k;
    uint32_t n_recv;
    const uint32_t* __restrict__ recv;
    T* __restrict__ ir;  // this step's receiver row
    T l2, c2;            // interior: p+ = c2 p + l2 sum - p-
    S l2s, lam, half_dt, s_n;
};

template <typename T>
__device__ __forceinline__ T interior_update(T c, T q, T xm, T xp, T ym, T yp, T zm, T zp, T l2, T c2) {
    if constexpr (sizeof(T) == 8) {
        // FP64: the reference's form, neighbour order -x, +x, -y, +y, -z, +z
        // (fdtd.rs:685-706).
        T s = xm;
        s += xp;
        s += ym;
        s += yp;
        s += zm;
        s += zp;
        return c2 * c + (l2 * s - q);
    } else {
        // FP32: the same update as p + (v + l2 * sum (p_nb - p)) with
        // v = p - p-. Neighbour differences and v are exact for close values
        // (Sterbenz), so a near-uniform field rounds only in the final add
        // instead of in l2 * sum ~ 2p: no correlated rounding drives the
        // room's DC mode.
        const T d = ((xm - c) + (xp - c)) + ((ym - c) + (yp - c)) + ((zm - c) + (zp - c));
        return c + ((c - q) + l2 * d);
    }
}

// Lane i takes v from lane i-1 (wave_shr:1); lane 0 keeps `fill`.
__device__ __forceinline__ int dpp_from_lower(int v, int fill) {
    return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xF, 0xF, false);
}
// Lane i takes v from lane i+1 (wave_shl:1); lane 63 keeps `fill`.
__device__ __forceinline__ int dpp_from_upper(int v, int fill) {
    return __builtin_amdgcn_update_dpp(fill, v, 0x130, 0xF, 0xF, false);
}
template <typename T>
__device__ __forceinline__ T from_lower_lane(T v, T fill) {
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(dpp_from_lower(__float_as_int(v), __float_as_int(fill)));
    } else {
        const long long a = __double_as_longlong(v), f = __double_as_longlong(fill);
        const int lo = dpp_from_lower((int)a, (int)f), hi = dpp_from_lower((int)(a >> 32), (int)(f >> 32));
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    }
}
template <typename T>
__device__ __forceinline__ T from_upper_lane(T v, T fill) {
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(dpp_from_upper(__float_as_int(v), __float_as_int(fill)));
    } else {
        const long long a = __double_as_longlong(v), f = __double_as_longlong(fill);
        const int lo = dpp_from_upper((int)a, (int)f), hi = dpp_from_upper((int)(a >> 32), (int)(f >> 32));
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    }
}

// W consecutive cells of a row per lane, moved as one 16-byte access.
template <typename T>
struct Cells {
    static constexpr int W = 16 / sizeof(T);
    T v[W];
    __device__ __forceinline__ void load(const T* __restrict__ p) {
        if constexpr (W == 4) {
            const float4 q = *reinterpret_cast<const float4*>(p);
            v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        } else {
            const double2 q = *reinterpret_cast<const double2*>(p);
            v[0] = q.x; v[1] = q.y;
        }
    }
    __device__ __forceinline__ void store(T* __restrict__ p) const {
        if constexpr (W == 4) *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
        else *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
    }
};

template <int W>
__device__ __forceinline__ void load_words(const uint32_t* __restrict__ p, uint32_t (&w)[W]) {
    if constexpr (W == 4) {
        const uint4 q = *reinterpret_cast<const uint4*>(p);
        w[0] = q.x; w[1] = q.y; w[2] = q.z; w[3] = q.w;
    } else {
        const uint2 q = *reinterpret_cast<const uint2*>(p);
        w[0] = q.x; w[1] = q.y;
    }
}

template <int W>
__device__ __forceinline__ uint32_t plane_bits(const uint32_t (&w)[W], uint32_t sh) {
    uint32_t b = 0;
#pragma unroll
    for (int e = 0; e < W; ++e) b |= ((w[e] >> sh) & 1u) << e;
    return b;
}

// Constants and state of slots jb .. jb+G-1 of a node (indices clamped to
// the last slot; the caller masks the extras).
template <typename S, uint32_t G>
struct SlotGroup {
    S inv_d[G], w[G], cv[G], u[G], kk[G], g[G];
    template <typename T>
    __device__ __forceinline__ void load(const StepParams<T, S>& a, uint32_t c0, uint32_t rank, uint32_t jb,
                                         uint32_t ns) {
#pragma unroll
        for (uint32_t k = 0; k < G; ++k) {
            const uint32_t j = min(jb + k, ns - 1);
            inv_d[k] = a.set_inv_d[c0 + j];
            w[k] = a.set_w[c0 + j];
            cv[k] = a.set_cv[c0 + j];
            u[k] = a.st_u[j * a.sp_pad + rank];
            kk[k] = 0;
            g[k] = 0;
            if (a.has_k) {
                kk[k] = a.set_k[c0 + j];
                g[k] = a.st_g[j * a.sp_pad + rank];
            }
        }
    }
    // The part of slot k's u+ that does not depend on p+ (fdtd.rs:589-592).
    __device__ __forceinline__ S rest(uint32_t k, S pc, S pp) const {
        return (cv[k] * u[k] + S(0.5) * pc + w[k] * pp - kk[k] * g[k]) * inv_d[k];
    }
    __device__ __forceinline__ void sums(uint32_t jb, uint32_t ns, S pc, S pp, S& s1, S& s2) const {
#pragma unroll
        for (uint32_t k = 0; k < G; ++k)
            if (jb + k < ns) {
                s1 += (S(0.5) - w[k]) * inv_d[k];
                s2 += rest(k, pc, pp) - u[k];
            }
    }
    template <typename T>
    __device__ __forceinline__ void advance(const StepParams<T, S>& a, uint32_t rank, uint32_t jb, uint32_t ns, S pc,
                                            S pp, S next) const {
#pragma unroll
        for (uint32_t k = 0; k < G; ++k)
            if (jb + k < ns) {
                const uint32_t j = jb + k;
                const S un = (S(0.5) - w[k]) * inv_d[k] * next + rest(k, pc, pp);
                if (a.has_k) a.st_g[j * a.sp_pad + rank] = g[k] + a.half_dt * (un + u[k]);
                a.st_u[j * a.sp_pad + rank] = un;
            }
    }
};

// One special cell: the air-neighbour Laplacian, the boundary solve of
// fdtd.rs:581-598 when it has slots, and its tap (injection after the
// update, fdtd.rs:718-720). nb = -x, +x, -y, +y, -z, +z; c0 is the first
// constant of the node's slot set and rank its index in the state arrays.
template <typename T, typename S>
__device__ __forceinline__ S special_update(const StepParams<T, S>& a, uint32_t meta, uint32_t c0, uint32_t rank,
                                            S pc, S pp, const S (&nb)[6]) {
    S next;
    if (meta & kMetaSolid) {
        next = pp;  // a tap on a solid cell: the reference only injects there
    } else {
        S lap;
        if constexpr (sizeof(S) == 8) {
            // fdtd.rs:685-696: l2 (sum_air - k p), neighbours in -x..+z order.
            S sum = 0, k = 0;
#pragma unroll
            for (int d = 0; d < 6; ++d)
                if ((meta >> d) & 1u) {
                    sum += nb[d];
                    k += 1;
                }
            lap = a.l2s * (sum - k * pc);
        } else {
            // FP32: sum of neighbour differences (see interior_update).
            S d = 0;
#pragma unroll
            for (int e = 0; e < 6; ++e)
                if ((meta >> e) & 1u) d += nb[e] - pc;
            lap = a.l2s * d;
        }
        const uint32_t ns = (meta >> kMetaSlotShift) & kMaxSlots;
        if (ns == 0) {
            if constexpr (sizeof(S) == 8) next = S(2) * pc - pp + lap;
            else next = pc + ((pc - pp) + lap);
        } else {
            // Each slot's u+ is linear in p+, so p+ is solved first from the
            // summed face flux, then the states advance. The first G slots
            // stay in registers (one pass for nodes with up to G slots);
            // further groups are read again for the second pass. G = 3 in
            // FP32 (a wall face with a three-branch admittance) keeps the
            // kernel within 64 VGPRs, 8 waves per SIMD.
            constexpr uint32_t G = sizeof(S) == 4 ? 3 : 2;
            SlotGroup<S, G> g0;
            g0.load(a, c0, rank, 0, ns);
            S s1 = 0, s2 = 0;
            g0.sums(0, ns, pc, pp, s1, s2);
            for (uint32_t jb = G; jb < ns; jb += G) {
                SlotGroup<S, G> gj;
                gj.load(a, c0, rank, jb, ns);
                gj.sums(jb, ns, pc, pp, s1, s2);
            }
            if constexpr (sizeof(S) == 8) {
                next = (S(2) * pc - pp + lap - a.lam * s2) / (S(1) + a.lam * s1);
            } else {
                // The same solve written as pc + (small terms).
                next = pc + ((pc - pp) + lap - a.lam * (s2 + s1 * pc)) / (S(1) + a.lam * s1);
            }
            g0.advance(a, rank, 0, ns, pc, pp, next);
            for (uint32_t jb = G; jb < ns; jb += G) {
                SlotGroup<S, G> gj;
                gj.load(a, c0, rank, jb, ns);
                gj.advance(a, rank, jb, ns, pc, pp, next);
            }
        }
    }
    if (meta & kMetaTap) next += a.sp_tapw[rank] * a.s_n;
    return next;
}

// A tile of 64*W cells (x) by TY rows (y), one 64-lane wave per row, each
// lane W consecutive cells (16 bytes: 4 FP32 or 2 FP64), marching z over
// [z0, z1). Per plane a lane loads p at z+1, prev, the two y-neighbour rows
// and, on lanes 0 and 63, the row-end cells of the neighbour tiles; x
// neighbours inside a lane are in its registers and move across lanes by
// wave shifts. All of a plane's loads are issued before any is used and the
// loop keeps few registers live, so many waves per SIMD hide the memory
// latency. Interior cells are written here; the wave then updates the
// special cells of its chunk, 64 at a time, while the rows they read are
// still in cache.
template <typename T, typename S, int TY>
__device__ __forceinline__ void tile_part(const StepParams<T, S>& a, uint32_t t) {
    constexpr int W = Cells<T>::W;
    constexpr uint32_t BX = kTileX * W;
    const uint32_t tx = t % a.tiles_x;
    const uint32_t r = t / a.tiles_x;
    const uint32_t ty = r % a.tiles_y;
    const uint32_t zc = r / a.tiles_y;
    const uint32_t lane = threadIdx.x & (kTileX - 1);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTileX);
    const uint32_t xl = tx * BX + lane * W;  // first cell of this lane
    const uint32_t y = ty * TY + wave;
    if (y >= a.ny) return;
    const uint32_t w_seg = ((zc * a.tiles_x + tx) * a.ny + y) * a.seg_stride;
    // Lanes past the row pitch (the last tile of a row) re-read the row's
    // last 16 bytes and update nothing: no traffic beyond the row.
    const bool live = xl < a.pitch;
    const uint32_t x = live ? xl : a.pitch - W;
    const uint32_t z0 = 1 + zc * a.zchunk;
    const uint32_t z1 = min(z0 + a.zchunk, a.nz - 1);
    const uint32_t col = y * a.pitch + x;
    const uint32_t ym_off = y > 0 ? a.pitch : 0u;
    const uint32_t yp_off = y + 1 < a.ny ? a.pitch : 0u;
    const uint32_t hl_off = x > 0 ? 1u : 0u;                     // lane 0: x - 1
    const uint32_t hr_off = x + W < a.pitch ? (uint32_t)W : 0u;  // lane 63: x + W
    const uint32_t full = (1u << W) - 1u;
    const uint32_t lane_mask = live ? full : 0u;
    // The chunk lies inside one 32-plane block of the z-packed bitmap
    // (z-chunks divide 32 and start at plane 1 + a multiple of the chunk).
    uint32_t ib[W];
    load_words<W>(a.zbits + (((z0 - 1) >> 5) * a.ny + y) * a.pitch + x, ib);

    Cells<T> zm, zcur;
    zm.load(a.p + (z0 - 1) * a.sz + col);
    zcur.load(a.p + z0 * a.sz + col);
    // The chunk is marched in sub-chunks of kSubPlanes planes; after each,
    // the wave updates that sub-chunk's special cells while the rows they
    // read are still in cache. Records are sorted by plane, so a sub-chunk's
    // records start at the compaction block of its first plane.
    const uint32_t row0 = z0 * a.sz + y * a.pitch + tx * BX;
    uint32_t s_begin = a.seg_base[w_seg];
    for (uint32_t zs = z0; zs < z1; zs += kSubPlanes) {
        const uint32_t ze = min(zs + kSubPlanes, z1);
        const uint32_t s_end =
            ze == z1 ? a.seg_base[w_seg + a.seg_stride] : a.seg_base[w_seg + (ze - z0) * BX / kBlock];
        // The first special records are requested now, so they have arrived
        // when the sub-chunk's z-march ends.
        uint2 rec = make_uint2(0u, 0u);
        if (s_begin + lane < s_end) rec = a.sp_rec[s_begin + lane];
        for (uint32_t z = zs; z < ze; ++z) {
            const uint32_t i = z * a.sz + col;
            T hl = 0, hr = 0;
            if (lane == 0) hl = a.p[i - hl_off];
            if (lane == kTileX - 1) hr = a.p[i + hr_off];
            Cells<T> zp, q, ym, yp;
            zp.load(a.p + i + a.sz);
            q.load(a.prev + i);
            ym.load(a.p + i - ym_off);
            yp.load(a.p + i + yp_off);
            const T lo = from_lower_lane(zcur.v[W - 1], hl);
            const T hi = from_upper_lane(zcur.v[0], hr);
            Cells<T> out;
#pragma unroll
            for (int e = 0; e < W; ++e) {
                const T xm = e == 0 ? lo : zcur.v[e - 1];
                const T xp = e == W - 1 ? hi : zcur.v[e + 1];
                out.v[e] =
                    interior_update(zcur.v[e], q.v[e], xm, xp, ym.v[e], yp.v[e], zm.v[e], zp.v[e], a.l2, a.c2);
            }
            const uint32_t bits = plane_bits<W>(ib, (z - 1) & 31u) & lane_mask;
            if (bits == full) {
                out.store(a.prev + i);
            } else if (bits) {
                // Row ends next to walls: only the interior cells.
#pragma unroll
                for (int e = 0; e < W; ++e)
                    if ((bits >> e) & 1u) a.prev[i + e] = out.v[e];
            }
            zm = zcur;
            zcur = zp;
        }

        // The sub-chunk's special cells, 64 at a time. A record's position
        // gives the cell; its neighbour loads are unconditional (a masked-out
        // neighbour reads the cell itself, and the mask never points outside
        // the grid).
        for (uint32_t s = s_begin + lane; s < s_end; s += kTileX) {
            uint2 next_rec = make_uint2(0u, 0u);
            if (s + kTileX < s_end) next_rec = a.sp_rec[s + kTileX];
            const uint32_t meta = rec.x;
            const uint32_t pos = meta >> kMetaPosShift;
            const uint32_t i = row0 + (pos / BX) * a.sz + pos % BX;
            const T c = a.p[i];
            const T q = a.prev[i];
            const S nb[6] = {(S)a.p[i - ((meta & 1u) ? 1u : 0u)], (S)a.p[i + ((meta & 2u) ? 1u : 0u)],
                             (S)a.p[i - ((meta & 4u) ? a.pitch : 0u)], (S)a.p[i + ((meta & 8u) ? a.pitch : 0u)],
                             (S)a.p[i - ((meta & 16u) ? a.sz : 0u)], (S)a.p[i + ((meta & 32u) ? a.sz : 0u)]};
            a.prev[i] = (T)special_update(a, meta, rec.y, s, (S)c, (S)q, nb);
            rec = next_rec;
        }
        s_begin = s_end;
    }
}

// One time step: tile blocks, then receiver blocks (p^n before its update,
// fdtd.rs:673-683). Consecutive tiles go to the 8 XCDs round-robin, so
// neighbouring tiles run at the same time and spread over all HBM stacks.
template <typename T, typename S, int TY>
__global__ __launch_bounds__(kTileX* TY) void k_step(StepParams<T, S> a) {
    if (blockIdx.x < a.main_blocks) {
        tile_part<T, S, TY>(a, blockIdx.x);
        return;
    }
    const uint32_t r = (blockIdx.x - a.main_blocks) * blockDim.x + threadIdx.x;
    if (r < a.n_recv) a.ir[r] = a.p[a.recv[r]];
}

// Marks cells that must not take the interior path (boundary nodes, taps).
__global__ void k_mark(uint32_t* __restrict__ force, const uint32_t* __restrict__ cells, uint32_t n,
                       uint32_t nx, uint32_t pitch) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t c = cells[t];
    const uint32_t q = (c / nx) * pitch + c % nx;
    atomicOr(&force[q >> 5], 1u << (q & 31u));
}

// Neighbour bits of cell (x, y, z) that stay inside the grid: flags from a
// caller can never send a read outside the field arrays.
__device__ __forceinline__ uint32_t inside_mask(uint32_t x, uint32_t y, uint32_t z, uint32_t nx, uint32_t ny,
                                                uint32_t nz) {
    return (x > 0 ? 1u : 0u) | (x + 1 < nx ? 2u : 0u) | (y > 0 ? 4u : 0u) | (y + 1 < ny ? 8u : 0u) |
           (z > 0 ? 16u : 0u) | (z + 1 < nz ? 32u : 0u);
}

struct Geo {
    uint64_t npitched;
    uint32_t pitch, nx, ny, nz;
};

struct Cls {
    bool interior, special;
    uint32_t mask;  // air neighbours inside the grid
    bool air;
};

__device__ __forceinline__ Cls classify(const uint8_t* __restrict__ flags, const uint32_t* __restrict__ force,
                                        uint32_t x, uint32_t y, uint32_t z, const Geo& g) {
    Cls c{false, false, 0u, false};
    if (x >= g.nx || y >= g.ny || z >= g.nz) return c;
    const uint64_t row = (uint64_t)z * g.ny + y;
    const uint64_t q = row * g.pitch + x;
    const uint8_t f = flags[row * g.nx + x];
    const bool forced = (force[q >> 5] >> (q & 31u)) & 1u;
    c.air = f & 64u;
    c.mask = f & 63u & inside_mask(x, y, z, g.nx, g.ny, g.nz);
    c.interior = c.air && c.mask == 63u && !forced;
    c.special = (c.air && !c.interior) || (forced && !c.air);
    return c;
}

// The z-packed interior bitmap: word (kb, y, x) holds the bits of planes
// 1 + 32 kb .. 32 + 32 kb of column (x, y), bit k for plane 1 + 32 kb + k.
__global__ __launch_bounds__(kBlock) void k_zbits(const uint8_t* __restrict__ flags, const uint32_t* __restrict__ force,
                                                  uint32_t* __restrict__ zbits, Geo g, uint32_t nzb) {
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (uint64_t)nzb * g.ny * g.pitch) return;
    const uint32_t x = (uint32_t)(t % g.pitch);
    const uint64_t r = t / g.pitch;
    const uint32_t y = (uint32_t)(r % g.ny), kb = (uint32_t)(r / g.ny);
    uint32_t wi = 0;
    for (uint32_t k = 0; k < 32; ++k) {
        const uint32_t z = 1 + kb * 32 + k;
        if (z + 1 >= g.nz) break;
        wi |= (uint32_t)classify(flags, force, x, y, z, g).interior << k;
    }
    zbits[t] = wi;
}

// The wave order of special cells: wave w = (zc * tiles_x + tx) * ny + y owns
// positions [w * seg, (w + 1) * seg); position (w, dz, xi) is cell
// (tx * BX + xi, y, 1 + zc * zchunk + dz). Every wave segment is a whole
// number of compaction blocks.
struct WaveOrder {
    uint32_t bx, zchunk, seg, tiles_x;
};

__device__ __forceinline__ bool wave_cell(uint64_t qw, const WaveOrder& o, const Geo& g, uint32_t* x, uint32_t* y,
                                          uint32_t* z) {
    const uint64_t w = qw / o.seg;
    const uint32_t rr = (uint32_t)(qw % o.seg);
    const uint32_t dz = rr / o.bx, xi = rr % o.bx;
    if (dz >= o.zchunk) return false;
    *y = (uint32_t)(w % g.ny);
    const uint64_t t2 = w / g.ny;
    const uint32_t tx = (uint32_t)(t2 % o.tiles_x), zc = (uint32_t)(t2 / o.tiles_x);
    *x = tx * o.bx + xi;
    *z = 1 + zc * o.zchunk + dz;
    return *x < g.nx && *z + 1 < g.nz;
}

// Pass 1 of the compaction: special cells per 256 positions of the wave order.
__global__ __launch_bounds__(kBlock) void k_sp_count(const uint8_t* __restrict__ flags,
                                                     const uint32_t* __restrict__ force, uint32_t* __restrict__ counts,
                                                     Geo g, WaveOrder o) {
    __shared__ uint32_t wc[kBlock / 64];
    const uint64_t qw = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t x, y, z;
    bool sp = false;
    if (wave_cell(qw, o, g, &x, &y, &z)) sp = classify(flags, force, x, y, z, g).special;
    const uint64_t b = __ballot(sp);
    if ((threadIdx.x & 63u) == 0) wc[threadIdx.x >> 6] = (uint32_t)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t w = 0; w < kBlock / 64; ++w) s += wc[w];
        counts[blockIdx.x] = s;
    }
}

// Exclusive scan in one work-group.
__global__ __launch_bounds__(1024) void k_scan(const uint32_t* __restrict__ counts, uint32_t* __restrict__ offs,
                                               uint32_t n, uint32_t* __restrict__ total) {
    __shared__ uint32_t part[1024];
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t lo = min(threadIdx.x * per, n), hi = min(lo + per, n);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += counts[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t base = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (uint32_t i = lo; i < hi; ++i) {
        offs[i] = base;
        base += counts[i];
    }
    if (threadIdx.x == 1023) *total = part[1023];
}

__device__ __forceinline__ int64_t find_sorted(const uint32_t* __restrict__ a, uint32_t n, uint32_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && a[lo] == key) ? (int64_t)lo : -1;
}

// Pass 2: each special cell's rank is its block's offset + earlier waves +
// earlier lanes (popcount of the ballot below the lane). Writes the
// record (metadata word with the cell's position in its wave's chunk, and
// the first constant of its slot set) and tap weight.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_sp_write(const uint8_t* __restrict__ flags,
                                                     const uint32_t* __restrict__ force,
                                                     const uint32_t* __restrict__ offs, Geo g, WaveOrder o,
                                                     const uint32_t* __restrict__ b_cell_sorted,
                                                     const uint32_t* __restrict__ b_set_sorted, uint32_t nb,
                                                     const uint32_t* __restrict__ set_n,
                                                     const uint32_t* __restrict__ set_off,
                                                     const uint32_t* __restrict__ tap_cell, const S* __restrict__ tap_w,
                                                     uint32_t n_tap, uint2* __restrict__ sp_rec,
                                                     S* __restrict__ sp_tapw) {
    __shared__ uint32_t wc[kBlock / 64];
    const uint64_t qw = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t x = 0, y = 0, z = 0;
    Cls c{false, false, 0u, false};
    if (wave_cell(qw, o, g, &x, &y, &z)) c = classify(flags, force, x, y, z, g);
    const uint64_t b = __ballot(c.special);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) wc[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    if (!c.special) return;
    uint32_t rank = offs[blockIdx.x];
    for (uint32_t w = 0; w < wave; ++w) rank += wc[w];
    rank += (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
    const uint32_t ref = (uint32_t)(((uint64_t)z * g.ny + y) * g.nx + x);
    uint32_t meta = (c.air ? c.mask : kMetaSolid) | ((uint32_t)(qw % o.seg) << kMetaPosShift);
    uint32_t c0 = 0;
    if (c.air) {
        const int64_t bi = find_sorted(b_cell_sorted, nb, ref);
        if (bi >= 0) {
            const uint32_t set = b_set_sorted[bi];
            meta |= set_n[set] << kMetaSlotShift;
            c0 = set_off[set];
        }
    }
    const int64_t t = find_sorted(tap_cell, n_tap, ref);
    S w = 0;
    if (t >= 0) {
        meta |= kMetaTap;
        w = tap_w[t];
    }
    sp_rec[rank] = make_uint2(meta, c0);
    sp_tapw[rank] = w;
}

// Receiver spectra from the recorded series, FP64: sum_n p^n e^{-j 2 pi fract(f dt n)}
// in n order (fdtd.rs:679-682).
template <typename T>
__global__ void k_dft(const T* __restrict__ ir, uint32_t steps, uint32_t n_recv, const double* __restrict__ freqs,
                      uint32_t nf, double dt, double* __restrict__ out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_recv * nf) return;
    const uint32_t r = t / nf, fi = t % nf;
    const double fdt = freqs[fi] * dt;
    double re = 0.0, im = 0.0;
    for (uint32_t n = 0; n < steps; ++n) {
        const double x = fdt * (double)n;
        const double ph = -2.0 * 3.141592653589793 * (x - trunc(x));
        double sv, cv;
        sincos(ph, &sv, &cv);
        const double v = (double)ir[(size_t)n * n_recv + r];
        re += v * cv;
        im += v * sv;
    }
    out[2 * t] = re;
    out[2 * t + 1] = im;
}

// Plain streaming copy: the calibration kernel for bandwidth and HBM
// traffic counters (reads n4 * 16 bytes, writes the same).
__global__ __launch_bounds__(kBlock) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n4) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n4) dst[i] = src[i];
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, size_t n, T v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Number of tile-kernel variants (rows per tile: see launch_step), the
// default one, and the default z-chunk.
constexpr int kVariants = 4;
constexpr int kDefaultVariant = 0;
constexpr uint32_t kDefaultZChunk = 4;

}  // namespace

// ------------------------------------------------------------------ plan

struct wv_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    int prec = WV_FP32;
    uint32_t nx = 0, ny = 0, nz = 0, pitch = 0, sz = 0;
    uint64_t npitched = 0, cells = 0, air = 0, nb = 0, ns = 0, ns_k = 0;
    double dx = 0, dt = 0;
    uint32_t zchunk = 0;      // requested (0 = default)
    uint32_t run_zchunk = 0;  // the current run's layout
    uint32_t nzb = 0;         // 32-plane blocks of the z-packed bitmaps
    int variant = kDefaultVariant;
    uint32_t max_slots = 0, n_sets = 0;

    DBuf flags, force_base, force, zbits, counts, offs;
    DBuf b_cell_sorted, b_set_sorted;
    DBuf set_n, set_off, set_inv_d, set_cv, set_k, set_w, st_u, st_g;
    DBuf fa, fb;
    DBuf sp_rec, sp_tapw;
    DBuf tap_cell, tap_w, recv, ir, freqs, spec;
    uint32_t n_special = 0, sp_pad = 0, seg_stride = 0;

    // Current run.
    bool active = false;
    size_t steps = 0, done = 0;
    uint32_t n_recv = 0;
    std::vector<double> signal;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    double last_ms = 0.0;
    uint64_t last_steps = 0, last_launches = 0;

    size_t tsize() const { return prec == WV_FP64 ? 8 : 4; }
    size_t ssize() const { return prec == WV_FP32 ? 4 : 8; }
    uint32_t lane_cells() const { return (uint32_t)(16 / tsize()); }
    ~wv_plan() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto e : {ev[0], ev[1], t0, t1})
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Host -> device copy of a vector (asynchronous on the plan's stream; the
// vector must outlive the next synchronize).
template <class V>
void upload(wv_plan* pl, DBuf& d, const V& v) {
    d.alloc(v.size() * sizeof(v[0]));
    if (!v.empty())
        WV_HIP(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice, pl->stream));
}

template <class S>
std::vector<S> cast_vec(const std::vector<double>& v) {
    return std::vector<S>(v.begin(), v.end());
}

Geo geo_of(const wv_plan* pl) { return Geo{pl->npitched, pl->pitch, pl->nx, pl->ny, pl->nz}; }

void plan_setup(wv_plan* pl, const wv_grid* g) {
    WV_HIP(hipGetDevice(&pl->device));
    WV_HIP(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking));
    WV_HIP(hipEventCreate(&pl->ev[0]));
    WV_HIP(hipEventCreate(&pl->ev[1]));
    WV_HIP(hipEventCreate(&pl->t0));
    WV_HIP(hipEventCreate(&pl->t1));
    pl->nx = (uint32_t)g->n[0];
    pl->ny = (uint32_t)g->n[1];
    pl->nz = (uint32_t)g->n[2];
    pl->cells = g->n[0] * g->n[1] * g->n[2];
    pl->pitch = (pl->nx + 63u) & ~63u;
    pl->sz = pl->pitch * pl->ny;
    pl->npitched = (uint64_t)pl->sz * pl->nz;
    pl->dx = g->dx;
    pl->dt = g->dt;
    const uint64_t cells = pl->cells;

    // Flags as given; the device layout is derived from them per run.
    pl->flags.alloc(cells);
    WV_HIP(hipMemcpyAsync(pl->flags.p, g->flags, cells, hipMemcpyHostToDevice, pl->stream));
    uint64_t air = 0;
    for (uint64_t i = 0; i < cells; ++i) air += (g->flags[i] >> 6) & 1u;
    pl->air = air;

    // Boundary nodes sorted by cell (the reference builds them in cell order;
    // any order is accepted), each with its slot set: nodes whose slot lists
    // are equal (same faces and materials) share one set of constants.
    const uint64_t nb = g->n_boundary;
    pl->nb = nb;
    std::vector<uint32_t> order(nb);
    std::iota(order.begin(), order.end(), 0u);
    bool sorted = true;
    for (uint64_t b = 1; b < nb; ++b)
        if (g->b_cell[b] <= g->b_cell[b - 1]) { sorted = false; break; }
    if (!sorted)
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return g->b_cell[a] < g->b_cell[b]; });
    std::vector<uint32_t> bcell(nb), bset(nb);
    std::unordered_map<std::string, uint32_t> sets;
    std::vector<uint32_t> set_n, set_off;
    std::vector<double> inv_d, cv, kk, w;
    uint64_t ns_k = 0;
    // Set 0 is the empty set (a node listed without slots).
    set_n.push_back(0);
    set_off.push_back(0);
    sets.emplace(std::string(), 0u);
    for (uint64_t k = 0; k < nb; ++k) {
        const uint32_t node = order[k];
        bcell[k] = g->b_cell[node];
        const uint32_t j0 = g->b_start[node], j1 = g->b_start[node + 1];
        if (j1 - j0 > kMaxSlots) throw DevErr{"a boundary node has more than 63 slots", WV_ERR_LIMIT};
        std::string key(reinterpret_cast<const char*>(g->slots + j0), (j1 - j0) * sizeof(wv_slot));
        auto it = sets.find(key);
        if (it == sets.end()) {
            const uint32_t id = (uint32_t)set_n.size();
            set_n.push_back(j1 - j0);
            set_off.push_back((uint32_t)inv_d.size());
            for (uint32_t j = j0; j < j1; ++j) {
                inv_d.push_back(g->slots[j].inv_d);
                cv.push_back(g->slots[j].cv);
                kk.push_back(g->slots[j].k);
                w.push_back(g->slots[j].w);
            }
            it = sets.emplace(std::move(key), id).first;
        }
        bset[k] = it->second;
        pl->max_slots = std::max(pl->max_slots, j1 - j0);
        for (uint32_t j = j0; j < j1; ++j) ns_k += g->slots[j].k != 0.0;
    }
    pl->ns = nb ? g->b_start[nb] : 0;
    pl->ns_k = ns_k;
    pl->n_sets = (uint32_t)set_n.size();
    upload(pl, pl->b_cell_sorted, bcell);
    upload(pl, pl->b_set_sorted, bset);
    upload(pl, pl->set_n, set_n);
    upload(pl, pl->set_off, set_off);
    std::vector<float> f_inv, f_cv, f_k, f_w;
    if (pl->prec == WV_FP32) {
        f_inv = cast_vec<float>(inv_d);
        f_cv = cast_vec<float>(cv);
        f_k = cast_vec<float>(kk);
        f_w = cast_vec<float>(w);
        upload(pl, pl->set_inv_d, f_inv);
        upload(pl, pl->set_cv, f_cv);
        upload(pl, pl->set_k, f_k);
        upload(pl, pl->set_w, f_w);
    } else {
        upload(pl, pl->set_inv_d, inv_d);
        upload(pl, pl->set_cv, cv);
        upload(pl, pl->set_k, kk);
        upload(pl, pl->set_w, w);
    }

    // Boundary marks, reused by every run.
    const uint64_t words = pl->npitched / 32;
    pl->force_base.alloc(words * 4);
    pl->force.alloc(words * 4);
    WV_HIP(hipMemsetAsync(pl->force_base.p, 0, words * 4, pl->stream));
    if (nb)
        hipLaunchKernelGGL(k_mark, dim3(div_up(nb, 256)), dim3(256), 0, pl->stream, pl->force_base.as<uint32_t>(),
                           pl->b_cell_sorted.as<uint32_t>(), (uint32_t)nb, pl->nx, pl->pitch);
    WV_HIP(hipGetLastError());
    pl->nzb = (pl->nz - 2 + 31) / 32;
    pl->zbits.alloc((size_t)pl->nzb * pl->ny * pl->pitch * 4);
    pl->fa.alloc(pl->npitched * pl->tsize());
    pl->fb.alloc(pl->npitched * pl->tsize());
    WV_HIP(hipStreamSynchronize(pl->stream));
}

template <class T>
void fill_async(wv_plan* pl, DBuf& d, size_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(div_up(n, 256)), dim3(256), 0, pl->stream, d.as<T>(), n, T(0));
    WV_HIP(hipGetLastError());
}

struct Layout {
    uint32_t zchunk, bx, tiles_x, nzc, seg;
    uint64_t nwaves, positions;
};

Layout layout_of(const wv_plan* pl) {
    Layout L;
    L.zchunk = pl->run_zchunk;
    L.bx = kTileX * pl->lane_cells();
    L.tiles_x = div_up(pl->nx, L.bx);
    L.nzc = div_up(pl->nz - 2, L.zchunk);
    L.seg = (L.zchunk * L.bx + kBlock - 1) / kBlock * kBlock;
    L.nwaves = (uint64_t)L.nzc * L.tiles_x * pl->ny;
    L.positions = L.nwaves * L.seg;
    return L;
}

template <class S>
void run_begin_impl(wv_plan* pl, const wv_tap* taps, size_t n_taps, const double* signal, size_t steps,
                    const uint64_t* receivers, size_t n_recv) {
    // Taps: one per cell, weights of a cell given twice added (gpu_fdtd.rs:67-75), sorted by cell.
    std::vector<std::pair<uint32_t, double>> merged;
    for (size_t t = 0; t < n_taps; ++t) {
        auto it = std::find_if(merged.begin(), merged.end(),
                               [&](const std::pair<uint32_t, double>& m) { return m.first == taps[t].cell; });
        if (it != merged.end()) it->second += taps[t].w;
        else merged.emplace_back((uint32_t)taps[t].cell, taps[t].w);
    }
    std::sort(merged.begin(), merged.end());
    std::vector<uint32_t> tc(merged.size());
    std::vector<S> tw(merged.size());
    for (size_t t = 0; t < merged.size(); ++t) {
        tc[t] = merged[t].first;
        tw[t] = (S)merged[t].second;
    }
    upload(pl, pl->tap_cell, tc);
    upload(pl, pl->tap_w, tw);
    const uint64_t words = pl->npitched / 32;
    WV_HIP(hipMemcpyAsync(pl->force.p, pl->force_base.p, words * 4, hipMemcpyDeviceToDevice, pl->stream));
    if (!tc.empty())
        hipLaunchKernelGGL(k_mark, dim3(div_up(tc.size(), 256)), dim3(256), 0, pl->stream, pl->force.as<uint32_t>(),
                           pl->tap_cell.as<uint32_t>(), (uint32_t)tc.size(), pl->nx, pl->pitch);
    const Geo geo = geo_of(pl);
    hipLaunchKernelGGL(k_zbits, dim3(div_up((uint64_t)pl->nzb * pl->ny * pl->pitch, kBlock)), dim3(kBlock), 0,
                       pl->stream, pl->flags.as<uint8_t>(), pl->force.as<uint32_t>(), pl->zbits.as<uint32_t>(),
                       geo, pl->nzb);
    // Special cells in wave order.
    pl->run_zchunk = pl->zchunk ? pl->zchunk : kDefaultZChunk;
    const Layout L = layout_of(pl);
    const WaveOrder wo{L.bx, L.zchunk, L.seg, L.tiles_x};
    const uint64_t nblocks = L.positions / kBlock;
    if (nblocks >= (1ull << 31)) throw DevErr{"grid too large for the special-cell layout", WV_ERR_LIMIT};
    // offs[b] = special cells before compaction block b; offs[nblocks] = total.
    pl->counts.alloc(nblocks * 4);
    pl->offs.alloc((nblocks + 1) * 4);
    pl->seg_stride = L.seg / kBlock;
    hipLaunchKernelGGL(k_sp_count, dim3((uint32_t)nblocks), dim3(kBlock), 0, pl->stream, pl->flags.as<uint8_t>(),
                       pl->force.as<uint32_t>(), pl->counts.as<uint32_t>(), geo, wo);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, pl->stream, pl->counts.as<uint32_t>(), pl->offs.as<uint32_t>(),
                       (uint32_t)nblocks, pl->offs.as<uint32_t>() + nblocks);
    WV_HIP(hipGetLastError());
    uint32_t n_special = 0;
    WV_HIP(hipMemcpyAsync(&n_special, pl->offs.as<uint32_t>() + nblocks, 4, hipMemcpyDeviceToHost, pl->stream));
    WV_HIP(hipStreamSynchronize(pl->stream));
    pl->n_special = n_special;
    pl->sp_pad = (n_special + 63u) & ~63u;
    pl->sp_rec.alloc((size_t)n_special * 8);
    pl->sp_tapw.alloc((size_t)n_special * sizeof(S));
    hipLaunchKernelGGL(k_sp_write<S>, dim3((uint32_t)nblocks), dim3(kBlock), 0, pl->stream, pl->flags.as<uint8_t>(),
                       pl->force.as<uint32_t>(), pl->offs.as<uint32_t>(), geo, wo, pl->b_cell_sorted.as<uint32_t>(),
                       pl->b_set_sorted.as<uint32_t>(), (uint32_t)pl->nb, pl->set_n.as<uint32_t>(),
                       pl->set_off.as<uint32_t>(), pl->tap_cell.as<uint32_t>(), pl->tap_w.as<S>(), (uint32_t)tc.size(),
                       pl->sp_rec.as<uint2>(), pl->sp_tapw.as<S>());
    WV_HIP(hipGetLastError());
    // Slot state, slot-major: state[j][rank].
    const size_t nstate = (size_t)pl->max_slots * pl->sp_pad;
    pl->st_u.alloc(nstate * pl->ssize());
    pl->st_g.alloc(nstate * pl->ssize());
    // Receivers: pitched cell indices.
    std::vector<uint32_t> rc(n_recv);
    for (size_t r = 0; r < n_recv; ++r) {
        const uint64_t c = receivers[r];
        rc[r] = (uint32_t)((c / pl->nx) * pl->pitch + c % pl->nx);
    }
    upload(pl, pl->recv, rc);
    pl->ir.alloc(std::max<size_t>(1, steps * n_recv) * pl->tsize());
    // Zeroed fields, state and records.
    if (pl->prec == WV_FP64) {
        fill_async<double>(pl, pl->fa, pl->npitched);
        fill_async<double>(pl, pl->fb, pl->npitched);
        fill_async<double>(pl, pl->ir, steps * n_recv);
    } else {
        fill_async<float>(pl, pl->fa, pl->npitched);
        fill_async<float>(pl, pl->fb, pl->npitched);
        fill_async<float>(pl, pl->ir, steps * n_recv);
    }
    fill_async<S>(pl, pl->st_u, nstate);
    fill_async<S>(pl, pl->st_g, nstate);
    WV_HIP(hipStreamSynchronize(pl->stream));
    pl->signal.assign(signal, signal + steps);
    pl->steps = steps;
    pl->done = 0;
    pl->n_recv = (uint32_t)n_recv;
    pl->active = true;
}

template <class T, class S>
StepParams<T, S> make_params(wv_plan* pl, uint32_t ty_rows) {
    StepParams<T, S> a{};
    const Layout L = layout_of(pl);
    a.zbits = pl->zbits.as<uint32_t>();
    a.nx = pl->nx;
    a.ny = pl->ny;
    a.nz = pl->nz;
    a.pitch = pl->pitch;
    a.sz = pl->sz;
    a.tiles_x = L.tiles_x;
    a.tiles_y = div_up(pl->ny, ty_rows);
    a.zchunk = L.zchunk;
    a.main_blocks = a.tiles_x * a.tiles_y * L.nzc;
    a.seg_base = pl->offs.as<uint32_t>();
    a.seg_stride = pl->seg_stride;
    a.sp_rec = pl->sp_rec.as<uint2>();
    a.sp_tapw = pl->sp_tapw.as<S>();
    a.sp_pad = pl->sp_pad;
    a.set_inv_d = pl->set_inv_d.as<S>();
    a.set_cv = pl->set_cv.as<S>();
    a.set_k = pl->set_k.as<S>();
    a.set_w = pl->set_w.as<S>();
    a.st_u = pl->st_u.as<S>();
    a.st_g = pl->st_g.as<S>();
    a.has_k = pl->ns_k > 0;
    a.n_recv = pl->n_recv;
    a.recv = pl->recv.as<uint32_t>();
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

// Tile-kernel variants: TY rows per tile (block = 64 x TY threads).
constexpr uint32_t kVariantRows[] = {4, 8, 16, 2};

template <class T, class S>
void launch_step(int variant, uint32_t grid, hipStream_t st, const StepParams<T, S>& a) {
    switch (variant) {
        case 0: hipLaunchKernelGGL((k_step<T, S, 4>), dim3(grid), dim3(64 * 4), 0, st, a); break;
        case 1: hipLaunchKernelGGL((k_step<T, S, 8>), dim3(grid), dim3(64 * 8), 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_step<T, S, 16>), dim3(grid), dim3(64 * 16), 0, st, a); break;
        default: hipLaunchKernelGGL((k_step<T, S, 2>), dim3(grid), dim3(64 * 2), 0, st, a); break;
    }
}

template <class T, class S>
wv_status advance_impl(wv_plan* pl, size_t k, const volatile int32_t* cancel, wv_progress_fn progress, void* ctx) {
    const uint32_t ty_rows = kVariantRows[pl->variant];
    StepParams<T, S> a = make_params<T, S>(pl, ty_rows);
    const uint32_t grid = a.main_blocks + div_up(pl->n_recv, kTileX * ty_rows);
    const size_t end = pl->done + k;
    T* A = pl->fa.as<T>();
    T* B = pl->fb.as<T>();
    T* ir = pl->ir.as<T>();
    uint64_t launches = 0;
    int blk = 0;
    bool pending = false;
    wv_status st = WV_OK;
    WV_HIP(hipEventRecord(pl->t0, pl->stream));
    while (pl->done < end) {
        if (cancel && *cancel) { st = WV_CANCELLED; break; }
        const size_t stop = std::min(end, pl->done + kStepsPerBlock);
        for (size_t n = pl->done; n < stop; ++n) {
            // Even steps: p = A, prev = B; odd steps swap (gpu_fdtd.rs:242-243).
            a.p = (n & 1) ? B : A;
            a.prev = (n & 1) ? A : B;
            a.s_n = (S)(sizeof(S) == 4 ? (double)(float)pl->signal[n] : pl->signal[n]);
            a.ir = ir + n * pl->n_recv;
            if (grid) launch_step<T, S>(pl->variant, grid, pl->stream, a);
            ++launches;
        }
        WV_HIP(hipGetLastError());
        WV_HIP(hipEventRecord(pl->ev[blk], pl->stream));
        // Wait for the previous block while this one runs, so the device
        // always has a block queued.
        if (pending) {
            WV_HIP(hipEventSynchronize(pl->ev[blk ^ 1]));
            if (progress) progress((double)(pl->done) / (double)pl->steps, ctx);
        }
        pending = true;
        pl->done = stop;
        blk ^= 1;
    }
    WV_HIP(hipEventRecord(pl->t1, pl->stream));
    WV_HIP(hipEventSynchronize(pl->t1));
    if (st == WV_OK && progress) progress((double)pl->done / (double)pl->steps, ctx);
    float ms = 0.f;
    WV_HIP(hipEventElapsedTime(&ms, pl->t0, pl->t1));
    pl->last_ms = ms;
    pl->last_steps = launches;
    pl->last_launches = launches;
    return st;
}

template <class T>
void finish_impl(wv_plan* pl, const double* freqs, size_t nf, int mode, double* spectra_out, double* ir_out) {
    const size_t R = pl->n_recv, Tn = pl->done;
    if ((mode & WV_MODE_SPECTRA) && spectra_out && R && nf) {
        pl->freqs.alloc(nf * 8);
        WV_HIP(hipMemcpyAsync(pl->freqs.p, freqs, nf * 8, hipMemcpyHostToDevice, pl->stream));
        pl->spec.alloc(R * nf * 16);
        hipLaunchKernelGGL(k_dft<T>, dim3(div_up(R * nf, 64)), dim3(64), 0, pl->stream, pl->ir.as<T>(),
                           (uint32_t)Tn, (uint32_t)R, pl->freqs.as<double>(), (uint32_t)nf, pl->dt,
                           pl->spec.as<double>());
        WV_HIP(hipGetLastError());
        WV_HIP(hipMemcpyAsync(spectra_out, pl->spec.p, R * nf * 16, hipMemcpyDeviceToHost, pl->stream));
    }
    if ((mode & WV_MODE_IR) && ir_out && R) {
        std::vector<T> h(Tn * R);
        if (Tn) WV_HIP(hipMemcpyAsync(h.data(), pl->ir.p, Tn * R * sizeof(T), hipMemcpyDeviceToHost, pl->stream));
        WV_HIP(hipStreamSynchronize(pl->stream));
        for (size_t r = 0; r < R; ++r)
            for (size_t n = 0; n < Tn; ++n) ir_out[r * pl->steps + n] = (double)h[n * R + r];
    }
    WV_HIP(hipStreamSynchronize(pl->stream));
}

wv_status check_grid(const wv_grid* g, char* err, size_t err_len) {
    if (!g || !g->flags || (g->n_boundary && (!g->b_cell || !g->b_start || !g->slots)) || !g->b_start) {
        set_err(err, err_len, "invalid grid: null arrays");
        return WV_ERR_INVALID;
    }
    for (int k = 0; k < 3; ++k)
        if (g->n[k] < 3) {
            set_err(err, err_len, "invalid grid: every axis needs at least 3 cells");
            return WV_ERR_INVALID;
        }
    const uint64_t cells = g->n[0] * g->n[1] * g->n[2];
    const uint64_t pitched = ((g->n[0] + 63) & ~63ull) * g->n[1] * g->n[2];
    if (cells > kMaxCells || pitched >= (1ull << 32)) {
        set_err(err, err_len, "grid of " + std::to_string(cells / 1000000) + " M cells exceeds the GPU layout limit");
        return WV_ERR_LIMIT;
    }
    if (g->n_boundary >= (1ull << 31)) {
        set_err(err, err_len, "too many boundary nodes for the GPU layout");
        return WV_ERR_LIMIT;
    }
    for (uint64_t b = 0; b < g->n_boundary; ++b)
        if (g->b_cell[b] >= cells || g->b_start[b + 1] < g->b_start[b]) {
            set_err(err, err_len, "invalid grid: boundary list out of range");
            return WV_ERR_INVALID;
        }
    return WV_OK;
}

template <class F>
wv_status guarded(char* err, size_t err_len, F f) {
    try {
        return f();
    } catch (const DevErr& e) {
        set_err(err, err_len, e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_err(err, err_len, "host allocation failed");
        return WV_ERR_NOMEM;
    }
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" wv_status wv_plan_create(const wv_grid* grid, int precision, wv_plan** out, char* err, size_t err_len) {
    if (!out || (precision != WV_FP32 && precision != WV_FP64 && precision != WV_FP32_STATE64)) {
        set_err(err, err_len, "wv_plan_create: invalid argument");
        return WV_ERR_INVALID;
    }
    wv_status s = check_grid(grid, err, err_len);
    if (s != WV_OK) return s;
    return guarded(err, err_len, [&]() -> wv_status {
        auto pl = std::make_unique<wv_plan>();
        pl->prec = precision;
        plan_setup(pl.get(), grid);
        *out = pl.release();
        return WV_OK;
    });
}

extern "C" void wv_plan_free(wv_plan* plan) { delete plan; }

extern "C" wv_status wv_plan_set_tile(wv_plan* plan, uint32_t zchunk, int32_t variant) {
    // z-chunks must divide the 32-plane blocks of the interior bitmap.
    if (!plan || zchunk > 32 || (zchunk & (zchunk - 1)) || variant >= kVariants) return WV_ERR_INVALID;
    plan->zchunk = zchunk;
    plan->variant = variant < 0 ? kDefaultVariant : variant;
    return WV_OK;
}

extern "C" wv_status wv_run_begin(wv_plan* pl, const wv_tap* taps, size_t n_taps, const double* signal, size_t steps,
                                  const uint64_t* receivers, size_t n_recv, char* err, size_t err_len) {
    if (!pl || (n_taps && !taps) || (steps && !signal) || (n_recv && !receivers)) {
        set_err(err, err_len, "wv_run_begin: invalid argument");
        return WV_ERR_INVALID;
    }
    // Taps must lie in the planes the step kernel updates (1 .. nz-2); every
    // air cell does. The reference injects anywhere (fdtd.rs:718-720), but a
    // tap on the outer padding planes could only reach a receiver there.
    const uint64_t plane = (uint64_t)pl->nx * pl->ny;
    for (size_t t = 0; t < n_taps; ++t)
        if (taps[t].cell >= pl->cells || taps[t].cell / plane == 0 || taps[t].cell / plane + 1 >= pl->nz) {
            set_err(err, err_len, "source tap outside the updated planes of the grid");
            return WV_ERR_INVALID;
        }
    for (size_t r = 0; r < n_recv; ++r)
        if (receivers[r] >= pl->cells) {
            set_err(err, err_len, "receiver outside the grid");
            return WV_ERR_INVALID;
        }
    if ((uint64_t)steps * n_recv >= (1ull << 34)) {
        set_err(err, err_len, "receiver record too large");
        return WV_ERR_LIMIT;
    }
    return guarded(err, err_len, [&]() -> wv_status {
        WV_HIP(hipSetDevice(pl->device));
        if (pl->prec == WV_FP32) run_begin_impl<float>(pl, taps, n_taps, signal, steps, receivers, n_recv);
        else run_begin_impl<double>(pl, taps, n_taps, signal, steps, receivers, n_recv);
        return WV_OK;
    });
}

extern "C" wv_status wv_run_advance(wv_plan* pl, size_t k, const volatile int32_t* cancel, wv_progress_fn progress,
                                    void* ctx, char* err, size_t err_len) {
    if (!pl || !pl->active || pl->done + k > pl->steps) {
        set_err(err, err_len, "wv_run_advance: no run, or more steps than the signal has");
        return WV_ERR_INVALID;
    }
    return guarded(err, err_len, [&]() -> wv_status {
        WV_HIP(hipSetDevice(pl->device));
        switch (pl->prec) {
            case WV_FP32: return advance_impl<float, float>(pl, k, cancel, progress, ctx);
            case WV_FP64: return advance_impl<double, double>(pl, k, cancel, progress, ctx);
            default: return advance_impl<float, double>(pl, k, cancel, progress, ctx);
        }
    });
}

extern "C" wv_status wv_run_finish(wv_plan* pl, const double* freqs, size_t nf, int mode, double* spectra_out,
                                   double* ir_out, char* err, size_t err_len) {
    if (!pl || !pl->active || (nf && !freqs)) {
        set_err(err, err_len, "wv_run_finish: invalid argument");
        return WV_ERR_INVALID;
    }
    return guarded(err, err_len, [&]() -> wv_status {
        WV_HIP(hipSetDevice(pl->device));
        if (pl->prec == WV_FP64) finish_impl<double>(pl, freqs, nf, mode, spectra_out, ir_out);
        else finish_impl<float>(pl, freqs, nf, mode, spectra_out, ir_out);
        pl->active = false;
        return WV_OK;
    });
}

extern "C" wv_status wv_plan_run(wv_plan* plan, const wv_tap* taps, size_t n_taps, const double* signal, size_t steps,
                                 const uint64_t* receivers, size_t n_recv, const double* freqs, size_t nf, int mode,
                                 const volatile int32_t* cancel, wv_progress_fn progress, void* ctx,
                                 double* spectra_out, double* ir_out, char* err, size_t err_len) {
    wv_status s = wv_run_begin(plan, taps, n_taps, signal, steps, receivers, n_recv, err, err_len);
    if (s != WV_OK) return s;
    s = wv_run_advance(plan, steps, cancel, progress, ctx, err, err_len);
    if (s != WV_OK) {
        plan->active = false;
        return s;
    }
    return wv_run_finish(plan, freqs, nf, mode, spectra_out, ir_out, err, err_len);
}

extern "C" wv_status wv_fdtd_run(const wv_grid* grid, const wv_tap* taps, size_t n_taps, const double* signal,
                                 size_t steps, const uint64_t* receivers, size_t n_recv, const double* freqs,
                                 size_t nf, int mode, int precision, const volatile int32_t* cancel,
                                 wv_progress_fn progress, void* ctx, double* spectra_out, double* ir_out, char* err,
                                 size_t err_len) {
    wv_plan* plan = nullptr;
    wv_status s = wv_plan_create(grid, precision, &plan, err, err_len);
    if (s != WV_OK) return s;
    s = wv_plan_run(plan, taps, n_taps, signal, steps, receivers, n_recv, freqs, nf, mode, cancel, progress, ctx,
                    spectra_out, ir_out, err, err_len);
    wv_plan_free(plan);
    return s;
}

extern "C" wv_status wv_plan_last_timing(const wv_plan* plan, double* ms, uint64_t* steps, uint64_t* launches) {
    if (!plan) return WV_ERR_INVALID;
    if (ms) *ms = plan->last_ms;
    if (steps) *steps = plan->last_steps;
    if (launches) *launches = plan->last_launches;
    return WV_OK;
}

extern "C" wv_status wv_plan_copy_probe(wv_plan* plan, uint64_t bytes, uint32_t reps, double* ms) {
    if (!plan || !ms || reps == 0) return WV_ERR_INVALID;
    char err[8];
    return guarded(err, sizeof err, [&]() -> wv_status {
        WV_HIP(hipSetDevice(plan->device));
        const size_t n4 = std::min<uint64_t>(bytes, plan->fa.bytes) / 16;
        if (n4 == 0) return WV_ERR_INVALID;
        WV_HIP(hipEventRecord(plan->t0, plan->stream));
        for (uint32_t r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_copy, dim3(div_up(n4, kBlock)), dim3(kBlock), 0, plan->stream,
                               plan->fa.as<float4>(), plan->fb.as<float4>(), n4);
        WV_HIP(hipGetLastError());
        WV_HIP(hipEventRecord(plan->t1, plan->stream));
        WV_HIP(hipEventSynchronize(plan->t1));
        float t = 0.f;
        WV_HIP(hipEventElapsedTime(&t, plan->t0, plan->t1));
        *ms = t / reps;
        return WV_OK;
    });
}

extern "C" wv_status wv_plan_stats(const wv_plan* plan, uint64_t out[7]) {
    if (!plan || !out) return WV_ERR_INVALID;
    out[0] = plan->cells;
    out[1] = plan->air;
    out[2] = plan->nb;
    out[3] = plan->ns;
    out[4] = plan->n_special;
    out[5] = plan->npitched;
    out[6] = plan->ns_k;
    return WV_OK;
}
