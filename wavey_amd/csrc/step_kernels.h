// Step kernel of the FDTD scheme on MI355X (gfx950): the per-step parameter
// block, the cell updates and the tile z-march, shared by the plain run and
// the batched launch.
//
// The scheme is Wavey's (core/wavey-solver/src/fdtd.rs:1-30): the 3D
// leapfrog p+ = (2 - K l2) p + l2 sum_air p_nb - p-, with locally reacting
// frequency-dependent boundaries solved per node in closed form
// (fdtd.rs:581-598).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace wv {

constexpr uint32_t kBlock = 256;          // setup kernels and compaction blocks
constexpr uint32_t kTileX = 64;           // one 64-lane wave per tile row
constexpr uint32_t kStepsPerBlock = 128;  // cancel/progress granularity, as gpu_fdtd.rs:35

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
    // Tiles are spans of 64 W consecutive cells of the pitched plane (pitch *
    // ny cells); a block holds TY of them and marches one z-chunk.
    uint32_t n_tiles, tile_blocks, zchunk, main_blocks;
    // Special cells, in wave order: wave w owns [seg_base[w * seg_stride],
    // seg_base[(w + 1) * seg_stride]).
    const uint32_t* __restrict__ seg_base;
    uint32_t seg_stride;
    const uint2* __restrict__ sp_rec;
    const S* __restrict__ sp_tapw;
    uint32_t sp_pad;  // row length of the slot-major state arrays
    const S* __restrict__ set_c;  // per slot {inv_d, cv, k, w}
    S* __restrict__ st_u;  // [slot][rank]
    S* __restrict__ st_g;
    uint32_t has_k;
    uint32_t max_slots;  // rows of the state arrays
    uint32_t n_recv;
    const uint32_t* __restrict__ recv;
    T* __restrict__ ir;  // this step's receiver row
    // FP32 fields hold the pressure less the run's offset *d0, which
    // re-centring moves out of them: the boundary slots add it back, and the
    // record (the stored value) gets it in FP64 when it is read out. FP64
    // fields hold the pressure.
    const double* __restrict__ d0;
    T l2, c2;            // interior: p+ = c2 p + l2 sum - p-
    S l2s, lam, half_dt, s_n;
};

// Index of the offset in force at step n: the number of re-centrings at
// steps 128, 256, 512, ... up to n.
__host__ __device__ __forceinline__ uint32_t offset_index(uint32_t n) {
    return n < 128 ? 0u : 31u - (uint32_t)__builtin_clz(n) - 6u;
}

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

// One row's update of the W cells of a lane: interior_update of each, with
// the FP32 neighbour differences shared. xm - c of a cell is the negation of
// xp - c of the cell before it, and zm - c at a plane the negation of zp - c
// at the plane below (dzm, kept by the caller; this plane's zp - c goes to
// dzp). Round to nearest is symmetric, so each sum of two differences is the
// one interior_update forms, bit for bit, with 5 subtractions per lane for x
// instead of 8 and 4 for z instead of 8. FP64 keeps interior_update's form.
template <typename T, int W>
__device__ __forceinline__ void row_update(const T (&c)[W], const T (&q)[W], T lo, T hi, const T (&ym)[W],
                                           const T (&yp)[W], const T (&zm)[W], const T (&zp)[W],
                                           const T (&dzm)[W], T (&dzp)[W], T l2, T c2, T (&out)[W]) {
    if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int e = 0; e < W; ++e)
            out[e] = interior_update(c[e], q[e], e == 0 ? lo : c[e - 1], e == W - 1 ? hi : c[e + 1], ym[e], yp[e],
                                     zm[e], zp[e], l2, c2);
    } else {
        T dx[W + 1];  // dx[e] = c[e] - (the cell before it)
        dx[0] = c[0] - lo;
#pragma unroll
        for (int e = 1; e < W; ++e) dx[e] = c[e] - c[e - 1];
        dx[W] = hi - c[W - 1];
#pragma unroll
        for (int e = 0; e < W; ++e) {
            const T zl = dzm[e];  // read before dzp[e] is written: the caller may pass one array as both
            const T zu = zp[e] - c[e];
            const T d = (dx[e + 1] - dx[e]) + ((ym[e] - c[e]) + (yp[e] - c[e])) + (zu - zl);
            out[e] = c[e] + ((c[e] - q[e]) + l2 * d);
            dzp[e] = zu;
        }
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

// W consecutive cells of a row per lane, moved as one access of W * sizeof(T)
// bytes (16 by default: 4 FP32 or 2 FP64 cells).
template <typename T, int Wd = 16 / (int)sizeof(T)>
struct Cells {
    static constexpr int W = Wd;
    T v[W];
    __device__ __forceinline__ void load(const T* __restrict__ p) {
        if constexpr (W * sizeof(T) == 16) {
            if constexpr (sizeof(T) == 4) {
                const float4 q = *reinterpret_cast<const float4*>(p);
                v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
            } else {
                const double2 q = *reinterpret_cast<const double2*>(p);
                v[0] = q.x; v[1] = q.y;
            }
        } else if constexpr (W * sizeof(T) == 8 && sizeof(T) == 4) {
            const float2 q = *reinterpret_cast<const float2*>(p);
            v[0] = q.x; v[1] = q.y;
        } else {
#pragma unroll
            for (int e = 0; e < W; ++e) v[e] = p[e];
        }
    }
    __device__ __forceinline__ void store(T* __restrict__ p) const {
        if constexpr (W * sizeof(T) == 16) {
            if constexpr (sizeof(T) == 4) *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
            else *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
        } else if constexpr (W * sizeof(T) == 8 && sizeof(T) == 4) {
            *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
        } else {
#pragma unroll
            for (int e = 0; e < W; ++e) p[e] = v[e];
        }
    }
};

template <int W>
__device__ __forceinline__ void load_words(const uint32_t* __restrict__ p, uint32_t (&w)[W]) {
    if constexpr (W == 4) {
        const uint4 q = *reinterpret_cast<const uint4*>(p);
        w[0] = q.x; w[1] = q.y; w[2] = q.z; w[3] = q.w;
    } else if constexpr (W == 2) {
        const uint2 q = *reinterpret_cast<const uint2*>(p);
        w[0] = q.x; w[1] = q.y;
    } else {
#pragma unroll
        for (int e = 0; e < W; ++e) w[e] = p[e];
    }
}

template <int W>
__device__ __forceinline__ uint32_t plane_bits(const uint32_t (&w)[W], uint32_t sh) {
    uint32_t b = 0;
#pragma unroll
    for (int e = 0; e < W; ++e) b |= ((w[e] >> sh) & 1u) << e;
    return b;
}

// Slots a special cell keeps in registers: the first group of its boundary
// solve. Three in FP32 (a wall face with a three-branch admittance) keeps the
// step kernel within 80 VGPRs; two with FP64 state.
template <typename S>
constexpr uint32_t kSlotGroup = sizeof(S) == 4 ? 3 : 2;

// Constants and state of slots jb .. jb+G-1 of a node (indices clamped to
// the last slot; the caller masks the extras). `upre`, when given, holds the
// states of rows jb .. jb+G-1 read earlier.
template <typename S, uint32_t G>
struct SlotGroup {
    S inv_d[G], w[G], cv[G], u[G], kk[G], g[G];
    template <typename T>
    __device__ __forceinline__ void load(const StepParams<T, S>& a, uint32_t c0, uint32_t rank, uint32_t jb,
                                         uint32_t ns, const S* upre = nullptr) {
#pragma unroll
        for (uint32_t k = 0; k < G; ++k) {
            const uint32_t j = min(jb + k, ns - 1);
            const S* c = a.set_c + 4 * (c0 + j);
            if constexpr (sizeof(S) == 4) {
                const float4 q = *reinterpret_cast<const float4*>(c);
                inv_d[k] = q.x;
                cv[k] = q.y;
                w[k] = q.w;
            } else {
                const double2 q0 = *reinterpret_cast<const double2*>(c);
                const double2 q1 = *reinterpret_cast<const double2*>(c + 2);
                inv_d[k] = q0.x;
                cv[k] = q0.y;
                w[k] = q1.y;
            }
            u[k] = upre ? upre[k] : a.st_u[j * a.sp_pad + rank];
            kk[k] = 0;
            g[k] = 0;
            if (a.has_k) {
                kk[k] = c[2];
                g[k] = a.st_g[j * a.sp_pad + rank];
            }
        }
    }
    // The part of slot k's u+ that does not depend on p+ (fdtd.rs:589-592).
    __device__ __forceinline__ S rest(uint32_t k, S pc, S pp) const {
#pragma clang fp contract(off)
        return (cv[k] * u[k] + S(0.5) * pc + w[k] * pp - kk[k] * g[k]) * inv_d[k];
    }
    __device__ __forceinline__ void sums(uint32_t jb, uint32_t ns, S pc, S pp, S& s1, S& s2) const {
#pragma clang fp contract(off)
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
#pragma clang fp contract(off)
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
// constant of the node's slot set, rank its index in the state arrays and u0
// the states of its first kSlotGroup rows, read ahead by the caller.
template <typename T, typename S>
__device__ __forceinline__ S special_update(const StepParams<T, S>& a, uint32_t meta, uint32_t c0, uint32_t rank,
                                            S pc, S pp, const S (&nb)[6], const S (&u0)[kSlotGroup<S>]) {
    // Every product and sum of the boundary solve is rounded on its own (no
    // fused multiply-adds): the compiler would otherwise fuse them
    // differently in the kernels that inline this (measured: FP32 series
    // with spring branches differed in the last bits between the single-step
    // kernel and a multi-step kernel that inlined the solve in its z-march),
    // and the pass modes, batch and server promise the same bits.
#pragma clang fp contract(off)
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
            // further groups are read again for the second pass.
            constexpr uint32_t G = kSlotGroup<S>;
            // The slots see the pressure itself: FP32 fields add the run's
            // offset (the Laplacian and the update's own terms do not
            // depend on it).
            S off = S(0), pa = pc, ppa = pp;
            if constexpr (sizeof(T) == 4) {
                off = (S)*a.d0;
                pa = pc + off;
                ppa = pp + off;
            }
            SlotGroup<S, G> g0;
            g0.load(a, c0, rank, 0, ns, u0);
            S s1 = 0, s2 = 0;
            g0.sums(0, ns, pa, ppa, s1, s2);
            for (uint32_t jb = G; jb < ns; jb += G) {
                SlotGroup<S, G> gj;
                gj.load(a, c0, rank, jb, ns);
                gj.sums(jb, ns, pa, ppa, s1, s2);
            }
            S nabs;
            if constexpr (sizeof(S) == 8) {
                nabs = (S(2) * pa - ppa + lap - a.lam * s2) / (S(1) + a.lam * s1);
                next = nabs;
                if constexpr (sizeof(T) == 4) next = nabs - off;
            } else {
                // The same solve written as pc + (small terms).
                next = pc + ((pc - pp) + lap - a.lam * (s2 + s1 * pa)) / (S(1) + a.lam * s1);
                nabs = next + off;
            }
            g0.advance(a, rank, 0, ns, pa, ppa, nabs);
            for (uint32_t jb = G; jb < ns; jb += G) {
                SlotGroup<S, G> gj;
                gj.load(a, c0, rank, jb, ns);
                gj.advance(a, rank, jb, ns, pa, ppa, nabs);
            }
        }
    }
    if (meta & kMetaTap) next += a.sp_tapw[rank] * a.s_n;
    return next;
}

// A tile is a span of 64*W consecutive cells of the pitched xy plane, one
// 64-lane wave, each lane W consecutive cells (16 bytes: 4 FP32 or 2 FP64),
// marching z over [z0, z1). Tiles run across row ends, so narrow grids keep
// their lanes busy: a cell next to a row end (x = 0, x >= nx - 1, the pitch
// padding) is never interior, so the x-neighbour values it would take from
// the adjacent row are never stored; interior cells (x in 1..nx-2, y in
// 1..ny-2) have both x-neighbours in their row and both y-neighbours at
// +-pitch. Per plane a lane loads p at z+1, prev, the two y-neighbour rows
// and, on lanes 0 and 63, the cells next to the tile; x neighbours inside a
// lane are in its registers and move across lanes by wave shifts. All of a
// plane's loads are issued before any is used and the loop keeps few
// registers live, so many waves per SIMD hide the memory latency. Only
// interior cells are written here.
template <typename T, typename S>
__device__ __forceinline__ void tile_march(const StepParams<T, S>& a, uint32_t tp, uint32_t z0) {
    constexpr int W = Cells<T>::W;
    constexpr uint32_t BX = kTileX * W;
    const uint32_t lane = threadIdx.x & (kTileX - 1);
    // Lanes past the plane (the last tile) re-read its last 16 bytes and
    // update nothing. Every neighbour offset below stays inside the field:
    // planes z0 - 1 .. z1 are within 0 .. nz - 1 and a plane has sz >= pitch
    // cells.
    const uint32_t fl = tp * BX + lane * W;  // first cell of this lane in the plane
    const bool live = fl < a.sz;
    const uint32_t col = live ? fl : a.sz - W;
    const uint32_t z1 = min(z0 + a.zchunk, a.nz - 1);
    const uint32_t full = (1u << W) - 1u;
    const uint32_t lane_mask = live ? full : 0u;
    // The chunk lies inside one 32-plane block of the z-packed bitmap
    // (z-chunks divide 32 and start at plane 1 + a multiple of the chunk).
    uint32_t ib[W];
    load_words<W>(a.zbits + ((z0 - 1) >> 5) * a.sz + col, ib);

    Cells<T> zm, zcur;
    zm.load(a.p + (z0 - 1) * a.sz + col);
    zcur.load(a.p + z0 * a.sz + col);
    T dz[W];  // zp - c at the plane below (FP32), carried from plane to plane
#pragma unroll
    for (int e = 0; e < W; ++e) dz[e] = zcur.v[e] - zm.v[e];
    for (uint32_t z = z0; z < z1; ++z) {
        const uint32_t i = z * a.sz + col;
        T hl = 0, hr = 0;
        if (lane == 0) hl = a.p[i - 1];                // the cell before the tile
        if (lane == kTileX - 1) hr = a.p[i + W];       // the cell after it
        Cells<T> zp, q, ym, yp;
        zp.load(a.p + i + a.sz);
        q.load(a.prev + i);
        ym.load(a.p + i - a.pitch);
        yp.load(a.p + i + a.pitch);
        const T lo = from_lower_lane(zcur.v[W - 1], hl);
        const T hi = from_upper_lane(zcur.v[0], hr);
        Cells<T> out;
        row_update<T, W>(zcur.v, q.v, lo, hi, ym.v, yp.v, zm.v, zp.v, dz, dz, a.l2, a.c2, out.v);
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
}

// Special cell s of a block (tiles tp0 .. tp0 + TY - 1, z-chunk from z0),
// given its record and the states of its first G slot rows: its tile is
// found from the block's per-tile record boundaries bnd[k] (first record of
// tile tp0 + k), its cell from the position the record holds.
// Special cell of record s at pitched index i: p from a.p, the cell's old
// value from a.prev, the result written to out[i]. The neighbour loads are
// unconditional: a masked-out neighbour reads the cell itself, and the mask
// never points outside the grid.
template <typename T, typename S>
__device__ __forceinline__ void special_at(const StepParams<T, S>& a, uint32_t s, uint32_t i, uint2 rec,
                                           const S (&u0)[kSlotGroup<S>], T* __restrict__ out) {
    const uint32_t meta = rec.x;
    const T c = a.p[i];
    const T q = a.prev[i];
    const S nb[6] = {(S)a.p[i - ((meta & 1u) ? 1u : 0u)], (S)a.p[i + ((meta & 2u) ? 1u : 0u)],
                     (S)a.p[i - ((meta & 4u) ? a.pitch : 0u)], (S)a.p[i + ((meta & 8u) ? a.pitch : 0u)],
                     (S)a.p[i - ((meta & 16u) ? a.sz : 0u)], (S)a.p[i + ((meta & 32u) ? a.sz : 0u)]};
    out[i] = (T)special_update(a, meta, rec.y, s, (S)c, (S)q, nb, u0);
}

template <typename T, typename S, int TY>
__device__ __forceinline__ void special_one(const StepParams<T, S>& a, const uint32_t (&bnd)[TY + 1], uint32_t tp0,
                                            uint32_t z0, uint32_t s, uint2 rec, const S (&u0)[kSlotGroup<S>]) {
    constexpr int W = Cells<T>::W;
    constexpr uint32_t BX = kTileX * W;
    uint32_t k = 0;
#pragma unroll
    for (int m = 1; m < TY; ++m) k += s >= bnd[m] ? 1u : 0u;
    const uint32_t pos = rec.x >> kMetaPosShift;
    const uint32_t i = z0 * a.sz + (tp0 + k) * BX + (pos / BX) * a.sz + pos % BX;
    special_at(a, s, i, rec, u0, a.prev);
}

// A batch of 64 special records: the record and the states of its first G
// slot rows (rows the state arrays have, whatever the cell's slot count).
template <typename S>
struct SpecialBatch {
    uint32_t base;
    uint2 rec;
    S u0[kSlotGroup<S>];
    template <typename T>
    __device__ __forceinline__ void load(const StepParams<T, S>& a, uint32_t first, uint32_t n, uint32_t lane) {
        rec = make_uint2(0u, 0u);
#pragma unroll
        for (uint32_t j = 0; j < kSlotGroup<S>; ++j) u0[j] = S(0);
        if (base + lane < n) {
            const uint32_t s = first + base + lane;
            rec = a.sp_rec[s];
#pragma unroll
            for (uint32_t j = 0; j < kSlotGroup<S>; ++j)
                if (j < a.max_slots) u0[j] = a.st_u[j * a.sp_pad + s];
        }
    }
};

// The special cells of a block's tiles and z-chunk: every wave of the block
// takes batches of 64 records from a counter in LDS until none are left. A
// wave claims and reads its next batch before it updates the current one, so
// the record reads overlap the neighbour reads.
template <typename T, typename S, int TY>
__device__ __forceinline__ void special_share(const StepParams<T, S>& a, const uint32_t (&bnd)[TY + 1], uint32_t tp0,
                                              uint32_t z0, uint32_t* handed) {
    const uint32_t lane = threadIdx.x & (kTileX - 1);
    const uint32_t first = bnd[0], n = bnd[TY] - bnd[0];
    auto claim = [&]() -> uint32_t {
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(handed, kTileX);
        return __builtin_amdgcn_readfirstlane(b);
    };
    SpecialBatch<S> cur;
    cur.base = claim();
    if (cur.base >= n) return;
    cur.load(a, first, n, lane);
    while (cur.base < n) {
        SpecialBatch<S> nxt;
        nxt.base = claim();
        nxt.load(a, first, n, lane);
        if (cur.base + lane < n) special_one<T, S, TY>(a, bnd, tp0, z0, first + cur.base + lane, cur.rec, cur.u0);
        cur = nxt;
    }
}

// Waves of a block that start on its special cells at once: one, and four in
// blocks of a single tile, the small grids' blocks, where the special cells
// of a floor or wall are the longest path of a launch. Per step against one
// special wave (FP32, 2-plane chunks): 36 K cells 4.2 to 3.3 us, 0.38 M 7.6
// to 6.0, 0.83 M 9.5 to 8.2-8.8, 1.53 M 11.4 to 10.6-10.9, 1.93 M 13.8 to
// 13.0-13.7 (two waves: 12.6-12.9 there but 4.2, 7.7, 9.2 and 11.1 below;
// three, five and seven: slower than four below 1.5 M cells). With three or
// seven tiles a block gains nothing from a second special wave.
constexpr int special_waves(int ty) { return ty == 1 ? 4 : 1; }

// One block: TY tiles of the plane over one z-chunk, a wave each, and the
// special-cell waves, which start on the block's special cells at once; the
// tile waves join them when their z-march ends, so special cells neither
// wait for the z-march nor hold it up, and blocks with many of them (floors,
// walls) share them among all their waves.
template <typename T, typename S, int TY>
__device__ __forceinline__ void block_part(const StepParams<T, S>& a, uint32_t t) {
    __shared__ uint32_t handed;  // special records handed out so far
    const uint32_t tb = t % a.tile_blocks;
    const uint32_t zc = t / a.tile_blocks;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTileX);
    const uint32_t tp0 = tb * TY;
    const uint32_t wbase = zc * a.n_tiles;
    uint32_t bnd[TY + 1];
#pragma unroll
    for (int k = 0; k <= TY; ++k) bnd[k] = a.seg_base[(wbase + min(tp0 + k, a.n_tiles)) * a.seg_stride];
    if (threadIdx.x == 0) handed = 0;
    __syncthreads();
    const uint32_t z0 = 1 + zc * a.zchunk;
    if (wave < TY && tp0 + wave < a.n_tiles) tile_march<T, S>(a, tp0 + wave, z0);
    special_share<T, S, TY>(a, bnd, tp0, z0, &handed);
}

// One time step: tile blocks, then receiver blocks (p^n before its update,
// fdtd.rs:673-683). Consecutive blocks go to the 8 XCDs round-robin, so
// neighbouring tiles run at the same time and spread over all HBM stacks.
template <typename T, typename S, int TY>
__global__ __launch_bounds__(kTileX*(TY + special_waves(TY))) void k_step(StepParams<T, S> a) {
    if (blockIdx.x < a.main_blocks) {
        block_part<T, S, TY>(a, blockIdx.x);
        return;
    }
    const uint32_t r = (blockIdx.x - a.main_blocks) * blockDim.x + threadIdx.x;
    if (r < a.n_recv) a.ir[r] = a.p[a.recv[r]];
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Step-kernel variants: TY tiles per block plus the special-cell waves, so
// blocks of 8, 4, 16 and 5 waves (the first three divide the 32 wave slots
// of a CU at the kernel's register use, at most 64 VGPRs: 8 waves per SIMD;
// six blocks of five fill 30). The default variant and z-chunk.
constexpr int kVariants = 4;
constexpr int kDefaultVariant = 0;
constexpr uint32_t kDefaultZChunk = 4;
// A plain run's default shape by grid size. A small grid runs few blocks, and
// the special cells of a block with a floor or wall in it are the longest
// path of its launch, so small grids take blocks of one tile (and four
// special-cell waves, special_waves) and short z-chunks. Measured per step
// against 7-tile blocks with 4-plane chunks, FP32: 36 K cells 7.7 to 3.3 us
// (1-plane chunks, under 128 Ki cells); 0.38 and 0.83 M cells 11.0 and 11.9
// to 6.0 and 8.0 us (2-plane chunks, under 1 Mi); 1.2-2.5 M cells, 24 rooms,
// 1.01-1.59 times as fast (4-plane chunks, under 2.5 Mi); from 2.5 M cells
// mixed (0.89-1.19), above 4 M slower. FP64 (tiles half as wide): 36 K cells
// 6.4 to 3.4 us (2-plane chunks), 0.38 and 0.83 M 10.1 and 14.4 to 9.9 and
// 12.5 us (4-plane chunks), slower from 1.2 M cells. Batches keep
// kDefaultVariant and kDefaultZChunk unless tuned: with the whole GPU busy
// the larger blocks win there.
constexpr uint64_t kTinyStepCells = 128ull << 10;
constexpr uint64_t kSmallStepCells = 1ull << 20;
constexpr uint64_t kMidStepCells = 5ull << 19;  // 2.5 Mi
inline int default_step_variant(uint64_t cells, bool fp64) {
    return cells < (fp64 ? kSmallStepCells : kMidStepCells) ? 3 : kDefaultVariant;
}
inline uint32_t default_step_zchunk(uint64_t cells, bool fp64) {
    if (fp64) return cells < kTinyStepCells ? 2u : kDefaultZChunk;
    return cells < kTinyStepCells ? 1u : cells < kSmallStepCells ? 2u : kDefaultZChunk;
}
// Default steps per pass by grid size (measured on the hall, on hall-shaped
// grids of 58, 29, 15 and 7.6 Mi cells and on rooms of 0.04-7.5 M cells):
// small grids are launch- and latency-bound, where one step per launch wins;
// two steps per pass win from 2.25 Mi cells in FP32 (on 26 rooms of 1.9-3.2
// M cells single steps were 4-31 % faster up to 2.3 M, two steps 2-23 %
// faster from 2.5 M) and from 0.75 Mi in FP64 (2-44 % faster from 0.8 M,
// slower below with 8-plane chunks), and three from 8 Mi cells in FP32 and
// from 7 Mi in FP64 (which gains
// more: 11-21 % against 0-17 % in FP32 on those grids), with z-chunks that
// keep enough tile blocks in flight (pass_zchunk).
constexpr uint64_t kPairMinCells = 9ull << 18;      // 2.25 Mi
constexpr uint64_t kPairMinCellsFp64 = 3ull << 18;  // 0.75 Mi
constexpr uint64_t kPairLongChunkCells = 64ull << 20;
constexpr uint64_t kTriMinCells = 8ull << 20;
constexpr uint64_t kTriMinCellsFp64 = 7ull << 20;
inline uint32_t default_steps_per_pass(uint64_t cells, bool fp64) {
    if (cells >= (fp64 ? kTriMinCellsFp64 : kTriMinCells)) return 3u;
    return cells >= (fp64 ? kPairMinCellsFp64 : kPairMinCells) ? 2u : 1u;
}
constexpr uint32_t kVariantTiles[] = {7, 3, 15, 1};

inline uint32_t variant_threads(int variant) {
    return kTileX * (kVariantTiles[variant] + special_waves((int)kVariantTiles[variant]));
}

template <class T, class S>
void launch_step(int variant, uint32_t grid, hipStream_t st, const StepParams<T, S>& a) {
    const dim3 blk(variant_threads(variant));
    switch (variant) {
        case 0: hipLaunchKernelGGL((k_step<T, S, 7>), dim3(grid), blk, 0, st, a); break;
        case 1: hipLaunchKernelGGL((k_step<T, S, 3>), dim3(grid), blk, 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_step<T, S, 15>), dim3(grid), blk, 0, st, a); break;
        default: hipLaunchKernelGGL((k_step<T, S, 1>), dim3(grid), blk, 0, st, a); break;
    }
}

}  // namespace wv
