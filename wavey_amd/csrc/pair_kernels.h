// Two time steps per pass over the fields (temporal blocking in z), for the
// FDTD step of step_kernels.h.
//
// A single step moves 12 bytes per FP32 air cell (p read, prev read and
// written) plus the halo re-reads of its tiles. Two steps n, n+1 in one
// z-march read p^n and p^{n-1} once and write p^{n+1} and p^{n+2}, with four
// field buffers so no input is overwritten while a neighbouring block still
// reads it:
//
//  * Pass 1, tile blocks: a block covers 256 FP32 (128 FP64) cells of a row
//    by kPairRows rows plus one halo row above and below (a wave per row, 16
//    bytes per lane) and marches a z-chunk. Per plane it computes p^{n+1} of
//    every cell of its
//    rows (the interior update; only interior cells are stored), shares the
//    plane's rows through LDS, and one plane behind computes p^{n+2} of its
//    deep cells: interior cells whose six neighbours are interior, so every
//    value they need is an interior p^{n+1} the block has itself. Tiles step
//    252 (FP64 124) columns and store p^{n+2} on columns 2 .. 253 of the
//    tile, where it needs no p^{n+1} beyond the tile's ends.
//  * Pass 1, special blocks: step n of every special cell (boundary solve,
//    tap injection), p^{n+1} written to the p^{n+1} buffer; and the
//    receivers' p^n.
//  * Pass 2: step n+1 of the special cells and of the ring (interior cells
//    that are not deep), from the complete p^{n+1}, in blocks of one kind
//    interleaved in proportion to the two counts (both lists run through
//    the grid in z order, so the blocks of a band, on one XCD, share the
//    cache lines around the same walls); and the receivers' p^{n+1}.
//
// Every cell's arithmetic is the single-step kernel's (interior_update and
// special_update with the same operands), so results are bit-identical to
// single steps.
#pragma once
#include "step_kernels.h"

namespace wv {

constexpr int kPairRows = 14;                           // output rows of a pass-1 tile block
constexpr int kPairWaves = kPairRows + 2;               // and a halo row above and below
constexpr uint32_t kPairThreads = kTileX * kPairWaves;  // 1024
constexpr uint32_t kPair2Block = 128;  // pass-2 items per block (specials or ring cells)
// Pass 2's item blocks run in kPair2Bands bands of consecutive blocks (the
// lists are in cell order), dealt to the eight XCDs in turn: a band's cells
// share cache lines with their neighbours in one L2, and each XCD's bands
// lie all along the lists, so dense regions (a floor) spread over the XCDs.
constexpr uint32_t kPair2Bands = 128;

template <typename T, typename S>
struct PairParams {
    StepParams<T, S> a;  // step n: p = p^n, prev = p^{n-1}; ir = row n; s_n = s[n]
    T* __restrict__ p1;  // p^{n+1}
    T* __restrict__ p2;  // p^{n+2}
    const uint32_t* __restrict__ dbits;    // deep bits, laid out as the interior bits
    const uint32_t* __restrict__ sp_cell;  // pitched index of each special record
    const uint32_t* __restrict__ ring_cell;
    uint32_t n_special, n_ring;
    uint32_t tiles_x, row_blocks;  // pass-1 tiles: 64 W cells by kPairRows rows
    uint32_t tile_blocks;          // tiles_x * row_blocks * z-chunks
    uint32_t tile_launch;          // launched tile blocks: tile_blocks / tiles_x rounded up to 8, times tiles_x
    uint32_t special_blocks;       // pass-1 blocks of kPairThreads special cells
};

// Cells per lane of a pass-1 tile: 16 bytes (4 FP32 or 2 FP64); 8-byte lanes
// fit two 16-wave blocks per CU but measured slower.
template <typename T>
constexpr int kPairLane = 16 / (int)sizeof(T);
// Columns a pass-1 tile steps by: the tile less two columns at each end
// (252 FP32, 124 FP64). A tile forms p^{n+1} on all its columns (from p^n of
// the column beyond each end) and the later levels on fewer.
template <typename T>
constexpr uint32_t kTileStep = kTileX * kPairLane<T> - 4;

template <typename T, typename S>
__device__ __forceinline__ void pair_tile(const PairParams<T, S>& pp, uint32_t t) {
    const StepParams<T, S>& a = pp.a;
    constexpr int W = kPairLane<T>;
    constexpr uint32_t BX = kTileX * W;
    using C = Cells<T, W>;
    // p^{n+1} rows of the last three planes (written at plane zq, read one
    // plane later; the third buffer lets a wave write the next plane while
    // others still read the previous one, with one barrier per plane).
    __shared__ __align__(16) T rows[3][kPairWaves][BX];
    // Blocks b, b+8, b+16, ... run on the same XCD (the hardware's
    // round-robin): one XCD takes all x-tiles of a (row block, z-chunk) in
    // turn, so the columns the tiles share at their ends stay in its L2;
    // neighbouring row blocks run on the other XCDs at the same time.
    const uint32_t xcd = t & 7u, q = t >> 3;
    const uint32_t tx = q % pp.tiles_x;
    const uint32_t r = (q / pp.tiles_x) * 8u + xcd;
    if (r >= pp.tile_blocks / pp.tiles_x) return;  // padding of the last group (the whole block)
    const uint32_t rb = r % pp.row_blocks;
    const uint32_t zc = r / pp.row_blocks;
    const uint32_t lane = threadIdx.x & (kTileX - 1);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTileX);
    const int ny = (int)a.ny, nz = (int)a.nz;
    // Row of this wave; rows outside the grid (the halo of the first and last
    // blocks) read the nearest row and store nothing.
    const int y = (int)(rb * kPairRows + wave) - 1;
    const bool out_row = wave >= 1 && wave <= (uint32_t)kPairRows && y < ny;
    const uint32_t yc = (uint32_t)min(max(y, 0), ny - 1);
    const uint32_t ymc = (uint32_t)min(max(y - 1, 0), ny - 1);
    const uint32_t ypc = (uint32_t)min(max(y + 1, 0), ny - 1);
    constexpr uint32_t SX = kTileStep<T>;
    const uint32_t x0 = tx * SX;
    const uint32_t xl = x0 + lane * W;
    const bool live = xl < a.pitch;
    const uint32_t x = live ? xl : a.pitch - W;
    const uint32_t col = yc * a.pitch + x, colm = ymc * a.pitch + x, colp = ypc * a.pitch + x;
    // Lanes 0 and 63: p^n of the column next to the tile (xh).
    const bool edge = lane == 0 || lane == kTileX - 1;
    const uint32_t xh = lane == 0 ? (x0 > 0 ? x0 - 1 : 0u) : min(x0 + BX, a.pitch - 1);
    const uint32_t hcol = yc * a.pitch + xh;
    const int z0 = 1 + (int)(zc * a.zchunk);
    const int z1 = min(z0 + (int)a.zchunk, nz - 1);
    const uint32_t full = (1u << W) - 1u;
    // Stored cells of this lane in an output row: p^{n+2} on columns x0 + 2
    // .. x0 + SX + 1 (each column of the grid in one tile), p^{n+1}, right on
    // every column of the tile, also on x0 + 1 and x0 + SX + 2, so that
    // columns 1 and nx - 2 are covered (the tiles beside write the same
    // values there).
    uint32_t own = 0, own1 = 0;
#pragma unroll
    for (int e = 0; e < W; ++e) {
        own |= (uint32_t)(xl + e >= x0 + 2 && xl + e <= x0 + SX + 1) << e;
        own1 |= (uint32_t)(xl + e >= x0 + 1 && xl + e <= x0 + SX + 2) << e;
    }
    const uint32_t lane_mask = (live && out_row) ? own : 0u;
    const uint32_t lane_mask1 = (live && out_row) ? own1 : 0u;
    auto plane = [&](int z) -> uint32_t { return (uint32_t)min(max(z, 0), nz - 1) * a.sz; };
    // The chunk lies inside one 32-plane block of the z-packed bitmaps.
    uint32_t ib[W], db[W];
    load_words<W>(a.zbits + ((z0 - 1) >> 5) * a.sz + col, ib);
    load_words<W>(pp.dbits + ((z0 - 1) >> 5) * a.sz + col, db);

    // A plane's loads: p^n of the row at zq + 1 and of the rows above and
    // below at zq, p^{n-1} of the row at zq, and (edge lanes) p^n of the
    // column next to the tile at zq + 1. They are issued one plane ahead, so
    // they are in flight during the previous plane's compute and barrier.
    struct PlaneLoads {
        C zp, q, ym, yp;
        T hp;
    };
    auto issue = [&](int zq, PlaneLoads& L) {
        const uint32_t pz = plane(zq), pzp = plane(zq + 1);
        L.zp.load(a.p + pzp + col);
        L.q.load(a.prev + pz + col);
        L.ym.load(a.p + pz + colm);
        L.yp.load(a.p + pz + colp);
        L.hp = T(0);
        if (edge) L.hp = a.p[pzp + hcol];
    };

    // p^n of the row at planes zq - 1 and zq, and of the column next to the
    // tile at zq.
    C zm, zcur;
    zm.load(a.p + plane(z0 - 2) + col);
    zcur.load(a.p + plane(z0 - 1) + col);
    T hc = 0;
    if (edge) hc = a.p[plane(z0 - 1) + hcol];
    // p^{n+1} of the row at planes zq - 2 and zq - 1.
    C p1m, p1c;
#pragma unroll
    for (int e = 0; e < W; ++e) p1m.v[e] = p1c.v[e] = T(0);
    // zp - c at the plane below of the p^n and p^{n+1} rows (FP32), carried
    // from plane to plane.
    T dz1[W], dz2[W];
#pragma unroll
    for (int e = 0; e < W; ++e) {
        dz1[e] = zcur.v[e] - zm.v[e];
        dz2[e] = T(0);
    }
    uint32_t bw = 0;  // LDS buffer of plane zq
    PlaneLoads nx;
    issue(z0 - 1, nx);
    // p^{n+1} is needed at planes z0 - 1 .. z1 for p^{n+2} at z0 .. z1 - 1.
    for (int zq = z0 - 1; zq <= z1; ++zq) {
        const PlaneLoads cur = nx;
        if (zq < z1) issue(zq + 1, nx);
        const uint32_t pz = plane(zq);
        // p^{n+1} at plane zq of the row.
        const T lo = from_lower_lane(zcur.v[W - 1], hc);
        const T hi = from_upper_lane(zcur.v[0], hc);
        C p1;
        row_update<T, W>(zcur.v, cur.q.v, lo, hi, cur.ym.v, cur.yp.v, zm.v, cur.zp.v, dz1, dz1, a.l2, a.c2, p1.v);
        p1.store(&rows[bw][wave][lane * W]);
        if (zq >= z0 && zq < z1) {
            const uint32_t bits = plane_bits<W>(ib, (uint32_t)(zq - 1) & 31u) & lane_mask1;
            if (bits == full) {
                p1.store(pp.p1 + pz + col);
            } else if (bits) {
#pragma unroll
                for (int e = 0; e < W; ++e)
                    if ((bits >> e) & 1u) pp.p1[pz + col + e] = p1.v[e];
            }
        }
        __syncthreads();
        // p^{n+2} at plane zo = zq - 1 of the deep cells of the output rows.
        const int zo = zq - 1;
        bool level2 = false;  // p^{n+2} formed this plane (uniform in the wave)
        if (zo >= z0 && out_row) {
            const uint32_t dmask = plane_bits<W>(db, (uint32_t)(zo - 1) & 31u) & lane_mask;
            if (__ballot(dmask != 0)) {
                const uint32_t br = bw == 0 ? 2u : bw - 1;  // buffer of plane zo
                C ym1, yp1;
                ym1.load(&rows[br][wave - 1][lane * W]);
                yp1.load(&rows[br][wave + 1][lane * W]);
                const T lo1 = from_lower_lane(p1c.v[W - 1], T(0));  // the tile's end columns are
                const T hi1 = from_upper_lane(p1c.v[0], T(0));      // not stored at this level
                C p2;
                row_update<T, W>(p1c.v, zm.v, lo1, hi1, ym1.v, yp1.v, p1m.v, p1.v, dz2, dz2, a.l2, a.c2, p2.v);
                level2 = true;
                const uint32_t po = (uint32_t)zo * a.sz;
                if (dmask == full) {
                    p2.store(pp.p2 + po + col);
                } else if (dmask) {
#pragma unroll
                    for (int e = 0; e < W; ++e)
                        if ((dmask >> e) & 1u) pp.p2[po + col + e] = p2.v[e];
                }
            }
        }
        if (!level2) {
            // The next plane's p^{n+2} reads this plane's difference.
#pragma unroll
            for (int e = 0; e < W; ++e) dz2[e] = p1.v[e] - p1c.v[e];
        }
        p1m = p1c;
        p1c = p1;
        zm = zcur;
        zcur = cur.zp;
        hc = cur.hp;
        bw = bw == 2 ? 0u : bw + 1;
    }
}

// The blocks of pass 1 after its tile blocks (block b counts from the first
// of them): step n of the special cells, p^{n+1} into p1; then the
// receivers' p^n.
template <typename T, typename S>
__device__ __forceinline__ void pass1_specials(const StepParams<T, S>& a, const uint32_t* __restrict__ sp_cell,
                                               uint32_t n_special, uint32_t special_blocks, T* __restrict__ p1,
                                               uint32_t b) {
    if (b < special_blocks) {
        const uint32_t s = b * kPairThreads + threadIdx.x;
        if (s >= n_special) return;
        const uint2 rec = a.sp_rec[s];
        S u0[kSlotGroup<S>];
#pragma unroll
        for (uint32_t j = 0; j < kSlotGroup<S>; ++j) u0[j] = j < a.max_slots ? a.st_u[j * a.sp_pad + s] : S(0);
        special_at(a, s, sp_cell[s], rec, u0, p1);
        return;
    }
    const uint32_t rr = (b - special_blocks) * kPairThreads + threadIdx.x;
    if (rr < a.n_recv) a.ir[rr] = a.p[a.recv[rr]];
}

// Pass 1: tile blocks, then blocks of special cells (step n), then the
// receivers' p^n.
template <typename T, typename S>
__global__ __launch_bounds__(kPairThreads) void k_pair1(PairParams<T, S> pp) {
    if (blockIdx.x < pp.tile_launch) {
        pair_tile<T, S>(pp, blockIdx.x);
        return;
    }
    pass1_specials<T, S>(pp.a, pp.sp_cell, pp.n_special, pp.special_blocks, pp.p1, blockIdx.x - pp.tile_launch);
}

// Pass 2: step n+1 of the special cells and the ring, then the receivers'
// p^{n+1}. Special and ring blocks are interleaved in proportion to their
// counts, so both lists advance through the grid together.
template <typename T, typename S>
__global__ __launch_bounds__(kPair2Block) void k_pair2(StepParams<T, S> a, T* __restrict__ p2,
                                               const uint32_t* __restrict__ sp_cell,
                                               const uint32_t* __restrict__ ring_cell, uint32_t n_special,
                                               uint32_t n_ring, uint32_t item_blocks) {
    // Launch block i runs on XCD i mod 8; its item block is in band
    // (i mod 8) + 8 k, k = (i / 8) / band.
    const uint32_t band = (item_blocks + kPair2Bands - 1u) / kPair2Bands;
    if (blockIdx.x < band * kPair2Bands) {
        const uint32_t q = blockIdx.x >> 3;
        const uint32_t b = ((blockIdx.x & 7u) + 8u * (q / band)) * band + q % band;
        if (b >= item_blocks) return;
        const uint32_t bs = (n_special + kPair2Block - 1) / kPair2Block;  // special blocks
        const uint32_t s0 = (uint32_t)(((uint64_t)b * bs) / item_blocks);
        const uint32_t s1 = (uint32_t)(((uint64_t)(b + 1) * bs) / item_blocks);
        if (s1 > s0) {
            const uint32_t t = s0 * kPair2Block + threadIdx.x;
            if (t < n_special) {
                const uint2 rec = a.sp_rec[t];
                S u0[kSlotGroup<S>];
#pragma unroll
                for (uint32_t j = 0; j < kSlotGroup<S>; ++j) u0[j] = j < a.max_slots ? a.st_u[j * a.sp_pad + t] : S(0);
                special_at(a, t, sp_cell[t], rec, u0, p2);
            }
        } else {
            const uint32_t r = (b - s1) * kPair2Block + threadIdx.x;  // ring blocks before this one: b - s1
            if (r < n_ring) {
                const uint32_t i = ring_cell[r];
                p2[i] = interior_update(a.p[i], a.prev[i], a.p[i - 1], a.p[i + 1], a.p[i - a.pitch], a.p[i + a.pitch],
                                        a.p[i - a.sz], a.p[i + a.sz], a.l2, a.c2);
            }
        }
        return;
    }
    const uint32_t rr = (blockIdx.x - band * kPair2Bands) * kPair2Block + threadIdx.x;
    if (rr < a.n_recv) a.ir[rr] = a.p[a.recv[rr]];
}

}  // namespace wv
