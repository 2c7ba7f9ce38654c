// Three time steps per pass over the fields (temporal blocking in z): the
// two-step pass of pair_kernels.h taken one step further.
//
// Two steps per pass read p^n and p^{n-1} once and write p^{n+1} and p^{n+2}
// (16 bytes per FP32 cell per two steps). Three steps read the same two
// fields and write p^{n+2} and p^{n+3} of the bulk; p^{n+1} is stored only
// where the later passes read it, near the walls (16 bytes per three steps),
// with five field buffers so no input is overwritten while a neighbouring
// block still reads it:
//
//  * Pass 1, tile blocks: a block covers 256 FP32 (128 FP64) cells of a row
//    by 16 rows (a wave per row) and marches a z-chunk. Per plane it computes
//    p^{n+1} of every row (from p^n of the tile and of the column next to
//    each end), one plane behind p^{n+2} of rows 1..14 from the p^{n+1} rows
//    shared through LDS, and two planes behind p^{n+3} of rows 2..13 from the
//    p^{n+2} rows shared through LDS. p^{n+1} is right on every column of the
//    tile, p^{n+2} on all but the two end columns, p^{n+3} on columns
//    2 .. 253 (2 .. 125 FP64): a block owns rows 2..13 and those columns, and
//    tiles step 252 (124) cells and 12 rows, so no tile forms a value beyond
//    its ends. It stores p^{n+1} of the interior cells that are not deep2
//    (also on columns 1 and 254 (126), which the tile beside writes with the
//    same value, so that columns 1 and nx - 2 are covered), p^{n+2} of the
//    deep cells and p^{n+3} of the deep2 cells (deep cells whose six
//    neighbours are deep: their third step needs only values the block
//    computes itself).
//  * Pass 1, special blocks and receivers: as with two steps per pass.
//  * Pass 2: step n+1 of the special cells and of the ring (interior, not
//    deep), as k_pair2.
//  * Pass 3: step n+2 of the special cells and of the interior cells that are
//    not deep2, as k_pair2 with that list.
//
// Receiver cells are kept out of the deep2 set, so their p^{n+1} is stored
// for pass 2's record. Every cell's arithmetic is the single-step kernel's,
// so results are bit-identical to single steps.
#pragma once
#include "pair_kernels.h"

namespace wv {

constexpr int kTriRows = kPairWaves - 4;  // rows a pass-1 block owns (12)

// Columns a pass-1 tile steps by: the tile less two columns at each end
// (252 FP32, 124 FP64). A tile forms p^{n+1} on all its columns (from p^n of
// the column beyond each end), p^{n+2} on all but its end columns and
// p^{n+3} on the stride's columns 2 .. stride + 1 of the tile.
template <typename T>
constexpr uint32_t kTriStride = kTileStep<T>;

template <typename T, typename S>
struct TriParams {
    StepParams<T, S> a;  // step n: p = p^n, prev = p^{n-1}; ir = row n; s_n = s[n]
    T* __restrict__ p1;  // p^{n+1}
    T* __restrict__ p2;  // p^{n+2}
    T* __restrict__ p3;  // p^{n+3}
    const uint32_t* __restrict__ dbits;   // deep bits, laid out as the interior bits
    const uint32_t* __restrict__ d2bits;  // deep2 bits (receiver cells cleared)
    const uint32_t* __restrict__ sp_cell;
    uint32_t n_special;
    uint32_t tiles_x, row_blocks;  // pass-1 tiles: kTriStride columns by kTriRows rows
    uint32_t tile_blocks;          // tiles_x * row_blocks * z-chunks
    uint32_t tile_launch;          // tile_blocks / tiles_x rounded up to 8, times tiles_x
    uint32_t special_blocks;       // pass-1 blocks of kPairThreads special cells
};

// Stores the cells of c whose bit is set (one access when all are).
template <typename T, int W>
__device__ __forceinline__ void store_masked(const Cells<T, W>& c, T* __restrict__ p, uint32_t bits) {
    if (bits == (1u << W) - 1u) {
        c.store(p);
    } else if (bits) {
#pragma unroll
        for (int e = 0; e < W; ++e)
            if ((bits >> e) & 1u) p[e] = c.v[e];
    }
}

template <typename T, typename S>
__device__ __forceinline__ void tri_tile(const TriParams<T, S>& tp, uint32_t t) {
    const StepParams<T, S>& a = tp.a;
    constexpr int W = kPairLane<T>;
    constexpr uint32_t BX = kTileX * W;
    constexpr uint32_t SX = kTriStride<T>;
    using C = Cells<T, W>;
    // p^{n+1} and p^{n+2} rows of the last three planes of each (written at
    // one plane, read one plane later; one barrier per plane).
    __shared__ __align__(16) T rows1[3][kPairWaves][BX];
    __shared__ __align__(16) T rows2[3][kPairWaves][BX];
    // XCD-ordered blocks, as the two-step pass: one XCD takes all x-tiles of
    // a (row block, z-chunk) in turn.
    const uint32_t xcd = t & 7u, q = t >> 3;
    const uint32_t tx = q % tp.tiles_x;
    const uint32_t r = (q / tp.tiles_x) * 8u + xcd;
    if (r >= tp.tile_blocks / tp.tiles_x) return;  // padding of the last group (the whole block)
    const uint32_t rb = r % tp.row_blocks;
    const uint32_t zc = r / tp.row_blocks;
    const uint32_t lane = threadIdx.x & (kTileX - 1);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTileX);
    // The waves of the owned rows form all three levels each plane, the four
    // halo rows one or two: the owned rows' waves issue first (hall 486 to
    // 508 Gcell-updates/s; the same in the two-step tile measured slower).
    if (wave >= 2 && wave < 2u + (uint32_t)kTriRows) __builtin_amdgcn_s_setprio(1);
    const int ny = (int)a.ny, nz = (int)a.nz;
    // Row of this wave: two rows of halo above and below the owned ones.
    const int y = (int)(rb * kTriRows + wave) - 2;
    const bool own_row = wave >= 2 && wave < 2u + (uint32_t)kTriRows && y < ny;
    const uint32_t yc = (uint32_t)min(max(y, 0), ny - 1);
    const uint32_t ymc = (uint32_t)min(max(y - 1, 0), ny - 1);
    const uint32_t ypc = (uint32_t)min(max(y + 1, 0), ny - 1);
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
    // Stored cells of this lane in an owned row: p^{n+2} and p^{n+3} on
    // columns x0 + 2 .. x0 + SX + 1 (each column of the grid in one tile), and
    // p^{n+1}, right on every column of the tile, also on x0 + 1 and
    // x0 + SX + 2, so that columns 1 and nx - 2 are covered (the tiles beside
    // write the same values there).
    uint32_t own = 0, own1 = 0;
#pragma unroll
    for (int e = 0; e < W; ++e) {
        own |= (uint32_t)(xl + e >= x0 + 2 && xl + e <= x0 + SX + 1) << e;
        own1 |= (uint32_t)(xl + e >= x0 + 1 && xl + e <= x0 + SX + 2) << e;
    }
    const uint32_t lane_mask = (live && own_row) ? own : 0u;
    const uint32_t lane_mask1 = (live && own_row) ? own1 : 0u;
    auto plane = [&](int z) -> uint32_t { return (uint32_t)min(max(z, 0), nz - 1) * a.sz; };
    // The chunk covers one 32-plane block of the z-packed bitmaps or, for
    // FP32 chunks of up to 64 planes (which start at a 64-plane boundary),
    // two: per cell the p^{n+1} store bits (interior, not deep2), the deep
    // and the deep2 bits of each block.
    constexpr int NW = sizeof(T) == 4 ? 2 : 1;
    const uint32_t kb0 = (uint32_t)(z0 - 1) >> 5;
    const uint32_t nzb = ((uint32_t)nz - 2 + 31) >> 5;
    uint32_t s1[NW][W], db[NW][W], d2b[NW][W];
#pragma unroll
    for (int h = 0; h < NW; ++h) {
        uint32_t ib[W];
        if (h == 0 || (a.zchunk > 32 && kb0 + h < nzb)) {
            const uint32_t wb = (kb0 + h) * a.sz + col;
            load_words<W>(a.zbits + wb, ib);
            load_words<W>(tp.dbits + wb, db[h]);
            load_words<W>(tp.d2bits + wb, d2b[h]);
        } else {
#pragma unroll
            for (int e = 0; e < W; ++e) ib[e] = db[h][e] = d2b[h][e] = 0u;
        }
#pragma unroll
        for (int e = 0; e < W; ++e) s1[h][e] = ib[e] & ~d2b[h][e];
    }
    // The W cells' bits of plane z (inside the chunk).
    auto bits_of = [&](const uint32_t (&w)[NW][W], int z) -> uint32_t {
        const uint32_t k = (uint32_t)(z - 1) - 32u * kb0;
        if constexpr (NW == 1) return plane_bits<W>(w[0], k);
        else return k < 32u ? plane_bits<W>(w[0], k) : plane_bits<W>(w[1], k - 32u);
    };

    // A plane's loads, issued one plane ahead: p^n of the row at zq + 1 and
    // of the rows above and below at zq, p^{n-1} of the row at zq, and (edge
    // lanes) p^n of the column next to the tile at zq + 1.
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
    zm.load(a.p + plane(z0 - 3) + col);
    zcur.load(a.p + plane(z0 - 2) + col);
    T hc = 0;
    if (edge) hc = a.p[plane(z0 - 2) + hcol];
    // p^{n+2} of the row at the last plane formed (for the next p^{n+3}).
    // The older rows of this wave
    // (p^{n+1} at planes zq - 1 and zq - 2, p^{n+2} at zq - 2 and zq - 3) are
    // read back from the LDS rings, which hold every wave's own row.
    C p2;
#pragma unroll
    for (int e = 0; e < W; ++e) p2.v[e] = T(0);
    // zp - c at the plane below of the p^n, p^{n+1} and p^{n+2} rows (FP32),
    // carried from plane to plane.
    T dz1[W], dz2[W], dz3[W];
#pragma unroll
    for (int e = 0; e < W; ++e) {
        dz1[e] = zcur.v[e] - zm.v[e];
        dz2[e] = dz3[e] = T(0);
    }
    uint32_t bw = 0;  // LDS buffer of plane zq
    PlaneLoads nx;
    issue(z0 - 2, nx);
    // p^{n+1} at planes z0 - 2 .. z1 + 1, p^{n+2} at z0 - 1 .. z1, p^{n+3} at z0 .. z1 - 1.
    for (int zq = z0 - 2; zq <= z1 + 1; ++zq) {
        const PlaneLoads cur = nx;
        if (zq <= z1) issue(zq + 1, nx);
        // p^{n+1} at plane zq of the row.
        const T lo = from_lower_lane(zcur.v[W - 1], hc);
        const T hi = from_upper_lane(zcur.v[0], hc);
        C p1;
        row_update<T, W>(zcur.v, cur.q.v, lo, hi, cur.ym.v, cur.yp.v, zm.v, cur.zp.v, dz1, dz1, a.l2, a.c2, p1.v);
        p1.store(&rows1[bw][wave][lane * W]);
        if (zq >= z0 && zq < z1) {
            store_masked<T, W>(p1, tp.p1 + plane(zq) + col, bits_of(s1, zq) & lane_mask1);
        }
        __syncthreads();
        const uint32_t b1 = bw == 0 ? 2u : bw - 1;  // LDS buffers of plane zq - 1
        const uint32_t b2 = b1 == 0 ? 2u : b1 - 1;  // and of plane zq - 2
        C p1c, p1m;  // this row's p^{n+1} at planes zq - 1 and zq - 2
        p1c.load(&rows1[b1][wave][lane * W]);
        p1m.load(&rows1[b2][wave][lane * W]);
        // p^{n+2} at plane zq - 1 of rows 1..14.
        if (zq >= z0 && wave >= 1 && wave <= (uint32_t)kPairWaves - 2) {
            C ym1, yp1;
            ym1.load(&rows1[b1][wave - 1][lane * W]);
            yp1.load(&rows1[b1][wave + 1][lane * W]);
            const T lo1 = from_lower_lane(p1c.v[W - 1], T(0));  // the tile's end columns are
            const T hi1 = from_upper_lane(p1c.v[0], T(0));      // not stored at this level
            row_update<T, W>(p1c.v, zm.v, lo1, hi1, ym1.v, yp1.v, p1m.v, p1.v, dz2, dz2, a.l2, a.c2, p2.v);
            p2.store(&rows2[b1][wave][lane * W]);
            const int zo = zq - 1;
            if (zo >= z0 && zo < z1)
                store_masked<T, W>(p2, tp.p2 + (uint32_t)zo * a.sz + col, bits_of(db, zo) & lane_mask);
        } else {
            // Planes and rows without p^{n+2} still carry the difference.
#pragma unroll
            for (int e = 0; e < W; ++e) dz2[e] = p1.v[e] - p1c.v[e];
        }
        // p^{n+3} at plane zq - 2 of rows 2..13 (the p^{n+2} rows of plane
        // zq - 2 were written before this plane's barrier).
        if (own_row) {
            C p2c;  // this row's p^{n+2} at plane zq - 2
            p2c.load(&rows2[b2][wave][lane * W]);
            if (zq >= z0 + 2) {
                C ym2, yp2, p2m, p3;
                ym2.load(&rows2[b2][wave - 1][lane * W]);
                yp2.load(&rows2[b2][wave + 1][lane * W]);
                p2m.load(&rows2[bw][wave][lane * W]);  // plane zq - 3 (FP64's form only)
                const T lo2 = from_lower_lane(p2c.v[W - 1], T(0));
                const T hi2 = from_upper_lane(p2c.v[0], T(0));
                row_update<T, W>(p2c.v, p1m.v, lo2, hi2, ym2.v, yp2.v, p2m.v, p2.v, dz3, dz3, a.l2, a.c2, p3.v);
                const int zo3 = zq - 2;
                store_masked<T, W>(p3, tp.p3 + (uint32_t)zo3 * a.sz + col, bits_of(d2b, zo3) & lane_mask);
            } else {
                // The first p^{n+3} plane reads the difference of the plane before.
#pragma unroll
                for (int e = 0; e < W; ++e) dz3[e] = p2.v[e] - p2c.v[e];
            }
        }
        zm = zcur;
        zcur = cur.zp;
        hc = cur.hp;
        bw = bw == 2 ? 0u : bw + 1;
    }
}

// Pass 1 of three steps: tile blocks, then blocks of special cells (step n),
// then the receivers' p^n.
template <typename T, typename S>
__global__ __launch_bounds__(kPairThreads) void k_tri1(TriParams<T, S> tp) {
    if (blockIdx.x < tp.tile_launch) {
        tri_tile<T, S>(tp, blockIdx.x);
        return;
    }
    pass1_specials<T, S>(tp.a, tp.sp_cell, tp.n_special, tp.special_blocks, tp.p1, blockIdx.x - tp.tile_launch);
}

}  // namespace wv
