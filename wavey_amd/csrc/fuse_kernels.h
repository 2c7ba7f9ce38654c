// Three time steps per pass with the boundary cells inside the march (the
// fused pass), for FP32 fields of grids whose rows are at most 768 cells.
//
// The three-step pass of tri_kernels.h updates the cells near walls in two
// more launches that gather them by index list. On small rooms those cells
// are a large share (a wall in every row) and the gathers move a cache line
// per few cells, so the fused pass does them in the same z-march:
//
//  * A block covers a span of 4096 consecutive cells of the pitched xy plane
//    (16 waves of 256 cells, 4 per lane), so narrow rows keep every lane
//    busy. y neighbours are +-pitch away inside the span. Level 1 (p^{n+1})
//    is formed on the whole span, level 2 on span positions pitch ..
//    4096 - pitch, level 3 on 2 pitch .. 4096 - 2 pitch: the block owns
//    those 4096 - 4 pitch cells of each plane, and spans step by that much.
//  * Per plane of the march (one barrier each): p^{n+1} of plane i, p^{n+2}
//    of plane i - 2 and p^{n+3} of plane i - 4. The tile part forms every
//    cell with the interior update and writes the planes to LDS rings (five
//    planes of p^{n+1}, four of p^{n+2}) except at special cells; the special
//    part, chunks of 64 special cells handed to the waves in turn, writes the
//    special cells of the same three (level, plane) pairs into the same
//    rings: level 1 from a pre-pass launch (step n of every special cell,
//    which needs only p^n and p^{n-1}), levels 2 and 3 by the boundary solve
//    from the rings, with the slot constants copied to LDS.
//  * Slot state: the pre-pass reads the state of step n (one array of a
//    pair) and writes that of step n+1 to a third array, level 2 to a fourth,
//    level 3 to the other array of the pair; every array has the plain
//    run's layout (slot-major, by special-cell rank). Blocks that overlap
//    form the same values from the same inputs.
//  * Stores: p^{n+2} and p^{n+3} of every owned air cell into the other pair
//    of field buffers (four in all), so the next pass reads them while no
//    block writes its inputs. A pass of two steps stores p^{n+1} and p^{n+2},
//    one of one step p^n and p^{n+1}, so the pairs alternate every pass.
//  * Receivers record p^n, p^{n+1} and p^{n+2} (by the block that owns the
//    cell; levels 1 and 2 from LDS).
//
// Every cell's arithmetic is the single-step kernel's (row_update and
// special_update with the same operands), so results are bit-identical to
// single steps.
#pragma once
#include "tri_kernels.h"

namespace wv {

constexpr int kFuseTileWaves = 12;                      // waves of the tile part
constexpr int kFuseSpecWaves = 4;                       // and of the special part
constexpr int kFuseWaves = kFuseTileWaves + kFuseSpecWaves;
constexpr uint32_t kFuseThreads = kTileX * kFuseWaves;  // 1024
constexpr uint32_t kFuseSeg = kTileX * 4;               // cells per tile wave
constexpr uint32_t kFuseSpan = kFuseSeg * kFuseTileWaves;  // 3072 cells per block
constexpr uint32_t kFuseR1 = 5;                         // LDS planes of p^{n+1}
constexpr uint32_t kFuseR2 = 4;                         // and of p^{n+2}
constexpr uint32_t kFusePasses = 43;                    // a 128-step block: 42 passes of three steps, one of two
constexpr uint32_t kFuseMaxPitch = 512;                 // a span owns 3072 - 4 pitch cells of a plane
constexpr uint32_t kFuseFirst = 6;                      // special-list bounds per (plane, span)
constexpr uint32_t kFuseChunk = 64;                     // special cells per chunk (a lane each)
static_assert(kFuseChunk == kTileX, "a special chunk is one wave");
// Slots of a boundary solve held at a time, and slot states read ahead with
// the cell's other inputs (up to two wall faces with three-branch
// admittances; further slots are read when solved).
constexpr uint32_t kFuseSlotGroup = 3;
constexpr uint32_t kFusePrefetch = 6;

// Cell class bytes (one per pitched cell, four to a word).
constexpr uint32_t kClsInterior = 1u;
constexpr uint32_t kClsSpecial = 2u;

template <typename S>
struct FuseEntry {
    StepParams<float, S> a;  // sizes, special records, slot constants, receivers, record base, scalars
    float* f[4];             // two pairs of field buffers (p, prev); pass k of block b reads pair (b + k) & 1
    S* su[2];                // slot state u of each pair's step
    S* sg[2];                // and g (springs)
    S* cu[2];                // state after the first and the second step of a pass
    S* cg[2];
    const uint32_t* cls;      // class bytes, four cells per word
    // Special cells in pitched order: pitched index of rank q (the plain
    // layout with one-plane chunks compacts them in that order, so a cell's
    // rank is its place in the list).
    const uint32_t* sq_cell;
    // sq_first[(z nblk + k) kFuseFirst + j]: first entry of the list at or
    // after span position 0, P, 2P, 4096 - 2P, 4096 - P, 4096 of span k in
    // plane z (P = pitch; positions clamped to the plane).
    const uint32_t* sq_first;
    const uint32_t* rx_first;  // [block]: first of its owned receivers; [blocks] = total
    const uint32_t* rx;        // owned receivers: receiver, plane, span position
    const S* sig;
    float2* p1s;  // {p^{n+1}, p^n} of each special cell (by rank), from the pre-pass
    uint32_t steps, n0;  // n0: first step of the current 128-step block
    uint32_t span_step, nblk, nzc, zchunk;
    uint32_t block0, blocks;  // first launch block and blocks per pass
    uint32_t n_special, n_setc;  // special cells; slot-constant records ({inv_d, cv, k, w} each)
    uint32_t pre0;               // first block of the pre-pass launch
};

// Interior and special bits of a lane's four cells from their class word.
__device__ __forceinline__ uint32_t cls_bits(uint32_t w) {
    return (w & 1u) | ((w >> 7) & 2u) | ((w >> 14) & 4u) | ((w >> 21) & 8u);
}

// LDS row store of the cells whose special bit is clear (the special part
// writes the others).
__device__ __forceinline__ void lds_store_tiles(float* __restrict__ d, const Cells<float, 4>& c, uint32_t special) {
    if (special == 0) {
        c.store(d);
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (!((special >> e) & 1u)) d[e] = c.v[e];
    }
}

// LDS copy of an instance's slot constants (the boundary solves of levels 2
// and 3 then read only their state from HBM); larger tables stay in HBM.
constexpr uint32_t kFuseSetBytes = 12288;

template <typename S>
struct FuseShared {
    float l1[kFuseR1][kFuseSpan];  // p^{n+1} of planes i - 4 .. i
    float l2[kFuseR2][kFuseSpan];  // p^{n+2} of planes i - 5 .. i - 2
    S setc[kFuseSetBytes / sizeof(S)];
};

__device__ __forceinline__ uint32_t fuse_slot1(int z) { return (uint32_t)(z + 5) % kFuseR1; }  // z >= -1
__device__ __forceinline__ uint32_t fuse_slot2(int z) { return (uint32_t)(z + 4) % kFuseR2; }  // z >= -1

// A lane's special cell of a fused pass, with what it reads from HBM
// loaded ahead: level lvl (step n + lvl - 1) at plane z, list entry q.
template <typename S>
struct FuseSpecial {
    uint32_t idx;
    uint2 rec;
    S u0[kFusePrefetch];
    float2 pv;  // the pre-pass's {p^{n+1}, p^n}
};

template <typename S>
__device__ __forceinline__ void fuse_special_load(const FuseEntry<S>& e, uint32_t lvl, uint32_t q, FuseSpecial<S>& f) {
    const StepParams<float, S>& a = e.a;
    f.idx = e.sq_cell[q];
    f.pv = lvl <= 2 ? e.p1s[q] : make_float2(0.f, 0.f);
    f.rec = make_uint2(0u, 0u);
#pragma unroll
    for (uint32_t j = 0; j < kFusePrefetch; ++j) f.u0[j] = S(0);
    if (lvl >= 2) {
        f.rec = a.sp_rec[q];
        const S* u_in = e.cu[lvl - 2];
#pragma unroll
        for (uint32_t j = 0; j < kFusePrefetch; ++j)
            if (j < a.max_slots) f.u0[j] = u_in[j * a.sp_pad + q];
    }
}

// Level 1 places the pre-pass's value; levels 2 and 3 solve from the LDS
// rings (the cell's span position and its neighbours lie inside the level's
// valid range by the list bounds) with the slot constants in setc.
template <typename S>
__device__ __forceinline__ void fuse_special_run(const FuseEntry<S>& e, FuseShared<S>& sh, const S* setc,
                                                 uint32_t lvl, int z, uint32_t q, const FuseSpecial<S>& f,
                                                 int span_lo, int z0, int z1, uint32_t n, uint32_t L, uint32_t par) {
    const StepParams<float, S>& a = e.a;
    const uint32_t idx = f.idx;
    const int P = (int)a.pitch;
    const int sp = (int)(idx - (uint32_t)z * a.sz) - span_lo;
    const bool owned = sp >= 2 * P && sp < (int)kFuseSpan - 2 * P && z >= z0 && z < z1;
    float* op = e.f[2 * (par ^ 1u)];
    float* oq = e.f[2 * (par ^ 1u) + 1];
    if (lvl == 1) {
        const float v = f.pv.x;
        if (L >= 2) sh.l1[fuse_slot1(z)][sp] = v;
        if (owned) {
            if (L == 1) {
                op[idx] = v;
                oq[idx] = f.pv.y;
            } else if (L == 2) {
                oq[idx] = v;
            }
        }
        return;
    }
    const uint32_t meta = f.rec.x;
    S pc, pp;
    S nb[6];
    SlotIo<S> io;
    if (lvl == 2) {
        const float* c = sh.l1[fuse_slot1(z)];
        pc = (S)c[sp];
        pp = (S)f.pv.y;
        nb[0] = (S)c[sp - ((meta & 1u) ? 1 : 0)];
        nb[1] = (S)c[sp + ((meta & 2u) ? 1 : 0)];
        nb[2] = (S)c[sp - ((meta & 4u) ? P : 0)];
        nb[3] = (S)c[sp + ((meta & 8u) ? P : 0)];
        nb[4] = (S)((meta & 16u) ? sh.l1[fuse_slot1(z - 1)][sp] : c[sp]);
        nb[5] = (S)((meta & 32u) ? sh.l1[fuse_slot1(z + 1)][sp] : c[sp]);
        io = SlotIo<S>{e.cu[0], e.cg[0], L == 2 ? e.su[par ^ 1u] : e.cu[1], L == 2 ? e.sg[par ^ 1u] : e.cg[1], setc};
    } else {
        const float* c = sh.l2[fuse_slot2(z)];
        pc = (S)c[sp];
        pp = (S)sh.l1[fuse_slot1(z)][sp];
        nb[0] = (S)c[sp - ((meta & 1u) ? 1 : 0)];
        nb[1] = (S)c[sp + ((meta & 2u) ? 1 : 0)];
        nb[2] = (S)c[sp - ((meta & 4u) ? P : 0)];
        nb[3] = (S)c[sp + ((meta & 8u) ? P : 0)];
        nb[4] = (S)((meta & 16u) ? sh.l2[fuse_slot2(z - 1)][sp] : c[sp]);
        nb[5] = (S)((meta & 32u) ? sh.l2[fuse_slot2(z + 1)][sp] : c[sp]);
        io = SlotIo<S>{e.cu[1], e.cg[1], e.su[par ^ 1u], e.sg[par ^ 1u], setc};
    }
    const float v = (float)special_update_io<float, S, kFuseSlotGroup, kFusePrefetch>(a, io, meta, f.rec.y, q, pc, pp,
                                                                                      nb, f.u0, e.sig[n + lvl - 1]);
    if (lvl == 2) {
        if (L == 3) sh.l2[fuse_slot2(z)][sp] = v;
        if (owned) (L == 3 ? oq : op)[idx] = v;
    } else if (owned) {
        op[idx] = v;
    }
}

// The pre-pass of fused pass kpass: step n of every special cell of every
// instance (as pass 1 of the three-step pass does), p^{n+1} into p1s and the
// slot state of step n+1 into the state after the pass's first step (or,
// for a pass of one step, into the pair's other state array).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_fuse_pre(const FuseEntry<S>* __restrict__ entries,
                                                     const uint32_t* __restrict__ pre_entry, uint32_t kpass) {
    const FuseEntry<S>& e = entries[pre_entry[blockIdx.x]];
    const uint32_t n = e.n0 + 3u * kpass;
    if (n >= e.steps) return;
    const uint32_t L = min(kpass + 1 < kFusePasses ? 3u : 2u, e.steps - n);
    const uint32_t par = ((e.n0 / kStepsPerBlock) + kpass) & 1u;
    const uint32_t q = (blockIdx.x - e.pre0) * kBlock + threadIdx.x;
    if (q >= e.n_special) return;
    const StepParams<float, S>& a = e.a;
    const float* p0 = e.f[2 * par];
    const float* q0 = e.f[2 * par + 1];
    const uint32_t i = e.sq_cell[q];
    const uint2 rec = a.sp_rec[q];
    const uint32_t meta = rec.x;
    const SlotIo<S> io{e.su[par], e.sg[par], L == 1 ? e.su[par ^ 1u] : e.cu[0], L == 1 ? e.sg[par ^ 1u] : e.cg[0]};
    S u0[kSlotGroup<S>];
#pragma unroll
    for (uint32_t j = 0; j < kSlotGroup<S>; ++j) u0[j] = j < a.max_slots ? io.u_in[j * a.sp_pad + q] : S(0);
    const S nb[6] = {(S)p0[i - ((meta & 1u) ? 1u : 0u)], (S)p0[i + ((meta & 2u) ? 1u : 0u)],
                     (S)p0[i - ((meta & 4u) ? a.pitch : 0u)], (S)p0[i + ((meta & 8u) ? a.pitch : 0u)],
                     (S)p0[i - ((meta & 16u) ? a.sz : 0u)], (S)p0[i + ((meta & 32u) ? a.sz : 0u)]};
    const float pc = p0[i];
    e.p1s[q] = make_float2((float)special_update_io(a, io, meta, rec.y, q, (S)pc, (S)q0[i], nb, u0, e.sig[n]), pc);
}

// Block lb of a fused pass over steps n .. n + L - 1 (L = 1, 2 or 3), from
// field pair par to pair par ^ 1. Waves 0 .. 11 form the tile part of each
// plane iteration and waves 12 .. 15 its special cells; the two run at the
// same time (an iteration's special cells read only what earlier iterations
// wrote), and the iteration ends with one barrier.
template <typename S>
__device__ __forceinline__ void fuse_block(const FuseEntry<S>& e, uint32_t lb, uint32_t n, uint32_t L, uint32_t par) {
    using C = Cells<float, 4>;
    __shared__ __align__(16) FuseShared<S> sh;
    const StepParams<float, S>& a = e.a;
    const uint32_t kb = lb % e.nblk, zc = lb / e.nblk;
    const int P = (int)a.pitch, sz = (int)a.sz, nz = (int)a.nz;
    const int span_lo = (int)(kb * e.span_step) - 2 * P;
    const int z0 = 1 + (int)(zc * e.zchunk);
    const int z1 = min(z0 + (int)e.zchunk, nz - 1);
    const uint32_t lane = threadIdx.x & (kTileX - 1);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTileX);
    const bool tile_wave = wave < (uint32_t)kFuseTileWaves;
    float* op = e.f[2 * (par ^ 1u)];
    float* oq = e.f[2 * (par ^ 1u) + 1];
    auto plane = [&](int z) -> uint32_t { return (uint32_t)min(max(z, 0), nz - 1) * (uint32_t)sz; };
    // The instance's slot constants into LDS when they fit (read before the
    // first barrier of the march).
    const S* setc = nullptr;
    if (e.n_setc * 4 <= kFuseSetBytes / sizeof(S)) {
        for (uint32_t k = threadIdx.x; k < e.n_setc * 4; k += kFuseThreads) sh.setc[k] = a.set_c[k];
        setc = sh.setc;
        __syncthreads();
    }
    const float* p0 = e.f[2 * par];

    // Owned receivers: one per thread (at most kFuseThreads per block); p^n
    // recorded now, p^{n+1} and p^{n+2} from LDS when their plane is done.
    const uint32_t r0 = e.rx_first[lb], nr = e.rx_first[lb + 1] - r0;
    uint32_t rr = ~0u, rs = 0;
    int rz = -100;
    if (threadIdx.x < nr) {
        const uint32_t* x = e.rx + 3 * (r0 + threadIdx.x);
        rr = x[0];
        rz = (int)x[1];
        rs = x[2];
        a.ir[(size_t)n * a.n_recv + rr] = p0[plane(rz) + (uint32_t)(span_lo + (int)rs)];
    }
    const int zlast = L == 3 ? z1 + 3 : L == 2 ? z1 + 2 : z1 + 1;

    if (tile_wave) {
        const int s = (int)(wave * kFuseSeg + lane * 4);  // span position of the lane's first cell
        const int o = span_lo + s;                        // its plane offset
        // Loads stay inside the plane (offsets clamped; cells outside it are
        // never owned and feed only cells that are not).
        const uint32_t ol = (uint32_t)min(max(o, 0), sz - 4);
        const uint32_t oym = (uint32_t)min(max(o - P, 0), sz - 4);
        const uint32_t oyp = (uint32_t)min(max(o + P, 0), sz - 4);
        const bool edge = lane == 0 || lane == kTileX - 1;
        const uint32_t oh = (uint32_t)(lane == 0 ? min(max(o - 1, 0), sz - 1) : min(max(o + 4, 0), sz - 1));
        // LDS positions of the row neighbours and of the cell next to the
        // lane, clamped into the span (only cells outside a level's valid
        // range read a clamped one).
        const int sym = max(s - P, 0), syp = min(s + P, (int)kFuseSpan - 4);
        const int sx = lane == 0 ? max(s - 1, 0) : min(s + 4, (int)kFuseSpan - 1);
        uint32_t own = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            own |= (uint32_t)(s + c >= 2 * P && s + c < (int)kFuseSpan - 2 * P && o + c >= 0 && o + c < sz) << c;
        const float* q0 = e.f[2 * par + 1];
        // Level-1 loads of plane zq, issued one plane ahead: p^n at zq + 1
        // and of the cells +-pitch at zq, p^{n-1} at zq, the cell next to
        // the lane (end lanes) at zq + 1, and the class word at zq.
        struct PlaneLoads {
            C zp, q, ym, yp;
            float hp;
            uint32_t cl;
        };
        auto issue = [&](int zq, PlaneLoads& X) {
            const uint32_t pz = plane(zq), pzp = plane(zq + 1);
            X.zp.load(p0 + pzp + ol);
            X.q.load(q0 + pz + ol);
            X.ym.load(p0 + pz + oym);
            X.yp.load(p0 + pz + oyp);
            X.hp = 0.f;
            if (edge) X.hp = p0[pzp + oh];
            X.cl = e.cls[(pz + ol) >> 2];
        };
        // p^n of the lane's cells at planes i - 2, i - 1, i and of the cell
        // next to it at i.
        C zmm, zm, zcur;
        zm.load(p0 + plane(z0 - 3) + ol);
        zcur.load(p0 + plane(z0 - 2) + ol);
#pragma unroll
        for (int c = 0; c < 4; ++c) zmm.v[c] = 0.f;
        float hc = 0.f;
        if (edge) hc = p0[plane(z0 - 2) + oh];
        // Interior (low nibble) and special (high nibble) bits of planes
        // i - 1 .. i - 4, a byte each.
        uint32_t clh = 0;
        PlaneLoads nx;
        issue(z0 - 2, nx);
        for (int i = z0 - 2; i <= zlast; ++i) {
            // Records of p^{n+1} (plane i - 1) and p^{n+2} (plane i - 3).
            if (rz == i - 1 && L >= 2) a.ir[(size_t)(n + 1) * a.n_recv + rr] = sh.l1[fuse_slot1(i - 1)][rs];
            if (rz == i - 3 && L == 3) a.ir[(size_t)(n + 2) * a.n_recv + rr] = sh.l2[fuse_slot2(i - 3)][rs];
            uint32_t cb0 = 0u;  // class bits of plane i
            const bool lev1 = i <= z1 + 1;
            C znext;
            float hnext = 0.f;
            // Level 1 at plane i. The z differences are formed from the
            // plane below (the values the single-step kernel carries).
            if (lev1) {
                const PlaneLoads cur = nx;
                if (i <= z1) issue(i + 1, nx);
                cb0 = cls_bits(cur.cl) | (cls_bits(cur.cl >> 1) << 4);
                znext = cur.zp;
                hnext = cur.hp;
                const float lo = from_lower_lane(zcur.v[3], hc);
                const float hi = from_upper_lane(zcur.v[0], hc);
                float dzm[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) dzm[c] = zcur.v[c] - zm.v[c];
                C p1;
                row_update<float, 4>(zcur.v, cur.q.v, lo, hi, cur.ym.v, cur.yp.v, zm.v, cur.zp.v, dzm, dzm, a.l2,
                                     a.c2, p1.v);
                if (L >= 2) lds_store_tiles(&sh.l1[fuse_slot1(i)][s], p1, cb0 >> 4);
                if (i >= z0 && i < z1 && L <= 2) {
                    const uint32_t bits = cb0 & 15u & own;
                    const uint32_t at = plane(i) + (uint32_t)o;
                    if (L == 1) {
                        store_masked<float, 4>(p1, op + at, bits);
                        store_masked<float, 4>(zcur, oq + at, bits);
                    } else {
                        store_masked<float, 4>(p1, oq + at, bits);
                    }
                }
            }
            // Level 2 at plane i - 2 (from p^{n+1} of planes i - 3 .. i - 1).
            const int z2 = i - 2;
            if (L >= 2 && z2 >= z0 - 1 && z2 <= z1) {
                const uint32_t cb = (clh >> 8) & 0xffu;
                const float* c1 = sh.l1[fuse_slot1(z2)];
                C cc, zb, zp, ym, yp;
                cc.load(c1 + s);
                zb.load(&sh.l1[fuse_slot1(z2 - 1)][s]);
                zp.load(&sh.l1[fuse_slot1(z2 + 1)][s]);
                ym.load(c1 + sym);
                yp.load(c1 + syp);
                float hv = 0.f;
                if (edge) hv = c1[sx];
                const float lo = from_lower_lane(cc.v[3], hv);
                const float hi = from_upper_lane(cc.v[0], hv);
                float dzm[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) dzm[c] = cc.v[c] - zb.v[c];
                C p2;
                row_update<float, 4>(cc.v, zmm.v, lo, hi, ym.v, yp.v, zb.v, zp.v, dzm, dzm, a.l2, a.c2, p2.v);
                if (L == 3) lds_store_tiles(&sh.l2[fuse_slot2(z2)][s], p2, cb >> 4);
                if (z2 >= z0 && z2 < z1)
                    store_masked<float, 4>(p2, (L == 3 ? oq : op) + plane(z2) + (uint32_t)o, cb & 15u & own);
            }
            // The march moves one plane (p^n at i - 2 was level 2's last use).
            if (lev1) {
                zmm = zm;
                zm = zcur;
                zcur = znext;
                hc = hnext;
            }
            // Level 3 at plane i - 4 (from p^{n+2} of planes i - 5 .. i - 3).
            const int z3 = i - 4;
            if (L == 3 && z3 >= z0 && z3 < z1) {
                const uint32_t cb = (clh >> 24) & 0xffu;
                const float* c2 = sh.l2[fuse_slot2(z3)];
                C cc, zb, zp, ym, yp, q;
                cc.load(c2 + s);
                zb.load(&sh.l2[fuse_slot2(z3 - 1)][s]);
                zp.load(&sh.l2[fuse_slot2(z3 + 1)][s]);
                ym.load(c2 + sym);
                yp.load(c2 + syp);
                q.load(&sh.l1[fuse_slot1(z3)][s]);
                float hv = 0.f;
                if (edge) hv = c2[sx];
                const float lo = from_lower_lane(cc.v[3], hv);
                const float hi = from_upper_lane(cc.v[0], hv);
                float dzm[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) dzm[c] = cc.v[c] - zb.v[c];
                C p3;
                row_update<float, 4>(cc.v, q.v, lo, hi, ym.v, yp.v, zb.v, zp.v, dzm, dzm, a.l2, a.c2, p3.v);
                store_masked<float, 4>(p3, op + plane(z3) + (uint32_t)o, cb & 15u & own);
            }
            clh = (clh << 8) | cb0;
            __syncthreads();
        }
    } else {
        // Entries of the special-cell list that iteration ii updates: level 1
        // at plane ii, level 2 at ii - 2, level 3 at ii - 4 (first entry and
        // count of each), in chunks of 64.
        struct SpRanges {
            uint32_t b[3], c[3];
            __device__ __forceinline__ uint32_t chunks(int k) const { return (c[k] + kFuseChunk - 1) / kFuseChunk; }
            __device__ __forceinline__ uint32_t count() const { return chunks(0) + chunks(1) + chunks(2); }
            __device__ __forceinline__ bool chunk(uint32_t j, uint32_t ln, int ii, uint32_t& lvl, int& zz,
                                                  uint32_t& q) const {
                const uint32_t c1n = chunks(0), c2n = chunks(1);
                uint32_t off;
                if (j < c1n) {
                    lvl = 1, off = j * kFuseChunk, zz = ii;
                } else if (j < c1n + c2n) {
                    lvl = 2, off = (j - c1n) * kFuseChunk, zz = ii - 2;
                } else {
                    lvl = 3, off = (j - c1n - c2n) * kFuseChunk, zz = ii - 4;
                }
                q = b[lvl - 1] + off + ln;
                return off + ln < c[lvl - 1];
            }
        };
        auto ranges = [&](int ii) -> SpRanges {
            SpRanges r{};
            const int y2 = ii - 2, y3 = ii - 4;
            if (ii >= 1 && ii <= nz - 2 && ii <= z1 + 1) {
                const uint32_t* t = e.sq_first + ((uint32_t)ii * e.nblk + kb) * kFuseFirst;
                r.b[0] = t[0];
                r.c[0] = t[5] - t[0];
            }
            if (L >= 2 && y2 >= max(z0 - 1, 1) && y2 <= min(z1, nz - 2)) {
                const uint32_t* t = e.sq_first + ((uint32_t)y2 * e.nblk + kb) * kFuseFirst;
                r.b[1] = t[1];
                r.c[1] = t[4] - t[1];
            }
            if (L == 3 && y3 >= z0 && y3 < z1) {
                const uint32_t* t = e.sq_first + ((uint32_t)y3 * e.nblk + kb) * kFuseFirst;
                r.b[2] = t[2];
                r.c[2] = t[3] - t[2];
            }
            return r;
        };
        const uint32_t sw = wave - (uint32_t)kFuseTileWaves;
        // The chunk being solved and the next one, whose loads are in flight
        // (the first chunk of an iteration is loaded during the iteration
        // before).
        SpRanges sr = ranges(z0 - 2);
        uint32_t c_lvl = 0, c_q = 0;
        int c_z = 0;
        FuseSpecial<S> cur;
        bool c_ok = sw < sr.count() && sr.chunk(sw, lane, z0 - 2, c_lvl, c_z, c_q);
        if (c_ok) fuse_special_load<S>(e, c_lvl, c_q, cur);
        for (int i = z0 - 2; i <= zlast; ++i) {
            // Records, as the tile waves do (receivers of threads 768 ..).
            if (rz == i - 1 && L >= 2) a.ir[(size_t)(n + 1) * a.n_recv + rr] = sh.l1[fuse_slot1(i - 1)][rs];
            if (rz == i - 3 && L == 3) a.ir[(size_t)(n + 2) * a.n_recv + rr] = sh.l2[fuse_slot2(i - 3)][rs];
            const SpRanges nsr = i < zlast ? ranges(i + 1) : SpRanges{};
            const uint32_t cn = sr.count();
            for (uint32_t j = sw; j < cn; j += kFuseSpecWaves) {
                uint32_t n_lvl = 0, n_q = 0;
                int n_z = 0;
                FuseSpecial<S> nxt;
                bool n_ok;
                if (j + kFuseSpecWaves < cn) {
                    n_ok = sr.chunk(j + kFuseSpecWaves, lane, i, n_lvl, n_z, n_q);
                } else {
                    // The next iteration's first chunk: its inputs were all
                    // written before this iteration began.
                    n_ok = i < zlast && sw < nsr.count() && nsr.chunk(sw, lane, i + 1, n_lvl, n_z, n_q);
                }
                if (n_ok) fuse_special_load<S>(e, n_lvl, n_q, nxt);
                if (c_ok) fuse_special_run<S>(e, sh, setc, c_lvl, c_z, c_q, cur, span_lo, z0, z1, n, L, par);
                cur = nxt;
                c_ok = n_ok;
                c_lvl = n_lvl;
                c_z = n_z;
                c_q = n_q;
            }
            if (sw >= cn && i < zlast) {
                // No chunk of this wave here: load the next iteration's first.
                c_ok = sw < nsr.count() && nsr.chunk(sw, lane, i + 1, c_lvl, c_z, c_q);
                if (c_ok) fuse_special_load<S>(e, c_lvl, c_q, cur);
            }
            sr = nsr;
            __syncthreads();
        }
    }
}

// One fused pass of every instance in the table: pass kpass (0 .. 42) of
// the current 128-step block, steps n0 + 3 kpass ..; instances past their
// last step return at once.
template <typename S>
__global__ __launch_bounds__(kFuseThreads) void k_fuse(const FuseEntry<S>* __restrict__ entries,
                                                       const uint32_t* __restrict__ block_entry, uint32_t kpass) {
    const FuseEntry<S>& e = entries[block_entry[blockIdx.x]];
    const uint32_t n = e.n0 + 3u * kpass;
    if (n >= e.steps) return;
    const uint32_t L = min(kpass + 1 < kFusePasses ? 3u : 2u, e.steps - n);
    const uint32_t par = ((e.n0 / kStepsPerBlock) + kpass) & 1u;
    fuse_block<S>(e, blockIdx.x - e.block0, n, L, par);
}

template <typename S>
__global__ void k_fuse_tick(FuseEntry<S>* __restrict__ entries, uint32_t n, uint32_t add) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) entries[i].n0 += add;
}

// Class bytes: interior (from the z-packed bitmap) for planes 1 .. nz - 2;
// the special bits follow from the special list (k_cls_special).
__global__ __launch_bounds__(kBlock) void k_cls(const uint32_t* __restrict__ zbits, uint32_t* __restrict__ cls,
                                                uint32_t sz, uint32_t nz, uint64_t nwords) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    const uint64_t i = w * 4;
    const uint32_t z = (uint32_t)(i / sz), off = (uint32_t)(i % sz);
    uint32_t v = 0;
    if (z >= 1 && z + 1 < nz) {
        const uint32_t k = z - 1;
        const uint32_t* zb = zbits + (size_t)(k >> 5) * sz + off;
#pragma unroll
        for (int c = 0; c < 4; ++c) v |= ((zb[c] >> (k & 31u)) & 1u) << (8 * c);
    }
    cls[w] = v;
}

// sq_first of every (plane, span, bound): binary search of the sorted list.
__global__ __launch_bounds__(kBlock) void k_sq_first(const uint32_t* __restrict__ cells, uint32_t n,
                                                     uint32_t* __restrict__ out, uint32_t sz, uint32_t nz,
                                                     uint32_t pitch, uint32_t span_step, uint32_t nblk) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nz * nblk * kFuseFirst) return;
    const uint32_t j = t % kFuseFirst, zk = t / kFuseFirst, k = zk % nblk, z = zk / nblk;
    const long long P = pitch, S = kFuseSpan;
    const long long at[kFuseFirst] = {0, P, 2 * P, S - 2 * P, S - P, S};
    const long long pos = min(max((long long)k * span_step - 2 * P + at[j], 0ll), (long long)sz);
    const uint32_t target = z * sz + (uint32_t)pos;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (cells[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    out[t] = lo;
}

// Sets *bad when the list is not strictly increasing.
__global__ __launch_bounds__(kBlock) void k_sq_check(const uint32_t* __restrict__ cells, uint32_t n,
                                                     uint32_t* __restrict__ bad) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t + 1 < n && cells[t] >= cells[t + 1]) *bad = 1u;
}

__global__ __launch_bounds__(kBlock) void k_cls_special(const uint32_t* __restrict__ sp_cell, uint32_t n,
                                                        uint32_t* __restrict__ cls) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t i = sp_cell[r];
    atomicOr(cls + (i >> 2), kClsSpecial << (8 * (i & 3u)));
}

}  // namespace wv
