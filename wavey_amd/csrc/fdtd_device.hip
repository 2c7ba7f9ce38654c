// MI355X (gfx950) FDTD wave-solver run: the plan's device layout, its set-up
// kernels, the plain run loop and its C ABI.
//
// Layout (the step kernel itself is in step_kernels.h):
//
//  * Fields p, prev live in a pitched grid: row pitch = nx rounded up to 32
//    cells, so rows start on 128-byte lines and a wave's 16-byte-per-lane
//    accesses cover aligned row segments.
//  * An interior bitmap, packed along z (one 32-bit word per column per 32
//    planes), marks plain interior air: six air neighbours, no boundary
//    slots, no source tap.
//  * Every other air cell, and every tap, is a special cell. Their records
//    are compacted (wave ballot + prefix sums) in the order the step
//    kernel's waves consume them: a wave's records are those of its row
//    segment and z-chunk, sorted by plane, and a record gives the cell's
//    position in the chunk.
//  * Boundary slot constants are deduplicated into a small table of slot
//    sets; slot state is stored slot-major, state[j][rank], so the reads of a
//    wave's special cells are contiguous.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "plan.h"

using namespace wv;

namespace {

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

__device__ __forceinline__ bool is_interior(const uint8_t* __restrict__ flags, const uint32_t* __restrict__ force,
                                            uint32_t x, uint32_t y, uint32_t z, const Geo& g) {
    return classify(flags, force, x, y, z, g).interior;
}

// A deep cell is interior and so are its six neighbours: its second step of
// a two-step pass needs no special cell's value.
__device__ __forceinline__ bool is_deep(const uint8_t* __restrict__ flags, const uint32_t* __restrict__ force,
                                        uint32_t x, uint32_t y, uint32_t z, const Geo& g) {
    return is_interior(flags, force, x, y, z, g) && is_interior(flags, force, x - 1, y, z, g) &&
           is_interior(flags, force, x + 1, y, z, g) && is_interior(flags, force, x, y - 1, z, g) &&
           is_interior(flags, force, x, y + 1, z, g) && is_interior(flags, force, x, y, z - 1, g) &&
           is_interior(flags, force, x, y, z + 1, g);
}

// The z-packed interior and deep bitmaps: word (kb, y, x) holds the bits of
// planes 1 + 32 kb .. 32 + 32 kb of column (x, y), bit k for plane
// 1 + 32 kb + k. (Coordinates 0 - 1 wrap to large values, which classify
// treats as outside the grid.)
__global__ __launch_bounds__(kBlock) void k_zbits(const uint8_t* __restrict__ flags, const uint32_t* __restrict__ force,
                                                  uint32_t* __restrict__ zbits, uint32_t* __restrict__ dbits, Geo g,
                                                  uint32_t nzb) {
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (uint64_t)nzb * g.ny * g.pitch) return;
    const uint32_t x = (uint32_t)(t % g.pitch);
    const uint64_t r = t / g.pitch;
    const uint32_t y = (uint32_t)(r % g.ny), kb = (uint32_t)(r / g.ny);
    uint32_t wi = 0, wd = 0;
    for (uint32_t k = 0; k < 32; ++k) {
        const uint32_t z = 1 + kb * 32 + k;
        if (z + 1 >= g.nz) break;
        wi |= (uint32_t)is_interior(flags, force, x, y, z, g) << k;
        wd |= (uint32_t)is_deep(flags, force, x, y, z, g) << k;
    }
    zbits[t] = wi;
    dbits[t] = wd;
}

// Deep2 bits: deep cells whose six neighbours are deep (the third step of a
// three-step pass needs nothing the block does not compute itself), from the
// deep bitmap; the z neighbours are the word's bits shifted by one, with the
// end bits from the words of the next 32-plane blocks.
__global__ __launch_bounds__(kBlock) void k_d2bits(const uint32_t* __restrict__ dbits, uint32_t* __restrict__ d2bits,
                                                   uint32_t pitch, uint32_t ny, uint32_t nzb) {
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t plane = (uint64_t)ny * pitch;
    if (t >= nzb * plane) return;
    const uint32_t x = (uint32_t)(t % pitch);
    const uint64_t r = t / pitch;
    const uint32_t y = (uint32_t)(r % ny), kb = (uint32_t)(r / ny);
    const uint32_t w = dbits[t];
    uint32_t m = w;
    m &= x > 0 ? dbits[t - 1] : 0u;
    m &= x + 1 < pitch ? dbits[t + 1] : 0u;
    m &= y > 0 ? dbits[t - pitch] : 0u;
    m &= y + 1 < ny ? dbits[t + pitch] : 0u;
    m &= (w << 1) | (kb > 0 ? dbits[t - plane] >> 31 : 0u);        // plane below
    m &= (w >> 1) | (kb + 1 < nzb ? dbits[t + plane] << 31 : 0u);  // plane above
    d2bits[t] = m;
}

// Receiver cells leave the deep2 set (pitched indices), so a three-step pass
// stores their p^{n+1}.
__global__ void k_unset_d2(uint32_t* __restrict__ d2bits, const uint32_t* __restrict__ cells, uint32_t n,
                           uint32_t pitch, uint32_t ny, uint32_t nz) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t plane = (uint64_t)ny * pitch;
    const uint32_t z = (uint32_t)(cells[i] / plane);
    if (z < 1 || z + 1 >= nz) return;
    const uint32_t rem = (uint32_t)(cells[i] % plane);
    const uint64_t word = ((uint64_t)((z - 1) >> 5) * ny + rem / pitch) * pitch + rem % pitch;
    atomicAnd(&d2bits[word], ~(1u << ((z - 1) & 31u)));
}

// Ring cells (interior, not deep) in pitched order, for the second half of a
// two-step pass: pass 1 counts them per 256 cells of planes 1 .. nz-2, pass 2
// writes their pitched indices.
struct RingAt {
    const uint8_t* __restrict__ flags;
    const uint32_t* __restrict__ force;
    Geo g;
    __device__ __forceinline__ bool operator()(uint64_t q) const {
        const uint32_t x = (uint32_t)(q % g.pitch);
        const uint64_t r = q / g.pitch;
        const uint32_t y = (uint32_t)(r % g.ny), z = (uint32_t)(r / g.ny);
        return is_interior(flags, force, x, y, z, g) && !is_deep(flags, force, x, y, z, g);
    }
};

// Interior cells that are not deep2 (the cells of a three-step pass's third
// launch besides the special ones), from the bitmaps.
struct Ring2At {
    const uint32_t* __restrict__ zbits;
    const uint32_t* __restrict__ d2bits;
    uint32_t pitch, ny;
    __device__ __forceinline__ bool operator()(uint64_t q) const {
        const uint64_t plane = (uint64_t)ny * pitch;
        const uint32_t z = (uint32_t)(q / plane);  // 1 .. nz-2
        const uint64_t rem = q % plane;
        const uint64_t word = (uint64_t)((z - 1) >> 5) * plane + rem;
        const uint32_t bit = 1u << ((z - 1) & 31u);
        return (zbits[word] & bit) && !(d2bits[word] & bit);
    }
};

template <class At>
__global__ __launch_bounds__(kBlock) void k_ring_count(At at, uint32_t* __restrict__ counts, Geo g) {
    __shared__ uint32_t wc[kBlock / 64];
    const uint64_t q = (uint64_t)g.pitch * g.ny + (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool rg = q < (uint64_t)g.pitch * g.ny * (g.nz - 1) && at(q);
    const uint64_t b = __ballot(rg);
    if ((threadIdx.x & 63u) == 0) wc[threadIdx.x >> 6] = (uint32_t)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t w = 0; w < kBlock / 64; ++w) s += wc[w];
        counts[blockIdx.x] = s;
    }
}

template <class At>
__global__ __launch_bounds__(kBlock) void k_ring_write(At at, const uint32_t* __restrict__ offs, Geo g,
                                                       uint32_t* __restrict__ ring_cell) {
    __shared__ uint32_t wc[kBlock / 64];
    const uint64_t q = (uint64_t)g.pitch * g.ny + (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool rg = q < (uint64_t)g.pitch * g.ny * (g.nz - 1) && at(q);
    const uint64_t b = __ballot(rg);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) wc[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    if (!rg) return;
    uint32_t rank = offs[blockIdx.x];
    for (uint32_t w = 0; w < wave; ++w) rank += wc[w];
    rank += (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
    ring_cell[rank] = (uint32_t)q;
}

// The wave order of special cells: wave w = zc * n_tiles + tp (tile tp of
// the plane, z-chunk zc) owns positions [w * seg, (w + 1) * seg); position
// (w, dz, xi) is the cell at plane offset tp * BX + xi in plane
// 1 + zc * zchunk + dz. Every wave segment is a whole number of compaction
// blocks.
struct WaveOrder {
    uint32_t bx, zchunk, seg, n_tiles;
};

__device__ __forceinline__ bool wave_cell(uint64_t qw, const WaveOrder& o, const Geo& g, uint32_t* x, uint32_t* y,
                                          uint32_t* z) {
    const uint64_t w = qw / o.seg;
    const uint32_t rr = (uint32_t)(qw % o.seg);
    const uint32_t dz = rr / o.bx, xi = rr % o.bx;
    if (dz >= o.zchunk) return false;
    const uint32_t tp = (uint32_t)(w % o.n_tiles), zc = (uint32_t)(w / o.n_tiles);
    const uint64_t f = (uint64_t)tp * o.bx + xi;
    if (f >= (uint64_t)g.pitch * g.ny) return false;
    *x = (uint32_t)(f % g.pitch);
    *y = (uint32_t)(f / g.pitch);
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

// Exclusive scan of many counts (one per compaction block) in three
// launches: the sum of each chunk of kScanChunk counts (coalesced reads), a
// one-group scan of those sums, then each chunk's scan from its base, eight
// consecutive counts per thread. The offsets are those of a serial scan.
constexpr uint32_t kScanThreads = 1024, kScanItems = 8, kScanChunk = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t group_exclusive(uint32_t v, uint32_t* part) {
    part[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t w = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += w;
        __syncthreads();
    }
    return part[threadIdx.x] - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_sums(const uint32_t* __restrict__ counts, uint32_t n,
                                                            uint32_t* __restrict__ sums) {
    __shared__ uint32_t part[kScanThreads];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    uint32_t s = 0;
    for (uint32_t k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kScanThreads + threadIdx.x;
        if (i < n) s += counts[i];
    }
    const uint32_t ex = group_exclusive(s, part);
    if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = ex + s;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_chunks(const uint32_t* __restrict__ counts, uint32_t n,
                                                              const uint32_t* __restrict__ bases,
                                                              uint32_t* __restrict__ offs) {
    __shared__ uint32_t part[kScanThreads];
    const uint64_t first = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanItems;
    uint32_t c[kScanItems];
    uint32_t s = 0;
    for (uint32_t k = 0; k < kScanItems; ++k) {
        c[k] = first + k < n ? counts[first + k] : 0u;
        s += c[k];
    }
    uint32_t o = bases[blockIdx.x] + group_exclusive(s, part);
    for (uint32_t k = 0; k < kScanItems; ++k)
        if (first + k < n) {
            offs[first + k] = o;
            o += c[k];
        }
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
                                                     S* __restrict__ sp_tapw, uint32_t* __restrict__ sp_cell) {
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
    sp_cell[rank] = (uint32_t)(((uint64_t)z * g.ny + y) * g.pitch + x);
    sp_tapw[rank] = w;
}

// Receiver spectra from the recorded series, FP64: sum_n p^n e^{-j 2 pi fract(f dt n)}
// in n order (fdtd.rs:679-682).
template <typename T>
__global__ void k_dft(const T* __restrict__ ir, uint32_t steps, uint32_t n_recv, const double* __restrict__ freqs,
                      uint32_t nf, double dt, const double* __restrict__ d0h, double* __restrict__ out) {
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
        double v = (double)ir[(size_t)n * n_recv + r];
        if (d0h) v += d0h[offset_index(n)];  // FP32: the pressure is the stored value plus the offset
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

// Re-centring of FP32 fields: the mean of p^n over the air cells in a fixed
// order (block b sums rows b, b + grid, ... thread by thread in x, then a
// tree in LDS), the final sum and the decision in one block, then every
// cell of p^n and p^{n-1} less the shift and the run's offset plus it.
constexpr uint32_t kMeanBlocks = 2048;
constexpr uint32_t kOffsets = 32;  // offset_index of any step below 2^31 is under 26

__global__ __launch_bounds__(256) void k_mean_part(const uint8_t* __restrict__ flags, const float* __restrict__ p,
                                                   uint32_t nx, uint32_t pitch, uint32_t rows,
                                                   double* __restrict__ part) {
    __shared__ double s[256];
    double acc = 0.0;
    for (uint32_t r = blockIdx.x; r < rows; r += gridDim.x)
        for (uint32_t x = threadIdx.x; x < nx; x += 256)
            if (flags[(uint64_t)r * nx + x] & 64u) acc += (double)p[(uint64_t)r * pitch + x];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_mean_final(const double* __restrict__ part, uint32_t n, double air,
                                                    float* __restrict__ shift, double* __restrict__ d0,
                                                    double* __restrict__ d0h, double* __restrict__ prev_mean,
                                                    uint32_t j) {
    __shared__ double s[256];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // The stored mean and the pressure's (plus the offset). The first
        // check only records the pressure's mean; later ones follow a mean
        // that persists (at least half of the last check's) and, where it
        // decays, move the offset back into the fields.
        const double stored = s[0] / air, mean = stored + *d0;
        float m = 0.0f;
        if (j > 1) m = fabs(mean) >= 0.5 * fabs(*prev_mean) ? (float)stored : (float)(-*d0);
        *prev_mean = mean;
        *shift = m;
        *d0 += (double)m;
        d0h[j] = *d0;
    }
}

__global__ __launch_bounds__(256) void k_shift(float4* __restrict__ a, float4* __restrict__ b, uint64_t n4,
                                               const float* __restrict__ shift) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const float m = *shift;
    if (i >= n4 || m == 0.0f) return;
    float4 u = a[i], v = b[i];
    u.x -= m; u.y -= m; u.z -= m; u.w -= m;
    v.x -= m; v.y -= m; v.z -= m; v.w -= m;
    a[i] = u;
    b[i] = v;
}

}  // namespace

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

}  // namespace

void wv::plan_setup(wv_plan* pl, const wv_grid* g) {
    WV_HIP(hipGetDevice(&pl->device));
    if (!pl->borrowed) {
        WV_HIP(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking));
        WV_HIP(hipEventCreate(&pl->ev[0]));
        WV_HIP(hipEventCreate(&pl->ev[1]));
        WV_HIP(hipEventCreate(&pl->t0));
        WV_HIP(hipEventCreate(&pl->t1));
    }
    pl->nx = (uint32_t)g->n[0];
    pl->ny = (uint32_t)g->n[1];
    pl->nz = (uint32_t)g->n[2];
    pl->cells = g->n[0] * g->n[1] * g->n[2];
    pl->pitch = (pl->nx + 31u) & ~31u;  // rows start on 128-byte lines (FP32; 256 bytes FP64)
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
    // Slot constants as one record per slot, {inv_d, cv, k, w}: one 16-byte
    // (FP32) or two (FP64 state) loads per slot.
    std::vector<float> f_c;
    std::vector<double> d_c;
    if (pl->prec == WV_FP32) {
        f_c.resize(4 * inv_d.size());
        for (size_t j = 0; j < inv_d.size(); ++j) {
            f_c[4 * j] = (float)inv_d[j];
            f_c[4 * j + 1] = (float)cv[j];
            f_c[4 * j + 2] = (float)kk[j];
            f_c[4 * j + 3] = (float)w[j];
        }
        upload(pl, pl->set_c, f_c);
    } else {
        d_c.resize(4 * inv_d.size());
        for (size_t j = 0; j < inv_d.size(); ++j) {
            d_c[4 * j] = inv_d[j];
            d_c[4 * j + 1] = cv[j];
            d_c[4 * j + 2] = kk[j];
            d_c[4 * j + 3] = w[j];
        }
        upload(pl, pl->set_c, d_c);
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
    pl->dbits.alloc((size_t)pl->nzb * pl->ny * pl->pitch * 4);
    pl->fa.alloc(pl->npitched * pl->tsize());
    pl->fb.alloc(pl->npitched * pl->tsize());
    WV_HIP(hipStreamSynchronize(pl->stream));
}

namespace {

template <class T>
void fill_async(wv_plan* pl, DBuf& d, size_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(div_up(n, 256)), dim3(256), 0, pl->stream, d.as<T>(), n, T(0));
    WV_HIP(hipGetLastError());
}

}  // namespace

void wv::recentre(wv_plan* pl, size_t n, hipStream_t stream) {
    if (!recentres_at(pl, n) || pl->recentred_at == n || pl->air == 0) return;
    pl->recentred_at = n;
    DBuf* fields[5] = {&pl->fa, &pl->fb, &pl->fc, &pl->fd, &pl->fe};
    const uint32_t rows = pl->ny * pl->nz;
    const uint32_t grid = std::min(rows, kMeanBlocks);
    hipLaunchKernelGGL(k_mean_part, dim3(grid), dim3(256), 0, stream, pl->flags.as<uint8_t>(),
                       fields[pl->buf_p]->as<float>(), pl->nx, pl->pitch, rows, pl->mean_part.as<double>());
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(256), 0, stream, pl->mean_part.as<double>(), grid,
                       (double)pl->air, pl->shift.as<float>(), pl->d0.as<double>(), pl->d0h.as<double>(),
                       pl->prev_mean.as<double>(), offset_index((uint32_t)n));
    const uint64_t n4 = pl->npitched / 4;  // the pitch is a multiple of 32 cells
    hipLaunchKernelGGL(k_shift, dim3(div_up(n4, 256)), dim3(256), 0, stream, fields[pl->buf_p]->as<float4>(),
                       fields[pl->buf_q]->as<float4>(), n4, pl->shift.as<float>());
    WV_HIP(hipGetLastError());
}

namespace {

// offs[0 .. n) = exclusive scan of counts[0 .. n), offs[n] = the total.
void scan_counts(wv_plan* pl, const uint32_t* counts, uint32_t* offs, uint32_t n) {
    const uint32_t chunks = std::max<uint32_t>(1, div_up(n, kScanChunk));
    pl->scan_tmp.alloc((size_t)2 * chunks * 4);
    uint32_t* sums = pl->scan_tmp.as<uint32_t>();
    uint32_t* bases = sums + chunks;
    hipLaunchKernelGGL(k_scan_sums, dim3(chunks), dim3(kScanThreads), 0, pl->stream, counts, n, sums);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, pl->stream, sums, bases, chunks, offs + n);
    hipLaunchKernelGGL(k_scan_chunks, dim3(chunks), dim3(kScanThreads), 0, pl->stream, counts, n, bases, offs);
    WV_HIP(hipGetLastError());
}

// Compacts the cells of planes 1 .. nz-2 that `at` selects into `cells`, in
// pitched order (counts and offsets per 256 cells in `counts`, `offs`);
// returns their number.
template <class At>
uint32_t compact_cells(wv_plan* pl, const At& at, const Geo& geo, DBuf& counts, DBuf& offs, DBuf& cells) {
    const uint64_t span = (uint64_t)pl->sz * (pl->nz - 2);
    const uint64_t rblocks = (span + kBlock - 1) / kBlock;
    if (rblocks >= (1ull << 31)) throw DevErr{"grid too large for the ring layout", WV_ERR_LIMIT};
    counts.alloc(rblocks * 4);
    offs.alloc((rblocks + 1) * 4);
    hipLaunchKernelGGL((k_ring_count<At>), dim3((uint32_t)rblocks), dim3(kBlock), 0, pl->stream, at,
                       counts.as<uint32_t>(), geo);
    scan_counts(pl, counts.as<uint32_t>(), offs.as<uint32_t>(), (uint32_t)rblocks);
    uint32_t n = 0;
    WV_HIP(hipMemcpyAsync(&n, offs.as<uint32_t>() + rblocks, 4, hipMemcpyDeviceToHost, pl->stream));
    WV_HIP(hipStreamSynchronize(pl->stream));
    cells.alloc((size_t)n * 4);
    hipLaunchKernelGGL((k_ring_write<At>), dim3((uint32_t)rblocks), dim3(kBlock), 0, pl->stream, at,
                       offs.as<uint32_t>(), geo, cells.as<uint32_t>());
    WV_HIP(hipGetLastError());
    return n;
}

// The state of step 0 of the plan's run (its layout already built): zeroed
// fields, records, slot state and re-centring offsets, p in the first field
// buffer and prev in the second. Waits for the fills.
template <class S>
void run_zero(wv_plan* pl, size_t steps, size_t n_recv) {
    const size_t nfields = pl->run_steps_per_pass == 3 ? 5 : pl->run_steps_per_pass == 2 ? 4 : 2;
    DBuf* fields[5] = {&pl->fa, &pl->fb, &pl->fc, &pl->fd, &pl->fe};
    for (size_t f = 0; f < nfields; ++f) {
        if (pl->prec == WV_FP64) fill_async<double>(pl, *fields[f], pl->npitched);
        else fill_async<float>(pl, *fields[f], pl->npitched);
    }
    if (pl->prec == WV_FP64) fill_async<double>(pl, pl->ir, steps * n_recv);
    else fill_async<float>(pl, pl->ir, steps * n_recv);
    const size_t nstate = (size_t)pl->max_slots * pl->sp_pad;
    fill_async<S>(pl, pl->st_u, nstate);
    fill_async<S>(pl, pl->st_g, nstate);
    fill_async<double>(pl, pl->d0, 1);
    fill_async<double>(pl, pl->prev_mean, 1);
    fill_async<double>(pl, pl->d0h, kOffsets);
    pl->recentred_at = 0;
    pl->buf_p = 0;
    pl->buf_q = 1;
    pl->done = 0;
    WV_HIP(hipStreamSynchronize(pl->stream));
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
                       pl->dbits.as<uint32_t>(), geo, pl->nzb);
    // Special cells in wave order.
    pl->run_zchunk = pl->zchunk ? pl->zchunk : default_step_zchunk(pl->cells, pl->prec == WV_FP64);
    pl->run_variant = pl->variant >= 0 ? pl->variant : default_step_variant(pl->cells, pl->prec == WV_FP64);
    const Layout L = layout_of(pl);
    const WaveOrder wo{L.bx, L.zchunk, L.seg, L.n_tiles};
    const uint64_t nblocks = L.positions / kBlock;
    if (nblocks >= (1ull << 31)) throw DevErr{"grid too large for the special-cell layout", WV_ERR_LIMIT};
    // offs[b] = special cells before compaction block b; offs[nblocks] = total.
    pl->counts.alloc(nblocks * 4);
    pl->offs.alloc((nblocks + 1) * 4);
    pl->seg_stride = L.seg / kBlock;
    hipLaunchKernelGGL(k_sp_count, dim3((uint32_t)nblocks), dim3(kBlock), 0, pl->stream, pl->flags.as<uint8_t>(),
                       pl->force.as<uint32_t>(), pl->counts.as<uint32_t>(), geo, wo);
    scan_counts(pl, pl->counts.as<uint32_t>(), pl->offs.as<uint32_t>(), (uint32_t)nblocks);
    uint32_t n_special = 0;
    WV_HIP(hipMemcpyAsync(&n_special, pl->offs.as<uint32_t>() + nblocks, 4, hipMemcpyDeviceToHost, pl->stream));
    WV_HIP(hipStreamSynchronize(pl->stream));
    pl->n_special = n_special;
    pl->sp_pad = (n_special + 63u) & ~63u;
    pl->sp_rec.alloc((size_t)n_special * 8);
    pl->sp_tapw.alloc((size_t)n_special * sizeof(S));
    pl->sp_cell.alloc((size_t)n_special * 4);
    hipLaunchKernelGGL(k_sp_write<S>, dim3((uint32_t)nblocks), dim3(kBlock), 0, pl->stream, pl->flags.as<uint8_t>(),
                       pl->force.as<uint32_t>(), pl->offs.as<uint32_t>(), geo, wo, pl->b_cell_sorted.as<uint32_t>(),
                       pl->b_set_sorted.as<uint32_t>(), (uint32_t)pl->nb, pl->set_n.as<uint32_t>(),
                       pl->set_off.as<uint32_t>(), pl->tap_cell.as<uint32_t>(), pl->tap_w.as<S>(), (uint32_t)tc.size(),
                       pl->sp_rec.as<uint2>(), pl->sp_tapw.as<S>(), pl->sp_cell.as<uint32_t>());
    WV_HIP(hipGetLastError());
    // Multi-step passes: the ring list and the two more field buffers (a
    // third for three steps).
    pl->run_steps_per_pass = pl->steps_per_pass ? pl->steps_per_pass : default_steps_per_pass(pl->cells, pl->prec == WV_FP64);
    pl->n_ring = pl->n_ring2 = 0;
    if (pl->run_steps_per_pass >= 2) {
        pl->n_ring = compact_cells(pl, RingAt{pl->flags.as<uint8_t>(), pl->force.as<uint32_t>(), geo}, geo,
                                   pl->ring_counts, pl->ring_offs, pl->ring_cell);
        pl->fc.alloc(pl->npitched * pl->tsize());
        pl->fd.alloc(pl->npitched * pl->tsize());
    }
    if (pl->run_steps_per_pass == 3) pl->fe.alloc(pl->npitched * pl->tsize());
    pl->buf_p = 0;
    pl->buf_q = 1;
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
    // Three-step passes: the deep2 bitmap without the receiver cells, and the
    // interior cells that are not deep2.
    if (pl->run_steps_per_pass == 3) {
        const uint64_t nwords = (uint64_t)pl->nzb * pl->ny * pl->pitch;
        pl->d2bits.alloc(nwords * 4);
        hipLaunchKernelGGL(k_d2bits, dim3((uint32_t)div_up(nwords, kBlock)), dim3(kBlock), 0, pl->stream,
                           pl->dbits.as<uint32_t>(), pl->d2bits.as<uint32_t>(), pl->pitch, pl->ny, pl->nzb);
        if (n_recv)
            hipLaunchKernelGGL(k_unset_d2, dim3(div_up(n_recv, 256)), dim3(256), 0, pl->stream,
                               pl->d2bits.as<uint32_t>(), pl->recv.as<uint32_t>(), (uint32_t)n_recv, pl->pitch,
                               pl->ny, pl->nz);
        WV_HIP(hipGetLastError());
        pl->n_ring2 = compact_cells(pl, Ring2At{pl->zbits.as<uint32_t>(), pl->d2bits.as<uint32_t>(), pl->pitch, pl->ny},
                                    geo, pl->ring2_counts, pl->ring2_offs, pl->ring2_cell);
    }
    pl->ir.alloc(std::max<size_t>(1, steps * n_recv) * pl->tsize());
    pl->d0.alloc(8);
    pl->prev_mean.alloc(8);
    pl->d0h.alloc(kOffsets * 8);
    pl->mean_part.alloc(kMeanBlocks * 8);
    pl->shift.alloc(4);
    run_zero<S>(pl, steps, n_recv);
    pl->signal.assign(signal, signal + steps);
    pl->steps = steps;
    pl->done = 0;
    pl->n_recv = (uint32_t)n_recv;
    pl->active = true;
}

// Parameters of a two-step pass: pass 1 (tiles of deep cells, special
// cells' step n, receivers' p^n) and pass 2 (ring and special cells' step
// n+1, receivers' p^{n+1}).
template <class T, class S>
PairParams<T, S> make_pair_params(wv_plan* pl, const StepParams<T, S>& base) {
    PairParams<T, S> pp{};
    pp.a = base;
    const uint32_t zch = pass_zchunk(pl, 2);
    pp.a.zchunk = zch;
    pp.dbits = pl->dbits.as<uint32_t>();
    pp.sp_cell = pl->sp_cell.as<uint32_t>();
    pp.ring_cell = pl->ring_cell.as<uint32_t>();
    pp.n_special = pl->n_special;
    pp.n_ring = pl->n_ring;
    // Tile k stores p^{n+2} on columns 2 + k * step .. (k + 1) * step + 1 and
    // p^{n+1} one column further on each side: columns 2 .. nx - 3 (a deep
    // cell is two from the grid's ends) and 1 .. nx - 2 (every interior cell).
    pp.tiles_x = std::max<uint32_t>(1, div_up(pl->nx > 4 ? pl->nx - 4 : 0, kTileStep<T>));
    pp.row_blocks = div_up(pl->ny, kPairRows);
    pp.tile_blocks = pp.tiles_x * pp.row_blocks * div_up(pl->nz - 2, zch);
    pp.tile_launch = div_up(pp.tile_blocks / pp.tiles_x, 8) * 8 * pp.tiles_x;
    pp.special_blocks = div_up(pl->n_special, kPairThreads);
    return pp;
}

// Parameters of a three-step pass's first launch (tiles of the bulk, the
// special cells' step n, the receivers' p^n).
template <class T, class S>
TriParams<T, S> make_tri_params(wv_plan* pl, const StepParams<T, S>& base) {
    TriParams<T, S> tp{};
    tp.a = base;
    const uint32_t zch = pass_zchunk(pl, 3);
    tp.a.zchunk = zch;
    tp.dbits = pl->dbits.as<uint32_t>();
    tp.d2bits = pl->d2bits.as<uint32_t>();
    tp.sp_cell = pl->sp_cell.as<uint32_t>();
    tp.n_special = pl->n_special;
    // Tile k stores p^{n+2} and p^{n+3} on columns 2 + k * stride ..
    // (k + 1) * stride + 1, and p^{n+1} one column further on each side: the
    // tiles cover columns 2 .. nx - 3 (a deep cell is two from the grid's
    // ends) and 1 .. nx - 2 for p^{n+1} (every interior cell).
    tp.tiles_x = std::max<uint32_t>(1, div_up(pl->nx > 4 ? pl->nx - 4 : 0, kTriStride<T>));
    tp.row_blocks = div_up(pl->ny, kTriRows);
    tp.tile_blocks = tp.tiles_x * tp.row_blocks * div_up(pl->nz - 2, zch);
    tp.tile_launch = div_up(tp.tile_blocks / tp.tiles_x, 8) * 8 * tp.tiles_x;
    tp.special_blocks = div_up(pl->n_special, kPairThreads);
    return tp;
}

// Enqueues steps [n, stop) of the plan's run on `stream`: three or two steps
// per pass where the run has them (a remainder of two steps as a two-step
// pass), single steps otherwise and for a last odd step. The plan's buf_p /
// buf_q follow the field buffers. Returns the number of launches.
template <class T, class S>
uint64_t enqueue_range(wv_plan* pl, size_t n, size_t stop, hipStream_t stream) {
    StepParams<T, S> a = make_params<T, S>(pl, kVariantTiles[pl->run_variant]);
    const uint32_t grid = a.main_blocks + div_up(pl->n_recv, variant_threads(pl->run_variant));
    const bool triples = pl->run_steps_per_pass == 3;
    const bool pairs = pl->run_steps_per_pass >= 2;
    PairParams<T, S> pp{};
    TriParams<T, S> tp{};
    uint32_t grid1 = 0, items2 = 0, grid2 = 0, gridt = 0, items3 = 0, grid3 = 0;
    if (pairs) {
        pp = make_pair_params<T, S>(pl, a);
        grid1 = pp.tile_launch + pp.special_blocks + div_up(pl->n_recv, kPairThreads);
        items2 = div_up(pl->n_special, kPair2Block) + div_up(pl->n_ring, kPair2Block);
        grid2 = div_up(items2, kPair2Bands) * kPair2Bands + div_up(pl->n_recv, kPair2Block);
    }
    if (triples) {
        tp = make_tri_params<T, S>(pl, a);
        gridt = tp.tile_launch + tp.special_blocks + div_up(pl->n_recv, kPairThreads);
        items3 = div_up(pl->n_special, kPair2Block) + div_up(pl->n_ring2, kPair2Block);
        grid3 = div_up(items3, kPair2Bands) * kPair2Bands + div_up(pl->n_recv, kPair2Block);
    }
    T* bufs[5] = {pl->fa.as<T>(), pl->fb.as<T>(), pl->fc.as<T>(), pl->fd.as<T>(), pl->fe.as<T>()};
    auto sig = [&](size_t m) -> S { return (S)(sizeof(S) == 4 ? (double)(float)pl->signal[m] : pl->signal[m]); };
    // The k-th field buffer that holds neither p nor prev.
    auto free_buf = [&](int k) -> int {
        for (int b = 0;; ++b)
            if (b != pl->buf_p && b != pl->buf_q && k-- == 0) return b;
    };
    T* ir = pl->ir.as<T>();
    uint64_t launches = 0;
    while (n < stop) {
        // Passes end at re-centring steps; the re-centring runs before the
        // step's update.
        if (recentres_at(pl, n) && pl->recentred_at != n) {
            recentre(pl, n, stream);
            launches += 3;
        }
        const size_t cap = std::min(stop, next_recentre(pl, n));
        T* P = bufs[pl->buf_p];
        T* Q = bufs[pl->buf_q];
        if (triples && n + 3 <= cap) {
            // Steps n, n+1, n+2 in three launches; outputs go to three free
            // buffers, which then hold p^{n+1} (read by passes 2 and 3 only),
            // p^{n+2} (prev) and p^{n+3} (p).
            const int o1 = free_buf(0), o2 = free_buf(1), o3 = free_buf(2);
            tp.a.p = P;
            tp.a.prev = Q;
            tp.a.s_n = sig(n);
            tp.a.ir = ir + n * pl->n_recv;
            tp.p1 = bufs[o1];
            tp.p2 = bufs[o2];
            tp.p3 = bufs[o3];
            hipLaunchKernelGGL((k_tri1<T, S>), dim3(gridt), dim3(kPairThreads), 0, stream, tp);
            StepParams<T, S> a2 = a;
            a2.p = bufs[o1];
            a2.prev = P;
            a2.s_n = sig(n + 1);
            a2.ir = ir + (n + 1) * pl->n_recv;
            hipLaunchKernelGGL((k_pair2<T, S>), dim3(grid2), dim3(kPair2Block), 0, stream, a2, bufs[o2],
                               pl->sp_cell.as<uint32_t>(), pl->ring_cell.as<uint32_t>(), pl->n_special, pl->n_ring,
                               items2);
            StepParams<T, S> a3 = a;
            a3.p = bufs[o2];
            a3.prev = bufs[o1];
            a3.s_n = sig(n + 2);
            a3.ir = ir + (n + 2) * pl->n_recv;
            hipLaunchKernelGGL((k_pair2<T, S>), dim3(grid3), dim3(kPair2Block), 0, stream, a3, bufs[o3],
                               pl->sp_cell.as<uint32_t>(), pl->ring2_cell.as<uint32_t>(), pl->n_special,
                               pl->n_ring2, items3);
            pl->buf_p = o3;
            pl->buf_q = o2;
            launches += 3;
            n += 3;
        } else if (pairs && n + 2 <= cap) {
            // Steps n, n+1 in two launches; outputs go to two free buffers,
            // which then hold p (p^{n+2}) and prev (p^{n+1}).
            const int o1 = free_buf(0), o2 = free_buf(1);
            pp.a.p = P;
            pp.a.prev = Q;
            pp.a.s_n = sig(n);
            pp.a.ir = ir + n * pl->n_recv;
            pp.p1 = bufs[o1];
            pp.p2 = bufs[o2];
            hipLaunchKernelGGL((k_pair1<T, S>), dim3(grid1), dim3(kPairThreads), 0, stream, pp);
            StepParams<T, S> a2 = a;
            a2.p = bufs[o1];
            a2.prev = P;
            a2.s_n = sig(n + 1);
            a2.ir = ir + (n + 1) * pl->n_recv;
            hipLaunchKernelGGL((k_pair2<T, S>), dim3(grid2), dim3(kPair2Block), 0, stream, a2, bufs[o2],
                               pl->sp_cell.as<uint32_t>(), pl->ring_cell.as<uint32_t>(), pl->n_special, pl->n_ring,
                               items2);
            pl->buf_p = o2;
            pl->buf_q = o1;
            launches += 2;
            n += 2;
        } else {
            // One step, updating prev in place (the reference swaps
            // buffers every step, gpu_fdtd.rs:242-243).
            a.p = P;
            a.prev = Q;
            a.s_n = sig(n);
            a.ir = ir + n * pl->n_recv;
            if (grid) launch_step<T, S>(pl->run_variant, grid, stream, a);
            std::swap(pl->buf_p, pl->buf_q);
            ++launches;
            ++n;
        }
    }
    WV_HIP(hipGetLastError());
    return launches;
}

template <class T, class S>
wv_status advance_impl(wv_plan* pl, size_t k, const volatile int32_t* cancel, wv_progress_fn progress, void* ctx) {
    const size_t end = pl->done + k;
    uint64_t launches = 0;
    int blk = 0;
    bool pending = false;
    wv_status st = WV_OK;
    WV_HIP(hipEventRecord(pl->t0, pl->stream));
    while (pl->done < end) {
        if (cancel && *cancel) { st = WV_CANCELLED; break; }
        const size_t stop = std::min(end, pl->done + kStepsPerBlock);
        launches += enqueue_range<T, S>(pl, pl->done, stop, pl->stream);
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
    // An empty signal reports no progress, as the reference's step loop
    // (gpu_fdtd.rs:200-260) makes no call then.
    if (st == WV_OK && progress && pl->steps) progress((double)pl->done / (double)pl->steps, ctx);
    float ms = 0.f;
    WV_HIP(hipEventElapsedTime(&ms, pl->t0, pl->t1));
    pl->last_ms = ms;
    pl->last_steps = k;
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
                           pl->prec == WV_FP64 ? nullptr : pl->d0h.as<double>(), pl->spec.as<double>());
        WV_HIP(hipGetLastError());
        WV_HIP(hipMemcpyAsync(spectra_out, pl->spec.p, R * nf * 16, hipMemcpyDeviceToHost, pl->stream));
    }
    if ((mode & WV_MODE_IR) && ir_out && R) {
        std::vector<T> h(Tn * R);
        std::vector<double> off(kOffsets, 0.0);
        if (Tn) WV_HIP(hipMemcpyAsync(h.data(), pl->ir.p, Tn * R * sizeof(T), hipMemcpyDeviceToHost, pl->stream));
        if (pl->prec != WV_FP64)
            WV_HIP(hipMemcpyAsync(off.data(), pl->d0h.p, kOffsets * 8, hipMemcpyDeviceToHost, pl->stream));
        WV_HIP(hipStreamSynchronize(pl->stream));
        for (size_t r = 0; r < R; ++r)
            for (size_t n = 0; n < Tn; ++n) {
                const double v = (double)h[n * R + r];
                // FP32: the stored value plus the offset in force at step n.
                ir_out[r * pl->steps + n] = pl->prec == WV_FP64 ? v : v + off[offset_index((uint32_t)n)];
            }
    }
    WV_HIP(hipStreamSynchronize(pl->stream));
}

}  // namespace

void wv::run_begin(wv_plan* pl, const wv_tap* taps, size_t n_taps, const double* signal, size_t steps,
                   const uint64_t* receivers, size_t n_recv) {
    if (pl->prec == WV_FP32) run_begin_impl<float>(pl, taps, n_taps, signal, steps, receivers, n_recv);
    else run_begin_impl<double>(pl, taps, n_taps, signal, steps, receivers, n_recv);
}

void wv::run_rewind(wv_plan* pl) {
    if (!pl->active) throw DevErr{"run_rewind: no run begun", WV_ERR_INVALID};
    if (pl->prec == WV_FP32) run_zero<float>(pl, pl->steps, pl->n_recv);
    else run_zero<double>(pl, pl->steps, pl->n_recv);
}

uint64_t wv::enqueue_steps(wv_plan* pl, size_t n, size_t stop, hipStream_t stream) {
    switch (pl->prec) {
        case WV_FP32: return enqueue_range<float, float>(pl, n, stop, stream);
        case WV_FP64: return enqueue_range<double, double>(pl, n, stop, stream);
        default: return enqueue_range<float, double>(pl, n, stop, stream);
    }
}

void wv::run_finish(wv_plan* pl, const double* freqs, size_t nf, int mode, double* spectra_out, double* ir_out) {
    if (pl->prec == WV_FP64) finish_impl<double>(pl, freqs, nf, mode, spectra_out, ir_out);
    else finish_impl<float>(pl, freqs, nf, mode, spectra_out, ir_out);
}

wv_status wv::check_run_args(const wv_plan* pl, const wv_tap* taps, size_t n_taps, const double* signal,
                             size_t steps, const uint64_t* receivers, size_t n_recv, char* err, size_t err_len) {
    if (!pl || (n_taps && !taps) || (steps && !signal) || (n_recv && !receivers)) {
        set_err(err, err_len, "invalid run arguments");
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
    if ((uint64_t)steps * n_recv >= (1ull << 34) || steps >= (1ull << 31)) {
        set_err(err, err_len, "receiver record too large");
        return WV_ERR_LIMIT;
    }
    return WV_OK;
}

uint64_t wv::plan_bytes(const wv_plan* pl) {
    uint64_t t = 0;
    for (const DBuf* d : {&pl->flags, &pl->force_base, &pl->force, &pl->zbits, &pl->dbits, &pl->d2bits, &pl->counts,
                          &pl->offs, &pl->b_cell_sorted, &pl->b_set_sorted, &pl->set_n, &pl->set_off, &pl->set_c,
                          &pl->st_u, &pl->st_g, &pl->fa, &pl->fb, &pl->fc, &pl->fd, &pl->fe, &pl->sp_rec,
                          &pl->sp_tapw, &pl->sp_cell, &pl->ring_cell, &pl->ring_counts, &pl->ring_offs,
                          &pl->ring2_cell, &pl->ring2_counts, &pl->ring2_offs, &pl->scan_tmp, &pl->d0, &pl->prev_mean, &pl->d0h, &pl->mean_part,
                          &pl->shift, &pl->tap_cell, &pl->tap_w, &pl->recv,
                          &pl->ir, &pl->freqs, &pl->spec})
        t += d->bytes;
    return t;
}

uint64_t wv::estimate_bytes(const wv_grid* g, int prec, size_t steps, size_t n_recv, size_t n_taps,
                            uint32_t steps_per_pass) {
    const uint64_t tsz = prec == WV_FP64 ? 8 : 4, ssz = prec == WV_FP32 ? 4 : 8;
    const uint64_t cells = g->n[0] * g->n[1] * g->n[2];
    const uint64_t pitch = (g->n[0] + 31) & ~31ull;
    const uint64_t npitched = pitch * g->n[1] * g->n[2];
    uint32_t max_slots = 0;
    for (uint64_t b = 0; b < g->n_boundary; ++b) max_slots = std::max(max_slots, g->b_start[b + 1] - g->b_start[b]);
    const uint64_t special = g->n_boundary + n_taps;
    const uint64_t sp_pad = (special + 63) & ~63ull;
    const uint64_t slots = g->n_boundary ? g->b_start[g->n_boundary] : 0;
    const uint32_t spp = steps_per_pass ? steps_per_pass : default_steps_per_pass(cells, prec == WV_FP64);
    const uint64_t fields = spp == 3 ? 5 : spp == 2 ? 4 : 2;
    uint64_t t = cells;                              // flags
    t += npitched / 4 + npitched / 4;                // boundary marks (two copies), interior and deep bitmaps
    t += spp == 3 ? npitched / 8 : 0;                // deep2 bitmap
    t += fields * npitched * tsz;                    // fields
    t += spp >= 2 ? 2 * 4 * special : 0;             // special cells' indices and the ring list (about as long)
    t += spp == 3 ? 2 * 4 * special : 0;             // the cells within two of a wall
    t += npitched / 16 + 1024;                       // compaction counts and offsets (z-chunk >= 2)
    t += 8 * g->n_boundary + slots * (4 * ssz + 8);  // sorted boundary, slot-set table (at most one set per node)
    t += special * (8 + ssz) + 2 * (uint64_t)max_slots * sp_pad * ssz;  // records, tap weights, slot state
    t += (uint64_t)steps * n_recv * tsz + steps * ssz + 16 * (n_recv + n_taps);
    return t + t / 20 + (1u << 16);
}

wv_status wv::check_grid(const wv_grid* g, char* err, size_t err_len) {
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
    const uint64_t pitched = ((g->n[0] + 31) & ~31ull) * g->n[1] * g->n[2];
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
    plan->variant = variant < 0 ? -1 : variant;
    return WV_OK;
}

extern "C" wv_status wv_plan_set_pass(wv_plan* plan, uint32_t steps_per_pass, uint32_t pair_zchunk) {
    // Pass z-chunks are powers of two up to 64 (pass_zchunk caps them per mode).
    if (!plan || steps_per_pass > 3 || pair_zchunk > 64 || (pair_zchunk & (pair_zchunk - 1))) return WV_ERR_INVALID;
    plan->steps_per_pass = steps_per_pass;
    plan->pair_zchunk = pair_zchunk;
    return WV_OK;
}

extern "C" wv_status wv_run_begin(wv_plan* pl, const wv_tap* taps, size_t n_taps, const double* signal, size_t steps,
                                  const uint64_t* receivers, size_t n_recv, char* err, size_t err_len) {
    wv_status st = check_run_args(pl, taps, n_taps, signal, steps, receivers, n_recv, err, err_len);
    if (st != WV_OK) return st;
    return guarded(err, err_len, [&]() -> wv_status {
        WV_HIP(hipSetDevice(pl->device));
        run_begin(pl, taps, n_taps, signal, steps, receivers, n_recv);
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
        run_finish(pl, freqs, nf, mode, spectra_out, ir_out);
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

// ------------------------------------------------ allocation-free entry point
//
// wv_fdtd_run_noalloc runs wv_fdtd_run's set-up and kernels with every device
// buffer of the run (fields, bitmaps, special-cell records, slot state,
// receiver record, spectra) carved from `scratch`, one static array per GPU,
// so the call allocates no device memory. The first call on a device looks
// up that device's address of the array and creates the device's stream and
// events; they are kept until the process ends. Calls on one device share the
// array and run one at a time (a per-device lock); every buffer the run reads
// is written by the run's own set-up (fields, state and record by fill
// kernels, marks and bitmaps by their kernels), so nothing depends on what
// the last call left in the array.

static constexpr size_t kNoAllocBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

struct NoAllocDevice {
    char* ptr = nullptr;  // this device's address of `scratch`
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::mutex mu;  // one call at a time on this device
};

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
NoAllocDevice* g_noalloc[kMaxDevices] = {};

// The current device's array address, stream and events, set up on the
// device's first call.
NoAllocDevice* noalloc_device() {
    int dev = 0;
    WV_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw DevErr{"device ordinal out of range", WV_ERR_DEVICE};
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc[dev]) {
        auto d = std::make_unique<NoAllocDevice>();
        void* ptr = nullptr;
        WV_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw DevErr{"no device address for the scratch array", WV_ERR_DEVICE};
        d->ptr = static_cast<char*>(ptr);
        WV_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
        WV_HIP(hipEventCreate(&d->ev[0]));
        WV_HIP(hipEventCreate(&d->ev[1]));
        WV_HIP(hipEventCreate(&d->t0));
        WV_HIP(hipEventCreate(&d->t1));
        g_noalloc[dev] = d.release();
    }
    return g_noalloc[dev];
}

// Installs an arena on the calling thread for the lifetime of the object.
struct ArenaScope {
    explicit ArenaScope(Arena* a) { t_arena = a; }
    ~ArenaScope() { t_arena = nullptr; }
};

}  // namespace

extern "C" wv_status wv_fdtd_run_noalloc(const wv_grid* grid, const wv_tap* taps, size_t n_taps,
                                         const double* signal, size_t steps, const uint64_t* receivers,
                                         size_t n_recv, const double* freqs, size_t nf, int mode, int precision,
                                         const volatile int32_t* cancel, wv_progress_fn progress, void* ctx,
                                         double* spectra_out, double* ir_out, char* err, size_t err_len) {
    if (precision != WV_FP32 && precision != WV_FP64 && precision != WV_FP32_STATE64) {
        set_err(err, err_len, "wv_fdtd_run_noalloc: invalid precision");
        return WV_ERR_INVALID;
    }
    wv_status s = check_grid(grid, err, err_len);
    if (s != WV_OK) return s;
    if ((nf && !freqs) || (n_taps && !taps) || (steps && !signal) || (n_recv && !receivers)) {
        set_err(err, err_len, "wv_fdtd_run_noalloc: invalid argument");
        return WV_ERR_INVALID;
    }
    uint64_t need = estimate_bytes(grid, precision, steps, n_recv, n_taps);
    uint32_t steps_per_pass = 0;
    if (need > kNoAllocBytes) {
        // Single steps hold two field buffers instead of four or five: the
        // run takes them where the default pass mode does not fit (the
        // results do not depend on the mode).
        const uint64_t single = estimate_bytes(grid, precision, steps, n_recv, n_taps, 1);
        if (single <= kNoAllocBytes) {
            need = single;
            steps_per_pass = 1;
        }
    }
    if (need > kNoAllocBytes) {
        set_err(err, err_len, "wv_fdtd_run_noalloc: the run needs about " + std::to_string((need + (1u << 20) - 1) >> 20) +
                                  " MiB of device memory; the allocation-free entry point holds " +
                                  std::to_string(kNoAllocBytes >> 20) + " MiB (use wv_fdtd_run)");
        return WV_ERR_LIMIT;
    }
    return guarded(err, err_len, [&]() -> wv_status {
        NoAllocDevice* d = noalloc_device();
        std::lock_guard<std::mutex> lk(d->mu);
        Arena arena;
        arena.base = d->ptr;
        arena.cap = kNoAllocBytes;
        ArenaScope scope(&arena);
        wv_plan pl;
        pl.prec = precision;
        pl.steps_per_pass = steps_per_pass;
        pl.borrowed = true;
        pl.stream = d->stream;
        pl.ev[0] = d->ev[0];
        pl.ev[1] = d->ev[1];
        pl.t0 = d->t0;
        pl.t1 = d->t1;
        plan_setup(&pl, grid);
        wv_status st = check_run_args(&pl, taps, n_taps, signal, steps, receivers, n_recv, err, err_len);
        if (st != WV_OK) return st;
        run_begin(&pl, taps, n_taps, signal, steps, receivers, n_recv);
        switch (pl.prec) {
            case WV_FP32: st = advance_impl<float, float>(&pl, steps, cancel, progress, ctx); break;
            case WV_FP64: st = advance_impl<double, double>(&pl, steps, cancel, progress, ctx); break;
            default: st = advance_impl<float, double>(&pl, steps, cancel, progress, ctx); break;
        }
        if (st != WV_OK) return st;
        run_finish(&pl, freqs, nf, mode, spectra_out, ir_out);
        return WV_OK;
    });
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

extern "C" wv_status wv_plan_pass_info(const wv_plan* plan, uint32_t out[4]) {
    if (!plan || !out) return WV_ERR_INVALID;
    const uint32_t spp = plan->steps_per_pass ? plan->steps_per_pass : default_steps_per_pass(plan->cells, plan->prec == WV_FP64);
    out[0] = plan->active ? plan->run_steps_per_pass : spp;
    out[1] = pass_zchunk(plan, out[0]);
    out[2] = plan->n_ring;
    out[3] = plan->n_ring2;
    return WV_OK;
}

extern "C" wv_status wv_footprint_estimate(const wv_grid* grid, int precision, size_t steps, size_t n_recv,
                                           size_t n_taps, uint64_t* bytes) {
    char err[8];
    if (!bytes || check_grid(grid, err, sizeof err) != WV_OK) return WV_ERR_INVALID;
    *bytes = estimate_bytes(grid, precision, steps, n_recv, n_taps);
    return WV_OK;
}
