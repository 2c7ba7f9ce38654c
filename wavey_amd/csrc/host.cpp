// Host side of the wave solver: box-room grid builder, source and receiver
// helpers, admittance models. These are the caller-side pieces of the
// reference path (core/wavey-solver/src/fdtd.rs) that produce the inputs the
// device run takes; the run itself is in fdtd_device.hip.
//
// Floating-point expressions that define outputs the reference also produces
// (slot constants, tap weights, source samples) keep the reference's
// evaluation order and are built with -ffp-contract=off, so the grids and
// taps match the reference's bit for bit.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "wavey_fdtd.h"
#include "grid_internal.h"
#include "wv_internal.h"

namespace wv {

const double kCourant = 0.5773502691896257;  // fdtd.rs:138
const uint64_t kMaxCells = 600000000ull;     // fdtd.rs:140
static const double kPi = 3.141592653589793;

void set_err(char* err, size_t len, const std::string& msg) {
    if (!err || len == 0) return;
    size_t n = std::min(len - 1, msg.size());
    std::memcpy(err, msg.data(), n);
    err[n] = 0;
}

int pick_threads(int threads) {
    if (threads > 0) return threads;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)std::min(h, 64u) : 4;
}

// fdtd.rs:539-571 one branch -> slot constants; count faces share the fit.
wv_slot slot_of(double r, double m, double k, double count, double dt) {
    r = r / count;
    m = m / count;
    k = k / count;
    wv_slot s;
    if (m > 0.0) {
        double d = m / dt + r / 2.0 + k * dt / 4.0;
        s.cv = m / dt - r / 2.0 - k * dt / 4.0;
        s.inv_d = 1.0 / d;
        s.w = 0.25;
    } else {
        double d = r + k * dt / 2.0;
        s.cv = -k * dt / 2.0;
        s.inv_d = 1.0 / d;
        s.w = 0.0;
    }
    s.k = k;
    return s;
}

bool fits_equal(const wv_admittance& a, const wv_admittance& b) {
    if (a.beta0 != b.beta0 || a.n_branches != b.n_branches) return false;
    for (uint32_t i = 0; i < a.n_branches; ++i)
        if (a.branches[i].r != b.branches[i].r || a.branches[i].m != b.branches[i].m ||
            a.branches[i].k != b.branches[i].k)
            return false;
    return a.band[0] == b.band[0] && a.band[1] == b.band[1] && a.rms_error == b.rms_error &&
           a.max_error == b.max_error;
}

std::vector<int32_t> canonical_fits(const wv_admittance* fits, uint32_t n) {
    // Candidates are bucketed by beta0 and the branch count (equal fits have
    // equal ones; -0 and +0 compare equal, so both hash as +0).
    std::vector<int32_t> canon(n);
    std::unordered_map<uint64_t, std::vector<int32_t>> by_head;
    for (uint32_t i = 0; i < n; ++i) {
        const wv_admittance& a = fits[i];
        if (!fits_equal(a, a)) {
            canon[i] = ~(int32_t)i;
            continue;
        }
        const double b0 = a.beta0 == 0.0 ? 0.0 : a.beta0;
        uint64_t h;
        std::memcpy(&h, &b0, 8);
        h ^= (uint64_t)a.n_branches * 0x9E3779B97F4A7C15ull;
        auto& bucket = by_head[h];
        canon[i] = (int32_t)i;
        for (int32_t j : bucket)
            if (fits_equal(fits[j], a)) { canon[i] = j; break; }
        if (canon[i] == (int32_t)i) bucket.push_back((int32_t)i);
    }
    return canon;
}

// fdtd.rs:539-571: faces grouped by fit in order of first appearance; each
// group gives its conductance slot (beta0 > 0) and one slot per branch, with
// the impedances divided by the group's face count. A negative canonical
// index ~i is fit i that equals no fit: each of its faces is a group.
void node_slots(const int* canon_fits, int n_faces, const wv_admittance* fits, double dt,
                std::vector<wv_slot>& v) {
    int grp[6], cnt[6], ng = 0;
    for (int f = 0; f < n_faces; ++f) {
        int fit = canon_fits[f];
        int j = 0;
        if (fit >= 0)
            while (j < ng && grp[j] != fit) ++j;
        else
            j = ng;
        if (j < ng) cnt[j]++;
        else { grp[ng] = fit; cnt[ng] = 1; ++ng; }
    }
    v.clear();
    for (int j = 0; j < ng; ++j) {
        const wv_admittance& a = fits[grp[j] >= 0 ? grp[j] : ~grp[j]];
        double n = (double)cnt[j];
        if (a.beta0 > 0.0) v.push_back(slot_of(1.0 / a.beta0, 0.0, 0.0, n, dt));
        for (uint32_t b = 0; b < a.n_branches; ++b)
            v.push_back(slot_of(a.branches[b].r, a.branches[b].m, a.branches[b].k, n, dt));
    }
}

void finish_view(wv_built_grid& g, const int64_t n[3], const double origin[3], double dx, double dt) {
    wv_grid& v = g.view;
    for (int k = 0; k < 3; ++k) {
        v.n[k] = (uint64_t)n[k];
        v.origin[k] = origin[k];
    }
    v.dx = dx;
    v.dt = dt;
    v.flags = g.flags.data();
    v.n_boundary = g.b_cell.size();
    v.b_cell = g.b_cell.data();
    v.b_start = g.b_start.data();
    v.slots = g.slots.data();
}

}  // namespace wv

using namespace wv;

namespace {

// Analytic description of a box room with pillars, on the grid of spacing dx.
struct BoxGrid {
    const wv_box_room* room;
    double dx, origin;  // origin is the same on every axis: 0 - dx/2
    int64_t n[3];
    std::vector<uint8_t> ax, ay, az;   // per-axis: inside the padding and the box
    std::vector<int32_t> pillar2d;     // [ny*nx]: pillar index or -1
    std::vector<int32_t> canon;        // fit index -> first equal fit

    double center(int64_t i) const { return origin + (double)i * dx; }

    bool air(int64_t x, int64_t y, int64_t z) const {
        return ax[x] && ay[y] && az[z] && pillar2d[y * n[0] + x] < 0;
    }

    int patch(uint32_t s, double u, double v, double u0, double u1, double v0, double v1) const {
        int gu = (int)room->patch_gu[s], gv = (int)room->patch_gv[s];
        int iu = std::clamp((int)std::floor((u - u0) / (u1 - u0) * gu), 0, gu - 1);
        int iv = std::clamp((int)std::floor((v - v0) / (v1 - v0) * gv), 0, gv - 1);
        return canon[room->patch_fit[room->patch_off[s] + (uint32_t)(iv * gu + iu)]];
    }

    // The surface crossed halfway from the node toward its missing neighbour
    // in direction d (the BVH face search of fdtd.rs:358-362 in a box room).
    int face_fit(int64_t x, int64_t y, int64_t z, int d) const {
        static const int DX[6] = {-1, 1, 0, 0, 0, 0}, DY[6] = {0, 0, -1, 1, 0, 0}, DZ[6] = {0, 0, 0, 0, -1, 1};
        double px = center(x), py = center(y), pz = center(z);
        double hx = px + DX[d] * dx * 0.5, hy = py + DY[d] * dx * 0.5, hz = pz + DZ[d] * dx * 0.5;
        const double* L = room->size;
        if (d < 4) {
            int64_t qx = x + DX[d], qy = y + DY[d];
            int p = pillar2d[qy * n[0] + qx];
            if (p >= 0) {
                const double* q = room->pillars + 4 * p;
                uint32_t s = 6u + 4u * (uint32_t)p + (uint32_t)(d ^ 1);
                return d < 2 ? patch(s, hy, hz, q[2], q[3], 0.0, L[2]) : patch(s, hx, hz, q[0], q[1], 0.0, L[2]);
            }
        }
        switch (d) {
            case 0: return patch(0, hy, hz, 0.0, L[1], 0.0, L[2]);
            case 1: return patch(1, hy, hz, 0.0, L[1], 0.0, L[2]);
            case 2: return patch(2, hx, hz, 0.0, L[0], 0.0, L[2]);
            case 3: return patch(3, hx, hz, 0.0, L[0], 0.0, L[2]);
            case 4: return patch(4, hx, hy, 0.0, L[0], 0.0, L[1]);
            default: return patch(5, hx, hy, 0.0, L[0], 0.0, L[1]);
        }
    }
};

// A node's missing-face fits in direction order, 8 bits per face (fit + 1),
// zero-terminated: the key of its slot list.
// A node's faces as their canonical fits, 9 bits a face: c + 1 for a fit
// index c (0..254), 257 + i for ~i (fit i equals no fit).
struct FaceSig {
    uint64_t key = 0;
    int n = 0;
    void push(int fit) { key |= (uint64_t)(fit >= 0 ? fit + 1 : 257 + ~fit) << (9 * n); ++n; }
    static int fit_of(uint64_t v) { return v <= 256 ? (int)v - 1 : ~(int)(v - 257); }
};

}  // namespace

static wv_status grid_build_box_impl(const wv_box_room* room, double c, double dx, int threads,
                                     wv_built_grid** out, char* err, size_t err_len) {
    if (!room || !out || !(dx > 0.0) || !(c > 0.0)) {
        set_err(err, err_len, "wv_grid_build_box: invalid argument");
        return WV_ERR_INVALID;
    }
    if (room->n_fits == 0 || room->n_fits > 255 || !room->fits || !room->patch_gu || !room->patch_gv ||
        !room->patch_off || !room->patch_fit || (room->n_pillars && !room->pillars)) {
        set_err(err, err_len, "wv_grid_build_box: room needs 1..255 fits and patch tables for every surface");
        return WV_ERR_INVALID;
    }
    threads = pick_threads(threads);
    BoxGrid B;
    B.room = room;
    B.dx = dx;
    B.origin = 0.0 - dx * 0.5;  // fdtd.rs:293 with lo = 0
    const double dt = kCourant * dx / c;  // fdtd.rs:290
    for (int k = 0; k < 3; ++k) B.n[k] = (int64_t)std::ceil(room->size[k] / dx) + 2;  // fdtd.rs:295-299
    const uint64_t cells = (uint64_t)B.n[0] * B.n[1] * B.n[2];
    if (cells > kMaxCells) {
        set_err(err, err_len, "the wave solver grid would have " + std::to_string(cells / 1000000) +
                                  " M cells; lower the crossover frequency");
        return WV_ERR_LIMIT;
    }
    B.canon = canonical_fits(room->fits, room->n_fits);
    for (int k = 0; k < 3; ++k) {
        auto& a = k == 0 ? B.ax : k == 1 ? B.ay : B.az;
        a.resize(B.n[k]);
        for (int64_t i = 0; i < B.n[k]; ++i) {
            double p = B.center(i);
            a[i] = i > 0 && i < B.n[k] - 1 && p > 0.0 && p < room->size[k];
        }
    }
    B.pillar2d.assign(B.n[0] * B.n[1], -1);
    for (uint32_t p = 0; p < room->n_pillars; ++p) {
        const double* q = room->pillars + 4 * p;
        for (int64_t y = 0; y < B.n[1]; ++y)
            for (int64_t x = 0; x < B.n[0]; ++x) {
                double px = B.center(x), py = B.center(y);
                if (B.pillar2d[y * B.n[0] + x] < 0 && px > q[0] && px < q[1] && py > q[2] && py < q[3])
                    B.pillar2d[y * B.n[0] + x] = (int32_t)p;
            }
    }

    auto g = std::make_unique<wv_built_grid>();
    g->flags.assign(cells, 0);
    const int64_t nx = B.n[0], ny = B.n[1], nz = B.n[2];
    // Per z-slab: flags, and the boundary nodes (cell, face signature) in cell order.
    int nthreads = threads;
    std::vector<std::vector<std::pair<uint32_t, uint64_t>>> part(nthreads);
    parallel_chunks(nz, nthreads, [&](int64_t z0, int64_t z1, int t) {
        auto& mine = part[t];
        for (int64_t z = z0; z < z1; ++z) {
            if (!B.az[z]) continue;
            for (int64_t y = 1; y < ny - 1; ++y) {
                if (!B.ay[y]) continue;
                const int64_t row = (z * ny + y) * nx;
                for (int64_t x = 1; x < nx - 1; ++x) {
                    if (!B.air(x, y, z)) continue;
                    const bool nb[6] = {B.air(x - 1, y, z), B.air(x + 1, y, z), B.air(x, y - 1, z),
                                        B.air(x, y + 1, z), B.air(x, y, z - 1), B.air(x, y, z + 1)};
                    uint8_t f = 64;
                    FaceSig sig;
                    for (int d = 0; d < 6; ++d) {
                        if (nb[d]) f |= (uint8_t)(1u << d);
                        else sig.push(B.face_fit(x, y, z, d));
                    }
                    g->flags[row + x] = f;
                    if (sig.n) mine.emplace_back((uint32_t)(row + x), sig.key);
                }
            }
        }
    });
    // Slot lists per distinct face signature (fdtd.rs:539-571).
    std::map<uint64_t, std::vector<wv_slot>> cache;
    auto slots_for = [&](uint64_t key) -> const std::vector<wv_slot>& {
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
        int faces[6], nf = 0;
        while (nf < 6 && ((key >> (9 * nf)) & 0x1ff)) {
            faces[nf] = FaceSig::fit_of((key >> (9 * nf)) & 0x1ff);
            ++nf;
        }
        std::vector<wv_slot> v;
        node_slots(faces, nf, room->fits, dt, v);
        return cache.emplace(key, std::move(v)).first->second;
    };
    size_t nb_total = 0;
    for (auto& p : part) nb_total += p.size();
    g->b_cell.reserve(nb_total);
    g->b_start.reserve(nb_total + 1);
    for (auto& p : part)
        for (auto& [cell, key] : p) {
            const auto& sl = slots_for(key);
            if (sl.empty()) continue;  // rigid faces only: no boundary node
            g->b_cell.push_back(cell);
            g->b_start.push_back((uint32_t)g->slots.size());
            g->slots.insert(g->slots.end(), sl.begin(), sl.end());
        }
    g->b_start.push_back((uint32_t)g->slots.size());
    const double org[3] = {B.origin, B.origin, B.origin};
    finish_view(*g, B.n, org, dx, dt);
    *out = g.release();
    return WV_OK;
}

extern "C" wv_status wv_grid_build_box(const wv_box_room* room, double c, double dx, int threads,
                                       wv_built_grid** out, char* err, size_t err_len) {
    return host_guarded(err, err_len, [&]() { return grid_build_box_impl(room, c, dx, threads, out, err, err_len); });
}

extern "C" const wv_grid* wv_built_grid_view(const wv_built_grid* g) { return g ? &g->view : nullptr; }
extern "C" void wv_grid_free(wv_built_grid* g) { delete g; }

// ------------------------------------------------------------- sources

extern "C" void wv_source_signal(double f_max, double dt, size_t steps, double* out) {
    // fdtd.rs:601-610: differentiated Gaussian, centred 6 sigma in.
    const double sigma = 0.342 / f_max;
    const double t0 = 6.0 * sigma;
    for (size_t n = 0; n < steps; ++n) {
        const double t = (double)n * dt - t0;
        out[n] = -t / (sigma * sigma) * std::exp(-(t * t) / (2.0 * sigma * sigma)) * dt;
    }
}

namespace {
inline bool in_grid(const wv_grid* g, int64_t x, int64_t y, int64_t z) {
    return x >= 0 && y >= 0 && z >= 0 && x < (int64_t)g->n[0] && y < (int64_t)g->n[1] && z < (int64_t)g->n[2];
}
inline uint64_t cell_of(const wv_grid* g, int64_t x, int64_t y, int64_t z) {
    return ((uint64_t)z * g->n[1] + (uint64_t)y) * g->n[0] + (uint64_t)x;
}
}  // namespace

extern "C" wv_status wv_nearest_air(const wv_grid* g, const double p[3], uint64_t* cell) {
    // fdtd.rs:245-272: growing cubes around the rounded node, up to 16 cells;
    // the first strictly nearest air node of the first cube that has one.
    int64_t c[3];
    for (int k = 0; k < 3; ++k) c[k] = (int64_t)std::round((p[k] - g->origin[k]) / g->dx);
    for (int64_t r = 0; r <= 16; ++r) {
        bool found = false;
        double best = 0.0;
        uint64_t bi = 0;
        for (int64_t dz = -r; dz <= r; ++dz)
            for (int64_t dy = -r; dy <= r; ++dy)
                for (int64_t dx = -r; dx <= r; ++dx) {
                    const int64_t x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
                    if (!in_grid(g, x, y, z)) continue;
                    const uint64_t i = cell_of(g, x, y, z);
                    if (!(g->flags[i] & 64)) continue;
                    const double ex = g->origin[0] + (double)x * g->dx - p[0];
                    const double ey = g->origin[1] + (double)y * g->dx - p[1];
                    const double ez = g->origin[2] + (double)z * g->dx - p[2];
                    const double d2 = ex * ex + ey * ey + ez * ez;
                    if (!found || d2 < best) { found = true; best = d2; bi = i; }
                }
        if (found) { *cell = bi; return WV_OK; }
    }
    return WV_ERR_INVALID;
}

extern "C" void wv_baffle_point(const wv_grid* g, const double p[3], const double axis[3], double out[3]) {
    // fdtd.rs:497-512: march from p along the axis in steps of dx/20 (for
    // up to 2 m) to the first point whose nearest node is air; the driver
    // sits halfway back from there, on the face between that node and the
    // solid one.
    const double step = g->dx * 0.05;
    const int64_t kmax = (int64_t)(2.0 / step);
    for (int64_t k = 0; k < kmax; ++k) {
        double q[3], r[3];
        bool neg = false;
        for (int i = 0; i < 3; ++i) {
            q[i] = p[i] + axis[i] * ((double)k * step);
            r[i] = std::round((q[i] - g->origin[i]) / g->dx);
            neg = neg || r[i] < 0.0;
        }
        if (neg) continue;
        const int64_t x = (int64_t)r[0], y = (int64_t)r[1], z = (int64_t)r[2];
        if (in_grid(g, x, y, z) && (g->flags[cell_of(g, x, y, z)] & 64)) {
            const double m = std::fmin(g->dx * 0.5, (double)k * step);
            for (int i = 0; i < 3; ++i) out[i] = q[i] - axis[i] * m;
            return;
        }
    }
    for (int i = 0; i < 3; ++i) out[i] = p[i];
}

extern "C" wv_status wv_air_along(const wv_grid* g, const double p[3], const double axis[3], uint64_t* cell) {
    // fdtd.rs:516-533: from p along the axis in steps of dx/2 (for up to
    // 2 m), the first point whose rounded node is air; else the nearest air
    // node to p.
    const double step = g->dx * 0.5;
    const int64_t kmax = (int64_t)(2.0 / step);
    for (int64_t k = 0; k < kmax; ++k) {
        double r[3];
        bool neg = false;
        for (int i = 0; i < 3; ++i) {
            const double q = p[i] + axis[i] * ((double)k * step);
            r[i] = std::round((q - g->origin[i]) / g->dx);
            neg = neg || r[i] < 0.0;
        }
        if (neg) continue;
        const int64_t x = (int64_t)r[0], y = (int64_t)r[1], z = (int64_t)r[2];
        if (in_grid(g, x, y, z) && (g->flags[cell_of(g, x, y, z)] & 64)) {
            *cell = cell_of(g, x, y, z);
            return WV_OK;
        }
    }
    return wv_nearest_air(g, p, cell);
}

extern "C" wv_status wv_wave_source_taps(const wv_grid* g, const double center[3], const double axis[3],
                                         int calibrated, uint32_t order, wv_tap comps[8], uint32_t comp_len[4],
                                         uint32_t* n_comps, int32_t* omni_fallback) {
    // run.rs:601-624. The source node: for a calibrated cabinet the first air
    // node along the axis, else the nearest air node, else the first air node
    // along the axis. The driver: on the baffle in front of the acoustic
    // centre with a calibration, else the centre itself. A calibrated or
    // order-0 source is the driver's trilinear taps (or the source node);
    // otherwise the source node's components, or the node alone when it is
    // too close to a surface for the dipoles.
    uint64_t node = 0;
    bool have = calibrated && wv_air_along(g, center, axis, &node) == WV_OK;
    if (!have) have = wv_nearest_air(g, center, &node) == WV_OK;
    if (!have) have = wv_air_along(g, center, axis, &node) == WV_OK;
    if (!have) return WV_ERR_INVALID;
    if (omni_fallback) *omni_fallback = 0;
    if (calibrated || order == 0) {
        double driver[3] = {center[0], center[1], center[2]};
        if (calibrated) wv_baffle_point(g, center, axis, driver);
        size_t nt = 0;
        if (wv_taps_at(g, driver, comps, &nt) != WV_OK) {
            comps[0] = wv_tap{node, 1.0};
            nt = 1;
        }
        comp_len[0] = (uint32_t)nt;
        *n_comps = 1;
        return WV_OK;
    }
    if (wv_source_components(g, node, order, comps, comp_len, n_comps) != WV_OK) {
        comps[0] = wv_tap{node, 1.0};
        comp_len[0] = 1;
        *n_comps = 1;
        if (omni_fallback) *omni_fallback = 1;
    }
    return WV_OK;
}

extern "C" void wv_blend(const double ism[2], const double wave[2], double w, double out[2]) {
    // fdtd.rs:847-862: the phase of the complex crossfade with the energy
    // (1 - w)|ism|^2 + w|wave|^2; where the two nearly cancel, the larger
    // weighted field's phase.
    if (w <= 0.0) { out[0] = ism[0]; out[1] = ism[1]; return; }
    if (w >= 1.0) { out[0] = wave[0]; out[1] = wave[1]; return; }
    const double ni = ism[0] * ism[0] + ism[1] * ism[1], nw = wave[0] * wave[0] + wave[1] * wave[1];
    const double target = (1.0 - w) * ni + w * nw;
    const double pr = ism[0] * (1.0 - w) + wave[0] * w, pi = ism[1] * (1.0 - w) + wave[1] * w;
    const double m2 = pr * pr + pi * pi;
    if (m2 < 1e-6 * target) {
        const double* d = nw * w >= ni * (1.0 - w) ? wave : ism;
        const double n = std::fmax(std::hypot(d[0], d[1]), 1e-300), st = std::sqrt(target);
        out[0] = d[0] / n * st;
        out[1] = d[1] / n * st;
        return;
    }
    const double sc = std::sqrt(target / m2);
    out[0] = pr * sc;
    out[1] = pi * sc;
}

extern "C" wv_status wv_taps_at(const wv_grid* g, const double p[3], wv_tap taps[8], size_t* n_taps) {
    // fdtd.rs:467-492: trilinear weights over the 8 surrounding nodes, solid
    // ones dropped and the rest renormalised; under 1/4 total weight the
    // nearest air node takes it all.
    double q[3], base[3], fr[3];
    for (int k = 0; k < 3; ++k) {
        q[k] = (p[k] - g->origin[k]) / g->dx;
        base[k] = std::floor(q[k]);
        fr[k] = q[k] - base[k];
    }
    size_t nt = 0;
    for (int ix = 0; ix < 2; ++ix)
        for (int iy = 0; iy < 2; ++iy)
            for (int iz = 0; iz < 2; ++iz) {
                const double wx = ix ? fr[0] : 1.0 - fr[0];
                const double wy = iy ? fr[1] : 1.0 - fr[1];
                const double wz = iz ? fr[2] : 1.0 - fr[2];
                const int64_t x = (int64_t)base[0] + ix, y = (int64_t)base[1] + iy, z = (int64_t)base[2] + iz;
                if (!in_grid(g, x, y, z)) continue;
                const uint64_t i = cell_of(g, x, y, z);
                const double w = wx * wy * wz;
                if ((g->flags[i] & 64) && w > 1e-9) taps[nt++] = wv_tap{i, w};
            }
    double sum = 0.0;
    for (size_t t = 0; t < nt; ++t) sum += taps[t].w;
    if (sum < 0.25) {
        uint64_t c;
        if (wv_nearest_air(g, p, &c) != WV_OK) return WV_ERR_INVALID;
        taps[0] = wv_tap{c, 1.0};
        *n_taps = 1;
        return WV_OK;
    }
    for (size_t t = 0; t < nt; ++t) taps[t].w = taps[t].w / sum;
    *n_taps = nt;
    return WV_OK;
}

extern "C" wv_status wv_source_components(const wv_grid* g, uint64_t c, uint32_t order, wv_tap comps[8],
                                          uint32_t comp_len[4], uint32_t* n_comps) {
    // fdtd.rs:730-743: the monopole, then +-1 dipoles along x, y, z.
    const uint64_t cells = g->n[0] * g->n[1] * g->n[2];
    comps[0] = wv_tap{c, 1.0};
    comp_len[0] = 1;
    *n_comps = 1;
    if (order < 1) return WV_OK;
    const uint64_t stride[3] = {1, g->n[0], g->n[0] * g->n[1]};
    for (int s = 0; s < 3; ++s) {
        if (c < stride[s]) return WV_ERR_INVALID;
        const uint64_t plus = c + stride[s], minus = c - stride[s];
        if (plus >= cells || !(g->flags[plus] & 64) || !(g->flags[minus] & 64)) return WV_ERR_INVALID;
        comps[2 * (s + 1)] = wv_tap{plus, 1.0};
        comps[2 * (s + 1) + 1] = wv_tap{minus, -1.0};
        comp_len[s + 1] = 2;
    }
    *n_comps = 4;
    return WV_OK;
}

namespace {
// Complex quotient in the textbook form (a c + b d, b c - a d) / |z|^2.
inline std::complex<double> cdiv(std::complex<double> a, std::complex<double> b) {
    const double n = b.real() * b.real() + b.imag() * b.imag();
    return {(a.real() * b.real() + a.imag() * b.imag()) / n, (a.imag() * b.real() - a.real() * b.imag()) / n};
}
inline std::complex<double> cmul(std::complex<double> a, std::complex<double> b) {
    return {a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real()};
}
}  // namespace

// fdtd.rs:61-78 (WaveAccuracy::headroom, cabinet_cells): Draft, Standard, High.
extern "C" double wv_accuracy_headroom(int accuracy) { return accuracy <= 0 ? 1.0 : accuracy == 1 ? 1.5 : 2.0; }
extern "C" double wv_accuracy_cabinet_cells(int accuracy) { return accuracy <= 0 ? 4.0 : accuracy == 1 ? 6.0 : 8.0; }

// fdtd.rs:144-146: the top of the one-octave crossover band.
extern "C" double wv_settings_f_max(double crossover_hz) { return crossover_hz * std::sqrt(2.0); }

// fdtd.rs:166-175: 1 below the crossover band, 0 above, a raised cosine over
// the octave centred on the crossover.
extern "C" double wv_crossover_weight(double crossover_hz, double f) {
    const double x = std::log2(f / crossover_hz);
    if (x <= -0.5) return 1.0;
    if (x >= 0.5) return 0.0;
    return 0.5 * (1.0 + std::cos(kPi * (x + 0.5)));
}

// fdtd.rs:148-154: the highest frequency the wave solver contributes to,
// times the accuracy's headroom, up to the top of the crossover band.
extern "C" double wv_grid_fmax(double crossover_hz, int accuracy, const double* freqs, size_t nf) {
    double top = 0.0;
    for (size_t i = 0; i < nf; ++i)
        if (wv_crossover_weight(crossover_hz, freqs[i]) > 0.0) top = std::fmax(top, freqs[i]);
    const double fm = wv_settings_f_max(crossover_hz);
    return top > 0.0 ? std::fmin(top * wv_accuracy_headroom(accuracy), fm) : fm;
}

// fdtd.rs:157-162: the grid-axis phase error of the scheme, degrees per metre.
extern "C" double wv_phase_error_deg_per_m(double f, double c, double dx) {
    const double dt = kCourant * dx / c;
    const double w = 2.0 * kPi * f;
    const double k_num = 2.0 / dx * std::asin(std::sin(w * dt / 2.0) / kCourant);
    return std::fabs(k_num - w / c) * (180.0 / kPi);
}

extern "C" void wv_first_order_directions(double out[1200]) {
    // fdtd.rs:753-760: 400 directions on a golden-angle spiral, even in z.
    const double golden = kPi * (3.0 - std::sqrt(5.0));
    for (int i = 0; i < 400; ++i) {
        const double z = 1.0 - 2.0 * ((double)i + 0.5) / 400.0;
        const double r = std::sqrt(1.0 - z * z);
        out[3 * i] = r * std::cos(golden * (double)i);
        out[3 * i + 1] = r * std::sin(golden * (double)i);
        out[3 * i + 2] = z;
    }
}

namespace {
// f64::total_cmp's order.
int64_t total_key(double v) {
    int64_t x;
    std::memcpy(&x, &v, 8);
    return x ^ (int64_t)(((uint64_t)(x >> 63)) >> 1);
}

// fdtd.rs:771-788: Gauss-Jordan on the 4x4 normal matrix with partial
// pivoting (the last of equal pivots), complex right-hand side.
std::array<std::complex<double>, 4> solve4(const double (&m0)[4][4], const std::array<std::complex<double>, 4>& rhs) {
    double m[4][4];
    std::memcpy(m, m0, sizeof(m));
    std::array<std::complex<double>, 4> x = rhs;
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (total_key(std::fabs(m[r][col])) >= total_key(std::fabs(m[piv][col]))) piv = r;
        for (int c = 0; c < 4; ++c) std::swap(m[col][c], m[piv][c]);
        std::swap(x[col], x[piv]);
        for (int row = 0; row < 4; ++row) {
            if (row == col) continue;
            const double k = m[row][col] / m[col][col];
            for (int c = 0; c < 4; ++c) m[row][c] -= k * m[col][c];
            x[row] = {x[row].real() - x[col].real() * k, x[row].imag() - x[col].imag() * k};
        }
    }
    for (int i = 0; i < 4; ++i) x[i] = {x[i].real() / m[i][i], x[i].imag() / m[i][i]};
    return x;
}
}  // namespace

extern "C" double wv_fit_first_order(const double* d, const double d_axis[2], const double to_room[9], double a0[2],
                                     double a[6]) {
    // fdtd.rs:749-807: least-squares D(u) ~ a0 + a.u over the spiral
    // directions (u in room coordinates), then moved along M^-1 e so that it
    // meets the on-axis value exactly.
    double dirs[1200];
    wv_first_order_directions(dirs);
    auto rot = [&](const double* v, double* o) {
        for (int r = 0; r < 3; ++r) o[r] = (to_room[3 * r] * v[0] + to_room[3 * r + 1] * v[1]) + to_room[3 * r + 2] * v[2];
    };
    double m[4][4] = {};
    std::array<std::complex<double>, 4> rhs{};
    std::vector<std::array<double, 4>> bs(400);
    for (int i = 0; i < 400; ++i) {
        double room[3];
        rot(dirs + 3 * i, room);
        const std::array<double, 4> b{1.0, room[0], room[1], room[2]};
        const std::complex<double> di(d[2 * i], d[2 * i + 1]);
        for (int p = 0; p < 4; ++p) {
            for (int c = 0; c < 4; ++c) m[p][c] += b[p] * b[c];
            rhs[p] = {rhs[p].real() + di.real() * b[p], rhs[p].imag() + di.imag() * b[p]};
        }
        bs[i] = b;
    }
    const auto ls = solve4(m, rhs);
    const double xax[3] = {1.0, 0.0, 0.0};
    double axis[3];
    rot(xax, axis);
    const double e[4] = {1.0, axis[0], axis[1], axis[2]};
    std::complex<double> got(0.0, 0.0);
    for (int i = 0; i < 4; ++i) got = {got.real() + ls[i].real() * e[i], got.imag() + ls[i].imag() * e[i]};
    const auto me = solve4(m, {std::complex<double>(e[0], 0.0), std::complex<double>(e[1], 0.0),
                               std::complex<double>(e[2], 0.0), std::complex<double>(e[3], 0.0)});
    double ete = -0.0;
    for (int i = 0; i < 4; ++i) ete += me[i].real() * e[i];
    const std::complex<double> q((d_axis[0] - got.real()) / ete, (d_axis[1] - got.imag()) / ete);
    std::complex<double> c[4];
    for (int i = 0; i < 4; ++i) {
        const std::complex<double> t = cmul(me[i], q);
        c[i] = {ls[i].real() + t.real(), ls[i].imag() + t.imag()};
    }
    double err = 0.0, tot = 0.0;
    for (int i = 0; i < 400; ++i) {
        const auto& b = bs[i];
        double fr = c[0].real(), fi = c[0].imag();
        for (int p = 1; p < 4; ++p) {
            fr = fr + c[p].real() * b[p];
            fi = fi + c[p].imag() * b[p];
        }
        const double er = fr - d[2 * i], ei = fi - d[2 * i + 1];
        err += er * er + ei * ei;
        tot += d[2 * i] * d[2 * i] + d[2 * i + 1] * d[2 * i + 1];
    }
    a0[0] = c[0].real();
    a0[1] = c[0].imag();
    for (int i = 0; i < 3; ++i) {
        a[2 * i] = c[i + 1].real();
        a[2 * i + 1] = c[i + 1].imag();
    }
    return std::sqrt(err / std::fmax(tot, 1e-30));
}

extern "C" void wv_component_weights(const double a0[2], const double a[6], double k, double dx, double out[8]) {
    // fdtd.rs:812-815: the +-dx dipole radiates 2j sin(k dx) cos(theta) times the monopole.
    const std::complex<double> dip(0.0, 2.0 * std::sin(k * dx));
    out[0] = a0[0];
    out[1] = a0[1];
    for (int i = 0; i < 3; ++i) {
        auto v = cdiv({a[2 * i], a[2 * i + 1]}, dip);
        out[2 + 2 * i] = v.real();
        out[3 + 2 * i] = v.imag();
    }
}

extern "C" void wv_to_transfer(const double* spectra, size_t n_spec, const double* signal, size_t steps,
                               double dt, double dx, const double* freqs, size_t nf, double c, double r_ref,
                               const double* h0, double* out) {
    // fdtd.rs:819-840: divide by the source spectrum, scale to pressure.
    std::vector<std::complex<double>> s(nf);
    for (size_t fi = 0; fi < nf; ++fi) {
        std::complex<double> acc(0.0, 0.0);
        for (size_t n = 0; n < steps; ++n) {
            const double x = freqs[fi] * dt * (double)n;
            const double ph = -2.0 * kPi * (x - std::trunc(x));
            acc += std::complex<double>(signal[n] * std::cos(ph), signal[n] * std::sin(ph));
        }
        s[fi] = acc;
    }
    const double scale = dt * dt / (dx * dx * dx) * 4.0 * kPi * c * c * r_ref;
    for (size_t i = 0; i < n_spec; ++i) {
        const size_t fi = i % nf;
        auto q = cdiv({spectra[2 * i], spectra[2 * i + 1]}, s[fi]);
        auto t = cmul({q.real() * scale, q.imag() * scale}, {h0[2 * fi], h0[2 * fi + 1]});
        out[2 * i] = t.real();
        out[2 * i + 1] = t.imag();
    }
}

// ------------------------------------------------------------ materials

extern "C" void wv_admittance_eval(const wv_admittance* fit, double f, double out[2]) {
    // impedance_fit.rs:48-55: beta0 + sum 1/(r + j w m + k/(j w)).
    const double w = 2.0 * kPi * std::max(f, 1e-6);
    std::complex<double> y(fit->beta0, 0.0);
    for (uint32_t b = 0; b < fit->n_branches; ++b) {
        const wv_branch& br = fit->branches[b];
        std::complex<double> den(br.r, w * br.m);
        den += cdiv({br.k, 0.0}, {0.0, w});
        y += cdiv({1.0, 0.0}, den);
    }
    out[0] = y.real();
    out[1] = y.imag();
}

extern "C" double wv_paris(double beta_re, double beta_im, uint32_t n) {
    // impedance_fit.rs:73-84: alpha_st = int (1 - |R|^2) sin 2t dt, midpoint rule.
    const double h = kPi / 2.0 / (double)n;
    double a = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        const double th = ((double)i + 0.5) * h;
        const double c = std::cos(th);
        auto r = cdiv({c - beta_re, -beta_im}, {c + beta_re, beta_im});
        a += (1.0 - std::norm(r)) * std::sin(2.0 * th) * h;
    }
    return a;
}

namespace {
double alpha_of_real_impedance(double z) {
    // material.rs:81-91 with R real.
    const int N = 256;
    const double h = (kPi / 2.0) / N;
    double s = 0.0;
    for (int i = 0; i < N; ++i) {
        const double th = ((double)i + 0.5) * h;
        const double c = std::cos(th);
        const double r = (z * c - 1.0) / (z * c + 1.0);
        s += (1.0 - r * r) * std::sin(2.0 * th) * h;
    }
    return s;
}
}  // namespace

extern "C" double wv_fit_real_impedance(double alpha) {
    // material.rs:96-111: bisection in log z on the high-impedance branch.
    double lo = 1.567, hi = 1e7;
    if (alpha >= alpha_of_real_impedance(lo)) return lo;
    if (alpha <= 0.0) return hi;
    for (int i = 0; i < 80; ++i) {
        const double mid = std::sqrt(lo * hi);
        if (alpha_of_real_impedance(mid) > alpha) lo = mid;
        else hi = mid;
    }
    return std::sqrt(lo * hi);
}

extern "C" const char* wv_version(void) { return "wavey-fdtd-mi355x 0.1 (gfx950)"; }
