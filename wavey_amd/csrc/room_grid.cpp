// Grid builder for general rooms: FdtdGrid::build_dx (fdtd.rs:289-377) with
// column_fill (fdtd.rs:385-422) for meshes, the plan test of extruded rooms
// (core/wavey-geom/src/room.rs:151-157) and rigid convex solids
// (fdtd.rs:424-461).
//
// Every ray this builder casts runs along a grid axis: one upward ray per
// (x, y) column for the inside test, and rays of at most 1.5 cells from a
// boundary node toward each missing neighbour for the surface behind it. So
// instead of a BVH the triangles are binned once on the cell grid: a 2D table
// over columns for the vertical rays, and per axis a 3D table of 8-cell bins
// for the face rays, each holding only triangles that a ray along that axis
// can hit. A triangle goes into every bin its bounding box, widened by one
// cell, touches (the 3D tables also drop bins its plane misses), so a ray
// meets every triangle that can report a hit on it. The hit test is the
// Moller-Trumbore test with the reference's tolerances (bvh.rs:209-226),
// so which triangles a ray hits, and at which t, follows from the ray alone.
// Where two triangles are hit at exactly the same t, the lower triangle index
// comes first (the reference's BVH takes them in tree order).
//
// Built with -ffp-contract=off: the geometry and slot arithmetic rounds as
// the reference's (glam's DVec3 products and sums, left to right).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "grid_internal.h"
#include "wavey_fdtd.h"
#include "wv_internal.h"

using namespace wv;

namespace {

struct V3 {
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double comp(V3 a, int k) { return k == 0 ? a.x : k == 1 ? a.y : a.z; }

// A triangle as v0 and its two edges from v0 (bvh.rs:55-58).
struct Tri {
    V3 v0, e1, e2;
};

// bvh.rs:209-226: two-sided Moller-Trumbore; t of the hit.
inline bool ray_hit(const Tri& T, V3 o, V3 d, double& t) {
    const V3 p = cross(d, T.e2);
    const double det = dot(T.e1, p);
    if (std::fabs(det) < 1e-14) return false;
    const double inv = 1.0 / det;
    const V3 s = sub(o, T.v0);
    const double u = dot(s, p) * inv;
    if (!(u >= -1e-10 && u <= 1.0 + 1e-10)) return false;
    const V3 q = cross(s, T.e1);
    const double v = dot(d, q) * inv;
    if (v < -1e-10 || u + v > 1.0 + 1e-10) return false;
    t = dot(T.e2, q) * inv;
    return true;
}

V3 axis_dir(int k, double sign) {
    V3 d{0.0, 0.0, 0.0};
    if (k == 0) d.x = sign;
    else if (k == 1) d.y = sign;
    else d.z = sign;
    return d;
}

// regions.rs:237-249
bool point_in_polygon(double px, double py, const double* xy, uint32_t n) {
    bool inside = false;
    uint32_t j = n - 1;
    for (uint32_t i = 0; i < n; ++i) {
        const double ax = xy[2 * i], ay = xy[2 * i + 1], bx = xy[2 * j], by = xy[2 * j + 1];
        if ((ay > py) != (by > py) && px < (bx - ax) * (py - ay) / (by - ay) + ax) inside = !inside;
        j = i;
    }
    return inside;
}

// Region::contains over the regions of an extruded room (room.rs:151-154,
// regions.rs:44-47), for the plan point (px, py).
bool in_plan(const wv_room_mesh* room, double px, double py) {
    for (uint32_t r = 0; r < room->n_regions; ++r) {
        const uint32_t c0 = room->region_cycles[r], c1 = room->region_cycles[r + 1];
        const uint32_t* off = room->cycle_off;
        if (!point_in_polygon(px, py, room->cycle_xy + 2 * (size_t)off[c0], off[c0 + 1] - off[c0])) continue;
        bool hole = false;
        for (uint32_t h = c0 + 1; h < c1 && !hole; ++h)
            hole = point_in_polygon(px, py, room->cycle_xy + 2 * (size_t)off[h], off[h + 1] - off[h]);
        if (!hole) return true;
    }
    return false;
}

// fdtd.rs:458-460
bool solid_contains(const wv_solid& s, V3 p) {
    if (!(p.x >= s.lo[0] && p.y >= s.lo[1] && p.z >= s.lo[2])) return false;
    if (!(p.x <= s.hi[0] && p.y <= s.hi[1] && p.z <= s.hi[2])) return false;
    for (uint32_t j = 0; j < s.n_planes; ++j) {
        const double* q = s.planes + 4 * j;
        if (!(dot(V3{q[0], q[1], q[2]}, p) <= q[3])) return false;
    }
    return true;
}

// Triangles binned on the cell grid: bin b of an axis covers cells
// [b*B, (b+1)*B). CSR: bin -> triangle indices in increasing order.
struct Bins {
    int64_t nb[3] = {1, 1, 1};
    int64_t B = 8;
    std::vector<uint32_t> start, items;
    int64_t bin(int64_t bx, int64_t by, int64_t bz) const { return (bz * nb[1] + by) * nb[0] + bx; }
};

struct Geometry {
    std::vector<Tri> tris;
    std::vector<V3> lo, hi;  // per-triangle bounding boxes
    double origin[3], dx;
    int64_t n[3];

    // Cell range [c0, c1] whose extents touch [a, b] along axis k, widened by
    // one cell on each side.
    void cell_range(int k, double a, double b, int64_t& c0, int64_t& c1) const {
        const double fa = std::floor((a - origin[k]) / dx + 0.5) - 1.0;
        const double fb = std::floor((b - origin[k]) / dx + 0.5) + 1.0;
        c0 = (int64_t)std::max(0.0, std::min(fa, (double)(n[k] - 1)));
        c1 = (int64_t)std::max(0.0, std::min(fb, (double)(n[k] - 1)));
    }
};

// The 2D column table: triangles a vertical ray can hit, by (x, y) bin.
Bins column_bins(const Geometry& G, int64_t B) {
    Bins t;
    t.B = B;
    t.nb[0] = (G.n[0] + B - 1) / B;
    t.nb[1] = (G.n[1] + B - 1) / B;
    t.nb[2] = 1;
    const V3 up{0.0, 0.0, 1.0};
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (size_t k = 0; k < G.tris.size(); ++k) {
        if (std::fabs(dot(G.tris[k].e1, cross(up, G.tris[k].e2))) < 1e-14) continue;  // never hit
        int64_t x0, x1, y0, y1;
        G.cell_range(0, G.lo[k].x, G.hi[k].x, x0, x1);
        G.cell_range(1, G.lo[k].y, G.hi[k].y, y0, y1);
        for (int64_t by = y0 / B; by <= y1 / B; ++by)
            for (int64_t bx = x0 / B; bx <= x1 / B; ++bx) pairs.emplace_back((uint32_t)t.bin(bx, by, 0), (uint32_t)k);
    }
    const int64_t nbins = t.nb[0] * t.nb[1];
    t.start.assign(nbins + 1, 0);
    for (auto& p : pairs) t.start[p.first + 1]++;
    for (int64_t b = 0; b < nbins; ++b) t.start[b + 1] += t.start[b];
    t.items.resize(pairs.size());
    std::vector<uint32_t> fill(t.start.begin(), t.start.end() - 1);
    for (auto& p : pairs) t.items[fill[p.first]++] = p.second;  // pairs are in triangle order
    return t;
}

// The 3D table of axis k: triangles a ray along k can hit, by 8-cell bin,
// keeping only bins the triangle's plane passes through.
Bins face_bins(const Geometry& G, int k, int64_t B) {
    Bins t;
    t.B = B;
    for (int a = 0; a < 3; ++a) t.nb[a] = (G.n[a] + B - 1) / B;
    const V3 d = axis_dir(k, 1.0);
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    const double half = 0.5 * (double)B * G.dx + G.dx;  // half a bin, plus a cell
    for (size_t j = 0; j < G.tris.size(); ++j) {
        const Tri& T = G.tris[j];
        if (std::fabs(dot(T.e1, cross(d, T.e2))) < 1e-14) continue;  // parallel to the axis: never hit
        int64_t c0[3], c1[3];
        G.cell_range(0, G.lo[j].x, G.hi[j].x, c0[0], c1[0]);
        G.cell_range(1, G.lo[j].y, G.hi[j].y, c0[1], c1[1]);
        G.cell_range(2, G.lo[j].z, G.hi[j].z, c0[2], c1[2]);
        const V3 N = cross(T.e1, T.e2);
        const double reach = (std::fabs(N.x) + std::fabs(N.y) + std::fabs(N.z)) * half;
        for (int64_t bz = c0[2] / B; bz <= c1[2] / B; ++bz)
            for (int64_t by = c0[1] / B; by <= c1[1] / B; ++by)
                for (int64_t bx = c0[0] / B; bx <= c1[0] / B; ++bx) {
                    // Bin centre: cells b*B .. b*B+B-1, centres origin + i dx.
                    const V3 c{G.origin[0] + ((double)(bx * B) + 0.5 * (double)(B - 1)) * G.dx,
                               G.origin[1] + ((double)(by * B) + 0.5 * (double)(B - 1)) * G.dx,
                               G.origin[2] + ((double)(bz * B) + 0.5 * (double)(B - 1)) * G.dx};
                    if (std::fabs(dot(N, sub(c, T.v0))) > reach * (1.0 + 1e-9)) continue;
                    pairs.emplace_back((uint32_t)t.bin(bx, by, bz), (uint32_t)j);
                }
    }
    const int64_t nbins = t.nb[0] * t.nb[1] * t.nb[2];
    t.start.assign(nbins + 1, 0);
    for (auto& p : pairs) t.start[p.first + 1]++;
    for (int64_t b = 0; b < nbins; ++b) t.start[b + 1] += t.start[b];
    t.items.resize(pairs.size());
    std::vector<uint32_t> fill(t.start.begin(), t.start.end() - 1);
    for (auto& p : pairs) t.items[fill[p.first]++] = p.second;
    return t;
}

struct Key {
    std::array<int32_t, 6> f;
    int32_t n;
    bool operator==(const Key& o) const { return n == o.n && f == o.f; }
};
struct KeyHash {
    size_t operator()(const Key& k) const {
        uint64_t h = 1469598103934665603ull ^ (uint64_t)k.n;
        for (int i = 0; i < k.n; ++i) h = (h ^ (uint64_t)(uint32_t)k.f[i]) * 1099511628211ull;
        return (size_t)h;
    }
};

bool finite_fit(const wv_admittance& a) {
    if (!std::isfinite(a.beta0)) return false;
    for (uint32_t b = 0; b < a.n_branches; ++b)
        if (!std::isfinite(a.branches[b].r) || !std::isfinite(a.branches[b].m) || !std::isfinite(a.branches[b].k))
            return false;
    return true;
}

wv_status check_room(const wv_room_mesh* room, const wv_solid* solids, size_t n_solids, std::string& why) {
    if (!room->positions || room->n_positions == 0) { why = "room needs positions"; return WV_ERR_INVALID; }
    if (room->n_triangles && (!room->triangles || !room->tri_surface)) {
        why = "room needs triangles and their surfaces";
        return WV_ERR_INVALID;
    }
    if (!room->fits && room->n_fits) { why = "fits missing"; return WV_ERR_INVALID; }
    for (uint32_t k = 0; k < room->n_triangles; ++k) {
        for (int j = 0; j < 3; ++j)
            if (room->triangles[3 * k + j] >= room->n_positions) {
                why = "triangle " + std::to_string(k) + " names a missing vertex";
                return WV_ERR_INVALID;
            }
        if (room->tri_surface[k] >= room->n_fits) {
            why = "triangle " + std::to_string(k) + " has a surface without a fit";
            return WV_ERR_INVALID;
        }
    }
    for (uint32_t i = 0; i < room->n_fits; ++i)
        if (!finite_fit(room->fits[i]) || (room->fits[i].n_branches && !room->fits[i].branches)) {
            why = "fit " + std::to_string(i) + " is not finite";
            return WV_ERR_INVALID;
        }
    for (uint32_t i = 0; i < 3 * room->n_positions; ++i)
        if (!std::isfinite(room->positions[i])) { why = "positions must be finite"; return WV_ERR_INVALID; }
    if (room->prism) {
        if (room->n_regions && (!room->region_cycles || !room->cycle_off || !room->cycle_xy)) {
            why = "prism room needs its plan regions";
            return WV_ERR_INVALID;
        }
        for (uint32_t r = 0; r < room->n_regions; ++r)
            if (room->region_cycles[r + 1] <= room->region_cycles[r]) {
                why = "every region needs an outer cycle";
                return WV_ERR_INVALID;
            }
        const uint32_t nc = room->n_regions ? room->region_cycles[room->n_regions] : 0;
        for (uint32_t c = 0; c < nc; ++c)
            if (room->cycle_off[c + 1] < room->cycle_off[c] + 1) {
                why = "every cycle needs a point";
                return WV_ERR_INVALID;
            }
    }
    if (n_solids && !solids) { why = "solids missing"; return WV_ERR_INVALID; }
    for (size_t s = 0; s < n_solids; ++s)
        if (solids[s].n_planes && !solids[s].planes) { why = "solid planes missing"; return WV_ERR_INVALID; }
    return WV_OK;
}

}  // namespace

extern "C" wv_status wv_solid_from_mesh(const double* verts, uint32_t n_verts, const uint32_t* tris,
                                        uint32_t n_tris, const double rot[9], const double pos[3],
                                        double* planes_out, uint32_t* n_planes, double lo[3], double hi[3]) {
    if (!verts || !tris || !rot || !pos || !planes_out || !n_planes || !lo || !hi || n_tris == 0)
        return WV_ERR_INVALID;
    for (uint32_t k = 0; k < 3 * n_tris; ++k)
        if (tris[k] >= n_verts) return WV_ERR_INVALID;
    // w = rot * v + pos, glam's DMat3 * DVec3: columns scaled and summed in order.
    std::vector<V3> w(n_verts);
    for (uint32_t i = 0; i < n_verts; ++i) {
        const double* v = verts + 3 * i;
        double o[3];
        for (int r = 0; r < 3; ++r) o[r] = (rot[3 * r] * v[0] + rot[3 * r + 1] * v[1] + rot[3 * r + 2] * v[2]) + pos[r];
        w[i] = {o[0], o[1], o[2]};
    }
    uint32_t np = 0;
    for (uint32_t k = 0; k < n_tris; ++k) {
        const V3 a = w[tris[3 * k]], b = w[tris[3 * k + 1]], c = w[tris[3 * k + 2]];
        V3 n = cross(sub(b, a), sub(c, a));
        if (std::sqrt(dot(n, n)) < 1e-12) continue;
        const double inv = 1.0 / std::sqrt(dot(n, n));
        n = {n.x * inv, n.y * inv, n.z * inv};
        const double d = dot(n, a);
        bool dup = false;
        for (uint32_t j = 0; j < np && !dup; ++j) {
            const double* m = planes_out + 4 * j;
            dup = dot(V3{m[0], m[1], m[2]}, n) > 1.0 - 1e-9 && std::fabs(m[3] - d) < 1e-9;
        }
        if (dup) continue;
        planes_out[4 * np] = n.x;
        planes_out[4 * np + 1] = n.y;
        planes_out[4 * np + 2] = n.z;
        planes_out[4 * np + 3] = d;
        ++np;
    }
    for (int r = 0; r < 3; ++r) {
        lo[r] = INFINITY;
        hi[r] = -INFINITY;
    }
    for (const V3& p : w)
        for (int r = 0; r < 3; ++r) {
            lo[r] = std::fmin(lo[r], comp(p, r));
            hi[r] = std::fmax(hi[r], comp(p, r));
        }
    *n_planes = np;
    return WV_OK;
}

extern "C" double wv_grid_spacing(double c, double f_max, double ppw, const wv_solid* solids, size_t n_solids) {
    double cab = INFINITY;
    for (size_t s = 0; s < n_solids; ++s) {
        double m = std::fmin(std::fmin(solids[s].hi[0] - solids[s].lo[0], solids[s].hi[1] - solids[s].lo[1]),
                             solids[s].hi[2] - solids[s].lo[2]);
        cab = std::fmin(cab, m / 8.0);
    }
    return std::fmin(c / (f_max * ppw), cab);
}

static wv_status grid_build_room_impl(const wv_room_mesh* room, const wv_solid* solids, size_t n_solids,
                                      double c, double dx, int threads, wv_built_grid** out, char* err,
                                      size_t err_len) {
    if (!room || !out || !(dx > 0.0) || !(c > 0.0)) {
        set_err(err, err_len, "wv_grid_build_room: invalid argument");
        return WV_ERR_INVALID;
    }
    {
        std::string why;
        wv_status st = check_room(room, solids, n_solids, why);
        if (st != WV_OK) {
            set_err(err, err_len, "wv_grid_build_room: " + why);
            return st;
        }
    }
    threads = pick_threads(threads);
    const double dt = kCourant * dx / c;  // fdtd.rs:290
    // Room::bounds (room.rs:186-191) and the padded grid (fdtd.rs:291-299).
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < room->n_positions; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], room->positions[3 * i + k]);
            hi[k] = std::fmax(hi[k], room->positions[3 * i + k]);
        }
    Geometry G;
    G.dx = dx;
    for (int k = 0; k < 3; ++k) {
        G.origin[k] = lo[k] - dx * 0.5;
        G.n[k] = (int64_t)std::ceil((hi[k] - lo[k]) / dx) + 2;
    }
    const int64_t nx = G.n[0], ny = G.n[1], nz = G.n[2];
    const uint64_t cells = (uint64_t)nx * ny * nz;
    if (cells > kMaxCells) {
        set_err(err, err_len, "the wave solver grid would have " + std::to_string(cells / 1000000) +
                                  " M cells; lower the crossover frequency");
        return WV_ERR_LIMIT;
    }
    G.tris.resize(room->n_triangles);
    G.lo.resize(room->n_triangles);
    G.hi.resize(room->n_triangles);
    for (uint32_t k = 0; k < room->n_triangles; ++k) {
        V3 v[3];
        for (int j = 0; j < 3; ++j) {
            const double* p = room->positions + 3 * (size_t)room->triangles[3 * k + j];
            v[j] = {p[0], p[1], p[2]};
        }
        G.tris[k] = {v[0], sub(v[1], v[0]), sub(v[2], v[0])};
        G.lo[k] = {std::min({v[0].x, v[1].x, v[2].x}), std::min({v[0].y, v[1].y, v[2].y}),
                   std::min({v[0].z, v[1].z, v[2].z})};
        G.hi[k] = {std::max({v[0].x, v[1].x, v[2].x}), std::max({v[0].y, v[1].y, v[2].y}),
                   std::max({v[0].z, v[1].z, v[2].z})};
    }
    auto center = [&](int64_t x, int64_t y, int64_t z) {
        return V3{G.origin[0] + (double)x * dx, G.origin[1] + (double)y * dx, G.origin[2] + (double)z * dx};
    };

    // Cell classes: bit 0 inside the room and off the padding, bit 1 solid.
    std::vector<uint8_t> cls(cells, 0);
    if (room->prism) {
        // Room::contains per cell; the plan test depends on (x, y) only.
        parallel_chunks(ny, threads, [&](int64_t y0, int64_t y1, int) {
            for (int64_t y = std::max<int64_t>(y0, 1); y < std::min(y1, ny - 1); ++y)
                for (int64_t x = 1; x < nx - 1; ++x) {
                    const V3 p0 = center(x, y, 0);
                    if (!in_plan(room, p0.x, p0.y)) continue;
                    for (int64_t z = 1; z < nz - 1; ++z) {
                        const double pz = center(x, y, z).z;
                        if (pz > 0.0 && pz < room->height) cls[(z * ny + y) * nx + x] = 1;
                    }
                }
        });
    } else {
        // column_fill (fdtd.rs:385-422).
        const int64_t CB = 16;
        const Bins colb = column_bins(G, CB);
        const double offx = 0.0001372 * dx, offy = 0.0002918 * dx;
        const double start = G.origin[2] - dx;
        const V3 up{0.0, 0.0, 1.0};
        parallel_chunks(ny, threads, [&](int64_t y0, int64_t y1, int) {
            struct Cross {
                double t;
                uint32_t k;
                bool up;
            };
            std::vector<Cross> hits, crossings;
            for (int64_t y = y0; y < y1; ++y)
                for (int64_t x = 1; x < nx - 1; ++x) {
                    if (y == 0 || y == ny - 1) break;  // padding columns hold no air
                    const V3 c0 = center(x, y, 0);
                    const V3 o{c0.x + offx, c0.y + offy, start + 0.0};
                    const int64_t b = colb.bin(x / CB, y / CB, 0);
                    hits.clear();
                    for (uint32_t s = colb.start[b]; s < colb.start[b + 1]; ++s) {
                        const uint32_t k = colb.items[s];
                        double t;
                        if (ray_hit(G.tris[k], o, up, t) && t > 0.0 && t < INFINITY) {
                            const V3 nrm = cross(G.tris[k].e1, G.tris[k].e2);
                            hits.push_back({t, k, nrm.z > 0.0});
                        }
                    }
                    std::sort(hits.begin(), hits.end(),
                              [](const Cross& a, const Cross& b) { return a.t < b.t || (a.t == b.t && a.k < b.k); });
                    crossings.clear();
                    for (const Cross& h : hits) {
                        if (!crossings.empty() && h.t - crossings.back().t < 1e-9 && crossings.back().up == h.up) continue;
                        crossings.push_back(h);
                    }
                    size_t k = 0;
                    for (int64_t z = 0; z < nz; ++z) {
                        const double tz = center(x, y, z).z - start;
                        while (k < crossings.size() && crossings[k].t < tz) ++k;
                        if (k % 2 == 1 && z > 0 && z < nz - 1) cls[(z * ny + y) * nx + x] = 1;
                    }
                }
        });
    }
    // Solids (fdtd.rs:305-316): cells whose centre a solid contains, looked
    // for inside each solid's box widened by a cell.
    for (size_t s = 0; s < n_solids; ++s) {
        const wv_solid& S = solids[s];
        if (!(S.lo[0] <= S.hi[0] && S.lo[1] <= S.hi[1] && S.lo[2] <= S.hi[2])) continue;  // contains nothing
        int64_t r0[3], r1[3];
        for (int k = 0; k < 3; ++k) G.cell_range(k, S.lo[k], S.hi[k], r0[k], r1[k]);
        parallel_chunks(r1[2] - r0[2] + 1, threads, [&](int64_t a, int64_t b, int) {
            for (int64_t z = r0[2] + a; z < r0[2] + b; ++z)
                for (int64_t y = r0[1]; y <= r1[1]; ++y)
                    for (int64_t x = r0[0]; x <= r1[0]; ++x)
                        if (solid_contains(S, center(x, y, z))) cls[(z * ny + y) * nx + x] |= 2;
        });
    }

    // Canonical fit of each surface: the first surface with an equal fit.
    const std::vector<int32_t> canon = canonical_fits(room->fits, room->n_fits);

    // Face rays need the 3D tables only when there are triangles.
    const int64_t FB = 8;
    Bins fb[3];
    if (room->n_triangles)
        for (int k = 0; k < 3; ++k) fb[k] = face_bins(G, k, FB);

    auto g = std::make_unique<wv_built_grid>();
    g->flags.assign(cells, 0);
    static const int DX[6] = {-1, 1, 0, 0, 0, 0}, DY[6] = {0, 0, -1, 1, 0, 0}, DZ[6] = {0, 0, 0, 0, -1, 1};
    const double tmax = dx * 1.5;
    using SlotMap = std::unordered_map<Key, std::vector<wv_slot>, KeyHash>;
    std::vector<SlotMap> caches(threads);
    std::vector<std::vector<std::pair<uint32_t, const std::vector<wv_slot>*>>> part(threads);
    parallel_chunks(nz, threads, [&](int64_t z0, int64_t z1, int th) {
        SlotMap& cache = caches[th];
        auto& mine = part[th];
        std::vector<wv_slot> scratch;
        for (int64_t z = std::max<int64_t>(z0, 1); z < std::min(z1, nz - 1); ++z)
            for (int64_t y = 1; y < ny - 1; ++y) {
                const int64_t row = (z * ny + y) * nx;
                for (int64_t x = 1; x < nx - 1; ++x) {
                    if (cls[row + x] != 1) continue;
                    const V3 p = center(x, y, z);
                    uint8_t f = 64;
                    Key key{};
                    for (int d = 0; d < 6; ++d) {
                        const int64_t j = row + x + (int64_t)DZ[d] * nx * ny + (int64_t)DY[d] * nx + DX[d];
                        if (cls[j] == 1) { f |= (uint8_t)(1u << d); continue; }
                        if (cls[j] & 2) continue;  // cabinet face: rigid
                        // Bvh::intersect(p, dir, 0, 1.5 dx): closest hit along the axis.
                        const int ax = d >> 1;
                        const double sg = (d & 1) ? 1.0 : -1.0;
                        const V3 dir = axis_dir(ax, sg);
                        const int64_t ci[3] = {x, y, z};
                        int64_t b0[3], b1[3];
                        for (int a = 0; a < 3; ++a) {
                            int64_t lo_c = ci[a], hi_c = ci[a];
                            if (a == ax) {
                                lo_c = std::max<int64_t>(0, ci[a] - 2);
                                hi_c = std::min<int64_t>(G.n[a] - 1, ci[a] + 2);
                            }
                            b0[a] = lo_c / FB;
                            b1[a] = hi_c / FB;
                        }
                        const Bins& T = fb[ax];
                        double best_t = tmax;
                        int64_t best = -1;
                        for (int64_t bz = b0[2]; bz <= b1[2]; ++bz)
                            for (int64_t by = b0[1]; by <= b1[1]; ++by)
                                for (int64_t bx = b0[0]; bx <= b1[0]; ++bx) {
                                    if (T.start.empty()) continue;
                                    const int64_t b = T.bin(bx, by, bz);
                                    for (uint32_t s = T.start[b]; s < T.start[b + 1]; ++s) {
                                        const uint32_t k = T.items[s];
                                        double t;
                                        if (!ray_hit(G.tris[k], p, dir, t) || !(t > 0.0)) continue;
                                        if (t < best_t || (t == best_t && best >= 0 && (int64_t)k < best)) {
                                            best_t = t;
                                            best = k;
                                        }
                                    }
                                }
                        if (best >= 0) key.f[key.n++] = canon[room->tri_surface[best]];
                    }
                    g->flags[row + x] = f;
                    if (key.n == 0) continue;
                    auto it = cache.find(key);
                    if (it == cache.end()) {
                        node_slots(key.f.data(), key.n, room->fits, dt, scratch);
                        it = cache.emplace(key, scratch).first;
                    }
                    if (!it->second.empty()) mine.emplace_back((uint32_t)(row + x), &it->second);
                }
            }
    });
    size_t nb_total = 0, ns_total = 0;
    for (auto& p : part) {
        nb_total += p.size();
        for (auto& e : p) ns_total += e.second->size();
    }
    g->b_cell.reserve(nb_total);
    g->b_start.reserve(nb_total + 1);
    g->slots.reserve(ns_total);
    for (auto& p : part)
        for (auto& [cell, sl] : p) {
            g->b_cell.push_back(cell);
            g->b_start.push_back((uint32_t)g->slots.size());
            g->slots.insert(g->slots.end(), sl->begin(), sl->end());
        }
    g->b_start.push_back((uint32_t)g->slots.size());
    finish_view(*g, G.n, G.origin, dx, dt);
    *out = g.release();
    return WV_OK;
}

extern "C" wv_status wv_grid_build_room(const wv_room_mesh* room, const wv_solid* solids, size_t n_solids,
                                        double c, double dx, int threads, wv_built_grid** out, char* err,
                                        size_t err_len) {
    return host_guarded(err, err_len, [&]() {
        return grid_build_room_impl(room, solids, n_solids, c, dx, threads, out, err, err_len);
    });
}
