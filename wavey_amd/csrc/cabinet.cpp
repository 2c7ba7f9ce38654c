// A cabinet's free-field calibration in the wave solver: calibrate_cabinet
// (core/wavey-solver/src/fdtd.rs:873-955). Two runs in a box with matched
// (rho c) walls, one with the cabinet as a rigid solid and one without, give
// the cabinet's on-axis gain over the bare driver and its horizontal
// directivity. Both grids come from the general-room builder; the two runs go
// to one batch (one launch per step for both) or one after another on their
// own plans, with the same results.
//
// Built on the library's public C ABI only, with -ffp-contract=off: the
// geometry and the ratio arithmetic round as the reference's.
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "wavey_fdtd.h"
#include "wv_internal.h"

using namespace wv;

namespace {

const double kPi = 3.141592653589793;

struct Cx {
    double re, im;
};
// num-complex 0.4 division: ((a c + b d) + (b c - a d) i) / |z|^2.
Cx cdiv(Cx a, Cx b) {
    const double n = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / n, (a.im * b.re - a.re * b.im) / n};
}

struct GridOwner {
    wv_built_grid* g = nullptr;
    ~GridOwner() { wv_grid_free(g); }
};

// Spectra [12][nf] of one run on its own plan.
wv_status run_plain(const wv_grid* g, const std::vector<wv_tap>& taps, const std::vector<double>& signal,
                    const std::vector<uint64_t>& rx, const double* freqs, size_t nf, int precision,
                    const volatile int32_t* cancel, std::vector<double>& spectra, char* err, size_t err_len) {
    wv_plan* plan = nullptr;
    wv_status st = wv_plan_create(g, precision, &plan, err, err_len);
    if (st != WV_OK) return st;
    spectra.assign(2 * rx.size() * nf, 0.0);
    st = wv_plan_run(plan, taps.data(), taps.size(), signal.data(), signal.size(), rx.data(), rx.size(), freqs, nf,
                     WV_MODE_SPECTRA, cancel, nullptr, nullptr, spectra.data(), nullptr, err, err_len);
    wv_plan_free(plan);
    return st;
}

}  // namespace

static wv_status calibrate_cabinet_impl(const double* verts, uint32_t n_verts, const uint32_t* tris,
                                        uint32_t n_tris, const double ac[3], double c, double f_max, double ppw,
                                        double cells_across, const double* freqs, size_t nf, int precision,
                                        int batched, const volatile int32_t* cancel, double* on_axis_out,
                                        double* angles_out, double* directivity_out, char* err, size_t err_len) {
    if (!verts || !n_verts || !tris || !n_tris || !ac || (nf && !freqs) || !on_axis_out || !angles_out ||
        !directivity_out || !(c > 0.0) || !(f_max > 0.0) || !(ppw > 0.0) || !(cells_across > 0.0)) {
        set_err(err, err_len, "wv_calibrate_cabinet: invalid argument");
        return WV_ERR_INVALID;
    }
    // The absorbing box around the cabinet (fdtd.rs:905-917): half a
    // wavelength at ka = 1/2, at least 1.5 m beyond the 3 m measuring circle,
    // at most 160 cells across.
    double reach = 0.1;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n_verts; ++i) {
        const double* v = verts + 3 * i;
        const double d[3] = {v[0] - ac[0], v[1] - ac[1], v[2] - ac[2]};
        reach = std::fmax(reach, std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], v[k]);
            hi[k] = std::fmax(hi[k], v[k]);
        }
    }
    const double f_cab = 0.5 * c / (2.0 * kPi * reach);
    const double ext = std::fmin(std::fmin(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    const double dx = std::fmin(c / (f_max * ppw), ext / cells_across);
    const double max_half = 0.5 * dx * 160.0;
    const double r_m = std::fmax(reach + 1.0, 3.0);
    const double half = std::fmax(std::fmin(0.5 * c / f_cab, max_half), r_m + 1.5);
    const double side = 2.0 * half;

    // A square plan extruded to `side` (fdtd.rs:914-916), every surface matched: beta0 = 1.
    double pos[8][3];
    for (int i = 0; i < 8; ++i) {
        pos[i][0] = (i & 1) ? side : 0.0;
        pos[i][1] = (i & 2) ? side : 0.0;
        pos[i][2] = (i & 4) ? side : 0.0;
    }
    static const uint32_t box_tris[12][3] = {{0, 2, 3}, {0, 3, 1}, {4, 5, 7}, {4, 7, 6}, {0, 1, 5}, {0, 5, 4},
                                             {2, 6, 7}, {2, 7, 3}, {0, 4, 6}, {0, 6, 2}, {1, 3, 7}, {1, 7, 5}};
    uint32_t tri_surface[12];
    for (int i = 0; i < 12; ++i) tri_surface[i] = (uint32_t)(i / 2);
    wv_admittance fits[6];
    for (auto& f : fits) f = {1.0, 0, nullptr, {0.0, 0.0}, 0.0, 0.0};
    const double plan_xy[8] = {0.0, 0.0, side, 0.0, side, side, 0.0, side};
    const uint32_t region_cycles[2] = {0, 1}, cycle_off[2] = {0, 4};
    wv_room_mesh room{1, side, 8, &pos[0][0], 12, &box_tris[0][0], tri_surface, 6, fits, 1,
                      region_cycles, cycle_off, plan_xy};

    // The cabinet, its acoustic centre at the box centre (fdtd.rs:918-919).
    const double p0[3] = {half, half, half};
    const double rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const double shift[3] = {p0[0] - ac[0], p0[1] - ac[1], p0[2] - ac[2]};
    std::vector<double> planes(4 * (size_t)n_tris);
    wv_solid cab{};
    uint32_t np = 0;
    if (wv_solid_from_mesh(verts, n_verts, tris, n_tris, rot, shift, planes.data(), &np, cab.lo, cab.hi) != WV_OK) {
        set_err(err, err_len, "wv_calibrate_cabinet: the cabinet mesh has no usable triangle");
        return WV_ERR_INVALID;
    }
    cab.n_planes = np;
    cab.planes = planes.data();
    GridOwner with, bare;
    wv_status st = wv_grid_build_room(&room, &cab, 1, c, dx, 0, &with.g, err, err_len);
    if (st != WV_OK) return st;
    st = wv_grid_build_room(&room, nullptr, 0, c, dx, 0, &bare.g, err, err_len);
    if (st != WV_OK) return st;
    const wv_grid* gw = wv_built_grid_view(with.g);
    const wv_grid* gb = wv_built_grid_view(bare.g);

    // The driver where the baffle meets the air, the receivers on a
    // horizontal circle of radius r_m every 30 degrees (fdtd.rs:922-932).
    double src[3];
    const double axis[3] = {1.0, 0.0, 0.0};
    wv_baffle_point(gw, p0, axis, src);
    wv_tap tw[8], tb[8];
    size_t ntw = 0, ntb = 0;
    if (wv_taps_at(gw, src, tw, &ntw) != WV_OK || wv_taps_at(gb, src, tb, &ntb) != WV_OK) {
        set_err(err, err_len, "wv_calibrate_cabinet: no air node around the driver");
        return WV_ERR_INVALID;
    }
    std::vector<uint64_t> rxw(12), rxb(12);
    for (int a = 0; a < 12; ++a) {
        angles_out[a] = (double)a * 30.0;
        const double r = angles_out[a] * (kPi / 180.0);
        const double pt[3] = {p0[0] + std::cos(r) * r_m, p0[1] + std::sin(r) * r_m, p0[2] + 0.0 * r_m};
        if (wv_nearest_air(gw, pt, &rxw[a]) != WV_OK || wv_nearest_air(gb, pt, &rxb[a]) != WV_OK) {
            set_err(err, err_len, "wv_calibrate_cabinet: a receiver has no air node nearby");
            return WV_ERR_INVALID;
        }
    }
    // Long enough to cross the box diagonal twice, plus the pulse (fdtd.rs:933-935).
    const double duration = side * std::sqrt(3.0) * 2.0 / c + 12.0 * 0.342 / f_max;
    const size_t steps = (size_t)std::ceil(duration / gw->dt);
    std::vector<double> signal(steps);
    wv_source_signal(f_max, gw->dt, steps, signal.data());
    const std::vector<wv_tap> taps_w(tw, tw + ntw), taps_b(tb, tb + ntb);

    std::vector<double> sw, sb;
    if (batched) {
        wv_batch* b = nullptr;
        st = wv_batch_create(precision, &b, err, err_len);
        if (st != WV_OK) return st;
        std::unique_ptr<wv_batch, void (*)(wv_batch*)> guard(b, wv_batch_free);
        int32_t iw = -1, ib = -1;
        st = wv_batch_add(b, gw, taps_w.data(), taps_w.size(), signal.data(), steps, rxw.data(), rxw.size(), &iw,
                          err, err_len);
        if (st == WV_OK)
            st = wv_batch_add(b, gb, taps_b.data(), taps_b.size(), signal.data(), steps, rxb.data(), rxb.size(),
                              &ib, err, err_len);
        if (st == WV_OK) st = wv_batch_run(b, 1, cancel, nullptr, nullptr, err, err_len);
        if (st != WV_OK) return st;
        sw.assign(2 * 12 * nf, 0.0);
        sb.assign(2 * 12 * nf, 0.0);
        st = wv_batch_result(b, iw, freqs, nf, WV_MODE_SPECTRA, sw.data(), nullptr, err, err_len);
        if (st == WV_OK) st = wv_batch_result(b, ib, freqs, nf, WV_MODE_SPECTRA, sb.data(), nullptr, err, err_len);
        if (st != WV_OK) return st;
    } else {
        st = run_plain(gw, taps_w, signal, rxw, freqs, nf, precision, cancel, sw, err, err_len);
        if (st != WV_OK) return st;
        st = run_plain(gb, taps_b, signal, rxb, freqs, nf, precision, cancel, sb, err, err_len);
        if (st != WV_OK) return st;
    }

    // Ratio with / bare per receiver and frequency, faded toward 1 below
    // f_cab as (f / f_cab)^2; directivity relative to on axis (fdtd.rs:940-953).
    for (size_t f = 0; f < nf; ++f) {
        const double x = freqs[f] / f_cab;
        const double w = std::fmin(x * x, 1.0);
        for (size_t r = 0; r < 12; ++r) {
            const size_t k = 2 * (r * nf + f);
            const Cx m = cdiv({sw[k], sw[k + 1]}, {sb[k], sb[k + 1]});
            const double mag = std::pow(std::hypot(m.re, m.im), w), arg = std::atan2(m.im, m.re) * w;
            directivity_out[k] = mag * std::cos(arg);
            directivity_out[k + 1] = mag * std::sin(arg);
        }
        on_axis_out[2 * f] = directivity_out[2 * f];
        on_axis_out[2 * f + 1] = directivity_out[2 * f + 1];
        const Cx on = {on_axis_out[2 * f], on_axis_out[2 * f + 1]};
        for (size_t r = 0; r < 12; ++r) {
            const size_t k = 2 * (r * nf + f);
            const Cx q = cdiv({directivity_out[k], directivity_out[k + 1]}, on);
            directivity_out[k] = q.re;
            directivity_out[k + 1] = q.im;
        }
    }
    return WV_OK;
}

extern "C" wv_status wv_calibrate_cabinet(const double* verts, uint32_t n_verts, const uint32_t* tris,
                                          uint32_t n_tris, const double ac[3], double c, double f_max, double ppw,
                                          double cells_across, const double* freqs, size_t nf, int precision,
                                          int batched, const volatile int32_t* cancel, double* on_axis_out,
                                          double* angles_out, double* directivity_out, char* err, size_t err_len) {
    return host_guarded(err, err_len, [&]() {
        return calibrate_cabinet_impl(verts, n_verts, tris, n_tris, ac, c, f_max, ppw, cells_across, freqs, nf,
                                      precision, batched, cancel, on_axis_out, angles_out, directivity_out, err,
                                      err_len);
    });
}
