// Surface admittance fits: fit_admittance (core/wavey-materials/src/
// impedance_fit.rs:158-235) and the absorption it fits, Material::absorption_at
// (core/wavey-materials/src/material.rs:72-93) over a Spectrum
// (core/wavey-materials/src/spectrum.rs:24-44).
//
// The fit is the reference's: 32 log-spaced frequencies over the band, the
// target absorption capped at 0.95, a conductance-only fit from the mean,
// then one and two resonant branches from nine deterministic starts each, two
// Nelder-Mead passes per start over (ln beta0, per branch ln m, ln k,
// logit(Q / 2)), the best by the Paris-formula RMS error, stopping once it is
// under 0.005; negligible branches dropped. Built with -ffp-contract=off and
// in the reference's operation order, so the fits come out as its do.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "wavey_fdtd.h"

namespace {

const double kPi = 3.141592653589793;
const double kMaxQ = 2.0;  // impedance_fit.rs:135
constexpr int kNf = 32;
constexpr int kMaxParams = 7;

struct Cx {
    double re, im;
};
Cx add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
Cx scale(Cx a, double s) { return {a.re * s, a.im * s}; }
// num-complex 0.4 division.
Cx div(Cx a, Cx b) {
    const double n = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / n, (a.im * b.re - a.re * b.im) / n};
}
double norm2(Cx a) { return a.re * a.re + a.im * a.im; }

double spectrum_eval(const wv_spectrum& s, double f) {
    if (s.n == 0) return s.values[0];
    if (f <= s.freqs_hz[0]) return s.values[0];
    if (f >= s.freqs_hz[s.n - 1]) return s.values[s.n - 1];
    const uint32_t i = (uint32_t)(std::upper_bound(s.freqs_hz, s.freqs_hz + s.n, f) - s.freqs_hz) - 1;
    const double f0 = std::log(s.freqs_hz[i]), f1 = std::log(s.freqs_hz[i + 1]);
    const double t = (std::log(f) - f0) / (f1 - f0);
    return s.values[i] + t * (s.values[i + 1] - s.values[i]);
}

// Paris integral by the midpoint rule over n angles of 1 - |R(cos th)|^2.
template <class R>
double paris_of(R refl, int n) {
    const double h = kPi / 2.0 / (double)n;
    double a = 0.0;
    for (int i = 0; i < n; ++i) {
        const double th = ((double)i + 0.5) * h;
        a += (1.0 - norm2(refl(std::cos(th)))) * std::sin(2.0 * th) * h;
    }
    return a;
}

// impedance_fit.rs:73-84 for admittance beta: R = (c - beta)/(c + beta).
double paris(Cx beta, int n) {
    return paris_of([&](double c) { return div({c - beta.re, 0.0 - beta.im}, {c + beta.re, beta.im}); }, n);
}

double absorption_at(const wv_material& m, double f) {
    if (m.impedance_re && m.impedance_im) {
        const Cx z{spectrum_eval(*m.impedance_re, f), spectrum_eval(*m.impedance_im, f)};
        return paris_of(
            [&](double c) {
                const Cx zc = scale(z, c);
                return div({zc.re - 1.0, zc.im}, {zc.re + 1.0, zc.im});
            },
            256);
    }
    return std::clamp(spectrum_eval(m.absorption, f), 0.0, 1.0);
}

struct Fit {
    double beta0 = 0.0;
    int nb = 0;
    wv_branch br[2]{};
};

// impedance_fit.rs:48-55
Cx admittance(const Fit& fit, double f) {
    const Cx jw{0.0, 2.0 * kPi * std::fmax(f, 1e-6)};
    Cx y{fit.beta0, 0.0};
    for (int b = 0; b < fit.nb; ++b) {
        const Cx den = add(add({fit.br[b].r, 0.0}, scale(jw, fit.br[b].m)), div({fit.br[b].k, 0.0}, jw));
        y = add(y, div({1.0, 0.0}, den));
    }
    return y;
}

Fit decode(const double* x, int nb) {
    Fit fit;
    fit.beta0 = std::exp(x[0]);
    fit.nb = nb;
    for (int b = 0; b < nb; ++b) {
        const double m = std::exp(x[1 + 3 * b]), k = std::exp(x[2 + 3 * b]);
        const double q = kMaxQ / (1.0 + std::exp(-x[3 + 3 * b]));
        fit.br[b] = {std::sqrt(m * k) / std::fmax(q, 1e-9), m, k};
    }
    return fit;
}

void encode_branch(double r, double m, double k, double* x) {
    double q = std::sqrt(m * k) / r / kMaxQ;
    q = q < 1e-6 ? 1e-6 : (q > 1.0 - 1e-6 ? 1.0 - 1e-6 : q);
    x[0] = std::log(m);
    x[1] = std::log(k);
    x[2] = std::log(q / (1.0 - q));
}

// The order f64::total_cmp gives.
int64_t total_key(double v) {
    int64_t x;
    std::memcpy(&x, &v, 8);
    return x ^ (int64_t)(((uint64_t)(x >> 63)) >> 1);
}

using Point = std::array<double, kMaxParams>;

// impedance_fit.rs:87-129
template <class F>
Point nelder_mead(const F& f, const Point& x0, int n, double step, int iters) {
    std::vector<Point> pts(n + 1, x0);
    for (int i = 0; i < n; ++i) pts[i + 1][i] += step;
    std::vector<double> vals(n + 1);
    for (int i = 0; i <= n; ++i) vals[i] = f(pts[i]);
    std::vector<int> idx(n + 1);
    for (int it = 0; it < iters; ++it) {
        for (int i = 0; i <= n; ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return total_key(vals[a]) < total_key(vals[b]); });
        std::vector<Point> sp(n + 1);
        std::vector<double> sv(n + 1);
        for (int i = 0; i <= n; ++i) {
            sp[i] = pts[idx[i]];
            sv[i] = vals[idx[i]];
        }
        pts.swap(sp);
        vals.swap(sv);
        if (std::fabs(vals[n] - vals[0]) < 1e-14) break;
        Point cen{};
        for (int d = 0; d < n; ++d) {
            double s = -0.0;
            for (int i = 0; i < n; ++i) s += pts[i][d];
            cen[d] = s / (double)n;
        }
        auto along = [&](double t) {
            Point p{};
            for (int d = 0; d < n; ++d) p[d] = cen[d] + t * (pts[n][d] - cen[d]);
            return p;
        };
        const Point xr = along(-1.0);
        const double fr = f(xr);
        if (fr < vals[0]) {
            const Point xe = along(-2.0);
            const double fe = f(xe);
            if (fe < fr) { pts[n] = xe; vals[n] = fe; }
            else { pts[n] = xr; vals[n] = fr; }
        } else if (fr < vals[n - 1]) {
            pts[n] = xr;
            vals[n] = fr;
        } else {
            const Point xc = along(fr < vals[n] ? -0.5 : 0.5);
            const double fc = f(xc);
            if (fc < std::fmin(vals[n], fr)) {
                pts[n] = xc;
                vals[n] = fc;
            } else {
                for (int i = 1; i <= n; ++i) {
                    for (int d = 0; d < n; ++d) pts[i][d] = pts[0][d] + 0.5 * (pts[i][d] - pts[0][d]);
                    vals[i] = f(pts[i]);
                }
            }
        }
    }
    int best = 0;
    for (int i = 1; i <= n; ++i)
        if (total_key(vals[i]) < total_key(vals[best])) best = i;
    return pts[best];
}

}  // namespace

extern "C" double wv_material_absorption(const wv_material* m, double f) { return absorption_at(*m, f); }

extern "C" wv_status wv_fit_admittance(const wv_material* material, double band_lo, double band_hi,
                                       uint32_t max_branches, wv_fit_result* out) {
    if (!material || !out || !material->absorption.values || max_branches > 2 ||
        (material->absorption.n && !material->absorption.freqs_hz))
        return WV_ERR_INVALID;
    const double lo = std::fmax(band_lo, 1.0), hi = std::fmax(band_hi, band_lo + 1.0);
    double freqs[kNf], target[kNf];
    bool all_small = true;
    for (int i = 0; i < kNf; ++i) {
        freqs[i] = lo * std::pow(hi / lo, (double)i / (double)(kNf - 1));
        target[i] = std::clamp(absorption_at(*material, freqs[i]), 0.0, 0.95);
        all_small = all_small && target[i] < 1e-6;
    }
    std::memset(out, 0, sizeof(*out));
    out->band[0] = lo;
    out->band[1] = hi;
    if (all_small) {  // rigid
        double m = 0.0;
        for (double t : target) m = std::fmax(m, t);
        out->rms_error = out->max_error = m;
        return WV_OK;
    }
    auto errors = [&](const Fit& fit, double& rms, double& mx) {
        double ss = -0.0;
        mx = 0.0;
        for (int i = 0; i < kNf; ++i) {
            const double e = paris(admittance(fit, freqs[i]), 96) - target[i];
            ss += e * e;
            mx = std::fmax(mx, std::fabs(e));
        }
        rms = std::sqrt(ss / (double)kNf);
    };
    double mean = -0.0;
    for (double t : target) mean += t;
    mean = mean / (double)kNf;
    const double beta_mean = 1.0 / wv_fit_real_impedance(std::fmax(mean, 1e-4));
    bool have = false;
    double best_rms = 0.0;
    Fit best;
    for (int nb = 0; nb <= (int)max_branches; ++nb) {
        auto cost = [&](const Point& x) {
            const Fit fit = decode(x.data(), nb);
            double s = 0.0;
            for (int i = 0; i < kNf; ++i) {
                const double e = paris(admittance(fit, freqs[i]), 48) - target[i];
                s += e * e;
            }
            return s;
        };
        std::vector<Point> starts;
        if (nb == 0) {
            Point x{};
            x[0] = std::log(beta_mean);
            starts.push_back(x);
        } else {
            for (double corner : {0.3, 1.0, 3.0})
                for (double level : {0.3, 1.0, 3.0}) {
                    Point x{};
                    x[0] = std::log(std::fmax(beta_mean * 0.3, 1e-6));
                    double p3 = 1.0;
                    for (int b = 0; b < nb; ++b) {
                        const double fc = std::sqrt(lo * hi) * corner * p3;
                        const double wc = 2.0 * kPi * fc;
                        const double r = 1.0 / std::fmax(beta_mean * level, 1e-6);
                        encode_branch(r, r / (wc * 3.0), r * wc, x.data() + 1 + 3 * b);
                        p3 *= 3.0;
                    }
                    starts.push_back(x);
                }
        }
        const int np = 1 + 3 * nb;
        for (const Point& x0 : starts) {
            const Point x1 = nelder_mead(cost, x0, np, 1.0, 1500 * (1 + nb));
            const Point x2 = nelder_mead(cost, x1, np, 0.2, 1000 * (1 + nb));
            const Fit fit = decode(x2.data(), nb);
            double rms, mx;
            errors(fit, rms, mx);
            if (!have || rms < best_rms - 1e-6) {
                have = true;
                best_rms = rms;
                best = fit;
            }
        }
        if (best_rms < 0.005) break;
    }
    Fit kept = best;
    kept.nb = 0;
    for (int b = 0; b < best.nb; ++b) {
        bool any = false;
        for (int i = 0; i < kNf && !any; ++i) {
            const Cx jw{0.0, 2.0 * kPi * freqs[i]};
            const Cx den = add(add({best.br[b].r, 0.0}, scale(jw, best.br[b].m)), div({best.br[b].k, 0.0}, jw));
            const Cx y = div({1.0, 0.0}, den);
            any = std::hypot(y.re, y.im) > 1e-5;
        }
        if (any) kept.br[kept.nb++] = best.br[b];
    }
    if (kept.beta0 < 1e-7) kept.beta0 = 0.0;
    out->beta0 = kept.beta0;
    out->n_branches = (uint32_t)kept.nb;
    for (int b = 0; b < kept.nb; ++b) out->branches[b] = kept.br[b];
    errors(kept, out->rms_error, out->max_error);
    return WV_OK;
}
