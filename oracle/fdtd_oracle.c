/*
 * FP64 CPU oracle for the Wavey FDTD wave-solver path -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of core/wavey-solver/src/fdtd.rs (and the pieces of
 * core/wavey-materials it needs), used only by tests/, the smoke check and
 * bench.py's CPU baseline, always as the checker. Arithmetic is written in
 * the reference's evaluation order and compiled with -ffp-contract=off, so
 * that no fused multiply-add changes a rounding the reference does not make.
 * Threading follows the reference's rayon structure: the interior update is
 * split over z-slabs (par_chunks_mut(nxy), fdtd.rs:699-708), the boundary
 * nodes are updated in parallel (fdtd.rs:710-713), the scatter, injection
 * and recording are serial. Every cell and node update is independent, so
 * results do not depend on the thread count.
 *
 * Pinning: every function is checked by tests/ against what fixes it without
 * the reference (closed forms, brute force, the reference's own tests and
 * data). Three parts are partly "parity unpinned", as only running the
 * reference could pin them exactly: the triangle a face ray takes where two
 * are hit at exactly the same distance (ob_build_room takes the lower index;
 * the reference's BVH takes its tree order), the exact Nelder-Mead
 * trajectory of ob_fit_admittance (the last digits of a fit), and the
 * composed outputs of ob_calibrate_cabinet (pinned through its pieces and
 * its exact limits).
 */
#include "fdtd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* fdtd.rs:137-140 */
static const double COURANT = 0.5773502691896257;
static const int64_t MAX_CELLS = 600000000;
/* fdtd.rs:209, :224 */
#define AIR ((uint8_t)(1u << 6))
#define NO_BOUNDARY 0xFFFFFFFFu

static const double PI = 3.141592653589793;

/* ---------------------------------------------------------------- complex */
typedef struct { double re, im; } cplx;
static cplx cmk(double re, double im) { cplx z = {re, im}; return z; }
static cplx cadd(cplx a, cplx b) { return cmk(a.re + b.re, a.im + b.im); }
static cplx cmul(cplx a, cplx b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cplx cscale(cplx a, double s) { return cmk(a.re * s, a.im * s); }
/* num-complex 0.4 Div<Complex>: (a c + b d)/|z|^2, (b c - a d)/|z|^2 */
static cplx cdiv(cplx a, cplx b) {
    double n = b.re * b.re + b.im * b.im;
    double re = a.re * b.re + a.im * b.im;
    double im = a.im * b.re - a.re * b.im;
    return cmk(re / n, im / n);
}
static double cnorm2(cplx a) { return a.re * a.re + a.im * a.im; }
static cplx from_polar(double r, double th) { return cmk(r * cos(th), r * sin(th)); }
/* Rust f64::fract: x - trunc(x) */
static double fract(double x) { return x - trunc(x); }

/* --------------------------------------------------- materials (L0 crate) */

/* impedance_fit.rs:48-55 */
void ob_admittance(const ob_fit* fit, double f, double out[2]) {
    double w = 2.0 * PI * (f > 1e-6 ? f : 1e-6);
    cplx jw = cmk(0.0, w);
    cplx y = cmk(fit->beta0, 0.0);
    for (int b = 0; b < fit->nb; ++b) {
        /* b.r + jw*b.m + b.k/jw */
        cplx den = cadd(cadd(cmk(fit->br[b].r, 0.0), cscale(jw, fit->br[b].m)),
                        cdiv(cmk(fit->br[b].k, 0.0), jw));
        y = cadd(y, cdiv(cmk(1.0, 0.0), den));
    }
    out[0] = y.re;
    out[1] = y.im;
}

/* impedance_fit.rs:73-84 (paris): midpoint rule over n angles. */
double ob_paris(double beta_re, double beta_im, int n) {
    double h = PI / 2.0 / (double)n;
    double a = 0.0;
    cplx beta = cmk(beta_re, beta_im);
    for (int i = 0; i < n; ++i) {
        double th = ((double)i + 0.5) * h;
        double c = cos(th);
        cplx r = cdiv(cmk(c - beta.re, -beta.im), cmk(c + beta.re, beta.im));
        a += (1.0 - cnorm2(r)) * sin(2.0 * th) * h;
    }
    return a;
}

/* material.rs:81-91 with R = (z c - 1)/(z c + 1) real. */
static double stat_abs_real(double z) {
    const int N = 256;
    double h = (PI / 2.0) / (double)N;
    double s = 0.0;
    for (int i = 0; i < N; ++i) {
        double th = ((double)i + 0.5) * h;
        double c = cos(th);
        double r = (z * c - 1.0) / (z * c + 1.0);
        s += (1.0 - r * r) * sin(2.0 * th) * h;
    }
    return s;
}

/* material.rs:96-111 */
double ob_fit_real_impedance(double alpha) {
    double lo = 1.567, hi = 1e7;
    if (alpha >= stat_abs_real(lo)) return lo;
    if (alpha <= 0.0) return hi;
    for (int i = 0; i < 80; ++i) {
        double mid = sqrt(lo * hi);
        if (stat_abs_real(mid) > alpha) lo = mid; else hi = mid;
    }
    return sqrt(lo * hi);
}

/* spectrum.rs:24-44 (Spectrum::eval): a scalar (n = 0, values[0]) or a table
 * interpolated linearly over ln f, held constant outside its range. */
double ob_spectrum_eval(const double* freqs_hz, const double* values, int32_t n, double f) {
    if (n == 0) return values[0];
    if (f <= freqs_hz[0]) return values[0];
    if (f >= freqs_hz[n - 1]) return values[n - 1];
    int32_t i = 0;
    while (i + 1 < n && freqs_hz[i + 1] <= f) ++i; /* partition_point(x <= f) - 1 */
    double f0 = log(freqs_hz[i]), f1 = log(freqs_hz[i + 1]);
    double t = (log(f) - f0) / (f1 - f0);
    return values[i] + t * (values[i + 1] - values[i]);
}

/* Rust's f64::total_cmp. */
static int total_cmp(double a, double b) {
    int64_t x, y;
    memcpy(&x, &a, 8);
    memcpy(&y, &b, 8);
    x ^= (int64_t)(((uint64_t)(x >> 63)) >> 1);
    y ^= (int64_t)(((uint64_t)(y >> 63)) >> 1);
    return (x > y) - (x < y);
}

#define OB_NM_MAX 7 /* 1 + 3 * 2 parameters */

typedef double (*ob_cost_fn)(const double* x, void* ctx);

/* impedance_fit.rs:87-129 (nelder_mead). */
static double nelder_mead(ob_cost_fn f, void* ctx, const double* x0, int n, double step, int iters, double* xbest) {
    double pts[OB_NM_MAX + 1][OB_NM_MAX], vals[OB_NM_MAX + 1];
    for (int i = 0; i <= n; ++i) {
        for (int d = 0; d < n; ++d) pts[i][d] = x0[d];
        if (i > 0) pts[i][i - 1] += step;
        vals[i] = f(pts[i], ctx);
    }
    for (int it = 0; it < iters; ++it) {
        /* Stable insertion sort of the vertices by value (sort_by is stable). */
        for (int i = 1; i <= n; ++i) {
            double pv = vals[i], pp[OB_NM_MAX];
            memcpy(pp, pts[i], sizeof(double) * n);
            int j = i - 1;
            while (j >= 0 && total_cmp(vals[j], pv) > 0) {
                vals[j + 1] = vals[j];
                memcpy(pts[j + 1], pts[j], sizeof(double) * n);
                --j;
            }
            vals[j + 1] = pv;
            memcpy(pts[j + 1], pp, sizeof(double) * n);
        }
        if (fabs(vals[n] - vals[0]) < 1e-14) break;
        double cen[OB_NM_MAX], xr[OB_NM_MAX], xe[OB_NM_MAX], xc[OB_NM_MAX];
        for (int d = 0; d < n; ++d) {
            double sum = -0.0;
            for (int i = 0; i < n; ++i) sum += pts[i][d];
            cen[d] = sum / (double)n;
        }
        for (int d = 0; d < n; ++d) xr[d] = cen[d] + -1.0 * (pts[n][d] - cen[d]);
        double fr = f(xr, ctx);
        if (fr < vals[0]) {
            for (int d = 0; d < n; ++d) xe[d] = cen[d] + -2.0 * (pts[n][d] - cen[d]);
            double fe = f(xe, ctx);
            if (fe < fr) { memcpy(pts[n], xe, sizeof(double) * n); vals[n] = fe; }
            else { memcpy(pts[n], xr, sizeof(double) * n); vals[n] = fr; }
        } else if (fr < vals[n - 1]) {
            memcpy(pts[n], xr, sizeof(double) * n);
            vals[n] = fr;
        } else {
            double t = fr < vals[n] ? -0.5 : 0.5;
            for (int d = 0; d < n; ++d) xc[d] = cen[d] + t * (pts[n][d] - cen[d]);
            double fc = f(xc, ctx);
            if (fc < fmin(vals[n], fr)) {
                memcpy(pts[n], xc, sizeof(double) * n);
                vals[n] = fc;
            } else {
                for (int i = 1; i <= n; ++i) {
                    for (int d = 0; d < n; ++d) pts[i][d] = pts[0][d] + 0.5 * (pts[i][d] - pts[0][d]);
                    vals[i] = f(pts[i], ctx);
                }
            }
        }
    }
    int best = 0;
    for (int i = 1; i <= n; ++i)
        if (total_cmp(vals[i], vals[best]) < 0) best = i;
    memcpy(xbest, pts[best], sizeof(double) * n);
    return vals[best];
}

#define OB_MAX_Q 2.0 /* impedance_fit.rs:135 */

/* impedance_fit.rs:139-150 (decode): ln beta0, then per branch (ln m, ln k, logit(Q / MAX_Q)). */
static void fit_decode(const double* x, int nb, ob_fit* out) {
    memset(out, 0, sizeof(*out));
    out->beta0 = exp(x[0]);
    out->nb = nb;
    for (int b = 0; b < nb; ++b) {
        double m = exp(x[1 + 3 * b]), k = exp(x[2 + 3 * b]);
        double q = OB_MAX_Q / (1.0 + exp(-x[3 + 3 * b]));
        out->br[b].r = sqrt(m * k) / fmax(q, 1e-9);
        out->br[b].m = m;
        out->br[b].k = k;
    }
}

/* impedance_fit.rs:152-155 (encode_branch), with Rust's clamp. */
static void fit_encode(double r, double m, double k, double* x) {
    double q = sqrt(m * k) / r / OB_MAX_Q;
    if (q < 1e-6) q = 1e-6;
    if (q > 1.0 - 1e-6) q = 1.0 - 1e-6;
    x[0] = log(m);
    x[1] = log(k);
    x[2] = log(q / (1.0 - q));
}

typedef struct { int nb; const double* freqs; const double* target; int nf; } ob_cost_ctx;

static double fit_cost(const double* x, void* vctx) {
    const ob_cost_ctx* c = (const ob_cost_ctx*)vctx;
    ob_fit fit;
    fit_decode(x, c->nb, &fit);
    double s = 0.0;
    for (int i = 0; i < c->nf; ++i) {
        double y[2];
        ob_admittance(&fit, c->freqs[i], y);
        double e = ob_paris(y[0], y[1], 48) - c->target[i];
        s += e * e;
    }
    return s;
}

static void fit_errors(const ob_fit* fit, const double* freqs, const double* target, int nf, double* rms,
                       double* mx) {
    double ss = -0.0, m = 0.0;
    for (int i = 0; i < nf; ++i) {
        double y[2];
        ob_admittance(fit, freqs[i], y);
        double e = ob_paris(y[0], y[1], 96) - target[i];
        ss += e * e;
        m = fmax(m, fabs(e));
    }
    *rms = sqrt(ss / (double)nf);
    *mx = m;
}

/* impedance_fit.rs:158-235 (fit_admittance). Writes the fit and returns the
 * number of branches kept; band_out = (lo, hi), err_out = (rms, max). */
/* material.rs:72-78 (absorption_at): from the impedance when the material
 * has one (Paris integral of R = (z c - 1)/(z c + 1), material.rs:83-93),
 * else the absorption spectrum clamped to 0..1. */
double ob_material_absorption(const ob_material* m, double f) {
    if (m->has_impedance) {
        cplx z = cmk(ob_spectrum_eval(m->imp_re.f, m->imp_re.v, m->imp_re.n, f),
                     ob_spectrum_eval(m->imp_im.f, m->imp_im.v, m->imp_im.n, f));
        const int N = 256;
        double h = (PI / 2.0) / (double)N, s = -0.0;
        for (int i = 0; i < N; ++i) {
            double th = ((double)i + 0.5) * h;
            double c = cos(th);
            cplx zc = cscale(z, c);
            cplx r = cdiv(cmk(zc.re - 1.0, zc.im), cmk(zc.re + 1.0, zc.im));
            s += (1.0 - cnorm2(r)) * sin(2.0 * th) * h;
        }
        return s;
    }
    double a = ob_spectrum_eval(m->absorption.f, m->absorption.v, m->absorption.n, f);
    return a < 0.0 ? 0.0 : (a > 1.0 ? 1.0 : a);
}

int ob_fit_admittance(const ob_material* mat, double band_lo, double band_hi, int max_branches, ob_fit* out,
                      double band_out[2], double err_out[2]) {
    const double lo = fmax(band_lo, 1.0), hi = fmax(band_hi, band_lo + 1.0);
    enum { NF = 32 };
    double freqs[NF], target[NF];
    int all_small = 1;
    for (int i = 0; i < NF; ++i) {
        freqs[i] = lo * pow(hi / lo, (double)i / (double)(NF - 1));
        double a = ob_material_absorption(mat, freqs[i]);
        target[i] = a < 0.0 ? 0.0 : (a > 0.95 ? 0.95 : a);
        all_small &= target[i] < 1e-6;
    }
    band_out[0] = lo;
    band_out[1] = hi;
    memset(out, 0, sizeof(*out));
    out->band[0] = lo;
    out->band[1] = hi;
    if (all_small) {
        double m = 0.0;
        for (int i = 0; i < NF; ++i) m = fmax(m, target[i]);
        err_out[0] = err_out[1] = m;
        out->rms_error = out->max_error = m;
        return 0;
    }
    double mean = -0.0;
    for (int i = 0; i < NF; ++i) mean += target[i];
    mean = mean / (double)NF;
    const double beta_mean = 1.0 / ob_fit_real_impedance(fmax(mean, 1e-4));
    int have = 0;
    double best_rms = 0.0;
    ob_fit best;
    memset(&best, 0, sizeof(best));
    for (int nb = 0; nb <= max_branches && nb <= 2; ++nb) {
        ob_cost_ctx ctx = {nb, freqs, target, NF};
        const int np = 1 + 3 * nb;
        double starts[9][OB_NM_MAX];
        int ns = 0;
        if (nb == 0) {
            starts[0][0] = log(beta_mean);
            ns = 1;
        } else {
            static const double corners[3] = {0.3, 1.0, 3.0}, levels[3] = {0.3, 1.0, 3.0};
            for (int ci = 0; ci < 3; ++ci)
                for (int li = 0; li < 3; ++li) {
                    double* x = starts[ns++];
                    x[0] = log(fmax(beta_mean * 0.3, 1e-6));
                    for (int b = 0; b < nb; ++b) {
                        double p3 = 1.0;
                        for (int e = 0; e < b; ++e) p3 *= 3.0;
                        double fc = sqrt(lo * hi) * corners[ci] * p3;
                        double wc = 2.0 * PI * fc;
                        double r = 1.0 / fmax(beta_mean * levels[li], 1e-6);
                        fit_encode(r, r / (wc * 3.0), r * wc, x + 1 + 3 * b);
                    }
                }
        }
        for (int si = 0; si < ns; ++si) {
            double x1[OB_NM_MAX], x2[OB_NM_MAX];
            nelder_mead(fit_cost, &ctx, starts[si], np, 1.0, 1500 * (1 + nb), x1);
            nelder_mead(fit_cost, &ctx, x1, np, 0.2, 1000 * (1 + nb), x2);
            ob_fit fit;
            fit_decode(x2, nb, &fit);
            double rms, mx;
            fit_errors(&fit, freqs, target, NF, &rms, &mx);
            if (!have || rms < best_rms - 1e-6) {
                have = 1;
                best_rms = rms;
                best = fit;
            }
        }
        if (best_rms < 0.005) break;
    }
    /* Drop branches with |1 / (r + j w m + k / (j w))| <= 1e-5 at every fitted frequency. */
    ob_fit kept = best;
    kept.nb = 0;
    for (int b = 0; b < best.nb; ++b) {
        int any = 0;
        for (int i = 0; i < NF && !any; ++i) {
            cplx jw = cmk(0.0, 2.0 * PI * freqs[i]);
            cplx den = cadd(cadd(cmk(best.br[b].r, 0.0), cscale(jw, best.br[b].m)), cdiv(cmk(best.br[b].k, 0.0), jw));
            cplx y = cdiv(cmk(1.0, 0.0), den);
            any = hypot(y.re, y.im) > 1e-5;
        }
        if (any) kept.br[kept.nb++] = best.br[b];
    }
    if (kept.beta0 < 1e-7) kept.beta0 = 0.0;
    fit_errors(&kept, freqs, target, NF, &err_out[0], &err_out[1]);
    kept.band[0] = lo;
    kept.band[1] = hi;
    kept.rms_error = err_out[0];
    kept.max_error = err_out[1];
    *out = kept;
    return kept.nb;
}

/* fdtd.rs:157-162 */
double ob_phase_error_deg_per_m(double f, double c, double dx) {
    double dt = COURANT * dx / c;
    double w = 2.0 * PI * f;
    double k_num = 2.0 / dx * asin(sin(w * dt / 2.0) / COURANT);
    return fabs(k_num - w / c) * (180.0 / PI);
}

/* ------------------------------------------------------------ boundary */

/* impedance_fit.rs:34-44: #[derive(PartialEq)] over beta0, branches, band,
 * rms_error and max_error. */
static int fit_eq(const ob_fit* a, const ob_fit* b) {
    if (a->beta0 != b->beta0 || a->nb != b->nb) return 0;
    for (int i = 0; i < a->nb; ++i)
        if (a->br[i].r != b->br[i].r || a->br[i].m != b->br[i].m || a->br[i].k != b->br[i].k) return 0;
    return a->band[0] == b->band[0] && a->band[1] == b->band[1] && a->rms_error == b->rms_error &&
           a->max_error == b->max_error;
}

/* fdtd.rs:539-571: faces with equal fits share slots (impedance / count). */
int ob_boundary_slots(const int32_t* face_fits, int n_faces, const ob_fit* fits, double dt,
                      ob_slot* out, int max_out) {
    int grp[6], cnt[6], ng = 0, ns = 0;
    for (int f = 0; f < n_faces; ++f) {
        int s = face_fits[f], found = -1;
        for (int g = 0; g < ng; ++g)
            if (fit_eq(&fits[grp[g]], &fits[s])) { found = g; break; }
        if (found >= 0) cnt[found]++;
        else { grp[ng] = s; cnt[ng] = 1; ng++; }
    }
    for (int g = 0; g < ng; ++g) {
        const ob_fit* fit = &fits[grp[g]];
        double n = (double)cnt[g];
        int total = (fit->beta0 > 0.0 ? 1 : 0) + fit->nb;
        for (int b = -1; b < fit->nb; ++b) {
            double r, m, k;
            if (b < 0) {
                if (!(fit->beta0 > 0.0)) continue;
                r = 1.0 / fit->beta0; m = 0.0; k = 0.0;
            } else {
                r = fit->br[b].r; m = fit->br[b].m; k = fit->br[b].k;
            }
            r = r / n; m = m / n; k = k / n;
            double d, cv, w;
            if (m > 0.0) {
                d = m / dt + r / 2.0 + k * dt / 4.0;
                cv = m / dt - r / 2.0 - k * dt / 4.0;
                w = 0.25;
            } else {
                d = r + k * dt / 2.0;
                cv = -k * dt / 2.0;
                w = 0.0;
            }
            if (ns >= max_out) return -1;
            out[ns].inv_d = 1.0 / d;
            out[ns].cv = cv;
            out[ns].k = k;
            out[ns].w = w;
            ns++;
        }
        (void)total;
    }
    return ns;
}

/* fdtd.rs:581-598 */
double ob_boundary_update(const ob_slot* slots, int n, double* st, double p, double p_prev,
                          double lap, double lam, double dt) {
    double s1 = 0.0, s2 = 0.0;
    for (int j = 0; j < n; ++j) {
        const ob_slot* sl = &slots[j];
        double u = st[2 * j], g = st[2 * j + 1];
        double e = (sl->cv * u - sl->k * g + 0.5 * p + sl->w * p_prev) * sl->inv_d;
        s1 += (0.5 - sl->w) * sl->inv_d;
        s2 += e - u;
    }
    double next = (2.0 * p - p_prev + lap - lam * s2) / (1.0 + lam * s1);
    for (int j = 0; j < n; ++j) {
        const ob_slot* sl = &slots[j];
        double u0 = st[2 * j], g = st[2 * j + 1];
        double e = (sl->cv * u0 - sl->k * g + 0.5 * p + sl->w * p_prev) * sl->inv_d;
        double u = (0.5 - sl->w) * sl->inv_d * next + e;
        st[2 * j + 1] = g + 0.5 * dt * (u + u0);
        st[2 * j] = u;
    }
    return next;
}

/* ---------------------------------------------------------------- grid */

/* Surface patch lookup for the box-room special case of the BVH face search
 * (fdtd.rs:351-363): the ray from an air node toward a missing neighbour
 * crosses the surface halfway between the two nodes. */
static int patch_fit(const ob_room* room, int s, double u, double v, double u0, double u1,
                     double v0, double v1) {
    int gu = room->gu[s], gv = room->gv[s];
    int iu = (int)floor((u - u0) / (u1 - u0) * gu);
    int iv = (int)floor((v - v0) / (v1 - v0) * gv);
    if (iu < 0) iu = 0;
    if (iu >= gu) iu = gu - 1;
    if (iv < 0) iv = 0;
    if (iv >= gv) iv = gv - 1;
    return room->patch_fit[room->patch_off[s] + iv * gu + iu];
}

static int in_pillar(const ob_room* room, double x, double y) {
    for (int p = 0; p < room->n_pillars; ++p) {
        const double* q = room->pillars + 4 * p;
        if (x > q[0] && x < q[1] && y > q[2] && y < q[3]) return p;
    }
    return -1;
}

/* Room::contains for an extruded box with pillar holes
 * (core/wavey-geom/src/room.rs:151-157). */
static int room_contains(const ob_room* room, double x, double y, double z) {
    return z > 0.0 && z < room->L[2] && x > 0.0 && x < room->L[0] && y > 0.0 && y < room->L[1] &&
           in_pillar(room, x, y) < 0;
}

/* The fit behind the missing face of direction d of the node at (x,y,z). */
static int face_fit(const ob_room* room, double x, double y, double z, int d, double dx) {
    static const int DX[6] = {-1, 1, 0, 0, 0, 0}, DY[6] = {0, 0, -1, 1, 0, 0}, DZ[6] = {0, 0, 0, 0, -1, 1};
    double hx = x + DX[d] * dx * 0.5, hy = y + DY[d] * dx * 0.5, hz = z + DZ[d] * dx * 0.5;
    double nx = x + DX[d] * dx, ny = y + DY[d] * dx;
    const double* L = room->L;
    int p = (d < 4) ? in_pillar(room, nx, ny) : -1;
    if (p >= 0) {
        const double* q = room->pillars + 4 * p;
        /* The side whose outward normal faces the node: -dir. */
        int side = d ^ 1;
        int s = 6 + 4 * p + side;
        if (d < 2) return patch_fit(room, s, hy, hz, q[2], q[3], 0.0, L[2]);
        return patch_fit(room, s, hx, hz, q[0], q[1], 0.0, L[2]);
    }
    switch (d) {
    case 0: return patch_fit(room, 0, hy, hz, 0.0, L[1], 0.0, L[2]);
    case 1: return patch_fit(room, 1, hy, hz, 0.0, L[1], 0.0, L[2]);
    case 2: return patch_fit(room, 2, hx, hz, 0.0, L[0], 0.0, L[2]);
    case 3: return patch_fit(room, 3, hx, hz, 0.0, L[0], 0.0, L[2]);
    case 4: return patch_fit(room, 4, hx, hy, 0.0, L[0], 0.0, L[1]);
    default: return patch_fit(room, 5, hx, hy, 0.0, L[0], 0.0, L[1]);
    }
}

void ob_grid_free(ob_grid* g) {
    free(g->flags);
    free(g->bidx);
    free(g->b_cell);
    free(g->b_start);
    free(g->slots);
    memset(g, 0, sizeof(*g));
}

/* fdtd.rs:289-377 (build_dx) for a box room with pillars and no cabinets. */
int ob_build_box(const ob_room* room, double c, double dx, ob_grid* g) {
    memset(g, 0, sizeof(*g));
    double dt = COURANT * dx / c;
    int64_t n[3];
    for (int k = 0; k < 3; ++k) n[k] = (int64_t)ceil(room->L[k] / dx) + 2;
    int64_t cells = n[0] * n[1] * n[2];
    if (cells > MAX_CELLS) return -1;
    g->dx = dx;
    g->dt = dt;
    for (int k = 0; k < 3; ++k) {
        g->n[k] = n[k];
        g->origin[k] = 0.0 - dx * 0.5;
    }
    uint8_t* air = (uint8_t*)calloc((size_t)cells, 1);
    g->flags = (uint8_t*)calloc((size_t)cells, 1);
    g->bidx = (uint32_t*)malloc((size_t)cells * 4);
    if (!air || !g->flags || !g->bidx) return -2;
    const int64_t nx = n[0], nxy = n[0] * n[1];
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < n[2]; ++z)
        for (int64_t y = 0; y < n[1]; ++y)
            for (int64_t x = 0; x < nx; ++x) {
                int64_t i = z * nxy + y * nx + x;
                double px = g->origin[0] + (double)x * dx, py = g->origin[1] + (double)y * dx,
                       pz = g->origin[2] + (double)z * dx;
                air[i] = x > 0 && y > 0 && z > 0 && x < n[0] - 1 && y < n[1] - 1 && z < n[2] - 1 &&
                         room_contains(room, px, py, pz);
            }
    /* Serial, in cell order, as fdtd.rs:368-375 builds the boundary list. */
    int64_t cap_b = 1024, cap_s = 4096;
    g->b_cell = (uint32_t*)malloc(cap_b * 4);
    g->b_start = (uint32_t*)malloc((cap_b + 1) * 4);
    g->slots = (ob_slot*)malloc(cap_s * sizeof(ob_slot));
    static const int64_t DXI[6] = {-1, 1, 0, 0, 0, 0}, DYI[6] = {0, 0, -1, 1, 0, 0}, DZI[6] = {0, 0, 0, 0, -1, 1};
    for (int64_t i = 0; i < cells; ++i) {
        g->bidx[i] = NO_BOUNDARY;
        if (!air[i]) continue;
        int64_t z = i / nxy, r = i % nxy, y = r / nx, x = r % nx;
        double px = g->origin[0] + (double)x * dx, py = g->origin[1] + (double)y * dx,
               pz = g->origin[2] + (double)z * dx;
        uint8_t fl = AIR;
        int32_t faces[6];
        int nf = 0;
        for (int k = 0; k < 6; ++k) {
            int64_t j = (z + DZI[k]) * nxy + (y + DYI[k]) * nx + (x + DXI[k]);
            if (air[j]) fl |= (uint8_t)(1u << k);
            else faces[nf++] = face_fit(room, px, py, pz, k, dx);
        }
        g->flags[i] = fl;
        ob_slot sl[6 * (OB_MAX_BRANCHES + 1)];
        int ns = ob_boundary_slots(faces, nf, room->fits, dt, sl, 6 * (OB_MAX_BRANCHES + 1));
        if (ns < 0) return -3;
        if (ns > 0) {
            if (g->nb + 1 >= cap_b) {
                cap_b *= 2;
                g->b_cell = (uint32_t*)realloc(g->b_cell, cap_b * 4);
                g->b_start = (uint32_t*)realloc(g->b_start, (cap_b + 1) * 4);
            }
            while (g->ns + ns > cap_s) {
                cap_s *= 2;
                g->slots = (ob_slot*)realloc(g->slots, cap_s * sizeof(ob_slot));
            }
            g->bidx[i] = (uint32_t)g->nb;
            g->b_cell[g->nb] = (uint32_t)i;
            g->b_start[g->nb] = (uint32_t)g->ns;
            memcpy(g->slots + g->ns, sl, ns * sizeof(ob_slot));
            g->ns += ns;
            g->nb++;
        }
    }
    g->b_start[g->nb] = (uint32_t)g->ns;
    free(air);
    return 0;
}

/* ---------------------------------------------------------- general rooms */

/* glam's DVec3 arithmetic: cross (a.y*b.z - b.y*a.z, ...), dot summed left
 * to right, normalize = v * (1 / sqrt(v.v)). */
static void v3sub(const double a[3], const double b[3], double o[3]) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
static void v3cross(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
}
static double v3dot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* bvh.rs:209-226 */
int ob_ray_triangle(const double o[3], const double d[3], const double a[3], const double b[3],
                    const double c[3], double* t) {
    double e1[3], e2[3], p[3], s[3], q[3];
    v3sub(b, a, e1);
    v3sub(c, a, e2);
    v3cross(d, e2, p);
    double det = v3dot(e1, p);
    if (fabs(det) < 1e-14) return 0;
    double inv = 1.0 / det;
    v3sub(o, a, s);
    double u = v3dot(s, p) * inv;
    if (!(u >= -1e-10 && u <= 1.0 + 1e-10)) return 0;
    v3cross(s, e1, q);
    double v = v3dot(d, q) * inv;
    if (v < -1e-10 || u + v > 1.0 + 1e-10) return 0;
    *t = v3dot(e2, q) * inv;
    return 1;
}

static const double* tri_vertex(const ob_mesh_room* room, int32_t k, int j) {
    return room->pos + 3 * (int64_t)room->tri[3 * k + j];
}

/* Bvh::intersect (bvh.rs:63-73): closest hit with t_min < t < t_max. */
static int32_t closest_hit(const ob_mesh_room* room, const double o[3], const double d[3], double t_min,
                           double t_max) {
    int32_t best = -1;
    double lim = t_max;
    for (int32_t k = 0; k < room->n_tri; ++k) {
        double t;
        if (ob_ray_triangle(o, d, tri_vertex(room, k, 0), tri_vertex(room, k, 1), tri_vertex(room, k, 2), &t) &&
            t > t_min && t < lim) {
            lim = t;
            best = k;
        }
    }
    return best;
}

typedef struct { double t; int32_t k; } ob_hit;

static int hit_cmp(const void* a, const void* b) {
    const ob_hit* x = (const ob_hit*)a;
    const ob_hit* y = (const ob_hit*)b;
    if (x->t < y->t) return -1;
    if (x->t > y->t) return 1;
    return (x->k > y->k) - (x->k < y->k);
}

/* Bvh::all_hits (bvh.rs:75-87): every hit with t_min < t < t_max, sorted.
 * hits must hold n_tri entries. */
static int32_t all_hits(const ob_mesh_room* room, const double o[3], const double d[3], double t_min,
                        double t_max, ob_hit* hits) {
    int32_t n = 0;
    for (int32_t k = 0; k < room->n_tri; ++k) {
        double t;
        if (ob_ray_triangle(o, d, tri_vertex(room, k, 0), tri_vertex(room, k, 1), tri_vertex(room, k, 2), &t) &&
            t > t_min && t < t_max) {
            hits[n].t = t;
            hits[n].k = k;
            n++;
        }
    }
    qsort(hits, (size_t)n, sizeof(ob_hit), hit_cmp);
    return n;
}

/* room.rs:355-367 */
int ob_parity_inside(const ob_mesh_room* room, const double p[3]) {
    static const double DIRS[3][3] = {
        {0.5772156649, 0.6180339887, 0.5338319},
        {-0.6931471806, 0.4142135624, -0.3010299957},
        {0.2360679775, -0.7320508076, 0.6457513111},
    };
    ob_hit* hits = (ob_hit*)malloc(sizeof(ob_hit) * (size_t)(room->n_tri + 1));
    int odd = 0;
    for (int r = 0; r < 3; ++r) {
        double inv = 1.0 / sqrt(v3dot(DIRS[r], DIRS[r]));
        double d[3] = {DIRS[r][0] * inv, DIRS[r][1] * inv, DIRS[r][2] * inv};
        if (all_hits(room, p, d, 1e-9, INFINITY, hits) % 2 == 1) odd++;
    }
    free(hits);
    return odd >= 2;
}

/* regions.rs:237-249 */
static int point_in_polygon(double px, double py, const double* xy, int32_t n) {
    int inside = 0;
    int32_t j = n - 1;
    for (int32_t i = 0; i < n; ++i) {
        double ax = xy[2 * i], ay = xy[2 * i + 1], bx = xy[2 * j], by = xy[2 * j + 1];
        if ((ay > py) != (by > py) && px < (bx - ax) * (py - ay) / (by - ay) + ax) inside = !inside;
        j = i;
    }
    return inside;
}

/* Room::contains for an extruded room (room.rs:151-154, regions.rs:44-47). */
static int prism_contains(const ob_mesh_room* room, const double p[3]) {
    if (!(p[2] > 0.0 && p[2] < room->height)) return 0;
    for (int32_t r = 0; r < room->n_regions; ++r) {
        int32_t c0 = room->region_cycles[r], c1 = room->region_cycles[r + 1];
        const int32_t* off = room->cycle_off;
        if (!point_in_polygon(p[0], p[1], room->cycle_xy + 2 * (int64_t)off[c0], off[c0 + 1] - off[c0])) continue;
        int in_hole = 0;
        for (int32_t h = c0 + 1; h < c1 && !in_hole; ++h)
            in_hole = point_in_polygon(p[0], p[1], room->cycle_xy + 2 * (int64_t)off[h], off[h + 1] - off[h]);
        if (!in_hole) return 1;
    }
    return 0;
}

/* fdtd.rs:433-456 */
int ob_solid_from_mesh(const double* verts, int32_t n_verts, const int32_t* tris, int32_t n_tris,
                       const double rot[9], const double pos[3], double* planes_out, double lo[3],
                       double hi[3]) {
    if (n_tris <= 0) return -1;
    double* w = (double*)malloc(sizeof(double) * 3 * (size_t)(n_verts > 0 ? n_verts : 1));
    for (int32_t i = 0; i < n_verts; ++i) {
        const double* v = verts + 3 * i;
        for (int r = 0; r < 3; ++r)
            w[3 * i + r] = (rot[3 * r] * v[0] + rot[3 * r + 1] * v[1] + rot[3 * r + 2] * v[2]) + pos[r];
    }
    int np = 0;
    for (int32_t k = 0; k < n_tris; ++k) {
        const double* a = w + 3 * tris[3 * k];
        const double* b = w + 3 * tris[3 * k + 1];
        const double* c = w + 3 * tris[3 * k + 2];
        double e1[3], e2[3], n[3];
        v3sub(b, a, e1);
        v3sub(c, a, e2);
        v3cross(e1, e2, n);
        if (sqrt(v3dot(n, n)) < 1e-12) continue;
        double inv = 1.0 / sqrt(v3dot(n, n));
        n[0] *= inv;
        n[1] *= inv;
        n[2] *= inv;
        double d = v3dot(n, a);
        int dup = 0;
        for (int j = 0; j < np && !dup; ++j) {
            const double* m = planes_out + 4 * j;
            dup = v3dot(m, n) > 1.0 - 1e-9 && fabs(m[3] - d) < 1e-9;
        }
        if (dup) continue;
        planes_out[4 * np] = n[0];
        planes_out[4 * np + 1] = n[1];
        planes_out[4 * np + 2] = n[2];
        planes_out[4 * np + 3] = d;
        np++;
    }
    for (int r = 0; r < 3; ++r) {
        lo[r] = INFINITY;
        hi[r] = -INFINITY;
    }
    for (int32_t i = 0; i < n_verts; ++i)
        for (int r = 0; r < 3; ++r) {
            lo[r] = fmin(lo[r], w[3 * i + r]);
            hi[r] = fmax(hi[r], w[3 * i + r]);
        }
    free(w);
    return np;
}

/* fdtd.rs:458-460 */
int ob_solid_contains(const ob_solid* s, const double p[3]) {
    for (int r = 0; r < 3; ++r)
        if (!(p[r] >= s->lo[r] && p[r] <= s->hi[r])) return 0;
    for (int32_t j = 0; j < s->n_planes; ++j)
        if (!(v3dot(s->planes + 4 * j, p) <= s->planes[4 * j + 3])) return 0;
    return 1;
}

/* fdtd.rs:385-422: per (x, y) column an upward ray from one cell below the
 * grid, moved off the cell centres; hits of one orientation less than 1e-9
 * apart count once; a cell is inside after an odd number of crossings. */
static void column_fill(const ob_mesh_room* room, const ob_grid* g, uint8_t* inside) {
    const int64_t nx = g->n[0], ny = g->n[1], nz = g->n[2];
    const double dx = g->dx;
    const double off[2] = {0.0001372 * dx, 0.0002918 * dx};
    const double start = g->origin[2] - dx;
#pragma omp parallel
    {
        ob_hit* hits = (ob_hit*)malloc(sizeof(ob_hit) * (size_t)(room->n_tri + 1));
        ob_hit* cross = (ob_hit*)malloc(sizeof(ob_hit) * (size_t)(room->n_tri + 1));
        int* up_of = (int*)malloc(sizeof(int) * (size_t)(room->n_tri + 1));
#pragma omp for schedule(dynamic, 16)
        for (int64_t col = 0; col < nx * ny; ++col) {
            int64_t x = col % nx, y = col / nx;
            double o[3] = {(g->origin[0] + (double)x * dx) + off[0], (g->origin[1] + (double)y * dx) + off[1],
                           start};
            const double up_dir[3] = {0.0, 0.0, 1.0};
            int32_t nh = all_hits(room, o, up_dir, 0.0, INFINITY, hits);
            int32_t nc = 0;
            for (int32_t h = 0; h < nh; ++h) {
                const double* a = tri_vertex(room, hits[h].k, 0);
                const double* b = tri_vertex(room, hits[h].k, 1);
                const double* c = tri_vertex(room, hits[h].k, 2);
                double e1[3], e2[3], n[3];
                v3sub(b, a, e1);
                v3sub(c, a, e2);
                v3cross(e1, e2, n);
                int up = n[2] > 0.0;
                if (nc > 0 && hits[h].t - cross[nc - 1].t < 1e-9 && up_of[nc - 1] == up) continue;
                cross[nc] = hits[h];
                up_of[nc] = up;
                nc++;
            }
            int32_t k = 0;
            for (int64_t z = 0; z < nz; ++z) {
                double tz = (g->origin[2] + (double)z * dx) - start;
                while (k < nc && cross[k].t < tz) k++;
                inside[z * nx * ny + col] = (uint8_t)(k % 2 == 1);
            }
        }
        free(hits);
        free(cross);
        free(up_of);
    }
}

/* fdtd.rs:289-377 (build_dx) for a general room with rigid solids. */
int ob_build_room(const ob_mesh_room* room, const ob_solid* solids, int32_t n_solids, double c,
                  double dx, ob_grid* g) {
    memset(g, 0, sizeof(*g));
    if (room->n_pos <= 0) return -4;
    double dt = COURANT * dx / c;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t i = 0; i < room->n_pos; ++i)
        for (int r = 0; r < 3; ++r) {
            lo[r] = fmin(lo[r], room->pos[3 * i + r]);
            hi[r] = fmax(hi[r], room->pos[3 * i + r]);
        }
    int64_t n[3];
    for (int k = 0; k < 3; ++k) {
        g->origin[k] = lo[k] - dx * 0.5;
        n[k] = (int64_t)ceil((hi[k] - lo[k]) / dx) + 2;
        g->n[k] = n[k];
    }
    int64_t cells = n[0] * n[1] * n[2];
    if (cells > MAX_CELLS) return -1;
    g->dx = dx;
    g->dt = dt;
    const int64_t nx = n[0], nxy = n[0] * n[1];
    uint8_t* solid = (uint8_t*)calloc((size_t)cells, 1);
    uint8_t* inside = (uint8_t*)calloc((size_t)cells, 1);
    uint8_t* air = (uint8_t*)calloc((size_t)cells, 1);
    int32_t* faces = (int32_t*)malloc(sizeof(int32_t) * 6 * (size_t)cells);
    g->flags = (uint8_t*)calloc((size_t)cells, 1);
    g->bidx = (uint32_t*)malloc((size_t)cells * 4);
    if (!solid || !inside || !air || !faces || !g->flags || !g->bidx) return -2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < cells; ++i) {
        int64_t z = i / nxy, r = i % nxy, y = r / nx, x = r % nx;
        double p[3] = {g->origin[0] + (double)x * dx, g->origin[1] + (double)y * dx, g->origin[2] + (double)z * dx};
        for (int32_t s = 0; s < n_solids && !solid[i]; ++s) solid[i] = (uint8_t)ob_solid_contains(&solids[s], p);
        if (room->prism) inside[i] = (uint8_t)prism_contains(room, p);
    }
    if (!room->prism) column_fill(room, g, inside);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < cells; ++i) {
        int64_t z = i / nxy, r = i % nxy, y = r / nx, x = r % nx;
        air[i] = !solid[i] && x > 0 && y > 0 && z > 0 && x < n[0] - 1 && y < n[1] - 1 && z < n[2] - 1 && inside[i];
    }
    static const int64_t DXI[6] = {-1, 1, 0, 0, 0, 0}, DYI[6] = {0, 0, -1, 1, 0, 0}, DZI[6] = {0, 0, 0, 0, -1, 1};
    /* Per air node: flags and the surface behind each missing face (-1: none). */
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < cells; ++i) {
        if (!air[i]) continue;
        int64_t z = i / nxy, r = i % nxy, y = r / nx, x = r % nx;
        double p[3] = {g->origin[0] + (double)x * dx, g->origin[1] + (double)y * dx, g->origin[2] + (double)z * dx};
        uint8_t fl = AIR;
        for (int k = 0; k < 6; ++k) {
            int64_t j = (z + DZI[k]) * nxy + (y + DYI[k]) * nx + (x + DXI[k]);
            faces[6 * i + k] = -1;
            if (air[j]) {
                fl |= (uint8_t)(1u << k);
            } else if (!solid[j]) {
                double d[3] = {(double)DXI[k], (double)DYI[k], (double)DZI[k]};
                int32_t h = closest_hit(room, p, d, 0.0, dx * 1.5);
                if (h >= 0) faces[6 * i + k] = room->tri_surface[h];
            }
        }
        g->flags[i] = fl;
    }
    /* Serial, in cell order, as fdtd.rs:368-375 builds the boundary list. */
    int64_t cap_b = 1024, cap_s = 4096;
    g->b_cell = (uint32_t*)malloc(cap_b * 4);
    g->b_start = (uint32_t*)malloc((cap_b + 1) * 4);
    g->slots = (ob_slot*)malloc(cap_s * sizeof(ob_slot));
    int rc = 0;
    for (int64_t i = 0; i < cells && rc == 0; ++i) {
        g->bidx[i] = NO_BOUNDARY;
        if (!air[i]) continue;
        int32_t fs[6];
        int nf = 0;
        for (int k = 0; k < 6; ++k)
            if (faces[6 * i + k] >= 0) fs[nf++] = faces[6 * i + k];
        ob_slot sl[6 * (OB_MAX_BRANCHES + 1)];
        int ns = ob_boundary_slots(fs, nf, room->fits, dt, sl, 6 * (OB_MAX_BRANCHES + 1));
        if (ns < 0) { rc = -3; break; }
        if (ns == 0) continue;
        if (g->nb + 1 >= cap_b) {
            cap_b *= 2;
            g->b_cell = (uint32_t*)realloc(g->b_cell, cap_b * 4);
            g->b_start = (uint32_t*)realloc(g->b_start, (cap_b + 1) * 4);
        }
        while (g->ns + ns > cap_s) {
            cap_s *= 2;
            g->slots = (ob_slot*)realloc(g->slots, cap_s * sizeof(ob_slot));
        }
        g->bidx[i] = (uint32_t)g->nb;
        g->b_cell[g->nb] = (uint32_t)i;
        g->b_start[g->nb] = (uint32_t)g->ns;
        memcpy(g->slots + g->ns, sl, ns * sizeof(ob_slot));
        g->ns += ns;
        g->nb++;
    }
    g->b_start[g->nb] = (uint32_t)g->ns;
    free(solid);
    free(inside);
    free(air);
    free(faces);
    return rc;
}

/* ---------------------------------------------------------------- sources */

/* fdtd.rs:601-610 */
void ob_source_signal(double f_max, double dt, int64_t steps, double* out) {
    double sigma = 0.342 / f_max;
    double t0 = 6.0 * sigma;
    for (int64_t n = 0; n < steps; ++n) {
        double t = (double)n * dt - t0;
        out[n] = -t / (sigma * sigma) * exp(-(t * t) / (2.0 * sigma * sigma)) * dt;
    }
}

static int64_t gidx(const ob_grid* g, int64_t x, int64_t y, int64_t z) {
    return (z * g->n[1] + y) * g->n[0] + x;
}

/* fdtd.rs:245-272 */
int64_t ob_nearest_air(const ob_grid* g, const double p[3]) {
    double q[3];
    int64_t c[3];
    for (int k = 0; k < 3; ++k) {
        q[k] = (p[k] - g->origin[k]) / g->dx;
        c[k] = (int64_t)round(q[k]);
    }
    for (int64_t r = 0; r < 17; ++r) {
        int have = 0;
        double bd = 0.0;
        int64_t bi = -1;
        for (int64_t dz = -r; dz <= r; ++dz)
            for (int64_t dy = -r; dy <= r; ++dy)
                for (int64_t ddx = -r; ddx <= r; ++ddx) {
                    int64_t x = c[0] + ddx, y = c[1] + dy, z = c[2] + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= g->n[0] || y >= g->n[1] || z >= g->n[2]) continue;
                    int64_t i = gidx(g, x, y, z);
                    if (g->flags[i] & AIR) {
                        double ex = g->origin[0] + (double)x * g->dx - p[0];
                        double ey = g->origin[1] + (double)y * g->dx - p[1];
                        double ez = g->origin[2] + (double)z * g->dx - p[2];
                        double d = ex * ex + ey * ey + ez * ez;
                        if (!have || d < bd) { have = 1; bd = d; bi = i; }
                    }
                }
        if (have) return bi;
    }
    return -1;
}

/* fdtd.rs:497-512: from p along axis in steps of dx/20 (up to 2 m) to the
 * first point whose nearest node is air; the face halfway back from there. */
void ob_baffle_point(const ob_grid* g, const double p[3], const double axis[3], double out[3]) {
    const double step = g->dx * 0.05;
    const int64_t kmax = (int64_t)(2.0 / step);
    for (int64_t k = 0; k < kmax; ++k) {
        double q[3], r[3];
        int neg = 0;
        for (int i = 0; i < 3; ++i) {
            q[i] = p[i] + axis[i] * ((double)k * step);
            r[i] = round((q[i] - g->origin[i]) / g->dx);
            neg |= r[i] < 0.0;
        }
        if (neg) continue;
        int64_t x = (int64_t)r[0], y = (int64_t)r[1], z = (int64_t)r[2];
        if (x < g->n[0] && y < g->n[1] && z < g->n[2] && (g->flags[gidx(g, x, y, z)] & AIR)) {
            double m = fmin(g->dx * 0.5, (double)k * step);
            for (int i = 0; i < 3; ++i) out[i] = q[i] - axis[i] * m;
            return;
        }
    }
    for (int i = 0; i < 3; ++i) out[i] = p[i];
}

/* fdtd.rs:516-533 (FdtdGrid::air_along): from p along axis in steps of
 * dx/2 (up to 2 m) the first point whose rounded node is air, else the
 * nearest air node; -1 for None. */
int64_t ob_air_along(const ob_grid* g, const double p[3], const double axis[3]) {
    const double step = g->dx * 0.5;
    const int64_t kmax = (int64_t)(2.0 / step);
    for (int64_t k = 0; k < kmax; ++k) {
        double r[3];
        int neg = 0;
        for (int i = 0; i < 3; ++i) {
            double q = p[i] + axis[i] * ((double)k * step);
            r[i] = round((q - g->origin[i]) / g->dx);
            neg |= r[i] < 0.0;
        }
        if (neg) continue;
        int64_t x = (int64_t)r[0], y = (int64_t)r[1], z = (int64_t)r[2];
        if (x < g->n[0] && y < g->n[1] && z < g->n[2] && (g->flags[gidx(g, x, y, z)] & AIR)) return gidx(g, x, y, z);
    }
    return ob_nearest_air(g, p);
}

/* run.rs:601-624: the source node (calibrated: air_along first), the driver
 * point (calibrated: baffle_point, else the centre), and the component taps:
 * a calibrated or order-0 source is taps_at(driver) (or the node alone);
 * otherwise source_components(node, order), or the node alone (*fallback =
 * 1). cells/w are [8]; len[4]. Returns the number of components, 0 when no
 * air node is found. */
int ob_wave_source_taps(const ob_grid* g, const double center[3], const double axis[3], int calibrated, int order,
                        int64_t* cells, double* w, int* len, int* fallback) {
    int64_t node = calibrated ? ob_air_along(g, center, axis) : -1;
    if (node < 0) node = ob_nearest_air(g, center);
    if (node < 0) node = ob_air_along(g, center, axis);
    if (node < 0) return 0;
    *fallback = 0;
    if (calibrated || order == 0) {
        double driver[3] = {center[0], center[1], center[2]};
        if (calibrated) ob_baffle_point(g, center, axis, driver);
        int nt = ob_taps_at(g, driver, cells, w);
        if (nt == 0) {
            cells[0] = node;
            w[0] = 1.0;
            nt = 1;
        }
        len[0] = nt;
        return 1;
    }
    int nc = ob_source_components(g, node, order, cells, w, len);
    if (nc == 0) {
        cells[0] = node;
        w[0] = 1.0;
        len[0] = 1;
        *fallback = 1;
        nc = 1;
    }
    return nc;
}

/* fdtd.rs:847-862: power-complementary crossfade of two complex fields. */
void ob_blend(const double ism[2], const double wave[2], double w, double out[2]) {
    if (w <= 0.0) { out[0] = ism[0]; out[1] = ism[1]; return; }
    if (w >= 1.0) { out[0] = wave[0]; out[1] = wave[1]; return; }
    double ni = ism[0] * ism[0] + ism[1] * ism[1], nw = wave[0] * wave[0] + wave[1] * wave[1];
    double target = (1.0 - w) * ni + w * nw;
    double pr = ism[0] * (1.0 - w) + wave[0] * w, pi = ism[1] * (1.0 - w) + wave[1] * w;
    double m2 = pr * pr + pi * pi;
    if (m2 < 1e-6 * target) {
        const double* d = nw * w >= ni * (1.0 - w) ? wave : ism;
        double n = fmax(hypot(d[0], d[1]), 1e-300), st = sqrt(target);
        out[0] = d[0] / n * st;
        out[1] = d[1] / n * st;
        return;
    }
    double s = sqrt(target / m2);
    out[0] = pr * s;
    out[1] = pi * s;
}

/* fdtd.rs:467-492 */
int ob_taps_at(const ob_grid* g, const double p[3], int64_t* cells, double* weights) {
    double q[3], f[3], b[3];
    for (int k = 0; k < 3; ++k) {
        q[k] = (p[k] - g->origin[k]) / g->dx;
        b[k] = floor(q[k]);
        f[k] = q[k] - b[k];
    }
    int nt = 0;
    for (int ix = 0; ix < 2; ++ix) {
        double wx = ix ? f[0] : 1.0 - f[0];
        for (int iy = 0; iy < 2; ++iy) {
            double wy = iy ? f[1] : 1.0 - f[1];
            for (int iz = 0; iz < 2; ++iz) {
                double wz = iz ? f[2] : 1.0 - f[2];
                int64_t x = (int64_t)b[0] + ix, y = (int64_t)b[1] + iy, z = (int64_t)b[2] + iz;
                if (x < 0 || y < 0 || z < 0 || x >= g->n[0] || y >= g->n[1] || z >= g->n[2]) continue;
                int64_t i = gidx(g, x, y, z);
                double w = wx * wy * wz;
                if ((g->flags[i] & AIR) && w > 1e-9) {
                    cells[nt] = i;
                    weights[nt] = w;
                    nt++;
                }
            }
        }
    }
    double sum = 0.0;
    for (int t = 0; t < nt; ++t) sum += weights[t];
    if (sum < 0.25) {
        int64_t i = ob_nearest_air(g, p);
        if (i < 0) return 0;
        cells[0] = i;
        weights[0] = 1.0;
        return 1;
    }
    for (int t = 0; t < nt; ++t) weights[t] = weights[t] / sum;
    return nt;
}

/* fdtd.rs:730-743 */
int ob_source_components(const ob_grid* g, int64_t c, int order, int64_t* cc, double* cw, int* len) {
    int64_t N = g->n[0] * g->n[1] * g->n[2];
    cc[0] = c;
    cw[0] = 1.0;
    len[0] = 1;
    if (order < 1) return 1;
    int64_t steps[3] = {1, g->n[0], g->n[0] * g->n[1]};
    for (int s = 0; s < 3; ++s) {
        int64_t plus = c + steps[s];
        if (c < steps[s]) return 0;
        int64_t minus = c - steps[s];
        if (plus >= N || !(g->flags[plus] & AIR) || !(g->flags[minus] & AIR)) return 0;
        cc[2 * (s + 1)] = plus;
        cw[2 * (s + 1)] = 1.0;
        cc[2 * (s + 1) + 1] = minus;
        cw[2 * (s + 1) + 1] = -1.0;
        len[s + 1] = 2;
    }
    return 4;
}

/* fdtd.rs:753-760: the N = 400 golden-spiral directions (local frame). */
void ob_first_order_directions(double* out) {
    const int N = 400;
    const double golden = PI * (3.0 - sqrt(5.0));
    for (int i = 0; i < N; ++i) {
        double z = 1.0 - 2.0 * ((double)i + 0.5) / (double)N;
        double r = sqrt(1.0 - z * z);
        out[3 * i] = r * cos(golden * (double)i);
        out[3 * i + 1] = r * sin(golden * (double)i);
        out[3 * i + 2] = z;
    }
}

/* Rust's f64::total_cmp > 0 for |a| vs |b| as in max_by (the last maximum wins). */
static void fo_solve(const double m0[4][4], const cplx rhs[4], cplx out[4]) {
    double m[4][4];
    cplx x[4];
    memcpy(m, m0, sizeof(m));
    for (int i = 0; i < 4; ++i) x[i] = rhs[i];
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (total_cmp(fabs(m[r][col]), fabs(m[piv][col])) >= 0) piv = r;
        for (int c = 0; c < 4; ++c) {
            double t = m[col][c];
            m[col][c] = m[piv][c];
            m[piv][c] = t;
        }
        cplx tx = x[col];
        x[col] = x[piv];
        x[piv] = tx;
        for (int row = 0; row < 4; ++row) {
            if (row == col) continue;
            double k = m[row][col] / m[col][col];
            for (int c = 0; c < 4; ++c) m[row][c] -= k * m[col][c];
            x[row] = cmk(x[row].re - x[col].re * k, x[row].im - x[col].im * k);
        }
    }
    for (int i = 0; i < 4; ++i) out[i] = cmk(x[i].re / m[i][i], x[i].im / m[i][i]);
}

/* fdtd.rs:749-807 (fit_first_order): d [400] complex directivity at the
 * spiral directions, d_axis the on-axis (+x local) value, to_room row-major.
 * out: a0 (2), a (6), returns the relative RMS residual. */
double ob_fit_first_order(const double* d, const double d_axis[2], const double to_room[9], double a0_out[2],
                          double a_out[6]) {
    const int N = 400;
    double dirs[400 * 3];
    ob_first_order_directions(dirs);
    double m[4][4];
    memset(m, 0, sizeof(m));
    cplx rhs[4];
    double bs[400][4];
    for (int a = 0; a < 4; ++a) rhs[a] = cmk(0.0, 0.0);
    for (int i = 0; i < N; ++i) {
        const double* v = dirs + 3 * i;
        double room[3];
        for (int r = 0; r < 3; ++r) room[r] = (to_room[3 * r] * v[0] + to_room[3 * r + 1] * v[1]) + to_room[3 * r + 2] * v[2];
        double b[4] = {1.0, room[0], room[1], room[2]};
        for (int a = 0; a < 4; ++a) {
            for (int c = 0; c < 4; ++c) m[a][c] += b[a] * b[c];
            rhs[a] = cadd(rhs[a], cscale(cmk(d[2 * i], d[2 * i + 1]), b[a]));
            bs[i][a] = b[a];
        }
    }
    cplx ls[4], me[4], e4[4];
    fo_solve(m, rhs, ls);
    const double x[3] = {1.0, 0.0, 0.0};
    double axis[3];
    for (int r = 0; r < 3; ++r) axis[r] = (to_room[3 * r] * x[0] + to_room[3 * r + 1] * x[1]) + to_room[3 * r + 2] * x[2];
    const double e[4] = {1.0, axis[0], axis[1], axis[2]};
    cplx got = cmk(0.0, 0.0);
    for (int i = 0; i < 4; ++i) got = cadd(got, cscale(ls[i], e[i]));
    for (int i = 0; i < 4; ++i) e4[i] = cmk(e[i], 0.0);
    fo_solve(m, e4, me);
    double ete = -0.0;
    for (int i = 0; i < 4; ++i) ete += me[i].re * e[i];
    const cplx diff = cmk(d_axis[0] - got.re, d_axis[1] - got.im);
    const cplx q = cmk(diff.re / ete, diff.im / ete);
    cplx a[4];
    for (int i = 0; i < 4; ++i) a[i] = cadd(ls[i], cmul(me[i], q));
    double err = 0.0, tot = 0.0;
    for (int i = 0; i < N; ++i) {
        cplx fit = cadd(cadd(cadd(a[0], cscale(a[1], bs[i][1])), cscale(a[2], bs[i][2])), cscale(a[3], bs[i][3]));
        cplx dv = cmk(d[2 * i], d[2 * i + 1]);
        err += cnorm2(cmk(fit.re - dv.re, fit.im - dv.im));
        tot += cnorm2(dv);
    }
    a0_out[0] = a[0].re;
    a0_out[1] = a[0].im;
    for (int i = 0; i < 3; ++i) {
        a_out[2 * i] = a[i + 1].re;
        a_out[2 * i + 1] = a[i + 1].im;
    }
    return sqrt(err / fmax(tot, 1e-30));
}

/* fdtd.rs:812-815 */
void ob_component_weights(const double a0[2], const double a[6], double k, double dx, double out[8]) {
    cplx dip = cmk(0.0, 2.0 * sin(k * dx));
    out[0] = a0[0];
    out[1] = a0[1];
    for (int i = 0; i < 3; ++i) {
        cplx v = cdiv(cmk(a[2 * i], a[2 * i + 1]), dip);
        out[2 + 2 * i] = v.re;
        out[3 + 2 * i] = v.im;
    }
}

/* Σ_n from_polar(x_n, -2π fract(f dt n)), summed in n order (fdtd.rs:680-681, :824-828). */
static cplx dft_at(const double* x, int64_t steps, double f, double dt) {
    cplx acc = cmk(0.0, 0.0);
    for (int64_t n = 0; n < steps; ++n) {
        double ph = -2.0 * PI * fract(f * dt * (double)n);
        acc = cadd(acc, from_polar(x[n], ph));
    }
    return acc;
}

/* fdtd.rs:819-840 */
void ob_to_transfer(const double* spectra, int64_t n_spec, const double* signal, int64_t steps,
                    double dt, double dx, const double* freqs, int64_t nf, double c, double r_ref,
                    const double* h0, double* out) {
    cplx* s = (cplx*)malloc((size_t)nf * sizeof(cplx));
    for (int64_t fi = 0; fi < nf; ++fi) s[fi] = dft_at(signal, steps, freqs[fi], dt);
    double scale = dt * dt / (dx * dx * dx) * 4.0 * PI * c * c * r_ref;
    for (int64_t i = 0; i < n_spec; ++i) {
        int64_t fi = i % nf;
        cplx v = cmk(spectra[2 * i], spectra[2 * i + 1]);
        cplx t = cmul(cscale(cdiv(v, s[fi]), scale), cmk(h0[2 * fi], h0[2 * fi + 1]));
        out[2 * i] = t.re;
        out[2 * i + 1] = t.im;
    }
    free(s);
}

/* ---------------------------------------------------------------- time loop */

static int run_loop(const ob_grid_view* g, const int64_t* tap_cells, const double* tap_w,
                    int64_t n_taps, const double* signal, int64_t steps, const int64_t* recv,
                    int64_t n_recv, double* rec, int threads, const volatile int32_t* cancel,
                    double* p_out, double* prev_out, double* state_out) {
    const int64_t nx = g->n[0], nxy = g->n[0] * g->n[1], cells = nxy * g->n[2];
    const double l2 = COURANT * COURANT, lam = COURANT;
    const int64_t offs[6] = {-1, 1, -nx, nx, -nxy, nxy};
    double* p = (double*)calloc((size_t)cells, sizeof(double));
    double* prev = (double*)calloc((size_t)cells, sizeof(double));
    uint32_t* bidx = (uint32_t*)malloc((size_t)cells * 4);
    int64_t ns = g->nb ? g->b_start[g->nb] : 0;
    double* st = (double*)calloc((size_t)(2 * ns + 2), sizeof(double));
    double* bnext = (double*)calloc((size_t)(g->nb + 1), sizeof(double));
    if (!p || !prev || !bidx || !st || !bnext) return -2;
    for (int64_t i = 0; i < cells; ++i) bidx[i] = NO_BOUNDARY;
    for (int64_t b = 0; b < g->nb; ++b) bidx[g->b_cell[b]] = (uint32_t)b;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const uint8_t* flags = g->flags;
    int status = 0;
    for (int64_t n = 0; n < steps; ++n) {
        if (n % 64 == 0 && cancel && *cancel) { status = 1; break; }
        /* Record p^n (fdtd.rs:673-683). */
        if (rec)
            for (int64_t ri = 0; ri < n_recv; ++ri) rec[ri * steps + n] = p[recv[ri]];
        /* Interior (fdtd.rs:699-708). */
#pragma omp parallel for schedule(static)
        for (int64_t z = 0; z < g->n[2]; ++z) {
            int64_t base = z * nxy;
            for (int64_t k = 0; k < nxy; ++k) {
                int64_t i = base + k;
                if (!(flags[i] & AIR) || bidx[i] != NO_BOUNDARY) continue;
                uint8_t f = flags[i];
                double sum = 0.0, kair = 0.0;
                for (int d = 0; d < 6; ++d)
                    if (f & (1u << d)) { sum += p[i + offs[d]]; kair += 1.0; }
                double lap = l2 * (sum - kair * p[i]);
                prev[i] = 2.0 * p[i] - prev[i] + lap;
            }
        }
        /* Boundary nodes (fdtd.rs:710-713). */
#pragma omp parallel for schedule(static)
        for (int64_t b = 0; b < g->nb; ++b) {
            int64_t i = g->b_cell[b];
            uint8_t f = flags[i];
            double sum = 0.0, kair = 0.0;
            for (int d = 0; d < 6; ++d)
                if (f & (1u << d)) { sum += p[i + offs[d]]; kair += 1.0; }
            double lap = l2 * (sum - kair * p[i]);
            int64_t j0 = g->b_start[b], j1 = g->b_start[b + 1];
            bnext[b] = ob_boundary_update(g->slots + j0, (int)(j1 - j0), st + 2 * j0, p[i], prev[i],
                                          lap, lam, g->dt);
        }
        /* Scatter, then inject (fdtd.rs:714-720). */
        for (int64_t b = 0; b < g->nb; ++b) prev[g->b_cell[b]] = bnext[b];
        for (int64_t t = 0; t < n_taps; ++t) prev[tap_cells[t]] += tap_w[t] * signal[n];
        double* tmp = p;
        p = prev;
        prev = tmp;
    }
    if (p_out) memcpy(p_out, p, (size_t)cells * sizeof(double));
    if (prev_out) memcpy(prev_out, prev, (size_t)cells * sizeof(double));
    if (state_out) memcpy(state_out, st, (size_t)(2 * ns) * sizeof(double));
    free(p);
    free(prev);
    free(bidx);
    free(st);
    free(bnext);
    return status;
}

int ob_run(const ob_grid_view* g, const int64_t* tap_cells, const double* tap_w, int64_t n_taps,
           const double* signal, int64_t steps, const int64_t* recv, int64_t n_recv,
           const double* freqs, int64_t nf, double* ir_out, double* spec_out, int threads,
           const volatile int32_t* cancel) {
    double* rec = ir_out;
    int own = 0;
    if (!rec && spec_out && n_recv > 0) {
        rec = (double*)malloc((size_t)(n_recv * steps) * sizeof(double));
        own = 1;
    }
    int st = run_loop(g, tap_cells, tap_w, n_taps, signal, steps, recv, n_recv, rec, threads, cancel,
                      NULL, NULL, NULL);
    if (st == 0 && spec_out) {
        /* The per-step DFT of fdtd.rs:679-682, evaluated from the recorded
         * series in the same n order, so the sums are identical. */
#pragma omp parallel for schedule(dynamic)
        for (int64_t k = 0; k < n_recv * nf; ++k) {
            int64_t ri = k / nf, fi = k % nf;
            cplx a = dft_at(rec + ri * steps, steps, freqs[fi], g->dt);
            spec_out[2 * k] = a.re;
            spec_out[2 * k + 1] = a.im;
        }
    }
    if (own) free(rec);
    return st;
}

int ob_run_fields(const ob_grid_view* g, const int64_t* tap_cells, const double* tap_w,
                  int64_t n_taps, const double* signal, int64_t steps, int threads, double* p_out,
                  double* prev_out, double* state_out) {
    return run_loop(g, tap_cells, tap_w, n_taps, signal, steps, NULL, 0, NULL, threads, NULL, p_out,
                    prev_out, state_out);
}

/* --------------------------------------------------- cabinet calibration */

static ob_grid_view view_of(const ob_grid* g) {
    ob_grid_view v;
    for (int k = 0; k < 3; ++k) v.n[k] = g->n[k];
    v.dx = g->dx;
    v.dt = g->dt;
    v.flags = g->flags;
    v.nb = g->nb;
    v.b_cell = g->b_cell;
    v.b_start = g->b_start;
    v.slots = g->slots;
    return v;
}

/* fdtd.rs:890-955 (calibrate_cabinet), with run_cpu_with as the runner.
 * on_axis_out [nf] and dir_out [12][nf] are complex (re, im); angles_out
 * [12]. Returns 0, -1 when a step of the reference returns None, -2 on
 * allocation failure. */
int ob_calibrate_cabinet(const double* verts, int32_t n_verts, const int32_t* tris, int32_t n_tris,
                         const double ac[3], double c, double f_max, double ppw, double cells_across,
                         const double* freqs, int32_t nf, int threads, double* on_axis_out,
                         double* angles_out, double* dir_out) {
    double reach = 0.1;
    for (int32_t i = 0; i < n_verts; ++i) {
        double d[3];
        v3sub(verts + 3 * i, ac, d);
        reach = fmax(reach, sqrt(v3dot(d, d)));
    }
    const double f_cab = 0.5 * c / (2.0 * PI * reach);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t i = 0; i < n_verts; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = fmin(lo[k], verts[3 * i + k]);
            hi[k] = fmax(hi[k], verts[3 * i + k]);
        }
    const double ext = fmin(fmin(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    const double dx = fmin(c / (f_max * ppw), ext / cells_across);
    const double max_half = 0.5 * dx * 160.0;
    const double r_m = fmax(reach + 1.0, 3.0);
    const double half = fmax(fmin(0.5 * c / f_cab, max_half), r_m + 1.5);
    const double side = 2.0 * half;
    /* The absorbing box: a square plan extruded to `side`, matched walls. */
    double pos[8][3];
    for (int i = 0; i < 8; ++i) {
        pos[i][0] = (i & 1) ? side : 0.0;
        pos[i][1] = (i & 2) ? side : 0.0;
        pos[i][2] = (i & 4) ? side : 0.0;
    }
    static const int32_t box_tris[12][3] = {{0, 2, 3}, {0, 3, 1}, {4, 5, 7}, {4, 7, 6}, {0, 1, 5}, {0, 5, 4},
                                            {2, 6, 7}, {2, 7, 3}, {0, 4, 6}, {0, 6, 2}, {1, 3, 7}, {1, 7, 5}};
    int32_t tri_surf[12];
    for (int i = 0; i < 12; ++i) tri_surf[i] = i / 2;
    ob_fit fits[6];
    memset(fits, 0, sizeof(fits));
    for (int i = 0; i < 6; ++i) fits[i].beta0 = 1.0;
    const double plan[8] = {0.0, 0.0, side, 0.0, side, side, 0.0, side};
    const int32_t region_cycles[2] = {0, 1}, cycle_off[2] = {0, 4};
    ob_mesh_room room = {1, side, 8, &pos[0][0], 12, &box_tris[0][0], tri_surf, 6, fits, 1, region_cycles, cycle_off, plan};
    const double p0[3] = {half, half, half};
    const double rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double shift[3];
    v3sub(p0, ac, shift);
    double* planes = (double*)malloc(sizeof(double) * 4 * (size_t)(n_tris > 0 ? n_tris : 1));
    ob_solid cab;
    cab.n_planes = ob_solid_from_mesh(verts, n_verts, tris, n_tris, rot, shift, planes, cab.lo, cab.hi);
    cab.planes = planes;
    if (cab.n_planes < 0) { free(planes); return -1; }
    ob_grid with, bare;
    int rc = ob_build_room(&room, &cab, 1, c, dx, &with);
    if (rc == 0) rc = ob_build_room(&room, NULL, 0, c, dx, &bare);
    free(planes);
    if (rc != 0) return rc == -2 ? -2 : -1;
    double src[3];
    const double axis[3] = {1.0, 0.0, 0.0};
    ob_baffle_point(&with, p0, axis, src);
    int64_t tw_c[8], tb_c[8], rx_w[12], rx_b[12];
    double tw_w[8], tb_w[8];
    int ntw = ob_taps_at(&with, src, tw_c, tw_w), ntb = ob_taps_at(&bare, src, tb_c, tb_w);
    int ok = ntw > 0 && ntb > 0;
    for (int a = 0; a < 12 && ok; ++a) {
        angles_out[a] = (double)a * 30.0;
        const double r = angles_out[a] * (PI / 180.0);
        const double d[3] = {cos(r) * r_m, sin(r) * r_m, 0.0 * r_m};
        const double pt[3] = {p0[0] + d[0], p0[1] + d[1], p0[2] + d[2]};
        rx_w[a] = ob_nearest_air(&with, pt);
        rx_b[a] = ob_nearest_air(&bare, pt);
        ok = rx_w[a] >= 0 && rx_b[a] >= 0;
    }
    if (!ok) {
        ob_grid_free(&with);
        ob_grid_free(&bare);
        return -1;
    }
    const double duration = side * sqrt(3.0) * 2.0 / c + 12.0 * 0.342 / f_max;
    const int64_t steps = (int64_t)ceil(duration / with.dt);
    double* signal = (double*)malloc(sizeof(double) * (size_t)steps);
    double* sw = (double*)malloc(sizeof(double) * 2 * 12 * (size_t)nf);
    double* sb = (double*)malloc(sizeof(double) * 2 * 12 * (size_t)nf);
    ob_source_signal(f_max, with.dt, steps, signal);
    ob_grid_view vw = view_of(&with), vb = view_of(&bare);
    ob_run(&vw, tw_c, tw_w, ntw, signal, steps, rx_w, 12, freqs, nf, NULL, sw, threads, NULL);
    ob_run(&vb, tb_c, tb_w, ntb, signal, steps, rx_b, 12, freqs, nf, NULL, sb, threads, NULL);
    /* Below f_cab the measured ratio fades toward 1 as (f / f_cab)^2. */
    for (int32_t f = 0; f < nf; ++f) {
        const double x = freqs[f] / f_cab;
        const double w = fmin(x * x, 1.0);
        for (int r = 0; r < 12; ++r) {
            cplx m = cdiv(cmk(sw[2 * (r * nf + f)], sw[2 * (r * nf + f) + 1]),
                          cmk(sb[2 * (r * nf + f)], sb[2 * (r * nf + f) + 1]));
            cplx q = from_polar(pow(hypot(m.re, m.im), w), atan2(m.im, m.re) * w);
            dir_out[2 * (r * nf + f)] = q.re;
            dir_out[2 * (r * nf + f) + 1] = q.im;
        }
        on_axis_out[2 * f] = dir_out[2 * f];
        on_axis_out[2 * f + 1] = dir_out[2 * f + 1];
        const cplx on = cmk(on_axis_out[2 * f], on_axis_out[2 * f + 1]);
        for (int r = 0; r < 12; ++r) {
            cplx q = cdiv(cmk(dir_out[2 * (r * nf + f)], dir_out[2 * (r * nf + f) + 1]), on);
            dir_out[2 * (r * nf + f)] = q.re;
            dir_out[2 * (r * nf + f) + 1] = q.im;
        }
    }
    free(signal);
    free(sw);
    free(sb);
    ob_grid_free(&with);
    ob_grid_free(&bare);
    return 0;
}
