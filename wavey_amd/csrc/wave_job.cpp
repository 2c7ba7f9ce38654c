// The wave solver's part of a wave job: every source's runs on one grid,
// combined into per-source transfer functions, as run.rs:593-661 does.
//
// A source is one or more runs (its components: an omni source's taps, or a
// directional source's monopole and three axis dipoles,
// fdtd.rs:730-743); each run's receiver spectra go through to_transfer
// (fdtd.rs:819-840), is divided by the cabinet calibration's on-axis
// response when the source has one (run.rs:651-656), and is summed with the
// source's per-frequency complex component weights (fdtd.rs:812-815,
// run.rs:625-661). The reference runs
// the components one after another; here they are either run the same way
// on one plan (the grid uploaded once) or handed to one batch (every run in
// one launch per step), which gives the same results bit for bit.
//
// Built on the library's public C ABI only.
#include <complex>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "wavey_fdtd.h"
#include "wv_internal.h"

using namespace wv;

namespace {

struct Run {
    size_t source, comp;
    std::vector<wv_tap> taps;
};

std::vector<Run> runs_of(const wv_wave_source* sources, size_t n_sources) {
    std::vector<Run> out;
    for (size_t s = 0; s < n_sources; ++s)
        for (uint32_t c = 0; c < sources[s].n_comps; ++c) {
            Run r{s, c, {}};
            r.taps.assign(sources[s].taps + sources[s].comp_off[c], sources[s].taps + sources[s].comp_off[c + 1]);
            out.push_back(std::move(r));
        }
    return out;
}

// Adds run r's transfer function, weighted, to its source's output.
void accumulate(const wv_wave_source& src, const Run& r, const std::vector<double>& spectra, const double* signal,
                size_t steps, const wv_grid* g, const double* freqs, size_t nf, double c, size_t n_recv,
                double* out) {
    std::vector<double> h0(2 * nf), tc(2 * n_recv * nf);
    for (size_t f = 0; f < nf; ++f) {
        h0[2 * f] = src.h0 ? src.h0[2 * f] : 1.0;
        h0[2 * f + 1] = src.h0 ? src.h0[2 * f + 1] : 0.0;
    }
    wv_to_transfer(spectra.data(), n_recv * nf, signal, steps, g->dt, g->dx, freqs, nf, c, src.r_ref, h0.data(),
                   tc.data());
    for (size_t k = 0; k < n_recv * nf; ++k) {
        const size_t f = k % nf;
        if (src.on_axis) {
            // run.rs:651-656: calibrated so the cabinet's free-field on-axis
            // response is the measured one; the quotient in the textbook form
            // (a c + b d, b c - a d) / |z|^2.
            const double cr = src.on_axis[2 * f], ci = src.on_axis[2 * f + 1];
            const double n = cr * cr + ci * ci;
            const double re = tc[2 * k] * cr + tc[2 * k + 1] * ci, im = tc[2 * k + 1] * cr - tc[2 * k] * ci;
            tc[2 * k] = re / n;
            tc[2 * k + 1] = im / n;
        }
        std::complex<double> w(1.0, 0.0);
        if (src.weights) {
            const double* wp = src.weights + 2 * (f * src.n_comps + r.comp);
            w = {wp[0], wp[1]};
        }
        const std::complex<double> v = w * std::complex<double>(tc[2 * k], tc[2 * k + 1]);
        out[2 * k] += v.real();
        out[2 * k + 1] += v.imag();
    }
}

}  // namespace

static wv_status wave_job_impl(const wv_grid* grid, const wv_wave_source* sources, size_t n_sources,
                               const double* signal, size_t steps, const uint64_t* receivers, size_t n_recv,
                               const double* freqs, size_t nf, double c, int precision, int batched,
                               const volatile int32_t* cancel, double* transfer_out, char* err, size_t err_len) {
    if (!grid || (n_sources && !sources) || !signal || !steps || (n_recv && !receivers) || (nf && !freqs) ||
        !transfer_out) {
        set_err(err, err_len, "wv_wave_job: invalid argument");
        return WV_ERR_INVALID;
    }
    for (size_t s = 0; s < n_sources; ++s)
        if (sources[s].n_comps == 0 || !sources[s].taps || !sources[s].comp_off ||
            (sources[s].n_comps > 1 && !sources[s].weights)) {
            set_err(err, err_len, "wv_wave_job: a source needs its components (and weights when it has several)");
            return WV_ERR_INVALID;
        }
    std::memset(transfer_out, 0, n_sources * n_recv * nf * 2 * sizeof(double));
    const std::vector<Run> runs = runs_of(sources, n_sources);
    const size_t per = n_recv * nf;
    std::vector<double> spectra(2 * per);
    if (!batched) {
        // One plan: the grid is uploaded once, the runs go one after another.
        wv_plan* plan = nullptr;
        wv_status st = wv_plan_create(grid, precision, &plan, err, err_len);
        if (st != WV_OK) return st;
        for (const Run& r : runs) {
            st = wv_plan_run(plan, r.taps.data(), r.taps.size(), signal, steps, receivers, n_recv, freqs, nf,
                             WV_MODE_SPECTRA, cancel, nullptr, nullptr, spectra.data(), nullptr, err, err_len);
            if (st != WV_OK) break;
            accumulate(sources[r.source], r, spectra, signal, steps, grid, freqs, nf, c, n_recv,
                       transfer_out + 2 * per * r.source);
        }
        wv_plan_free(plan);
        return st;
    }
    // One batch: every run advances in the same launches.
    wv_batch* b = nullptr;
    wv_status st = wv_batch_create(precision, &b, err, err_len);
    if (st != WV_OK) return st;
    std::unique_ptr<wv_batch, void (*)(wv_batch*)> guard(b, wv_batch_free);
    std::vector<int32_t> ids(runs.size());
    for (size_t i = 0; i < runs.size(); ++i) {
        st = wv_batch_add(b, grid, runs[i].taps.data(), runs[i].taps.size(), signal, steps, receivers, n_recv,
                          &ids[i], err, err_len);
        if (st != WV_OK) return st;
    }
    st = wv_batch_run(b, 1, cancel, nullptr, nullptr, err, err_len);
    if (st != WV_OK) return st;
    for (size_t i = 0; i < runs.size(); ++i) {
        st = wv_batch_result(b, ids[i], freqs, nf, WV_MODE_SPECTRA, spectra.data(), nullptr, err, err_len);
        if (st != WV_OK) return st;
        accumulate(sources[runs[i].source], runs[i], spectra, signal, steps, grid, freqs, nf, c, n_recv,
                   transfer_out + 2 * per * runs[i].source);
    }
    return WV_OK;
}

extern "C" wv_status wv_wave_job(const wv_grid* grid, const wv_wave_source* sources, size_t n_sources,
                                 const double* signal, size_t steps, const uint64_t* receivers, size_t n_recv,
                                 const double* freqs, size_t nf, double c, int precision, int batched,
                                 const volatile int32_t* cancel, double* transfer_out, char* err, size_t err_len) {
    return host_guarded(err, err_len, [&]() {
        return wave_job_impl(grid, sources, n_sources, signal, steps, receivers, n_recv, freqs, nf, c, precision,
                             batched, cancel, transfer_out, err, err_len);
    });
}
