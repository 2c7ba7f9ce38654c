// Result storage of a solve, as `wavey simulate` writes it
// (core/wavey-cli/src/main.rs:513-516 -> SimulationResult::save,
// core/wavey-solver/src/result.rs:444-461): results.json holds the metadata
// (serde's field order of SimulationResult, MicResult, Microphone, RunStats
// and FdtdSettings, with the skipped transfer and late fields left out) and
// results.bin the f32 data: every microphone's transfer functions
// [mic][speaker][freq] as complex f32 (re, im), then the late energies when
// the result has them. This build fills the wave solver's part: no cut
// planes, no late field (has_late false), so results.bin is the transfers.
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <system_error>
#include <vector>

#include "wavey_fdtd.h"
#include "wv_internal.h"

using namespace wv;

namespace {

// A JSON number: the shortest decimal that reads back as the same double;
// non-finite values are null (serde_json writes NaN and infinities as null).
void num(std::string& o, double v) {
    if (!std::isfinite(v)) {
        o += "null";
        return;
    }
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof buf, v);
    std::string s(buf, r.ptr);
    // A whole number keeps a fraction digit, so it reads back as a float.
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    o += s;
}

void num(std::string& o, uint64_t v) { o += std::to_string(v); }

// A JSON string: quote, backslash and control characters escaped; other
// bytes (UTF-8) as they are.
void str(std::string& o, const char* s) {
    o += '"';
    for (const unsigned char* p = (const unsigned char*)(s ? s : ""); *p; ++p) {
        switch (*p) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            case '\b': o += "\\b"; break;
            case '\f': o += "\\f"; break;
            default:
                if (*p < 0x20) {
                    char u[8];
                    std::snprintf(u, sizeof u, "\\u%04x", *p);
                    o += u;
                } else {
                    o += (char)*p;
                }
        }
    }
    o += '"';
}

void key(std::string& o, const char* k, bool first = false) {
    if (!first) o += ',';
    str(o, k);
    o += ':';
}

void bands(std::string& o, const double* v) {
    o += '[';
    for (int b = 0; b < WV_RESULT_BANDS; ++b) {
        if (b) o += ',';
        num(o, v ? v[b] : 0.0);
    }
    o += ']';
}

bool write_file(const std::filesystem::path& p, const void* data, size_t n, std::string& why) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) {
        why = p.string() + ": " + std::strerror(errno);
        return false;
    }
    const bool ok = (n == 0 || std::fwrite(data, 1, n, f) == n);
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) why = p.string() + ": write failed";
    return ok && closed;
}

}  // namespace

static wv_status results_save_impl(const char* dir, const wv_results* r, char* err, size_t err_len) {
    if (!dir || !r || (r->nf && !r->freqs) || (r->n_speakers && !r->speaker_ids) || (r->n_mics && !r->mics) ||
        (r->n_mics && r->n_speakers && r->nf && !r->transfer) || (r->n_notes && !r->notes)) {
        set_err(err, err_len, "wv_results_save: invalid argument");
        return WV_ERR_INVALID;
    }
    // result.rs:444-445: the directory and its parents.
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) {
        set_err(err, err_len, std::string("wv_results_save: ") + dir + ": " + ec.message());
        return WV_ERR_INVALID;
    }
    // results.json: SimulationResult (result.rs:16-38) in field order.
    std::string o;
    o += '{';
    key(o, "freqs", true);
    o += '[';
    for (size_t f = 0; f < r->nf; ++f) {
        if (f) o += ',';
        num(o, r->freqs[f]);
    }
    o += ']';
    key(o, "speaker_ids");
    o += '[';
    for (size_t s = 0; s < r->n_speakers; ++s) {
        if (s) o += ',';
        num(o, r->speaker_ids[s]);
    }
    o += ']';
    key(o, "mics");
    o += '[';
    for (size_t m = 0; m < r->n_mics; ++m) {
        // MicResult (result.rs:40-56): mic, late_gain, echograms; the
        // transfer and late fields are skipped (they go to results.bin).
        const wv_microphone& mic = r->mics[m];
        if (m) o += ',';
        o += '{';
        key(o, "mic", true);
        o += '{';
        key(o, "id", true);
        num(o, mic.id);
        key(o, "name");
        str(o, mic.name);
        key(o, "position");
        o += '[';
        for (int k = 0; k < 3; ++k) {
            if (k) o += ',';
            num(o, mic.position[k]);
        }
        o += "]}";
        key(o, "late_gain");
        o += "[]";
        key(o, "echograms");
        o += "[]}";
    }
    o += ']';
    key(o, "planes");
    o += "[]";
    // RunStats (result.rs:80-90).
    key(o, "stats");
    o += '{';
    key(o, "backend", true);
    str(o, r->backend);
    key(o, "seconds");
    num(o, r->seconds);
    key(o, "images");
    num(o, r->images);
    key(o, "paths");
    num(o, r->paths);
    key(o, "diffracted_paths");
    num(o, r->diffracted_paths);
    key(o, "receivers");
    num(o, r->receivers);
    key(o, "notes");
    o += '[';
    for (size_t n = 0; n < r->n_notes; ++n) {
        if (n) o += ',';
        str(o, r->notes[n]);
    }
    o += "]}";
    key(o, "fingerprint");
    num(o, r->fingerprint);
    key(o, "rt_sabine");
    bands(o, r->rt_sabine);
    key(o, "rt_eyring");
    bands(o, r->rt_eyring);
    key(o, "has_late");
    o += "false";
    key(o, "fdtd");
    if (!r->fdtd) {
        o += "null";
    } else {
        // FdtdSettings (fdtd.rs:85-114); accuracy is left out when High.
        const wv_fdtd_settings& s = *r->fdtd;
        o += '{';
        key(o, "crossover_hz", true);
        num(o, s.crossover_hz);
        key(o, "duration_s");
        num(o, s.duration_s);
        key(o, "points_per_wavelength");
        num(o, s.points_per_wavelength);
        key(o, "source_order");
        num(o, (uint64_t)s.source_order);
        key(o, "cabinets");
        o += s.cabinets ? "true" : "false";
        if (s.accuracy == WV_ACCURACY_DRAFT || s.accuracy == WV_ACCURACY_STANDARD) {
            key(o, "accuracy");
            str(o, s.accuracy == WV_ACCURACY_DRAFT ? "draft" : "standard");
        }
        o += '}';
    }
    o += '}';
    std::string why;
    const std::filesystem::path d(dir);
    if (!write_file(d / "results.json", o.data(), o.size(), why)) {
        set_err(err, err_len, "wv_results_save: " + why);
        return WV_ERR_INVALID;
    }
    // results.bin (result.rs:447-460): the transfers as complex f32.
    const size_t n = r->n_mics * r->n_speakers * r->nf;
    std::vector<float> bin(2 * n);
    for (size_t i = 0; i < 2 * n; ++i) bin[i] = (float)r->transfer[i];
    if (!write_file(d / "results.bin", bin.data(), bin.size() * sizeof(float), why)) {
        set_err(err, err_len, "wv_results_save: " + why);
        return WV_ERR_INVALID;
    }
    return WV_OK;
}

extern "C" wv_status wv_results_save(const char* dir, const wv_results* r, char* err, size_t err_len) {
    return host_guarded(err, err_len, [&]() { return results_save_impl(dir, r, err, err_len); });
}

static wv_status results_read_transfer_impl(const char* dir, size_t n_mics, size_t n_speakers, size_t nf,
                                            float* out, char* err, size_t err_len) {
    if (!dir || (n_mics * n_speakers * nf != 0 && !out)) {
        set_err(err, err_len, "wv_results_read_transfer: invalid argument");
        return WV_ERR_INVALID;
    }
    // result.rs:463-476: each microphone's [speaker][freq] complex f32, in
    // microphone order, from the start of results.bin.
    const std::string p = (std::filesystem::path(dir) / "results.bin").string();
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) {
        set_err(err, err_len, "wv_results_read_transfer: " + p + ": " + std::strerror(errno));
        return WV_ERR_INVALID;
    }
    const size_t n = 2 * n_mics * n_speakers * nf;
    const size_t got = n ? std::fread(out, sizeof(float), n, f) : 0;
    std::fclose(f);
    if (got != n) {
        set_err(err, err_len, "wv_results_read_transfer: " + p + " holds fewer values than the counts need");
        return WV_ERR_INVALID;
    }
    return WV_OK;
}

extern "C" wv_status wv_results_read_transfer(const char* dir, size_t n_mics, size_t n_speakers, size_t nf,
                                              float* out, char* err, size_t err_len) {
    return host_guarded(err, err_len, [&]() {
        return results_read_transfer_impl(dir, n_mics, n_speakers, nf, out, err, err_len);
    });
}
