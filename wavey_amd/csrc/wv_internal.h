// Internal helpers shared by the host and device translation units.
#pragma once
#include <cstddef>
#include <cstdint>
#include <exception>
#include <new>
#include <string>

#include "wavey_fdtd.h"

namespace wv {
extern const double kCourant;
extern const uint64_t kMaxCells;
void set_err(char* err, size_t len, const std::string& msg);
int pick_threads(int threads);

// Runs the body of a host entry point so that no C++ exception leaves the C
// ABI: a failed host allocation is WV_ERR_NOMEM, any other exception (a
// thread that cannot start, a file stream error) WV_ERR_INVALID with its
// message.
template <class F>
wv_status host_guarded(char* err, size_t err_len, F f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        set_err(err, err_len, "host allocation failed");
        return WV_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_err(err, err_len, std::string("internal error: ") + e.what());
        return WV_ERR_INVALID;
    }
}
}  // namespace wv
