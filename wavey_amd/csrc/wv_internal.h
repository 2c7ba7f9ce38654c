// Internal helpers shared by the host and device translation units.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

namespace wv {
extern const double kCourant;
extern const uint64_t kMaxCells;
void set_err(char* err, size_t len, const std::string& msg);
int pick_threads(int threads);
}  // namespace wv
