// Pieces shared by the grid builders (host.cpp: box rooms, room_grid.cpp:
// general rooms): the built grid's storage, slot constants of one branch,
// admittance equality and a plain thread split.
#pragma once
#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

#include "wavey_fdtd.h"

struct wv_built_grid {
    wv_grid view{};
    std::vector<uint8_t> flags;
    std::vector<uint32_t> b_cell, b_start;
    std::vector<wv_slot> slots;
};

namespace wv {

// fdtd.rs:539-571: one branch (r, m, k) of `count` equal faces in parallel.
wv_slot slot_of(double r, double m, double k, double count, double dt);

// The slot list of one node: faces in direction order, each given as its
// canonical fit (equal fits share one index; see canonical_fits),
// fdtd.rs:539-571.
void node_slots(const int* canon_fits, int n_faces, const wv_admittance* fits, double dt,
                std::vector<wv_slot>& out);

// AdmittanceFit equality as the reference's derived PartialEq
// (impedance_fit.rs:34-44): beta0, the branches, band and both errors, each
// with == (NaN equals nothing).
bool fits_equal(const wv_admittance& a, const wv_admittance& b);

// Canonical index of each fit: the first equal fit, or ~i for a fit that
// equals no fit (not even itself), whose faces never share slots.
std::vector<int32_t> canonical_fits(const wv_admittance* fits, uint32_t n);

// Runs f(lo, hi, t) over [0, n) split into contiguous chunks, one per thread.
template <class F>
void parallel_chunks(int64_t n, int threads, F f) {
    int t = std::max(1, std::min<int>(threads, (int)std::max<int64_t>(1, n)));
    if (t == 1) { f(0, n, 0); return; }
    std::vector<std::thread> pool;
    for (int k = 0; k < t; ++k) {
        int64_t lo = n * k / t, hi = n * (k + 1) / t;
        pool.emplace_back([=] { f(lo, hi, k); });
    }
    for (auto& th : pool) th.join();
}

// Fills the view of a built grid from its vectors.
void finish_view(wv_built_grid& g, const int64_t n[3], const double origin[3], double dx, double dt);

}  // namespace wv
