// Batched and served FDTD runs on one MI355X.
//
// Many instances (rooms, per-speaker solves) of different sizes advance
// together: one launch per time step covers the tiles of every active
// instance (a block table maps launch blocks to instances), and each
// instance's step counter lives in device memory, so a block of 128 steps is
// captured once as a HIP graph (128 step launches and one counter advance)
// and replayed until the set of instances changes. Instances share nothing,
// and each runs the plain run's kernels and arithmetic, so its receiver
// series is bit-identical alone and in a batch.
//
// Tile shapes come from an autotuner that times the whole batch: first the
// block variant (one template for the launch), then the z-chunk of each size
// class of instances, coordinate by coordinate.
//
// An instance that advances alone (the last of a batch, or a large request
// the server runs by itself) runs the plain run's kernels, two or three steps
// per pass where the plain run has them; the results are the same bit for bit.
//
// The server admits queued requests by HBM footprint into a running batch at
// the next 128-step boundary, so requests that arrive while others run share
// the device (continuous batching); finished instances leave at the same
// boundaries and their memory is freed. Each block advances either every
// admitted instance or those with the least work left, up to a fill of air
// cells, so small requests do not wait in step with large ones. By default
// large requests run on a second stream beside the batch (their own lane, two
// steps per pass), one at a time in admission order; the device runs both
// streams' blocks at once.
//
// The reference runs a wave job's sources and components one after another
// (core/wavey-solver/src/run.rs:645-661); each instance here is one such run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "plan.h"

using namespace wv;

namespace {

constexpr uint32_t kGraphSteps = kStepsPerBlock;  // steps per captured block
constexpr uint64_t kDefaultFillCells = 16ull << 20;  // the server's fill: air cells per block
constexpr uint32_t kTuneZChunks[] = {4, 8};
constexpr int kTuneVariants[] = {0, 1, 2};

template <typename T, typename S>
struct BatchEntry {
    StepParams<T, S> a;  // even-step buffers (p = fa, prev = fb) and the record base
    const S* sig;        // the signal in the slot arithmetic's type
    uint32_t steps;
    uint32_t n0;       // first step of the current 128-step block
    uint32_t block0;   // first launch block of this instance (a launch per step)
    uint32_t blocks;   // launch blocks per step: tile blocks, then receiver blocks
    uint32_t block0x;  // first block in its XCD share's list (persistent launch)
    uint32_t pad;
};

// Local block lb of an instance at step n: tile blocks, then receiver blocks.
template <typename T, typename S, int TY>
__device__ __forceinline__ void batch_item(const BatchEntry<T, S>& be, uint32_t n, uint32_t lb) {
    StepParams<T, S> a = be.a;
    if (n & 1u) {
        // Odd steps swap the two buffers (gpu_fdtd.rs:242-243).
        a.p = be.a.prev;
        a.prev = const_cast<T*>(be.a.p);
    }
    a.s_n = be.sig[n];
    a.ir = be.a.ir + (size_t)n * a.n_recv;
    if (lb < a.main_blocks) {
        block_part<T, S, TY>(a, lb);  // the whole block takes this branch
        return;
    }
    // p^n before its update (fdtd.rs:673-683).
    const uint32_t r = (lb - a.main_blocks) * blockDim.x + threadIdx.x;
    if (r < a.n_recv) a.ir[r] = a.p[a.recv[r]];
}

// One time step of every instance in the table: block b runs local block
// b - block0 of instance block_entry[b] at step n0 + k. Instances past their
// last step return at once.
template <typename T, typename S, int TY>
__global__ __launch_bounds__(kTileX*(TY + special_waves(TY))) void k_batch_step(const BatchEntry<T, S>* __restrict__ entries,
                                                           const uint32_t* __restrict__ block_entry, uint32_t k) {
    const BatchEntry<T, S>& be = entries[block_entry[blockIdx.x]];
    const uint32_t n = be.n0 + k;
    if (n >= be.steps) return;
    batch_item<T, S, TY>(be, n, blockIdx.x - be.block0);
}

// The persistent form: one launch runs steps kbase .. kbase+K-1 of a block
// for every instance. Each of the 8 XCDs, each with its own L2, takes a share
// of the instances (balanced by air cells on the host); the XCD's workgroups
// hand out its items, (step, instance, local block) in step-major order,
// kClaimRun at a time from a counter of their own, and an item of step k
// first waits until every item of its instance's step k-1 has been counted
// done. A workgroup reads its XCD from the XCC_ID hardware register, so an
// instance's fields are written and read through one L2 for the whole
// launch, which keeps them coherent without L2 write-backs or invalidations
// (what an agent-scope fence costs here, per item): a workgroup's stores are
// acknowledged by the L2 before it counts its items, and after a wait it
// invalidates its CU's L1 before reading. The counters are device-scope
// atomics. A workgroup counts its finished items at the end of each run, or
// before it waits, and skips the wait when a count it read before already
// covers the item. Items are claimed in order by running workgroups and wait
// only on items claimed before them, so the launch drains whatever the
// residency. A wait is still bounded: past the bound (or once another wait
// has failed) it sets the error word, which the host checks, and the launch
// runs on to its end.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kLineWords = 32;  // 128 bytes: one cache line per counter
constexpr uint32_t kSpinLimit = 1u << 21;
constexpr uint32_t kClaimRun = 4;

struct XcdShare {
    uint32_t first[kXcds];   // the share's first block in the block list
    uint32_t blocks[kXcds];  // the share's blocks per step
};

// Counter words: the XCDs' claim counters (a line each), the error word's
// line, then one line per table entry (its items done in this launch).
inline size_t persist_ctl_words(size_t entries) { return (kXcds + 1 + entries) * kLineWords; }

__global__ void k_persist_reset(uint32_t* __restrict__ ctl, uint32_t n_entries) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kXcds) ctl[i * kLineWords] = 0;
    if (i < n_entries) ctl[(kXcds + 1 + i) * kLineWords] = 0;
}

__device__ __forceinline__ uint32_t xcc_id() {
    uint32_t x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
    return x;
}

template <typename T, typename S, int TY>
__global__ __launch_bounds__(kTileX*(TY + special_waves(TY))) void k_batch_persist(const BatchEntry<T, S>* __restrict__ entries,
                                                              const uint32_t* __restrict__ pblock,
                                                              uint32_t* __restrict__ ctl, XcdShare xs,
                                                              uint32_t kbase, uint32_t K) {
    __shared__ uint32_t s_run, s_seen;
    const uint32_t x = xcc_id() & (kXcds - 1);
    const uint32_t nb = xs.blocks[x], first = xs.first[x];
    const uint32_t total = K * nb;  // below 2^31 (checked on the host)
    uint32_t* claim = ctl + x * kLineWords;
    uint32_t* err = ctl + kXcds * kLineWords;
    uint32_t* done = ctl + (kXcds + 1) * kLineWords;
    // Items finished but not yet counted (all of one entry), and the last
    // count read (of entry seen_e).
    uint32_t pend_e = 0, pend_n = 0, seen_e = ~0u, seen_v = 0;
    auto flush = [&]() {
        if (pend_n == 0) return;
        __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are in the L2
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_fetch_add(done + pend_e * kLineWords, pend_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        pend_n = 0;
    };
    if (threadIdx.x == 0) s_run = __hip_atomic_fetch_add(claim, kClaimRun, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    uint32_t run = __builtin_amdgcn_readfirstlane(s_run);
    while (run < total) {
        __syncthreads();  // every wave has read s_run
        if (threadIdx.x == 0)  // the next run, claimed ahead
            s_run = __hip_atomic_fetch_add(claim, kClaimRun, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t run_end = min(run + kClaimRun, total);
        for (uint32_t item = run; item < run_end; ++item) {
            const uint32_t k = item / nb, j = item - k * nb;
            const uint32_t ei = pblock[first + j];
            const BatchEntry<T, S>& be = entries[ei];
            if (pend_n && pend_e != ei) flush();
            const uint32_t need = k * be.blocks;
            if (k > 0 && !(seen_e == ei && seen_v >= need)) {
                flush();
                if (threadIdx.x == 0) {
                    uint32_t* cnt = done + ei * kLineWords;
                    uint32_t v, spins = 0;
                    while ((v = __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < need) {
                        if ((spins & 255u) == 0 &&
                            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                            break;
                        if (++spins >= kSpinLimit) {
                            __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                    s_seen = v;
                    // This CU's L1 may hold lines of the instance's fields
                    // from before the wait: an agent-scope acquire
                    // invalidates it (and the L2's unmodified lines; the
                    // other items' writes are in this XCD's L2).
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    __builtin_amdgcn_s_waitcnt(0);
                }
                __syncthreads();
                seen_e = ei;
                seen_v = __builtin_amdgcn_readfirstlane(s_seen);
            }
            const uint32_t n = be.n0 + kbase + k;
            if (n < be.steps) batch_item<T, S, TY>(be, n, j - be.block0x);
            // Every wave is done with the item's shared memory (the tile
            // block's hand-out counter) before the next item resets it.
            __syncthreads();
            pend_e = ei;
            ++pend_n;
        }
        flush();
        run = __builtin_amdgcn_readfirstlane(s_run);
    }
}

// The XCC_ID of each of 64 blocks (the persistent form's check that the
// register names the 8 XCDs).
__global__ void k_xcc_probe(uint32_t* __restrict__ out) {
    if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

template <typename T, typename S>
__global__ void k_batch_tick(BatchEntry<T, S>* __restrict__ entries, uint32_t n, uint32_t add) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) entries[i].n0 += add;
}

template <typename T, typename S>
void launch_batch_step(int variant, uint32_t grid, hipStream_t st, const BatchEntry<T, S>* e, const uint32_t* be,
                       uint32_t k) {
    const dim3 blk(variant_threads(variant));
    switch (variant) {
        case 0: hipLaunchKernelGGL((k_batch_step<T, S, 7>), dim3(grid), blk, 0, st, e, be, k); break;
        case 1: hipLaunchKernelGGL((k_batch_step<T, S, 3>), dim3(grid), blk, 0, st, e, be, k); break;
        case 2: hipLaunchKernelGGL((k_batch_step<T, S, 15>), dim3(grid), blk, 0, st, e, be, k); break;
        default: hipLaunchKernelGGL((k_batch_step<T, S, 1>), dim3(grid), blk, 0, st, e, be, k); break;
    }
}

template <typename T, typename S>
void launch_batch_persist(int variant, uint32_t grid, hipStream_t st, const BatchEntry<T, S>* e, const uint32_t* pb,
                          uint32_t* ctl, const XcdShare& xs, uint32_t kbase, uint32_t K) {
    const dim3 blk(variant_threads(variant));
    switch (variant) {
        case 0: hipLaunchKernelGGL((k_batch_persist<T, S, 7>), dim3(grid), blk, 0, st, e, pb, ctl, xs, kbase, K); break;
        case 1: hipLaunchKernelGGL((k_batch_persist<T, S, 3>), dim3(grid), blk, 0, st, e, pb, ctl, xs, kbase, K); break;
        case 2: hipLaunchKernelGGL((k_batch_persist<T, S, 15>), dim3(grid), blk, 0, st, e, pb, ctl, xs, kbase, K); break;
        default: hipLaunchKernelGGL((k_batch_persist<T, S, 1>), dim3(grid), blk, 0, st, e, pb, ctl, xs, kbase, K); break;
    }
}

// Workgroups of a persistent launch: as many as the CUs hold at once (the
// occupancy of the variant's kernel), a multiple of the 8 XCDs.
template <typename T, typename S>
uint32_t persist_grid(int variant) {
    int per_cu = 0;
    const int threads = (int)variant_threads(variant);
    switch (variant) {
        case 0: WV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_persist<T, S, 7>, threads, 0)); break;
        case 1: WV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_persist<T, S, 3>, threads, 0)); break;
        case 2: WV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_persist<T, S, 15>, threads, 0)); break;
        default: WV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_persist<T, S, 1>, threads, 0)); break;
    }
    int dev = 0, cus = 0;
    WV_HIP(hipGetDevice(&dev));
    WV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t g = (uint32_t)std::max(1, per_cu) * (uint32_t)std::max(1, cus);
    return std::max(kXcds, g / kXcds * kXcds);
}

// Whether the XCC_ID register names 8 XCDs on this device (each of 0..7
// seen, nothing else), checked once per device: the persistent form needs
// every share's workgroups.
bool xcds_ok() {
    static std::mutex mu;
    static std::map<int, bool> known;
    int dev = 0;
    WV_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(dev);
    if (it != known.end()) return it->second;
    DBuf d;
    d.alloc(64 * sizeof(uint32_t));
    hipLaunchKernelGGL(k_xcc_probe, dim3(64), dim3(64), 0, 0, d.as<uint32_t>());
    WV_HIP(hipGetLastError());
    uint32_t h[64];
    WV_HIP(hipMemcpy(h, d.p, sizeof h, hipMemcpyDeviceToHost));
    uint32_t seen = 0;
    bool ok = true;
    for (uint32_t v : h) {
        if (v >= kXcds) ok = false;
        else seen |= 1u << v;
    }
    ok = ok && seen == (1u << kXcds) - 1;
    known[dev] = ok;
    return ok;
}

// Size class of an instance: floor(log2(air cells)).
int size_class_of(uint64_t air) {
    int c = 0;
    while (air > 1) {
        air >>= 1;
        ++c;
    }
    return c;
}

// Autotuned shapes, shared by every batch and server of the process:
// z-chunk per (precision, size class) and block variant per precision.
std::mutex g_tune_mu;
std::map<std::pair<int, int>, uint32_t> g_tune_zchunk;
std::map<int, int> g_tune_variant;
std::map<int, int> g_tune_launch;

uint32_t tuned_zchunk(int prec, int cls) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_zchunk.find({prec, cls});
    return it == g_tune_zchunk.end() ? 0u : it->second;
}

int tuned_variant(int prec) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_variant.find(prec);
    return it == g_tune_variant.end() ? kDefaultVariant : it->second;
}

int tuned_launch(int prec) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_launch.find(prec);
    return it == g_tune_launch.end() ? WV_BATCH_STEPS : it->second;
}

struct Instance {
    int32_t id = -1;
    std::unique_ptr<wv_plan> plan;
    std::vector<wv_tap> taps;
    std::vector<double> signal;
    std::vector<uint64_t> recv;
    DBuf sig;
    uint32_t steps = 0;
    uint32_t n0 = 0;  // host mirror of the device block start
    bool done = false;
    bool paused = false;  // left out of the launch table by the server's selection
    int lane = -1;        // the server lane with a block of this instance in flight, or -1
    int size_class = 0;
    uint64_t footprint = 0;
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct wv_batch {
    int prec = WV_FP32;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int variant = kDefaultVariant;
    // The current table's variant: variant, or blocks of one tile for a
    // small table (small_table) unless the variant was set or is being tuned.
    int table_variant = kDefaultVariant;
    bool variant_set = false;
    bool tuning = false;
    int launch = WV_BATCH_STEPS;  // the launch form of a block
    std::vector<std::unique_ptr<Instance>> inst;
    DBuf entries, block_entry;
    std::vector<Instance*> table;
    uint32_t total_blocks = 0;
    // The persistent form: the XCD shares' block lists and counters.
    DBuf pblock, ctl;
    XcdShare share{};
    std::vector<uint32_t> entry_blocks;  // blocks per step of each table entry
    uint32_t max_share_blocks = 0;
    bool last_persist = false;  // the last enqueued block was a persistent launch
    bool dirty = true;
    bool graph_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double last_ms = 0.0;
    uint64_t last_steps = 0, last_cell_updates = 0;
    std::atomic<uint64_t> graph_launches{0}, captures{0}, solo_blocks{0};
    std::mutex add_mu;  // wv_batch_add may be called from several threads at once

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_valid = false;
    }
    ~wv_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        inst.clear();
        entries.release();
        block_entry.release();
        pblock.release();
        ctl.release();
        for (auto e : {t0, t1})
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void batch_init(wv_batch* b, int prec) {
    b->prec = prec;
    WV_HIP(hipGetDevice(&b->device));
    WV_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    WV_HIP(hipEventCreate(&b->t0));
    WV_HIP(hipEventCreate(&b->t1));
    b->variant = tuned_variant(prec);
    b->launch = tuned_launch(prec);
}

// Starts (or restarts from rest) an instance's run with its z-chunk (0: the
// batch's default, not the plain run's by size).
void instance_begin(Instance* in, uint32_t zchunk) {
    wv_plan* pl = in->plan.get();
    pl->zchunk = zchunk ? zchunk : kDefaultZChunk;
    run_begin(pl, in->taps.data(), in->taps.size(), in->signal.data(), in->signal.size(), in->recv.data(),
              in->recv.size());
    in->n0 = 0;
    in->done = false;
}

// Returns an instance's run to step 0 with the layout it has (no
// reclassification or compaction).
void instance_rewind(Instance* in) {
    run_rewind(in->plan.get());
    in->n0 = 0;
    in->done = false;
}

// Restarts an instance from rest with z-chunk zchunk: a rewind when its
// layout already has it.
void instance_restart(Instance* in, uint32_t zchunk) {
    if (in->plan->run_zchunk == (zchunk ? zchunk : kDefaultZChunk)) instance_rewind(in);
    else instance_begin(in, zchunk);
}

template <typename S>
std::unique_ptr<Instance> make_instance(int prec, const wv_grid* g, const wv_tap* taps, size_t n_taps,
                                        const double* signal, size_t steps, const uint64_t* recv, size_t n_recv) {
    auto in = std::make_unique<Instance>();
    in->plan = std::make_unique<wv_plan>();
    in->plan->prec = prec;
    // The batched launch is single-step; the plan keeps the plain run's
    // two-step passes (ring list, two more field buffers) for blocks in which
    // the instance advances alone.
    in->plan->steps_per_pass = 0;
    plan_setup(in->plan.get(), g);
    in->taps.assign(taps, taps + n_taps);
    in->signal.assign(signal, signal + steps);
    in->recv.assign(recv, recv + n_recv);
    in->steps = (uint32_t)steps;
    in->size_class = size_class_of(in->plan->air);
    std::vector<S> sig(std::max<size_t>(steps, 1), S(0));
    for (size_t n = 0; n < steps; ++n) sig[n] = (S)signal[n];
    in->sig.alloc(sig.size() * sizeof(S));
    WV_HIP(hipMemcpy(in->sig.p, sig.data(), sig.size() * sizeof(S), hipMemcpyHostToDevice));
    instance_begin(in.get(), tuned_zchunk(prec, in->size_class));
    in->footprint = plan_bytes(in->plan.get()) + in->sig.bytes;
    return in;
}

// A table of few air cells launches blocks of one tile, as a plain run of a
// grid that size does (default_step_variant): the server's small-request
// lane often holds one or two small rooms, where the larger blocks leave the
// device waiting on their special cells (measured: one 0.38 M-cell room 6.9
// against 9.3-11.0 us per step).
bool small_table(const wv_batch* b) {
    if (b->variant_set || b->tuning || b->launch != WV_BATCH_STEPS) return false;
    uint64_t air = 0;
    for (auto& up : b->inst) {
        const Instance* in = up.get();
        if (!in->done && !in->paused && in->n0 < in->steps) air += in->plan->air;
    }
    return default_step_variant(air, b->prec == WV_FP64) == 3;
}

// Rebuilds the launch table from the instances that still have steps to
// run and are not paused. The stream must be idle. Every block starts at an
// even step (a multiple of 128), where p is in the plan's buffer buf_p and
// prev in buf_q: the batched step updates prev in place, so after a block p
// is back in buf_p.
template <typename T, typename S>
void build_table(wv_batch* b) {
    b->table_variant = small_table(b) ? 3 : b->variant;
    const uint32_t ty = kVariantTiles[b->table_variant];
    std::vector<BatchEntry<T, S>> es;
    std::vector<uint32_t> be;
    b->table.clear();
    b->entry_blocks.clear();
    for (auto& up : b->inst) {
        Instance* in = up.get();
        if (in->done || in->paused || in->n0 >= in->steps) continue;
        wv_plan* pl = in->plan.get();
        const DBuf* fields[5] = {&pl->fa, &pl->fb, &pl->fc, &pl->fd, &pl->fe};
        BatchEntry<T, S> e{};
        e.a = make_params<T, S>(pl, ty);
        e.a.p = fields[pl->buf_p]->as<T>();
        e.a.prev = fields[pl->buf_q]->as<T>();
        e.a.ir = pl->ir.as<T>();
        e.sig = in->sig.as<S>();
        e.steps = in->steps;
        e.n0 = in->n0;
        e.block0 = (uint32_t)be.size();
        const uint32_t blocks = e.a.main_blocks + div_up(pl->n_recv, variant_threads(b->table_variant));
        if ((uint64_t)be.size() + blocks >= (1ull << 31)) throw DevErr{"batch launch too large", WV_ERR_LIMIT};
        e.blocks = blocks;
        be.insert(be.end(), blocks, (uint32_t)es.size());
        es.push_back(e);
        b->table.push_back(in);
        b->entry_blocks.push_back(blocks);
    }
    b->total_blocks = (uint32_t)be.size();
    // The persistent form's XCD shares: instances largest first, each to the
    // share with the fewest air cells so far; each share's blocks listed
    // instance by instance.
    std::vector<uint32_t> order(es.size());
    for (uint32_t i = 0; i < (uint32_t)order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return b->table[x]->plan->air > b->table[y]->plan->air; });
    std::vector<uint32_t> xcd_of(es.size(), 0);
    uint64_t load[kXcds] = {};
    for (uint32_t i : order) {
        const uint32_t x = (uint32_t)(std::min_element(load, load + kXcds) - load);
        xcd_of[i] = x;
        load[x] += b->table[i]->plan->air;
    }
    std::vector<uint32_t> pb;
    pb.reserve(be.size());
    b->max_share_blocks = 0;
    for (uint32_t x = 0; x < kXcds; ++x) {
        b->share.first[x] = (uint32_t)pb.size();
        for (uint32_t i = 0; i < (uint32_t)es.size(); ++i)
            if (xcd_of[i] == x) {
                es[i].block0x = (uint32_t)pb.size() - b->share.first[x];
                pb.insert(pb.end(), es[i].blocks, i);
            }
        b->share.blocks[x] = (uint32_t)pb.size() - b->share.first[x];
        b->max_share_blocks = std::max(b->max_share_blocks, b->share.blocks[x]);
    }
    if (!es.empty()) {
        b->entries.alloc(es.size() * sizeof(es[0]));
        b->block_entry.alloc(be.size() * 4);
        b->pblock.alloc(pb.size() * 4);
        b->ctl.alloc(persist_ctl_words(es.size()) * 4);
        WV_HIP(hipMemcpy(b->entries.p, es.data(), es.size() * sizeof(es[0]), hipMemcpyHostToDevice));
        WV_HIP(hipMemcpy(b->block_entry.p, be.data(), be.size() * 4, hipMemcpyHostToDevice));
        WV_HIP(hipMemcpy(b->pblock.p, pb.data(), pb.size() * 4, hipMemcpyHostToDevice));
        // The error word starts clear; the claim and done counters are
        // cleared by every launch.
        const uint32_t zero = 0;
        WV_HIP(hipMemcpy(b->ctl.as<uint32_t>() + kXcds * kLineWords, &zero, 4, hipMemcpyHostToDevice));
    }
    b->drop_graph();
    b->dirty = false;
}

// Whether a block of K steps of the current table may take the persistent
// form: the 8 XCDs found, and every share's items counted in 31 bits.
bool persist_usable(const wv_batch* b, uint32_t K) {
    return b->launch == WV_BATCH_PERSISTENT && (uint64_t)K * b->max_share_blocks < (1ull << 31) && xcds_ok();
}

// Steps kbase .. kbase+K-1 of the current table (each instance from its n0)
// in the persistent form: the counters cleared, then the one launch.
template <typename T, typename S>
void enqueue_persist(wv_batch* b, uint32_t grid, uint32_t kbase, uint32_t K) {
    const uint32_t n = (uint32_t)b->table.size();
    hipLaunchKernelGGL(k_persist_reset, dim3(div_up(std::max(n, kXcds), 256)), dim3(256), 0, b->stream,
                       b->ctl.as<uint32_t>(), n);
    launch_batch_persist<T, S>(b->table_variant, grid, b->stream, b->entries.as<BatchEntry<T, S>>(), b->pblock.as<uint32_t>(),
                               b->ctl.as<uint32_t>(), b->share, kbase, K);
}

// After the stream has drained a persistent launch of K steps: every
// instance's items done, no wait past its bound.
void check_persist(wv_batch* b, uint32_t K) {
    const size_t n = b->table.size();
    std::vector<uint32_t> h(persist_ctl_words(n));
    WV_HIP(hipMemcpy(h.data(), b->ctl.p, h.size() * 4, hipMemcpyDeviceToHost));
    if (h[kXcds * kLineWords]) throw DevErr{"persistent batch launch: a step wait passed its bound", WV_ERR_DEVICE};
    for (size_t i = 0; i < n; ++i)
        if (h[(kXcds + 1 + i) * kLineWords] != K * b->entry_blocks[i])
            throw DevErr{"persistent batch launch: an instance's items did not all run", WV_ERR_DEVICE};
}

// Enqueues one block of kGraphSteps steps of the current table: the captured
// graph (captured on first use after a table change), or the same launches
// issued one by one.
template <typename T, typename S>
void enqueue_block(wv_batch* b, bool use_graph) {
    const uint32_t n = (uint32_t)b->table.size();
    if (n == 0) return;
    // Re-centring of the instances whose block starts at one of its steps
    // (their p and prev are in the plan's buf_p, buf_q at a block start).
    for (Instance* in : b->table) recentre(in->plan.get(), in->n0, b->stream);
    auto* e = b->entries.as<BatchEntry<T, S>>();
    const uint32_t* be = b->block_entry.as<uint32_t>();
    const bool persist = persist_usable(b, kGraphSteps);
    const uint32_t pgrid = persist ? persist_grid<T, S>(b->table_variant) : 0;
    auto issue = [&]() {
        if (persist) enqueue_persist<T, S>(b, pgrid, 0, kGraphSteps);
        else
            for (uint32_t k = 0; k < kGraphSteps; ++k)
                launch_batch_step<T, S>(b->table_variant, b->total_blocks, b->stream, e, be, k);
        hipLaunchKernelGGL((k_batch_tick<T, S>), dim3(div_up(n, 256)), dim3(256), 0, b->stream, e, n, kGraphSteps);
    };
    if (use_graph) {
        if (!b->graph_valid) {
            WV_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            issue();
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(b->stream, &g);
            WV_HIP(ce);
            b->graph = g;
            WV_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graph_valid = true;
            ++b->captures;
        }
        WV_HIP(hipGraphLaunch(b->exec, b->stream));
        ++b->graph_launches;
    } else {
        issue();
        WV_HIP(hipGetLastError());
    }
    for (Instance* in : b->table) in->n0 += kGraphSteps;
    b->last_persist = persist;
}

// Enqueues the next block (up to kGraphSteps steps) of one instance alone,
// with the plain run's kernels.
void enqueue_solo(wv_batch* b, Instance* in) {
    const size_t stop = std::min<size_t>((size_t)in->n0 + kGraphSteps, in->steps);
    enqueue_steps(in->plan.get(), in->n0, stop, b->stream);
    in->n0 = (uint32_t)stop;
    b->dirty = true;  // its p and prev may have moved to other buffers
    ++b->solo_blocks;
}

// Marks instances whose last step has run (after a synchronize). Returns how
// many finished.
uint32_t collect_done(wv_batch* b) {
    uint32_t k = 0;
    for (auto& up : b->inst) {
        Instance* in = up.get();
        if (!in->done && in->n0 >= in->steps) {
            in->done = true;
            in->plan->done = in->steps;
            in->plan->steps = in->steps;
            in->plan->active = true;
            b->dirty = true;
            ++k;
        }
    }
    return k;
}

template <typename T, typename S>
wv_status batch_run(wv_batch* b, bool use_graph, const volatile int32_t* cancel, wv_progress_fn progress, void* ctx) {
    uint64_t total = 0, cell_updates = 0;
    for (auto& in : b->inst)
        if (!in->done) {
            total += in->steps - std::min(in->n0, in->steps);
            cell_updates += (uint64_t)in->plan->air * (in->steps - std::min(in->n0, in->steps));
        }
    uint64_t done_steps = 0;
    WV_HIP(hipStreamSynchronize(b->stream));
    WV_HIP(hipEventRecord(b->t0, b->stream));
    uint64_t max_steps = 0;
    while (true) {
        if (cancel && *cancel) return WV_CANCELLED;
        if (b->dirty) build_table<T, S>(b);
        if (b->table.empty()) break;
        // Blocks until the first instance of the table finishes: the table
        // does not change before then, so they are enqueued back to back.
        uint32_t blocks = ~0u;
        for (Instance* in : b->table) blocks = std::min(blocks, div_up(in->steps - in->n0, kGraphSteps));
        Instance* solo = b->table.size() == 1 && b->table[0]->plan->run_steps_per_pass >= 2 ? b->table[0] : nullptr;
        uint32_t issued = 0;
        while (issued < blocks) {
            if (solo) enqueue_solo(b, solo);
            else enqueue_block<T, S>(b, use_graph);
            ++issued;
            if (cancel && *cancel) break;
        }
        max_steps += (uint64_t)issued * kGraphSteps;
        WV_HIP(hipStreamSynchronize(b->stream));
        // (No block is issued while an instance of no steps heads the table.)
        if (!solo && issued && b->last_persist) check_persist(b, kGraphSteps);
        done_steps += (uint64_t)b->table.size() * issued * kGraphSteps;
        collect_done(b);
        if (progress && total) progress(std::min(1.0, (double)done_steps / (double)total), ctx);
    }
    WV_HIP(hipEventRecord(b->t1, b->stream));
    WV_HIP(hipEventSynchronize(b->t1));
    float ms = 0.f;
    WV_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
    b->last_ms = ms;
    b->last_steps = max_steps;
    b->last_cell_updates = cell_updates;
    return WV_OK;
}

// Device time per step of the whole table, eager launches, after warm-up.
template <typename T, typename S>
double time_table(wv_batch* b, uint32_t warmup, uint32_t steps) {
    build_table<T, S>(b);
    if (b->table.empty()) return 0.0;
    auto* e = b->entries.as<BatchEntry<T, S>>();
    const uint32_t* be = b->block_entry.as<uint32_t>();
    const bool persist = persist_usable(b, warmup + steps);
    const uint32_t pgrid = persist ? persist_grid<T, S>(b->table_variant) : 0;
    if (persist && warmup) enqueue_persist<T, S>(b, pgrid, 0, warmup);
    for (uint32_t k = 0; !persist && k < warmup; ++k)
        launch_batch_step<T, S>(b->table_variant, b->total_blocks, b->stream, e, be, k);
    WV_HIP(hipEventRecord(b->t0, b->stream));
    if (persist) enqueue_persist<T, S>(b, pgrid, warmup, steps);
    for (uint32_t k = 0; !persist && k < steps; ++k)
        launch_batch_step<T, S>(b->table_variant, b->total_blocks, b->stream, e, be, warmup + k);
    WV_HIP(hipGetLastError());
    WV_HIP(hipEventRecord(b->t1, b->stream));
    WV_HIP(hipEventSynchronize(b->t1));
    if (persist) check_persist(b, steps);
    float ms = 0.f;
    WV_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
    return ms / std::max(steps, 1u);
}

// Coordinate search over the batch's tile shape: the block variant
// (one template for the launch), then each size class's z-chunk, largest
// class first, each timed on the whole batch. Instances restart from rest.
template <typename T, typename S>
void batch_autotune(wv_batch* b, uint32_t trial_steps) {
    WV_HIP(hipStreamSynchronize(b->stream));
    struct Tuning {
        wv_batch* b;
        explicit Tuning(wv_batch* bb) : b(bb) { b->tuning = true; }
        ~Tuning() { b->tuning = false; }
    } tuning(b);
    for (auto& in : b->inst) instance_rewind(in.get());
    int best_v = b->variant;
    double best = 1e300;
    for (int v : kTuneVariants) {
        b->variant = v;
        const double t = time_table<T, S>(b, 2, trial_steps);
        if (t < best) {
            best = t;
            best_v = v;
        }
    }
    b->variant = best_v;
    std::map<int, uint64_t, std::greater<int>> classes;
    for (auto& in : b->inst) classes[in->size_class] += in->plan->air;
    std::map<int, uint32_t> chosen;
    uint64_t total_air = 0;
    for (auto& [cls, air] : classes) total_air += air;
    for (auto& [cls, air] : classes) {
        // A class with under 5 % of the air cells keeps its z-chunk: its
        // trials (each sets its instances up again) cost more than its
        // choice can win (room mix: autotune 0.19 to 0.16 s, run unchanged).
        if (air * 20 < total_air) {
            const uint32_t zc = tuned_zchunk(b->prec, cls);
            chosen[cls] = zc ? zc : kDefaultZChunk;
            continue;
        }
        uint32_t best_z = 0;
        double best_t = 1e300;
        for (uint32_t zc : kTuneZChunks) {
            // Timing needs the layout, not a run from rest.
            for (auto& in : b->inst)
                if (in->size_class == cls && in->plan->run_zchunk != zc) instance_begin(in.get(), zc);
            const double t = time_table<T, S>(b, 2, trial_steps);
            if (t < best_t) {
                best_t = t;
                best_z = zc;
            }
        }
        chosen[cls] = best_z;
        for (auto& in : b->inst)
            if (in->size_class == cls && in->plan->run_zchunk != best_z) instance_begin(in.get(), best_z);
    }
    // The launch form, with the shapes chosen.
    int best_l = WV_BATCH_STEPS;
    if (xcds_ok()) {
        double tl[2];
        for (int l : {WV_BATCH_STEPS, WV_BATCH_PERSISTENT}) {
            b->launch = l;
            tl[l] = time_table<T, S>(b, 2, trial_steps);
        }
        best_l = tl[WV_BATCH_PERSISTENT] < tl[WV_BATCH_STEPS] ? WV_BATCH_PERSISTENT : WV_BATCH_STEPS;
    }
    b->launch = best_l;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_variant[b->prec] = best_v;
        g_tune_launch[b->prec] = best_l;
        for (auto& [cls, zc] : chosen) g_tune_zchunk[{b->prec, cls}] = zc;
    }
    // Every instance restarts from rest with its class's z-chunk.
    for (auto& in : b->inst) instance_restart(in.get(), chosen[in->size_class]);
    b->dirty = true;
}

template <class F>
wv_status dispatch_prec(int prec, F f) {
    switch (prec) {
        case WV_FP32: return f(float(), float());
        case WV_FP64: return f(double(), double());
        default: return f(float(), double());
    }
}

Instance* find_instance(wv_batch* b, int32_t id) {
    if (id < 0 || (size_t)id >= b->inst.size()) return nullptr;
    return b->inst[id].get();
}

}  // namespace

// ------------------------------------------------------------------ batch C ABI

extern "C" wv_status wv_batch_create(int precision, wv_batch** out, char* err, size_t err_len) {
    if (!out || (precision != WV_FP32 && precision != WV_FP64 && precision != WV_FP32_STATE64)) {
        set_err(err, err_len, "wv_batch_create: invalid argument");
        return WV_ERR_INVALID;
    }
    return guarded(err, err_len, [&]() -> wv_status {
        auto b = std::make_unique<wv_batch>();
        batch_init(b.get(), precision);
        *out = b.release();
        return WV_OK;
    });
}

extern "C" void wv_batch_free(wv_batch* b) { delete b; }

extern "C" wv_status wv_batch_add(wv_batch* b, const wv_grid* grid, const wv_tap* taps, size_t n_taps,
                                  const double* signal, size_t steps, const uint64_t* receivers, size_t n_recv,
                                  int32_t* id, char* err, size_t err_len) {
    if (!b || !id) {
        set_err(err, err_len, "wv_batch_add: invalid argument");
        return WV_ERR_INVALID;
    }
    wv_status s = check_grid(grid, err, err_len);
    if (s != WV_OK) return s;
    return guarded(err, err_len, [&]() -> wv_status {
        WV_HIP(hipSetDevice(b->device));
        // Argument checks against the grid's plan layout (tap planes, receivers).
        wv_plan probe;
        probe.nx = (uint32_t)grid->n[0];
        probe.ny = (uint32_t)grid->n[1];
        probe.nz = (uint32_t)grid->n[2];
        probe.cells = grid->n[0] * grid->n[1] * grid->n[2];
        wv_status st = check_run_args(&probe, taps, n_taps, signal, steps, receivers, n_recv, err, err_len);
        if (st != WV_OK) return st;
        std::unique_ptr<Instance> in;
        if (b->prec == WV_FP32) in = make_instance<float>(b->prec, grid, taps, n_taps, signal, steps, receivers, n_recv);
        else in = make_instance<double>(b->prec, grid, taps, n_taps, signal, steps, receivers, n_recv);
        std::lock_guard<std::mutex> lk(b->add_mu);
        in->id = (int32_t)b->inst.size();
        *id = in->id;
        b->inst.push_back(std::move(in));
        b->dirty = true;
        return WV_OK;
    });
}

extern "C" wv_status wv_batch_autotune(wv_batch* b, uint32_t trial_steps, char* err, size_t err_len) {
    if (!b || trial_steps == 0) {
        set_err(err, err_len, "wv_batch_autotune: invalid argument");
        return WV_ERR_INVALID;
    }
    return guarded(err, err_len, [&]() -> wv_status {
        WV_HIP(hipSetDevice(b->device));
        return dispatch_prec(b->prec, [&](auto t, auto s) -> wv_status {
            batch_autotune<decltype(t), decltype(s)>(b, trial_steps);
            return WV_OK;
        });
    });
}

extern "C" wv_status wv_batch_set_tile(wv_batch* b, int32_t variant, uint32_t zchunk, char* err, size_t err_len) {
    if (!b || variant >= kVariants || zchunk > 32 || (zchunk & (zchunk - 1))) {
        set_err(err, err_len, "wv_batch_set_tile: invalid argument");
        return WV_ERR_INVALID;
    }
    return guarded(err, err_len, [&]() -> wv_status {
        WV_HIP(hipSetDevice(b->device));
        WV_HIP(hipStreamSynchronize(b->stream));
        b->variant = variant < 0 ? kDefaultVariant : variant;
        b->variant_set = variant >= 0;
        for (auto& in : b->inst) instance_begin(in.get(), zchunk);
        b->dirty = true;
        return WV_OK;
    });
}

extern "C" wv_status wv_batch_set_launch(wv_batch* b, int form, char* err, size_t err_len) {
    if (!b || (form != WV_BATCH_STEPS && form != WV_BATCH_PERSISTENT)) {
        set_err(err, err_len, "wv_batch_set_launch: invalid argument");
        return WV_ERR_INVALID;
    }
    return guarded(err, err_len, [&]() -> wv_status {
        WV_HIP(hipSetDevice(b->device));
        if (form == WV_BATCH_PERSISTENT && !xcds_ok())
            throw DevErr{"wv_batch_set_launch: the device's XCC_ID register does not name 8 XCDs", WV_ERR_LIMIT};
        WV_HIP(hipStreamSynchronize(b->stream));
        b->launch = form;
        b->drop_graph();
        b->dirty = true;  // the table's variant depends on the launch form
        return WV_OK;
    });
}

extern "C" wv_status wv_batch_run(wv_batch* b, int use_graphs, const volatile int32_t* cancel,
                                  wv_progress_fn progress, void* ctx, char* err, size_t err_len) {
    if (!b) {
        set_err(err, err_len, "wv_batch_run: invalid argument");
        return WV_ERR_INVALID;
    }
    return guarded(err, err_len, [&]() -> wv_status {
        WV_HIP(hipSetDevice(b->device));
        return dispatch_prec(b->prec, [&](auto t, auto s) -> wv_status {
            return batch_run<decltype(t), decltype(s)>(b, use_graphs != 0, cancel, progress, ctx);
        });
    });
}

extern "C" wv_status wv_batch_result(wv_batch* b, int32_t id, const double* freqs, size_t nf, int mode,
                                     double* spectra_out, double* ir_out, char* err, size_t err_len) {
    Instance* in = b ? find_instance(b, id) : nullptr;
    if (!in || (nf && !freqs)) {
        set_err(err, err_len, "wv_batch_result: invalid argument");
        return WV_ERR_INVALID;
    }
    if (!in->done) {
        set_err(err, err_len, "wv_batch_result: instance has not finished");
        return WV_ERR_INVALID;
    }
    return guarded(err, err_len, [&]() -> wv_status {
        WV_HIP(hipSetDevice(b->device));
        run_finish(in->plan.get(), freqs, nf, mode, spectra_out, ir_out);
        return WV_OK;
    });
}

extern "C" wv_status wv_batch_info(const wv_batch* b, int32_t id, uint64_t out[6]) {
    if (!b || !out || id < 0 || (size_t)id >= b->inst.size()) return WV_ERR_INVALID;
    const Instance* in = b->inst[id].get();
    out[0] = in->plan->cells;
    out[1] = in->plan->air;
    out[2] = in->steps;
    out[3] = in->plan->run_zchunk;
    out[4] = (uint64_t)in->size_class;
    out[5] = in->footprint;
    return WV_OK;
}

extern "C" wv_status wv_batch_timing(const wv_batch* b, double* ms, uint64_t out[7]) {
    if (!b || !ms || !out) return WV_ERR_INVALID;
    *ms = b->last_ms;
    out[0] = b->last_steps;
    out[1] = b->last_cell_updates;
    out[2] = b->graph_launches;
    out[3] = b->captures;
    out[4] = (uint64_t)b->variant;
    out[5] = b->solo_blocks;
    out[6] = (uint64_t)b->launch;
    return WV_OK;
}

// ------------------------------------------------------------------ server

namespace {

struct Request {
    uint64_t id = 0;
    const wv_grid* grid = nullptr;  // borrowed until wv_serve_wait returns
    std::vector<wv_tap> taps;
    std::vector<double> signal;
    std::vector<uint64_t> recv;
    std::vector<double> freqs;
    int mode = WV_MODE_IR;
    double t_submit = 0, t_admit = 0, t_done = 0;
    bool done = false;
    wv_status status = WV_OK;
    std::string err;
    std::vector<double> ir, spectra;
    Instance* inst = nullptr;
    uint64_t reserved = 0;
};

}  // namespace

struct wv_server {
    int prec = WV_FP32;
    uint64_t cap = 0;
    std::atomic<int> policy{WV_SERVE_LANES};
    std::atomic<uint64_t> fill{kDefaultFillCells};
    std::atomic<uint64_t> lane_cells{0};  // 0: three-step plans take the large lane
    std::atomic<uint64_t> submitted{0};  // submissions so far (wakes the worker's wait)
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    std::unique_ptr<wv_batch> batch;
    hipStream_t lane_stream = nullptr;  // the large-request lane
    hipEvent_t lane_ev[2] = {nullptr, nullptr};
    std::vector<std::shared_ptr<Request>> running;
    uint64_t used = 0, peak_used = 0, cell_updates = 0, completed = 0, failed = 0, lane_blocks = 0;
    uint32_t max_active = 0;
    double t_first = -1.0, t_last = 0.0;
    std::vector<double> latencies;
    std::string fatal;

    ~wv_server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        if (lane_stream) (void)hipStreamSynchronize(lane_stream);
        for (auto e : lane_ev)
            if (e) (void)hipEventDestroy(e);
        if (lane_stream) (void)hipStreamDestroy(lane_stream);
    }
};

namespace {

// A stream with the block in flight on it: lane 0 is the batch's stream
// (the batched launch, or one instance alone), lane 1 the large-request lane.
struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<Instance*> members;  // instances the block in flight advances
};

bool lane_finished(const Lane& l) {
    const hipError_t e = hipEventQuery(l.ev);
    if (e == hipErrorNotReady) return false;
    WV_HIP(e);
    return true;
}

uint64_t work_left(const Instance* in) { return in->plan->air * (uint64_t)(in->steps - in->n0); }

// Whether a request belongs to the large-request lane (policy
// WV_SERVE_LANES): with lane_cells 0 (the default), the plans that take two
// or three steps per pass (the plain run's pass mode), whose extra field
// buffers the lane's multi-step kernels use, on grids of at least 2.25 Mi
// cells (every such FP32 plan; FP64 plans take two steps from 0.75 Mi, and
// those below 2.25 Mi stay in the batch); else multi-step plans of at least
// lane_cells air cells. On the 40-request trace the default took the p50 /
// p95 / p99 latency from 0.48 / 2.69 / 2.92 s (three-step plans only) to
// 0.38 / 2.62 / 2.90 s at 2.08 requests/s, and burst capacity from 3.48 to
// 3.42 solves/s.
bool in_large_lane(const wv_server* sv, const wv_plan* pl) {
    const uint64_t lane = sv->lane_cells.load();
    if (pl->run_steps_per_pass < 2) return false;
    return lane ? pl->air >= lane : pl->cells >= kPairMinCells;
}

// The instance for the large-request lane's next block (policy
// WV_SERVE_LANES): of the instances in the lane (in_large_lane), the one the
// lane has started, else the earliest
// admitted (so the lane serves large requests first come, first served, as
// a queue of plain runs would), or null.
Instance* pick_lane_instance(wv_server* sv) {
    Instance* first = nullptr;
    for (auto& up : sv->batch->inst) {
        Instance* in = up.get();
        if (in->done || in->n0 >= in->steps || in->lane >= 0) continue;
        if (!in_large_lane(sv, in->plan.get())) continue;
        if (in->n0 > 0) return in;
        if (!first) first = in;
    }
    return first;
}

// Chooses the instances the batch stream's next block advances and pauses
// the others. WV_SERVE_ALL: every admitted instance. WV_SERVE_SHORTEST and
// WV_SERVE_LANES: least remaining work (air cells x steps left) first,
// earliest admitted among equals; an instance joins while the chosen air
// cells stay within the fill (the first always runs). WV_SERVE_LANES leaves
// out the instances that belong to the large-request lane. Instances in
// flight on the other lane are never chosen. Returns the chosen instances.
std::vector<Instance*> select_batch(wv_server* sv) {
    wv_batch* b = sv->batch.get();
    const int policy = sv->policy.load();
    std::vector<Instance*> live;
    for (auto& up : b->inst) {
        Instance* in = up.get();
        if (in->done || in->n0 >= in->steps) continue;
        if (in->lane >= 0) {
            if (!in->paused) b->dirty = true;
            in->paused = true;
            continue;
        }
        live.push_back(in);
    }
    std::vector<char> run(live.size(), 1);
    if (policy != WV_SERVE_ALL) {
        std::vector<size_t> order;
        for (size_t k = 0; k < live.size(); ++k) {
            const wv_plan* pl = live[k]->plan.get();
            if (policy == WV_SERVE_LANES && in_large_lane(sv, pl)) run[k] = 0;
            else order.push_back(k);
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t x, size_t y) { return work_left(live[x]) < work_left(live[y]); });
        const uint64_t fill = sv->fill.load();
        uint64_t cells = 0;
        for (size_t j = 0; j < order.size(); ++j) {
            const uint64_t air = live[order[j]]->plan->air;
            run[order[j]] = j == 0 || cells + air <= fill;
            if (run[order[j]]) cells += air;
        }
    }
    std::vector<Instance*> chosen;
    for (size_t k = 0; k < live.size(); ++k) {
        const bool paused = !run[k];
        if (live[k]->paused != paused) {
            live[k]->paused = paused;
            b->dirty = true;
        }
        if (!paused) chosen.push_back(live[k]);
    }
    return chosen;
}

void fail_request(wv_server* sv, const std::shared_ptr<Request>& r, wv_status st, const std::string& msg) {
    r->status = st;
    r->err = msg;
    r->done = true;
    r->t_done = now_s();
    ++sv->failed;
}

// Admits queued requests whose footprint fits under the cap (the lock is
// held on entry and exit; released while an instance is set up).
template <typename S>
void admit(wv_server* sv, std::unique_lock<std::mutex>& lk) {
    wv_batch* b = sv->batch.get();
    while (!sv->pending.empty()) {
        auto r = sv->pending.front();
        const uint64_t est = estimate_bytes(r->grid, sv->prec, r->signal.size(), r->recv.size(), r->taps.size());
        if (est > sv->cap) {
            sv->pending.pop_front();
            fail_request(sv, r, WV_ERR_LIMIT, "request needs more device memory than the admission cap");
            sv->cv_done.notify_all();
            continue;
        }
        if (sv->used + est > sv->cap) break;  // waits for running requests to free memory
        sv->pending.pop_front();
        sv->used += est;
        r->reserved = est;
        lk.unlock();
        std::unique_ptr<Instance> in;
        std::string msg;
        wv_status st = WV_OK;
        try {
            in = make_instance<S>(sv->prec, r->grid, r->taps.data(), r->taps.size(), r->signal.data(),
                                  r->signal.size(), r->recv.data(), r->recv.size());
        } catch (const DevErr& e) {
            st = e.code;
            msg = e.msg;
        } catch (const std::bad_alloc&) {
            st = WV_ERR_NOMEM;
            msg = "host allocation failed";
        }
        lk.lock();
        if (st != WV_OK) {
            sv->used -= r->reserved;
            fail_request(sv, r, st, msg);
            sv->cv_done.notify_all();
            continue;
        }
        sv->used += in->footprint;
        sv->used -= r->reserved;
        r->reserved = in->footprint;
        sv->peak_used = std::max(sv->peak_used, sv->used);
        r->t_admit = now_s();
        in->id = (int32_t)r->id;
        r->inst = in.get();
        b->inst.push_back(std::move(in));
        b->dirty = true;
        sv->running.push_back(r);
        sv->max_active = std::max<uint32_t>(sv->max_active, (uint32_t)sv->running.size());
    }
}

// Finished instances that no lane has in flight: results to host memory,
// device memory freed, waiters woken. Called without the lock.
void finish_done(wv_server* sv) {
    wv_batch* b = sv->batch.get();
    std::vector<std::shared_ptr<Request>> fin;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        for (auto& r : sv->running)
            if (r->inst && r->inst->lane < 0 && r->inst->n0 >= r->inst->steps) fin.push_back(r);
    }
    for (auto& r : fin) {
        Instance* in = r->inst;
        in->done = true;
        in->plan->done = in->steps;
        in->plan->steps = in->steps;
        in->plan->active = true;
        const size_t R = in->recv.size(), Tn = in->steps, nf = r->freqs.size();
        if (r->mode & WV_MODE_IR) r->ir.assign(R * Tn, 0.0);
        if ((r->mode & WV_MODE_SPECTRA) && nf) r->spectra.assign(2 * R * nf, 0.0);
        run_finish(in->plan.get(), r->freqs.data(), nf, r->mode, r->spectra.empty() ? nullptr : r->spectra.data(),
                   r->ir.empty() ? nullptr : r->ir.data());
        const uint64_t air = in->plan->air;
        auto it = std::find_if(b->inst.begin(), b->inst.end(),
                               [&](const std::unique_ptr<Instance>& p) { return p.get() == in; });
        b->inst.erase(it);  // frees the instance's device memory
        b->dirty = true;
        std::lock_guard<std::mutex> lk(sv->mu);
        r->inst = nullptr;
        r->done = true;
        r->t_done = now_s();
        sv->t_last = r->t_done;
        sv->used -= r->reserved;
        sv->cell_updates += air * Tn;
        sv->latencies.push_back(r->t_done - r->t_submit);
        ++sv->completed;
        sv->running.erase(std::find(sv->running.begin(), sv->running.end(), r));
    }
    if (!fin.empty()) sv->cv_done.notify_all();
}

// The worker: admission, then each idle lane gets its next block (up to 128
// steps), then a wait until a lane's block ends or a request arrives. Blocks
// on the two lanes run at the same time on the device.
template <typename T, typename S>
void serve_loop(wv_server* sv) {
    wv_batch* b = sv->batch.get();
    Lane lanes[2];
    lanes[0].stream = b->stream;
    lanes[0].ev = sv->lane_ev[0];
    lanes[1].stream = sv->lane_stream;
    lanes[1].ev = sv->lane_ev[1];
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        seen = sv->submitted.load();
        admit<S>(sv, lk);
        lk.unlock();
        for (Lane& l : lanes)
            if (l.busy && lane_finished(l)) {
                for (Instance* in : l.members) in->lane = -1;
                l.members.clear();
                l.busy = false;
            }
        finish_done(sv);
        // The large-request lane.
        Lane& big = lanes[1];
        if (!big.busy && sv->policy.load() == WV_SERVE_LANES) {
            if (Instance* in = pick_lane_instance(sv)) {
                const size_t stop = std::min<size_t>((size_t)in->n0 + kGraphSteps, in->steps);
                enqueue_steps(in->plan.get(), in->n0, stop, big.stream);
                in->n0 = (uint32_t)stop;
                in->lane = 1;
                big.members.push_back(in);
                WV_HIP(hipEventRecord(big.ev, big.stream));
                big.busy = true;
                ++sv->lane_blocks;
            }
        }
        // The batch stream: the batched launch, or one instance alone.
        Lane& bat = lanes[0];
        if (!bat.busy) {
            std::vector<Instance*> chosen = select_batch(sv);
            if (!chosen.empty()) {
                if (chosen.size() == 1 && chosen[0]->plan->run_steps_per_pass >= 2) {
                    enqueue_solo(b, chosen[0]);
                } else {
                    if (b->dirty) build_table<T, S>(b);
                    enqueue_block<T, S>(b, true);
                }
                for (Instance* in : chosen) in->lane = 0;
                bat.members = std::move(chosen);
                WV_HIP(hipEventRecord(bat.ev, bat.stream));
                bat.busy = true;
            }
        }
        const bool busy = lanes[0].busy || lanes[1].busy;
        if (busy) {
            // Until a block ends or a request is submitted.
            for (uint32_t spin = 0;; ++spin) {
                if ((lanes[0].busy && lane_finished(lanes[0])) || (lanes[1].busy && lane_finished(lanes[1]))) break;
                if (sv->submitted.load() != seen) break;
                if (spin < 256) std::this_thread::yield();
                else std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
        lk.lock();
        if (!busy && sv->running.empty()) {
            if (sv->stop && sv->pending.empty()) break;
            // Idle: sleep until a request arrives (with nothing running, any
            // pending request fits under the cap and is admitted above).
            if (sv->pending.empty()) sv->cv_work.wait(lk);
        }
    }
}

// The worker thread. Whatever ends it early (a device error, a failed host
// allocation, any other exception) fails every running and queued request,
// so no wv_serve_wait is left waiting, and later submissions are refused.
void serve_main(wv_server* sv) {
    wv_status code = WV_OK;
    std::string msg;
    try {
        WV_HIP(hipSetDevice(sv->device));
        auto run = [&](auto t, auto s) -> wv_status {
            serve_loop<decltype(t), decltype(s)>(sv);
            return WV_OK;
        };
        dispatch_prec(sv->prec, run);
    } catch (const DevErr& e) {
        code = e.code;
        msg = e.msg;
    } catch (const std::bad_alloc&) {
        code = WV_ERR_NOMEM;
        msg = "host allocation failed";
    } catch (const std::exception& e) {
        code = WV_ERR_DEVICE;
        msg = std::string("server worker: ") + e.what();
    }
    if (code != WV_OK) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->fatal = msg;
        for (auto& r : sv->running) fail_request(sv, r, code, msg);
        for (auto& r : sv->pending) fail_request(sv, r, code, msg);
        sv->running.clear();
        sv->pending.clear();
    }
    sv->cv_done.notify_all();
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

// ------------------------------------------------------------------ server C ABI

extern "C" wv_status wv_serve_create(int precision, uint64_t hbm_cap_bytes, wv_server** out, char* err,
                                     size_t err_len) {
    if (!out || hbm_cap_bytes == 0 ||
        (precision != WV_FP32 && precision != WV_FP64 && precision != WV_FP32_STATE64)) {
        set_err(err, err_len, "wv_serve_create: invalid argument");
        return WV_ERR_INVALID;
    }
    return guarded(err, err_len, [&]() -> wv_status {
        auto sv = std::make_unique<wv_server>();
        sv->prec = precision;
        sv->cap = hbm_cap_bytes;
        WV_HIP(hipGetDevice(&sv->device));
        sv->batch = std::make_unique<wv_batch>();
        batch_init(sv->batch.get(), precision);
        // A launch per step: the blocks share the device with the large-request
        // lane, which a persistent launch holding every CU's slots would keep
        // waiting until the block ends.
        sv->batch->launch = WV_BATCH_STEPS;
        WV_HIP(hipStreamCreateWithFlags(&sv->lane_stream, hipStreamNonBlocking));
        for (auto& e : sv->lane_ev) WV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        wv_server* p = sv.get();
        sv->worker = std::thread([p]() { serve_main(p); });
        *out = sv.release();
        return WV_OK;
    });
}

extern "C" void wv_serve_free(wv_server* sv) { delete sv; }

extern "C" wv_status wv_serve_submit(wv_server* sv, const wv_grid* grid, const wv_tap* taps, size_t n_taps,
                                     const double* signal, size_t steps, const uint64_t* receivers, size_t n_recv,
                                     const double* freqs, size_t nf, int mode, uint64_t* id, char* err,
                                     size_t err_len) {
    if (!sv || !id || (nf && !freqs) || !(mode & (WV_MODE_IR | WV_MODE_SPECTRA))) {
        set_err(err, err_len, "wv_serve_submit: invalid argument");
        return WV_ERR_INVALID;
    }
    wv_status s = check_grid(grid, err, err_len);
    if (s != WV_OK) return s;
    wv_plan probe;
    probe.nx = (uint32_t)grid->n[0];
    probe.ny = (uint32_t)grid->n[1];
    probe.nz = (uint32_t)grid->n[2];
    probe.cells = grid->n[0] * grid->n[1] * grid->n[2];
    s = check_run_args(&probe, taps, n_taps, signal, steps, receivers, n_recv, err, err_len);
    if (s != WV_OK) return s;
    try {
        auto r = std::make_shared<Request>();
        r->grid = grid;
        r->taps.assign(taps, taps + n_taps);
        r->signal.assign(signal, signal + steps);
        r->recv.assign(receivers, receivers + n_recv);
        if (nf) r->freqs.assign(freqs, freqs + nf);
        r->mode = mode;
        r->t_submit = now_s();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            if (!sv->fatal.empty()) {
                set_err(err, err_len, "server stopped: " + sv->fatal);
                return WV_ERR_DEVICE;
            }
            r->id = sv->next_id++;
            if (sv->t_first < 0) sv->t_first = r->t_submit;
            sv->reqs[r->id] = r;
            sv->pending.push_back(r);
            *id = r->id;
        }
        sv->submitted.fetch_add(1);
        sv->cv_work.notify_all();
        return WV_OK;
    } catch (const std::bad_alloc&) {
        set_err(err, err_len, "host allocation failed");
        return WV_ERR_NOMEM;
    }
}

extern "C" wv_status wv_serve_wait(wv_server* sv, uint64_t id, double* spectra_out, double* ir_out,
                                   double* latency_s, char* err, size_t err_len) {
    if (!sv) return WV_ERR_INVALID;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) {
        set_err(err, err_len, "wv_serve_wait: unknown request id");
        return WV_ERR_INVALID;
    }
    auto r = it->second;
    sv->cv_done.wait(lk, [&]() { return r->done; });
    sv->reqs.erase(id);
    if (r->status != WV_OK) {
        set_err(err, err_len, r->err);
        return r->status;
    }
    if (latency_s) *latency_s = r->t_done - r->t_submit;
    const size_t R = r->recv.size(), Tn = r->signal.size();
    if (ir_out && !r->ir.empty()) std::memcpy(ir_out, r->ir.data(), R * Tn * sizeof(double));
    if (spectra_out && !r->spectra.empty()) std::memcpy(spectra_out, r->spectra.data(), r->spectra.size() * 8);
    return WV_OK;
}

extern "C" wv_status wv_serve_set_policy(wv_server* sv, int policy, uint64_t fill_cells, uint64_t lane_cells) {
    if (!sv || (policy != WV_SERVE_ALL && policy != WV_SERVE_SHORTEST && policy != WV_SERVE_LANES))
        return WV_ERR_INVALID;
    sv->policy.store(policy);
    sv->fill.store(fill_cells ? fill_cells : kDefaultFillCells);
    sv->lane_cells.store(lane_cells);
    return WV_OK;
}

extern "C" wv_status wv_serve_stats(wv_server* sv, double out[13]) {
    if (!sv || !out) return WV_ERR_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    out[0] = (double)sv->completed;
    out[1] = percentile(sv->latencies, 0.50);
    out[2] = percentile(sv->latencies, 0.95);
    out[3] = percentile(sv->latencies, 0.99);
    out[4] = sv->t_first >= 0 ? sv->t_last - sv->t_first : 0.0;
    out[5] = (double)sv->cell_updates;
    out[6] = (double)sv->max_active;
    out[7] = (double)sv->peak_used;
    out[8] = (double)sv->failed;
    out[9] = (double)(sv->batch ? sv->batch->captures.load() : 0);
    out[10] = (double)(sv->batch ? sv->batch->solo_blocks.load() : 0);
    out[11] = (double)(sv->batch ? sv->batch->graph_launches.load() : 0);
    out[12] = (double)sv->lane_blocks;
    return WV_OK;
}
