// common.hpp — shared host/device helpers of libbolap (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "bolap.h"

struct bolap_trace_rec {
    char name[40];
    hipEvent_t start = nullptr, stop = nullptr;  // null for counts
    int64_t count = 0;
};

// ---------------------------------------------------------------------------
// Context.
// ---------------------------------------------------------------------------
struct bolap_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    int lds_per_cu = 160 * 1024;
    int64_t hbm_bytes = 0;
    char arch[32] = {0};

    // Growable device scratch (partials, hash tables, sort buffers).  Carved
    // per call by WsArena; grown only at a call boundary after a stream sync.
    void* ws = nullptr;
    size_t ws_size = 0;

    // Small device area of self-resetting counters for last-workgroup
    // reductions: zeroed once at creation, every kernel that increments one
    // returns it to zero before exiting.
    unsigned* counters = nullptr;
    static constexpr int kNumCounters = 64;

    // Join state kept between calls: the key list's segment counters and
    // result slots, zeroed at creation and returned to zero by the join's
    // final kernel; join_state_clean says the last join call got that far.
    void* join_state = nullptr;
    static constexpr size_t kJoinStateBytes = 8192;
    bool join_state_clean = true;

    // Device staging for small results that are copied back to the host.
    void* result_dev = nullptr;
    static constexpr size_t kResultBytes = size_t(1) << 20;

    // Pinned host staging for results copied back at the end of a pipeline.
    void* pinned = nullptr;
    size_t pinned_size = 0;

    // Tuning knobs (bolap_tune_set).
    std::map<std::string, int64_t> knobs;
    int64_t knob(const char* name, int64_t dflt = 0) const {
        auto it = knobs.find(name);
        return (it == knobs.end() || it->second == 0) ? dflt : it->second;
    }

    // Batched CEB workload: its own device buffer (per-query bitmaps and
    // count arrays can reach gigabytes, sized exactly per batch) and the
    // instantiated hipGraph of the last batch shape.
    void* ceb_buf = nullptr;
    size_t ceb_size = 0;
    hipGraph_t ceb_graph = nullptr;
    hipGraphExec_t ceb_exec = nullptr;
    std::string ceb_key;
    hipStream_t capture_stream = nullptr;  // capture cannot run on the null stream

    // Device-visible address of `pinned` (kernels write small results there
    // directly, with system-scope stores).
    void* pinned_dev = nullptr;
    // Sequence number of the last result row a kernel was asked to publish
    // in pinned memory (wait_pinned_seq).
    int64_t pinned_seq = 0;

    // Stage tracing (bolap_ctx_set_trace).
    bool trace = false;
    std::vector<bolap_trace_rec> trace_recs;
    std::vector<hipEvent_t> event_pool;
};

namespace bolap {

// Thread-local error text returned by bolap_last_error().
void set_error(const char* fmt, ...);

#define BOLAP_HIP_TRY(expr)                                                        \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            ::bolap::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                               __FILE__, __LINE__);                                \
            return BOLAP_ERR_HIP;                                                  \
        }                                                                          \
    } while (0)

#define BOLAP_ARG_CHECK(cond, ...)                                                 \
    do {                                                                           \
        if (!(cond)) {                                                             \
            ::bolap::set_error(__VA_ARGS__);                                       \
            return BOLAP_ERR_ARG;                                                  \
        }                                                                          \
    } while (0)

#define BOLAP_TRY(expr)                                                            \
    do {                                                                           \
        int _s = (expr);                                                           \
        if (_s != BOLAP_OK) return _s;                                             \
    } while (0)

// Ensures ctx->ws holds at least `bytes`.  When it has to grow, the first
// `keep` bytes are copied into the new buffer (regions planned earlier in
// the same call stay valid as offsets; raw pointers must be re-derived).
int ws_reserve(bolap_ctx* ctx, size_t bytes, size_t keep = 0);
// Resident workgroups per CU of `kern` at `threads` threads and `lds` bytes
// of dynamic LDS on the context's device, with the kernel's dynamic-LDS
// limit raised to `lds` first when needed.  Both runtime queries are made
// once per (device, kernel, threads, LDS) and cached for the process: they
// cost host time on every call of a synchronous pipeline otherwise.
int kernel_occupancy(bolap_ctx* ctx, const void* kern, int threads, size_t lds, int* per_cu);
// Ensures ctx->pinned holds at least `bytes`.
int pinned_reserve(bolap_ctx* ctx, size_t bytes);
// Waits until the pinned int64 at `slot` holds `seq` (written last, with a
// system-scope release store, by the stream's final kernel): a host spin
// instead of a stream synchronisation, with a synchronisation after a
// second as the fallback.  Later work on the stream stays ordered after it.
int wait_pinned_seq(bolap_ctx* ctx, size_t slot, int64_t seq);
// Checks the last kernel launch.
int launch_check(const char* what);
// Looks up this device's copy of the allocation-free join's scratch array
// (join_sum.hip); called by bolap_ctx_create with `device` current.
int noalloc_setup(int device);

// Grouped MIN/MAX in steps (group_minmax.hip), for pipelines that build the
// join rows in the workspace themselves: the expanded row count (one
// synchronisation; it also packs the item columns into the workspace at
// `base`, up to minmax_pack_end, which the caller keeps for the call), the
// workspace end of a call whose own arrays start at `base`, and the call
// (rows in (group, key) order, or with first_seen in the order of the first
// join row and value that reach each group).  A caller whose arrays live in
// the workspace reserves it up to minmax_pack_end before taking their
// addresses: minmax_expanded may not move it then.
size_t minmax_pack_end(const bolap_minmax_args* a, size_t base);
int minmax_expanded(bolap_ctx* ctx, const bolap_minmax_args* a, size_t base, int64_t* expanded);
size_t minmax_ws_end(bolap_ctx* ctx, const bolap_minmax_args* a, int64_t expanded,
                     bool first_seen, size_t base);
int minmax_run(bolap_ctx* ctx, const bolap_minmax_args* a, int64_t expanded, bool first_seen,
               size_t base, bolap_minmax_row* rows_out, int64_t max_rows, int64_t* n_out,
               bolap_minmax_stats* stats);

// An ILIKE pattern as the reference classifies it (query_q10a.cpp:135-173):
// lower-cased, and for patterns without '_' whose '%' signs sit only at the
// ends, the literal token of the whole-string / prefix / suffix / substring
// test (like.hip).
enum class LikeKind { kAll, kExact, kPrefix, kSuffix, kContains, kGeneral };
struct LikePattern {
    LikeKind kind = LikeKind::kGeneral;
    std::string p;      // lower-cased pattern
    std::string token;  // literal part of the whole-string / prefix / suffix / substring forms
};
LikePattern like_classify(const char* raw);

// Records a pair of events around a stage when tracing is on.
struct TraceScope {
    bolap_ctx* ctx;
    int idx = -1;
    TraceScope(bolap_ctx* c, const char* name);
    ~TraceScope();
};
// Records a COUNT entry when tracing is on.
void trace_count(bolap_ctx* ctx, const char* name, int64_t n);

// Bump allocator over the context workspace.  Plan first (sizes only), then
// reserve, then carve with the same sequence of sizes.
struct WsArena {
    size_t off = 0;
    char* base = nullptr;
    static size_t align(size_t v) { return (v + 255) & ~size_t(255); }
    size_t plan(size_t bytes) {
        const size_t at = off;
        off += align(bytes);
        return at;
    }
    template <class T>
    T* at(size_t offset) const {
        return reinterpret_cast<T*>(base + offset);
    }
};

inline bool aligned(const void* p, size_t a) {
    return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0;
}

inline int dtype_size(int32_t dt) {
    switch (dt) {
        case BOLAP_U8: return 1;
        case BOLAP_I16: return 2;
        case BOLAP_I32: return 4;
        case BOLAP_I64: return 8;
        default: return 0;
    }
}

}  // namespace bolap

// ---------------------------------------------------------------------------
// Device helpers.
// ---------------------------------------------------------------------------
#ifdef __HIP_DEVICE_COMPILE__
#define BOLAP_DEV __device__ __forceinline__
#else
#define BOLAP_DEV __device__ inline
#endif

namespace bolap {

constexpr int kWave = 64;

template <typename T, int R>
struct alignas(sizeof(T) * R) VecT {
    T v[R];
};

// One vector load of R consecutive elements starting at element unit*R.
template <typename T, int R>
BOLAP_DEV VecT<T, R> load_unit(const T* __restrict__ p, int64_t unit) {
    return reinterpret_cast<const VecT<T, R>*>(p)[unit];
}

// ---- wave reductions with DPP (VALU only, no LDS round trips) -------------
// Inclusive prefix within each 16-lane row by row_shr 1, 2, 4, 8 (lanes whose
// source falls outside the row read 0); lanes 15, 31, 47 and 63 then hold the
// four row totals, which readlane moves to scalar registers.  Every lane of
// the wave must be active.
template <int CTRL>
BOLAP_DEV uint32_t dpp_u32(uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
BOLAP_DEV uint64_t dpp_u64(uint64_t v) {
    const uint32_t lo = dpp_u32<CTRL>(uint32_t(v));
    const uint32_t hi = dpp_u32<CTRL>(uint32_t(v >> 32));
    return (uint64_t(hi) << 32) | lo;
}
BOLAP_DEV uint64_t readlane_u64(uint64_t v, int lane) {
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v)), lane));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v >> 32)), lane));
    return (uint64_t(hi) << 32) | lo;
}

// Sum over the 64 lanes (mod 2^64); the result is wave-uniform.
BOLAP_DEV uint64_t wave_sum_u64(uint64_t x) {
    x += dpp_u64<0x111>(x);
    x += dpp_u64<0x112>(x);
    x += dpp_u64<0x114>(x);
    x += dpp_u64<0x118>(x);
    return readlane_u64(x, 15) + readlane_u64(x, 31) + readlane_u64(x, 47) +
           readlane_u64(x, 63);
}
// Bitwise OR over the 64 lanes; the result is wave-uniform.
BOLAP_DEV uint32_t wave_or(uint32_t x) {
    x |= dpp_u32<0x111>(x);
    x |= dpp_u32<0x112>(x);
    x |= dpp_u32<0x114>(x);
    x |= dpp_u32<0x118>(x);
    return uint32_t(__builtin_amdgcn_readlane(int(x), 15)) |
           uint32_t(__builtin_amdgcn_readlane(int(x), 31)) |
           uint32_t(__builtin_amdgcn_readlane(int(x), 47)) |
           uint32_t(__builtin_amdgcn_readlane(int(x), 63));
}
BOLAP_DEV int64_t wave_sum(int64_t v) { return int64_t(wave_sum_u64(uint64_t(v))); }
BOLAP_DEV uint64_t wave_sum(uint64_t v) { return wave_sum_u64(v); }

// Two's-complement 128-bit accumulator as (lo, hi).
struct Acc128 {
    uint64_t lo;
    int64_t hi;
};

BOLAP_DEV void add128(Acc128& a, int64_t v) {
    const uint64_t nlo = a.lo + static_cast<uint64_t>(v);
    a.hi += (v >> 63) + (nlo < a.lo ? 1 : 0);
    a.lo = nlo;
}

BOLAP_DEV void add128(Acc128& a, const Acc128& b) {
    const uint64_t nlo = a.lo + b.lo;
    a.hi += b.hi + (nlo < a.lo ? 1 : 0);
    a.lo = nlo;
}

template <int CTRL>
BOLAP_DEV void dpp_step128(Acc128& a) {
    Acc128 o;
    o.lo = dpp_u64<CTRL>(a.lo);
    o.hi = int64_t(dpp_u64<CTRL>(uint64_t(a.hi)));
    add128(a, o);
}

BOLAP_DEV Acc128 wave_sum128(Acc128 a) {
    dpp_step128<0x111>(a);
    dpp_step128<0x112>(a);
    dpp_step128<0x114>(a);
    dpp_step128<0x118>(a);
    Acc128 r{readlane_u64(a.lo, 15), int64_t(readlane_u64(uint64_t(a.hi), 15))};
    for (int lane = 31; lane < 64; lane += 16)
        add128(r, Acc128{readlane_u64(a.lo, lane), int64_t(readlane_u64(uint64_t(a.hi), lane))});
    return r;
}

// Coherent (agent-scope) loads/stores for data exchanged between workgroups
// that may sit on different XCDs (separate L2s).
template <typename T>
BOLAP_DEV T ld_agent(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
BOLAP_DEV void st_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Returns true in every thread of the workgroup that finished last.  Every
// workgroup publishes with a release RMW on the counter (write-back of its
// own stores, no cache invalidation); only the last one runs the acquire
// fence.  All threads must call it after writing their results.  The last
// workgroup must return the counter to zero (reset_counter) when done.
BOLAP_DEV bool last_block_done(unsigned* counter) {
    __shared__ bool s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned t =
            __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    const bool last = s_last;
    if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return last;
}
BOLAP_DEV void reset_counter(unsigned* counter) {
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-wave row counter: ballots accumulate in a wave-uniform register and
// the wave adds its total to a global counter once, at the end of the kernel
// (one atomic per wave instead of one per loop iteration).
struct WaveCount {
    uint64_t n = 0;
    BOLAP_DEV void add(bool p) { n += uint64_t(__popcll(__ballot(p))); }
    BOLAP_DEV void flush(int64_t* dst) const {
        if ((threadIdx.x & (kWave - 1)) == 0 && n)
            atomicAdd(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)n);
    }
};

// Slot of this lane's item in a list whose length is *counter, for lanes
// with `p`; one atomic per wave.  Every lane of the wave must call it;
// lanes without `p` get ~0ull.
BOLAP_DEV unsigned long long wave_append(bool p, unsigned long long* counter) {
    const uint64_t m = __ballot(p);
    if (m == 0) return ~0ull;
    const int leader = __ffsll((long long)m) - 1;
    const uint32_t rank =
        __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
    unsigned long long base = 0;
    if (int(__lane_id()) == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(base)), leader));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(base >> 32)), leader));
    const unsigned long long b = (uint64_t(hi) << 32) | lo;
    return p ? b + rank : ~0ull;
}

// Appends `cnt` entries staged in LDS to `out` (one atomic), then empties
// the stage.  Every thread of the BS-thread workgroup calls it with the same
// cnt.
template <int BS, typename T, typename C>
BOLAP_DEV void flush_stage(T* stage, C* s_n, unsigned long long* s_base, uint32_t cnt,
                           T* __restrict__ out, unsigned long long* __restrict__ n_out) {
    if (threadIdx.x == 0) *s_base = cnt ? atomicAdd(n_out, (unsigned long long)cnt) : 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < cnt; i += BS) out[*s_base + i] = stage[i];
    __syncthreads();
    if (threadIdx.x == 0) *s_n = 0;
    __syncthreads();
}

// 64-bit mixing hash for open-addressing tables (SplitMix64 output function).
BOLAP_DEV uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// 32-bit key hash for LDS/HBM tables: multiplicative (Fibonacci) hashing.
BOLAP_DEV uint32_t hash32(uint32_t k) { return k * 0x9E3779B1u; }

}  // namespace bolap
