// join_sum.hip — filtered dimension -> hash build -> fact probe -> SUM.
//
// Replaces the part scan + lineitem probe of run_q19
// (bespoke_tpch/query_q19.cpp:210-345) and run_q17's bitset semi-join
// (query_q17.cpp:161-262): the dimension filter's surviving keys form the
// build side, every fact row probes it, and matching rows add
// price * (100 - disc) into one exact int64.
//
// Kernels (all on the context stream; nothing returns to the host before
// the result):
//   compact  predicate pass over the dimension (R rows per lane, one vector
//            load per predicate column), surviving keys compacted with wave
//            ballots into a key list of 64 segments; workgroup b appends to
//            segment b % 64 with one atomic per wave that has a survivor, so
//            no counter sees more than a 64th of the appends.  The sum of
//            the segment lengths is the exact build size.
//   hbm_init / hbm_build  only when the build side is too large for LDS:
//            an open-addressing table (linear probing, load <= 1/2) and a
//            blocked Bloom filter (two bits in one 32-bit word per key) in
//            HBM, sized on the device from the list length.  Both kernels
//            exit at once when the build side goes to LDS.
//   probe    every workgroup reads the list length, sizes its table from it
//            and, when it fits, builds its own copy in LDS (load <= 1/2,
//            128 filter bits per key) from the L2-resident key list; then it
//            streams the fact key column with 128-bit loads.  For the keys
//            a lane holds it first issues all Bloom-word reads, forms a hit
//            mask, and probes the table only for Bloom hits.  Matches gather
//            price/disc and accumulate per lane; DPP wave sums, LDS and
//            per-block partials, summed by a one-workgroup merge kernel
//            that also writes the result row into pinned host memory.
// An estimate (capacity_hint) fixes the table geometry at launch instead;
// a build side larger than half its slots raises the overflow flag and the
// host reruns with device sizing.
#include <cstring>
#include <mutex>

#include "preds.hpp"

namespace bolap {
namespace {

constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / kWave;
constexpr int kCompactThreads = 256;
constexpr int32_t kEmpty = INT32_MIN;
constexpr uint32_t kSlotMul = 0x9E3779B1u;   // Fibonacci hashing
constexpr uint32_t kBitMul = 0x85EBCA77u;    // independent multiplier for the bitmap
constexpr int kLdsBudget = 96 * 1024;
// Device-sized LDS tables: 2 slots and 2^lds_bits_log2 filter bits per key
// (knob "join.lds_bloom_bits", 7 by default); up to 2048 keys with placement
// AUTO (48 KB at 128 bits per key), up to 4096 keys when LDS is forced.
constexpr int64_t kAutoLdsKeys = 2048;
constexpr int64_t kForcedLdsKeys = 4096;
// Largest device-sized build (placement AUTO) whose filter stays in LDS
// with the table in HBM (32 bits per key at the limit).
constexpr int64_t kHybridKeys = 8192;
// Bloom hits a probe workgroup collects in LDS before checking them
// (beyond it, hits are checked in the loop).
constexpr uint32_t kCandMax = 2048;

// Result slot layout (int64): revenue, matches, build rows, flags
// (bit 1: an INT32_MIN key passed the filter).
enum { kRev = 0, kMatch = 1, kBuild = 2, kFlags = 4, kResN = 8 };

// Segmented key list: segment r holds keys[r * seg_cap .. + cnt[r * kSegStride]).
constexpr int kSegs = 64;
constexpr int kSegStride = 16;  // counters 64 bytes apart
struct KList {
    int32_t* keys;
    uint32_t* cnt;
    int64_t seg_cap;
};

struct Table {
    int32_t* slots;
    uint32_t* bits;
};

// How the probe side sizes its table: from the device count (fixed == 0)
// or from an estimate fixed at launch (fixed == 1).
struct Geo {
    int32_t fixed;
    int32_t placement;      // bolap_placement
    int32_t cap_log2;       // fixed only
    int32_t bb_log2;        // fixed only
    int64_t lds_keys_max;   // device sizing: largest build that goes to LDS
    int32_t lds_bits_log2;  // device sizing: log2 filter bits per key in LDS
    // device sizing, placement AUTO: first attempt that plans no HBM table
    // (no table clearing and build launches); a build larger than
    // lds_keys_max is flagged 4 and the call reruns with the HBM table
    int32_t lds_only;
};

// Table geometry for n build keys.  flag: 0 ok, 1 estimate too small
// (n > slots / 2), 4 forced LDS placement cannot hold n keys.
struct Shape {
    bool lds;
    int32_t cap_log2, bb_log2;
    int32_t flag;
    bool lds_filter;  // table in HBM, filter built in LDS by every probe workgroup
};

__host__ __device__ inline int clog2(int64_t v) {
    int l = 0;
    while ((int64_t(1) << l) < v) ++l;
    return l;
}
__host__ __device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

// Filter bits of an HBM table: 128 per key up to 2^24 bits (2 MB, so that
// the filter stays resident in one XCD's L2 while the fact side streams).
__host__ __device__ inline int hbm_bloom_log2(int64_t m) {
    return clampi(clog2(m * 128), 10, 24);
}

__host__ __device__ inline Shape shape_of(int64_t n, const Geo& g) {
    if (g.fixed) {
        Shape s{g.placement == BOLAP_PLACE_LDS, g.cap_log2, g.bb_log2, 0};
        if (n > (int64_t(1) << g.cap_log2) / 2) s.flag = 1;
        return s;
    }
    const int64_t m = n < 1 ? 1 : n;
    if (g.placement != BOLAP_PLACE_HBM && n <= g.lds_keys_max)
        return Shape{true, clog2(m * 2 < 64 ? 64 : m * 2),
                     clampi(clog2(m << g.lds_bits_log2), 10, 18), 0};
    if (g.placement == BOLAP_PLACE_LDS || g.lds_only) return Shape{true, 0, 0, 4};
    // Mid-size builds keep the filter in LDS (32 KB) and only the table in
    // HBM: the filter test per fact key stays an LDS read, and only the
    // collected candidates probe the HBM table.
    if (g.placement == BOLAP_PLACE_AUTO && n <= kHybridKeys)
        return Shape{false, clog2(m * 2 < 64 ? 64 : m * 2), clampi(clog2(m << 7), 10, 18), 0,
                     true};
    return Shape{false, clog2(m * 2 < 64 ? 64 : m * 2), hbm_bloom_log2(m), 0};
}

__host__ __device__ inline size_t lds_bytes_of(int cap_log2, int bb_log2) {
    return (size_t(1) << cap_log2) * 4 + (size_t(1) << bb_log2) / 8;
}

BOLAP_DEV uint32_t slot_of(int32_t key, int cap_log2) {
    return (uint32_t(key) * kSlotMul) >> (32 - cap_log2);
}
// Blocked Bloom filter: one 32-bit word per key, chosen by the top
// bb_log2 - 5 bits of a multiplicative hash, with two bits set at positions
// taken from the ten hash bits just below (10 <= bb_log2 <= 27 keeps the
// fields disjoint and inside the word).  For lambda keys per word the
// false-positive rate is about (lambda + lambda^2) / 256.  The test costs
// a handful of VALU operations per key, which bounds the probe loop as
// much as the HBM stream does: a three-bit variant (second multiply) cut
// false positives fourfold but made the loop slower.
BOLAP_DEV uint32_t bloom_hash(int32_t key) { return uint32_t(key) * kBitMul; }
BOLAP_DEV uint32_t bloom_word(int32_t key, int bb_log2) {
    return bloom_hash(key) >> (32 - (bb_log2 - 5));
}
BOLAP_DEV uint32_t bloom_mask(int32_t key, int bb_log2) {
    const uint32_t h = bloom_hash(key);
    const int sh = 32 - bb_log2;
    return (1u << ((h >> sh) & 31)) | (1u << ((h >> (sh - 5)) & 31));
}
BOLAP_DEV bool bloom_test(uint32_t w, int32_t key, int bb_log2) {
    const uint32_t h = bloom_hash(key);
    const int sh = 32 - bb_log2;
    return ((w >> ((h >> sh) & 31)) & (w >> ((h >> (sh - 5)) & 31)) & 1u) != 0;
}

// Inserts a key (not the empty marker) with linear probing; the table is
// sized so that every key finds a slot.
BOLAP_DEV void table_insert(int32_t* T, int32_t key, int cap_log2) {
    const uint32_t mask = (uint32_t(1) << cap_log2) - 1;
    uint32_t s = slot_of(key, cap_log2);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const int32_t prev = atomicCAS(T + s, kEmpty, key);
        if (prev == kEmpty || prev == key) return;
        s = (s + 1) & mask;
    }
}

// Exact membership test of a key that passed the Bloom filter.
BOLAP_DEV bool table_has(int32_t key, const int32_t* T, int cap_log2, uint32_t mask) {
    uint32_t s = slot_of(key, cap_log2);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const int32_t t = T[s];
        if (t == key) return true;
        if (t == kEmpty) return false;
        s = (s + 1) & mask;
    }
    return false;
}

__global__ __launch_bounds__(256) void zero_res_k(int64_t* res, uint32_t* seg_cnt) {
    if (threadIdx.x < kResN) res[threadIdx.x] = 0;
    if (threadIdx.x < kSegs) seg_cnt[threadIdx.x * kSegStride] = 0;
}

// Workgroup barrier that orders LDS accesses only: global loads issued
// before it stay in flight (a full __syncthreads waits for them).
BOLAP_DEV void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// U units of 4 fact keys held by a lane.
template <int U>
struct KeyTile {
    VecT<int32_t, 4> v[U];
};

// Inclusive prefix sums of the segment lengths into `ends` (LDS), computed
// by wave 0; returns the build size to every thread of the workgroup.
BOLAP_DEV int64_t klist_prefix(const KList& kl, uint32_t* ends) {
    __shared__ int64_t s_total;
    if (threadIdx.x < kWave) {
        uint32_t c = kl.cnt[threadIdx.x * kSegStride];
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t o = __shfl_up(c, d, kWave);
            if (int(threadIdx.x) >= d) c += o;
        }
        ends[threadIdx.x] = c;
        if (threadIdx.x == kWave - 1) s_total = int64_t(c);
    }
    lds_barrier();
    return s_total;
}

// Key i of the segmented list (0 <= i < build size), found by a binary
// search over the segment end offsets.
BOLAP_DEV int32_t klist_key(const KList& kl, const uint32_t* ends, int64_t i) {
    int lo = 0, hi = kSegs - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (int64_t(ends[mid]) > i) hi = mid; else lo = mid + 1;
    }
    const int64_t start = lo ? int64_t(ends[lo - 1]) : 0;
    return kl.keys[int64_t(lo) * kl.seg_cap + (i - start)];
}

// Build size: the sum of the segment lengths, computed by wave 0 (all 64
// lanes) and returned to every thread of the workgroup; `seg` (LDS, may be
// null) receives the lengths.
BOLAP_DEV int64_t klist_size(const KList& kl, uint32_t* seg) {
    __shared__ int64_t s_n;
    if (threadIdx.x < kWave) {
        const uint32_t c = kl.cnt[threadIdx.x * kSegStride];
        if (seg) seg[threadIdx.x] = c;
        const uint64_t n = wave_sum(uint64_t(c));
        if (threadIdx.x == 0) s_n = int64_t(n);
    }
    __syncthreads();
    return s_n;
}

// Filter + compaction of the dimension keys.  Lane `u` evaluates rows
// u*R .. u*R+R-1; for each of the R row slots one ballot gives the wave's
// survivors, so a lane's offset is the survivors of lower lanes in earlier
// slots plus those of lower lanes in its own slot.  Workgroup b appends to
// segment b % kSegs (the tail rows past the last whole unit go to segment 0).
template <int R>
__global__ __launch_bounds__(kCompactThreads) void dim_compact_k(PredSet ps, int64_t rows,
                                                                 const int32_t* __restrict__ keys,
                                                                 KList kl, int64_t* res) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int64_t units = rows / R;
    const int seg = int(blockIdx.x % kSegs);
    int32_t* __restrict__ out = kl.keys + int64_t(seg) * kl.seg_cap;
    uint32_t* cnt = kl.cnt + seg * kSegStride;
    bool has_min = false;
    for (int64_t base = int64_t(blockIdx.x) * kCompactThreads; base < units;
         base += int64_t(gridDim.x) * kCompactThreads) {
        const int64_t u = base + threadIdx.x;
        const uint32_t m = u < units ? eval_preds_unit<R>(ps, u) : 0u;
        uint64_t bal[R];
        uint32_t total = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            bal[r] = __ballot((m >> r) & 1u);
            total += uint32_t(__popcll(bal[r]));
        }
        if (total == 0) continue;  // wave-uniform
        VecT<int32_t, R> k{};
        if (m) k = load_unit<int32_t, R>(keys, u);
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(cnt, total);
        at = uint32_t(__builtin_amdgcn_readfirstlane(int(at)));
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if ((m >> r) & 1u) {
                const int32_t key = k.v[r];
                out[at + uint32_t(__popcll(bal[r] & lt))] = key;
                has_min |= key == kEmpty;
            }
            at += uint32_t(__popcll(bal[r]));
        }
    }
    // Rows past the last whole unit (fewer than R).
    if (blockIdx.x == 0 && int64_t(threadIdx.x) < rows - units * R) {
        const int64_t row = units * R + threadIdx.x;
        if (eval_preds(ps, row)) {
            const int32_t key = keys[row];
            out[atomicAdd(cnt, 1u)] = key;
            has_min |= key == kEmpty;
        }
    }
    if (__ballot(has_min) && lane == 0)
        atomicOr(reinterpret_cast<unsigned long long*>(res + kFlags), 2ull);
}

// HBM table for a build side that does not go to LDS.  The per-slot
// aggregates agg_sum/agg_cnt (when given) are cleared for any geometry:
// callers that aggregate per build key index them by HBM slot, or by the
// key's list position when the table goes to LDS.
__global__ __launch_bounds__(256) void hbm_init_k(Geo g, Table t, KList kl,
                                                  uint64_t* agg_sum, uint32_t* agg_cnt) {
    const Shape s = shape_of(klist_size(kl, nullptr), g);
    if (s.flag || (s.lds && !agg_sum)) return;
    const int64_t cap = int64_t(1) << s.cap_log2;
    const int64_t words = (int64_t(1) << s.bb_log2) / 32;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < cap;
         i += int64_t(gridDim.x) * 256) {
        if (!s.lds) t.slots[i] = kEmpty;
        if (agg_sum) {
            agg_sum[i] = 0;
            agg_cnt[i] = 0;
        }
    }
    if (s.lds || s.lds_filter) return;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < words;
         i += int64_t(gridDim.x) * 256)
        t.bits[i] = 0;
}

__global__ __launch_bounds__(256) void hbm_build_k(Geo g, Table t, KList kl) {
    __shared__ uint32_t seg[kSegs];
    const Shape s = shape_of(klist_size(kl, seg), g);
    if (s.flag || s.lds) return;
    const int r = int(blockIdx.x % kSegs);
    const int64_t per = gridDim.x / kSegs;
    const int32_t* __restrict__ keys = kl.keys + int64_t(r) * kl.seg_cap;
    for (int64_t i = int64_t(blockIdx.x / kSegs) * 256 + threadIdx.x; i < seg[r]; i += per * 256) {
        const int32_t key = keys[i];
        if (!s.lds_filter)
            atomicOr(t.bits + bloom_word(key, s.bb_log2), bloom_mask(key, s.bb_log2));
        if (key != kEmpty) table_insert(t.slots, key, s.cap_log2);
    }
}

template <typename TP, typename TD>
BOLAP_DEV void add_match(int64_t row, const TP* __restrict__ price, const TD* __restrict__ disc,
                         int64_t& rev, uint32_t& cnt) {
    rev += int64_t(uint64_t(int64_t(price[row])) * uint64_t(100 - int64_t(disc[row])));
    cnt += 1;
}

// Sums the per-block (revenue, matches) partials with one workgroup of NT
// threads into res[kRev], res[kMatch], records the build size in
// res[kBuild], and writes the whole result slot to `host` (pinned host
// memory) with system-scope stores.
template <int NT>
BOLAP_DEV void join_merge_rows(const int64_t* __restrict__ partials, int nblocks, KList kl,
                               int64_t* res, int64_t* host, int64_t seq) {
    __shared__ uint64_t s_a[NT / kWave], s_c[NT / kWave];
    const int64_t n = klist_size(kl, nullptr);
    uint64_t a = 0, c = 0;
    for (int b = threadIdx.x; b < nblocks; b += NT) {
        const VecT<int64_t, 2> v = reinterpret_cast<const VecT<int64_t, 2>*>(partials)[b];
        a += uint64_t(v.v[0]);
        c += uint64_t(v.v[1]);
    }
    a = wave_sum(a);
    c = wave_sum(c);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s_a[threadIdx.x / kWave] = a;
        s_c[threadIdx.x / kWave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t ta = 0, tc = 0;
        for (int w = 0; w < NT / kWave; ++w) {
            ta += s_a[w];
            tc += s_c[w];
        }
        res[kRev] = int64_t(ta);
        res[kMatch] = int64_t(tc);
        res[kBuild] = n;
        for (int i = 0; i < kResN; ++i) {
            const int64_t v = i == kRev ? int64_t(ta) : i == kMatch ? int64_t(tc)
                            : i == kBuild ? n : res[i];
            __hip_atomic_store(host + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        // Published last: the host waits for this value.
        __hip_atomic_store(host + kResN, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        // The next call starts from zeroed result slots and counters.
        for (int i = 0; i < kResN; ++i) res[i] = 0;
    }
    if (threadIdx.x < kSegs) kl.cnt[threadIdx.x * kSegStride] = 0;
}

// The same merge as a separate kernel after the probe on the same stream.
__global__ __launch_bounds__(256) void join_merge_k(const int64_t* __restrict__ partials,
                                                    int nblocks, KList kl, int64_t* res,
                                                    int64_t* host, int64_t seq) {
    join_merge_rows<256>(partials, nblocks, kl, res, host, seq);
}

// The fact-side loop of one workgroup over the Bloom filter B: every key
// that passes the filter is handed to on_cand(row, key), which does the
// exact lookup (the filter and the table live in LDS or in HBM depending on
// where the pointers come from; each call site is inlined with its own
// address space).
// The lane's first tile (units base0, base0 + kThreads, ...), clamped into
// bounds; loaded before the workgroup builds its table so that the HBM
// latency of the first keys overlaps the build.
template <int kProbeU>
BOLAP_DEV KeyTile<kProbeU> first_tile(const int32_t* __restrict__ fk, int64_t rows) {
    const int64_t n_units = rows / 4;
    const int64_t base = int64_t(blockIdx.x) * kThreads * kProbeU + threadIdx.x;
    KeyTile<kProbeU> t{};
    if (n_units > 0) {
#pragma unroll
        for (int k = 0; k < kProbeU; ++k) {
            int64_t u = base + int64_t(k) * kThreads;
            u = u < n_units ? u : n_units - 1;
            t.v[k] = load_unit<int32_t, 4>(fk, u);
        }
    }
    return t;
}

// Fills an LDS filter of 2^bb_log2 bits with the n keys of the list (one
// key per thread per round); the caller orders it with a barrier.
BOLAP_DEV void build_lds_filter(uint32_t* Bl, int bb_log2, const KList& kl, const uint32_t* ends,
                                int64_t n) {
    const int64_t words = (int64_t(1) << bb_log2) / 32;
    for (int64_t i = threadIdx.x; i < words; i += blockDim.x) Bl[i] = 0;
    lds_barrier();
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const int32_t key = klist_key(kl, ends, i);
        atomicOr(Bl + bloom_word(key, bb_log2), bloom_mask(key, bb_log2));
    }
}

// Bloom hits collected by a probe loop (kCollect): (row, key) entries in
// LDS, appended with one LDS atomic per wave; entries past `cap` go to the
// loop's inline on_cand instead.  The caller zeroes *n before the loop
// (ordered by a barrier) and checks the entries after it.
struct CandBuf {
    uint2* e;
    uint32_t* n;
    uint32_t cap;
};

// Appends the lane's hit keys (bit i of `hits` for key i of the tile) to
// cb: a ballot per key slot gives each lane its offset (mbcnt), so the
// wave's only divergent work is the predicated LDS stores.
template <int kKeys, class Row, class OnCand>
BOLAP_DEV void collect_hits(const CandBuf& cb, const int32_t* keys, uint32_t hits, Row&& row_of,
                            OnCand&& on_cand) {
    uint64_t m[kKeys];
    uint32_t pre[kKeys];
    uint32_t tot = 0;
#pragma unroll
    for (int i = 0; i < kKeys; ++i) {
        m[i] = __ballot((hits >> i) & 1u);
        pre[i] = tot;
        tot += uint32_t(__popcll(m[i]));
    }
    const uint64_t act = __ballot(1);
    const int leader = __ffsll(static_cast<long long>(act)) - 1;
    uint32_t base = 0;
    if (int(threadIdx.x & (kWave - 1)) == leader) base = atomicAdd(cb.n, tot);
    base = __shfl(base, leader, kWave);
#pragma unroll
    for (int i = 0; i < kKeys; ++i) {
        if ((hits >> i) & 1u) {
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                uint32_t(m[i] >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m[i]), 0u));
            const uint32_t at = base + pre[i] + below;
            const int64_t row = row_of(i);
            if (at < cb.cap) cb.e[at] = make_uint2(uint32_t(row), uint32_t(keys[i]));
            else on_cand(row, keys[i]);
        }
    }
}

template <int kProbeU, int kPipe, bool kCollect = false, class OnCand>
BOLAP_DEV void probe_rows(const int32_t* __restrict__ fk, int64_t rows, const uint32_t* B,
                          int bb_log2, OnCand&& on_cand,
                          const KeyTile<kProbeU>* first = nullptr, CandBuf cb = {}) {
    const int64_t n_units = rows / 4;
    const int64_t stride = int64_t(gridDim.x) * kThreads * kProbeU;
    using Tile = KeyTile<kProbeU>;
    auto load_tile = [&](Tile& x, int64_t at) {
#pragma unroll
        for (int k = 0; k < kProbeU; ++k) {
            int64_t u = at + int64_t(k) * kThreads;
            u = u < n_units ? u : n_units - 1;  // in bounds; masked when consumed
            x.v[k] = load_unit<int32_t, 4>(fk, u);
        }
    };
    auto use_tile = [&](const Tile& x, int64_t at) {
        // Phase 1: all Bloom words first (independent reads), then a
        // per-lane hit mask over the kProbeU * 4 keys.
        constexpr int kKeys = kProbeU * 4;
        uint32_t wv[kKeys];
#pragma unroll
        for (int i = 0; i < kKeys; ++i) wv[i] = B[bloom_word(x.v[i / 4].v[i % 4], bb_log2)];
        uint32_t hits = 0;
#pragma unroll
        for (int i = 0; i < kKeys; ++i)
            hits |= (bloom_test(wv[i], x.v[i / 4].v[i % 4], bb_log2) ? 1u : 0u) << i;
        // Units past the end were clamped to the last unit: drop them.
        if (at + int64_t(kProbeU - 1) * kThreads >= n_units) {
#pragma unroll
            for (int k = 0; k < kProbeU; ++k)
                if (at + int64_t(k) * kThreads >= n_units) hits &= ~(0xFu << (4 * k));
        }
        // Phase 2: Bloom hits (rare) are collected for an exact check after
        // the loop, or checked at once.
        if constexpr (kCollect) {
            if (__ballot(hits != 0)) {
                int32_t keys[kKeys];
#pragma unroll
                for (int i = 0; i < kKeys; ++i) keys[i] = x.v[i / 4].v[i % 4];
                collect_hits<kKeys>(cb, keys, hits, [&](int i) {
                    return (at + int64_t(i / 4) * kThreads) * 4 + (i % 4);
                }, on_cand);
            }
        } else if (hits) {
#pragma unroll
            for (int i = 0; i < kKeys; ++i) {
                if (hits & (1u << i)) {
                    const int32_t key = x.v[i / 4].v[i % 4];
                    const int64_t row = (at + int64_t(i / 4) * kThreads) * 4 + (i % 4);
                    on_cand(row, key);
                }
            }
        }
    };
    int64_t base = int64_t(blockIdx.x) * kThreads * kProbeU + threadIdx.x;
    if constexpr (!kPipe) {
        // Whole tiles without clamping or masking, then the remaining units
        // one at a time.
        for (; base + int64_t(kProbeU - 1) * kThreads < n_units; base += stride) {
            Tile A;
#pragma unroll
            for (int k = 0; k < kProbeU; ++k)
                A.v[k] = load_unit<int32_t, 4>(fk, base + int64_t(k) * kThreads);
            constexpr int kKeys = kProbeU * 4;
            uint32_t wv[kKeys];
#pragma unroll
            for (int i = 0; i < kKeys; ++i) wv[i] = B[bloom_word(A.v[i / 4].v[i % 4], bb_log2)];
            uint32_t hits = 0;
#pragma unroll
            for (int i = 0; i < kKeys; ++i)
                hits |= (bloom_test(wv[i], A.v[i / 4].v[i % 4], bb_log2) ? 1u : 0u) << i;
            if (hits) {
#pragma unroll
                for (int i = 0; i < kKeys; ++i)
                    if (hits & (1u << i))
                        on_cand((base + int64_t(i / 4) * kThreads) * 4 + (i % 4),
                                A.v[i / 4].v[i % 4]);
            }
        }
        for (; base < n_units; base += kThreads) {
            const VecT<int32_t, 4> v = load_unit<int32_t, 4>(fk, base);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (bloom_test(B[bloom_word(v.v[i], bb_log2)], v.v[i], bb_log2))
                    on_cand(base * 4 + i, v.v[i]);
            }
        }
    } else {
        // Wave-uniform trip count (from the wave's first lane; a lane past
        // the end loads a clamped unit whose hits are masked), so the loop
        // is not structured around a divergent exit, which would copy the
        // loop-carried tile registers at the latch and wait for their loads.
        const int64_t wave_base = int64_t(blockIdx.x) * kThreads * kProbeU +
                                  int64_t(__builtin_amdgcn_readfirstlane(threadIdx.x & ~(kWave - 1)));
        const int64_t trips = n_units > wave_base ? (n_units - wave_base + stride - 1) / stride : 0;
        if constexpr (kPipe == 2) {
            // The scheduling barriers keep the next tile's loads issued
            // before the wait for the current tile (otherwise the compiler
            // hoists the current tile's hashing above them and waits for
            // everything first, halving the loads in flight).
            Tile A, Bt;
            if (trips > 0) {
                if (first) A = *first;
                else load_tile(A, base);
            }
            for (int64_t t = 0; t < trips; t += 2) {
                load_tile(Bt, base + stride);
                __builtin_amdgcn_sched_barrier(0);
                use_tile(A, base);
                if (t + 1 >= trips) break;
                load_tile(A, base + 2 * stride);
                __builtin_amdgcn_sched_barrier(0);
                use_tile(Bt, base + stride);
                base += 2 * stride;
            }
        } else {
            // Three register buffers: two tiles in flight while one is
            // probed.
            Tile A, Bt, C;
            if (trips > 0) {
                if (first) A = *first;
                else load_tile(A, base);
                load_tile(Bt, base + stride);
            }
            for (int64_t t = 0; t < trips; t += 3) {
                load_tile(C, base + 2 * stride);
                __builtin_amdgcn_sched_barrier(0);
                use_tile(A, base);
                if (t + 1 >= trips) break;
                load_tile(A, base + 3 * stride);
                __builtin_amdgcn_sched_barrier(0);
                use_tile(Bt, base + stride);
                if (t + 2 >= trips) break;
                load_tile(Bt, base + 4 * stride);
                __builtin_amdgcn_sched_barrier(0);
                use_tile(C, base + 2 * stride);
                base += 3 * stride;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) {
        const int64_t row = n_units * 4 + threadIdx.x;
        if (row < rows) on_cand(row, fk[row]);
    }
}

template <typename TP, typename TD, int kProbeU, int kPipe, bool kLastBlock, bool kCollect>
__global__ __launch_bounds__(kThreads, 8) void probe_sum_k(
    const int32_t* __restrict__ fk, const TP* __restrict__ price, const TD* __restrict__ disc,
    int64_t rows, Geo g, Table hbm, KList kl, int64_t* res, int64_t* partials,
    unsigned* counter, int64_t* host, int64_t seq) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t ends[kSegs];
    const KeyTile<kProbeU> first = first_tile<kProbeU>(fk, rows);
    const bool has_min = (res[kFlags] & 2) != 0;
    const int64_t n = klist_prefix(kl, ends);
    const Shape s = shape_of(n, g);
    int64_t rev = 0;
    uint32_t cnt = 0;
    // Bloom hits are collected in LDS (kCollect, rows below 2^32) and
    // checked and gathered after the streaming loop; past kCandMax, or
    // without collection, a lane checks and gathers at once.
    __shared__ uint32_t s_cn;
    __shared__ uint2 s_cand[kCandMax];
    const bool collect = kCollect && rows <= int64_t(UINT32_MAX);
    const CandBuf cb{s_cand, &s_cn, collect ? kCandMax : 0u};
    // No build keys (or a table that cannot hold them): nothing matches.
    if (s.flag == 0 && n > 0) {
        const uint32_t mask = (uint32_t(1) << s.cap_log2) - 1;
        const int32_t* T = hbm.slots;
        if (s.lds) {
            const int64_t words = (int64_t(1) << s.bb_log2) / 32;
            const int64_t cap = int64_t(1) << s.cap_log2;
            uint32_t* Bl = lds;
            int32_t* Tl = reinterpret_cast<int32_t*>(lds + words);
            for (int64_t i = threadIdx.x; i < words; i += kThreads) Bl[i] = 0;
            for (int64_t i = threadIdx.x; i < cap; i += kThreads) Tl[i] = kEmpty;
            lds_barrier();
            // One key per thread per round (independent loads), not one
            // segment after another.
            for (int64_t i = threadIdx.x; i < n; i += kThreads) {
                const int32_t key = klist_key(kl, ends, i);
                atomicOr(Bl + bloom_word(key, s.bb_log2), bloom_mask(key, s.bb_log2));
                if (key != kEmpty) table_insert(Tl, key, s.cap_log2);
            }
            if (threadIdx.x == 0) s_cn = 0;
            lds_barrier();
            T = Tl;
            probe_rows<kProbeU, kPipe, kCollect>(fk, rows, Bl, s.bb_log2,
                                                 [&](int64_t row, int32_t key) {
                if (key == kEmpty ? has_min : table_has(key, Tl, s.cap_log2, mask))
                    add_match<TP, TD>(row, price, disc, rev, cnt);
            }, &first, cb);
        } else {
            auto check = [&](int64_t row, int32_t key) {
                if (key == kEmpty ? has_min : table_has(key, T, s.cap_log2, mask))
                    add_match<TP, TD>(row, price, disc, rev, cnt);
            };
            if (s.lds_filter) build_lds_filter(lds, s.bb_log2, kl, ends, n);
            if (threadIdx.x == 0) s_cn = 0;
            lds_barrier();
            // Separate calls keep the LDS filter's reads LDS reads.
            if (s.lds_filter)
                probe_rows<kProbeU, kPipe, kCollect>(fk, rows, lds, s.bb_log2, check, &first, cb);
            else
                probe_rows<kProbeU, kPipe, kCollect>(fk, rows, hbm.bits, s.bb_log2, check, &first,
                                                     cb);
        }
        if constexpr (kCollect) {
            // The collected candidates: exact checks and payload gathers of
            // all threads together instead of stalling the streaming loop.
            __syncthreads();
            const uint32_t nc = s_cn < cb.cap ? s_cn : cb.cap;
            for (uint32_t i = threadIdx.x; i < nc; i += kThreads) {
                const uint2 e = s_cand[i];
                const int32_t key = int32_t(e.y);
                if (key == kEmpty ? has_min : table_has(key, T, s.cap_log2, mask))
                    add_match<TP, TD>(e.x, price, disc, rev, cnt);
            }
        }
    }

    __shared__ int64_t s_rev[kWaves];
    __shared__ int64_t s_cnt[kWaves];
    const int64_t wr = wave_sum(rev);
    const int64_t wc = int64_t(wave_sum(uint64_t(cnt)));
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    if (lane == 0) {
        s_rev[wave] = wr;
        s_cnt[wave] = wc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t a = 0, c = 0;
        for (int w = 0; w < kWaves; ++w) {
            a += s_rev[w];
            c += s_cnt[w];
        }
        partials[2 * int64_t(blockIdx.x)] = a;
        partials[2 * int64_t(blockIdx.x) + 1] = c;
    }
    if constexpr (kLastBlock) {
        if (!last_block_done(counter)) return;
        join_merge_rows<kThreads>(partials, int(gridDim.x), kl, res, host, seq);
        if (threadIdx.x == 0) reset_counter(counter);
    }
}

// Exact slot of a key, or -1.
BOLAP_DEV int32_t table_find(int32_t key, const int32_t* T, int cap_log2, uint32_t mask) {
    uint32_t s = slot_of(key, cap_log2);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const int32_t t = T[s];
        if (t == key) return int32_t(s);
        if (t == kEmpty) return -1;
        s = (s + 1) & mask;
    }
    return -1;
}

// Q17 (bespoke_tpch/query_q17.cpp:119-309): pass 1 adds quantity and 1
// into the part's aggregate for every lineitem row of a passing part; pass
// 2 sums extendedprice over the rows with quantity * 5 * cnt < sum (0.2 *
// average, :288-291).  The aggregates are indexed by the key's position in
// the key list when the table goes to LDS (every workgroup builds its own
// table there; a key listed twice, a duplicated part row, takes its
// smallest position, which does not depend on insertion order) and by the
// HBM slot otherwise.  Pass 1 also records each match as (row, index) in
// the workgroup's own region of `mlist` (list_cap entries) and its count in
// mcnt[block]; pass 2, launched with the same grid, walks that list instead
// of the fact keys, and re-scans the workgroup's rows (the same grid-stride
// share as pass 1) only when its list overflowed.  Part keys are never the
// empty marker (negative part keys are filtered out before the build).
constexpr uint32_t kListOverflow = 0xffffffffu;
constexpr int kQ17Pipe = 2;          // register buffers of the fact-key loop
constexpr uint32_t kQ17CandMax = 1024;  // Bloom hits collected per workgroup

// Slot holding `key` after inserting it (the table must have room).
BOLAP_DEV uint32_t table_insert_at(int32_t* T, int32_t key, int cap_log2) {
    const uint32_t mask = (uint32_t(1) << cap_log2) - 1;
    uint32_t s = slot_of(key, cap_log2);
    for (uint32_t probe = 0; probe < mask; ++probe) {
        const int32_t prev = atomicCAS(T + s, kEmpty, key);
        if (prev == kEmpty || prev == key) return s;
        s = (s + 1) & mask;
    }
    return s;
}

// Dynamic LDS of the Q17 passes: filter, key slots and aggregate indices.
__host__ __device__ inline size_t q17_lds_bytes(int cap_log2, int bb_log2) {
    return lds_bytes_of(cap_log2, bb_log2) + (size_t(1) << cap_log2) * 4;
}

template <int kPass>
__global__ __launch_bounds__(kThreads, 8) void q17_probe_k(
    const int32_t* __restrict__ pk, const int16_t* __restrict__ qty,
    const int32_t* __restrict__ price, int64_t rows, Geo g, Table hbm, KList kl,
    uint64_t* agg_sum, uint32_t* agg_cnt, uint2* mlist, uint32_t* mcnt, int64_t list_cap,
    int64_t* partials) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t ends[kSegs];
    __shared__ uint32_t s_n;
    KeyTile<2> first{};
    if (kPass == 1) first = first_tile<2>(pk, rows);
    const int64_t n = klist_prefix(kl, ends);
    const Shape s = shape_of(n, g);
    uint64_t acc = 0;
    uint32_t cnt = 0;
    uint2* my = mlist + int64_t(blockIdx.x) * list_cap;
    bool scan = s.flag == 0 && n > 0;
    if (kPass == 1) {
        if (threadIdx.x == 0) s_n = 0;
    } else if (scan) {
        const uint32_t m = mcnt[blockIdx.x];
        if (m != kListOverflow) {
            scan = false;
            for (uint32_t i = threadIdx.x; i < m; i += kThreads) {
                const uint2 e = my[i];
                const int64_t q = int64_t(qty[e.x]);
                const int64_t c = int64_t(agg_cnt[e.y]);
                if (q * 5 * c < int64_t(agg_sum[e.y])) {
                    acc += uint64_t(int64_t(price[e.x]));
                    cnt += 1;
                }
            }
        }
    }
    if (scan) {
        const uint32_t mask = (uint32_t(1) << s.cap_log2) - 1;
        auto on_match = [&](int64_t row, uint32_t at) {
            if (kPass == 1) {
                atomicAdd(reinterpret_cast<unsigned long long*>(agg_sum + at),
                          (unsigned long long)int64_t(qty[row]));
                atomicAdd(agg_cnt + at, 1u);
                cnt += 1;
                const uint32_t i = atomicAdd(&s_n, 1u);
                if (int64_t(i) < list_cap) my[i] = make_uint2(uint32_t(row), at);
            } else {
                const int64_t q = int64_t(qty[row]);
                const int64_t c = int64_t(agg_cnt[at]);
                if (q * 5 * c < int64_t(agg_sum[at])) {
                    acc += uint64_t(int64_t(price[row]));
                    cnt += 1;
                }
            }
        };
        // Aggregate index of a key, or -1: the key's list position (LDS
        // table) or its HBM slot.  The empty marker is never a build key (no
        // part key is negative).
        const int32_t* T = hbm.slots;
        const uint32_t* Ix = nullptr;
        if (s.lds) {
            const int64_t words = (int64_t(1) << s.bb_log2) / 32;
            const int64_t cap = int64_t(1) << s.cap_log2;
            uint32_t* Bl = lds;
            int32_t* Tl = reinterpret_cast<int32_t*>(lds + words);
            uint32_t* Il = reinterpret_cast<uint32_t*>(Tl + cap);
            for (int64_t i = threadIdx.x; i < words; i += kThreads) Bl[i] = 0;
            for (int64_t i = threadIdx.x; i < cap; i += kThreads) {
                Tl[i] = kEmpty;
                Il[i] = 0xffffffffu;
            }
            lds_barrier();
            for (int64_t i = threadIdx.x; i < n; i += kThreads) {
                const int32_t key = klist_key(kl, ends, i);
                atomicOr(Bl + bloom_word(key, s.bb_log2), bloom_mask(key, s.bb_log2));
                atomicMin(Il + table_insert_at(Tl, key, s.cap_log2), uint32_t(i));
            }
            T = Tl;
            Ix = Il;
        } else if (s.lds_filter) {
            build_lds_filter(lds, s.bb_log2, kl, ends, n);
        }
        // The filter sits at the start of LDS for both LDS geometries.
        const bool lds_bf = s.lds || s.lds_filter;
        auto find_at = [&](int32_t key) -> int64_t {
            if (key == kEmpty) return -1;
            const int32_t slot = table_find(key, T, s.cap_log2, mask);
            if (slot < 0) return -1;
            return Ix ? int64_t(Ix[slot]) : int64_t(slot);
        };
        auto on_cand = [&](int64_t row, int32_t key) {
            const int64_t at = find_at(key);
            if (at >= 0) on_match(row, uint32_t(at));
        };
        if constexpr (kPass == 1) {
            // Bloom hits collected in LDS, checked after the loop.
            __shared__ uint32_t s_cn;
            __shared__ uint2 s_cand[kQ17CandMax];
            const CandBuf cb{s_cand, &s_cn, rows <= int64_t(UINT32_MAX) ? kQ17CandMax : 0u};
            if (threadIdx.x == 0) s_cn = 0;
            lds_barrier();
            // Separate calls keep an LDS filter's reads LDS reads.
            if (lds_bf)
                probe_rows<2, kQ17Pipe, true>(pk, rows, lds, s.bb_log2, on_cand, &first, cb);
            else
                probe_rows<2, kQ17Pipe, true>(pk, rows, hbm.bits, s.bb_log2, on_cand, &first, cb);
            __syncthreads();
            const uint32_t nc = s_cn < cb.cap ? s_cn : cb.cap;
            for (uint32_t i = threadIdx.x; i < nc; i += kThreads) {
                const uint2 e = s_cand[i];
                on_cand(e.x, int32_t(e.y));
            }
        } else {
            lds_barrier();
            if (lds_bf)
                probe_rows<2, kQ17Pipe>(pk, rows, lds, s.bb_log2, on_cand);
            else
                probe_rows<2, kQ17Pipe>(pk, rows, hbm.bits, s.bb_log2, on_cand);
        }
    }
    __shared__ uint64_t s_a[kWaves], s_c[kWaves];
    const uint64_t wa = wave_sum_u64(acc), wc = wave_sum_u64(uint64_t(cnt));
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s_a[threadIdx.x / kWave] = wa;
        s_c[threadIdx.x / kWave] = wc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t a = 0, c = 0;
        for (int w = 0; w < kWaves; ++w) {
            a += s_a[w];
            c += s_c[w];
        }
        partials[2 * int64_t(blockIdx.x)] = int64_t(a);
        partials[2 * int64_t(blockIdx.x) + 1] = int64_t(c);
        if (kPass == 1) mcnt[blockIdx.x] = int64_t(s_n) <= list_cap ? s_n : kListOverflow;
    }
}

// Sums the two passes' partials: host row = (SUM(extendedprice), rows
// emitted, part rows, matches).
__global__ __launch_bounds__(256) void q17_merge_k(const int64_t* __restrict__ p1,
                                                   const int64_t* __restrict__ p2, int nblocks,
                                                   KList kl, int64_t* host, int64_t seq) {
    const int64_t n = klist_size(kl, nullptr);
    uint64_t m = 0, sum = 0, rows = 0;
    for (int b = threadIdx.x; b < nblocks; b += 256) {
        m += uint64_t(p1[2 * b + 1]);
        sum += uint64_t(p2[2 * b]);
        rows += uint64_t(p2[2 * b + 1]);
    }
    __shared__ uint64_t sh[3][4];
    m = wave_sum_u64(m);
    sum = wave_sum_u64(sum);
    rows = wave_sum_u64(rows);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sh[0][threadIdx.x / kWave] = sum;
        sh[1][threadIdx.x / kWave] = rows;
        sh[2][threadIdx.x / kWave] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // One thread writes the row, then the sequence number (host slot 4)
        // with release semantics: the host waits for that value.
        for (int i = 0; i < 4; ++i) {
            int64_t v = n;
            if (i != 2) {
                const int f = i == 3 ? 2 : i;
                v = int64_t(sh[f][0] + sh[f][1] + sh[f][2] + sh[f][3]);
            }
            __hip_atomic_store(host + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __hip_atomic_store(host + 4, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Table geometry fixed from an expected build size (capacity_hint): LDS
// tables at load <= 1/2 with a 128-bit-per-key Bloom filter; HBM
// tables at load <= 1/2 with 16 bits per key.  Either overflows past
// `cap / 2`.
int size_from_hint(int64_t n, int placement, Geo& g, size_t& lds_bytes) {
    if (n < 1) n = 1;
    const int lds_cap = clog2(n * 2 < 64 ? 64 : n * 2);
    const int lds_bb = clampi(clog2(n * 128), 10, 18);  // at most 32 KB of filter
    const size_t lb = lds_bytes_of(lds_cap, lds_bb);
    const bool fits = lds_cap <= 30 && lb <= size_t(kLdsBudget);
    if (placement == BOLAP_PLACE_LDS && !fits) {
        set_error("bolap_join_sum: %lld keys do not fit the LDS budget", (long long)n);
        return BOLAP_ERR_CAPACITY;
    }
    if (fits && placement != BOLAP_PLACE_HBM) {
        g = Geo{1, BOLAP_PLACE_LDS, lds_cap, lds_bb, 0};
        lds_bytes = lb;
        return BOLAP_OK;
    }
    const int cap = clog2(n * 2 < 64 ? 64 : n * 2);
    if (cap > 31) {
        set_error("bolap_join_sum: %lld keys exceed the table limit", (long long)n);
        return BOLAP_ERR_CAPACITY;
    }
    g = Geo{1, BOLAP_PLACE_HBM, cap, hbm_bloom_log2(n), 0};
    lds_bytes = 0;
    return BOLAP_OK;
}

// Rows per lane of the compaction: 4 when the key and predicate columns are
// aligned for vector loads, else 1.  Knob "join.compact_rows" = 16 gives one
// 16-byte load of a u8 column per lane; measured slower on 2.5M-row
// dimensions (fewer waves in flight), so 4 is the default.
int compact_rows(bolap_ctx* ctx, const int32_t* keys, const PredSet& ps) {
    const int64_t cap = ctx->knob("join.compact_rows", 4);
    if (cap >= 16 && aligned(keys, 64) && preds_aligned(ps, 16)) return 16;
    if (cap >= 4 && aligned(keys, 16) && preds_aligned(ps, 4)) return 4;
    return 1;
}

// Everything one run of the pipeline launches with (device pointers carved
// from the workspace, grid sizes, table geometry).
struct Plan {
    Geo g;
    size_t lds;              // dynamic LDS of the probe kernel
    Table t;                 // HBM table (unused when the build goes to LDS)
    KList kl;
    int64_t* res;
    int64_t* partials;
    int64_t* host;           // device address of the pinned result row
    int64_t seq;             // published after the row (host slot kResN)
    bool state_clean;        // counters and result slots are already zero
    int64_t max_blocks;
    int64_t hbm_cap;
    bool may_hbm;
    int compact_r;           // rows per lane of the compaction (16, 4 or 1)
    unsigned compact_grid;
};

template <typename TP, typename TD, int kProbeU, int kPipe, bool kLast, bool kDefer>
int launch_probe_d(bolap_ctx* ctx, const bolap_join_sum_args* a, const Plan& p) {
    auto kern = probe_sum_k<TP, TD, kProbeU, kPipe, kLast, kDefer>;
    int per_cu = 0;
    BOLAP_TRY(kernel_occupancy(ctx, reinterpret_cast<const void*>(kern), kThreads, p.lds,
                               &per_cu));
    int64_t want = ctx->knob("join.blocks_per_cu", 2);
    if (per_cu > 0 && want > per_cu) want = per_cu;
    int64_t blocks = int64_t(ctx->num_cus) * want;
    const int64_t need = (a->fact_rows / 4 + int64_t(kThreads) * kProbeU - 1) /
                         (int64_t(kThreads) * kProbeU);
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    if (blocks > p.max_blocks) blocks = p.max_blocks;
    {
        TraceScope ts(ctx, "join_probe");
        hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(kThreads), p.lds, ctx->stream,
                           a->fact_fk, static_cast<const TP*>(a->fact_price),
                           static_cast<const TD*>(a->fact_disc), a->fact_rows, p.g, p.t, p.kl,
                           p.res, p.partials, ctx->counters + 1, p.host, p.seq);
    }
    BOLAP_TRY(launch_check("probe_sum_k"));
    if (kLast) return BOLAP_OK;
    TraceScope ts(ctx, "join_merge");
    hipLaunchKernelGGL(join_merge_k, dim3(1), dim3(256), 0, ctx->stream, p.partials, int(blocks),
                       p.kl, p.res, p.host, p.seq);
    return launch_check("join_merge_k");
}

// Collected Bloom hits (knob "join.defer"): 1 = on, 2 = off (checked and
// gathered in the loop).
template <typename TP, typename TD, int kProbeU, int kPipe, bool kLast>
int launch_probe_f(bolap_ctx* ctx, const bolap_join_sum_args* a, const Plan& p) {
    if (ctx->knob("join.defer", 1) == 2)
        return launch_probe_d<TP, TD, kProbeU, kPipe, kLast, false>(ctx, a, p);
    return launch_probe_d<TP, TD, kProbeU, kPipe, kLast, true>(ctx, a, p);
}

// Finish (knob "join.finish"): 1 = last workgroup merges, 2 = merge kernel.
template <typename TP, typename TD, int kProbeU, int kPipe>
int launch_probe_v(bolap_ctx* ctx, const bolap_join_sum_args* a, const Plan& p) {
    if (ctx->knob("join.finish", 2) == 1)
        return launch_probe_f<TP, TD, kProbeU, kPipe, true>(ctx, a, p);
    return launch_probe_f<TP, TD, kProbeU, kPipe, false>(ctx, a, p);
}

// Probe variants (knob "join.probe_variant"): 1 plain U=4, 2 plain U=8,
// 3 double-buffered U=2, 4 double-buffered U=4, 5 triple-buffered U=2
// (4-key units per lane per tile).
template <typename TP, typename TD>
int launch_probe(bolap_ctx* ctx, const bolap_join_sum_args* a, const Plan& p) {
    switch (ctx->knob("join.probe_variant", 3)) {
        case 1: return launch_probe_v<TP, TD, 4, 0>(ctx, a, p);
        case 2: return launch_probe_v<TP, TD, 8, 0>(ctx, a, p);
        case 3: return launch_probe_v<TP, TD, 2, 2>(ctx, a, p);
        case 5: return launch_probe_v<TP, TD, 2, 3>(ctx, a, p);
        default: return launch_probe_v<TP, TD, 4, 2>(ctx, a, p);
    }
}

int dispatch_probe(bolap_ctx* ctx, const bolap_join_sum_args* a, const Plan& p) {
    const bool p64 = a->price_dtype == BOLAP_I64;
    const bool d64 = a->disc_dtype == BOLAP_I64;
    if (p64 && d64) return launch_probe<int64_t, int64_t>(ctx, a, p);
    if (p64) return launch_probe<int64_t, uint8_t>(ctx, a, p);
    if (d64) return launch_probe<int32_t, int64_t>(ctx, a, p);
    return launch_probe<int32_t, uint8_t>(ctx, a, p);
}

inline unsigned grid_for(bolap_ctx* ctx, int64_t items, int threads, int per_cu) {
    int64_t b = (items + threads - 1) / threads;
    const int64_t cap = int64_t(ctx->num_cus) * per_cu;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

// The launches of one run.
int enqueue_join(bolap_ctx* ctx, const bolap_join_sum_args* a, const PredSet& ps,
                 const Plan& p) {
    if (!p.state_clean) {
        hipLaunchKernelGGL(zero_res_k, dim3(1), dim3(256), 0, ctx->stream, p.res, p.kl.cnt);
        BOLAP_TRY(launch_check("zero_res_k"));
    }
    if (a->dim_rows > 0) {
        TraceScope ts(ctx, "join_compact");
        if (p.compact_r == 16)
            hipLaunchKernelGGL(dim_compact_k<16>, dim3(p.compact_grid), dim3(kCompactThreads), 0,
                               ctx->stream, ps, a->dim_rows, a->dim_key, p.kl, p.res);
        else if (p.compact_r == 4)
            hipLaunchKernelGGL(dim_compact_k<4>, dim3(p.compact_grid), dim3(kCompactThreads), 0,
                               ctx->stream, ps, a->dim_rows, a->dim_key, p.kl, p.res);
        else
            hipLaunchKernelGGL(dim_compact_k<1>, dim3(p.compact_grid), dim3(kCompactThreads), 0,
                               ctx->stream, ps, a->dim_rows, a->dim_key, p.kl, p.res);
        BOLAP_TRY(launch_check("dim_compact_k"));
    }
    if (p.may_hbm && a->dim_rows > 0) {
        // Small grids (grid-stride loops): both exit at once when the build
        // side goes to LDS, which is the common case.
        TraceScope ts(ctx, "join_hbm_build");
        hipLaunchKernelGGL(hbm_init_k, dim3(grid_for(ctx, p.hbm_cap, 256, 1)), dim3(256), 0,
                           ctx->stream, p.g, p.t, p.kl, nullptr, nullptr);
        BOLAP_TRY(launch_check("hbm_init_k"));
        hipLaunchKernelGGL(hbm_build_k, dim3(kSegs * 4), dim3(256), 0, ctx->stream, p.g, p.t,
                           p.kl);
        BOLAP_TRY(launch_check("hbm_build_k"));
    }
    return dispatch_probe(ctx, a, p);
}

// Device regions of one run (byte offsets from the start of the region the
// run is carved from) and the plan fields that do not depend on where the
// regions live.
struct JoinLayout {
    size_t o_keys, o_slots, o_bits, o_part;
    size_t bytes;  // total
    int64_t seg_cap;
};

JoinLayout plan_join(bolap_ctx* ctx, const bolap_join_sum_args* a, const PredSet& ps,
                     const Geo& g, size_t lds, Plan& p, WsArena& ws) {
    p = Plan{};
    p.g = g;
    p.lds = lds;
    // HBM table: the largest the build side can need under `g`.
    p.may_hbm = g.fixed ? g.placement == BOLAP_PLACE_HBM
                        : g.placement != BOLAP_PLACE_LDS && !g.lds_only;
    int cap_log2 = 0, bb_log2 = 0;
    if (p.may_hbm) {
        const Shape worst = g.fixed ? Shape{false, g.cap_log2, g.bb_log2, 0}
                                    : shape_of(a->dim_rows, Geo{0, BOLAP_PLACE_HBM, 0, 0, 0});
        cap_log2 = worst.cap_log2;
        bb_log2 = worst.bb_log2;
    }
    p.hbm_cap = p.may_hbm ? int64_t(1) << cap_log2 : 0;
    const int64_t words = p.may_hbm ? (int64_t(1) << bb_log2) / 32 : 0;
    p.max_blocks = int64_t(ctx->num_cus) * 8;
    // Compaction geometry and the segment capacity it implies: workgroup b
    // covers at most ceil(units / (grid * threads)) * threads units of R rows,
    // segment r takes the workgroups b = r (mod kSegs), and segment 0 also
    // takes the fewer than R tail rows.
    p.compact_r = compact_rows(ctx, a->dim_key, ps);
    const int64_t units = a->dim_rows / p.compact_r;
    p.compact_grid = grid_for(ctx, units, kCompactThreads, 16);
    const int64_t per_block =
        (units + int64_t(p.compact_grid) * kCompactThreads - 1) /
        (int64_t(p.compact_grid) * kCompactThreads) * kCompactThreads * p.compact_r;
    const int64_t blocks_per_seg = (int64_t(p.compact_grid) + kSegs - 1) / kSegs;
    JoinLayout L{};
    L.seg_cap = blocks_per_seg * per_block + p.compact_r;
    L.o_keys = ws.plan(size_t(L.seg_cap) * kSegs * 4);
    L.o_slots = ws.plan(size_t(p.hbm_cap) * 4);
    L.o_bits = ws.plan(size_t(words) * 4);
    L.o_part = ws.plan(size_t(p.max_blocks) * 2 * sizeof(int64_t));
    L.bytes = ws.off;
    return L;
}

// Points the plan's regions into `base` (laid out by plan_join) and the
// segment counters and result slots into `state` (kJoinStateBytes).
void carve_join(Plan& p, const JoinLayout& L, char* base, char* state) {
    static_assert(size_t(kSegs) * kSegStride * 4 + kResN * 8 <= bolap_ctx::kJoinStateBytes,
                  "join state too small");
    WsArena ws;
    ws.base = base;
    p.res = reinterpret_cast<int64_t*>(state + size_t(kSegs) * kSegStride * 4);
    p.kl = KList{ws.at<int32_t>(L.o_keys), reinterpret_cast<uint32_t*>(state), L.seg_cap};
    p.t = Table{ws.at<int32_t>(L.o_slots), ws.at<uint32_t>(L.o_bits)};
    p.partials = ws.at<int64_t>(L.o_part);
}

// One run of the pipeline with table geometry `g`; host_res receives the
// result slot.
int join_attempt(bolap_ctx* ctx, const bolap_join_sum_args* a, const PredSet& ps, const Geo& g,
                 size_t lds, int64_t* host_res) {
    Plan p;
    WsArena ws;
    const JoinLayout L = plan_join(ctx, a, ps, g, lds, p, ws);
    BOLAP_TRY(ws_reserve(ctx, L.bytes));
    BOLAP_TRY(pinned_reserve(ctx, (kResN + 1) * sizeof(int64_t)));
    // Segment counters (64 bytes apart) and result slots in the context's
    // join state, which the merge leaves zeroed for the next call.
    carve_join(p, L, static_cast<char*>(ctx->ws), static_cast<char*>(ctx->join_state));
    p.host = static_cast<int64_t*>(ctx->pinned_dev);
    p.seq = ++ctx->pinned_seq;
    p.state_clean = ctx->join_state_clean;
    ctx->join_state_clean = false;  // until this call's merge is known to have run

    BOLAP_TRY(enqueue_join(ctx, a, ps, p));
    // The merge publishes the row and then p.seq; spinning on it returns
    // sooner than a stream synchronisation.
    BOLAP_TRY(wait_pinned_seq(ctx, kResN, p.seq));
    ctx->join_state_clean = true;
    std::memcpy(host_res, ctx->pinned, kResN * sizeof(int64_t));
    return BOLAP_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free variant (bolap_join_sum_noalloc).  Its key list, HBM table,
// filter, partials, counters and result row live in `scratch`, a file-scope
// device array that every GPU the library runs on holds until the process
// ends.  Host code never passes the array's name to a kernel: it is a
// host-side stand-in, not a device address.  The device address of each
// GPU's copy is looked up once per device (noalloc_setup, called by
// bolap_ctx_create with that device current) and kept in g_scratch.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// Scratch layout: join state (counters, result slots), then the published
// result row (kResN slots + sequence), then the run's regions.
constexpr size_t kScratchState = 0;
constexpr size_t kScratchRow = bolap_ctx::kJoinStateBytes;
constexpr size_t kScratchRuns = kScratchRow + 256;

void* scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

int join_noalloc_attempt(bolap_ctx* ctx, const bolap_join_sum_args* a, const PredSet& ps,
                         const Geo& g, size_t lds, int64_t* host_res) {
    char* base = static_cast<char*>(scratch_of(ctx->device));
    if (!base) {
        set_error("bolap_join_sum_noalloc: no scratch address for device %d", ctx->device);
        return BOLAP_ERR_HIP;
    }
    Plan p;
    WsArena ws;
    const JoinLayout L = plan_join(ctx, a, ps, g, lds, p, ws);
    if (L.bytes > kScratchBytes - kScratchRuns) {
        set_error("bolap_join_sum_noalloc: %lld dimension rows need %zu scratch bytes; the "
                  "maximum is %zu (use bolap_join_sum)",
                  (long long)a->dim_rows, L.bytes, kScratchBytes - kScratchRuns);
        return BOLAP_ERR_CAPACITY;
    }
    carve_join(p, L, base + kScratchRuns, base + kScratchState);
    int64_t* row = reinterpret_cast<int64_t*>(base + kScratchRow);
    p.host = row;
    p.seq = 1;
    // The array keeps the previous call's contents: counters and result
    // slots are cleared by the first kernel of every call.
    p.state_clean = false;
    BOLAP_TRY(enqueue_join(ctx, a, ps, p));
    BOLAP_HIP_TRY(hipMemcpyAsync(host_res, row, kResN * sizeof(int64_t), hipMemcpyDeviceToHost,
                                 ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BOLAP_OK;
}

}  // namespace

// Looks up the device address of this device's `scratch` (the current
// device must be `device`).  Called once per device by bolap_ctx_create.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("noalloc_setup: device %d beyond %d", device, kMaxDevices);
        return BOLAP_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return BOLAP_OK;
    void* ptr = nullptr;
    BOLAP_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device] = ptr;
    return BOLAP_OK;
}

namespace {

// Validates the arguments shared by bolap_join_sum and its allocation-free
// variant and builds the predicate set.
int join_check_args(const bolap_join_sum_args* a, PredSet& ps, const char* who) {
    BOLAP_ARG_CHECK(a->dim_rows >= 0 && a->fact_rows >= 0, "%s: negative rows", who);
    BOLAP_ARG_CHECK(a->dim_rows <= (int64_t(1) << 30), "%s: %lld dimension rows "
                    "exceed 2^30", who, (long long)a->dim_rows);
    BOLAP_ARG_CHECK(a->dim_rows == 0 || a->dim_key, "%s: null dim_key", who);
    BOLAP_ARG_CHECK(a->fact_rows == 0 || (a->fact_fk && a->fact_price && a->fact_disc),
                    "%s: null fact column", who);
    BOLAP_ARG_CHECK(a->price_dtype == BOLAP_I64 || a->price_dtype == BOLAP_I32,
                    "%s: price dtype must be i64 or i32", who);
    BOLAP_ARG_CHECK(a->disc_dtype == BOLAP_I64 || a->disc_dtype == BOLAP_U8,
                    "%s: disc dtype must be i64 or u8", who);
    BOLAP_ARG_CHECK(a->placement >= BOLAP_PLACE_AUTO && a->placement <= BOLAP_PLACE_HBM,
                    "%s: placement %d", who, a->placement);
    BOLAP_ARG_CHECK(aligned(a->fact_fk, 16) || a->fact_rows == 0,
                    "%s: fact_fk must be 16-byte aligned", who);
    return make_predset(ps, a->n_dim_preds, a->dim_preds, a->dim_rows, who);
}

// Device-sized geometry (the exact path) and the probe's dynamic LDS.
Geo exact_geo(bolap_ctx* ctx, int32_t placement, size_t& lds) {
    const int64_t lds_keys = placement == BOLAP_PLACE_LDS ? kForcedLdsKeys : kAutoLdsKeys;
    const Geo g{0, placement, 0, 0, lds_keys,
                int32_t(clampi(int(ctx->knob("join.lds_bloom_bits", 7)), 5, 8))};
    lds = 0;
    if (placement != BOLAP_PLACE_HBM) {
        const Shape big = shape_of(lds_keys, g);
        lds = lds_bytes_of(big.cap_log2, big.bb_log2);
    }
    if (placement == BOLAP_PLACE_AUTO) {
        // Room for the LDS filter of a mid-size build as well.
        const Shape mid = shape_of(kHybridKeys, g);
        const size_t fb = (size_t(1) << mid.bb_log2) / 8;
        if (lds < fb) lds = fb;
    }
    return g;
}

void fill_result(bolap_ctx* ctx, const bolap_join_sum_args* a, const int64_t* r, const Geo& g,
                 int fallbacks, bolap_join_sum_result* out) {
    const Shape s = shape_of(r[kBuild], g);
    trace_count(ctx, "join_build_rows", r[kBuild]);
    trace_count(ctx, "join_probe_rows_in", a->fact_rows);
    trace_count(ctx, "join_rows_emitted", r[kMatch]);
    out->revenue = r[kRev];
    out->matches = r[kMatch];
    out->build_rows = r[kBuild];
    out->capacity = int64_t(1) << s.cap_log2;
    out->placement = s.lds ? BOLAP_PLACE_LDS : BOLAP_PLACE_HBM;
    out->fallbacks = fallbacks;
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" int bolap_join_sum(bolap_ctx* ctx, const bolap_join_sum_args* a,
                              bolap_join_sum_result* out) {
    BOLAP_ARG_CHECK(ctx && a && out, "bolap_join_sum: null argument");
    PredSet ps;
    BOLAP_TRY(join_check_args(a, ps, "bolap_join_sum"));

    int64_t r[kResN];
    Geo g{};
    int fallbacks = 0;
    bool done = false;
    if (a->capacity_hint > 0) {
        size_t lds = 0;
        BOLAP_TRY(size_from_hint(a->capacity_hint, a->placement, g, lds));
        BOLAP_TRY(join_attempt(ctx, a, ps, g, lds, r));
        if (shape_of(r[kBuild], g).flag & 1)
            fallbacks = 1;  // estimate too small: device-sized run below
        else
            done = true;
    }
    if (!done) {
        // Exact sizing on the device from the compacted build size.  With
        // placement AUTO the first attempt assumes the build goes to LDS
        // (the common case: no HBM table to clear and build); a larger build
        // is flagged and the call reruns with the HBM table planned.
        size_t lds = 0;
        g = exact_geo(ctx, a->placement, lds);
        if (a->placement == BOLAP_PLACE_AUTO) {
            Geo spec = g;
            spec.lds_only = 1;
            BOLAP_TRY(join_attempt(ctx, a, ps, spec, lds, r));
            if (!(shape_of(r[kBuild], spec).flag & 4)) {
                fill_result(ctx, a, r, spec, fallbacks, out);
                return BOLAP_OK;
            }
        }
        BOLAP_TRY(join_attempt(ctx, a, ps, g, lds, r));
        if (shape_of(r[kBuild], g).flag & 4) {
            set_error("bolap_join_sum: %lld keys do not fit the LDS budget", (long long)r[kBuild]);
            return BOLAP_ERR_CAPACITY;
        }
    }
    fill_result(ctx, a, r, g, fallbacks, out);
    return BOLAP_OK;
}

extern "C" int bolap_join_sum_noalloc(bolap_ctx* ctx, const bolap_join_sum_args* a,
                                      bolap_join_sum_result* out) {
    BOLAP_ARG_CHECK(ctx && a && out, "bolap_join_sum_noalloc: null argument");
    PredSet ps;
    BOLAP_TRY(join_check_args(a, ps, "bolap_join_sum_noalloc"));
    // Exact device sizing only: an estimate could not change the scratch
    // bound, which is the worst case of the dimension size.
    size_t lds = 0;
    const Geo g = exact_geo(ctx, a->placement, lds);
    int64_t r[kResN];
    BOLAP_TRY(join_noalloc_attempt(ctx, a, ps, g, lds, r));
    if (shape_of(r[kBuild], g).flag & 4) {
        set_error("bolap_join_sum_noalloc: %lld keys do not fit the LDS budget",
                  (long long)r[kBuild]);
        return BOLAP_ERR_CAPACITY;
    }
    fill_result(ctx, a, r, g, 0, out);
    return BOLAP_OK;
}

extern "C" int bolap_q17(bolap_ctx* ctx, const bolap_q17_args* a, bolap_q17_result* out) {
    BOLAP_ARG_CHECK(ctx && a && out, "bolap_q17: null argument");
    BOLAP_ARG_CHECK(a->part_rows >= 0 && a->li_rows >= 0, "bolap_q17: negative rows");
    BOLAP_ARG_CHECK(a->part_rows <= (int64_t(1) << 30), "bolap_q17: too many parts");
    BOLAP_ARG_CHECK(a->part_rows == 0 || a->p_partkey, "bolap_q17: null p_partkey");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (a->l_partkey && a->l_quantity && a->l_price),
                    "bolap_q17: null lineitem column");
    BOLAP_ARG_CHECK(a->li_rows == 0 || aligned(a->l_partkey, 16),
                    "bolap_q17: l_partkey must be 16-byte aligned");
    BOLAP_ARG_CHECK(a->n_part_preds >= 0 && a->n_part_preds < BOLAP_MAX_PREDS,
                    "bolap_q17: %d part predicates (max %d)", a->n_part_preds,
                    BOLAP_MAX_PREDS - 1);
    // The reference's part scan skips negative part keys (query_q17.cpp:168-171).
    bolap_pred preds[BOLAP_MAX_PREDS];
    for (int i = 0; i < a->n_part_preds; ++i) preds[i] = a->part_preds[i];
    bolap_pred& nonneg = preds[a->n_part_preds];
    std::memset(&nonneg, 0, sizeof(nonneg));
    nonneg.col = a->p_partkey;
    nonneg.dtype = BOLAP_I32;
    nonneg.op = BOLAP_OP_RANGE;
    nonneg.lo = 0;
    nonneg.hi = INT32_MAX;
    PredSet ps;
    BOLAP_TRY(make_predset(ps, a->n_part_preds + 1, preds, a->part_rows, "bolap_q17"));

    const Geo g{0, BOLAP_PLACE_AUTO, 0, 0, kAutoLdsKeys, 7};
    const Shape worst = shape_of(a->part_rows, Geo{0, BOLAP_PLACE_HBM, 0, 0, 0});
    const Shape lds_max = shape_of(kAutoLdsKeys, g);
    const size_t lds = q17_lds_bytes(lds_max.cap_log2, lds_max.bb_log2);
    int cap_log2 = worst.cap_log2 > lds_max.cap_log2 ? worst.cap_log2 : lds_max.cap_log2;
    int bb_log2 = worst.bb_log2 > lds_max.bb_log2 ? worst.bb_log2 : lds_max.bb_log2;
    const int64_t cap = int64_t(1) << cap_log2;
    const int64_t words = (int64_t(1) << bb_log2) / 32;
    const int64_t max_blocks = int64_t(ctx->num_cus) * 8;
    const int compact_r = compact_rows(ctx, a->p_partkey, ps);
    const int64_t units = a->part_rows / compact_r;
    const unsigned compact_grid = grid_for(ctx, units, kCompactThreads, 16);
    const int64_t per_block = (units + int64_t(compact_grid) * kCompactThreads - 1) /
                              (int64_t(compact_grid) * kCompactThreads) * kCompactThreads *
                              compact_r;
    const int64_t seg_cap = (int64_t(compact_grid) + kSegs - 1) / kSegs * per_block + compact_r;
    WsArena ws;
    const size_t o_cnt = ws.plan(size_t(kSegs) * kSegStride * 4);
    const size_t o_keys = ws.plan(size_t(seg_cap) * kSegs * 4);
    const size_t o_slots = ws.plan(size_t(cap) * 4);
    const size_t o_bits = ws.plan(size_t(words) * 4);
    const size_t o_sum = ws.plan(size_t(cap) * 8);
    const size_t o_acnt = ws.plan(size_t(cap) * 4);
    const size_t o_part = ws.plan(size_t(max_blocks) * 2 * sizeof(int64_t) * 2);
    int64_t blocks = int64_t(ctx->num_cus) * 2;
    const int64_t need = (a->li_rows / 4 + int64_t(kThreads) * 2 - 1) / (int64_t(kThreads) * 2);
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    // Match list: a 32nd of each workgroup's rows (at least 256 entries);
    // row indices are 32-bit, so larger tables always re-scan.
    const int64_t rows_per_block = (a->li_rows + blocks - 1) / blocks;
    int64_t list_cap = rows_per_block / 32 > 256 ? rows_per_block / 32 : 256;
    const int64_t forced_cap = ctx->knob("q17.list_cap", 0);
    if (forced_cap > 0) list_cap = forced_cap;
    if (a->li_rows >= (int64_t(1) << 32)) list_cap = 0;
    const size_t o_list = ws.plan(size_t(blocks) * size_t(list_cap) * sizeof(uint2));
    const size_t o_mcnt = ws.plan(size_t(blocks) * sizeof(uint32_t));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    BOLAP_TRY(pinned_reserve(ctx, kResN * sizeof(int64_t)));  // row + sequence (slot 4)
    ws.base = static_cast<char*>(ctx->ws);
    int64_t* res = static_cast<int64_t*>(ctx->result_dev);
    const KList kl{ws.at<int32_t>(o_keys), ws.at<uint32_t>(o_cnt), seg_cap};
    const Table t{ws.at<int32_t>(o_slots), ws.at<uint32_t>(o_bits)};
    uint64_t* agg_sum = ws.at<uint64_t>(o_sum);
    uint32_t* agg_cnt = ws.at<uint32_t>(o_acnt);
    int64_t* part1 = ws.at<int64_t>(o_part);
    int64_t* part2 = part1 + 2 * max_blocks;

    hipLaunchKernelGGL(zero_res_k, dim3(1), dim3(256), 0, ctx->stream, res, kl.cnt);
    BOLAP_TRY(launch_check("zero_res_k"));
    if (a->part_rows > 0) {
        TraceScope ts(ctx, "q17_part_scan");
        if (compact_r == 16)
            hipLaunchKernelGGL(dim_compact_k<16>, dim3(compact_grid), dim3(kCompactThreads), 0,
                               ctx->stream, ps, a->part_rows, a->p_partkey, kl, res);
        else if (compact_r == 4)
            hipLaunchKernelGGL(dim_compact_k<4>, dim3(compact_grid), dim3(kCompactThreads), 0,
                               ctx->stream, ps, a->part_rows, a->p_partkey, kl, res);
        else
            hipLaunchKernelGGL(dim_compact_k<1>, dim3(compact_grid), dim3(kCompactThreads), 0,
                               ctx->stream, ps, a->part_rows, a->p_partkey, kl, res);
        BOLAP_TRY(launch_check("dim_compact_k"));
    }
    {
        TraceScope ts(ctx, "q17_build");
        hipLaunchKernelGGL(hbm_init_k, dim3(grid_for(ctx, cap, 256, 1)), dim3(256), 0,
                           ctx->stream, g, t, kl, agg_sum, agg_cnt);
        BOLAP_TRY(launch_check("hbm_init_k"));
        hipLaunchKernelGGL(hbm_build_k, dim3(kSegs * 4), dim3(256), 0, ctx->stream, g, t, kl);
        BOLAP_TRY(launch_check("hbm_build_k"));
    }
    uint2* mlist = ws.at<uint2>(o_list);
    uint32_t* mcnt = ws.at<uint32_t>(o_mcnt);
    for (int pass = 1; pass <= 2; ++pass) {
        TraceScope ts(ctx, pass == 1 ? "q17_sum_scan" : "q17_filter_scan");
        auto kern = pass == 1 ? q17_probe_k<1> : q17_probe_k<2>;
        int per_cu = 0;  // raises the kernel's dynamic-LDS limit (once)
        BOLAP_TRY(kernel_occupancy(ctx, reinterpret_cast<const void*>(kern), kThreads, lds,
                                   &per_cu));
        hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(kThreads), lds, ctx->stream,
                           a->l_partkey, a->l_quantity, a->l_price, a->li_rows, g, t, kl,
                           agg_sum, agg_cnt, mlist, mcnt, list_cap,
                           pass == 1 ? part1 : part2);
        BOLAP_TRY(launch_check("q17_probe_k"));
    }
    const int64_t seq = ++ctx->pinned_seq;
    hipLaunchKernelGGL(q17_merge_k, dim3(1), dim3(256), 0, ctx->stream, part1, part2,
                       int(blocks), kl, static_cast<int64_t*>(ctx->pinned_dev), seq);
    BOLAP_TRY(launch_check("q17_merge_k"));
    BOLAP_TRY(wait_pinned_seq(ctx, 4, seq));
    const int64_t* h = static_cast<const int64_t*>(ctx->pinned);
    out->sum_extendedprice = h[0];
    out->rows_emitted = h[1];
    out->parts = h[2];
    out->matches = h[3];
    trace_count(ctx, "q17_part_rows_emitted", out->parts);
    trace_count(ctx, "q17_lineitem_sum_rows_emitted", out->matches);
    trace_count(ctx, "q17_lineitem_filter_rows_emitted", out->rows_emitted);
    return BOLAP_OK;
}
