// join_sum.hip — filtered dimension -> hash build -> fact probe -> SUM.
//
// Replaces the part scan + lineitem probe of run_q19
// (bespoke_tpch/query_q19.cpp:210-345) and run_q17's bitset semi-join
// (query_q17.cpp:161-262): the dimension filter's surviving keys form the
// build side, every fact row probes it, and matching rows add
// price * (100 - disc) into one exact int64.
//
// Kernels (all on the context stream):
//   count  (exact path only) predicate pass over the dimension, one atomic
//          per wave; the host reads the count to size the table.
//   init   fills the table image with EMPTY, clears bitmap and result slot.
//   build  predicate pass + open-addressing insert (atomicCAS, linear
//          probing) of surviving keys into an HBM table image, plus a
//          blocked Bloom filter (2 bits in one 32-bit word per key); an
//          overflow flag is raised when the
//          distinct keys exceed half the slots (estimate-sized tables).
//   probe  streams the fact key column with 128-bit loads; for the 16 keys
//          a lane holds it first issues all 16 Bloom-word reads, forms a
//          hit mask, and probes the table only for Bloom hits; filter and
//          table are staged in LDS when they fit (one copy per workgroup)
//          and read through L2 otherwise.  Matches
//          gather price/disc and accumulate per lane; DPP wave sums, LDS
//          and per-block partials, summed by a one-workgroup merge kernel.
#include <cstring>

#include "preds.hpp"

namespace bolap {
namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / kWave;
constexpr int32_t kEmpty = INT32_MIN;
constexpr uint32_t kSlotMul = 0x9E3779B1u;   // Fibonacci hashing
constexpr uint32_t kBitMul = 0x85EBCA77u;    // independent multiplier for the bitmap
constexpr int kLdsBudget = 96 * 1024;

// Result slot layout (int64): revenue, matches, build rows, distinct keys,
// flags (bit 0 overflow, bit 1 INT32_MIN key present), counted rows.
enum { kRev = 0, kMatch = 1, kBuild = 2, kDistinct = 3, kFlags = 4, kCount = 5, kResN = 8 };

struct Table {
    int32_t* slots;
    uint32_t* bits;
    int32_t cap_log2;
    int32_t bb_log2;
    uint32_t max_probe;  // insert gives up (overflow) after this many slots
};

BOLAP_DEV uint32_t slot_of(int32_t key, int cap_log2) {
    return (uint32_t(key) * kSlotMul) >> (32 - cap_log2);
}
// Blocked Bloom filter: one 32-bit word per key (chosen by the top bits of
// a second multiplicative hash) with two bits set from middle bits of the
// slot hash.  bb_log2 counts bits; the word index has bb_log2 - 5 bits.
BOLAP_DEV uint32_t bloom_hash(int32_t key) { return uint32_t(key) * kBitMul; }
BOLAP_DEV uint32_t bloom_word(uint32_t h, int bb_log2) { return h >> (32 - (bb_log2 - 5)); }
// Two bit positions from the ten hash bits just below the word index (the
// top bb_log2 - 5 bits pick the word; 10 <= bb_log2 <= 22 keeps the three
// fields disjoint and inside the word).
BOLAP_DEV bool bloom_test(uint32_t w, uint32_t h, int bb_log2) {
    const int sh = 32 - bb_log2;
    const uint32_t b1 = (h >> sh) & 31, b2 = (h >> (sh - 5)) & 31;
    return ((w >> b1) & (w >> b2) & 1u) != 0;
}
BOLAP_DEV uint32_t bloom_bits(uint32_t h, int bb_log2) {
    const int sh = 32 - bb_log2;
    const uint32_t b1 = (h >> sh) & 31, b2 = (h >> (sh - 5)) & 31;
    return (1u << b1) | (1u << b2);
}

__global__ __launch_bounds__(256) void zero_res_k(int64_t* res) {
    if (threadIdx.x < kResN) res[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kThreads) void dim_count_k(PredSet ps, int64_t rows, int64_t* res) {
    WaveCount wc;
    for (int64_t base = int64_t(blockIdx.x) * kThreads; base < rows;
         base += int64_t(gridDim.x) * kThreads) {
        const int64_t row = base + threadIdx.x;
        wc.add(row < rows && eval_preds(ps, row));
    }
    wc.flush(res + kCount);
}

__global__ __launch_bounds__(256) void table_init_k(Table t, int64_t* res) {
    const int64_t cap = int64_t(1) << t.cap_log2;
    const int64_t words = (int64_t(1) << t.bb_log2) / 32;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < cap;
         i += int64_t(gridDim.x) * 256)
        t.slots[i] = kEmpty;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < words;
         i += int64_t(gridDim.x) * 256)
        t.bits[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x < kResN && threadIdx.x != kCount) res[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kThreads) void dim_build_k(PredSet ps, int64_t rows,
                                                        const int32_t* __restrict__ keys,
                                                        Table t, int64_t* res) {
    const uint32_t mask = (uint32_t(1) << t.cap_log2) - 1;
    WaveCount passed;
    uint64_t inserted = 0;
    bool failed = false, has_min = false;
    for (int64_t base = int64_t(blockIdx.x) * kThreads; base < rows;
         base += int64_t(gridDim.x) * kThreads) {
        const int64_t row = base + threadIdx.x;
        const bool pass = row < rows && eval_preds(ps, row);
        passed.add(pass);
        if (!pass || failed) continue;
        const int32_t key = keys[row];
        const uint32_t h = bloom_hash(key);
        atomicOr(t.bits + bloom_word(h, t.bb_log2), bloom_bits(h, t.bb_log2));
        if (key == kEmpty) {  // the table's empty marker: recorded as a flag
            has_min = true;
            continue;
        }
        uint32_t s = slot_of(key, t.cap_log2);
        bool placed = false;
        for (uint32_t probe = 0; probe < t.max_probe; ++probe) {
            const int32_t prev = atomicCAS(t.slots + s, kEmpty, key);
            if (prev == kEmpty || prev == key) {
                inserted += prev == kEmpty;
                placed = true;
                break;
            }
            s = (s + 1) & mask;
        }
        failed = !placed;  // estimate-sized table too small: overflow
    }
    passed.flush(res + kBuild);
    const uint64_t wins = wave_sum(inserted);
    const bool lead = (threadIdx.x & (kWave - 1)) == 0;
    if (lead && wins)
        atomicAdd(reinterpret_cast<unsigned long long*>(res + kDistinct), (unsigned long long)wins);
    const uint64_t flags = (__ballot(failed) ? 1ull : 0ull) | (__ballot(has_min) ? 2ull : 0ull);
    if (lead && flags) atomicOr(reinterpret_cast<unsigned long long*>(res + kFlags), flags);
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

template <typename TP, typename TD>
BOLAP_DEV void add_match(int64_t row, const TP* __restrict__ price, const TD* __restrict__ disc,
                         int64_t& rev, uint32_t& cnt) {
    rev += int64_t(uint64_t(int64_t(price[row])) * uint64_t(100 - int64_t(disc[row])));
    cnt += 1;
}

// Sums the per-block (revenue, matches) partials with one workgroup of NT
// threads into res[kRev], res[kMatch].
template <int NT>
BOLAP_DEV void join_merge_rows(const int64_t* __restrict__ partials, int nblocks, int64_t* res) {
    __shared__ uint64_t s_a[NT / kWave], s_c[NT / kWave];
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
    }
}

// The same merge as a separate kernel after the probe on the same stream.
__global__ __launch_bounds__(256) void join_merge_k(const int64_t* __restrict__ partials,
                                                    int nblocks, int64_t* res) {
    join_merge_rows<256>(partials, nblocks, res);
}

template <bool kLds, typename TP, typename TD, int kProbeU, bool kPipe, bool kLastBlock>
__global__ __launch_bounds__(kThreads) void probe_sum_k(
    const int32_t* __restrict__ fk, const TP* __restrict__ price, const TD* __restrict__ disc,
    int64_t rows, Table t, int64_t* res, int64_t* partials, unsigned* counter) {
    extern __shared__ uint32_t lds[];
    const int64_t words = (int64_t(1) << t.bb_log2) / 32;
    const int64_t cap = int64_t(1) << t.cap_log2;
    const uint32_t* B = t.bits;
    const int32_t* T = t.slots;
    if constexpr (kLds) {
        for (int64_t i = threadIdx.x; i < words; i += kThreads) lds[i] = t.bits[i];
        int32_t* ts = reinterpret_cast<int32_t*>(lds + words);
        for (int64_t i = threadIdx.x; i < cap; i += kThreads) ts[i] = t.slots[i];
        __syncthreads();
        B = lds;
        T = ts;
    }
    const bool has_min = (res[kFlags] & 2) != 0;
    const uint32_t mask = uint32_t(cap - 1);
    int64_t rev = 0;
    uint32_t cnt = 0;
    const int64_t n_units = rows / 4;
    const int64_t stride = int64_t(gridDim.x) * kThreads * kProbeU;
    struct Tile {
        VecT<int32_t, 4> v[kProbeU];
    };
    auto load_tile = [&](Tile& x, int64_t at) {
#pragma unroll
        for (int k = 0; k < kProbeU; ++k) {
            int64_t u = at + int64_t(k) * kThreads;
            u = u < n_units ? u : n_units - 1;  // in bounds; masked when consumed
            x.v[k] = load_unit<int32_t, 4>(fk, u);
        }
    };
    auto use_tile = [&](const Tile& x, int64_t at) {
        // Phase 1: all Bloom words first (independent LDS reads), then a
        // per-lane hit mask over the kProbeU * 4 keys.
        constexpr int kKeys = kProbeU * 4;
        uint32_t hv[kKeys], wv[kKeys];
#pragma unroll
        for (int i = 0; i < kKeys; ++i) {
            hv[i] = bloom_hash(x.v[i / 4].v[i % 4]);
            wv[i] = B[bloom_word(hv[i], t.bb_log2)];
        }
        uint32_t hits = 0;
#pragma unroll
        for (int i = 0; i < kKeys; ++i)
            hits |= (bloom_test(wv[i], hv[i], t.bb_log2) ? 1u : 0u) << i;
        // Units past the end were clamped to the last unit: drop them.
        if (at + int64_t(kProbeU - 1) * kThreads >= n_units) {
#pragma unroll
            for (int k = 0; k < kProbeU; ++k)
                if (at + int64_t(k) * kThreads >= n_units) hits &= ~(0xFu << (4 * k));
        }
        // Phase 2: exact probes only for Bloom hits (rare, divergent).
        if (hits) {
#pragma unroll
            for (int i = 0; i < kKeys; ++i) {
                if (hits & (1u << i)) {
                    const int32_t key = x.v[i / 4].v[i % 4];
                    const int64_t row = (at + int64_t(i / 4) * kThreads) * 4 + (i % 4);
                    if (key == kEmpty ? has_min : table_has(key, T, t.cap_log2, mask))
                        add_match<TP, TD>(row, price, disc, rev, cnt);
                }
            }
        }
    };
    int64_t base = int64_t(blockIdx.x) * kThreads * kProbeU + threadIdx.x;
    if constexpr (!kPipe) {
        for (; base < n_units; base += stride) {
            Tile A;
            load_tile(A, base);
            use_tile(A, base);
        }
    } else if (n_units > 0 && base < n_units) {
        // Two register buffers: the next tile's keys are in flight while
        // the current tile is probed.
        Tile A, Bt;
        load_tile(A, base);
        for (;;) {
            load_tile(Bt, base + stride);
            use_tile(A, base);
            base += stride;
            if (base >= n_units) break;
            load_tile(A, base + stride);
            use_tile(Bt, base);
            base += stride;
            if (base >= n_units) break;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) {
        const int64_t row = n_units * 4 + threadIdx.x;
        if (row < rows) {
            const int32_t key = fk[row];
            const bool h = key == kEmpty ? has_min : table_has(key, T, t.cap_log2, mask);
            if (h) add_match<TP, TD>(row, price, disc, rev, cnt);
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
        join_merge_rows<kThreads>(partials, int(gridDim.x), res);
        if (threadIdx.x == 0) reset_counter(counter);
    }
}

int ceil_log2(int64_t v) {
    int l = 0;
    while ((int64_t(1) << l) < v) ++l;
    return l;
}

struct Sizing {
    int cap_log2, bb_log2;
    int placement;
    size_t lds_bytes;
};

// Table geometry for `n` expected build keys: LDS tables run at load <= 1/4
// with a 128-bit-per-key Bloom filter; HBM tables at load <= 1/2 with 16 bits
// per key.  Either overflows (and the caller re-runs exactly) past `cap / 2`.
int size_table(int64_t n, int placement, Sizing& z) {
    if (n < 1) n = 1;
    const int lds_cap = ceil_log2(n * 4 < 64 ? 64 : n * 4);
    int lds_bb = ceil_log2(n * 128);
    if (lds_bb < 10) lds_bb = 10;
    if (lds_bb > 18) lds_bb = 18;  // 32 KB of LDS
    const size_t lds_bytes = (size_t(1) << lds_cap) * 4 + (size_t(1) << lds_bb) / 8;
    const bool fits = lds_cap <= 30 && lds_bytes <= size_t(kLdsBudget);
    if (placement == BOLAP_PLACE_LDS && !fits) {
        set_error("bolap_join_sum: %lld keys do not fit the LDS budget", (long long)n);
        return BOLAP_ERR_CAPACITY;
    }
    if (fits && placement != BOLAP_PLACE_HBM) {
        z = Sizing{lds_cap, lds_bb, BOLAP_PLACE_LDS, lds_bytes};
        return BOLAP_OK;
    }
    int cap = ceil_log2(n * 2 < 64 ? 64 : n * 2);
    int bb = ceil_log2(n * 16);
    if (bb < 10) bb = 10;
    if (bb > 22) bb = 22;  // 512 KB filter, L2-resident
    if (cap > 31) {
        set_error("bolap_join_sum: %lld keys exceed the table limit", (long long)n);
        return BOLAP_ERR_CAPACITY;
    }
    z = Sizing{cap, bb, BOLAP_PLACE_HBM, 0};
    return BOLAP_OK;
}

template <bool kLds, typename TP, typename TD, int kProbeU, bool kPipe, bool kLast>
int launch_probe_f(bolap_ctx* ctx, const bolap_join_sum_args* a, const Table& t, size_t lds,
                   int64_t* res, WsArena& ws, size_t o_part, int64_t max_blocks) {
    auto kern = probe_sum_k<kLds, TP, TD, kProbeU, kPipe, kLast>;
    if (kLds) {
        BOLAP_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          int(lds)));
    }
    int per_cu = 0;
    BOLAP_HIP_TRY(
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads, kLds ? lds : 0));
    int64_t want = ctx->knob("join.blocks_per_cu", 2);
    if (per_cu > 0 && want > per_cu) want = per_cu;
    int64_t blocks = int64_t(ctx->num_cus) * want;
    const int64_t need = (a->fact_rows / 4 + int64_t(kThreads) * kProbeU - 1) /
                         (int64_t(kThreads) * kProbeU);
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    if (blocks > max_blocks) blocks = max_blocks;
    {
        TraceScope ts(ctx, "join_probe");
        hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(kThreads), kLds ? lds : 0,
                           ctx->stream, a->fact_fk, static_cast<const TP*>(a->fact_price),
                           static_cast<const TD*>(a->fact_disc), a->fact_rows, t, res,
                           ws.at<int64_t>(o_part), ctx->counters + 1);
    }
    BOLAP_TRY(launch_check("probe_sum_k"));
    if (kLast) return BOLAP_OK;
    TraceScope ts(ctx, "join_merge");
    hipLaunchKernelGGL(join_merge_k, dim3(1), dim3(256), 0, ctx->stream, ws.at<int64_t>(o_part),
                       int(blocks), res);
    return launch_check("join_merge_k");
}

// Finish (knob "join.finish"): 1 = last workgroup merges, 2 = merge kernel.
template <bool kLds, typename TP, typename TD, int kProbeU, bool kPipe>
int launch_probe_v(bolap_ctx* ctx, const bolap_join_sum_args* a, const Table& t, size_t lds,
                   int64_t* res, WsArena& ws, size_t o_part, int64_t max_blocks) {
    if (ctx->knob("join.finish", 2) == 1)
        return launch_probe_f<kLds, TP, TD, kProbeU, kPipe, true>(ctx, a, t, lds, res, ws, o_part,
                                                                  max_blocks);
    return launch_probe_f<kLds, TP, TD, kProbeU, kPipe, false>(ctx, a, t, lds, res, ws, o_part,
                                                               max_blocks);
}

// Probe variants (knob "join.probe_variant"): 1 plain U=4, 2 plain U=8,
// 3 pipelined U=2, 4 pipelined U=4 (4-key units per lane per tile).
template <bool kLds, typename TP, typename TD>
int launch_probe(bolap_ctx* ctx, const bolap_join_sum_args* a, const Table& t, size_t lds,
                 int64_t* res, WsArena& ws, size_t o_part, int64_t max_blocks) {
    switch (ctx->knob("join.probe_variant", 3)) {
        case 1: return launch_probe_v<kLds, TP, TD, 4, false>(ctx, a, t, lds, res, ws, o_part, max_blocks);
        case 2: return launch_probe_v<kLds, TP, TD, 8, false>(ctx, a, t, lds, res, ws, o_part, max_blocks);
        case 3: return launch_probe_v<kLds, TP, TD, 2, true>(ctx, a, t, lds, res, ws, o_part, max_blocks);
        default: return launch_probe_v<kLds, TP, TD, 4, true>(ctx, a, t, lds, res, ws, o_part, max_blocks);
    }
}

template <bool kLds>
int dispatch_probe(bolap_ctx* ctx, const bolap_join_sum_args* a, const Table& t, size_t lds,
                   int64_t* res, WsArena& ws, size_t o_part, int64_t max_blocks) {
    const bool p64 = a->price_dtype == BOLAP_I64;
    const bool d64 = a->disc_dtype == BOLAP_I64;
    if (p64 && d64) return launch_probe<kLds, int64_t, int64_t>(ctx, a, t, lds, res, ws, o_part, max_blocks);
    if (p64) return launch_probe<kLds, int64_t, uint8_t>(ctx, a, t, lds, res, ws, o_part, max_blocks);
    if (d64) return launch_probe<kLds, int32_t, int64_t>(ctx, a, t, lds, res, ws, o_part, max_blocks);
    return launch_probe<kLds, int32_t, uint8_t>(ctx, a, t, lds, res, ws, o_part, max_blocks);
}

inline unsigned small_grid(bolap_ctx* ctx, int64_t rows, int threads) {
    int64_t b = (rows + threads - 1) / threads;
    const int64_t cap = int64_t(ctx->num_cus) * 8;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

// One attempt with a given expected build size; exact == sized from a count.
int join_attempt(bolap_ctx* ctx, const bolap_join_sum_args* a, const PredSet& ps,
                 int64_t n_expect, bool exact, int64_t* host_res, Sizing& z) {
    BOLAP_TRY(size_table(n_expect, a->placement, z));
    const int64_t cap = int64_t(1) << z.cap_log2;
    const int64_t words = (int64_t(1) << z.bb_log2) / 32;
    const int64_t max_blocks = int64_t(ctx->num_cus) * 8;
    WsArena ws;
    const size_t o_slots = ws.plan(size_t(cap) * 4);
    const size_t o_bits = ws.plan(size_t(words) * 4);
    const size_t o_part = ws.plan(size_t(max_blocks) * 2 * sizeof(int64_t));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    int64_t* res = static_cast<int64_t*>(ctx->result_dev);
    // Exact tables (load <= 1/2) always place every key; estimate-sized ones
    // give up after 128 slots and fall back to exact sizing.
    Table t{ws.at<int32_t>(o_slots), ws.at<uint32_t>(o_bits), z.cap_log2, z.bb_log2,
            exact ? uint32_t(cap) : 128u};
    int64_t init_blocks = (cap + 255) / 256;
    if (init_blocks > 4096) init_blocks = 4096;
    {
        TraceScope ts(ctx, "join_init");
        hipLaunchKernelGGL(table_init_k, dim3(unsigned(init_blocks)), dim3(256), 0, ctx->stream,
                           t, res);
    }
    BOLAP_TRY(launch_check("table_init_k"));
    if (a->dim_rows > 0) {
        TraceScope ts(ctx, "join_build");
        hipLaunchKernelGGL(dim_build_k, dim3(small_grid(ctx, a->dim_rows, kThreads)),
                           dim3(kThreads), 0, ctx->stream, ps, a->dim_rows, a->dim_key, t, res);
        BOLAP_TRY(launch_check("dim_build_k"));
    }
    if (z.placement == BOLAP_PLACE_LDS)
        BOLAP_TRY(dispatch_probe<true>(ctx, a, t, z.lds_bytes, res, ws, o_part, max_blocks));
    else
        BOLAP_TRY(dispatch_probe<false>(ctx, a, t, 0, res, ws, o_part, max_blocks));
    BOLAP_TRY(pinned_reserve(ctx, kResN * sizeof(int64_t)));
    BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, res, kResN * sizeof(int64_t),
                                 hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::memcpy(host_res, ctx->pinned, kResN * sizeof(int64_t));
    if (host_res[kDistinct] > cap / 2) host_res[kFlags] |= 1;  // load factor above 1/2
    return BOLAP_OK;
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" int bolap_join_sum(bolap_ctx* ctx, const bolap_join_sum_args* a,
                              bolap_join_sum_result* out) {
    BOLAP_ARG_CHECK(ctx && a && out, "bolap_join_sum: null argument");
    BOLAP_ARG_CHECK(a->dim_rows >= 0 && a->fact_rows >= 0, "bolap_join_sum: negative rows");
    BOLAP_ARG_CHECK(a->dim_rows == 0 || a->dim_key, "bolap_join_sum: null dim_key");
    BOLAP_ARG_CHECK(a->fact_rows == 0 || (a->fact_fk && a->fact_price && a->fact_disc),
                    "bolap_join_sum: null fact column");
    BOLAP_ARG_CHECK(a->price_dtype == BOLAP_I64 || a->price_dtype == BOLAP_I32,
                    "bolap_join_sum: price dtype must be i64 or i32");
    BOLAP_ARG_CHECK(a->disc_dtype == BOLAP_I64 || a->disc_dtype == BOLAP_U8,
                    "bolap_join_sum: disc dtype must be i64 or u8");
    BOLAP_ARG_CHECK(a->placement >= BOLAP_PLACE_AUTO && a->placement <= BOLAP_PLACE_HBM,
                    "bolap_join_sum: placement %d", a->placement);
    BOLAP_ARG_CHECK(aligned(a->fact_fk, 16) || a->fact_rows == 0,
                    "bolap_join_sum: fact_fk must be 16-byte aligned");
    PredSet ps;
    BOLAP_TRY(make_predset(ps, a->n_dim_preds, a->dim_preds, a->dim_rows, "bolap_join_sum"));

    int64_t r[kResN];
    Sizing z{};
    int fallbacks = 0;
    bool done = false;
    if (a->capacity_hint > 0) {
        BOLAP_TRY(join_attempt(ctx, a, ps, a->capacity_hint, false, r, z));
        if (r[kFlags] & 1)
            fallbacks = 1;  // estimate too small: exact path below
        else
            done = true;
    }
    if (!done) {
        // Exact sizing: count the surviving dimension rows first.
        int64_t* res = static_cast<int64_t*>(ctx->result_dev);
        hipLaunchKernelGGL(zero_res_k, dim3(1), dim3(256), 0, ctx->stream, res);
        BOLAP_TRY(launch_check("zero_res_k"));
        if (a->dim_rows > 0) {
            TraceScope ts(ctx, "join_count");
            hipLaunchKernelGGL(dim_count_k, dim3(small_grid(ctx, a->dim_rows, kThreads)),
                               dim3(kThreads), 0, ctx->stream, ps, a->dim_rows, res);
            BOLAP_TRY(launch_check("dim_count_k"));
        }
        int64_t count = 0;
        BOLAP_TRY(pinned_reserve(ctx, sizeof(int64_t)));
        BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, res + kCount, sizeof(int64_t),
                                     hipMemcpyDeviceToHost, ctx->stream));
        BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::memcpy(&count, ctx->pinned, sizeof(int64_t));
        BOLAP_TRY(join_attempt(ctx, a, ps, count, true, r, z));
        if (r[kFlags] & 1) {
            set_error("bolap_join_sum: exact-sized table overflowed");
            return BOLAP_ERR_CAPACITY;
        }
    }
    trace_count(ctx, "join_build_rows", r[kBuild]);
    trace_count(ctx, "join_probe_rows_in", a->fact_rows);
    trace_count(ctx, "join_rows_emitted", r[kMatch]);
    out->revenue = r[kRev];
    out->matches = r[kMatch];
    out->build_rows = r[kBuild];
    out->capacity = int64_t(1) << z.cap_log2;
    out->placement = z.placement;
    out->fallbacks = fallbacks;
    return BOLAP_OK;
}
