// star.hip — star join + hash group-by: a fact table semi-joined with 1..3
// filtered dimensions, grouped on one fact column, SUM(price * (100 - disc))
// per group.
//
// Reference shapes: the CEB star queries build per-dimension masks over the
// key domain and then scan the fact table (bespoke_ceb/query_q1a.cpp:69-152,
// 241-621); the TPC-H Q20 pipeline sizes its group map from an estimated
// count and grows it on overflow (bespoke_tpch/query_q20.cpp:132-223,
// 405-417); CEB Q11b sizes linear-probing maps at 2x the expected count
// (bespoke_ceb/query_q11b.cpp:273-362).
//
// Per dimension the passing keys go into either
//   * a 2048-slot open-addressing set (keys stored as key - key_min + 1,
//     0 = empty) that every probe workgroup stages into LDS, when at most
//     BOLAP_STAR_LDS_KEYS keys pass, or
//   * an exact bitmap over [key_min, key_max] in HBM (probed through L2).
// The group table is open addressing in HBM (u32 key - group_min + 1, u64
// revenue); each workgroup pre-aggregates into an LDS table first and spills
// to HBM only on an LDS miss.  Groups come out in key order through a
// bitmap over the key range (popcount prefix sums give each group its rank)
// or, for key ranges above 2^26, through the LSD radix sort.
//
// Sizing: the plain ("exact") path counts each dimension filter, then counts
// the join rows with a probe pass, and sizes from those counts.  The hinted
// path takes the counts from the caller's estimates and skips both passes
// and both host round trips; a table that ends up over half full or an
// insert that exceeds the probe bound flags an overflow, and the query reruns
// exactly (fallbacks = 1).
#include <algorithm>
#include <cstring>

#include "fold.hpp"
#include "preds.hpp"
#include "prims.hpp"
#include "radix.hpp"

namespace bolap {
namespace {

constexpr int kT = 256;
constexpr int kMaxDims = BOLAP_STAR_MAX_DIMS;
constexpr int kSetLog2 = 11;
constexpr int kSetCap = 1 << kSetLog2;  // 2 x BOLAP_STAR_LDS_KEYS
// LDS pre-aggregation table per workgroup: 4096 slots when every group is
// expected to fit (<= 1024 groups), else 2048 slots that catch the hot keys
// (Zipf-skewed group keys would otherwise serialise on a few HBM atomics).
constexpr int kGlLog2Small = 12;
constexpr int kGlLog2Large = 11;
constexpr int64_t kGlMaxGroups = 1024;
constexpr int kGlProbe = 8;
constexpr uint32_t kMul = 0x9E3779B1u;
constexpr int kPlaceSet = BOLAP_PLACE_LDS, kPlaceBits = BOLAP_PLACE_HBM;

static_assert(kSetCap == 2 * BOLAP_STAR_LDS_KEYS, "dimension set sized for load 1/2");

enum {
    sPass0 = 0, sSet0 = kMaxDims, sMatch = 2 * kMaxDims, sGroups, sOver, sRange, sRev, sN
};
enum { kFlagOverflow = 1, kFlagRange = 2 };  // per-lane flags; folded as counts

struct DimK {
    const int32_t* key;
    const int32_t* fk;
    uint32_t* set;     // kSetCap slots or bitmap words
    int64_t rows;
    int64_t span;      // key_max - key_min + 1
    int32_t kmin;
    int32_t place;
    int32_t lds_off;   // offset of the staged set in LDS (u32 words)
    PredSet ps;
};

struct StarK {
    int32_t n;
    int32_t n_lds;     // dimensions staged in LDS
    DimK d[kMaxDims];
};

struct GTab {
    uint32_t* keys;    // cap slots, 0 = empty
    uint64_t* rev;
    int32_t cap_log2;
    uint32_t max_probe;
};

BOLAP_DEV uint32_t slot_of(uint32_t key1, int log2) { return (key1 * kMul) >> (32 - log2); }

static_assert(sN <= kStatLine, "statistics fit a replica line");

__global__ void st_init_k(int64_t* st) {
    if (threadIdx.x < kStatWords) st[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kT) void fill32_k(uint32_t* p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        p[i] = 0u;
}

__global__ __launch_bounds__(kT) void fill64_k(uint64_t* p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        p[i] = 0u;
}

__global__ __launch_bounds__(kT) void dim_count_k(DimK d, int64_t* st, int idx) {
    WaveCount wc;
    for (int64_t base = int64_t(blockIdx.x) * kT; base < d.rows;
         base += int64_t(gridDim.x) * kT) {
        const int64_t r = base + threadIdx.x;
        wc.add(r < d.rows && eval_preds(d.ps, r));
    }
    block_fold<kT>(st, sPass0 + idx, int64_t(wc.n), 0);
}

// Passing keys into the dimension's set or bitmap.  Counts passing rows and
// distinct set entries; flags keys outside [key_min, key_max] and set
// inserts that find no free slot.
__global__ __launch_bounds__(kT) void dim_build_k(DimK d, int64_t* st, int idx) {
    WaveCount wc;
    uint32_t inserted = 0;
    int flags = 0;
    for (int64_t base = int64_t(blockIdx.x) * kT; base < d.rows;
         base += int64_t(gridDim.x) * kT) {
        const int64_t r = base + threadIdx.x;
        const bool p = r < d.rows && eval_preds(d.ps, r);
        wc.add(p);
        if (!p) continue;
        const uint32_t v = uint32_t(d.key[r] - d.kmin);
        if (int64_t(v) >= d.span) {
            flags |= kFlagRange;
            continue;
        }
        if (d.place == kPlaceSet) {
            const uint32_t key1 = v + 1;
            uint32_t s = slot_of(key1, kSetLog2);
            bool done = false;
            for (int probe = 0; probe < kSetCap && !done; ++probe) {
                const uint32_t prev = atomicCAS(d.set + s, 0u, key1);
                if (prev == 0u) {
                    ++inserted;
                    done = true;
                } else if (prev == key1) {
                    done = true;
                }
                s = (s + 1) & (kSetCap - 1);
            }
            if (!done) flags |= kFlagOverflow;
        } else {
            atomicOr(d.set + (v >> 5), 1u << (v & 31));
        }
    }
    block_fold<kT>(st, sPass0 + idx, int64_t(wc.n), 0);
    block_fold<kT>(st, sSet0 + idx, int64_t(wave_sum(uint64_t(inserted))), 0);
    block_fold<kT>(st, sOver, __ballot(flags & kFlagOverflow) ? 1 : 0, 0);
    block_fold<kT>(st, sRange, __ballot(flags & kFlagRange) ? 1 : 0, 0);
}

BOLAP_DEV bool member(const DimK& d, const uint32_t* lds, int32_t k) {
    const uint32_t v = uint32_t(k - d.kmin);
    if (int64_t(v) >= d.span) return false;
    if (d.place == kPlaceSet) {
        const uint32_t key1 = v + 1;
        const uint32_t* set = lds + d.lds_off;
        uint32_t s = slot_of(key1, kSetLog2);
        for (int probe = 0; probe < kSetCap; ++probe) {
            const uint32_t x = set[s];
            if (x == key1) return true;
            if (x == 0u) return false;
            s = (s + 1) & (kSetCap - 1);
        }
        return false;
    }
    return (d.set[v >> 5] >> (v & 31)) & 1u;
}

// Adds rv to group key1 in the HBM table; returns false when no slot was
// found within max_probe.  `fresh` counts the groups this lane created.
BOLAP_DEV bool global_add(const GTab& g, uint32_t key1, uint64_t rv, uint32_t& fresh) {
    const uint32_t mask = (1u << g.cap_log2) - 1u;
    uint32_t s = slot_of(key1, g.cap_log2);
    for (uint32_t probe = 0; probe < g.max_probe; ++probe) {
        uint32_t x = g.keys[s];  // keys only go 0 -> key, so a non-zero read is final
        if (x == 0u) {
            x = atomicCAS(g.keys + s, 0u, key1);
            if (x == 0u) {
                ++fresh;
                x = key1;
            }
        }
        if (x == key1) {
            atomicAdd(reinterpret_cast<unsigned long long*>(g.rev + s), (unsigned long long)rv);
            return true;
        }
        s = (s + 1) & mask;
    }
    return false;
}

template <typename T>
BOLAP_DEV int64_t ld(const T* p, int64_t r) { return int64_t(p[r]); }

// The fact pass.  kMode 0: count join rows only.  kMode 1: aggregate into
// the HBM table.  kMode 2: pre-aggregate in LDS, spill to HBM.
template <typename TP, typename TD, int kMode>
__global__ __launch_bounds__(kT) void star_probe_k(StarK s, GTab g,
                                                   const int32_t* __restrict__ gkey,
                                                   int32_t gmin, int64_t gspan,
                                                   const TP* __restrict__ price,
                                                   const TD* __restrict__ disc, int64_t rows,
                                                   int gl_log2, int64_t* st) {
    extern __shared__ uint32_t lds[];
    const int set_words = s.n_lds * kSetCap;
    const int gl_cap = 1 << gl_log2;
    uint32_t* lkeys = lds + set_words;
    uint64_t* lrev = reinterpret_cast<uint64_t*>(lkeys + gl_cap);
    for (int j = 0; j < s.n; ++j) {
        const DimK& d = s.d[j];
        if (d.place != kPlaceSet) continue;
        for (int i = threadIdx.x; i < kSetCap; i += kT) lds[d.lds_off + i] = d.set[i];
    }
    if (kMode == 2) {
        for (int i = threadIdx.x; i < gl_cap; i += kT) {
            lkeys[i] = 0u;
            lrev[i] = 0u;
        }
    }
    __syncthreads();

    constexpr int U = 4;
    WaveCount matched;
    uint32_t fresh = 0;
    int flags = 0;
    const int64_t stride = int64_t(gridDim.x) * kT;
    for (int64_t base = int64_t(blockIdx.x) * kT; base < rows; base += U * stride) {
        int32_t k[U][kMaxDims];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t r = base + u * stride + threadIdx.x;
            in[u] = r < rows;
#pragma unroll
            for (int j = 0; j < kMaxDims; ++j)
                k[u][j] = (j < s.n && in[u]) ? s.d[j].fk[r] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            bool alive = in[u];
#pragma unroll
            for (int j = 0; j < kMaxDims; ++j)
                if (j < s.n) alive = alive && member(s.d[j], lds, k[u][j]);
            matched.add(alive);
            if (kMode == 0 || !alive) continue;
            const int64_t r = base + u * stride + threadIdx.x;
            const uint32_t v = uint32_t(gkey[r] - gmin);
            if (int64_t(v) >= gspan) {
                flags |= kFlagRange;
                continue;
            }
            const uint32_t key1 = v + 1;
            const uint64_t rv = uint64_t(ld(price, r) * (100 - ld(disc, r)));
            bool done = false;
            if (kMode == 2) {
                uint32_t sl = slot_of(key1, gl_log2);
                for (int probe = 0; probe < kGlProbe && !done; ++probe) {
                    uint32_t x = lkeys[sl];  // keys only go 0 -> key
                    if (x == 0u) x = atomicCAS(lkeys + sl, 0u, key1);
                    if (x == 0u || x == key1) {
                        atomicAdd(reinterpret_cast<unsigned long long*>(lrev + sl),
                                  (unsigned long long)rv);
                        done = true;
                    }
                    sl = (sl + 1) & uint32_t(gl_cap - 1);
                }
            }
            if (!done && !global_add(g, key1, rv, fresh)) flags |= kFlagOverflow;
        }
    }
    if (kMode == 2) {
        __syncthreads();
        for (int i = threadIdx.x; i < gl_cap; i += kT) {
            const uint32_t key1 = lkeys[i];
            if (key1 && !global_add(g, key1, lrev[i], fresh)) flags |= kFlagOverflow;
        }
    }
    block_fold<kT>(st, sMatch, int64_t(matched.n), 0);
    if (kMode != 0) {
        block_fold<kT>(st, sGroups, int64_t(wave_sum(uint64_t(fresh))), 0);
        block_fold<kT>(st, sOver, __ballot(flags & kFlagOverflow) ? 1 : 0, 0);
        block_fold<kT>(st, sRange, __ballot(flags & kFlagRange) ? 1 : 0, 0);
    }
}

// Occupied slots -> (key - group_min, slot) pairs; one atomic per workgroup
// over a contiguous chunk of slots.  Also folds the revenue total.
constexpr int kCT = 1024;
__global__ __launch_bounds__(kCT) void star_compact_k(GTab g, int64_t chunk,
                                                      uint64_t* __restrict__ okeys,
                                                      uint32_t* __restrict__ oslot,
                                                      unsigned long long* __restrict__ n_out,
                                                      uint32_t* __restrict__ bits, int64_t* st) {
    constexpr int kW = kCT / kWave;
    __shared__ uint32_t s_w[kW];
    __shared__ unsigned long long s_base;
    const int64_t cap = int64_t(1) << g.cap_log2;
    const int64_t b0 = int64_t(blockIdx.x) * chunk;
    const int64_t b1 = b0 + chunk < cap ? b0 + chunk : cap;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    uint32_t mine = 0;
    uint64_t rsum = 0;
    for (int64_t i = b0 + threadIdx.x; i < b1; i += kCT)
        if (g.keys[i]) {
            ++mine;
            rsum += g.rev[i];
        }
    const uint32_t wsum = uint32_t(wave_sum(uint64_t(mine)));
    if (lane == 0) s_w[wave] = wsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < kW; ++w) tot += s_w[w];
        s_base = tot ? atomicAdd(n_out, (unsigned long long)tot) : 0;
    }
    __syncthreads();
    unsigned long long base = s_base;
    for (int64_t t0 = b0; t0 < b1; t0 += kCT) {
        const int64_t i = t0 + threadIdx.x;
        const uint32_t key1 = i < b1 ? g.keys[i] : 0u;
        const uint64_t ball = __ballot(key1 != 0u);
        const uint32_t before_lane = __popcll(ball & ((uint64_t(1) << lane) - 1));
        __syncthreads();
        if (lane == 0) s_w[wave] = __popcll(ball);
        __syncthreads();
        uint32_t before = 0, tile = 0;
        for (int w = 0; w < kW; ++w) {
            before += w < wave ? s_w[w] : 0;
            tile += s_w[w];
        }
        if (key1) {
            const unsigned long long at = base + before + before_lane;
            okeys[at] = uint64_t(key1 - 1);
            oslot[at] = uint32_t(i);
            if (bits) atomicOr(bits + ((key1 - 1) >> 5), 1u << ((key1 - 1) & 31));
        }
        base += tile;
    }
    block_fold<kCT>(st, sRev, int64_t(wave_sum(rsum)), 0);
}

__global__ __launch_bounds__(kT) void star_gather_k(const uint64_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ slots,
                                                    const uint64_t* __restrict__ rev, int64_t n,
                                                    int32_t gmin, int32_t* __restrict__ kout,
                                                    int64_t* __restrict__ rout) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        if (kout) kout[i] = int32_t(int64_t(gmin) + int64_t(keys[i]));
        if (rout) rout[i] = int64_t(rev[slots[i]]);
    }
}

__global__ __launch_bounds__(kT) void word_pop_k(const uint32_t* __restrict__ bits, int64_t w,
                                                 int64_t* __restrict__ cnt) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < w;
         i += int64_t(gridDim.x) * kT)
        cnt[i] = __popc(bits[i]);
}

// Group i goes to its rank among the present keys: the exclusive prefix of
// the word popcounts plus the set bits below it in its word.
__global__ __launch_bounds__(kT) void rank_scatter_k(const uint64_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ slots,
                                                     const uint64_t* __restrict__ rev, int64_t n,
                                                     const uint32_t* __restrict__ bits,
                                                     const int64_t* __restrict__ prefix,
                                                     int32_t gmin, int32_t* __restrict__ kout,
                                                     int64_t* __restrict__ rout) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const uint64_t v = keys[i];
        const int64_t w = int64_t(v >> 5);
        const int64_t pos = prefix[w] + __popc(bits[w] & ((1u << (v & 31)) - 1u));
        if (kout) kout[pos] = int32_t(int64_t(gmin) + int64_t(v));
        if (rout) rout[pos] = int64_t(rev[slots[i]]);
    }
}

inline unsigned grid_for(bolap_ctx* ctx, int64_t n, int per_cu = 8) {
    int64_t b = (n + kT - 1) / kT;
    const int64_t cap = int64_t(ctx->num_cus) * per_cu;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

int ceil_log2(int64_t v) {
    int l = 0;
    while ((int64_t(1) << l) < v) ++l;
    return l;
}

int read_stats(bolap_ctx* ctx, int64_t* dst) {
    const size_t bytes = kStatWords * 8;
    BOLAP_TRY(pinned_reserve(ctx, bytes));
    BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, ctx->result_dev, bytes, hipMemcpyDeviceToHost,
                                 ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const int64_t* rep = static_cast<const int64_t*>(ctx->pinned);
    for (int i = 0; i < sN; ++i) {
        uint64_t v = 0;
        for (int r = 0; r < kRep; ++r) v += uint64_t(rep[stat_slot(i, r)]);
        dst[i] = int64_t(v);
    }
    return BOLAP_OK;
}

int validate(const bolap_star_args* a) {
    BOLAP_ARG_CHECK(a->n_dims >= 1 && a->n_dims <= kMaxDims, "bolap_star: %d dimensions (1..%d)",
                    a->n_dims, kMaxDims);
    BOLAP_ARG_CHECK(a->fact_rows >= 0, "bolap_star: fact_rows < 0");
    BOLAP_ARG_CHECK(a->group_min <= a->group_max, "bolap_star: group_min > group_max");
    BOLAP_ARG_CHECK(a->price_dtype == BOLAP_I64 || a->price_dtype == BOLAP_I32,
                    "bolap_star: price dtype %d", a->price_dtype);
    BOLAP_ARG_CHECK(a->disc_dtype == BOLAP_I64 || a->disc_dtype == BOLAP_U8,
                    "bolap_star: disc dtype %d", a->disc_dtype);
    BOLAP_ARG_CHECK(a->group_hint >= 0, "bolap_star: group_hint < 0");
    BOLAP_ARG_CHECK(int64_t(a->group_max) - a->group_min < int64_t(0xFFFFFFFF),
                    "bolap_star: the group key range must hold fewer than 2^32 - 1 keys");
    BOLAP_ARG_CHECK(a->fact_rows == 0 || (a->group_key && a->fact_price && a->fact_disc),
                    "bolap_star: null fact column");
    for (int j = 0; j < a->n_dims; ++j) {
        const bolap_star_dim& d = a->dims[j];
        BOLAP_ARG_CHECK(d.rows >= 0 && d.key_min <= d.key_max && d.pass_hint >= 0,
                        "bolap_star: dimension %d: bad rows, key range or hint", j);
        BOLAP_ARG_CHECK(int64_t(d.key_max) - d.key_min < (int64_t(1) << 31),
                        "bolap_star: dimension %d: key range above 2^31 keys", j);
        BOLAP_ARG_CHECK(d.rows == 0 || d.key, "bolap_star: dimension %d has no key column", j);
        BOLAP_ARG_CHECK(a->fact_rows == 0 || d.fact_fk,
                        "bolap_star: dimension %d has no fact key column", j);
    }
    return BOLAP_OK;
}

struct StarRun {
    int64_t st[sN];
    StarK k;
    int64_t cap = 0;
    int64_t n_groups = 0;
    int32_t group_place = kPlaceBits;
    int32_t count_passes = 0;
    size_t o_keys = 0, o_slots = 0, o_rev = 0;
    size_t o_bits = 0, o_cnt = 0;
    int64_t bit_words = 0;     // > 0: key-order output through the bitmap
    bool overflow = false;
};

constexpr int64_t kMaxBitmapSpan = int64_t(1) << 26;

template <int kMode>
void launch_probe(bolap_ctx* ctx, const bolap_star_args* a, const StarK& k, const GTab& g,
                  unsigned grid, size_t lds, int gl_log2, int64_t* st) {
    const int64_t gspan = int64_t(a->group_max) - a->group_min + 1;
    const bool p64 = a->price_dtype == BOLAP_I64, d64 = a->disc_dtype == BOLAP_I64;
    // More than 64 KB of dynamic LDS needs the per-kernel limit raised
    // (once per kernel and size: kernel_occupancy caches it).
#define BOLAP_STAR_LAUNCH(TP, TD)                                                            \
    do {                                                                                       \
        int per_cu_ = 0;                                                                       \
        (void)kernel_occupancy(ctx,                                                            \
                               reinterpret_cast<const void*>(star_probe_k<TP, TD, kMode>),     \
                               kT, lds, &per_cu_);                                             \
        hipLaunchKernelGGL((star_probe_k<TP, TD, kMode>), dim3(grid), dim3(kT), lds,          \
                           ctx->stream, k, g, a->group_key, a->group_min, gspan,              \
                           static_cast<const TP*>(a->fact_price),                             \
                           static_cast<const TD*>(a->fact_disc), a->fact_rows, gl_log2, st);  \
    } while (0)
    if (p64 && d64) BOLAP_STAR_LAUNCH(int64_t, int64_t);
    else if (p64) BOLAP_STAR_LAUNCH(int64_t, uint8_t);
    else if (d64) BOLAP_STAR_LAUNCH(int32_t, int64_t);
    else BOLAP_STAR_LAUNCH(int32_t, uint8_t);
#undef BOLAP_STAR_LAUNCH
}

// Probe grid: workgroups per CU limited by the LDS the kernel stages.
unsigned probe_grid(bolap_ctx* ctx, int64_t rows, size_t lds) {
    int per_cu = 8;
    if (lds > 0) {
        const int fit = int(ctx->lds_per_cu / lds);
        per_cu = std::max(1, std::min(per_cu, fit));
    }
    per_cu = int(ctx->knob("star.blocks_per_cu", per_cu));
    return grid_for(ctx, rows / 4 + 1, per_cu);
}

int star_run(bolap_ctx* ctx, const bolap_star_args* a, bool exact, bool want_rows, WsArena& ws,
             StarRun& run) {
    int64_t* st = static_cast<int64_t*>(ctx->result_dev);
    hipLaunchKernelGGL(st_init_k, dim3(1), dim3(kStatWords), 0, ctx->stream, st);
    BOLAP_TRY(launch_check("st_init_k"));
    std::memset(run.st, 0, sizeof(run.st));
    StarK& k = run.k;
    std::memset(&k, 0, sizeof(k));
    k.n = a->n_dims;
    for (int j = 0; j < a->n_dims; ++j) {
        const bolap_star_dim& d = a->dims[j];
        DimK& dk = k.d[j];
        dk.key = d.key;
        dk.fk = d.fact_fk;
        dk.rows = d.rows;
        dk.span = int64_t(d.key_max) - d.key_min + 1;
        dk.kmin = d.key_min;
        BOLAP_TRY(make_predset(dk.ps, d.n_preds, d.preds, d.rows, "bolap_star"));
    }

    // 1. Dimension sizes: exact counts where there is no estimate.
    int64_t est[kMaxDims] = {0, 0, 0};
    bool counted = false;
    for (int j = 0; j < a->n_dims; ++j) {
        est[j] = a->dims[j].pass_hint;
        if ((exact || est[j] == 0) && k.d[j].rows > 0) {
            TraceScope ts(ctx, "star_dim_count");
            hipLaunchKernelGGL(dim_count_k, dim3(grid_for(ctx, k.d[j].rows)), dim3(kT), 0,
                               ctx->stream, k.d[j], st, j);
            BOLAP_TRY(launch_check("dim_count_k"));
            counted = true;
        }
    }
    if (counted) {
        BOLAP_TRY(read_stats(ctx, run.st));
        ++run.count_passes;
        for (int j = 0; j < a->n_dims; ++j)
            if (exact || est[j] == 0) est[j] = run.st[sPass0 + j];
        hipLaunchKernelGGL(st_init_k, dim3(1), dim3(kStatWords), 0, ctx->stream, st);
    }
    size_t o_set[kMaxDims];
    int64_t words[kMaxDims];
    for (int j = 0; j < a->n_dims; ++j) {
        DimK& dk = k.d[j];
        dk.place = est[j] <= BOLAP_STAR_LDS_KEYS ? kPlaceSet : kPlaceBits;
        if (dk.place == kPlaceSet) {
            dk.lds_off = k.n_lds * kSetCap;
            ++k.n_lds;
            words[j] = kSetCap;
        } else {
            words[j] = (dk.span + 31) / 32;
        }
        o_set[j] = ws.plan(size_t(words[j]) * 4);
    }
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    for (int j = 0; j < a->n_dims; ++j) {
        DimK& dk = k.d[j];
        dk.set = ws.at<uint32_t>(o_set[j]);
        hipLaunchKernelGGL(fill32_k, dim3(grid_for(ctx, words[j])), dim3(kT), 0, ctx->stream,
                           dk.set, words[j]);
        if (dk.rows > 0) {
            TraceScope ts(ctx, "star_dim_build");
            hipLaunchKernelGGL(dim_build_k, dim3(grid_for(ctx, dk.rows)), dim3(kT), 0,
                               ctx->stream, dk, st, j);
        }
        BOLAP_TRY(launch_check("dim_build_k"));
    }

    // 2. Group table size: from the estimate, or from the exact join count.
    const int64_t gspan = int64_t(a->group_max) - a->group_min + 1;
    const size_t set_lds = size_t(k.n_lds) * kSetCap * 4;
    GTab g{nullptr, nullptr, 10, 0};
    int64_t est_g = a->group_hint;
    bool group_read = false;  // run.st holds the builds' counts and the join count
    if (exact || est_g == 0) {
        if (a->fact_rows > 0) {
            TraceScope ts(ctx, "star_count");
            launch_probe<0>(ctx, a, k, g, probe_grid(ctx, a->fact_rows, set_lds), set_lds, 0,
                            st);
            BOLAP_TRY(launch_check("star_probe_k(count)"));
        }
        BOLAP_TRY(read_stats(ctx, run.st));
        ++run.count_passes;
        group_read = true;
        est_g = std::min(run.st[sMatch], gspan);
    }
    est_g = std::max<int64_t>(est_g, 1);
    const int cap_log2 = std::max(10, ceil_log2(2 * est_g));
    BOLAP_ARG_CHECK(cap_log2 <= 31, "bolap_star: %lld groups exceed the table limit",
                    (long long)est_g);
    const int64_t cap = int64_t(1) << cap_log2;
    run.cap = cap;
    const bool gl = ctx->knob("star.group_lds", 1) == 1;
    const int gl_log2 = est_g <= kGlMaxGroups ? kGlLog2Small : kGlLog2Large;
    run.group_place = gl ? kPlaceSet : kPlaceBits;
    const size_t keep = ws.off;
    const size_t o_gk = ws.plan(size_t(cap) * 4);
    const size_t o_gr = ws.plan(size_t(cap) * 8);
    run.o_keys = ws.plan(size_t(cap) * 8);
    run.o_slots = ws.plan(size_t(cap) * 4);
    run.o_rev = o_gr;
    const size_t o_cnt = ws.plan(64);
    if (want_rows && gspan <= kMaxBitmapSpan) {
        run.bit_words = (gspan + 31) / 32;
        run.o_bits = ws.plan(size_t(run.bit_words) * 4);
        run.o_cnt = ws.plan(size_t(run.bit_words) * 8);
    }
    BOLAP_TRY(ws_reserve(ctx, ws.off, keep));
    ws.base = static_cast<char*>(ctx->ws);
    for (int j = 0; j < a->n_dims; ++j) k.d[j].set = ws.at<uint32_t>(o_set[j]);
    // Exact tables (cap >= 2 x an upper bound) always place every key;
    // estimate-sized ones give up after 128 slots and fall back.
    g = GTab{ws.at<uint32_t>(o_gk), ws.at<uint64_t>(o_gr), cap_log2,
             exact ? uint32_t(cap) : 128u};
    int64_t pre[sN];
    const bool had_read = group_read;
    if (had_read) {
        std::memcpy(pre, run.st, sizeof(pre));
        hipLaunchKernelGGL(st_init_k, dim3(1), dim3(kStatWords), 0, ctx->stream, st);
    }
    {
        TraceScope ts(ctx, "star_init");
        hipLaunchKernelGGL(fill32_k, dim3(grid_for(ctx, cap)), dim3(kT), 0, ctx->stream, g.keys,
                           cap);
        hipLaunchKernelGGL(fill64_k, dim3(grid_for(ctx, cap)), dim3(kT), 0, ctx->stream, g.rev,
                           cap);
        hipLaunchKernelGGL(fill32_k, dim3(1), dim3(kT), 0, ctx->stream, ws.at<uint32_t>(o_cnt),
                           int64_t(16));
        if (run.bit_words)
            hipLaunchKernelGGL(fill32_k, dim3(grid_for(ctx, run.bit_words)), dim3(kT), 0,
                               ctx->stream, ws.at<uint32_t>(run.o_bits), run.bit_words);
    }
    BOLAP_TRY(launch_check("star_init"));
    if (a->fact_rows > 0) {
        TraceScope ts(ctx, "star_probe");
        const size_t lds = set_lds + (gl ? (size_t(12) << gl_log2) : 0);
        const unsigned grid = probe_grid(ctx, a->fact_rows, lds);
        if (gl) launch_probe<2>(ctx, a, k, g, grid, lds, gl_log2, st);
        else launch_probe<1>(ctx, a, k, g, grid, lds, 0, st);
    }
    BOLAP_TRY(launch_check("star_probe_k"));
    auto* n_rows = ws.at<unsigned long long>(o_cnt);
    {
        TraceScope ts(ctx, "star_compact");
        int64_t cblocks = int64_t(ctx->num_cus);
        int64_t chunk = (cap + cblocks - 1) / cblocks;
        chunk = (chunk + kCT - 1) / kCT * kCT;
        cblocks = (cap + chunk - 1) / chunk;
        hipLaunchKernelGGL(star_compact_k, dim3(unsigned(cblocks)), dim3(kCT), 0, ctx->stream, g,
                           chunk, ws.at<uint64_t>(run.o_keys), ws.at<uint32_t>(run.o_slots),
                           n_rows, run.bit_words ? ws.at<uint32_t>(run.o_bits) : nullptr, st);
    }
    BOLAP_TRY(launch_check("star_compact_k"));
    BOLAP_TRY(read_stats(ctx, run.st));
    if (had_read) {
        for (int j = 0; j < kMaxDims; ++j) {
            run.st[sPass0 + j] += pre[sPass0 + j];
            run.st[sSet0 + j] += pre[sSet0 + j];
        }
        run.st[sOver] += pre[sOver];
        run.st[sRange] += pre[sRange];
    }
    BOLAP_ARG_CHECK(run.st[sRange] == 0,
                    "bolap_star: a key lies outside the [min, max] range passed in the args");
    run.n_groups = run.st[sGroups];  // each group is created by exactly one CAS
    bool over = run.st[sOver] != 0 || run.st[sGroups] > cap / 2;
    for (int j = 0; j < a->n_dims; ++j)
        if (k.d[j].place == kPlaceSet && run.st[sSet0 + j] > BOLAP_STAR_LDS_KEYS) over = true;
    run.overflow = over;
    return BOLAP_OK;
}

// Sorts the compacted groups by key and writes the caller's arrays.
int star_emit(bolap_ctx* ctx, const bolap_star_args* a, WsArena& ws, StarRun& run,
              int32_t* keys_out, int64_t* rev_out, int64_t max_rows) {
    const int64_t n = run.n_groups;
    BOLAP_ARG_CHECK(n <= max_rows, "bolap_star_groupby: %lld groups exceed max_rows %lld",
                    (long long)n, (long long)max_rows);
    if (n == 0) return BOLAP_OK;
    if (run.bit_words) {
        TraceScope ts(ctx, "star_order");
        const size_t keep = ws.off;
        const size_t o_scr = ws.plan(scan_scratch_bytes(run.bit_words));
        BOLAP_TRY(ws_reserve(ctx, ws.off, keep));
        ws.base = static_cast<char*>(ctx->ws);
        int64_t* cnt = ws.at<int64_t>(run.o_cnt);
        hipLaunchKernelGGL(word_pop_k, dim3(grid_for(ctx, run.bit_words)), dim3(kT), 0,
                           ctx->stream, ws.at<uint32_t>(run.o_bits), run.bit_words, cnt);
        BOLAP_TRY(exclusive_scan_i64(ctx, cnt, run.bit_words, ws.at<void>(o_scr)));
        hipLaunchKernelGGL(rank_scatter_k, dim3(grid_for(ctx, n)), dim3(kT), 0, ctx->stream,
                           ws.at<uint64_t>(run.o_keys), ws.at<uint32_t>(run.o_slots),
                           ws.at<uint64_t>(run.o_rev), n, ws.at<uint32_t>(run.o_bits), cnt,
                           a->group_min, keys_out, rev_out);
        return launch_check("rank_scatter_k");
    }
    const int64_t gspan = int64_t(a->group_max) - a->group_min + 1;
    const int bits = std::max(1, ceil_log2(gspan));
    const size_t keep = ws.off;
    const size_t o_scr = ws.plan(radix_sort_scratch_bytes(ctx, n));
    BOLAP_TRY(ws_reserve(ctx, ws.off, keep));
    ws.base = static_cast<char*>(ctx->ws);
    int passes = 0;
    {
        TraceScope ts(ctx, "star_sort");
        BOLAP_TRY(radix_sort_pairs(ctx, ws.at<uint64_t>(run.o_keys), ws.at<uint32_t>(run.o_slots),
                                   n, bits, ws.at<void>(o_scr), &passes));
    }
    hipLaunchKernelGGL(star_gather_k, dim3(grid_for(ctx, n)), dim3(kT), 0, ctx->stream,
                       ws.at<uint64_t>(run.o_keys), ws.at<uint32_t>(run.o_slots),
                       ws.at<uint64_t>(run.o_rev), n, a->group_min, keys_out, rev_out);
    return launch_check("star_gather_k");
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" int bolap_star_groupby(bolap_ctx* ctx, const bolap_star_args* a, int32_t* keys_out,
                                  int64_t* rev_out, int64_t max_rows, int64_t* n_out,
                                  bolap_star_stats* stats) {
    BOLAP_ARG_CHECK(ctx && a, "bolap_star_groupby: null argument");
    BOLAP_TRY(validate(a));
    BOLAP_ARG_CHECK(max_rows >= 0, "bolap_star_groupby: max_rows < 0");
    bool exact = a->group_hint == 0;
    for (int j = 0; j < a->n_dims; ++j) exact = exact && a->dims[j].pass_hint == 0;
    StarRun run;
    WsArena ws;
    const bool want_rows = keys_out || rev_out;
    BOLAP_TRY(star_run(ctx, a, exact, want_rows, ws, run));
    int32_t fallbacks = 0;
    if (run.overflow) {
        BOLAP_ARG_CHECK(!exact, "bolap_star_groupby: exact sizing overflowed");
        fallbacks = 1;
        const int32_t passes = run.count_passes;
        ws = WsArena();
        run = StarRun();
        BOLAP_TRY(star_run(ctx, a, true, want_rows, ws, run));
        run.count_passes += passes;
        BOLAP_ARG_CHECK(!run.overflow, "bolap_star_groupby: exact sizing overflowed");
    }
    if (want_rows) BOLAP_TRY(star_emit(ctx, a, ws, run, keys_out, rev_out, max_rows));
    if (n_out) *n_out = run.n_groups;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        for (int j = 0; j < a->n_dims; ++j) {
            stats->dim_pass[j] = run.st[sPass0 + j];
            stats->dim_placement[j] = run.k.d[j].place;
        }
        stats->group_placement = run.group_place;
        stats->matches = run.st[sMatch];
        stats->groups = run.n_groups;
        stats->group_capacity = run.cap;
        stats->revenue = run.st[sRev];
        stats->fallbacks = fallbacks;
        stats->count_passes = run.count_passes;
    }
    return BOLAP_OK;
}
