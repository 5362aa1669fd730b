// group_minmax.hip — grouped MIN/MAX of a nullable value over a join whose
// rows expand through per-item value ranges.
//
// Reference: bespoke_ceb/query_q10a.cpp:838-1107.  Every (person, movie)
// join row whose person has a name group expands into the movie's info
// values (a CSR range, :1075-1092); each (name group, info) pair is a group
// holding {has_year, min_year, max_year}.  A group first reached through a
// movie without a year is created as {false, 0, 0} and takes the first year
// that arrives later (:880-900).  The reference keeps a dense matrix when
// groups x infos <= 5,000,000 slots and an unordered_map otherwise (:843-856)
// and emits rows in first-seen / hash order (:1094-1120); here rows come out
// in (group, key) order, the canonical form of that unordered set.
//
// Device plan:
//   count   one lane per join row: validates its group, item and range and
//           sums the expanded rows (sizes the hashed table exactly: it can
//           never hold more groups than expanded rows).
//   dense   slots = group * keys + key: int32 min / max under atomics, a
//           used byte and a has byte; a bitmap of used slots, a popcount
//           scan and a rank scatter give the rows in slot order.
//   hashed  open addressing on the 64-bit key (group << 32 | key) with CAS
//           insertion and linear probing, the same per-slot fields; occupied
//           slots are compacted (one atomic per wave) and LSD radix sorted by
//           key.
#include <algorithm>
#include <cstring>

#include "common.hpp"
#include "prims.hpp"
#include "radix.hpp"

namespace bolap {
namespace {

constexpr int kT = 256;
constexpr uint64_t kEmpty = ~0ull;
constexpr int64_t kDenseDefault = 5000000;  // query_q10a.cpp:843

enum { mExpanded = 0, mErrors = 1, mGroups = 2, mN = 4 };

unsigned grid_of(bolap_ctx* ctx, int64_t n) {
    int64_t b = (n + kT - 1) / kT;
    const int64_t cap = int64_t(ctx->num_cus) * 8;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

int log2_ceil(int64_t v) {
    int b = 0;
    while ((int64_t(1) << b) < v) ++b;
    return b;
}

struct Args {
    int64_t pairs, items, values;
    const int32_t *pg, *pi, *off, *cnt, *val, *rv;
    const uint8_t* has;
    int32_t groups, keys;
};

__global__ __launch_bounds__(kT) void mm_zero_k(int64_t* __restrict__ st) {
    if (threadIdx.x < mN) st[threadIdx.x] = 0;
}

// The join row's item range, or false when the row is dropped (group < 0)
// or malformed (counted as an error).
BOLAP_DEV bool pair_range(const Args& a, int64_t p, int32_t& g, int64_t& lo, int64_t& hi,
                          bool& has, int32_t& v, bool& bad) {
    g = a.pg[p];
    bad = false;
    if (g < 0) return false;
    const int32_t it = a.pi[p];
    if (g >= a.groups || it < 0 || int64_t(it) >= a.items) {
        bad = true;
        return false;
    }
    lo = a.off[it];
    const int64_t c = a.cnt[it];
    hi = lo + c;
    if (lo < 0 || c < 0 || hi > a.values) {
        bad = true;
        return false;
    }
    has = a.has ? a.has[it] != 0 : true;
    v = a.val[it];
    return true;
}

__global__ __launch_bounds__(kT) void mm_count_k(Args a, int64_t* __restrict__ st) {
    uint64_t rows = 0, errs = 0;
    for (int64_t p = int64_t(blockIdx.x) * kT + threadIdx.x; p < a.pairs;
         p += int64_t(gridDim.x) * kT) {
        int32_t g, v;
        int64_t lo, hi;
        bool has, bad;
        if (pair_range(a, p, g, lo, hi, has, v, bad)) rows += uint64_t(hi - lo);
        errs += bad ? 1u : 0u;
    }
    // every lane of the wave reaches the reductions (the loop above is
    // per lane, the sums are taken after it)
    rows = wave_sum(rows);
    errs = wave_sum(errs);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (rows) atomicAdd(reinterpret_cast<unsigned long long*>(st + mExpanded), rows);
        if (errs) atomicAdd(reinterpret_cast<unsigned long long*>(st + mErrors), errs);
    }
}

__global__ __launch_bounds__(kT) void mm_init_k(int32_t* __restrict__ mn, int32_t* __restrict__ mx,
                                                uint8_t* __restrict__ used,
                                                uint8_t* __restrict__ has,
                                                uint64_t* __restrict__ keys, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        mn[i] = INT32_MAX;
        mx[i] = INT32_MIN;
        has[i] = 0;
        if (used) used[i] = 0;
        if (keys) keys[i] = kEmpty;
    }
}

BOLAP_DEV void fold(int32_t* mn, int32_t* mx, uint8_t* hs, int64_t s, bool has, int32_t v) {
    if (!has) return;
    hs[s] = 1;
    atomicMin(mn + s, v);
    atomicMax(mx + s, v);
}

__global__ __launch_bounds__(kT) void mm_dense_k(Args a, int32_t* __restrict__ mn,
                                                 int32_t* __restrict__ mx,
                                                 uint8_t* __restrict__ used,
                                                 uint8_t* __restrict__ hs,
                                                 int64_t* __restrict__ st) {
    uint64_t errs = 0;
    for (int64_t p = int64_t(blockIdx.x) * kT + threadIdx.x; p < a.pairs;
         p += int64_t(gridDim.x) * kT) {
        int32_t g, v;
        int64_t lo, hi;
        bool has, bad;
        if (!pair_range(a, p, g, lo, hi, has, v, bad)) continue;
        const int64_t base = int64_t(g) * a.keys;
        for (int64_t i = lo; i < hi; ++i) {
            const int32_t k = a.rv[i];
            if (k < 0 || k >= a.keys) {
                ++errs;
                continue;
            }
            const int64_t s = base + k;
            used[s] = 1;
            fold(mn, mx, hs, s, has, v);
        }
    }
    errs = wave_sum(errs);
    if ((threadIdx.x & (kWave - 1)) == 0 && errs)
        atomicAdd(reinterpret_cast<unsigned long long*>(st + mErrors), errs);
}

// Used bytes -> bitmap words and per-word popcounts; the used-slot total.
__global__ __launch_bounds__(kT) void mm_bits_k(const uint8_t* __restrict__ used, int64_t slots,
                                                uint32_t* __restrict__ bits,
                                                int64_t* __restrict__ cnt,
                                                int64_t* __restrict__ st) {
    const int64_t words = (slots + 31) / 32;
    uint64_t tot = 0;
    for (int64_t w = int64_t(blockIdx.x) * kT + threadIdx.x; w < words;
         w += int64_t(gridDim.x) * kT) {
        uint32_t b = 0;
        for (int i = 0; i < 32; ++i) {
            const int64_t s = w * 32 + i;
            if (s < slots && used[s]) b |= 1u << i;
        }
        bits[w] = b;
        cnt[w] = __popc(b);
        tot += __popc(b);
    }
    tot = wave_sum(tot);
    if ((threadIdx.x & (kWave - 1)) == 0 && tot)
        atomicAdd(reinterpret_cast<unsigned long long*>(st + mGroups), tot);
}

__global__ __launch_bounds__(kT) void mm_dense_emit_k(const uint32_t* __restrict__ bits,
                                                      const int64_t* __restrict__ cnt,
                                                      int64_t slots, int32_t keys,
                                                      const int32_t* __restrict__ mn,
                                                      const int32_t* __restrict__ mx,
                                                      const uint8_t* __restrict__ hs,
                                                      bolap_minmax_row* __restrict__ out) {
    for (int64_t s = int64_t(blockIdx.x) * kT + threadIdx.x; s < slots;
         s += int64_t(gridDim.x) * kT) {
        const uint32_t b = bits[s >> 5];
        if (!((b >> (s & 31)) & 1u)) continue;
        const int64_t pos = cnt[s >> 5] + __popc(b & ((1u << (s & 31)) - 1u));
        const bool h = hs[s] != 0;
        out[pos] = bolap_minmax_row{int32_t(s / keys), int32_t(s % keys), h ? 1 : 0,
                                    h ? mn[s] : 0, h ? mx[s] : 0};
    }
}

__global__ __launch_bounds__(kT) void mm_hash_k(Args a, uint64_t* __restrict__ keys,
                                                uint64_t mask, int32_t* __restrict__ mn,
                                                int32_t* __restrict__ mx,
                                                uint8_t* __restrict__ hs,
                                                int64_t* __restrict__ st) {
    uint64_t errs = 0;
    for (int64_t p = int64_t(blockIdx.x) * kT + threadIdx.x; p < a.pairs;
         p += int64_t(gridDim.x) * kT) {
        int32_t g, v;
        int64_t lo, hi;
        bool has, bad;
        if (!pair_range(a, p, g, lo, hi, has, v, bad)) continue;
        for (int64_t i = lo; i < hi; ++i) {
            const int32_t k = a.rv[i];
            if (k < 0 || k >= a.keys) {
                ++errs;
                continue;
            }
            const uint64_t key = (uint64_t(uint32_t(g)) << 32) | uint32_t(k);
            // the table holds at most half as many keys as slots: a free
            // slot is always reached within mask + 1 probes
            uint64_t s = mix64(key) & mask;
            for (uint64_t step = 0; step <= mask; ++step) {
                uint64_t cur = keys[s];
                if (cur == kEmpty) {
                    cur = atomicCAS(reinterpret_cast<unsigned long long*>(keys + s),
                                    (unsigned long long)kEmpty, (unsigned long long)key);
                    if (cur == kEmpty) cur = key;
                }
                if (cur == key) break;
                s = (s + 1) & mask;
            }
            fold(mn, mx, hs, int64_t(s), has, v);
        }
    }
    errs = wave_sum(errs);
    if ((threadIdx.x & (kWave - 1)) == 0 && errs)
        atomicAdd(reinterpret_cast<unsigned long long*>(st + mErrors), errs);
}

// Occupied slots -> (key, slot) pairs, one atomic per wave.
__global__ __launch_bounds__(kT) void mm_compact_k(const uint64_t* __restrict__ keys, int64_t cap,
                                                   uint64_t* __restrict__ ck,
                                                   uint32_t* __restrict__ cs,
                                                   int64_t* __restrict__ st) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t stride = int64_t(gridDim.x) * kT;
    // trip count uniform across the wave: the ballot needs every lane
    for (int64_t b0 = int64_t(blockIdx.x) * kT + (threadIdx.x & ~(kWave - 1)); b0 < cap;
         b0 += stride) {
        const int64_t s = b0 + lane;
        const uint64_t k = s < cap ? keys[s] : kEmpty;
        const bool occ = k != kEmpty;
        const uint64_t bal = __ballot(occ);
        if (!bal) continue;
        unsigned long long base = 0;
        if (lane == 0)
            base = atomicAdd(reinterpret_cast<unsigned long long*>(st + mGroups),
                             (unsigned long long)__popcll(bal));
        base = readlane_u64(base, 0);
        if (occ) {
            const uint64_t at = base + __popcll(bal & ((1ull << lane) - 1ull));
            ck[at] = k;
            cs[at] = uint32_t(s);
        }
    }
}

__global__ __launch_bounds__(kT) void mm_hash_emit_k(const uint64_t* __restrict__ ck,
                                                     const uint32_t* __restrict__ cs, int64_t n,
                                                     const int32_t* __restrict__ mn,
                                                     const int32_t* __restrict__ mx,
                                                     const uint8_t* __restrict__ hs,
                                                     bolap_minmax_row* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const uint32_t s = cs[i];
        const uint64_t k = ck[i];
        const bool h = hs[s] != 0;
        out[i] = bolap_minmax_row{int32_t(k >> 32), int32_t(uint32_t(k)), h ? 1 : 0,
                                  h ? mn[s] : 0, h ? mx[s] : 0};
    }
}

int read_st(bolap_ctx* ctx, int64_t* dst) {
    BOLAP_TRY(pinned_reserve(ctx, mN * 8));
    BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, ctx->result_dev, mN * 8, hipMemcpyDeviceToHost,
                                 ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::memcpy(dst, ctx->pinned, mN * 8);
    return BOLAP_OK;
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" int bolap_group_minmax(bolap_ctx* ctx, const bolap_minmax_args* a,
                                  bolap_minmax_row* rows_out, int64_t max_rows, int64_t* n_out,
                                  bolap_minmax_stats* stats) {
    BOLAP_ARG_CHECK(ctx && a, "bolap_group_minmax: null argument");
    BOLAP_ARG_CHECK(a->pairs >= 0 && a->items >= 0 && a->values >= 0,
                    "bolap_group_minmax: negative size");
    BOLAP_ARG_CHECK(a->groups >= 1 && a->keys >= 1, "bolap_group_minmax: groups %d, keys %d",
                    a->groups, a->keys);
    BOLAP_ARG_CHECK(a->pairs == 0 || (a->pair_group && a->pair_item),
                    "bolap_group_minmax: null join-row column");
    BOLAP_ARG_CHECK(a->pairs == 0 || (a->item_offset && a->item_count && a->item_value),
                    "bolap_group_minmax: null item column");
    BOLAP_ARG_CHECK(a->values == 0 || a->range_value, "bolap_group_minmax: null range values");
    BOLAP_ARG_CHECK(max_rows >= 0, "bolap_group_minmax: max_rows < 0");
    TraceScope ts_all(ctx, "minmax");
    Args k;
    k.pairs = a->pairs;
    k.items = a->items;
    k.values = a->values;
    k.pg = a->pair_group;
    k.pi = a->pair_item;
    k.off = a->item_offset;
    k.cnt = a->item_count;
    k.val = a->item_value;
    k.rv = a->range_value;
    k.has = a->item_has;
    k.groups = a->groups;
    k.keys = a->keys;
    int64_t* st = static_cast<int64_t*>(ctx->result_dev);
    int64_t h[mN];
    hipLaunchKernelGGL(mm_zero_k, dim3(1), dim3(kT), 0, ctx->stream, st);
    {
        TraceScope ts(ctx, "minmax_count");
        hipLaunchKernelGGL(mm_count_k, dim3(grid_of(ctx, a->pairs)), dim3(kT), 0, ctx->stream, k,
                           st);
        BOLAP_TRY(launch_check("mm_count_k"));
    }
    BOLAP_TRY(read_st(ctx, h));
    BOLAP_ARG_CHECK(h[mErrors] == 0,
                    "bolap_group_minmax: %lld join rows with a group, item or range out of bounds",
                    (long long)h[mErrors]);
    const int64_t expanded = h[mExpanded];
    const int64_t slots = int64_t(a->groups) * a->keys;
    const int64_t dense_max = a->dense_max_slots > 0 ? a->dense_max_slots : kDenseDefault;
    const bool dense = slots <= dense_max;
    int64_t groups_n = 0;
    WsArena ws;
    if (dense) {
        const int64_t words = (slots + 31) / 32;
        const size_t o_mn = ws.plan(size_t(slots) * 4), o_mx = ws.plan(size_t(slots) * 4);
        const size_t o_used = ws.plan(size_t(slots)), o_has = ws.plan(size_t(slots));
        const size_t o_bits = ws.plan(size_t(words) * 4), o_cnt = ws.plan(size_t(words) * 8);
        const size_t o_scr = ws.plan(scan_scratch_bytes(words));
        BOLAP_TRY(ws_reserve(ctx, ws.off));
        ws.base = static_cast<char*>(ctx->ws);
        int32_t* mn = ws.at<int32_t>(o_mn);
        int32_t* mx = ws.at<int32_t>(o_mx);
        uint8_t* used = ws.at<uint8_t>(o_used);
        uint8_t* hs = ws.at<uint8_t>(o_has);
        {
            TraceScope ts(ctx, "minmax_dense");
            hipLaunchKernelGGL(mm_init_k, dim3(grid_of(ctx, slots)), dim3(kT), 0, ctx->stream, mn,
                               mx, used, hs, static_cast<uint64_t*>(nullptr), slots);
            hipLaunchKernelGGL(mm_dense_k, dim3(grid_of(ctx, a->pairs)), dim3(kT), 0, ctx->stream,
                               k, mn, mx, used, hs, st);
            hipLaunchKernelGGL(mm_bits_k, dim3(grid_of(ctx, words)), dim3(kT), 0, ctx->stream,
                               used, slots, ws.at<uint32_t>(o_bits), ws.at<int64_t>(o_cnt), st);
            BOLAP_TRY(launch_check("mm_dense_k"));
        }
        BOLAP_TRY(read_st(ctx, h));
        BOLAP_ARG_CHECK(h[mErrors] == 0,
                        "bolap_group_minmax: %lld range values outside [0, keys)",
                        (long long)h[mErrors]);
        groups_n = h[mGroups];
        if (rows_out && groups_n > 0) {
            BOLAP_ARG_CHECK(groups_n <= max_rows,
                            "bolap_group_minmax: %lld groups exceed max_rows %lld",
                            (long long)groups_n, (long long)max_rows);
            TraceScope ts(ctx, "minmax_emit");
            BOLAP_TRY(exclusive_scan_i64(ctx, ws.at<int64_t>(o_cnt), words, ws.at<void>(o_scr)));
            hipLaunchKernelGGL(mm_dense_emit_k, dim3(grid_of(ctx, slots)), dim3(kT), 0,
                               ctx->stream, ws.at<uint32_t>(o_bits), ws.at<int64_t>(o_cnt), slots,
                               a->keys, mn, mx, hs, rows_out);
            BOLAP_TRY(launch_check("mm_dense_emit_k"));
        }
    } else {
        // exact bound on the distinct keys: expanded rows (and slots)
        const int64_t bound = std::max<int64_t>(1, std::min(expanded, slots));
        const int lg = std::max(10, log2_ceil(bound * 2));
        BOLAP_ARG_CHECK(lg <= 32, "bolap_group_minmax: %lld expanded rows need a table beyond "
                        "2^32 slots", (long long)expanded);
        const int64_t cap = int64_t(1) << lg;
        const size_t o_keys = ws.plan(size_t(cap) * 8), o_mn = ws.plan(size_t(cap) * 4);
        const size_t o_mx = ws.plan(size_t(cap) * 4), o_has = ws.plan(size_t(cap));
        const size_t o_ck = ws.plan(size_t(bound) * 8), o_cs = ws.plan(size_t(bound) * 4);
        const size_t o_scr = ws.plan(radix_sort_scratch_bytes(ctx, bound));
        BOLAP_TRY(ws_reserve(ctx, ws.off));
        ws.base = static_cast<char*>(ctx->ws);
        uint64_t* keys = ws.at<uint64_t>(o_keys);
        int32_t* mn = ws.at<int32_t>(o_mn);
        int32_t* mx = ws.at<int32_t>(o_mx);
        uint8_t* hs = ws.at<uint8_t>(o_has);
        {
            TraceScope ts(ctx, "minmax_hash");
            hipLaunchKernelGGL(mm_init_k, dim3(grid_of(ctx, cap)), dim3(kT), 0, ctx->stream, mn,
                               mx, static_cast<uint8_t*>(nullptr), hs, keys, cap);
            hipLaunchKernelGGL(mm_hash_k, dim3(grid_of(ctx, a->pairs)), dim3(kT), 0, ctx->stream,
                               k, keys, uint64_t(cap - 1), mn, mx, hs, st);
            hipLaunchKernelGGL(mm_compact_k, dim3(grid_of(ctx, cap)), dim3(kT), 0, ctx->stream,
                               keys, cap, ws.at<uint64_t>(o_ck), ws.at<uint32_t>(o_cs), st);
            BOLAP_TRY(launch_check("mm_hash_k"));
        }
        BOLAP_TRY(read_st(ctx, h));
        BOLAP_ARG_CHECK(h[mErrors] == 0,
                        "bolap_group_minmax: %lld range values outside [0, keys)",
                        (long long)h[mErrors]);
        groups_n = h[mGroups];
        if (rows_out && groups_n > 0) {
            BOLAP_ARG_CHECK(groups_n <= max_rows,
                            "bolap_group_minmax: %lld groups exceed max_rows %lld",
                            (long long)groups_n, (long long)max_rows);
            TraceScope ts(ctx, "minmax_emit");
            int passes = 0;
            BOLAP_TRY(radix_sort_pairs(ctx, ws.at<uint64_t>(o_ck), ws.at<uint32_t>(o_cs),
                                       groups_n, 32 + std::max(1, log2_ceil(a->groups)),
                                       ws.at<void>(o_scr), &passes));
            hipLaunchKernelGGL(mm_hash_emit_k, dim3(grid_of(ctx, groups_n)), dim3(kT), 0,
                               ctx->stream, ws.at<uint64_t>(o_ck), ws.at<uint32_t>(o_cs),
                               groups_n, mn, mx, hs, rows_out);
            BOLAP_TRY(launch_check("mm_hash_emit_k"));
        }
    }
    if (rows_out) BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n_out) *n_out = groups_n;
    if (stats) {
        stats->expanded_rows = expanded;
        stats->groups = groups_n;
        stats->placement = dense ? BOLAP_MINMAX_DENSE : BOLAP_MINMAX_HASH;
        stats->slots = dense ? slots : 0;
    }
    return BOLAP_OK;
}
