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
//   expand  a wave takes 64 join rows and spreads their concatenated value
//           ranges over its lanes; each expanded row reads its slot and
//           issues an atomic min / max only when its value lies outside
//           the (min, max) read.
//   dense   slots = group * keys + key: int32 min / max side by side, and a
//           used byte set only by rows without a value (a slot is used when
//           that byte is set or min <= max); a bitmap of used slots, a
//           popcount scan and a rank scatter give the rows in slot order.
//   hashed  open addressing on the 64-bit slot number group * keys + key,
//           16-byte slots {key, min, max} so that an expanded row touches one
//           cache line (CAS insertion, linear probing); occupied slots are
//           compacted (one atomic per wave) and LSD radix sorted on the
//           slot number's bits.
// A group has a value exactly when min <= max: both start at the empty
// extremes (INT32_MAX, INT32_MIN) and every folded value lands between them.
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

struct HSlot {  // hashed table slot
    uint64_t key;
    int32_t mn, mx;
};

__global__ __launch_bounds__(kT) void mm_init_dense_k(int2* __restrict__ mm,
                                                      uint8_t* __restrict__ used, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        mm[i] = make_int2(INT32_MAX, INT32_MIN);
        used[i] = 0;
    }
}

__global__ __launch_bounds__(kT) void fill_u64_k(unsigned long long* __restrict__ p, int64_t n,
                                                 unsigned long long v) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        p[i] = v;
}

__global__ __launch_bounds__(kT) void mm_init_hash_k(HSlot* __restrict__ t, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        t[i] = HSlot{kEmpty, INT32_MAX, INT32_MIN};
}

// Folds v into a slot's (min, max) after a plain read `cur` of the slot:
// the minimum only falls and the maximum only rises, so a value inside a
// (possibly stale) [cur.x, cur.y] is inside the current range too and needs
// no atomic.
BOLAP_DEV void fold_seen(int32_t* mn, int32_t* mx, int2 cur, int32_t v) {
    if (v < cur.x) atomicMin(mn, v);
    if (v > cur.y) atomicMax(mx, v);
}

// The item ranges of join rows p0 .. p0 + 63 expanded across one wave: lane
// L takes expanded rows L, L + 64, ... of the concatenated ranges and calls
// body(group, value, has, i, pos) for value index i, the j-th value of join
// row p (pos = p << 32 | j: the reference's loop order), so the lanes stay busy
// whatever the range lengths (a lane per join row idles on short ranges
// while one long range runs).  p0 is uniform across the wave; every lane
// reaches each shuffle.  Malformed join rows are skipped here (mm_count_k
// counts them).
template <typename F>
BOLAP_DEV void expand_wave(const Args& a, int64_t p0, F&& body) {
    const int lane = int(threadIdx.x & (kWave - 1));
    const int64_t p = p0 + lane;
    int32_t g = 0, v = 0;
    int64_t lo = 0, hi = 0;
    bool has = false, bad = false;
    const bool ok = p < a.pairs && pair_range(a, p, g, lo, hi, has, v, bad);
    const int64_t c = ok ? hi - lo : 0;
    int64_t inc = c;
    for (int d = 1; d < kWave; d <<= 1) {
        const int64_t t = __shfl_up(inc, d, kWave);
        if (lane >= d) inc += t;
    }
    const int64_t total = __shfl(inc, kWave - 1, kWave);
    const int64_t exc = inc - c;
    const int hv = has ? 1 : 0;
    for (int64_t e0 = 0; e0 < total; e0 += kWave) {
        const int64_t e = e0 + lane;
        // owner: the last lane whose range starts at or before e; it has a
        // non-empty range holding e (exclusive starts never decrease)
        int o = 0;
        for (int step = kWave / 2; step >= 1; step >>= 1) {
            const int cand = o + step;
            if (__shfl(exc, cand, kWave) <= e) o = cand;
        }
        const int64_t o_exc = __shfl(exc, o, kWave);
        const int64_t o_lo = __shfl(lo, o, kWave);
        const int32_t o_g = __shfl(g, o, kWave);
        const int32_t o_v = __shfl(v, o, kWave);
        const int o_has = __shfl(hv, o, kWave);
        const uint64_t pos = (uint64_t(p0 + o) << 32) | uint64_t(e - o_exc);
        if (e < total) body(o_g, o_v, o_has != 0, o_lo + (e - o_exc), pos);
    }
}

// Wave-uniform start of the wave's first batch of join rows, and the stride.
BOLAP_DEV int64_t wave_first() { return (int64_t(blockIdx.x) * kT + threadIdx.x) & ~int64_t(kWave - 1); }
BOLAP_DEV int64_t wave_stride() { return int64_t(gridDim.x) * kT; }

// With FS, first[s] also takes the smallest loop position that reaches slot
// s (the reference appends a dense slot to its output when first used,
// query_q10a.cpp:866-872), read first so that most rows skip the atomic.
template <bool FS>
__global__ __launch_bounds__(kT) void mm_dense_k(Args a, int2* __restrict__ mm,
                                                 uint8_t* __restrict__ used,
                                                 unsigned long long* __restrict__ first,
                                                 int64_t* __restrict__ st) {
    uint64_t errs = 0;
    for (int64_t p0 = wave_first(); p0 < a.pairs; p0 += wave_stride()) {
        expand_wave(a, p0, [&](int32_t g, int32_t v, bool has, int64_t i, uint64_t pos) {
            const int32_t k = a.rv[i];
            if (k < 0 || k >= a.keys) {
                ++errs;
                return;
            }
            const int64_t s = int64_t(g) * a.keys + k;
            if (FS && pos < first[s]) atomicMin(first + s, (unsigned long long)pos);
            // a slot with a value is marked by min <= max; only value-less
            // rows mark `used`
            if (!has)
                used[s] = 1;
            else
                fold_seen(&mm[s].x, &mm[s].y, mm[s], v);
        });
    }
    errs = wave_sum(errs);
    if ((threadIdx.x & (kWave - 1)) == 0 && errs)
        atomicAdd(reinterpret_cast<unsigned long long*>(st + mErrors), errs);
}

// Used bytes -> bitmap words and per-word popcounts; the used-slot total.
__global__ __launch_bounds__(kT) void mm_bits_k(const uint8_t* __restrict__ used,
                                                const int2* __restrict__ mm, int64_t slots,
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
            if (s < slots && (used[s] || mm[s].x <= mm[s].y)) b |= 1u << i;
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
                                                      const int2* __restrict__ mm,
                                                      bolap_minmax_row* __restrict__ out) {
    for (int64_t s = int64_t(blockIdx.x) * kT + threadIdx.x; s < slots;
         s += int64_t(gridDim.x) * kT) {
        const uint32_t b = bits[s >> 5];
        if (!((b >> (s & 31)) & 1u)) continue;
        const int64_t pos = cnt[s >> 5] + __popc(b & ((1u << (s & 31)) - 1u));
        const int2 v = mm[s];
        const bool h = v.x <= v.y;
        out[pos] = bolap_minmax_row{int32_t(s / keys), int32_t(s % keys), h ? 1 : 0,
                                    h ? v.x : 0, h ? v.y : 0};
    }
}

__global__ __launch_bounds__(kT) void mm_hash_k(Args a, HSlot* __restrict__ t, uint64_t mask,
                                                int64_t* __restrict__ st) {
    uint64_t errs = 0;
    for (int64_t p0 = wave_first(); p0 < a.pairs; p0 += wave_stride()) {
        expand_wave(a, p0, [&](int32_t g, int32_t v, bool has, int64_t i, uint64_t) {
            const int32_t k = a.rv[i];
            if (k < 0 || k >= a.keys) {
                ++errs;
                return;
            }
            const uint64_t key = uint64_t(g) * uint64_t(a.keys) + uint64_t(k);
            // the table holds at most half as many keys as slots: a free
            // slot is always reached within mask + 1 probes
            uint64_t s = mix64(key) & mask;
            int2 seen = make_int2(INT32_MAX, INT32_MIN);
            for (uint64_t step = 0; step <= mask; ++step) {
                const HSlot cur_slot = t[s];
                uint64_t cur = cur_slot.key;
                if (cur == kEmpty) {
                    cur = atomicCAS(reinterpret_cast<unsigned long long*>(&t[s].key),
                                    (unsigned long long)kEmpty, (unsigned long long)key);
                    if (cur == kEmpty) cur = key;
                } else if (cur == key) {
                    seen = make_int2(cur_slot.mn, cur_slot.mx);
                }
                if (cur == key) break;
                s = (s + 1) & mask;
            }
            if (has) fold_seen(&t[s].mn, &t[s].mx, seen, v);
        });
    }
    errs = wave_sum(errs);
    if ((threadIdx.x & (kWave - 1)) == 0 && errs)
        atomicAdd(reinterpret_cast<unsigned long long*>(st + mErrors), errs);
}

// Occupied slots -> (key, slot) pairs in two passes over per-workgroup
// chunks of the table: occupied counts per chunk, a device scan, then each
// workgroup writes its chunk's pairs from its offset (no shared counter).
__global__ __launch_bounds__(kT) void mm_occ_count_k(const HSlot* __restrict__ t, int64_t cap,
                                                     int64_t chunk, int64_t* __restrict__ cnt,
                                                     int64_t* __restrict__ st) {
    __shared__ uint64_t s_w[kT / kWave];
    const int64_t c0 = int64_t(blockIdx.x) * chunk;
    const int64_t c1 = c0 + chunk < cap ? c0 + chunk : cap;
    uint64_t n = 0;
    for (int64_t s = c0 + threadIdx.x; s < c1; s += kT) n += t[s].key != kEmpty ? 1u : 0u;
    n = wave_sum(n);
    if ((threadIdx.x & (kWave - 1)) == 0) s_w[threadIdx.x / kWave] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t tot = 0;
        for (int w = 0; w < kT / kWave; ++w) tot += s_w[w];
        cnt[blockIdx.x] = int64_t(tot);
        if (tot) atomicAdd(reinterpret_cast<unsigned long long*>(st + mGroups), tot);
    }
}

__global__ __launch_bounds__(kT) void mm_compact_k(const HSlot* __restrict__ t, int64_t cap,
                                                   int64_t chunk, const int64_t* __restrict__ off,
                                                   uint64_t* __restrict__ ck,
                                                   uint32_t* __restrict__ cs) {
    __shared__ uint32_t s_w[kT / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t c0 = int64_t(blockIdx.x) * chunk;
    const int64_t c1 = c0 + chunk < cap ? c0 + chunk : cap;
    uint64_t base = uint64_t(off[blockIdx.x]);
    // the trip count is the same for every thread of the workgroup
    for (int64_t b0 = c0; b0 < c1; b0 += kT) {
        const int64_t s = b0 + threadIdx.x;
        const uint64_t k = s < c1 ? t[s].key : kEmpty;
        const bool occ = k != kEmpty;
        const uint64_t bal = __ballot(occ);
        if (lane == 0) s_w[wave] = uint32_t(__popcll(bal));
        __syncthreads();
        uint32_t before = 0, tile = 0;
#pragma unroll
        for (int w = 0; w < kT / kWave; ++w) {
            before += w < wave ? s_w[w] : 0u;
            tile += s_w[w];
        }
        __syncthreads();  // s_w is rewritten by the next tile
        if (occ) {
            const uint64_t at = base + before + __popcll(bal & ((1ull << lane) - 1ull));
            ck[at] = k;
            cs[at] = uint32_t(s);
        }
        base += tile;
    }
}

__global__ __launch_bounds__(kT) void mm_hash_emit_k(const uint64_t* __restrict__ ck,
                                                     const uint32_t* __restrict__ cs, int64_t n,
                                                     int32_t keys, const HSlot* __restrict__ t,
                                                     bolap_minmax_row* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const HSlot v = t[cs[i]];
        const uint64_t k = ck[i];
        const bool h = v.mn <= v.mx;
        out[i] = bolap_minmax_row{int32_t(k / uint64_t(keys)), int32_t(k % uint64_t(keys)),
                                  h ? 1 : 0, h ? v.mn : 0, h ? v.mx : 0};
    }
}

// Sort keys of the dense rows: the first loop position of each row's slot.
__global__ __launch_bounds__(kT) void mm_first_keys_k(const bolap_minmax_row* __restrict__ rows,
                                                      int64_t n, int32_t keys,
                                                      const unsigned long long* __restrict__ first,
                                                      uint64_t* __restrict__ k,
                                                      uint32_t* __restrict__ v) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        k[i] = first[int64_t(rows[i].group) * keys + rows[i].key];
        v[i] = uint32_t(i);
    }
}

__global__ __launch_bounds__(kT) void mm_permute_k(const bolap_minmax_row* __restrict__ src,
                                                   const uint32_t* __restrict__ idx, int64_t n,
                                                   bolap_minmax_row* __restrict__ dst) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        dst[i] = src[idx[i]];
}

int read_st(bolap_ctx* ctx, int64_t* dst) {
    BOLAP_TRY(pinned_reserve(ctx, mN * 8));
    BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, ctx->result_dev, mN * 8, hipMemcpyDeviceToHost,
                                 ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::memcpy(dst, ctx->pinned, mN * 8);
    return BOLAP_OK;
}

// Workspace plan of one call (offsets from `base`), shared by the size query
// and the call so that they agree.
struct Plan {
    bool dense = false;
    int64_t slots = 0, words = 0, bound = 0, cap = 0, nb = 0, chunk = 0, rows_cap = 0;
    size_t o_mm = 0, o_used = 0, o_bits = 0, o_cnt = 0, o_scr = 0, o_first = 0, o_tmp = 0,
           o_k = 0, o_v = 0, o_sort = 0;
    size_t o_t = 0, o_ck = 0, o_cs = 0, o_occ = 0, o_oscr = 0;
    size_t end = 0;
};

Plan make_plan(bolap_ctx* ctx, const bolap_minmax_args* a, int64_t expanded, bool first_seen,
               size_t base) {
    Plan p;
    WsArena ws;
    ws.off = WsArena::align(base);
    p.slots = int64_t(a->groups) * a->keys;
    const int64_t dense_max = a->dense_max_slots > 0 ? a->dense_max_slots : kDenseDefault;
    p.dense = p.slots <= dense_max;
    if (p.dense) {
        p.words = (p.slots + 31) / 32;
        p.o_mm = ws.plan(size_t(p.slots) * 8);
        p.o_used = ws.plan(size_t(p.slots));
        p.o_bits = ws.plan(size_t(p.words) * 4);
        p.o_cnt = ws.plan(size_t(p.words) * 8);
        p.o_scr = ws.plan(scan_scratch_bytes(p.words));
        if (first_seen) {
            // rows: at most one per slot and per expanded row
            p.rows_cap = std::max<int64_t>(1, std::min(p.slots, expanded));
            p.o_first = ws.plan(size_t(p.slots) * 8);
            p.o_tmp = ws.plan(size_t(p.rows_cap) * sizeof(bolap_minmax_row));
            p.o_k = ws.plan(size_t(p.rows_cap) * 8);
            p.o_v = ws.plan(size_t(p.rows_cap) * 4);
            p.o_sort = ws.plan(radix_sort_scratch_bytes(ctx, p.rows_cap));
        }
    } else {
        // exact bound on the distinct keys: expanded rows (and slots)
        p.bound = std::max<int64_t>(1, std::min(expanded, p.slots));
        const int lg = std::max(10, log2_ceil(p.bound * 2));
        p.cap = int64_t(1) << std::min(lg, 40);
        p.o_t = ws.plan(size_t(p.cap) * sizeof(HSlot));
        p.o_ck = ws.plan(size_t(p.bound) * 8);
        p.o_cs = ws.plan(size_t(p.bound) * 4);
        p.nb = int64_t(grid_of(ctx, p.cap));
        p.chunk = (p.cap + p.nb - 1) / p.nb;
        p.o_occ = ws.plan(size_t(p.nb) * 8);
        p.o_oscr = ws.plan(scan_scratch_bytes(p.nb));
        p.o_sort = ws.plan(radix_sort_scratch_bytes(ctx, p.bound));
    }
    p.end = ws.off;
    return p;
}

int check_args(const bolap_minmax_args* a, int64_t max_rows) {
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
    return BOLAP_OK;
}

Args args_of(const bolap_minmax_args* a) {
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
    return k;
}

}  // namespace

int minmax_expanded(bolap_ctx* ctx, const bolap_minmax_args* a, int64_t* expanded) {
    BOLAP_TRY(check_args(a, 0));
    const Args k = args_of(a);
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
    *expanded = h[mExpanded];
    return BOLAP_OK;
}

size_t minmax_ws_end(bolap_ctx* ctx, const bolap_minmax_args* a, int64_t expanded,
                     bool first_seen, size_t base) {
    return make_plan(ctx, a, expanded, first_seen, base).end;
}

int minmax_run(bolap_ctx* ctx, const bolap_minmax_args* a, int64_t expanded, bool first_seen,
               size_t base, bolap_minmax_row* rows_out, int64_t max_rows, int64_t* n_out,
               bolap_minmax_stats* stats) {
    BOLAP_TRY(check_args(a, max_rows));
    const Args k = args_of(a);
    int64_t* st = static_cast<int64_t*>(ctx->result_dev);
    int64_t h[mN];
    const Plan p = make_plan(ctx, a, expanded, first_seen, base);
    BOLAP_ARG_CHECK(p.dense || p.cap <= (int64_t(1) << 32),
                    "bolap_group_minmax: %lld expanded rows need a table beyond 2^32 slots",
                    (long long)expanded);
    // the caller's arrays below `base` stay where they are: a workspace that
    // must grow keeps them, but only a caller that reserved p.end already may
    // pass pointers into the workspace
    BOLAP_TRY(ws_reserve(ctx, p.end, base));
    WsArena ws;
    ws.base = static_cast<char*>(ctx->ws);
    int64_t groups_n = 0;
    if (p.dense) {
        int2* mm = ws.at<int2>(p.o_mm);
        uint8_t* used = ws.at<uint8_t>(p.o_used);
        auto* first = first_seen ? ws.at<unsigned long long>(p.o_first) : nullptr;
        {
            TraceScope ts(ctx, "minmax_dense");
            hipLaunchKernelGGL(mm_init_dense_k, dim3(grid_of(ctx, p.slots)), dim3(kT), 0,
                               ctx->stream, mm, used, p.slots);
            if (first_seen) {
                hipLaunchKernelGGL(fill_u64_k, dim3(grid_of(ctx, p.slots)), dim3(kT), 0,
                                   ctx->stream, first, p.slots, ~0ull);
                hipLaunchKernelGGL(mm_dense_k<true>, dim3(grid_of(ctx, a->pairs)), dim3(kT), 0,
                                   ctx->stream, k, mm, used, first, st);
            } else {
                hipLaunchKernelGGL(mm_dense_k<false>, dim3(grid_of(ctx, a->pairs)), dim3(kT), 0,
                                   ctx->stream, k, mm, used, first, st);
            }
            hipLaunchKernelGGL(mm_bits_k, dim3(grid_of(ctx, p.words)), dim3(kT), 0, ctx->stream,
                               used, mm, p.slots, ws.at<uint32_t>(p.o_bits),
                               ws.at<int64_t>(p.o_cnt), st);
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
            BOLAP_TRY(exclusive_scan_i64(ctx, ws.at<int64_t>(p.o_cnt), p.words,
                                         ws.at<void>(p.o_scr)));
            bolap_minmax_row* dst = first_seen ? ws.at<bolap_minmax_row>(p.o_tmp) : rows_out;
            hipLaunchKernelGGL(mm_dense_emit_k, dim3(grid_of(ctx, p.slots)), dim3(kT), 0,
                               ctx->stream, ws.at<uint32_t>(p.o_bits), ws.at<int64_t>(p.o_cnt),
                               p.slots, a->keys, mm, dst);
            BOLAP_TRY(launch_check("mm_dense_emit_k"));
            if (first_seen) {
                // rows in the order their slots were first reached
                uint64_t* fk = ws.at<uint64_t>(p.o_k);
                uint32_t* fv = ws.at<uint32_t>(p.o_v);
                hipLaunchKernelGGL(mm_first_keys_k, dim3(grid_of(ctx, groups_n)), dim3(kT), 0,
                                   ctx->stream, dst, groups_n, a->keys, first, fk, fv);
                int passes = 0;
                BOLAP_TRY(radix_sort_pairs(ctx, fk, fv, groups_n,
                                           32 + std::max(1, log2_ceil(a->pairs)),
                                           ws.at<void>(p.o_sort), &passes));
                hipLaunchKernelGGL(mm_permute_k, dim3(grid_of(ctx, groups_n)), dim3(kT), 0,
                                   ctx->stream, dst, fv, groups_n, rows_out);
                BOLAP_TRY(launch_check("mm_permute_k"));
            }
        }
    } else {
        HSlot* t = ws.at<HSlot>(p.o_t);
        {
            TraceScope ts(ctx, "minmax_hash");
            hipLaunchKernelGGL(mm_init_hash_k, dim3(grid_of(ctx, p.cap)), dim3(kT), 0,
                               ctx->stream, t, p.cap);
            hipLaunchKernelGGL(mm_hash_k, dim3(grid_of(ctx, a->pairs)), dim3(kT), 0, ctx->stream,
                               k, t, uint64_t(p.cap - 1), st);
            int64_t* occ = ws.at<int64_t>(p.o_occ);
            hipLaunchKernelGGL(mm_occ_count_k, dim3(unsigned(p.nb)), dim3(kT), 0, ctx->stream, t,
                               p.cap, p.chunk, occ, st);
            BOLAP_TRY(exclusive_scan_i64(ctx, occ, p.nb, ws.at<void>(p.o_oscr)));
            hipLaunchKernelGGL(mm_compact_k, dim3(unsigned(p.nb)), dim3(kT), 0, ctx->stream, t,
                               p.cap, p.chunk, occ, ws.at<uint64_t>(p.o_ck),
                               ws.at<uint32_t>(p.o_cs));
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
            BOLAP_TRY(radix_sort_pairs(ctx, ws.at<uint64_t>(p.o_ck), ws.at<uint32_t>(p.o_cs),
                                       groups_n, std::max(1, log2_ceil(p.slots)),
                                       ws.at<void>(p.o_sort), &passes));
            hipLaunchKernelGGL(mm_hash_emit_k, dim3(grid_of(ctx, groups_n)), dim3(kT), 0,
                               ctx->stream, ws.at<uint64_t>(p.o_ck), ws.at<uint32_t>(p.o_cs),
                               groups_n, a->keys, t, rows_out);
            BOLAP_TRY(launch_check("mm_hash_emit_k"));
        }
    }
    if (rows_out) BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n_out) *n_out = groups_n;
    if (stats) {
        stats->expanded_rows = expanded;
        stats->groups = groups_n;
        stats->placement = p.dense ? BOLAP_MINMAX_DENSE : BOLAP_MINMAX_HASH;
        stats->slots = p.dense ? p.slots : 0;
    }
    return BOLAP_OK;
}

}  // namespace bolap

extern "C" int bolap_group_minmax(bolap_ctx* ctx, const bolap_minmax_args* a,
                                  bolap_minmax_row* rows_out, int64_t max_rows, int64_t* n_out,
                                  bolap_minmax_stats* stats) {
    BOLAP_ARG_CHECK(ctx && a, "bolap_group_minmax: null argument");
    bolap::TraceScope ts_all(ctx, "minmax");
    int64_t expanded = 0;
    BOLAP_TRY(bolap::minmax_expanded(ctx, a, &expanded));
    return bolap::minmax_run(ctx, a, expanded, false, 0, rows_out, max_rows, n_out, stats);
}
