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
//   sorted  beyond the dense limit nearly every expanded row is a group of
//           its own (q10a: 14.1 M groups from 14.4 M expanded rows), so a
//           table would only scatter them: the expanded rows are written as
//           (slot << 1 | has, value) pairs (one cursor reservation per
//           wave), radix sorted on the slot bits, and each wave reduces the
//           runs that start in its chunk of the sorted rows (a segmented
//           scan across the wave, the open run carried from tile to tile and
//           past the chunk's end); run heads counted per wave and scanned
//           give each run its output index, so rows come out in slot order.
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
    // optional: the item columns packed per item as {offset, count, value,
    // has}, so that a join row's item costs one 16-byte gather, not four
    const int4* pk;
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
    int64_t c;
    if (a.pk) {
        const int4 q = a.pk[it];
        lo = q.x;
        c = q.y;
        v = q.z;
        has = q.w != 0;
    } else {
        lo = a.off[it];
        c = a.cnt[it];
        has = a.has ? a.has[it] != 0 : true;
        v = a.val[it];
    }
    hi = lo + c;
    if (lo < 0 || c < 0 || hi > a.values) {
        bad = true;
        return false;
    }
    return true;
}

__global__ __launch_bounds__(kT) void mm_pack_k(Args a, int4* __restrict__ pk) {
    for (int64_t it = int64_t(blockIdx.x) * kT + threadIdx.x; it < a.items;
         it += int64_t(gridDim.x) * kT)
        pk[it] = make_int4(a.off[it], a.cnt[it], a.val[it], a.has ? (a.has[it] != 0) : 1);
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
// body(group, value, has, i, pos, out) for value index i, the j-th value of
// join row p (pos = p << 32 | j: the reference's loop order), so the lanes
// stay busy whatever the range lengths (a lane per join row idles on short
// ranges while one long range runs).  reserve(total) is called by every
// lane with the wave's total of expanded rows and returns the (uniform)
// output index of its first one; out = that index + the row's rank in the
// batch.  p0 is uniform across the wave; every lane reaches each shuffle.
// Malformed join rows are skipped here (mm_count_k counts them).
struct WaveRanges {
    int64_t lo = 0, exc = 0, total = 0;  // total is uniform across the wave
    int32_t g = 0, v = 0;
    int hv = 0;
};

// Lane-local range of join row p0 + lane and the wave's prefix sums.
BOLAP_DEV WaveRanges wave_ranges(const Args& a, int64_t p0) {
    const int lane = int(threadIdx.x & (kWave - 1));
    const int64_t p = p0 + lane;
    WaveRanges w;
    int64_t hi = 0;
    bool has = false, bad = false;
    const bool ok = p < a.pairs && pair_range(a, p, w.g, w.lo, hi, has, w.v, bad);
    const int64_t c = ok ? hi - w.lo : 0;
    int64_t inc = c;
    for (int d = 1; d < kWave; d <<= 1) {
        const int64_t t = __shfl_up(inc, d, kWave);
        if (lane >= d) inc += t;
    }
    w.total = __shfl(inc, kWave - 1, kWave);
    w.exc = inc - c;
    w.hv = has ? 1 : 0;
    return w;
}

// The expanded rows of the wave's batch: lane L takes rows L, L + 64, ...
template <typename F>
BOLAP_DEV void wave_emit(const WaveRanges& w, int64_t p0, int64_t base, F&& body) {
    const int lane = int(threadIdx.x & (kWave - 1));
    for (int64_t e0 = 0; e0 < w.total; e0 += kWave) {
        const int64_t e = e0 + lane;
        // owner: the last lane whose range starts at or before e; it has a
        // non-empty range holding e (exclusive starts never decrease)
        int o = 0;
        for (int step = kWave / 2; step >= 1; step >>= 1) {
            const int cand = o + step;
            if (__shfl(w.exc, cand, kWave) <= e) o = cand;
        }
        const int64_t o_exc = __shfl(w.exc, o, kWave);
        const int64_t o_lo = __shfl(w.lo, o, kWave);
        const int32_t o_g = __shfl(w.g, o, kWave);
        const int32_t o_v = __shfl(w.v, o, kWave);
        const int o_has = __shfl(w.hv, o, kWave);
        const uint64_t pos = (uint64_t(p0 + o) << 32) | uint64_t(e - o_exc);
        if (e < w.total) body(o_g, o_v, o_has != 0, o_lo + (e - o_exc), pos, base + e);
    }
}

template <typename R, typename F>
BOLAP_DEV void expand_core(const Args& a, int64_t p0, R&& reserve, F&& body) {
    const WaveRanges w = wave_ranges(a, p0);
    wave_emit(w, p0, reserve(w.total), body);
}

template <typename F>
BOLAP_DEV void expand_wave(const Args& a, int64_t p0, F&& body) {
    expand_core(a, p0, [](int64_t) { return int64_t(0); },
                [&](int32_t g, int32_t v, bool has, int64_t i, uint64_t pos, int64_t) {
                    body(g, v, has, i, pos);
                });
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

// Expanded rows as sort pairs: key (slot << 1 | has), value the item's
// value.  A workgroup of kET threads takes kET join rows per round (a
// 64-row batch per wave) and reserves the round's rows with one atomic on
// `cursor` (one atomic per wave serialised 62,500 same-address atomics on
// 4 M join rows).  A range value outside [0, keys) is counted as an error
// (the call then fails) and its row gets the largest key.
constexpr int kET = 1024;

__global__ __launch_bounds__(kET) void mm_expand_k(Args a, uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals,
                                                   unsigned long long* __restrict__ cursor,
                                                   int64_t cap, uint64_t bad_key,
                                                   int64_t* __restrict__ st) {
    constexpr int kEW = kET / kWave;
    __shared__ int64_t s_tot[kEW];
    __shared__ int64_t s_base;
    const int lane = int(threadIdx.x & (kWave - 1));
    const int wave = int(threadIdx.x / kWave);
    uint64_t errs = 0;
    // the trip count is uniform across the workgroup (barriers below)
    for (int64_t r0 = int64_t(blockIdx.x) * kET; r0 < a.pairs; r0 += int64_t(gridDim.x) * kET) {
        const int64_t p0 = r0 + int64_t(wave) * kWave;
        const WaveRanges w = wave_ranges(a, p0);
        if (lane == 0) s_tot[wave] = w.total;
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t sum = 0;
            for (int k = 0; k < kEW; ++k) sum += s_tot[k];
            s_base = sum ? int64_t(atomicAdd(cursor, (unsigned long long)sum)) : 0;
        }
        __syncthreads();
        int64_t base = s_base;
        for (int k = 0; k < wave; ++k) base += s_tot[k];
        wave_emit(w, p0, base,
                  [&](int32_t g, int32_t v, bool has, int64_t i, uint64_t, int64_t out) {
                      const int32_t k = a.rv[i];
                      uint64_t key = bad_key;
                      if (k < 0 || k >= a.keys)
                          ++errs;
                      else
                          key = ((uint64_t(g) * uint64_t(a.keys) + uint64_t(k)) << 1) |
                                (has ? 1u : 0u);
                      if (out < cap) {  // cap = the counted rows
                          keys[out] = key;
                          vals[out] = uint32_t(v);
                      }
                  });
        __syncthreads();  // s_tot and s_base are rewritten by the next round
    }
    errs = wave_sum(errs);
    if (lane == 0 && errs) atomicAdd(reinterpret_cast<unsigned long long*>(st + mErrors), errs);
}

// Rows [c0, c1) of the sorted pairs belong to wave w: chunk is a multiple of
// the wave width, so tiles of 64 rows never straddle two chunks.
BOLAP_DEV bool run_head(const uint64_t* __restrict__ keys, int64_t i, int64_t n) {
    return i < n && (i == 0 || (keys[i] >> 1) != (keys[i - 1] >> 1));
}

// Run heads (rows whose slot differs from the previous row's) per wave
// chunk; cnt[waves] is set to 0, so that after an exclusive scan over
// waves + 1 entries it holds the total (the group count).
__global__ __launch_bounds__(kT) void mm_heads_k(const uint64_t* __restrict__ keys, int64_t n,
                                                 int64_t chunk, int64_t waves,
                                                 int64_t* __restrict__ cnt) {
    const int lane = int(threadIdx.x & (kWave - 1));
    const int64_t w = (int64_t(blockIdx.x) * kT + threadIdx.x) / kWave;
    const int64_t c0 = w * chunk;
    const int64_t c1 = c0 + chunk < n ? c0 + chunk : n;
    uint64_t heads = 0;
    constexpr int kU = 4;  // tiles per round, their loads in flight together
    for (int64_t t0 = c0; t0 < c1; t0 += kU * kWave) {
        bool h[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) h[u] = run_head(keys, t0 + u * kWave + lane, c1);
#pragma unroll
        for (int u = 0; u < kU; ++u) heads += uint64_t(__popcll(__ballot(h[u])));
    }
    if (lane == 0) cnt[w] = int64_t(heads);
    if (w == 0 && lane == 0) cnt[waves] = 0;
}

__global__ void mm_total_k(const int64_t* __restrict__ scanned, int64_t waves,
                           int64_t* __restrict__ st) {
    st[mGroups] = scanned[waves];
}

struct RunAcc {
    int32_t mn, mx, has;
};

BOLAP_DEV RunAcc run_join(RunAcc x, RunAcc y) {
    return RunAcc{x.mn < y.mn ? x.mn : y.mn, x.mx > y.mx ? x.mx : y.mx, x.has | y.has};
}

BOLAP_DEV void emit_run(bolap_minmax_row* out, int64_t id, int64_t max_rows, uint64_t key,
                        int32_t keys_n, RunAcc r) {
    if (id < 0 || id >= max_rows) return;  // only a broken input could reach here
    const uint64_t slot = key >> 1;
    out[id] = bolap_minmax_row{int32_t(slot / uint64_t(keys_n)), int32_t(slot % uint64_t(keys_n)),
                               r.has ? 1 : 0, r.has ? r.mn : 0, r.has ? r.mx : 0};
}

// Reduces the runs that start in wave w's chunk of the sorted pairs.  Rows
// of the chunk before its first head continue the previous wave's run and
// are skipped; the run still open at the chunk's end is followed into the
// next chunks until it ends.  Inside a tile a segmented scan across the
// wave folds each run's rows into the run's last lane in the tile; a run
// reaching the tile's last lane is carried (uniformly, through readlane)
// into the next tile.  base[w] is the output index of the chunk's first run.
__global__ __launch_bounds__(kT) void mm_reduce_k(const uint64_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ vals, int64_t n,
                                                  int64_t chunk, const int64_t* __restrict__ base,
                                                  int32_t keys_n, int64_t max_rows,
                                                  bolap_minmax_row* __restrict__ out) {
    const int lane = int(threadIdx.x & (kWave - 1));
    const int64_t w = (int64_t(blockIdx.x) * kT + threadIdx.x) / kWave;
    const int64_t c0 = w * chunk;
    if (c0 >= n) return;  // uniform across the wave
    const int64_t c1 = c0 + chunk < n ? c0 + chunk : n;
    const RunAcc none{INT32_MAX, INT32_MIN, 0};
    int64_t rid = base[w];        // output index of the next head
    bool started = c0 == 0;       // a head of this chunk has been seen
    bool open = false;            // a run reaches into the next tile
    RunAcc carry = none;
    uint64_t carry_key = 0;
    int64_t carry_id = 0;
    // rounds of kR tiles: each tile's key, the previous row's key and the
    // value are loaded for the whole round before the first is used
    constexpr int kR = 4;
    bool done = false;
    for (int64_t r0 = c0; r0 < n && !done; r0 += kR * kWave) {
        uint64_t kk[kR], pp[kR];
        uint32_t vv[kR];
    #pragma unroll
        for (int u = 0; u < kR; ++u) {
            const int64_t i = r0 + u * kWave + lane;
            kk[u] = i < n ? keys[i] : 0;
            pp[u] = i < n && i > 0 ? keys[i - 1] : 0;
            vv[u] = i < n ? vals[i] : 0u;
        }
    #pragma unroll
        for (int u = 0; u < kR; ++u) {
            const int64_t t0 = r0 + u * kWave;
            const bool ext = t0 >= c1;  // past the chunk: only the open run
            if (t0 >= n || (ext && !open)) {
                done = true;
                break;
            }
            const uint64_t key_c = kk[u], prev_c = pp[u];
            const uint32_t val_c = vv[u];
            const int64_t i = t0 + lane;
            const bool valid = i < n;
            const uint64_t key = key_c;
            const bool head = valid && (i == 0 || (key_c >> 1) != (prev_c >> 1));
            const uint64_t hmask = __ballot(head);
            const int first = hmask ? __ffsll((unsigned long long)hmask) - 1 : kWave;
            // live rows: in the chunk once started; before `started`, from the
            // first head on; past the chunk, up to the first head
            bool live;
            if (ext)
                live = valid && lane < first;
            else if (!started)
                live = valid && lane >= first;
            else
                live = valid;
            const bool lhead = head && live && !ext;
            const uint64_t lmask = __ballot(lhead);
            RunAcc r = none;
            if (live) {
                const bool h = key & 1u;
                const int32_t v = int32_t(val_c);
                r = RunAcc{h ? v : INT32_MAX, h ? v : INT32_MIN, h ? 1 : 0};
            }
            // segmented inclusive scan: f says a head lies in (lane - d, lane]
            bool f = lhead;
    #pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                const RunAcc o{__shfl_up(r.mn, d, kWave), __shfl_up(r.mx, d, kWave),
                               __shfl_up(r.has, d, kWave)};
                const bool of = __shfl_up(f ? 1 : 0, d, kWave) != 0;
                if (lane >= d) {
                    if (!f) r = run_join(r, o);
                    f = f || of;
                }
            }
            // the open run goes on into this tile when its first row is live and
            // not a head; otherwise it ended with the previous tile
            const bool cont = open && (__shfl(live && !lhead ? 1 : 0, 0, kWave) != 0);
            if (open && !cont && lane == 0) emit_run(out, carry_id, max_rows, carry_key, keys_n, carry);
            // lanes before the tile's first live head continue the open run
            if (!f && cont) r = run_join(r, carry);
            // a live lane ends its run when the next row is a live head or is
            // not live (end of the input, or past the chunk's run)
            const bool next_live = __shfl_down(live ? 1 : 0, 1, kWave) != 0;
            const bool next_head = __shfl_down(lhead ? 1 : 0, 1, kWave) != 0;
            int top = live ? lane : -1;
    #pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                const int o = __shfl_xor(top, d, kWave);
                top = o > top ? o : top;
            }
            const int64_t my_id = f ? rid + __popcll(lmask & ((2ull << lane) - 1ull)) - 1 : carry_id;
            const uint64_t my_key = f ? key : carry_key;
            const bool ends_here = live && lane < kWave - 1 && (next_head || !next_live);
            // the run at the top live lane stays open only when every lane is
            // live and the input goes on
            const bool tile_open = top == kWave - 1 && t0 + kWave < n;
            if (live && (f || cont) && (ends_here || (lane == top && !tile_open)))
                emit_run(out, my_id, max_rows, my_key, keys_n, r);
            if (tile_open) {
                const bool top_f = __shfl(f ? 1 : 0, kWave - 1, kWave) != 0;
                carry = RunAcc{__shfl(r.mn, kWave - 1, kWave), __shfl(r.mx, kWave - 1, kWave),
                               __shfl(r.has, kWave - 1, kWave)};
                if (top_f) {
                    carry_key = uint64_t(__shfl((long long)key, kWave - 1, kWave));
                    carry_id = int64_t(__shfl((long long)my_id, kWave - 1, kWave));
                }
                open = top_f || cont;
            } else {
                open = false;
            }
            if (lmask) started = true;
            rid += __popcll(lmask);
        }
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
    size_t o_ck = 0, o_cs = 0, o_occ = 0, o_oscr = 0, o_cur = 0, o_pk = 0;
    int kbits = 0;
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
    p.o_pk = ws.plan(size_t(std::max<int64_t>(1, a->items)) * sizeof(int4));
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
        // sorted: one (key, value) pair per expanded row
        p.bound = std::max<int64_t>(1, expanded);
        p.kbits = log2_ceil(p.slots) + 1;
        p.o_ck = ws.plan(size_t(p.bound) * 8);
        p.o_cs = ws.plan(size_t(p.bound) * 4);
        p.o_sort = ws.plan(radix_sort_scratch_bytes(ctx, p.bound));
        p.nb = int64_t(grid_of(ctx, p.bound));
        // one chunk of whole 64-row tiles per wave
        const int64_t waves = p.nb * (kT / kWave);
        p.chunk = ((p.bound + waves - 1) / waves + kWave - 1) / kWave * kWave;
        p.o_occ = ws.plan(size_t(waves + 1) * 8);
        p.o_oscr = ws.plan(scan_scratch_bytes(waves + 1));
        p.o_cur = ws.plan(8);
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
    k.pk = nullptr;
    return k;
}

}  // namespace

size_t minmax_pack_end(const bolap_minmax_args* a, size_t base) {
    return WsArena::align(base) +
           WsArena::align(size_t(std::max<int64_t>(1, a->items)) * sizeof(int4));
}

int minmax_expanded(bolap_ctx* ctx, const bolap_minmax_args* a, size_t base, int64_t* expanded) {
    BOLAP_TRY(check_args(a, 0));
    Args k = args_of(a);
    // the item columns packed at align(base), where the call's plan puts
    // them (its first region): the count and the call read one 16-byte row
    // per join row
    BOLAP_TRY(ws_reserve(ctx, minmax_pack_end(a, base), base));
    if (a->items > 0) {
        int4* pk = reinterpret_cast<int4*>(static_cast<char*>(ctx->ws) + WsArena::align(base));
        hipLaunchKernelGGL(mm_pack_k, dim3(grid_of(ctx, a->items)), dim3(kT), 0, ctx->stream, k,
                           pk);
        BOLAP_TRY(launch_check("mm_pack_k"));
        k.pk = pk;
    }
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
    Args k = args_of(a);
    int64_t* st = static_cast<int64_t*>(ctx->result_dev);
    int64_t h[mN];
    const Plan p = make_plan(ctx, a, expanded, first_seen, base);
    BOLAP_ARG_CHECK(p.dense || p.slots <= (int64_t(1) << 62),
                    "bolap_group_minmax: %lld slots exceed 2^62", (long long)p.slots);
    // the caller's arrays below `base` and the items minmax_expanded packed
    // stay where they are: a workspace that must grow keeps them, but only a
    // caller that reserved p.end already may pass pointers into the workspace
    BOLAP_TRY(ws_reserve(ctx, p.end, minmax_pack_end(a, base)));
    WsArena ws;
    ws.base = static_cast<char*>(ctx->ws);
    if (a->items > 0) k.pk = ws.at<int4>(p.o_pk);  // packed by minmax_expanded
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
        uint64_t* ck = ws.at<uint64_t>(p.o_ck);
        uint32_t* cv = ws.at<uint32_t>(p.o_cs);
        int64_t* wcnt = ws.at<int64_t>(p.o_occ);
        const int64_t waves = p.nb * (kT / kWave);
        {
            TraceScope ts(ctx, "minmax_sorted");
            auto* cur = ws.at<unsigned long long>(p.o_cur);
            hipLaunchKernelGGL(fill_u64_k, dim3(1), dim3(kT), 0, ctx->stream, cur, int64_t(1),
                               0ull);
            const uint64_t bad_key = p.kbits >= 64 ? ~0ull : ((uint64_t(1) << p.kbits) - 1);
            const int64_t eb = std::max<int64_t>(
                1, std::min<int64_t>((a->pairs + kET - 1) / kET, int64_t(ctx->num_cus) * 2));
            hipLaunchKernelGGL(mm_expand_k, dim3(unsigned(eb)), dim3(kET), 0, ctx->stream,
                               k, ck, cv, cur, expanded, bad_key, st);
            BOLAP_TRY(launch_check("mm_expand_k"));
            int passes = 0;
            BOLAP_TRY(radix_sort_pairs(ctx, ck, cv, expanded, p.kbits, ws.at<void>(p.o_sort),
                                       &passes));
            hipLaunchKernelGGL(mm_heads_k, dim3(unsigned(p.nb)), dim3(kT), 0, ctx->stream, ck,
                               expanded, p.chunk, waves, wcnt);
            BOLAP_TRY(launch_check("mm_heads_k"));
            BOLAP_TRY(exclusive_scan_i64(ctx, wcnt, waves + 1, ws.at<void>(p.o_oscr)));
            hipLaunchKernelGGL(mm_total_k, dim3(1), dim3(1), 0, ctx->stream, wcnt, waves, st);
            BOLAP_TRY(launch_check("mm_total_k"));
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
            hipLaunchKernelGGL(mm_reduce_k, dim3(unsigned(p.nb)), dim3(kT), 0, ctx->stream, ck, cv,
                               expanded, p.chunk, wcnt, a->keys, groups_n, rows_out);
            BOLAP_TRY(launch_check("mm_reduce_k"));
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
    BOLAP_TRY(bolap::minmax_expanded(ctx, a, 0, &expanded));
    return bolap::minmax_run(ctx, a, expanded, false, 0, rows_out, max_rows, n_out, stats);
}
