// chain_join.hip — customer filter -> orders filter + join -> fact join ->
// GROUP BY order key SUM(price * (100 - disc)) -> ORDER BY revenue DESC,
// day ASC, key ASC (top-k or all groups).
//
// Replaces run_q3 (bespoke_tpch/query_q3.cpp:153-365): the customer
// segment mask (:182-205), the date-filtered orders (:207-210, :306-318), the
// per-order revenue over the order's fact rows (:219-260, which the reference
// reads through precomputed contiguous ranges) and the final std::sort
// (:348-356).  Here the fact table may be in any order:
//   stats    min/max of the customer and order keys (key ranges for bitmaps)
//   cust     bitmap over the customer key range, bit = segment matches
//   count    (exact sizing only) orders passing day < limit and the bitmap
//   init     group table (open addressing, EMPTY keys) + side arrays cleared
//   build    passing orders inserted (key -> slot, day stored), an exact
//            bitmap over the order key range marks them for the fact probe
//   probe    fact keys streamed with 128-bit loads; bitmap test; runs of the
//            same key inside a lane are merged before the slot lookup; one
//            64-bit atomic per run adds the revenue, a plain byte store marks
//            the group as hit
//   compact  hit groups -> (revenue, day, key) rows (wave-aggregated appends)
//   order    rows packed into one u64 key (revenue desc, day, key) when the
//            value ranges fit 64 bits, then radix select (top-k) or a full
//            LSD radix sort; otherwise three stable radix passes.
#include <algorithm>
#include <cstring>
#include <vector>

#include "fold.hpp"
#include "prims.hpp"
#include "radix.hpp"

namespace bolap {
namespace {

constexpr int kT = 256;
constexpr int32_t kEmpty = INT32_MIN;
constexpr uint32_t kSlotMul = 0x9E3779B1u;

// Statistics slots (replicated, see fold.hpp); read_stats folds them.
enum {
    sCMin = 0, sCMax, sOMin, sOMax, sCustPass, sOrderPass, sCount, sDMin, sDMax, sMatch,
    sGroups, sFlags, sRMin, sRMax, sBad, sN = 16
};

BOLAP_DEV bool is_min_stat(int i) { return i == sCMin || i == sOMin || i == sDMin || i == sRMin; }
BOLAP_DEV bool is_max_stat(int i) { return i == sCMax || i == sOMax || i == sDMax || i == sRMax; }

struct alignas(16) RowT {
    int64_t rev;
    int32_t day;
    int32_t key;
};

static_assert(sN <= kStatLine, "statistics fit a replica line");

__global__ void stats_init_k(int64_t* st) {
    if (threadIdx.x < kStatWords) {
        const int i = threadIdx.x % kStatLine;
        st[threadIdx.x] = is_min_stat(i) ? INT64_MAX : is_max_stat(i) ? INT64_MIN : 0;
    }
}

// ---------------------------------------------------------------------------
// Top-k of the group rows for small k, without reading anything back first.
// The rows' revenues are bucketed by a non-decreasing 12-bit code (sign,
// position of the leading one, five bits after it).  A histogram of the
// codes gives the smallest code with k rows at or above it; those
// candidates (at most kSelCap) are gathered and one workgroup sorts them by
// (revenue desc, day, key).  With more, the call takes the radix path.
// ---------------------------------------------------------------------------
constexpr int kTopMax = 64;      // largest k taken by this path
constexpr int kCodes = 4096;     // revenue codes
constexpr int kSelCap = 2048;    // candidates one workgroup sorts
constexpr int kSortT = 1024;     // threads of the sorting workgroup

BOLAP_DEV bool row_before(const RowT& a, const RowT& b) {
    if (a.rev != b.rev) return a.rev > b.rev;
    if (a.day != b.day) return a.day < b.day;
    return a.key < b.key;
}

// Magnitude code of v >= 0: 0 for 0, else 1 + 32 * (position of the
// leading one) + the five bits after it (at most 2016).
BOLAP_DEV uint32_t mag_code(uint64_t v) {
    if (v == 0) return 0;
    const int e = 63 - __clzll((long long)v);
    const uint32_t m = e >= 5 ? uint32_t(v >> (e - 5)) & 31u : uint32_t(v << (5 - e)) & 31u;
    return 1u + uint32_t(e) * 32u + m;
}

// Non-decreasing in rev: negative revenues take codes below 2048.
BOLAP_DEV uint32_t rev_code(int64_t rev) {
    if (rev >= 0) return 2048u + mag_code(uint64_t(rev));
    return 2047u - mag_code(uint64_t(-(rev + 1)));
}

// Exclusive prefix sum of one u32 per thread over the workgroup (kT
// threads); *total = the sum.  Every thread must call it.
BOLAP_DEV uint32_t block_exscan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t s_w[kT / kWave];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += y;
    }
    __syncthreads();
    if (lane == kWave - 1) s_w[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kT / kWave; ++w) {
        before += w < wave ? s_w[w] : 0u;
        all += s_w[w];
    }
    *total = all;
    return before + incl - v;
}

// The smallest code c with at least k rows at or above it in the histogram
// s_h (LDS, kCodes bins), or 0 when it holds fewer than k rows.  Every
// thread of the workgroup calls it and gets the same answer.
BOLAP_DEV uint32_t find_cut(const uint32_t* s_h, int32_t k) {
    __shared__ uint32_t s_rev[kT];
    __shared__ uint32_t s_cut;
    constexpr int kPer = kCodes / kT;  // thread t holds codes [t * kPer, (t + 1) * kPer)
    uint32_t h[kPer], mine = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        h[q] = s_h[threadIdx.x * kPer + q];
        mine += h[q];
    }
    if (threadIdx.x == 0) s_cut = 0;
    // rows in the codes of higher threads: a scan in reversed thread order
    s_rev[kT - 1 - threadIdx.x] = mine;
    __syncthreads();
    const uint32_t r_in = s_rev[threadIdx.x];
    uint32_t total = 0;
    const uint32_t ex = block_exscan(r_in, &total);
    __syncthreads();
    s_rev[kT - 1 - threadIdx.x] = ex;
    __syncthreads();
    const uint32_t above = s_rev[threadIdx.x];
    if (total >= uint32_t(k) && above < uint32_t(k) && above + mine >= uint32_t(k)) {
        uint32_t acc = above;
        int q = kPer - 1;
        while (acc + h[q] < uint32_t(k)) acc += h[q--];
        s_cut = uint32_t(threadIdx.x * kPer + q);
    }
    __syncthreads();
    return s_cut;
}

// Histogram of the rows' revenue codes: per workgroup in LDS, then one
// atomic per non-empty bin.
__global__ __launch_bounds__(kT) void topk_hist_k(const RowT* __restrict__ rows,
                                                  const unsigned long long* __restrict__ n_rows,
                                                  uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_h[kCodes];
    for (int i = threadIdx.x; i < kCodes; i += kT) s_h[i] = 0;
    __syncthreads();
    const int64_t n = int64_t(*n_rows);
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        atomicAdd(&s_h[rev_code(rows[i].rev)], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kCodes; i += kT)
        if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// Rows with a code at or above the cut, staged in LDS and appended with one
// atomic per workgroup; rows beyond kSelCap are counted but not written.
__global__ __launch_bounds__(kT) void topk_select_k(const RowT* __restrict__ rows,
                                                    const unsigned long long* __restrict__ n_rows,
                                                    const uint32_t* __restrict__ hist, int32_t k,
                                                    RowT* __restrict__ cand,
                                                    unsigned long long* __restrict__ n_cand) {
    __shared__ uint32_t s_n;
    __shared__ unsigned long long s_base;
    __shared__ RowT s_rows[kSelCap];
    __shared__ uint32_t s_h[kCodes];
    if (threadIdx.x == 0) s_n = 0;
    for (int i = threadIdx.x; i < kCodes; i += kT) s_h[i] = hist[i];
    __syncthreads();
    const uint32_t c = find_cut(s_h, k);
    const int64_t n = int64_t(*n_rows);
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = int64_t(blockIdx.x) * per;
    const int64_t b1 = b0 + per < n ? b0 + per : n;
    for (int64_t i = b0 + threadIdx.x; i < b1; i += kT) {
        const RowT x = rows[i];
        if (rev_code(x.rev) >= c) {
            const uint32_t at = atomicAdd(&s_n, 1u);
            if (at < uint32_t(kSelCap)) s_rows[at] = x;
        }
    }
    __syncthreads();
    const uint32_t cnt = s_n;
    if (threadIdx.x == 0 && cnt) s_base = atomicAdd(n_cand, (unsigned long long)cnt);
    __syncthreads();
    if (cnt == 0) return;
    const unsigned long long base = s_base;
    const uint32_t mine = cnt < uint32_t(kSelCap) ? cnt : uint32_t(kSelCap);
    for (uint32_t q = threadIdx.x; q < mine; q += kT)
        if (base + q < uint64_t(kSelCap)) cand[base + q] = s_rows[q];
}

// One workgroup: sorts the candidates (bitonic, padded with rows that sort
// last) and writes the first k; *overflow = 1 when there are too many.
__global__ __launch_bounds__(kSortT) void topk_sort_k(const RowT* __restrict__ cand,
                                                      const unsigned long long* __restrict__ n_cand,
                                                      int32_t k,
                                                      bolap_group_row* __restrict__ out,
                                                      unsigned* __restrict__ overflow) {
    __shared__ RowT s[kSelCap];
    if (*n_cand > uint64_t(kSelCap)) {
        if (threadIdx.x == 0) *overflow = 1u;
        return;
    }
    const uint32_t n = uint32_t(*n_cand);
    int size = 2;  // a power of two holding the candidates and the k output rows
    while (size < kSelCap && (uint32_t(size) < n || size < k)) size <<= 1;
    for (int i = threadIdx.x; i < size; i += kSortT)
        s[i] = uint32_t(i) < n ? cand[i] : RowT{INT64_MIN, INT32_MAX, INT32_MAX};
    __syncthreads();
    for (int kk = 2; kk <= size; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < size; i += kSortT) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const RowT a = s[i], b = s[ixj];
                    const bool up = (i & kk) == 0;
                    if (up ? row_before(b, a) : row_before(a, b)) {
                        s[i] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < k; i += kSortT) {
        const RowT x = s[i];
        out[i] = bolap_group_row{x.key, x.day, x.rev};
    }
}

// Min/max of an int32 column: four keys per lane per 16-byte load, two
// loads in flight per iteration.
BOLAP_DEV void col_minmax(const int32_t* __restrict__ k, int64_t n, int64_t& lo, int64_t& hi) {
    const int64_t units = n / 4;
    const int64_t stride = int64_t(gridDim.x) * kT;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < units; u += 2 * stride) {
        const VecT<int32_t, 4> a = load_unit<int32_t, 4>(k, u);
        const int64_t u2 = u + stride < units ? u + stride : u;
        const VecT<int32_t, 4> b = load_unit<int32_t, 4>(k, u2);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            lo = min(lo, int64_t(min(a.v[j], b.v[j])));
            hi = max(hi, int64_t(max(a.v[j], b.v[j])));
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < n - units * 4) {
        const int64_t v = k[units * 4 + threadIdx.x];
        lo = min(lo, v);
        hi = max(hi, v);
    }
}

__global__ __launch_bounds__(kT) void key_minmax_k(const int32_t* __restrict__ ck, int64_t nc,
                                                   const int32_t* __restrict__ ok, int64_t no,
                                                   int64_t* st) {
    int64_t cl = INT64_MAX, ch = INT64_MIN, ol = INT64_MAX, oh = INT64_MIN;
    col_minmax(ck, nc, cl, ch);
    col_minmax(ok, no, ol, oh);
    fold_minmax<kT>(cl, ch, st, sCMin, sCMax);
    fold_minmax<kT>(ol, oh, st, sOMin, sOMax);
}

__global__ __launch_bounds__(kT) void fill_u32_k(uint32_t* p, int64_t n, uint32_t v) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT)
        p[i] = v;
}

struct Ranges {
    int64_t cmin, cmax, omin, omax;
};

BOLAP_DEV bool test_bit(const uint32_t* __restrict__ bits, int64_t rel, int64_t span) {
    if (rel < 0 || rel >= span) return false;
    return (bits[rel >> 5] >> (rel & 31)) & 1u;
}

// Customers of the segment set their bit over the customer key range.  The
// OR of a wave's bits is gathered in an LDS window of 64 words when its
// passing keys span fewer than 64 words (one atomic per word for sorted
// keys); otherwise one atomic per passing customer.
__global__ __launch_bounds__(kT) void cust_bits_k(const int32_t* __restrict__ key,
                                                  const uint8_t* __restrict__ seg, int64_t n,
                                                  int32_t segment, Ranges r,
                                                  uint32_t* __restrict__ bits, int64_t* st) {
    __shared__ uint32_t s_win[kT / kWave][kWave];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    WaveCount wc;
    uint64_t bad = 0;  // keys outside [cmin, cmax] (only with caller-given ranges)
    const int64_t cspan = r.cmax - r.cmin + 1;
    const int64_t units = (n + 3) / 4;
    const int64_t full = n / 4;
    const int64_t stride = int64_t(gridDim.x) * kT;
    const int64_t rounds = (units + stride - 1) / stride;  // the same for every lane
    for (int64_t rd = 0; rd < rounds; ++rd) {
        const int64_t u = rd * stride + int64_t(blockIdx.x) * kT + threadIdx.x;
        int32_t kv[4] = {0, 0, 0, 0};
        uint8_t sv[4] = {0, 0, 0, 0};
        int cnt = 0;
        if (u < full) {
            const VecT<int32_t, 4> k4 = load_unit<int32_t, 4>(key, u);
            const VecT<uint8_t, 4> s4 = load_unit<uint8_t, 4>(seg, u);
            for (int j = 0; j < 4; ++j) {
                kv[j] = k4.v[j];
                sv[j] = s4.v[j];
            }
            cnt = 4;
        } else if (u < units) {
            cnt = int(n - u * 4);
            for (int j = 0; j < 4; ++j) {
                kv[j] = j < cnt ? key[u * 4 + j] : 0;
                sv[j] = j < cnt ? seg[u * 4 + j] : 0;
            }
        }
        uint32_t pm = 0;
        int64_t wlo = INT64_MAX, whi = INT64_MIN;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool pass = j < cnt && int32_t(sv[j]) == segment;
            wc.add(pass);
            const bool inr = int64_t(kv[j]) >= r.cmin && int64_t(kv[j]) <= r.cmax;
            if (j < cnt && !inr) ++bad;
            if (pass && inr) {
                pm |= 1u << j;
                const int64_t w = (int64_t(kv[j]) - r.cmin) >> 5;
                wlo = w < wlo ? w : wlo;
                whi = w > whi ? w : whi;
            }
        }
        const int64_t lo = wave_min(wlo), hi = wave_max(whi);
        if (hi >= lo && hi - lo < kWave) {
            s_win[wave][lane] = 0;
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (!((pm >> j) & 1u)) continue;
                const int64_t rel = int64_t(kv[j]) - r.cmin;
                atomicOr(&s_win[wave][(rel >> 5) - lo], 1u << (rel & 31));
            }
            __builtin_amdgcn_wave_barrier();
            const uint32_t v = s_win[wave][lane];
            if (v) atomicOr(bits + lo + lane, v);
        } else if (hi >= lo) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (!((pm >> j) & 1u)) continue;
                const int64_t rel = int64_t(kv[j]) - r.cmin;
                if (rel < cspan) atomicOr(bits + (rel >> 5), 1u << (rel & 31));
            }
        }
    }
    block_fold<kT>(st, sCustPass, int64_t(wc.n), 0);
    block_fold<kT>(st, sBad, int64_t(wave_sum(bad)), 0);
}

// Four order rows per lane: day and customer columns read with 16-byte loads.
struct Orders4 {
    int32_t day[4], cust[4], key[4];
    int cnt;
};

BOLAP_DEV Orders4 load_orders4(const int32_t* __restrict__ okey, const int32_t* __restrict__ day,
                               const int32_t* __restrict__ cust, int64_t n, int64_t u,
                               bool with_key) {
    Orders4 o;
    if ((u + 1) * 4 <= n) {
        const VecT<int32_t, 4> d = load_unit<int32_t, 4>(day, u);
        const VecT<int32_t, 4> c = load_unit<int32_t, 4>(cust, u);
        VecT<int32_t, 4> k{};
        if (with_key) k = load_unit<int32_t, 4>(okey, u);
        for (int j = 0; j < 4; ++j) {
            o.day[j] = d.v[j];
            o.cust[j] = c.v[j];
            o.key[j] = k.v[j];
        }
        o.cnt = 4;
    } else {
        o.cnt = int(n - u * 4);
        for (int j = 0; j < 4; ++j) {
            const bool in = j < o.cnt;
            o.day[j] = in ? day[u * 4 + j] : 0;
            o.cust[j] = in ? cust[u * 4 + j] : 0;
            o.key[j] = (in && with_key) ? okey[u * 4 + j] : 0;
        }
    }
    return o;
}

BOLAP_DEV bool order_pass_v(int32_t day, int32_t cust, int32_t day_lt,
                            const uint32_t* __restrict__ cbits, int64_t cspan, const Ranges& r) {
    return day < day_lt && test_bit(cbits, int64_t(cust) - r.cmin, cspan);
}

__global__ __launch_bounds__(kT) void order_count_k(const int32_t* __restrict__ day,
                                                    const int32_t* __restrict__ cust, int64_t n,
                                                    int32_t day_lt,
                                                    const uint32_t* __restrict__ cbits,
                                                    int64_t cspan, Ranges r, int64_t* st) {
    WaveCount wc;
    const int64_t units = (n + 3) / 4;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < units;
         u += int64_t(gridDim.x) * kT) {
        const Orders4 o = load_orders4(nullptr, day, cust, n, u, false);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            wc.add(j < o.cnt && order_pass_v(o.day[j], o.cust[j], day_lt, cbits, cspan, r));
    }
    block_fold<kT>(st, sCount, int64_t(wc.n), 0);
}

// Group table.  Hashed: open addressing over `cap` slots.  Direct (dense
// order keys, span <= 8 x orders): slot = key - kbase, no key array, no
// probing; only the hit bytes are cleared per call, days and revenue are
// written for each passing order by the build.
struct RowK {
    int32_t key, day;
};

// Direct mode (dense order keys): a group's slot is the rank of its key's
// bit in `obits` (passing orders only), so revenue and hit arrays hold one
// entry per passing order; rank = toff[word / kRankWords] + wrank[word] +
// popcount of the word's lower bits.  The passing orders' (key, day) pairs
// are appended to `plist` by the build and read back by the compaction.
constexpr int kRankWords = kT * 8;  // bitmap words per rank tile

struct GTable {
    int32_t* keys;   // cap slots, kEmpty = free (hashed only)
    int32_t* days;   // hashed only
    uint64_t* rev;   // revenue per slot (two's complement)
    uint8_t* hit;    // 1 when the group got at least one fact row
    uint32_t* obits; // exact bitmap over the order key range
    int32_t cap_log2;
    uint32_t max_probe;  // insert gives up (overflow) after this many slots
    int32_t direct;
    int64_t kbase;       // key of bit 0 (the smallest order key)
    // direct only
    const uint32_t* wrank;  // per bitmap word: set bits before it in its tile
    const uint32_t* toff;   // per tile: set bits before it
    RowK* plist;            // passing orders (key, day): region b holds build
                            // workgroup b's, from b * list_chunk on
    uint32_t* list_n;       // per build workgroup: pairs in its region
    int64_t list_chunk;     // orders per build workgroup (a multiple of kT * 8)
    int32_t list_regions;   // build workgroups
};

BOLAP_DEV uint32_t rank_in(const GTable& t, int64_t w, uint32_t word, uint32_t bit) {
    return t.toff[w / kRankWords] + t.wrank[w] + __popc(word & ((1u << bit) - 1u));
}

BOLAP_DEV uint32_t gslot(int32_t key, int cap_log2) {
    return (uint32_t(key) * kSlotMul) >> (32 - cap_log2);
}

__global__ __launch_bounds__(kT) void gtable_init_k(GTable t, int64_t owords) {
    const int64_t cap = int64_t(1) << t.cap_log2;
    const int64_t stride = int64_t(gridDim.x) * kT;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < cap; i += stride) {
        if (!t.direct) {
            t.keys[i] = kEmpty;
            t.rev[i] = 0;
        }
        t.hit[i] = 0;
    }
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < owords; i += stride)
        t.obits[i] = 0;
}

__global__ __launch_bounds__(kT) void order_build_k(const int32_t* __restrict__ okey,
                                                    const int32_t* __restrict__ day,
                                                    const int32_t* __restrict__ cust, int64_t n,
                                                    int32_t day_lt,
                                                    const uint32_t* __restrict__ cbits,
                                                    int64_t cspan, Ranges r, GTable t,
                                                    int64_t* st) {
    const uint32_t mask = (uint32_t(1) << t.cap_log2) - 1;
    int64_t dl = INT64_MAX, dh = INT64_MIN;
    WaveCount passed;
    uint64_t inserted = 0, bad = 0;
    bool failed = false;
    const int64_t units = (n + 3) / 4;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < units;
         u += int64_t(gridDim.x) * kT) {
        const Orders4 o = load_orders4(okey, day, cust, n, u, true);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool pass =
                j < o.cnt && order_pass_v(o.day[j], o.cust[j], day_lt, cbits, cspan, r);
            passed.add(pass);
            if (j < o.cnt && (int64_t(o.key[j]) < r.omin || int64_t(o.key[j]) > r.omax)) {
                ++bad;  // only with caller-given ranges
                continue;
            }
            if (!pass || failed) continue;
            const int32_t key = o.key[j];
            const int32_t d = o.day[j];
            dl = d < dl ? d : dl;
            dh = d > dh ? d : dh;
            const int64_t rel = int64_t(key) - r.omin;
            atomicOr(t.obits + (rel >> 5), 1u << (rel & 31));
            if (t.direct) {
                const int64_t slot = int64_t(key) - t.kbase;
                t.days[slot] = d;
                t.rev[slot] = 0;
                ++inserted;
                continue;
            }
            uint32_t s = gslot(key, t.cap_log2);
            bool placed = false;
            for (uint32_t probe = 0; probe < t.max_probe; ++probe) {
                const int32_t prev = atomicCAS(t.keys + s, kEmpty, key);
                if (prev == kEmpty || prev == key) {
                    t.days[s] = d;
                    inserted += prev == kEmpty;
                    placed = true;
                    break;
                }
                s = (s + 1) & mask;
            }
            failed = !placed;  // table too small: stop inserting, report overflow
        }
    }
    block_fold<kT>(st, sOrderPass, int64_t(passed.n), 0);
    block_fold<kT>(st, sGroups, int64_t(wave_sum(inserted)), 0);
    block_fold<kT>(st, sBad, int64_t(wave_sum(bad)), 0);
    block_fold<kT>(st, sFlags, __ballot(failed) ? 1 : 0, 0);
    fold_minmax<kT>(dl, dh, st, sDMin, sDMax);
}

// Direct-mode build: workgroup b filters orders [b * list_chunk, ...) and
// writes the (key, day) pairs of the passing ones to its own region of the
// pass list (staged in LDS, plain stores, no shared counter); passing orders
// set their bit (the OR of a wave's bits is gathered in an LDS window of 64
// words when its passing keys span fewer than 64 words, so a wave over
// sorted keys issues one atomic per word).
constexpr int kMarkStage = 4096;  // staged pairs; a tile adds at most kT * 8

__global__ __launch_bounds__(kT) void order_mark_k(const int32_t* __restrict__ okey,
                                                   const int32_t* __restrict__ day,
                                                   const int32_t* __restrict__ cust, int64_t n,
                                                   int32_t day_lt,
                                                   const uint32_t* __restrict__ cbits,
                                                   int64_t cspan, Ranges r, GTable t,
                                                   int64_t* st) {
    __shared__ uint32_t s_win[kT / kWave][kWave];
    __shared__ RowK s_list[kMarkStage];
    __shared__ uint32_t s_n;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    int64_t dl = INT64_MAX, dh = INT64_MIN;
    WaveCount passed;
    uint64_t bad = 0;
    const int64_t b0 = int64_t(blockIdx.x) * t.list_chunk;
    const int64_t b1 = b0 + t.list_chunk < n ? b0 + t.list_chunk : n;
    RowK* region = t.plist + b0;
    uint32_t written = 0;  // block-uniform
    for (int64_t t0 = b0; t0 < b1; t0 += int64_t(kT) * 8) {
        // two units of four orders per lane, both loaded before either is used
        Orders4 o[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int64_t u = t0 / 4 + q * kT + threadIdx.x;
            o[q].cnt = 0;
            if (u * 4 < b1) o[q] = load_orders4(okey, day, cust, b1, u, true);
        }
        uint32_t pm = 0;  // bit 4q + j: order j of unit q passes
        int64_t wlo = INT64_MAX, whi = INT64_MIN;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool pass = j < o[q].cnt &&
                                  order_pass_v(o[q].day[j], o[q].cust[j], day_lt, cbits, cspan, r);
                passed.add(pass);
                const bool inr =
                    int64_t(o[q].key[j]) >= r.omin && int64_t(o[q].key[j]) <= r.omax;
                if (j < o[q].cnt && !inr) ++bad;  // only with caller-given ranges
                if (pass && inr) {
                    pm |= 1u << (4 * q + j);
                    dl = o[q].day[j] < dl ? o[q].day[j] : dl;
                    dh = o[q].day[j] > dh ? o[q].day[j] : dh;
                    const int64_t w = (int64_t(o[q].key[j]) - r.omin) >> 5;
                    wlo = w < wlo ? w : wlo;
                    whi = w > whi ? w : whi;
                    s_list[atomicAdd(&s_n, 1u)] = RowK{o[q].key[j], o[q].day[j]};
                }
            }
        }
        const int64_t lo = wave_min(wlo), hi = wave_max(whi);
        if (hi >= lo && hi - lo < kWave) {
            s_win[wave][lane] = 0;
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                if (!((pm >> b) & 1u)) continue;
                const int64_t rel = int64_t(o[b >> 2].key[b & 3]) - r.omin;
                atomicOr(&s_win[wave][(rel >> 5) - lo], 1u << (rel & 31));
            }
            __builtin_amdgcn_wave_barrier();
            const uint32_t v = s_win[wave][lane];
            if (v) atomicOr(t.obits + lo + lane, v);
        } else if (hi >= lo) {
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                if (!((pm >> b) & 1u)) continue;
                const int64_t rel = int64_t(o[b >> 2].key[b & 3]) - r.omin;
                atomicOr(t.obits + (rel >> 5), 1u << (rel & 31));
            }
        }
        __syncthreads();
        const uint32_t cnt = s_n;
        __syncthreads();
        if (cnt > uint32_t(kMarkStage - kT * 8)) {
            for (uint32_t q = threadIdx.x; q < cnt; q += kT) region[written + q] = s_list[q];
            written += cnt;
            __syncthreads();
            if (threadIdx.x == 0) s_n = 0;
            __syncthreads();
        }
    }
    __syncthreads();
    const uint32_t cnt = s_n;
    for (uint32_t q = threadIdx.x; q < cnt; q += kT) region[written + q] = s_list[q];
    if (threadIdx.x == 0) t.list_n[blockIdx.x] = written + cnt;
    block_fold<kT>(st, sOrderPass, int64_t(passed.n), 0);
    block_fold<kT>(st, sBad, int64_t(wave_sum(bad)), 0);
    fold_minmax<kT>(dl, dh, st, sDMin, sDMax);
}

// Rank tiles of the order bitmap: wrank[w] = set bits before word w in its
// tile, tsum[tile] = the tile's set bits; the last workgroup turns tsum into
// toff (exclusive prefix) and writes the total (= passing orders) to *total.
__global__ __launch_bounds__(kT) void rank_tiles_k(const uint32_t* __restrict__ obits,
                                                   int64_t owords, uint32_t* __restrict__ wrank,
                                                   uint32_t* __restrict__ tsum,
                                                   uint32_t* __restrict__ toff, int64_t tiles,
                                                   unsigned* counter,
                                                   unsigned long long* __restrict__ total) {
    __shared__ uint32_t s_w[kT / kWave];
    __shared__ uint32_t s_carry;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    // one wave-then-workgroup exclusive scan of a per-thread value
    auto block_scan = [&](uint32_t v, uint32_t* block_total) -> uint32_t {
        uint32_t incl = v;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, kWave);
            if (lane >= off) incl += y;
        }
        __syncthreads();
        if (lane == kWave - 1) s_w[wave] = incl;
        __syncthreads();
        uint32_t before = 0, all = 0;
        for (int w = 0; w < kT / kWave; ++w) {
            before += w < wave ? s_w[w] : 0;
            all += s_w[w];
        }
        *block_total = all;
        return before + incl - v;
    };
    const int64_t w0 = int64_t(blockIdx.x) * kRankWords + int64_t(threadIdx.x) * 8;
    uint32_t c[8], mine = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        c[q] = w0 + q < owords ? uint32_t(__popc(obits[w0 + q])) : 0u;
        mine += c[q];
    }
    uint32_t tile_total = 0;
    uint32_t at = block_scan(mine, &tile_total);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (w0 + q < owords) wrank[w0 + q] = at;
        at += c[q];
    }
    if (threadIdx.x == 0) tsum[blockIdx.x] = tile_total;
    if (!last_block_done(counter)) return;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t t0 = 0; t0 < tiles; t0 += kT) {
        const int64_t i = t0 + threadIdx.x;
        const uint32_t v = i < tiles ? ld_agent(tsum + i) : 0u;
        uint32_t chunk_total = 0;
        const uint32_t ex = block_scan(v, &chunk_total);
        if (i < tiles) toff[i] = s_carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += chunk_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *total = s_carry;
        reset_counter(counter);
    }
}

// Clears revenue and hit entries [0, *n) of the rank-indexed group arrays.
__global__ __launch_bounds__(kT) void zero_ranked_k(uint64_t* __restrict__ rev,
                                                    uint8_t* __restrict__ hit,
                                                    const unsigned long long* __restrict__ n) {
    const int64_t m = int64_t(*n);
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < m;
         i += int64_t(gridDim.x) * kT) {
        rev[i] = 0;
        hit[i] = 0;
    }
}

// Direct-mode compaction: every listed passing order whose group got a
// fact row becomes a (revenue, day, key) row.  Workgroup b walks a
// contiguous range of the build's list regions; rows are staged in LDS and
// appended with one atomic per flush.
constexpr int kRowStage = 1024;  // staged rows; a round adds at most kT

__global__ __launch_bounds__(kT) void compact_list_k(GTable t, Ranges r,
                                                     RowT* __restrict__ out,
                                                     unsigned long long* __restrict__ n_out,
                                                     int64_t* st) {
    __shared__ RowT s_rows[kRowStage];
    __shared__ uint32_t s_n;
    __shared__ unsigned long long s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int32_t per = (t.list_regions + int32_t(gridDim.x) - 1) / int32_t(gridDim.x);
    const int32_t g0 = int32_t(blockIdx.x) * per;
    const int32_t g1 = g0 + per < t.list_regions ? g0 + per : t.list_regions;
    int64_t rl = INT64_MAX, rh = INT64_MIN;
    for (int32_t g = g0; g < g1; ++g) {
        const int64_t n = int64_t(t.list_n[g]);
        const RowK* region = t.plist + int64_t(g) * t.list_chunk;
        for (int64_t t0 = 0; t0 < n; t0 += 2 * kT) {
            // two entries per thread, each step's loads issued for both first
            RowK e[2];
            bool in[2];
            uint32_t word[2], slot[2];
            uint8_t hit[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int64_t i = t0 + q * kT + threadIdx.x;
                in[q] = i < n;
                e[q] = in[q] ? region[i] : RowK{int32_t(r.omin), 0};
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) word[q] = t.obits[(int64_t(e[q].key) - r.omin) >> 5];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int64_t rel = int64_t(e[q].key) - r.omin;
                slot[q] = rank_in(t, rel >> 5, word[q], uint32_t(rel & 31));
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) hit[q] = in[q] ? t.hit[slot[q]] : 0;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                if (!hit[q]) continue;
                const RowT row{int64_t(t.rev[slot[q]]), e[q].day, e[q].key};
                rl = row.rev < rl ? row.rev : rl;
                rh = row.rev > rh ? row.rev : rh;
                s_rows[atomicAdd(&s_n, 1u)] = row;
            }
            __syncthreads();
            const uint32_t cnt = s_n;
            __syncthreads();
            if (cnt > uint32_t(kRowStage - 2 * kT))
                flush_stage<kT>(s_rows, &s_n, &s_base, cnt, out, n_out);
        }
    }
    __syncthreads();
    const uint32_t cnt = s_n;
    __syncthreads();
    flush_stage<kT>(s_rows, &s_n, &s_base, cnt, out, n_out);
    fold_minmax<kT>(rl, rh, st, sRMin, sRMax);
}

// Slot of a key known to be in the table.
BOLAP_DEV uint32_t find_slot(int32_t key, const int32_t* __restrict__ keys, int cap_log2) {
    const uint32_t mask = (uint32_t(1) << cap_log2) - 1;
    uint32_t s = gslot(key, cap_log2);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        if (keys[s] == key) return s;
        s = (s + 1) & mask;
    }
    return 0xFFFFFFFFu;
}

template <typename TP, typename TD>
__global__ __launch_bounds__(kT) void fact_probe_k(const int32_t* __restrict__ fk,
                                                   const TP* __restrict__ price,
                                                   const TD* __restrict__ disc,
                                                   const int32_t* __restrict__ dprice,
                                                   const int16_t* __restrict__ ship,
                                                   int32_t ship_gt, int64_t n, Ranges r,
                                                   GTable t, int64_t* st) {
    const int64_t ospan = r.omax - r.omin + 1;
    const int64_t n_units = n / 4;
    uint64_t matched = 0;
    auto row_value = [&](int64_t row) -> int64_t {
        if (dprice) return int64_t(dprice[row]);
        return int64_t(uint64_t(int64_t(price[row])) * uint64_t(100 - int64_t(disc[row])));
    };
    // Processes rows [row0, row0 + cnt) whose keys are in kv, merging runs of
    // equal keys before one atomic per run.
    auto run_rows = [&](const int32_t* kv, int64_t row0, int cnt) {
        int32_t cur_key = 0;
        uint32_t cur_slot = 0xFFFFFFFFu;
        uint64_t acc = 0;
        for (int j = 0; j < cnt; ++j) {
            const int32_t key = kv[j];
            const int64_t rel = int64_t(key) - r.omin;
            if (rel < 0 || rel >= ospan) continue;
            const uint32_t word = t.obits[rel >> 5];
            if (!((word >> (rel & 31)) & 1u)) continue;
            if (ship && ship[row0 + j] <= ship_gt) continue;  // run_q3's shipdate test
            ++matched;
            const uint64_t v = uint64_t(row_value(row0 + j));
            if (cur_slot != 0xFFFFFFFFu && key == cur_key) {
                acc += v;
                continue;
            }
            if (cur_slot != 0xFFFFFFFFu) {
                atomicAdd(reinterpret_cast<unsigned long long*>(t.rev + cur_slot),
                          (unsigned long long)acc);
                t.hit[cur_slot] = 1;
            }
            cur_key = key;
            // a key missing from an overflowed (estimate-sized) hashed table
            // finds no slot (0xFFFFFFFF): its run is skipped, and the exact
            // rerun discards this table's results
            cur_slot = t.direct ? rank_in(t, rel >> 5, word, uint32_t(rel & 31))
                                : find_slot(key, t.keys, t.cap_log2);
            acc = v;
        }
        if (cur_slot != 0xFFFFFFFFu) {
            atomicAdd(reinterpret_cast<unsigned long long*>(t.rev + cur_slot),
                      (unsigned long long)acc);
            t.hit[cur_slot] = 1;
        }
    };
    // Two 4-key units per lane per iteration (both loads in flight), then the
    // units' runs in row order.
    const int64_t stride = int64_t(gridDim.x) * kT;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < n_units; u += 2 * stride) {
        const int64_t u2 = u + stride;
        const VecT<int32_t, 4> v1 = load_unit<int32_t, 4>(fk, u);
        const VecT<int32_t, 4> v2 = load_unit<int32_t, 4>(fk, u2 < n_units ? u2 : u);
        run_rows(v1.v, u * 4, 4);
        if (u2 < n_units) run_rows(v2.v, u2 * 4, 4);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int32_t kv[4];
        const int rem = int(n - n_units * 4);
        for (int j = 0; j < rem; ++j) kv[j] = fk[n_units * 4 + j];
        run_rows(kv, n_units * 4, rem);
    }
    block_fold<kT>(st, sMatch, int64_t(wave_sum(matched)), 0);
}

constexpr int kCompactR = 8;    // slots per lane per tile
constexpr int kCT = 1024;       // compaction workgroup
constexpr int kCTile = kCT * kCompactR;

BOLAP_DEV uint32_t hit_mask8(const uint8_t* __restrict__ hit, int64_t i0, int64_t end) {
    uint32_t m = 0;
    if (i0 + kCompactR <= end) {
        const VecT<uint8_t, kCompactR> h = load_unit<uint8_t, kCompactR>(hit, i0 / kCompactR);
#pragma unroll
        for (int j = 0; j < kCompactR; ++j) m |= (h.v[j] ? 1u : 0u) << j;
    }
    return m;
}

// Orders whose fact rows are given as contiguous ranges (the reference's
// lineitem_ranges, query_q3.cpp:219-260): no group table, two passes.
//   filter  workgroup b walks orders [b * chunk, (b + 1) * chunk), sixteen
//           per lane per tile (four 16-byte loads per column), and appends
//           the rows of passing orders to an LDS list that is flushed to
//           `pass_idx` with one atomic per flush (when it could not take
//           another tile, and at the end);
//   sum     the passing orders, one per lane with every lane busy: range,
//           key and day gathers, then the range's fact rows (eight
//           independent loads per round); orders with rows append their
//           group row through an LDS buffer the same way.
constexpr int kRangeBuf = 2048;  // staged group rows; a sum tile adds at most kT * 4
constexpr int kFilterPer = 16;   // orders per lane per filter tile
constexpr int kPassBuf = kT * kFilterPer;  // orders per filter tile
constexpr int kPassCap = 2 * kPassBuf;     // staged passing orders (32 KB)

struct RangeArgs {
    const int32_t* okey;
    const int32_t* day;
    const int32_t* cust;
    const uint32_t* start;
    const uint32_t* end;
    int64_t n;
    int32_t day_lt;
    int64_t chunk;
    const int16_t* ship;  // optional: sum only fact rows with ship > ship_gt
    int32_t ship_gt;
    // Optional: the fact rows' order keys when ranges are min/max spans over
    // unsorted rows; a row counts for an order only when its key matches.
    const int32_t* forder;
};

__global__ __launch_bounds__(kT) void range_filter_k(RangeArgs a, const uint32_t* __restrict__ cbits,
                                                     int64_t cspan, Ranges r,
                                                     uint32_t* __restrict__ pass_idx,
                                                     unsigned long long* __restrict__ n_pass,
                                                     int64_t* st) {
    __shared__ uint32_t s_idx[kPassCap];
    __shared__ uint32_t s_n;
    __shared__ unsigned long long s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    WaveCount passed;
    const int64_t b0 = int64_t(blockIdx.x) * a.chunk;
    const int64_t b1 = b0 + a.chunk < a.n ? b0 + a.chunk : a.n;
    for (int64_t t0 = b0; t0 < b1; t0 += int64_t(kPassBuf)) {
        // Lane l takes units t0/4 + l + q * kT, q < 4 (coalesced per q).
        Orders4 o[kFilterPer / 4];
#pragma unroll
        for (int q = 0; q < kFilterPer / 4; ++q) {
            const int64_t i0 = t0 + (int64_t(threadIdx.x) + int64_t(q) * kT) * 4;
            if (i0 < b1) o[q] = load_orders4(a.okey, a.day, a.cust, b1, i0 / 4, false);
            else o[q].cnt = 0;
        }
#pragma unroll
        for (int q = 0; q < kFilterPer / 4; ++q) {
            const int64_t i0 = t0 + (int64_t(threadIdx.x) + int64_t(q) * kT) * 4;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool pass = j < o[q].cnt && order_pass_v(o[q].day[j], o[q].cust[j],
                                                               a.day_lt, cbits, cspan, r);
                passed.add(pass);
                if (pass) s_idx[atomicAdd(&s_n, 1u)] = uint32_t(i0 + j);
            }
        }
        __syncthreads();
        const uint32_t cnt = s_n;
        __syncthreads();
        if (cnt > uint32_t(kPassCap - kPassBuf))
            flush_stage<kT>(s_idx, &s_n, &s_base, cnt, pass_idx, n_pass);
    }
    __syncthreads();
    const uint32_t cnt = s_n;
    __syncthreads();
    flush_stage<kT>(s_idx, &s_n, &s_base, cnt, pass_idx, n_pass);
    block_fold<kT>(st, sOrderPass, int64_t(passed.n), 0);
}

template <typename TP, typename TD>
__global__ __launch_bounds__(kT) void range_groups_k(RangeArgs a, const TP* __restrict__ price,
                                                     const TD* __restrict__ disc,
                                                     const int32_t* __restrict__ dprice,
                                                     const uint32_t* __restrict__ pass_idx,
                                                     const unsigned long long* __restrict__ n_pass,
                                                     RowT* __restrict__ out,
                                                     unsigned long long* __restrict__ n_out,
                                                     int64_t* st) {
    __shared__ RowT s_rows[kRangeBuf];
    __shared__ uint32_t s_n;
    __shared__ unsigned long long s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    int64_t dl = INT64_MAX, dh = INT64_MIN, kl = INT64_MAX, kh = INT64_MIN;
    int64_t rl = INT64_MAX, rh = INT64_MIN;
    uint64_t matched = 0;
    const int64_t np = int64_t(*n_pass);
    // Contiguous share of the list per workgroup, whole tiles of kT * 4.
    int64_t chunk = (np + gridDim.x - 1) / gridDim.x;
    chunk = (chunk + kT * 4 - 1) / (kT * 4) * (kT * 4);
    const int64_t b0 = int64_t(blockIdx.x) * chunk;
    const int64_t b1 = b0 + chunk < np ? b0 + chunk : np;
    for (int64_t t0 = b0; t0 < b1; t0 += int64_t(kT) * 4) {
        uint32_t oi[4], s[4], e[4];
        int32_t key[4], day[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t p = t0 + threadIdx.x + int64_t(q) * kT;
            oi[q] = p < b1 ? pass_idx[p] : 0xffffffffu;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            s[q] = e[q] = 0;
            if (oi[q] != 0xffffffffu) {
                s[q] = a.start[oi[q]];
                e[q] = a.end[oi[q]];
                key[q] = a.okey[oi[q]];
                day[q] = a.day[oi[q]];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (e[q] <= s[q]) continue;
            uint64_t acc = 0;
            uint32_t took = 0;
            for (uint32_t r0 = s[q]; r0 < e[q]; r0 += 8) {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint32_t row = r0 + k;
                    if (row < e[q] && (!a.forder || a.forder[row] == key[q]) &&
                        (!a.ship || a.ship[row] > a.ship_gt)) {
                        acc += dprice ? uint64_t(int64_t(dprice[row]))
                                      : uint64_t(int64_t(price[row])) *
                                            uint64_t(100 - int64_t(disc[row]));
                        ++took;
                    }
                }
            }
            if (took == 0) continue;  // no row shipped after the date: no group
            matched += took;
            const int64_t rev = int64_t(acc);
            rl = rev < rl ? rev : rl;
            rh = rev > rh ? rev : rh;
            dl = day[q] < dl ? day[q] : dl;
            dh = day[q] > dh ? day[q] : dh;
            kl = key[q] < kl ? key[q] : kl;
            kh = key[q] > kh ? key[q] : kh;
            s_rows[atomicAdd(&s_n, 1u)] = RowT{rev, day[q], key[q]};
        }
        // Every thread reads the fill level before any appends again, so
        // that all of them take the same decision.
        __syncthreads();
        const uint32_t cnt = s_n;
        __syncthreads();
        if (cnt > uint32_t(kRangeBuf - kT * 4))
            flush_stage<kT>(s_rows, &s_n, &s_base, cnt, out, n_out);
    }
    __syncthreads();
    const uint32_t cnt = s_n;
    __syncthreads();
    flush_stage<kT>(s_rows, &s_n, &s_base, cnt, out, n_out);
    block_fold<kT>(st, sMatch, int64_t(wave_sum(matched)), 0);
    fold_minmax<kT>(dl, dh, st, sDMin, sDMax);
    fold_minmax<kT>(kl, kh, st, sOMin, sOMax);
    fold_minmax<kT>(rl, rh, st, sRMin, sRMax);
}

// Each workgroup compacts one contiguous chunk of slots (a multiple of
// kCTile): pass 1 counts the hit slots, one atomic reserves the output
// range, pass 2 writes the rows in slot order.
__global__ __launch_bounds__(kCT) void compact_groups_k(GTable t, int64_t chunk,
                                                        RowT* __restrict__ out,
                                                        unsigned long long* __restrict__ n_out,
                                                        int64_t* st) {
    constexpr int kW = kCT / kWave;
    __shared__ uint32_t s_w[kW];
    __shared__ unsigned long long s_base;
    const int64_t cap = int64_t(1) << t.cap_log2;
    const int64_t b0 = int64_t(blockIdx.x) * chunk;
    const int64_t b1 = b0 + chunk < cap ? b0 + chunk : cap;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    uint32_t mine = 0;
    for (int64_t tb = b0; tb < b1; tb += kCTile)
        mine += __popc(hit_mask8(t.hit, tb + int64_t(threadIdx.x) * kCompactR, b1));
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
    int64_t rl = INT64_MAX, rh = INT64_MIN;
    for (int64_t tb = b0; tb < b1; tb += kCTile) {
        const int64_t i0 = tb + int64_t(threadIdx.x) * kCompactR;
        uint32_t hits = hit_mask8(t.hit, i0, b1);
        const uint32_t c = __popc(hits);
        uint32_t incl = c;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, kWave);
            if (lane >= off) incl += y;
        }
        __syncthreads();
        if (lane == kWave - 1) s_w[wave] = incl;
        __syncthreads();
        uint32_t before = 0, tile = 0;
        for (int w = 0; w < kW; ++w) {
            before += w < wave ? s_w[w] : 0;
            tile += s_w[w];
        }
        unsigned long long at = base + before + (incl - c);
        while (hits) {
            const int j = __ffs(hits) - 1;
            hits &= hits - 1;
            const int64_t i = i0 + j;
            const int64_t rv = int64_t(t.rev[i]);
            rl = rv < rl ? rv : rl;
            rh = rv > rh ? rv : rh;
            out[at++] = RowT{rv, t.days[i], t.direct ? int32_t(t.kbase + i) : t.keys[i]};
        }
        base += tile;
    }
    fold_minmax<kCT>(rl, rh, st, sRMin, sRMax);
}

int bits_for(uint64_t span) {  // bits to hold values 0..span
    int b = 0;
    while (b < 64 && (span >> b) != 0) ++b;
    return b;
}

struct Pack {
    int64_t rmax, dmin, kmin;
    int kb, db;
};

// Ascending order of the packed key = revenue desc, day asc, key asc.
__global__ __launch_bounds__(kT) void pack_rows_k(const RowT* __restrict__ rows, int64_t n,
                                                  Pack p, uint64_t* __restrict__ keys,
                                                  uint32_t* __restrict__ idx) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const RowT x = rows[i];
        const int hs = p.db + p.kb;  // fields wider than 63 bits hold only zeros
        const uint64_t rf = uint64_t(p.rmax) - uint64_t(x.rev);
        const uint64_t df = uint64_t(int64_t(x.day) - p.dmin);
        keys[i] = (hs >= 64 ? 0 : rf << hs) | (p.kb >= 64 ? 0 : df << p.kb) |
                  uint64_t(int64_t(x.key) - p.kmin);
        idx[i] = uint32_t(i);
    }
}

// Field keys for the fallback ordering (one stable pass per field).
__global__ __launch_bounds__(kT) void field_keys_k(const RowT* __restrict__ rows,
                                                   const uint32_t* __restrict__ idx, int64_t n,
                                                   int field, uint64_t* __restrict__ keys) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const RowT x = rows[idx[i]];
        uint64_t k;
        if (field == 0) k = uint64_t(uint32_t(x.key) ^ 0x80000000u);
        else if (field == 1) k = uint64_t(uint32_t(x.day) ^ 0x80000000u);
        else k = ~(uint64_t(x.rev) ^ 0x8000000000000000ull);  // descending
        keys[i] = k;
    }
}

__global__ __launch_bounds__(kT) void iota_k(uint32_t* idx, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        idx[i] = uint32_t(i);
}

__global__ __launch_bounds__(kT) void gather_rows_k(const RowT* __restrict__ rows,
                                                    const uint32_t* __restrict__ idx, int64_t n,
                                                    bolap_group_row* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const RowT x = rows[idx[i]];
        out[i] = bolap_group_row{x.key, x.day, x.rev};
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

struct ChainRun {
    int64_t st[sN];
    int64_t n_groups;
    int cap_log2;
    size_t o_cnt;   // workspace offset of the device row count
    bool synced;    // st and n_groups were read back
    Ranges r;
};

bool host_is_min(int i) { return i == sCMin || i == sOMin || i == sDMin || i == sRMin; }
bool host_is_max(int i) { return i == sCMax || i == sOMax || i == sDMax || i == sRMax; }

void fold_stats(const int64_t* rep, int64_t* dst);

// Copies the statistics back and folds the replicas.
int read_stats(bolap_ctx* ctx, int64_t* dst) {
    const size_t bytes = kStatWords * 8;
    BOLAP_TRY(pinned_reserve(ctx, bytes));
    BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, ctx->result_dev, bytes, hipMemcpyDeviceToHost,
                                 ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    fold_stats(static_cast<const int64_t*>(ctx->pinned), dst);
    return BOLAP_OK;
}

// Folds the replicas of every statistic (sum, or min / max).
void fold_stats(const int64_t* rep, int64_t* dst) {
    for (int i = 0; i < sN; ++i) {
        int64_t v = rep[stat_slot(i, 0)];
        for (int r = 1; r < kRep; ++r) {
            const int64_t x = rep[stat_slot(i, r)];
            if (host_is_min(i)) v = x < v ? x : v;
            else if (host_is_max(i)) v = x > v ? x : v;
            else v = int64_t(uint64_t(v) + uint64_t(x));
        }
        dst[i] = v;
    }
}

// The fact probe for the column types of the call.
int launch_probe(bolap_ctx* ctx, const bolap_chain_args* a, const Ranges& r, const GTable& t,
                 int64_t* st) {
    const unsigned g = grid_for(ctx, a->fact_rows / 4 + 1, 8);
    const bool p64 = a->price_dtype == BOLAP_I64, d64 = a->disc_dtype == BOLAP_I64;
    const int32_t* dp = a->fact_disc_price;
    const dim3 b(kT);
    if (p64 && d64)
        hipLaunchKernelGGL((fact_probe_k<int64_t, int64_t>), dim3(g), b, 0, ctx->stream,
                           a->fact_order, static_cast<const int64_t*>(a->fact_price),
                           static_cast<const int64_t*>(a->fact_disc), dp, a->fact_ship,
                           a->ship_gt, a->fact_rows, r, t, st);
    else if (p64)
        hipLaunchKernelGGL((fact_probe_k<int64_t, uint8_t>), dim3(g), b, 0, ctx->stream,
                           a->fact_order, static_cast<const int64_t*>(a->fact_price),
                           static_cast<const uint8_t*>(a->fact_disc), dp, a->fact_ship,
                           a->ship_gt, a->fact_rows, r, t, st);
    else if (d64)
        hipLaunchKernelGGL((fact_probe_k<int32_t, int64_t>), dim3(g), b, 0, ctx->stream,
                           a->fact_order, static_cast<const int32_t*>(a->fact_price),
                           static_cast<const int64_t*>(a->fact_disc), dp, a->fact_ship,
                           a->ship_gt, a->fact_rows, r, t, st);
    else
        hipLaunchKernelGGL((fact_probe_k<int32_t, uint8_t>), dim3(g), b, 0, ctx->stream,
                           a->fact_order, static_cast<const int32_t*>(a->fact_price),
                           static_cast<const uint8_t*>(a->fact_disc), dp, a->fact_ship,
                           a->ship_gt, a->fact_rows, r, t, st);
    return launch_check("fact_probe_k");
}

// Reads the statistics and the device row count back (one copy, one sync),
// and checks for keys outside caller-given ranges.
int finish_groups(bolap_ctx* ctx, const bolap_chain_args* a, const unsigned long long* n_rows,
                  ChainRun& run, bool given_key_ranges) {
    const size_t sbytes = kStatWords * 8;
    BOLAP_TRY(pinned_reserve(ctx, sbytes + 8));
    char* pin = static_cast<char*>(ctx->pinned);
    BOLAP_HIP_TRY(hipMemcpyAsync(pin, ctx->result_dev, sbytes, hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipMemcpyAsync(pin + sbytes, n_rows, 8, hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    fold_stats(reinterpret_cast<const int64_t*>(pin), run.st);
    std::memcpy(&run.n_groups, pin + sbytes, 8);
    run.synced = true;
    if (given_key_ranges && a->key_ranges_known) {  // no min/max pass ran: the ordering packs keys by these
        run.st[sCMin] = run.r.cmin;
        run.st[sCMax] = run.r.cmax;
        run.st[sOMin] = run.r.omin;
        run.st[sOMax] = run.r.omax;
    }
    BOLAP_ARG_CHECK(run.st[sBad] == 0,
                    "bolap_chain: %lld customer or order keys outside the key ranges passed in "
                    "the args", (long long)run.st[sBad]);
    return BOLAP_OK;
}

// Dense order keys: the group of a passing order is the rank of its key's
// bit, so the group arrays hold one entry per passing order (no table, no
// probing, no per-key clearing).  With `defer` the statistics and the row
// count stay on the device (the caller reads them with its results).
int chain_groups_direct(bolap_ctx* ctx, const bolap_chain_args* a, const Ranges& r,
                        int64_t cspan, int64_t ospan, size_t o_cb, WsArena& ws,
                        size_t* o_rows_out, ChainRun& run, int32_t fused_k) {
    int64_t* st = static_cast<int64_t*>(ctx->result_dev);
    const int64_t owords = (ospan + 31) / 32 + 1;
    const int64_t tiles = (owords + kRankWords - 1) / kRankWords;
    const int64_t maxg = a->order_rows > 0 ? a->order_rows : 1;
    // build workgroups over contiguous order chunks (whole tiles of kT * 8)
    const int64_t want = int64_t(ctx->num_cus) * 8;
    int64_t chunk = (maxg + want - 1) / want;
    chunk = (chunk + kT * 8 - 1) / (kT * 8) * (kT * 8);
    const int32_t regions = int32_t((maxg + chunk - 1) / chunk);
    const int32_t cgroups = int32_t(ctx->num_cus) * 2;
    const size_t keep = ws.off;
    const size_t o_ob = ws.plan(size_t(owords) * 4);
    const size_t o_wr = ws.plan(size_t(owords) * 4);
    const size_t o_ts = ws.plan(size_t(tiles) * 4);
    const size_t o_to = ws.plan(size_t(tiles) * 4);
    const size_t o_pl = ws.plan(size_t(maxg) * sizeof(RowK));
    const size_t o_ln = ws.plan(size_t(regions) * 4);
    const size_t o_rev = ws.plan(size_t(maxg) * 8);
    const size_t o_hit = ws.plan(size_t(maxg));
    const size_t o_rows = ws.plan(size_t(maxg) * sizeof(RowT));
    const size_t o_cnt = ws.plan(64);  // rows, -, ranked orders
    BOLAP_TRY(ws_reserve(ctx, ws.off, keep));
    ws.base = static_cast<char*>(ctx->ws);
    auto* cnt = ws.at<unsigned long long>(o_cnt);
    GTable t{nullptr, nullptr, ws.at<uint64_t>(o_rev), ws.at<uint8_t>(o_hit),
             ws.at<uint32_t>(o_ob), 0, 0, 1, r.omin, ws.at<uint32_t>(o_wr),
             ws.at<uint32_t>(o_to), ws.at<RowK>(o_pl), ws.at<uint32_t>(o_ln), chunk, regions};
    {
        TraceScope ts(ctx, "chain_init");
        hipLaunchKernelGGL(fill_u32_k, dim3(grid_for(ctx, owords)), dim3(kT), 0, ctx->stream,
                           ws.at<uint32_t>(o_ob), owords, 0u);
        hipLaunchKernelGGL(fill_u32_k, dim3(1), dim3(kT), 0, ctx->stream,
                           reinterpret_cast<uint32_t*>(cnt), 16, 0u);
        hipLaunchKernelGGL(fill_u32_k, dim3(grid_for(ctx, regions)), dim3(kT), 0, ctx->stream,
                           t.list_n, int64_t(regions), 0u);
    }
    if (a->order_rows > 0) {
        TraceScope ts(ctx, "chain_build");
        hipLaunchKernelGGL(order_mark_k, dim3(unsigned(regions)), dim3(kT), 0, ctx->stream,
                           a->order_key, a->order_day, a->order_cust, a->order_rows, a->day_lt,
                           ws.at<uint32_t>(o_cb), cspan, r, t, st);
        BOLAP_TRY(launch_check("order_mark_k"));
    }
    {
        TraceScope ts(ctx, "chain_rank");
        hipLaunchKernelGGL(rank_tiles_k, dim3(unsigned(tiles)), dim3(kT), 0, ctx->stream,
                           ws.at<uint32_t>(o_ob), owords, ws.at<uint32_t>(o_wr),
                           ws.at<uint32_t>(o_ts), ws.at<uint32_t>(o_to), tiles,
                           ctx->counters + 2, cnt + 2);
        hipLaunchKernelGGL(zero_ranked_k, dim3(grid_for(ctx, maxg)), dim3(kT), 0, ctx->stream,
                           t.rev, t.hit, cnt + 2);
        BOLAP_TRY(launch_check("zero_ranked_k"));
    }
    if (a->fact_rows > 0 && ospan > 0) {
        TraceScope ts(ctx, "chain_probe");
        BOLAP_TRY(launch_probe(ctx, a, r, t, st));
    }
    {
        TraceScope ts(ctx, "chain_compact");
        hipLaunchKernelGGL(compact_list_k, dim3(unsigned(cgroups)), dim3(kT), 0, ctx->stream, t,
                           r, ws.at<RowT>(o_rows), cnt, st);
        BOLAP_TRY(launch_check("compact_list_k"));
    }
    run.cap_log2 = -1;
    run.o_cnt = o_cnt;
    run.r = r;
    run.synced = false;
    *o_rows_out = o_rows;
    if (fused_k >= 0) return BOLAP_OK;
    return finish_groups(ctx, a, cnt, run, true);
}

// Runs the join + group-by; the hit groups end up as RowT rows at workspace
// offset *o_rows and the statistics in `run`.
int chain_groups(bolap_ctx* ctx, const bolap_chain_args* a, int64_t n_expect, bool exact,
                 WsArena& ws, size_t* o_rows_out, ChainRun& run, int32_t fused_k) {
    int64_t* st = static_cast<int64_t*>(ctx->result_dev);
    hipLaunchKernelGGL(stats_init_k, dim3(1), dim3(kStatWords), 0, ctx->stream, st);
    Ranges r;
    if (a->key_ranges_known) {
        r = Ranges{a->cust_key_min, a->cust_key_max, a->order_key_min, a->order_key_max};
    } else {
        {
            TraceScope ts(ctx, "chain_stats");
            hipLaunchKernelGGL(key_minmax_k, dim3(grid_for(ctx, a->order_rows)), dim3(kT), 0,
                               ctx->stream, a->cust_key, a->cust_rows, a->order_key,
                               a->order_rows, st);
        }
        BOLAP_TRY(launch_check("key_minmax_k"));
        BOLAP_TRY(read_stats(ctx, run.st));
        r = Ranges{run.st[sCMin], run.st[sCMax], run.st[sOMin], run.st[sOMax]};
    }
    const int64_t cspan = a->cust_rows > 0 ? r.cmax - r.cmin + 1 : 0;
    const int64_t ospan = a->order_rows > 0 ? r.omax - r.omin + 1 : 0;
    const int64_t cwords = (cspan + 31) / 32 + 1;
    const int64_t owords = (ospan + 31) / 32 + 1;

    // Customer bitmap first; the group table is planned once its size is known.
    const size_t o_cb = ws.plan(size_t(cwords) * 4);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    hipLaunchKernelGGL(fill_u32_k, dim3(grid_for(ctx, cwords)), dim3(kT), 0, ctx->stream,
                       ws.at<uint32_t>(o_cb), cwords, 0u);
    if (a->cust_rows > 0) {
        TraceScope ts(ctx, "chain_cust");
        hipLaunchKernelGGL(cust_bits_k, dim3(grid_for(ctx, a->cust_rows)), dim3(kT), 0,
                           ctx->stream, a->cust_key, a->cust_segment, a->cust_rows, a->segment,
                           r, ws.at<uint32_t>(o_cb), st);
    }
    BOLAP_TRY(launch_check("cust_bits_k"));
    // Dense order keys: a directly addressed table over the key span, no
    // sizing count, no probing.
    const bool direct = ospan > 0 && ospan <= (int64_t(1) << 27) &&
                        ospan <= 8 * a->order_rows + 1024 && ctx->knob("chain.direct", 1) == 1;
    if (exact && !direct) {
        if (a->order_rows > 0) {
            TraceScope ts(ctx, "chain_count");
            hipLaunchKernelGGL(order_count_k, dim3(grid_for(ctx, a->order_rows)), dim3(kT), 0,
                               ctx->stream, a->order_day, a->order_cust, a->order_rows,
                               a->day_lt, ws.at<uint32_t>(o_cb), cspan, r, st);
        }
        BOLAP_TRY(launch_check("order_count_k"));
        BOLAP_TRY(read_stats(ctx, run.st));
        n_expect = run.st[sCount];
    }
    if (direct)
        return chain_groups_direct(ctx, a, r, cspan, ospan, o_cb, ws, o_rows_out, run, fused_k);
    const int cap_log2 = ceil_log2(n_expect * 2 < 1024 ? 1024 : n_expect * 2);
    BOLAP_ARG_CHECK(cap_log2 <= 31, "bolap_chain: %lld groups exceed the table limit",
                    (long long)n_expect);
    const int64_t cap = int64_t(1) << cap_log2;
    const int64_t rows_cap = cap;
    const size_t keep = ws.off;
    const size_t o_keys = ws.plan(size_t(cap) * 4);
    const size_t o_days = ws.plan(size_t(cap) * 4);
    const size_t o_rev = ws.plan(size_t(cap) * 8);
    const size_t o_hit = ws.plan(size_t(cap));
    const size_t o_ob = ws.plan(size_t(owords) * 4);
    const size_t o_rows = ws.plan(size_t(rows_cap) * sizeof(RowT));
    const size_t o_cnt = ws.plan(64);
    BOLAP_TRY(ws_reserve(ctx, ws.off, keep));
    ws.base = static_cast<char*>(ctx->ws);
    const uint32_t* cbits = ws.at<uint32_t>(o_cb);
    // Exact tables (cap >= 2 x count) always place every key; estimate-sized
    // ones give up after 128 slots and fall back to exact sizing.
    const uint32_t max_probe = exact ? uint32_t(cap) : 128u;
    GTable t{ws.at<int32_t>(o_keys), ws.at<int32_t>(o_days), ws.at<uint64_t>(o_rev),
             ws.at<uint8_t>(o_hit), ws.at<uint32_t>(o_ob), cap_log2, max_probe, 0, r.omin,
             nullptr, nullptr, nullptr, nullptr};
    {
        TraceScope ts(ctx, "chain_init");
        hipLaunchKernelGGL(gtable_init_k, dim3(grid_for(ctx, cap, 8)), dim3(kT), 0, ctx->stream,
                           t, owords);
    }
    BOLAP_TRY(launch_check("gtable_init_k"));
    if (a->order_rows > 0) {
        TraceScope ts(ctx, "chain_build");
        hipLaunchKernelGGL(order_build_k, dim3(grid_for(ctx, a->order_rows)), dim3(kT), 0,
                           ctx->stream, a->order_key, a->order_day, a->order_cust, a->order_rows,
                           a->day_lt, cbits, cspan, r, t, st);
    }
    BOLAP_TRY(launch_check("order_build_k"));
    if (a->fact_rows > 0 && ospan > 0) {
        TraceScope ts(ctx, "chain_probe");
        BOLAP_TRY(launch_probe(ctx, a, r, t, st));
    }
    auto* n_rows = ws.at<unsigned long long>(o_cnt);
    hipLaunchKernelGGL(fill_u32_k, dim3(1), dim3(kT), 0, ctx->stream,
                       reinterpret_cast<uint32_t*>(n_rows), 2, 0u);
    {
        TraceScope ts(ctx, "chain_compact");
        int64_t cblocks = int64_t(ctx->num_cus);
        int64_t chunk = (cap + cblocks - 1) / cblocks;
        chunk = (chunk + kCTile - 1) / kCTile * kCTile;
        cblocks = (cap + chunk - 1) / chunk;
        hipLaunchKernelGGL(compact_groups_k, dim3(unsigned(cblocks)), dim3(kCT), 0, ctx->stream,
                           t, chunk, ws.at<RowT>(o_rows), n_rows, st);
    }
    BOLAP_TRY(launch_check("compact_groups_k"));
    BOLAP_TRY(read_stats(ctx, run.st));
    BOLAP_ARG_CHECK(run.st[sBad] == 0,
                    "bolap_chain: %lld customer or order keys outside the key ranges passed in "
                    "the args", (long long)run.st[sBad]);
    if (a->key_ranges_known) {  // no min/max pass ran: the ordering packs keys by these
        run.st[sCMin] = r.cmin;
        run.st[sCMax] = r.cmax;
        run.st[sOMin] = r.omin;
        run.st[sOMax] = r.omax;
    }
    if (run.st[sGroups] > cap / 2) run.st[sFlags] |= 1;  // load factor above 1/2
    BOLAP_TRY(pinned_reserve(ctx, 8));
    BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, n_rows, 8, hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::memcpy(&run.n_groups, ctx->pinned, 8);
    run.cap_log2 = cap_log2;
    run.o_cnt = o_cnt;
    run.synced = true;
    run.r = r;
    *o_rows_out = o_rows;
    return BOLAP_OK;
}

// The group-by when each order's fact rows are a given contiguous range:
// customer bitmap, then one pass over the orders that sums the ranges of
// the passing ones.  Rows at workspace offset *o_rows_out.
int chain_groups_ranges(bolap_ctx* ctx, const bolap_chain_args* a, WsArena& ws,
                        size_t* o_rows_out, ChainRun& run, bool defer) {
    int64_t* st = static_cast<int64_t*>(ctx->result_dev);
    Ranges r{a->cust_key_min, a->cust_key_max, 0, 0};
    if (!a->key_ranges_known) {
        hipLaunchKernelGGL(stats_init_k, dim3(1), dim3(kStatWords), 0, ctx->stream, st);
        {
            TraceScope ts(ctx, "chain_stats");
            hipLaunchKernelGGL(key_minmax_k, dim3(grid_for(ctx, a->cust_rows / 4 + 1)), dim3(kT),
                               0, ctx->stream, a->cust_key, a->cust_rows, nullptr, int64_t(0), st);
        }
        BOLAP_TRY(launch_check("key_minmax_k"));
        int64_t cst[sN];
        BOLAP_TRY(read_stats(ctx, cst));
        r = Ranges{cst[sCMin], cst[sCMax], 0, 0};
    }
    const int64_t cspan = a->cust_rows > 0 ? r.cmax - r.cmin + 1 : 0;
    const int64_t cwords = (cspan + 31) / 32 + 1;
    // Rows and passing orders: at most one per order.  Filter chunks are
    // whole tiles of kPassBuf orders.
    const int64_t blocks_max = int64_t(ctx->num_cus) * 4;
    int64_t chunk = (a->order_rows + blocks_max - 1) / blocks_max;
    chunk = (chunk + kPassBuf - 1) / kPassBuf * kPassBuf;
    const int64_t blocks = a->order_rows > 0 ? (a->order_rows + chunk - 1) / chunk : 0;
    const size_t o_cb = ws.plan(size_t(cwords) * 4);
    const size_t o_rows = ws.plan(size_t(a->order_rows > 0 ? a->order_rows : 1) * sizeof(RowT));
    const size_t o_pass = ws.plan(size_t(a->order_rows > 0 ? a->order_rows : 1) * 4);
    const size_t o_cnt = ws.plan(64);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    hipLaunchKernelGGL(stats_init_k, dim3(1), dim3(kStatWords), 0, ctx->stream, st);
    hipLaunchKernelGGL(fill_u32_k, dim3(grid_for(ctx, cwords)), dim3(kT), 0, ctx->stream,
                       ws.at<uint32_t>(o_cb), cwords, 0u);
    auto* n_rows = ws.at<unsigned long long>(o_cnt);
    auto* n_pass = n_rows + 1;
    hipLaunchKernelGGL(fill_u32_k, dim3(1), dim3(kT), 0, ctx->stream,
                       reinterpret_cast<uint32_t*>(n_rows), 4, 0u);
    if (a->cust_rows > 0) {
        TraceScope ts(ctx, "chain_cust");
        hipLaunchKernelGGL(cust_bits_k, dim3(grid_for(ctx, a->cust_rows)), dim3(kT), 0,
                           ctx->stream, a->cust_key, a->cust_segment, a->cust_rows, a->segment,
                           r, ws.at<uint32_t>(o_cb), st);
    }
    BOLAP_TRY(launch_check("cust_bits_k"));
    if (blocks > 0) {
        const RangeArgs ra{a->order_key, a->order_day, a->order_cust, a->order_fact_start,
                           a->order_fact_end, a->order_rows, a->day_lt, chunk,
                           a->fact_ship, a->ship_gt,
                           a->fact_order_sorted ? nullptr : a->fact_order};
        const uint32_t* cb = ws.at<uint32_t>(o_cb);
        uint32_t* pass = ws.at<uint32_t>(o_pass);
        {
            TraceScope ts(ctx, "chain_range_filter");
            hipLaunchKernelGGL(range_filter_k, dim3(unsigned(blocks)), dim3(kT), 0, ctx->stream,
                               ra, cb, cspan, r, pass, n_pass, st);
            BOLAP_TRY(launch_check("range_filter_k"));
        }
        TraceScope ts(ctx, "chain_ranges");
        RowT* rows = ws.at<RowT>(o_rows);
        const bool p64 = a->price_dtype == BOLAP_I64, d64 = a->disc_dtype == BOLAP_I64;
        const int32_t* dp = a->fact_disc_price;
        const dim3 g = dim3(unsigned(blocks_max)), b = dim3(kT);
        if (p64 && d64)
            hipLaunchKernelGGL((range_groups_k<int64_t, int64_t>), g, b, 0, ctx->stream, ra,
                               static_cast<const int64_t*>(a->fact_price),
                               static_cast<const int64_t*>(a->fact_disc), dp, pass, n_pass, rows,
                               n_rows, st);
        else if (p64)
            hipLaunchKernelGGL((range_groups_k<int64_t, uint8_t>), g, b, 0, ctx->stream, ra,
                               static_cast<const int64_t*>(a->fact_price),
                               static_cast<const uint8_t*>(a->fact_disc), dp, pass, n_pass, rows,
                               n_rows, st);
        else if (d64)
            hipLaunchKernelGGL((range_groups_k<int32_t, int64_t>), g, b, 0, ctx->stream, ra,
                               static_cast<const int32_t*>(a->fact_price),
                               static_cast<const int64_t*>(a->fact_disc), dp, pass, n_pass, rows,
                               n_rows, st);
        else
            hipLaunchKernelGGL((range_groups_k<int32_t, uint8_t>), g, b, 0, ctx->stream, ra,
                               static_cast<const int32_t*>(a->fact_price),
                               static_cast<const uint8_t*>(a->fact_disc), dp, pass, n_pass, rows,
                               n_rows, st);
    }
    BOLAP_TRY(launch_check("range_groups_k"));
    run.cap_log2 = -1;
    run.o_cnt = o_cnt;
    run.r = r;
    run.synced = false;
    *o_rows_out = o_rows;
    if (defer) return BOLAP_OK;
    BOLAP_TRY(finish_groups(ctx, a, n_rows, run, false));
    run.st[sGroups] = run.n_groups;
    return BOLAP_OK;
}

// Orders the n rows at workspace offset o_rows and writes the first `limit`
// as bolap_group_row at workspace offset *o_out.
int order_rows(bolap_ctx* ctx, WsArena& ws, size_t o_rows, int64_t n, int64_t limit,
               const ChainRun& run, size_t* o_out, int* passes) {
    *passes = 0;
    const int64_t m = std::min(limit, n);
    const size_t keep = ws.off;
    const size_t o_k = ws.plan(size_t(n > 0 ? n : 1) * 8);
    const size_t o_i = ws.plan(size_t(n > 0 ? n : 1) * 4);
    const size_t o_ko = ws.plan(size_t(m > 0 ? m : 1) * 8);
    const size_t o_io = ws.plan(size_t(m > 0 ? m : 1) * 4);
    *o_out = ws.plan(size_t(m > 0 ? m : 1) * sizeof(bolap_group_row));
    const size_t o_scr =
        ws.plan(std::max(radix_sort_scratch_bytes(ctx, n), topk_scratch_bytes(n)));
    BOLAP_TRY(ws_reserve(ctx, ws.off, keep));
    ws.base = static_cast<char*>(ctx->ws);
    if (m <= 0) return BOLAP_OK;
    const RowT* rows = ws.at<RowT>(o_rows);
    uint64_t* keys = ws.at<uint64_t>(o_k);
    uint32_t* idx = ws.at<uint32_t>(o_i);
    const int64_t rmin = run.st[sRMin], rmax = run.st[sRMax];
    const int64_t dmin = run.st[sDMin], dmax = run.st[sDMax];
    const int64_t kmin = run.st[sOMin], kmax = run.st[sOMax];
    const int rb = bits_for(uint64_t(rmax) - uint64_t(rmin));
    const int db = bits_for(uint64_t(dmax) - uint64_t(dmin));
    const int kb = bits_for(uint64_t(kmax) - uint64_t(kmin));
    // Revenue spans beyond int64 (wrapped sums) take the fallback.
    const bool rev_ok = rmax >= rmin && (rmax - rmin) >= 0;
    TraceScope ts(ctx, "chain_order");
    const unsigned g = grid_for(ctx, n);
    const uint32_t* sel = nullptr;
    if (rev_ok && rb + db + kb <= 64) {
        Pack p{rmax, dmin, kmin, kb, db};
        hipLaunchKernelGGL(pack_rows_k, dim3(g), dim3(kT), 0, ctx->stream, rows, n, p, keys, idx);
        BOLAP_TRY(launch_check("pack_rows_k"));
        if (m < n && m <= 4096) {
            BOLAP_TRY(topk_pairs(ctx, keys, idx, n, int32_t(m), ws.at<uint64_t>(o_ko),
                                 ws.at<uint32_t>(o_io), ws.at<void>(o_scr)));
            sel = ws.at<uint32_t>(o_io);
        } else {
            BOLAP_TRY(radix_sort_pairs(ctx, keys, idx, n, rb + db + kb, ws.at<void>(o_scr),
                                       passes));
            sel = idx;
        }
    } else {
        // Three stable passes: key, then day, then revenue descending.
        hipLaunchKernelGGL(iota_k, dim3(g), dim3(kT), 0, ctx->stream, idx, n);
        int total = 0;
        for (int field = 0; field < 3; ++field) {
            hipLaunchKernelGGL(field_keys_k, dim3(g), dim3(kT), 0, ctx->stream, rows, idx, n,
                               field, keys);
            BOLAP_TRY(launch_check("field_keys_k"));
            int p = 0;
            BOLAP_TRY(radix_sort_pairs(ctx, keys, idx, n, field == 2 ? 64 : 32,
                                       ws.at<void>(o_scr), &p));
            total += p;
        }
        *passes = total;
        sel = idx;
    }
    hipLaunchKernelGGL(gather_rows_k, dim3(grid_for(ctx, m)), dim3(kT), 0, ctx->stream, rows, sel,
                       m, ws.at<bolap_group_row>(*o_out));
    return launch_check("gather_rows_k");
}

int chain_validate(const bolap_chain_args* a) {
    BOLAP_ARG_CHECK(a, "bolap_chain: null args");
    BOLAP_ARG_CHECK(a->cust_rows >= 0 && a->order_rows >= 0 && a->fact_rows >= 0,
                    "bolap_chain: negative rows");
    BOLAP_ARG_CHECK(a->order_rows < (int64_t(1) << 31), "bolap_chain: order_rows >= 2^31");
    BOLAP_ARG_CHECK(a->cust_rows == 0 || (a->cust_key && a->cust_segment),
                    "bolap_chain: null customer column");
    BOLAP_ARG_CHECK(a->order_rows == 0 || (a->order_key && a->order_day && a->order_cust),
                    "bolap_chain: null order column");
    BOLAP_ARG_CHECK(a->fact_rows == 0 || ((a->fact_order || a->order_fact_start) &&
                                          (a->fact_disc_price || (a->fact_price && a->fact_disc))),
                    "bolap_chain: null fact column");
    BOLAP_ARG_CHECK((a->order_fact_start == nullptr) == (a->order_fact_end == nullptr),
                    "bolap_chain: order_fact_start and order_fact_end go together");
    BOLAP_ARG_CHECK(!a->key_ranges_known || (a->cust_key_min <= a->cust_key_max &&
                                             a->order_key_min <= a->order_key_max),
                    "bolap_chain: key ranges with min > max");
    BOLAP_ARG_CHECK(a->order_fact_start || a->fact_rows == 0 || aligned(a->fact_order, 16),
                    "bolap_chain: fact_order must be 16-byte aligned");
    BOLAP_ARG_CHECK(!a->order_fact_start || a->fact_rows < (int64_t(1) << 32),
                    "bolap_chain: fact ranges are 32-bit row numbers");
    BOLAP_ARG_CHECK(!a->order_fact_start || a->fact_order_sorted || a->fact_rows == 0 ||
                        a->fact_order,
                    "bolap_chain: order ranges over unsorted fact rows need fact_order");
    BOLAP_ARG_CHECK(a->order_rows == 0 || (aligned(a->order_key, 16) && aligned(a->order_day, 16) &&
                                           aligned(a->order_cust, 16)),
                    "bolap_chain: order columns must be 16-byte aligned");
    BOLAP_ARG_CHECK(a->cust_rows == 0 || (aligned(a->cust_key, 16) && aligned(a->cust_segment, 4)),
                    "bolap_chain: customer columns must be 16/4-byte aligned");
    BOLAP_ARG_CHECK(a->fact_disc_price || a->price_dtype == BOLAP_I64 ||
                        a->price_dtype == BOLAP_I32,
                    "bolap_chain: price dtype must be i64 or i32");
    BOLAP_ARG_CHECK(a->fact_disc_price || a->disc_dtype == BOLAP_I64 || a->disc_dtype == BOLAP_U8,
                    "bolap_chain: disc dtype must be i64 or u8");
    return BOLAP_OK;
}

// The k best (k <= kTopMax) of the rows at o_rows, count on the device at
// o_cnt, as bolap_group_row at *o_out; *o_flag (a device u32) is set when
// the revenue cut held more than kSelCap rows (the caller then orders the
// rows by the radix path).  No host round trip.
int topk_by_cut(bolap_ctx* ctx, WsArena& ws, size_t o_rows, size_t o_cnt, int32_t k,
                size_t* o_out, size_t* o_flag) {
    const size_t keep = ws.off;
    const size_t o_hist = ws.plan(size_t(kCodes) * 4);
    *o_flag = ws.plan(64);  // overflow flag, candidate count
    const size_t o_cand = ws.plan(size_t(kSelCap) * sizeof(RowT));
    *o_out = ws.plan(size_t(kTopMax) * sizeof(bolap_group_row));
    BOLAP_TRY(ws_reserve(ctx, ws.off, keep));
    ws.base = static_cast<char*>(ctx->ws);
    TraceScope ts(ctx, "chain_order");
    auto* flag = ws.at<unsigned>(*o_flag);
    auto* n_cand = ws.at<unsigned long long>(*o_flag + 8);
    hipLaunchKernelGGL(fill_u32_k, dim3(grid_for(ctx, kCodes)), dim3(kT), 0, ctx->stream,
                       ws.at<uint32_t>(o_hist), int64_t(kCodes), 0u);
    hipLaunchKernelGGL(fill_u32_k, dim3(1), dim3(kT), 0, ctx->stream, ws.at<uint32_t>(*o_flag),
                       int64_t(16), 0u);
    const unsigned g = unsigned(ctx->num_cus);
    const auto* rows = ws.at<RowT>(o_rows);
    const auto* n_rows = ws.at<unsigned long long>(o_cnt);
    hipLaunchKernelGGL(topk_hist_k, dim3(g), dim3(kT), 0, ctx->stream, rows, n_rows,
                       ws.at<uint32_t>(o_hist));
    hipLaunchKernelGGL(topk_select_k, dim3(g * 2), dim3(kT), 0, ctx->stream, rows, n_rows,
                       ws.at<uint32_t>(o_hist), k, ws.at<RowT>(o_cand), n_cand);
    hipLaunchKernelGGL(topk_sort_k, dim3(1), dim3(kSortT), 0, ctx->stream, ws.at<RowT>(o_cand),
                       n_cand, k, ws.at<bolap_group_row>(*o_out), flag);
    return launch_check("topk_sort_k");
}

// Full pipeline: group-by (estimate-sized with exact fallback), ordering,
// and the first `limit` rows (all when limit < 0) copied to `host_out`.
// A small top-k (limit <= kTopMax) on a path without a fallback keeps the
// statistics and counts on the device until the rows are copied back: one
// host synchronisation per call.
int chain_run(bolap_ctx* ctx, const bolap_chain_args* a, int64_t limit,
              bolap_group_row* host_out, int64_t max_rows, int64_t* n_out,
              bolap_chain_stats* stats) {
    BOLAP_TRY(chain_validate(a));
    ChainRun run{};
    WsArena ws;
    size_t o_rows = 0;
    int fallbacks = 0;
    const bool ranged = a->order_fact_start != nullptr;
    const bool exact = a->capacity_hint <= 0;
    const bool fused = limit >= 1 && limit <= kTopMax && limit <= max_rows &&
                       ctx->knob("chain.fused_topk", 1) == 1;
    const int32_t fused_k = fused ? int32_t(limit) : -1;
    if (ranged)
        BOLAP_TRY(chain_groups_ranges(ctx, a, ws, &o_rows, run, fused));
    else
        BOLAP_TRY(chain_groups(ctx, a, a->capacity_hint, exact, ws, &o_rows, run, fused_k));
    if (run.synced && !ranged && !exact && (run.st[sFlags] != 0)) {
        fallbacks = 1;  // the estimate was too small: exact sizing
        ws = WsArena{};
        BOLAP_TRY(chain_groups(ctx, a, 0, true, ws, &o_rows, run, fused_k));
    }
    if (run.synced && run.st[sFlags] != 0) {
        set_error("bolap_chain: exact-sized group table overflowed");
        return BOLAP_ERR_CAPACITY;
    }
    int passes = 0;
    int64_t m = 0;
    if (fused) {
        size_t o_out = 0, o_flag = 0;
        BOLAP_TRY(topk_by_cut(ctx, ws, o_rows, run.o_cnt, int32_t(limit), &o_out, &o_flag));
        // statistics, row count, overflow flag and the k rows in one copy back
        const size_t sbytes = kStatWords * 8;
        const size_t rbytes = size_t(limit) * sizeof(bolap_group_row);
        BOLAP_TRY(pinned_reserve(ctx, sbytes + 16 + rbytes));
        char* pin = static_cast<char*>(ctx->pinned);
        if (!run.synced) {
            BOLAP_HIP_TRY(hipMemcpyAsync(pin, ctx->result_dev, sbytes, hipMemcpyDeviceToHost,
                                         ctx->stream));
            BOLAP_HIP_TRY(hipMemcpyAsync(pin + sbytes, ws.at<void>(run.o_cnt), 8,
                                         hipMemcpyDeviceToHost, ctx->stream));
        }
        BOLAP_HIP_TRY(hipMemcpyAsync(pin + sbytes + 8, ws.at<void>(o_flag), 8,
                                     hipMemcpyDeviceToHost, ctx->stream));
        BOLAP_HIP_TRY(hipMemcpyAsync(pin + sbytes + 16, ws.at<void>(o_out), rbytes,
                                     hipMemcpyDeviceToHost, ctx->stream));
        BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        uint32_t overflow = 0;
        std::memcpy(&overflow, pin + sbytes + 8, 4);
        if (!run.synced) {
            fold_stats(reinterpret_cast<const int64_t*>(pin), run.st);
            std::memcpy(&run.n_groups, pin + sbytes, 8);
            run.synced = true;
            if (ranged) run.st[sGroups] = run.n_groups;
            BOLAP_ARG_CHECK(run.st[sBad] == 0,
                            "bolap_chain: %lld customer or order keys outside the key ranges "
                            "passed in the args", (long long)run.st[sBad]);
        }
        m = std::min<int64_t>(limit, run.n_groups);
        if (!overflow) {
            if (m > 0)
                std::memcpy(host_out, pin + sbytes + 16, size_t(m) * sizeof(bolap_group_row));
        } else {  // the cut held too many ties: the radix path orders the rows
            size_t o_rad = 0;
            BOLAP_TRY(order_rows(ctx, ws, o_rows, run.n_groups, m, run, &o_rad, &passes));
            if (m > 0) {
                const size_t bytes = size_t(m) * sizeof(bolap_group_row);
                BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, ws.at<bolap_group_row>(o_rad), bytes,
                                             hipMemcpyDeviceToHost, ctx->stream));
                BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
                std::memcpy(host_out, ctx->pinned, bytes);
            }
        }
    } else {
        const int64_t n = run.n_groups;
        m = limit < 0 ? n : std::min<int64_t>(limit, n);
        if (m > max_rows) {
            set_error("bolap_chain: %lld rows exceed max_rows %lld", (long long)m,
                      (long long)max_rows);
            return BOLAP_ERR_CAPACITY;
        }
        size_t o_out = 0;
        BOLAP_TRY(order_rows(ctx, ws, o_rows, n, m, run, &o_out, &passes));
        if (m > 0) {
            const size_t bytes = size_t(m) * sizeof(bolap_group_row);
            BOLAP_TRY(pinned_reserve(ctx, bytes));
            BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, ws.at<bolap_group_row>(o_out), bytes,
                                         hipMemcpyDeviceToHost, ctx->stream));
            BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
            std::memcpy(host_out, ctx->pinned, bytes);
        }
    }
    const int64_t n = run.n_groups;
    *n_out = m;
    trace_count(ctx, "chain_customer_rows_emitted", run.st[sCustPass]);
    trace_count(ctx, "chain_orders_rows_emitted", run.st[sOrderPass]);
    trace_count(ctx, "chain_join_rows_emitted", run.st[sMatch]);
    trace_count(ctx, "chain_agg_groups_created", n);
    if (stats) {
        stats->cust_pass = run.st[sCustPass];
        stats->order_pass = run.st[sOrderPass];
        stats->fact_matches = run.st[sMatch];
        stats->groups = n;
        stats->capacity = run.cap_log2 >= 0 ? int64_t(1) << run.cap_log2 : 0;
        stats->fallbacks = fallbacks;
        stats->sort_passes = passes;
    }
    return BOLAP_OK;
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" {

int bolap_chain_topk(bolap_ctx* ctx, const bolap_chain_args* args, int32_t k,
                     bolap_group_row* out, int32_t* n_out, bolap_chain_stats* stats) {
    BOLAP_ARG_CHECK(ctx && args && n_out && (out || k == 0) && k >= 0,
                    "bolap_chain_topk: bad arguments");
    int64_t n = 0;
    BOLAP_TRY(chain_run(ctx, args, k, out, k, &n, stats));
    *n_out = int32_t(n);
    return BOLAP_OK;
}

int bolap_chain_groups_sorted(bolap_ctx* ctx, const bolap_chain_args* args,
                              bolap_group_row* out, int64_t max_rows, int64_t* n_out,
                              bolap_chain_stats* stats) {
    BOLAP_ARG_CHECK(ctx && args && n_out && max_rows >= 0,
                    "bolap_chain_groups_sorted: bad arguments");
    return chain_run(ctx, args, -1, out, max_rows, n_out, stats);
}

}  // extern "C"
