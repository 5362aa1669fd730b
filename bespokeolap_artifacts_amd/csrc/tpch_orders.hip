// tpch_orders.hip — TPC-H Q4, Q5, Q10 (orders by date -> lineitem ranges),
// Q12 (lineitem filter -> order priority) and Q19 (part mask -> lineitem
// probe) over the reference Database layout.
//
// Orders are sorted by date at build (bespoke_tpch/builder_impl.cpp:597-608),
// so a date window is a contiguous run of order rows.  Every workgroup finds
// that run itself with a workgroup-wide lower_bound (each round probes one
// position per thread, three rounds for 15 M orders), then walks its share
// of the window one order per lane through lineitem_ranges.  Group counters
// with few keys (priorities, nations, ship modes) are summed in LDS and
// added to the device result once per workgroup; Q10's per-customer revenue
// goes to a dense array over the customer keys, compacted in key order and
// sorted by the stable LSD radix sort.
//
// Q12 and Q19 stream the lineitem columns four rows per lane (one 4-byte
// load of a u8 column, 8 of an i16 column, 16 of an i32 column) and gather
// the order / part side only for rows that pass the row predicates.
#include <cstring>

#include "radix.hpp"

namespace bolap {
namespace {

constexpr int kT = 256;

// First index i in [0, n) with v[i] >= x (n when none) of a non-decreasing
// column, computed by the whole workgroup: each round probes kT evenly
// spaced positions of the remaining interval.  Every thread must call it.
BOLAP_DEV int64_t wg_lower_bound(const int16_t* __restrict__ v, int64_t n, int32_t x) {
    int64_t lo = 0, hi = n;  // the answer lies in [lo, hi]
    while (lo < hi) {
        const int64_t step = (hi - lo + kT - 1) / kT;
        const int64_t p = lo + int64_t(threadIdx.x) * step;
        const bool below = p < hi && int32_t(v[p]) < x;
        const int k = __syncthreads_count(below);
        // positions lo + t * step, t < k, hold values below x
        const int64_t nlo = k > 0 ? lo + int64_t(k - 1) * step + 1 : lo;
        const int64_t top = lo + int64_t(k) * step;
        hi = top < hi ? top : hi;
        lo = nlo;
    }
    return lo;
}

struct Window {
    int64_t s, e;  // order rows [s, e)
};

// The order rows of [date_lo, date_hi) and this workgroup's contiguous part.
BOLAP_DEV Window wg_window(const int16_t* __restrict__ date, int64_t n, int32_t lo, int32_t hi) {
    const int64_t s = wg_lower_bound(date, n, lo);
    const int64_t e = hi > lo ? wg_lower_bound(date, n, hi) : s;
    const int64_t len = e > s ? e - s : 0;
    const int64_t per = (len + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = s + per * blockIdx.x;
    const int64_t b1 = b0 + per < e ? b0 + per : e;
    return Window{b0, b1 > b0 ? b1 : b0};
}

// The lineitem side of an order-range query.
struct LiSide {
    const uint32_t* start;
    const uint32_t* end;
    const int32_t* okey;   // null when lineitem is sorted by order key
};

// ---------------------------------------------------------------------------
// Q4: orders with a late lineitem, counted per priority code.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void q4_k(const int16_t* __restrict__ odate, int64_t o_rows,
                                           const int32_t* __restrict__ okey,
                                           const uint8_t* __restrict__ prio, int32_t n_prio,
                                           LiSide li, const uint32_t* __restrict__ cr,
                                           int32_t date_lo, int32_t date_hi,
                                           unsigned long long* __restrict__ counts) {
    __shared__ unsigned s_cnt[256];
    s_cnt[threadIdx.x] = 0;
    const Window w = wg_window(odate, o_rows, date_lo, date_hi);
    __syncthreads();
    for (int64_t o = w.s + threadIdx.x; o < w.e; o += kT) {
        const uint32_t e = li.end[o];
        if (e == 0) continue;
        const int32_t key = li.okey ? okey[o] : 0;
        bool late = false;
        for (uint32_t r = li.start[o]; r < e && !late; ++r) {
            if (li.okey && li.okey[r] != key) continue;
            const uint32_t p = cr[r];
            late = uint16_t(p) < uint16_t(p >> 16);  // q4_is_late, query_q4.cpp:70-72
        }
        if (late) {
            const uint32_t c = prio[o];
            if (c < uint32_t(n_prio)) atomicAdd(&s_cnt[c], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < n_prio && s_cnt[threadIdx.x])
        atomicAdd(counts + threadIdx.x, (unsigned long long)s_cnt[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// Q5: revenue of lineitems whose supplier nation is the customer's, per nation.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void q5_k(const int16_t* __restrict__ odate, int64_t o_rows,
                                           const int32_t* __restrict__ okey,
                                           const int16_t* __restrict__ cnation, LiSide li,
                                           const int16_t* __restrict__ snation,
                                           const int32_t* __restrict__ dprice,
                                           const uint8_t* __restrict__ in_region,
                                           int32_t n_nation, int32_t date_lo, int32_t date_hi,
                                           unsigned long long* __restrict__ revenue) {
    __shared__ unsigned long long s_rev[256];
    __shared__ uint8_t s_in[256];
    s_rev[threadIdx.x] = 0;
    s_in[threadIdx.x] = threadIdx.x < n_nation ? in_region[threadIdx.x] : 0;
    const Window w = wg_window(odate, o_rows, date_lo, date_hi);
    __syncthreads();
    for (int64_t o = w.s + threadIdx.x; o < w.e; o += kT) {
        const int32_t cn = cnation[o];
        if (cn < 0 || cn >= n_nation || !s_in[cn]) continue;
        const uint32_t e = li.end[o];
        if (e == 0) continue;
        const int32_t key = li.okey ? okey[o] : 0;
        int64_t local = 0;
        for (uint32_t r = li.start[o]; r < e; ++r) {
            if (li.okey && li.okey[r] != key) continue;
            if (snation[r] == cn) local += dprice[r];
        }
        if (local) atomicAdd(&s_rev[cn], (unsigned long long)local);
    }
    __syncthreads();
    if (threadIdx.x < n_nation && s_rev[threadIdx.x])
        atomicAdd(revenue + threadIdx.x, s_rev[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// Q10: per-order revenue of returned lineitems, summed per customer key.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void q10_orders_k(
    const int16_t* __restrict__ odate, int64_t o_rows, const int32_t* __restrict__ okey,
    const int32_t* __restrict__ ocust, LiSide li, const uint8_t* __restrict__ code,
    int32_t code_lo, int32_t code_span, const int32_t* __restrict__ dprice, int32_t date_lo,
    int32_t date_hi, int64_t cust_size, unsigned long long* __restrict__ rev,
    uint8_t* __restrict__ has) {
    const Window w = wg_window(odate, o_rows, date_lo, date_hi);
    for (int64_t o = w.s + threadIdx.x; o < w.e; o += kT) {
        const int32_t ck = ocust[o];
        if (ck < 0 || ck >= cust_size) continue;
        const uint32_t e = li.end[o];
        if (e == 0) continue;
        const int32_t key = li.okey ? okey[o] : 0;
        int64_t order_rev = 0;
        for (uint32_t r = li.start[o]; r < e; ++r) {
            if (li.okey && li.okey[r] != key) continue;
            if (uint32_t(int32_t(code[r]) - code_lo) >= uint32_t(code_span)) continue;
            order_rev += dprice[r];
        }
        if (order_rev != 0) {
            atomicAdd(rev + ck, (unsigned long long)order_rev);
            has[ck] = 1;
        }
    }
}

// custkey -> (row + 1) of the last customer row with that key (atomic max).
__global__ __launch_bounds__(kT) void cust_map_k(const int32_t* __restrict__ ck, int64_t rows,
                                                 int64_t size, int32_t* __restrict__ map) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = ck[i];
        if (k >= 0 && k < size) atomicMax(map + k, int32_t(i + 1));
    }
}

// A kept Q10 group: has revenue, a customer row, a nation key in range.
BOLAP_DEV bool q10_keep(int64_t k, const uint8_t* has, const int32_t* map,
                        const int32_t* nation, int32_t n_names) {
    if (!has[k]) return false;
    const int32_t row = map[k] - 1;
    if (row < 0) return false;
    const int32_t nk = nation[row];
    return nk >= 0 && nk < n_names;
}

// Ordered compaction of the kept keys, pass 1: kept keys per workgroup tile.
constexpr int kTile = kT * 16;
__global__ __launch_bounds__(kT) void q10_count_k(int64_t size, const uint8_t* __restrict__ has,
                                                  const int32_t* __restrict__ map,
                                                  const int32_t* __restrict__ nation,
                                                  int32_t n_names, uint32_t* __restrict__ tiles) {
    const int64_t t0 = int64_t(blockIdx.x) * kTile;
    uint32_t n = 0;
    for (int j = 0; j < 16; ++j) {
        const int64_t k = t0 + int64_t(j) * kT + threadIdx.x;
        if (k < size && q10_keep(k, has, map, nation, n_names)) ++n;
    }
    __shared__ uint32_t s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    atomicAdd(&s, n);
    __syncthreads();
    if (threadIdx.x == 0) tiles[blockIdx.x] = s;
}

// Exclusive scan of the tile counts (one workgroup); total in tiles[n].
__global__ __launch_bounds__(1024) void scan_tiles_k(uint32_t* __restrict__ tiles, int64_t n) {
    __shared__ uint32_t s_carry;
    __shared__ uint32_t s_w[1024 / kWave];
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t b = 0; b < n; b += 1024) {
        const int64_t i = b + threadIdx.x;
        const uint32_t v = i < n ? tiles[i] : 0;
        // wave inclusive scan through LDS by wave totals
        uint32_t x = v;
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, kWave);
            if ((threadIdx.x & (kWave - 1)) >= d) x += y;
        }
        if ((threadIdx.x & (kWave - 1)) == kWave - 1) s_w[threadIdx.x / kWave] = x;
        __syncthreads();
        uint32_t before = 0;
        for (int w = 0; w < int(threadIdx.x / kWave); ++w) before += s_w[w];
        const uint32_t carry = s_carry;
        if (i < n) tiles[i] = carry + before + x - v;
        __syncthreads();
        if (threadIdx.x == 1023) s_carry = carry + before + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) tiles[n] = s_carry;
}

// Pass 2: each tile writes its kept keys in key order at its offset, as
// (~revenue, custkey) pairs for the stable sort.
__global__ __launch_bounds__(kT) void q10_write_k(int64_t size, const uint8_t* __restrict__ has,
                                                  const int32_t* __restrict__ map,
                                                  const int32_t* __restrict__ nation,
                                                  int32_t n_names,
                                                  const unsigned long long* __restrict__ rev,
                                                  const uint32_t* __restrict__ tiles,
                                                  uint64_t* __restrict__ keys,
                                                  uint32_t* __restrict__ vals,
                                                  unsigned long long* __restrict__ maxrev) {
    __shared__ uint32_t s_w[kT / kWave];
    __shared__ unsigned long long s_m[kT / kWave];
    const int64_t t0 = int64_t(blockIdx.x) * kTile;
    uint32_t base = tiles[blockIdx.x];
    unsigned long long mx = 0;  // the largest kept revenue (sort width)
    for (int j = 0; j < 16; ++j) {
        const int64_t k = t0 + int64_t(j) * kT + threadIdx.x;
        const bool keep = k < size && q10_keep(k, has, map, nation, n_names);
        const uint64_t bal = __ballot(keep);
        const uint32_t lane = threadIdx.x & (kWave - 1);
        const uint32_t before_lane = __popcll(bal & ((uint64_t(1) << lane) - 1));
        if (lane == 0) s_w[threadIdx.x / kWave] = __popcll(bal);
        __syncthreads();
        uint32_t before = 0;
        for (int w = 0; w < int(threadIdx.x / kWave); ++w) before += s_w[w];
        uint32_t total = 0;
        for (int w = 0; w < kT / kWave; ++w) total += s_w[w];
        if (keep) {
            const uint32_t at = base + before + before_lane;
            const unsigned long long r = rev[k];
            keys[at] = ~uint64_t(r);
            vals[at] = uint32_t(k);
            mx = r > mx ? r : mx;
        }
        base += total;
        __syncthreads();
    }
    for (int d = 1; d < kWave; d <<= 1) {
        const unsigned long long o = __shfl_xor(mx, d, kWave);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) s_m[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kT / kWave; ++w) mx = s_m[w] > mx ? s_m[w] : mx;
        if (mx) atomicMax(maxrev, mx);
    }
}

// Sorted pairs -> output rows {custkey, cust_row, revenue}.
// The group count and the largest revenue side by side, for one read-back.
__global__ void q10_stage_k(const uint32_t* __restrict__ total,
                            const unsigned long long* __restrict__ maxrev,
                            unsigned long long* __restrict__ stage) {
    stage[0] = *total;
    stage[1] = *maxrev;
}

__global__ __launch_bounds__(kT) void q10_rows_k(int64_t n, const uint64_t* __restrict__ keys,
                                                 const uint32_t* __restrict__ vals,
                                                 const int32_t* __restrict__ map,
                                                 bolap_q10_row* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const uint32_t k = vals[i];
        out[i] = bolap_q10_row{int32_t(k), map[k] - 1, int64_t(~keys[i])};
    }
}

// ---------------------------------------------------------------------------
// Q12: lineitems by ship mode, split by the priority of their order.
// ---------------------------------------------------------------------------
struct Q12K {
    int64_t rows;
    const uint8_t* mode;
    const uint32_t* cr;
    const int16_t* ship;
    const int32_t* okey;
    int32_t mode1, mode2;
    int32_t date_lo;
    uint32_t span;
    const int32_t* map;
    int64_t map_size;
    const uint8_t* prio;
    int32_t high1, high2;
};

// Adds one passing row to the lane's four counters {high, low} x {mode1, mode2}.
BOLAP_DEV void q12_row(const Q12K& a, int64_t i, int32_t m, uint32_t p, int32_t ship,
                       uint32_t (&c)[4]) {
    if (m != a.mode1 && m != a.mode2) return;
    const int32_t commit = int16_t(p & 0xffff), receipt = int16_t(p >> 16);
    if (commit >= receipt || ship >= commit) return;
    if (uint32_t(receipt - a.date_lo) >= a.span) return;
    const int32_t k = a.okey[i];
    if (uint32_t(k) >= uint64_t(a.map_size)) return;
    const int32_t row = a.map[k];
    if (row < 0) return;
    const int32_t pr = a.prio[row];
    const int slot = m == a.mode1 ? 0 : 2;
    c[slot + ((pr == a.high1 || pr == a.high2) ? 0 : 1)] += 1;
}

__global__ __launch_bounds__(kT) void q12_k(Q12K a, unsigned long long* __restrict__ out) {
    uint32_t c[4] = {0, 0, 0, 0};
    const int64_t units = a.rows / 4;
    const int64_t stride = int64_t(gridDim.x) * kT;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < units; u += stride) {
        const uint32_t m4 = reinterpret_cast<const uint32_t*>(a.mode)[u];
        bool any = false;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int32_t m = (m4 >> (8 * j)) & 0xff;
            any |= m == a.mode1 || m == a.mode2;
        }
        if (!any) continue;
        const uint4 p4 = reinterpret_cast<const uint4*>(a.cr)[u];
        const VecT<int16_t, 4> s4 = load_unit<int16_t, 4>(a.ship, u);
        const uint32_t pv[4] = {p4.x, p4.y, p4.z, p4.w};
#pragma unroll
        for (int j = 0; j < 4; ++j)
            q12_row(a, u * 4 + j, int32_t((m4 >> (8 * j)) & 0xff), pv[j], s4.v[j], c);
    }
    // tail rows
    for (int64_t i = units * 4 + int64_t(blockIdx.x) * kT + threadIdx.x; i < a.rows; i += stride)
        q12_row(a, i, a.mode[i], a.cr[i], a.ship[i], c);
    __shared__ unsigned s[4];
    if (threadIdx.x < 4) s[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint64_t t = wave_sum_u64(c[j]);
        if ((threadIdx.x & (kWave - 1)) == 0 && t) atomicAdd(&s[j], unsigned(t));
    }
    __syncthreads();
    if (threadIdx.x < 4 && s[threadIdx.x])
        atomicAdd(out + threadIdx.x, (unsigned long long)s[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// Q19: part mask (last part row of a key wins), then the lineitem probe.
// ---------------------------------------------------------------------------
struct Q19Part {
    int64_t rows;
    const int32_t* key;
    const void* brand;
    int32_t brand_dtype;
    const void* container;
    int32_t container_dtype;
    const int32_t* size;
    bolap_q19_group g[3];
};

BOLAP_DEV int32_t code_at(const void* col, int32_t dtype, int64_t i) {
    return dtype == BOLAP_U8 ? int32_t(static_cast<const uint8_t*>(col)[i])
                             : int32_t(static_cast<const int16_t*>(col)[i]);
}

__global__ __launch_bounds__(kT) void q19_part_k(Q19Part a, int64_t size,
                                                 int32_t* __restrict__ owner) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < a.rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = a.key[i];
        if (k < 0 || k >= size) continue;
        const int32_t b = code_at(a.brand, a.brand_dtype, i);
        const int32_t c = code_at(a.container, a.container_dtype, i);
        const int32_t sz = a.size[i];
        uint32_t m = 0;
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            const bolap_q19_group& G = a.g[g];
            const bool in = c == G.containers[0] || c == G.containers[1] ||
                            c == G.containers[2] || c == G.containers[3];
            if (G.enabled && b == G.brand && c >= 0 && in && sz >= 1 && sz <= G.size_hi)
                m |= 1u << g;
        }
        // rows numbered from 1 so that 0 means "no part row"
        atomicMax(owner + k, int32_t((uint32_t(i + 1) << 3) | m));
    }
}

__global__ __launch_bounds__(kT) void q19_narrow_k(const int32_t* __restrict__ owner,
                                                   int64_t size, uint8_t* __restrict__ mask) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < size;
         i += int64_t(gridDim.x) * kT)
        mask[i] = uint8_t(owner[i] & 7);
}

struct Q19Li {
    int64_t rows;
    const uint8_t* mode;
    const uint8_t* instruct;
    const int32_t* pkey;
    const int16_t* qty;
    const int32_t* price;
    const uint8_t* disc;
    int32_t mode1, mode2, instr;
    int32_t qlo[3];
    const uint8_t* mask;
    int64_t size;
};

BOLAP_DEV int64_t q19_row(const Q19Li& a, int64_t i, int32_t pk) {
    if (pk < 0 || pk >= a.size) return 0;
    const uint32_t m = a.mask[pk];
    if (!m) return 0;
    const int32_t q = a.qty[i];
    bool ok = false;
#pragma unroll
    for (int g = 0; g < 3; ++g)
        ok |= ((m >> g) & 1) && q >= a.qlo[g] && q <= a.qlo[g] + 1000;
    if (!ok) return 0;
    return int64_t(a.price[i]) * int64_t(100 - int32_t(a.disc[i]));
}

BOLAP_DEV bool q19_codes(const Q19Li& a, int32_t m, int32_t ins) {
    return (m == a.mode1 || m == a.mode2) && ins == a.instr;
}

__global__ __launch_bounds__(kT) void q19_li_k(Q19Li a, unsigned long long* __restrict__ out) {
    int64_t acc = 0;
    const int64_t units = a.rows / 4;
    const int64_t stride = int64_t(gridDim.x) * kT;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < units; u += stride) {
        const uint32_t m4 = reinterpret_cast<const uint32_t*>(a.mode)[u];
        const uint32_t i4 = reinterpret_cast<const uint32_t*>(a.instruct)[u];
        uint32_t pass = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            pass |= uint32_t(q19_codes(a, (m4 >> (8 * j)) & 0xff, (i4 >> (8 * j)) & 0xff)) << j;
        if (!pass) continue;
        const VecT<int32_t, 4> k4 = load_unit<int32_t, 4>(a.pkey, u);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if ((pass >> j) & 1) acc += q19_row(a, u * 4 + j, k4.v[j]);
    }
    for (int64_t i = units * 4 + int64_t(blockIdx.x) * kT + threadIdx.x; i < a.rows; i += stride)
        if (q19_codes(a, a.mode[i], a.instruct[i])) acc += q19_row(a, i, a.pkey[i]);
    __shared__ uint64_t s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const uint64_t t = wave_sum_u64(uint64_t(acc));
    if ((threadIdx.x & (kWave - 1)) == 0 && t) atomicAdd((unsigned long long*)&s, t);
    __syncthreads();
    if (threadIdx.x == 0 && s) atomicAdd(out, (unsigned long long)s);
}

// ---------------------------------------------------------------------------
// Small helpers.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void fill_bytes_k(uint32_t* __restrict__ p, int64_t words) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < words;
         i += int64_t(gridDim.x) * kT)
        p[i] = 0;
}

// Maximum of an i32 column into *out (which starts at INT32_MIN).
// Workgroup maximum of m; every thread gets it.
BOLAP_DEV int32_t block_max_i32(int32_t m) {
    __shared__ int32_t s_m[kT / kWave];
    for (int d = kWave / 2; d > 0; d >>= 1) {
        const int32_t o = __shfl_xor(m, d, kWave);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) s_m[threadIdx.x / kWave] = m;
    __syncthreads();
    m = s_m[0];
    for (int w = 1; w < kT / kWave; ++w) m = s_m[w] > m ? s_m[w] : m;
    return m;
}

// One atomic per workgroup (same-address atomics serialise).
__global__ __launch_bounds__(kT) void max_i32_k(const int32_t* __restrict__ v, int64_t n,
                                                int32_t* __restrict__ out) {
    int32_t m = INT32_MIN;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        m = v[i] > m ? v[i] : m;
    m = block_max_i32(m);
    if (threadIdx.x == 0 && m != INT32_MIN) atomicMax(out, m);
}

inline unsigned grid_of(bolap_ctx* ctx, int64_t items, int per_cu = 4) {
    int64_t b = (items + kT - 1) / kT;
    const int64_t cap = int64_t(ctx->num_cus) * per_cu;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

// Zeroes `bytes` (a multiple of 4) of device memory with a kernel (so the
// calls stay free of memset nodes when a caller captures them).
int zero_dev(bolap_ctx* ctx, void* p, size_t bytes) {
    const int64_t words = int64_t((bytes + 3) / 4);
    hipLaunchKernelGGL(fill_bytes_k, dim3(grid_of(ctx, words)), dim3(kT), 0, ctx->stream,
                       static_cast<uint32_t*>(p), words);
    return launch_check("fill_bytes_k");
}

// Copies `bytes` of device results to the host and waits for them.
int fetch(bolap_ctx* ctx, void* host, const void* dev, size_t bytes) {
    BOLAP_HIP_TRY(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BOLAP_OK;
}

// Maximum of an i32 column (INT32_MIN when empty), read back to the host.
int max_i32(bolap_ctx* ctx, const int32_t* v, int64_t n, int32_t* slot, int32_t* out) {
    const int32_t init = INT32_MIN;
    BOLAP_HIP_TRY(hipMemcpyAsync(slot, &init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
    if (n > 0) {
        hipLaunchKernelGGL(max_i32_k, dim3(grid_of(ctx, n, 1)), dim3(kT), 0, ctx->stream, v, n,
                           slot);
        BOLAP_TRY(launch_check("max_i32_k"));
    }
    return fetch(ctx, out, slot, sizeof(int32_t));
}

int check_ranges(const bolap_order_side* o, const char* who) {
    BOLAP_ARG_CHECK(o, "%s: null order side", who);
    BOLAP_ARG_CHECK(o->order_rows >= 0 && o->order_rows < (int64_t(1) << 31),
                    "%s: order rows %lld", who, (long long)o->order_rows);
    BOLAP_ARG_CHECK(o->order_rows == 0 || (o->o_orderdate && o->range_start && o->range_end),
                    "%s: null order column", who);
    BOLAP_ARG_CHECK(o->l_orderkey_sorted || o->order_rows == 0 ||
                        (o->l_orderkey && o->o_orderkey),
                    "%s: unsorted lineitem needs l_orderkey and o_orderkey", who);
    return BOLAP_OK;
}

LiSide li_side(const bolap_order_side* o) {
    return LiSide{o->range_start, o->range_end, o->l_orderkey_sorted ? nullptr : o->l_orderkey};
}

unsigned window_grid(bolap_ctx* ctx, int64_t rows) {
    // enough workgroups for a window of any size; each searches its window
    return grid_of(ctx, rows, 2);
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" int bolap_q4(bolap_ctx* ctx, const bolap_q4_args* a, int64_t* counts) {
    BOLAP_ARG_CHECK(ctx && a && counts, "bolap_q4: null argument");
    BOLAP_TRY(check_ranges(&a->orders, "bolap_q4"));
    BOLAP_ARG_CHECK(a->n_priority >= 0 && a->n_priority <= 256, "bolap_q4: %d priority codes",
                    a->n_priority);
    const bolap_order_side& o = a->orders;
    BOLAP_ARG_CHECK(o.order_rows == 0 || (a->o_orderpriority && a->l_commit_receipt),
                    "bolap_q4: null column");
    auto* res = static_cast<unsigned long long*>(ctx->result_dev);
    BOLAP_TRY(zero_dev(ctx, res, 256 * 8));
    if (o.order_rows > 0 && a->n_priority > 0) {
        TraceScope ts(ctx, "q4_orders");
        hipLaunchKernelGGL(q4_k, dim3(window_grid(ctx, o.order_rows)), dim3(kT), 0, ctx->stream,
                           o.o_orderdate, o.order_rows, o.o_orderkey, a->o_orderpriority,
                           a->n_priority, li_side(&o), a->l_commit_receipt, a->date_lo,
                           a->date_hi, res);
        BOLAP_TRY(launch_check("q4_k"));
    }
    if (a->n_priority > 0) BOLAP_TRY(fetch(ctx, counts, res, size_t(a->n_priority) * 8));
    return BOLAP_OK;
}

extern "C" int bolap_q5(bolap_ctx* ctx, const bolap_q5_args* a, int64_t* revenue) {
    BOLAP_ARG_CHECK(ctx && a && revenue, "bolap_q5: null argument");
    BOLAP_TRY(check_ranges(&a->orders, "bolap_q5"));
    BOLAP_ARG_CHECK(a->n_nation >= 0 && a->n_nation <= 256, "bolap_q5: %d nations (at most 256)",
                    a->n_nation);
    const bolap_order_side& o = a->orders;
    BOLAP_ARG_CHECK(o.order_rows == 0 || (a->o_cust_nation && a->l_supp_nation &&
                                          a->l_disc_price && a->nation_in_region),
                    "bolap_q5: null column");
    auto* res = static_cast<unsigned long long*>(ctx->result_dev);
    uint8_t* inreg = reinterpret_cast<uint8_t*>(res + 256);
    BOLAP_TRY(zero_dev(ctx, res, 256 * 8));
    if (o.order_rows > 0 && a->n_nation > 0) {
        BOLAP_HIP_TRY(hipMemcpyAsync(inreg, a->nation_in_region, size_t(a->n_nation),
                                     hipMemcpyHostToDevice, ctx->stream));
        TraceScope ts(ctx, "q5_orders");
        hipLaunchKernelGGL(q5_k, dim3(window_grid(ctx, o.order_rows)), dim3(kT), 0, ctx->stream,
                           o.o_orderdate, o.order_rows, o.o_orderkey, a->o_cust_nation,
                           li_side(&o), a->l_supp_nation, a->l_disc_price, inreg, a->n_nation,
                           a->date_lo, a->date_hi, res);
        BOLAP_TRY(launch_check("q5_k"));
    }
    if (a->n_nation > 0) BOLAP_TRY(fetch(ctx, revenue, res, size_t(a->n_nation) * 8));
    return BOLAP_OK;
}

extern "C" int bolap_q10(bolap_ctx* ctx, const bolap_q10_args* a, bolap_q10_row* rows,
                         int64_t max_rows, int64_t* n_rows) {
    BOLAP_ARG_CHECK(ctx && a && n_rows && (rows || max_rows == 0), "bolap_q10: null argument");
    BOLAP_TRY(check_ranges(&a->orders, "bolap_q10"));
    const bolap_order_side& o = a->orders;
    BOLAP_ARG_CHECK(o.order_rows == 0 || (a->o_custkey && a->l_code && a->l_disc_price),
                    "bolap_q10: null column");
    BOLAP_ARG_CHECK(a->cust_rows >= 0 && (a->cust_rows == 0 || (a->c_custkey && a->c_nationkey)),
                    "bolap_q10: null customer column");
    *n_rows = 0;
    if (a->code_lo < 0 || a->code_span <= 0) return BOLAP_OK;  // no 'R' in the dictionary
    // customer key domain [0, max key] (q10::max_value; 0 for no customers)
    int32_t* slot = static_cast<int32_t*>(ctx->result_dev);
    int32_t kmax = 0;
    if (a->cust_rows > 0) BOLAP_TRY(max_i32(ctx, a->c_custkey, a->cust_rows, slot, &kmax));
    if (kmax < 0) return BOLAP_OK;
    const int64_t size = int64_t(kmax) + 1;
    const int64_t tiles = (size + kTile - 1) / kTile;
    WsArena ws;
    const size_t o_rev = ws.plan(size_t(size) * 8);
    const size_t o_has = ws.plan(size_t(size));
    const size_t o_map = ws.plan(size_t(size) * 4);
    const size_t o_max = ws.plan(8);  // cleared with rev, has and map
    const size_t o_tiles = ws.plan(size_t(tiles + 1) * 4);
    const size_t o_keys = ws.plan(size_t(size) * 8);
    const size_t o_vals = ws.plan(size_t(size) * 4);
    const size_t o_out = ws.plan(size_t(size) * sizeof(bolap_q10_row));
    const size_t o_sort = ws.plan(radix_sort_scratch_bytes(ctx, size));
    const size_t o_stage = ws.plan(16);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    auto* rev = ws.at<unsigned long long>(o_rev);
    auto* has = ws.at<uint8_t>(o_has);
    auto* map = ws.at<int32_t>(o_map);
    auto* tl = ws.at<uint32_t>(o_tiles);
    // rev, has and map are contiguous regions: one clearing launch
    BOLAP_TRY(zero_dev(ctx, rev, o_tiles - o_rev));
    {
        TraceScope ts(ctx, "q10_orders");
        hipLaunchKernelGGL(cust_map_k, dim3(grid_of(ctx, a->cust_rows)), dim3(kT), 0,
                           ctx->stream, a->c_custkey, a->cust_rows, size, map);
        BOLAP_TRY(launch_check("cust_map_k"));
        if (o.order_rows > 0) {
            hipLaunchKernelGGL(q10_orders_k, dim3(window_grid(ctx, o.order_rows)), dim3(kT), 0,
                               ctx->stream, o.o_orderdate, o.order_rows, o.o_orderkey, a->o_custkey,
                               li_side(&o), a->l_code, a->code_lo, a->code_span, a->l_disc_price,
                               a->date_lo, a->date_hi, size, rev, has);
            BOLAP_TRY(launch_check("q10_orders_k"));
        }
    }
    {
        TraceScope ts(ctx, "q10_compact");
        hipLaunchKernelGGL(q10_count_k, dim3(unsigned(tiles)), dim3(kT), 0, ctx->stream, size, has,
                           map, a->c_nationkey, a->n_nation_names, tl);
        hipLaunchKernelGGL(scan_tiles_k, dim3(1), dim3(1024), 0, ctx->stream, tl, tiles);
        hipLaunchKernelGGL(q10_write_k, dim3(unsigned(tiles)), dim3(kT), 0, ctx->stream, size, has,
                           map, a->c_nationkey, a->n_nation_names, rev, tl,
                           ws.at<uint64_t>(o_keys), ws.at<uint32_t>(o_vals),
                           ws.at<unsigned long long>(o_max));
        hipLaunchKernelGGL(q10_stage_k, dim3(1), dim3(1), 0, ctx->stream, tl + tiles,
                           ws.at<unsigned long long>(o_max),
                           ws.at<unsigned long long>(o_stage));
        BOLAP_TRY(launch_check("q10_write_k"));
    }
    unsigned long long staged[2] = {0, 0};
    BOLAP_TRY(fetch(ctx, staged, ws.at<unsigned long long>(o_stage), sizeof(staged)));
    const uint32_t n = uint32_t(staged[0]);
    // Every kept revenue is at most the largest, mx < 2^b, so the low b bits
    // of ~revenue are 2^b - 1 - revenue: sorting on them alone gives the
    // full 64-bit order (ties stay in customer-key order, the passes being
    // stable) in ceil(b / 8) passes instead of eight.
    const int sort_bits = staged[1] ? 64 - __builtin_clzll(staged[1]) : 1;
    *n_rows = n;
    trace_count(ctx, "q10_agg_groups_created", n);
    if (n > max_rows) {
        set_error("bolap_q10: %u result rows do not fit %lld", n, (long long)max_rows);
        return BOLAP_ERR_CAPACITY;
    }
    if (n == 0) return BOLAP_OK;
    {
        TraceScope ts(ctx, "q10_sort");
        int passes = 0;
        BOLAP_TRY(radix_sort_pairs(ctx, ws.at<uint64_t>(o_keys), ws.at<uint32_t>(o_vals), n,
                                   sort_bits, ws.at<void>(o_sort), &passes));
        hipLaunchKernelGGL(q10_rows_k, dim3(grid_of(ctx, n)), dim3(kT), 0, ctx->stream, int64_t(n),
                           ws.at<uint64_t>(o_keys), ws.at<uint32_t>(o_vals), map,
                           ws.at<bolap_q10_row>(o_out));
        BOLAP_TRY(launch_check("q10_rows_k"));
    }
    return fetch(ctx, rows, ws.at<bolap_q10_row>(o_out), size_t(n) * sizeof(bolap_q10_row));
}

extern "C" int bolap_q12(bolap_ctx* ctx, const bolap_q12_args* a, bolap_q12_result* out) {
    BOLAP_ARG_CHECK(ctx && a && out, "bolap_q12: null argument");
    BOLAP_ARG_CHECK(a->li_rows >= 0, "bolap_q12: negative rows");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (a->l_shipmode && a->l_commit_receipt && a->l_shipdate &&
                                        a->l_orderkey && a->order_map && a->o_orderpriority),
                    "bolap_q12: null column");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (aligned(a->l_shipmode, 4) &&
                                        aligned(a->l_commit_receipt, 16) &&
                                        aligned(a->l_shipdate, 8)),
                    "bolap_q12: l_shipmode / l_commit_receipt / l_shipdate must be 4 / 16 / 8-"
                    "byte aligned");
    std::memset(out, 0, sizeof(*out));
    if (a->mode1 < 0 && a->mode2 < 0) return BOLAP_OK;
    auto* res = static_cast<unsigned long long*>(ctx->result_dev);
    BOLAP_TRY(zero_dev(ctx, res, 4 * 8));
    if (a->li_rows > 0) {
        const Q12K k{a->li_rows, a->l_shipmode, a->l_commit_receipt, a->l_shipdate, a->l_orderkey,
                     a->mode1, a->mode2, a->date_lo, uint32_t(a->date_hi - a->date_lo),
                     a->order_map, a->map_size, a->o_orderpriority, a->high1, a->high2};
        TraceScope ts(ctx, "q12_lineitem");
        hipLaunchKernelGGL(q12_k, dim3(grid_of(ctx, (a->li_rows + 3) / 4, 8)), dim3(kT), 0,
                           ctx->stream, k, res);
        BOLAP_TRY(launch_check("q12_k"));
    }
    unsigned long long c[4];
    BOLAP_TRY(fetch(ctx, c, res, sizeof(c)));
    out->high[0] = int64_t(c[0]);
    out->low[0] = int64_t(c[1]);
    out->high[1] = int64_t(c[2]);
    out->low[1] = int64_t(c[3]);
    return BOLAP_OK;
}

extern "C" int bolap_q19(bolap_ctx* ctx, const bolap_q19_args* a, int64_t* revenue) {
    BOLAP_ARG_CHECK(ctx && a && revenue, "bolap_q19: null argument");
    BOLAP_ARG_CHECK(a->part_rows >= 0 && a->li_rows >= 0, "bolap_q19: negative rows");
    BOLAP_ARG_CHECK(a->part_rows == 0 || (a->p_partkey && a->p_brand && a->p_container &&
                                          a->p_size),
                    "bolap_q19: null part column");
    BOLAP_ARG_CHECK((a->p_brand_dtype == BOLAP_U8 || a->p_brand_dtype == BOLAP_I16) &&
                        (a->p_container_dtype == BOLAP_U8 || a->p_container_dtype == BOLAP_I16),
                    "bolap_q19: brand / container codes must be u8 or i16");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (a->l_partkey && a->l_quantity && a->l_price &&
                                        a->l_disc && a->l_shipmode && a->l_shipinstruct),
                    "bolap_q19: null lineitem column");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (aligned(a->l_shipmode, 4) &&
                                        aligned(a->l_shipinstruct, 4) &&
                                        aligned(a->l_partkey, 16)),
                    "bolap_q19: l_shipmode / l_shipinstruct / l_partkey must be 4 / 4 / 16-byte "
                    "aligned");
    BOLAP_ARG_CHECK(a->part_rows < (int64_t(1) << 28) - 1, "bolap_q19: %lld part rows (at most "
                    "2^28 - 2)", (long long)a->part_rows);
    *revenue = 0;
    const bool any = a->groups[0].enabled || a->groups[1].enabled || a->groups[2].enabled;
    if (!any || a->shipinstruct < 0 || (a->mode1 < 0 && a->mode2 < 0) || a->part_rows == 0 ||
        a->li_rows == 0)
        return BOLAP_OK;
    int32_t* slot = static_cast<int32_t*>(ctx->result_dev);
    int32_t kmax = 0;
    BOLAP_TRY(max_i32(ctx, a->p_partkey, a->part_rows, slot, &kmax));
    if (kmax < 0) return BOLAP_OK;
    const int64_t size = int64_t(kmax) + 1;
    WsArena ws;
    const size_t o_owner = ws.plan(size_t(size) * 4);
    const size_t o_mask = ws.plan(size_t(size));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    auto* res = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->result_dev) + 256);
    BOLAP_TRY(zero_dev(ctx, res, 8));
    BOLAP_TRY(zero_dev(ctx, ws.at<int32_t>(o_owner), size_t(size) * 4));
    Q19Part p{a->part_rows, a->p_partkey, a->p_brand, a->p_brand_dtype, a->p_container,
              a->p_container_dtype, a->p_size, {a->groups[0], a->groups[1], a->groups[2]}};
    {
        TraceScope ts(ctx, "q19_part_scan");
        hipLaunchKernelGGL(q19_part_k, dim3(grid_of(ctx, a->part_rows)), dim3(kT), 0, ctx->stream,
                           p, size, ws.at<int32_t>(o_owner));
        hipLaunchKernelGGL(q19_narrow_k, dim3(grid_of(ctx, size)), dim3(kT), 0, ctx->stream,
                           ws.at<int32_t>(o_owner), size, ws.at<uint8_t>(o_mask));
        BOLAP_TRY(launch_check("q19_narrow_k"));
    }
    Q19Li l{a->li_rows, a->l_shipmode, a->l_shipinstruct, a->l_partkey, a->l_quantity, a->l_price,
            a->l_disc, a->mode1, a->mode2, a->shipinstruct,
            {a->groups[0].qty_lo, a->groups[1].qty_lo, a->groups[2].qty_lo},
            ws.at<uint8_t>(o_mask), size};
    {
        TraceScope ts(ctx, "q19_lineitem_scan");
        hipLaunchKernelGGL(q19_li_k, dim3(grid_of(ctx, (a->li_rows + 3) / 4, 8)), dim3(kT), 0,
                           ctx->stream, l, res);
        BOLAP_TRY(launch_check("q19_li_k"));
    }
    return fetch(ctx, revenue, res, 8);
}

// ---------------------------------------------------------------------------
// Q21 and Q22: the anti-join queries.
// ---------------------------------------------------------------------------
namespace bolap {
namespace {

// suppkey -> (row + 1) << 1 | in target nation, of the last supplier row of
// the key (atomic max over row numbers).
__global__ __launch_bounds__(kT) void q21_owner_k(const int32_t* __restrict__ sk,
                                                  const int32_t* __restrict__ sn, int64_t rows,
                                                  int64_t n, int32_t nation,
                                                  int32_t* __restrict__ owner) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = sk[i];
        if (k >= 0 && k < n)
            atomicMax(owner + k, int32_t((uint32_t(i + 1) << 1) | uint32_t(sn[i] == nation)));
    }
}

// One order per lane over every order: a final-status order with more than
// one lineitem, several suppliers and exactly one late supplier adds its
// late-line count to that supplier when the supplier is a target.
__global__ __launch_bounds__(kT) void q21_k(int64_t o_rows, const uint8_t* __restrict__ status,
                                            int32_t final_code, const int32_t* __restrict__ okey,
                                            LiSide li, const int32_t* __restrict__ sk,
                                            const uint32_t* __restrict__ cr,
                                            const int32_t* __restrict__ owner, int64_t n,
                                            unsigned long long* __restrict__ wait) {
    for (int64_t o = int64_t(blockIdx.x) * kT + threadIdx.x; o < o_rows;
         o += int64_t(gridDim.x) * kT) {
        if (int32_t(status[o]) != final_code) continue;
        const uint32_t e = li.end[o];
        if (e == 0) continue;
        const uint32_t s = li.start[o];
        if (e - s <= 1) continue;
        const int32_t key = li.okey ? okey[o] : 0;
        int32_t first = -1, late_supp = -1, late_n = 0;
        bool multi = false, multi_late = false;
        for (uint32_t r = s; r < e; ++r) {
            if (li.okey && li.okey[r] != key) continue;
            const int32_t k = sk[r];
            if (uint32_t(k) >= uint64_t(n)) continue;
            if (first < 0) first = k;
            else if (k != first) multi = true;
            const uint32_t p = cr[r];
            if ((p >> 16) > (p & 0xffffu)) {
                if (late_supp == -1 || late_supp == k) {
                    late_supp = k;
                    ++late_n;
                } else {
                    multi_late = true;
                    break;
                }
            }
        }
        if (multi_late || !multi || late_supp < 0) continue;
        if (owner[late_supp] & 1) atomicAdd(wait + late_supp, (unsigned long long)late_n);
    }
}

struct Q22F {
    uint64_t allowed[2];  // two-digit prefix codes 0..99
    int32_t n_extra;
    int32_t extra[7];
};

// Prefix slot of a customer: the code (0..99), 100 + extra index, or -1.
BOLAP_DEV int32_t q22_slot(const Q22F& f, uint8_t code, int32_t pair) {
    if (code != 255) return code < 100 && ((f.allowed[code >> 6] >> (code & 63)) & 1) ? code : -1;
    if (pair < 0) return -1;
    for (int e = 0; e < f.n_extra; ++e)
        if (f.extra[e] == pair) return 100 + e;
    return -1;
}

__global__ __launch_bounds__(kT) void q22_mark_k(const int32_t* __restrict__ ock, int64_t rows,
                                                 int64_t size, uint8_t* __restrict__ has) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = ock[i];
        if (k >= 0 && k < size) has[k] = 1;
    }
}

// Sum and count of the candidates' acctbal (acctbal > 0) into tot[0], tot[1].
__global__ __launch_bounds__(kT) void q22_avg_k(Q22F f, int64_t rows,
                                                const uint8_t* __restrict__ code,
                                                const int32_t* __restrict__ pair,
                                                const int32_t* __restrict__ bal,
                                                unsigned long long* __restrict__ tot) {
    uint64_t s = 0, c = 0;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        if (q22_slot(f, code[i], pair[i]) >= 0 && bal[i] > 0) {
            s += uint64_t(int64_t(bal[i]));
            ++c;
        }
    }
    s = wave_sum_u64(s);
    c = wave_sum_u64(c);
    if ((threadIdx.x & (kWave - 1)) == 0 && c) {
        atomicAdd(tot, (unsigned long long)s);
        atomicAdd(tot + 1, (unsigned long long)c);
    }
}

// Candidates above the floor of the average and without orders, per slot.
__global__ __launch_bounds__(kT) void q22_agg_k(Q22F f, int64_t rows,
                                                const int32_t* __restrict__ ck,
                                                const uint8_t* __restrict__ code,
                                                const int32_t* __restrict__ pair,
                                                const int32_t* __restrict__ bal,
                                                const uint8_t* __restrict__ has, int64_t size,
                                                unsigned long long* __restrict__ tot) {
    __shared__ unsigned long long s_agg[107 * 2];
    for (int i = threadIdx.x; i < 107 * 2; i += kT) s_agg[i] = 0;
    __syncthreads();
    const uint64_t cnt = tot[1];
    if (cnt > 0) {
        const int64_t avg = int64_t(tot[0]) / int64_t(cnt);
        for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
             i += int64_t(gridDim.x) * kT) {
            const int32_t slot = q22_slot(f, code[i], pair[i]);
            if (slot < 0 || bal[i] <= 0 || int64_t(bal[i]) <= avg) continue;
            const int32_t k = ck[i];
            if (k >= 0 && k < size && has[k]) continue;
            atomicAdd(&s_agg[2 * slot], 1ull);
            atomicAdd(&s_agg[2 * slot + 1], (unsigned long long)int64_t(bal[i]));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 107 * 2; i += kT)
        if (s_agg[i]) atomicAdd(tot + 2 + i, s_agg[i]);
}

}  // namespace
}  // namespace bolap

extern "C" int bolap_q21(bolap_ctx* ctx, const bolap_q21_args* a, int64_t* wait,
                         int64_t n_suppkeys) {
    BOLAP_ARG_CHECK(ctx && a && (wait || n_suppkeys == 0), "bolap_q21: null argument");
    BOLAP_ARG_CHECK(n_suppkeys >= 0 && n_suppkeys < (int64_t(1) << 31),
                    "bolap_q21: %lld supplier keys", (long long)n_suppkeys);
    BOLAP_ARG_CHECK(a->order_rows >= 0 && a->supp_rows >= 0 && a->supp_rows < (int64_t(1) << 30),
                    "bolap_q21: rows out of range");
    BOLAP_ARG_CHECK(a->order_rows == 0 || (a->o_orderstatus && a->range_start && a->range_end &&
                                           a->l_suppkey && a->l_commit_receipt),
                    "bolap_q21: null column");
    BOLAP_ARG_CHECK(a->l_orderkey_sorted || a->order_rows == 0 || (a->o_orderkey && a->l_orderkey),
                    "bolap_q21: unsorted lineitem needs l_orderkey and o_orderkey");
    BOLAP_ARG_CHECK(a->supp_rows == 0 || (a->s_suppkey && a->s_nationkey),
                    "bolap_q21: null supplier column");
    if (n_suppkeys == 0) return BOLAP_OK;
    WsArena ws;
    const size_t o_owner = ws.plan(size_t(n_suppkeys) * 4);
    const size_t o_wait = ws.plan(size_t(n_suppkeys) * 8);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    auto* owner = ws.at<int32_t>(o_owner);
    auto* w = ws.at<unsigned long long>(o_wait);
    BOLAP_TRY(zero_dev(ctx, owner, ws.off));
    TraceScope ts(ctx, "q21_orders");
    if (a->supp_rows > 0) {
        hipLaunchKernelGGL(q21_owner_k, dim3(grid_of(ctx, a->supp_rows)), dim3(kT), 0, ctx->stream,
                           a->s_suppkey, a->s_nationkey, a->supp_rows, n_suppkeys, a->nation,
                           owner);
        BOLAP_TRY(launch_check("q21_owner_k"));
    }
    if (a->order_rows > 0 && a->final_code >= 0) {
        const LiSide li{a->range_start, a->range_end,
                        a->l_orderkey_sorted ? nullptr : a->l_orderkey};
        hipLaunchKernelGGL(q21_k, dim3(grid_of(ctx, a->order_rows, 8)), dim3(kT), 0, ctx->stream,
                           a->order_rows, a->o_orderstatus, a->final_code, a->o_orderkey, li,
                           a->l_suppkey, a->l_commit_receipt, owner, n_suppkeys, w);
        BOLAP_TRY(launch_check("q21_k"));
    }
    return fetch(ctx, wait, w, size_t(n_suppkeys) * 8);
}

extern "C" int bolap_q22(bolap_ctx* ctx, const bolap_q22_args* a, bolap_q22_result* out) {
    BOLAP_ARG_CHECK(ctx && a && out, "bolap_q22: null argument");
    BOLAP_ARG_CHECK(a->cust_rows >= 0 && a->order_rows >= 0, "bolap_q22: negative rows");
    BOLAP_ARG_CHECK(a->n_extra >= 0 && a->n_extra <= 7, "bolap_q22: %d extra prefixes",
                    a->n_extra);
    BOLAP_ARG_CHECK(a->cust_rows == 0 || (a->c_custkey && a->c_phone_code && a->c_phone_pair &&
                                          a->c_acctbal),
                    "bolap_q22: null customer column");
    BOLAP_ARG_CHECK(a->order_rows == 0 || a->o_custkey, "bolap_q22: null o_custkey");
    std::memset(out, 0, sizeof(*out));
    if (a->cust_rows == 0) return BOLAP_OK;
    Q22F f{};
    for (int c = 0; c < 100; ++c)
        if (a->numeric_allowed[c]) f.allowed[c >> 6] |= uint64_t(1) << (c & 63);
    f.n_extra = a->n_extra;
    for (int e = 0; e < a->n_extra; ++e) f.extra[e] = a->extra_pairs[e];
    int32_t* slot = static_cast<int32_t*>(ctx->result_dev);
    int32_t kmax = 0;
    BOLAP_TRY(max_i32(ctx, a->c_custkey, a->cust_rows, slot, &kmax));
    const int64_t size = kmax >= 0 ? int64_t(kmax) + 1 : 0;
    WsArena ws;
    const size_t o_has = ws.plan(size_t(size > 0 ? size : 1));
    const size_t o_tot = ws.plan((2 + 107 * 2) * 8);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    auto* has = ws.at<uint8_t>(o_has);
    auto* tot = ws.at<unsigned long long>(o_tot);
    BOLAP_TRY(zero_dev(ctx, has, ws.off));
    {
        TraceScope ts(ctx, "q22_customers");
        if (a->order_rows > 0 && size > 0) {
            hipLaunchKernelGGL(q22_mark_k, dim3(grid_of(ctx, a->order_rows, 8)), dim3(kT), 0,
                               ctx->stream, a->o_custkey, a->order_rows, size, has);
            BOLAP_TRY(launch_check("q22_mark_k"));
        }
        const unsigned g = grid_of(ctx, a->cust_rows, 4);
        hipLaunchKernelGGL(q22_avg_k, dim3(g), dim3(kT), 0, ctx->stream, f, a->cust_rows,
                           a->c_phone_code, a->c_phone_pair, a->c_acctbal, tot);
        hipLaunchKernelGGL(q22_agg_k, dim3(g), dim3(kT), 0, ctx->stream, f, a->cust_rows,
                           a->c_custkey, a->c_phone_code, a->c_phone_pair, a->c_acctbal, has, size,
                           tot);
        BOLAP_TRY(launch_check("q22_agg_k"));
    }
    uint64_t h[2 + 107 * 2];
    BOLAP_TRY(fetch(ctx, h, tot, sizeof(h)));
    for (int c = 0; c < 100; ++c) {
        out->numcust[c] = int64_t(h[2 + 2 * c]);
        out->totacctbal[c] = int64_t(h[3 + 2 * c]);
    }
    for (int e = 0; e < 7; ++e) {
        out->extra_numcust[e] = int64_t(h[2 + 2 * (100 + e)]);
        out->extra_totacctbal[e] = int64_t(h[3 + 2 * (100 + e)]);
    }
    return BOLAP_OK;
}

// ---------------------------------------------------------------------------
// Q7, Q8, Q9: the multi-way joins.
// ---------------------------------------------------------------------------
namespace bolap {
namespace {

// Q7 supplier tags: the last supplier row of a key with nation n1 (tag 0)
// or, when n1 != n2, n2 (tag 1): owner = (row + 1) << 1 | tag, 0 untagged.
__global__ __launch_bounds__(kT) void q7_tag_k(const int32_t* __restrict__ sk,
                                               const int32_t* __restrict__ sn, int64_t rows,
                                               int64_t n, int32_t n1, int32_t n2,
                                               int32_t* __restrict__ owner) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = sk[i];
        if (k < 0 || k >= n) continue;
        const int32_t nat = sn[i];
        int tag = -1;
        if (nat == n1) tag = 0;
        else if (n1 != n2 && nat == n2) tag = 1;
        if (tag >= 0) atomicMax(owner + k, int32_t((uint32_t(i + 1) << 1) | uint32_t(tag)));
    }
}

struct Q7K {
    int64_t o_rows;
    const int16_t* cnation;
    const int32_t* okey;
    LiSide li;
    const int16_t* ship;
    const int32_t* sk;
    const int32_t* dprice;
    const int32_t* owner;
    int64_t n;
    int32_t n1, n2, ship_lo;
    uint32_t ship_span;
    int32_t split;
};

__global__ __launch_bounds__(kT) void q7_k(Q7K a, unsigned long long* __restrict__ out) {
    int64_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const bool same = a.n1 == a.n2;
    for (int64_t o = int64_t(blockIdx.x) * kT + threadIdx.x; o < a.o_rows;
         o += int64_t(gridDim.x) * kT) {
        const int32_t cn = a.cnation[o];
        int ci;
        if (cn == a.n1) ci = 0;
        else if (!same && cn == a.n2) ci = 1;
        else continue;
        const uint32_t s = a.li.start[o], e = a.li.end[o];
        if (e <= s) continue;
        const int32_t key = a.li.okey ? a.okey[o] : 0;
        for (uint32_t r = s; r < e; ++r) {
            if (a.li.okey && a.li.okey[r] != key) continue;
            const int32_t ship = a.ship[r];
            if (uint32_t(ship - a.ship_lo) > a.ship_span) continue;
            const int32_t k = a.sk[r];
            if (k < 0 || k >= a.n) continue;
            const int32_t w = a.owner[k];
            if (w == 0) continue;
            const int si = w & 1;
            if (same ? (si != 0 || ci != 0) : si == ci) continue;
            const int idx = si * 4 + ci * 2 + (ship >= a.split ? 1 : 0);
            const int64_t v = a.dprice[r];
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[j] += j == idx ? v : 0;
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint64_t t = wave_sum_u64(uint64_t(acc[j]));
        if ((threadIdx.x & (kWave - 1)) == 0 && t) atomicAdd(out + j, (unsigned long long)t);
    }
}

// Marks keys of rows that pass (a byte per key; marks are never cleared).
__global__ __launch_bounds__(kT) void mark_keys_k(const int32_t* __restrict__ keys,
                                                  const void* __restrict__ col, int32_t dtype,
                                                  int64_t rows, int64_t n, int32_t lo,
                                                  int32_t hi, const uint8_t* __restrict__ lut,
                                                  int32_t lut_n, uint8_t* __restrict__ mark) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = keys[i];
        if (k < 0 || k >= n) continue;
        int32_t v;
        switch (dtype) {
            case BOLAP_U8: v = static_cast<const uint8_t*>(col)[i]; break;
            case BOLAP_I16: v = static_cast<const int16_t*>(col)[i]; break;
            default: v = static_cast<const int32_t*>(col)[i]; break;
        }
        // lut: value passes when lut[v] (v in [0, lut_n)); else lo <= v <= hi
        const bool pass = lut ? (v >= 0 && v < lut_n && lut[v]) : (v >= lo && v <= hi);
        if (pass) mark[k] = 1;
    }
}

struct Q8K {
    const int16_t* odate;
    int64_t o_rows;
    const int32_t* okey;
    const int32_t* ocust;
    LiSide li;
    const int32_t* pk;
    const int32_t* sk;
    const int32_t* price;
    const uint8_t* disc;
    const uint8_t *cmark, *smark, *pmark;
    int64_t nc, ns, np;
    int32_t date_lo, date_hi, split;
};

__global__ __launch_bounds__(kT) void q8_k(Q8K a, unsigned long long* __restrict__ out) {
    int64_t tot[2] = {0, 0}, nat[2] = {0, 0};
    const Window w = wg_window(a.odate, a.o_rows, a.date_lo, a.date_hi + 1);
    for (int64_t o = w.s + threadIdx.x; o < w.e; o += kT) {
        const int32_t ck = a.ocust[o];
        if (ck < 0 || ck >= a.nc || !a.cmark[ck]) continue;
        const uint32_t s = a.li.start[o], e = a.li.end[o];
        if (e <= s) continue;
        const int y = a.odate[o] >= a.split ? 1 : 0;
        const int32_t key = a.li.okey ? a.okey[o] : 0;
        for (uint32_t r = s; r < e; ++r) {
            if (a.li.okey && a.li.okey[r] != key) continue;
            const int32_t p = a.pk[r];
            if (p < 0 || p >= a.np || !a.pmark[p]) continue;
            const int64_t vol = int64_t(a.price[r]) * int64_t(100 - int32_t(a.disc[r]));
            const int32_t k = a.sk[r];
            const bool hit = k >= 0 && k < a.ns && a.smark[k];
            if (y) {
                tot[1] += vol;
                nat[1] += hit ? vol : 0;
            } else {
                tot[0] += vol;
                nat[0] += hit ? vol : 0;
            }
        }
    }
    const int64_t v[4] = {tot[0], tot[1], nat[0], nat[1]};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint64_t t = wave_sum_u64(uint64_t(v[j]));
        if ((threadIdx.x & (kWave - 1)) == 0 && t) atomicAdd(out + j, (unsigned long long)t);
    }
}

// Q9: partsupp rows grouped by partkey.  The builder's layout keeps
// partsupp sorted by partkey; then the offsets come straight from the run
// boundaries and a key's rows are the contiguous range off[k]..off[k + 1].
// Row i writes off[k] = i for the keys k in (pk[i - 1], pk[i]] within
// [0, n], so each offset has one writer; a descent anywhere sets *unsorted.
__global__ __launch_bounds__(kT) void ps_offsets_k(const int32_t* __restrict__ pk, int64_t rows,
                                                   int64_t n, uint32_t* __restrict__ off,
                                                   uint32_t* __restrict__ unsorted) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i <= rows;
         i += int64_t(gridDim.x) * kT) {
        const int64_t prev = i == 0 ? INT64_MIN / 2 : int64_t(pk[i - 1]);
        const int64_t cur = i == rows ? INT64_MAX / 2 : int64_t(pk[i]);
        if (cur < prev) *unsorted = 1u;
        const int64_t lo = prev + 1 > 0 ? prev + 1 : 0;
        const int64_t hi = cur < n ? cur : n;
        for (int64_t k = lo; k <= hi; ++k) off[k] = uint32_t(i);
    }
}

// Otherwise a CSR in three steps (counts, scan, fill); a key's rows land in
// any order and the lookup takes the smallest row of a match, which is the
// reference's first match in row order.
__global__ __launch_bounds__(kT) void ps_count_k(const int32_t* __restrict__ pk, int64_t rows,
                                                 int64_t n, uint32_t* __restrict__ cnt) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = pk[i];
        if (k >= 0 && k < n) atomicAdd(cnt + k, 1u);
    }
}

// Exclusive scan of n u32 counts in place (tiles of kTile per workgroup:
// tile sums, scan_tiles_k over them, then each tile's local scan).
__global__ __launch_bounds__(kT) void tile_sum_k(const uint32_t* __restrict__ v, int64_t n,
                                                 uint32_t* __restrict__ tiles) {
    const int64_t t0 = int64_t(blockIdx.x) * kTile;
    uint32_t s = 0;
    for (int j = 0; j < 16; ++j) {
        const int64_t i = t0 + int64_t(j) * kT + threadIdx.x;
        if (i < n) s += v[i];
    }
    __shared__ uint32_t sh;
    if (threadIdx.x == 0) sh = 0;
    __syncthreads();
    atomicAdd(&sh, s);
    __syncthreads();
    if (threadIdx.x == 0) tiles[blockIdx.x] = sh;
}

__global__ __launch_bounds__(kT) void tile_scan_k(uint32_t* __restrict__ v, int64_t n,
                                                  const uint32_t* __restrict__ tiles) {
    __shared__ uint32_t s_w[kT / kWave];
    const int64_t t0 = int64_t(blockIdx.x) * kTile;
    uint32_t base = tiles[blockIdx.x];
    for (int j = 0; j < 16; ++j) {
        const int64_t i = t0 + int64_t(j) * kT + threadIdx.x;
        const uint32_t x = i < n ? v[i] : 0;
        uint32_t inc = x;
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t y = __shfl_up(inc, d, kWave);
            if ((threadIdx.x & (kWave - 1)) >= d) inc += y;
        }
        if ((threadIdx.x & (kWave - 1)) == kWave - 1) s_w[threadIdx.x / kWave] = inc;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int w = 0; w < kT / kWave; ++w) {
            if (w < int(threadIdx.x / kWave)) before += s_w[w];
            total += s_w[w];
        }
        if (i < n) v[i] = base + before + inc - x;
        base += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kT) void ps_fill_k(const int32_t* __restrict__ pk, int64_t rows,
                                                int64_t n, uint32_t* __restrict__ cursor,
                                                uint32_t* __restrict__ idx) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = pk[i];
        if (k >= 0 && k < n) idx[atomicAdd(cursor + k, 1u)] = uint32_t(i);
    }
}

// Year of a day offset from 1992-01-01 (Hinnant's civil_from_days).
BOLAP_DEV int32_t year_of_offset(int32_t off) {
    const int32_t z = off + 8035 + 719468;  // 1992-01-01 is day 8035 since 1970-01-01
    const int32_t era = (z >= 0 ? z : z - 146096) / 146097;
    const uint32_t doe = uint32_t(z - era * 146097);
    const uint32_t yoe = (doe - doe / 1460u + doe / 36524u - doe / 146096u) / 365u;
    const uint32_t doy = doe - (365u * yoe + yoe / 4u - yoe / 100u);
    const uint32_t mp = (5u * doy + 2u) / 153u;
    const uint32_t m = mp < 10u ? mp + 3u : mp - 9u;
    return int32_t(yoe) + era * 400 + (m <= 2u ? 1 : 0);
}

constexpr int kQ9Groups = 2048;

struct Q9K {
    int64_t rows;
    const int32_t* pk;
    const int32_t* sk;
    const int32_t* ok;
    const int16_t* qty;
    const int32_t* dprice;
    const int16_t* snation;
    const uint8_t* pmark;
    int64_t np;
    const uint32_t* ps_off;  // np + 1 offsets
    const uint32_t* ps_idx;
    const int32_t* ps_sk;
    const int32_t* ps_cost;
    const int32_t* omap;
    int64_t map_size;
    const int16_t* odate;
    int32_t n_nation;
    int32_t min_year, span;
};

BOLAP_DEV void q9_row(const Q9K& a, int64_t i, int32_t p, unsigned long long* s_tot,
                      uint8_t* s_used) {
    const int32_t k = a.sk[i];
    uint32_t best = 0xffffffffu;
    if (a.ps_idx) {
        for (uint32_t j = a.ps_off[p]; j < a.ps_off[p + 1]; ++j) {
            const uint32_t row = a.ps_idx[j];
            if (a.ps_sk[row] == k && row < best) best = row;
        }
    } else {  // sorted partsupp: the first match in the key's range
        for (uint32_t j = a.ps_off[p]; j < a.ps_off[p + 1]; ++j) {
            if (a.ps_sk[j] == k) {
                best = j;
                break;
            }
        }
    }
    if (best == 0xffffffffu) return;
    const int32_t nk = a.snation[i];
    if (nk < 0 || nk >= a.n_nation) return;
    const int32_t okey = a.ok[i];
    if (okey < 0 || okey >= a.map_size) return;
    const int32_t orow = a.omap[okey];
    if (orow < 0) return;
    const int32_t yi = year_of_offset(a.odate[orow]) - a.min_year;
    const int64_t profit =
        int64_t(a.dprice[i]) - int64_t(a.ps_cost[best]) * int64_t(a.qty[i]) / 100;
    const int g = nk * a.span + yi;
    atomicAdd(s_tot + g, (unsigned long long)profit);
    s_used[g] = 1;
}

// A workgroup takes tiles of kQ9Tile lineitem rows: the rows of marked parts
// are gathered into LDS first (about one in twenty for one colour), then
// every lane of the workgroup takes one of them through the partsupp,
// order and date lookups, so the dependent loads run with full waves.
constexpr int kQ9Tile = kT * 16;

__global__ __launch_bounds__(kT) void q9_k(Q9K a, unsigned long long* __restrict__ tot,
                                           uint8_t* __restrict__ used) {
    __shared__ unsigned long long s_tot[kQ9Groups];
    __shared__ uint8_t s_used[kQ9Groups];
    __shared__ uint32_t s_rows[kQ9Tile];
    __shared__ unsigned long long s_n;
    const int groups = a.n_nation * a.span;
    for (int g = threadIdx.x; g < groups; g += kT) {
        s_tot[g] = 0;
        s_used[g] = 0;
    }
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int64_t tiles = (a.rows + kQ9Tile - 1) / kQ9Tile;
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t base = t * kQ9Tile;
#pragma unroll 4
        for (int j = 0; j < kQ9Tile / kT; ++j) {
            const int64_t i = base + int64_t(j) * kT + threadIdx.x;
            bool hit = false;
            if (i < a.rows) {
                const int32_t p = a.pk[i];
                hit = p >= 0 && p < a.np && a.pmark[p];
            }
            const unsigned long long slot = wave_append(hit, &s_n);
            if (hit) s_rows[slot] = uint32_t(i - base);
        }
        __syncthreads();
        const uint32_t n = uint32_t(s_n);
        for (uint32_t x = threadIdx.x; x < n; x += kT) {
            const int64_t i = base + s_rows[x];
            q9_row(a, i, a.pk[i], s_tot, s_used);
        }
        __syncthreads();
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
    }
    for (int g = threadIdx.x; g < groups; g += kT) {
        if (s_tot[g]) atomicAdd(tot + g, s_tot[g]);
        if (s_used[g]) used[g] = 1;
    }
}

}  // namespace
}  // namespace bolap

extern "C" int bolap_q7(bolap_ctx* ctx, const bolap_q7_args* a, int64_t* rev) {
    BOLAP_ARG_CHECK(ctx && a && rev, "bolap_q7: null argument");
    const bolap_order_side& o = a->orders;
    BOLAP_ARG_CHECK(o.order_rows >= 0 && o.order_rows < (int64_t(1) << 31),
                    "bolap_q7: order rows");
    BOLAP_ARG_CHECK(o.order_rows == 0 || (a->o_cust_nation && o.range_start && o.range_end &&
                                          a->l_shipdate && a->l_suppkey && a->l_disc_price),
                    "bolap_q7: null column");
    BOLAP_ARG_CHECK(o.l_orderkey_sorted || o.order_rows == 0 || (o.l_orderkey && o.o_orderkey),
                    "bolap_q7: unsorted lineitem needs l_orderkey and o_orderkey");
    BOLAP_ARG_CHECK(a->supp_rows >= 0 && a->supp_rows < (int64_t(1) << 30) &&
                        (a->supp_rows == 0 || (a->s_suppkey && a->s_nationkey)),
                    "bolap_q7: supplier columns");
    BOLAP_ARG_CHECK(a->n_suppkeys >= 0 && a->n_suppkeys < (int64_t(1) << 31),
                    "bolap_q7: n_suppkeys");
    for (int i = 0; i < 8; ++i) rev[i] = 0;
    if (o.order_rows == 0 || a->n_suppkeys == 0) return BOLAP_OK;
    WsArena ws;
    const size_t o_owner = ws.plan(size_t(a->n_suppkeys) * 4);
    const size_t o_out = ws.plan(8 * 8);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    auto* owner = ws.at<int32_t>(o_owner);
    auto* out = ws.at<unsigned long long>(o_out);
    BOLAP_TRY(zero_dev(ctx, owner, ws.off));
    TraceScope ts(ctx, "q7_orders");
    if (a->supp_rows > 0) {
        hipLaunchKernelGGL(q7_tag_k, dim3(grid_of(ctx, a->supp_rows)), dim3(kT), 0, ctx->stream,
                           a->s_suppkey, a->s_nationkey, a->supp_rows, a->n_suppkeys, a->nation1,
                           a->nation2, owner);
        BOLAP_TRY(launch_check("q7_tag_k"));
    }
    const Q7K k{o.order_rows, a->o_cust_nation, o.o_orderkey, li_side(&o), a->l_shipdate,
                a->l_suppkey, a->l_disc_price, owner, a->n_suppkeys, a->nation1, a->nation2,
                a->ship_lo, uint32_t(a->ship_hi - a->ship_lo), a->split};
    hipLaunchKernelGGL(q7_k, dim3(grid_of(ctx, o.order_rows, 8)), dim3(kT), 0, ctx->stream, k, out);
    BOLAP_TRY(launch_check("q7_k"));
    return fetch(ctx, rev, out, 8 * 8);
}

extern "C" int bolap_q8(bolap_ctx* ctx, const bolap_q8_args* a, int64_t* total, int64_t* nation) {
    BOLAP_ARG_CHECK(ctx && a && total && nation, "bolap_q8: null argument");
    BOLAP_TRY(check_ranges(&a->orders, "bolap_q8"));
    const bolap_order_side& o = a->orders;
    BOLAP_ARG_CHECK(o.order_rows == 0 || (a->o_custkey && a->l_partkey && a->l_suppkey &&
                                          a->l_price && a->l_disc),
                    "bolap_q8: null column");
    BOLAP_ARG_CHECK(a->n_nation >= 0 && a->n_nation <= 256, "bolap_q8: n_nation %d", a->n_nation);
    BOLAP_ARG_CHECK(a->n_cust_keys >= 0 && a->n_supp_keys >= 0 && a->n_part_keys >= 0,
                    "bolap_q8: key domains");
    BOLAP_ARG_CHECK((a->cust_rows == 0 || (a->c_custkey && a->c_nationkey)) &&
                        (a->supp_rows == 0 || (a->s_suppkey && a->s_nationkey)) &&
                        (a->part_rows == 0 || (a->p_partkey && a->p_type)),
                    "bolap_q8: null dimension column");
    total[0] = total[1] = nation[0] = nation[1] = 0;
    if (o.order_rows == 0) return BOLAP_OK;
    WsArena ws;
    const size_t o_c = ws.plan(size_t(a->n_cust_keys) + 1);
    const size_t o_s = ws.plan(size_t(a->n_supp_keys) + 1);
    const size_t o_p = ws.plan(size_t(a->n_part_keys) + 1);
    const size_t o_out = ws.plan(4 * 8);
    const size_t o_lut = ws.plan(256);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    BOLAP_TRY(zero_dev(ctx, ws.base, o_lut));
    auto* lut = ws.at<uint8_t>(o_lut);
    BOLAP_HIP_TRY(hipMemcpyAsync(lut, a->nation_in_region, size_t(a->n_nation),
                                 hipMemcpyHostToDevice, ctx->stream));
    TraceScope ts(ctx, "q8_orders");
    if (a->cust_rows > 0)
        hipLaunchKernelGGL(mark_keys_k, dim3(grid_of(ctx, a->cust_rows)), dim3(kT), 0, ctx->stream,
                           a->c_custkey, static_cast<const void*>(a->c_nationkey), int32_t(BOLAP_I32),
                           a->cust_rows, a->n_cust_keys, 0, -1, lut, a->n_nation,
                           ws.at<uint8_t>(o_c));
    if (a->supp_rows > 0)
        hipLaunchKernelGGL(mark_keys_k, dim3(grid_of(ctx, a->supp_rows)), dim3(kT), 0, ctx->stream,
                           a->s_suppkey, static_cast<const void*>(a->s_nationkey), int32_t(BOLAP_I32),
                           a->supp_rows, a->n_supp_keys, a->target_nation, a->target_nation,
                           static_cast<const uint8_t*>(nullptr), 0, ws.at<uint8_t>(o_s));
    if (a->part_rows > 0 && a->type_code >= 0)
        hipLaunchKernelGGL(mark_keys_k, dim3(grid_of(ctx, a->part_rows)), dim3(kT), 0, ctx->stream,
                           a->p_partkey, static_cast<const void*>(a->p_type), int32_t(BOLAP_U8),
                           a->part_rows, a->n_part_keys, a->type_code, a->type_code,
                           static_cast<const uint8_t*>(nullptr), 0, ws.at<uint8_t>(o_p));
    BOLAP_TRY(launch_check("mark_keys_k"));
    const Q8K k{o.o_orderdate, o.order_rows, o.o_orderkey, a->o_custkey, li_side(&o), a->l_partkey,
                a->l_suppkey, a->l_price, a->l_disc, ws.at<uint8_t>(o_c), ws.at<uint8_t>(o_s),
                ws.at<uint8_t>(o_p), a->n_cust_keys, a->n_supp_keys, a->n_part_keys, a->date_lo,
                a->date_hi, a->split};
    auto* out = ws.at<unsigned long long>(o_out);
    hipLaunchKernelGGL(q8_k, dim3(window_grid(ctx, o.order_rows)), dim3(kT), 0, ctx->stream, k, out);
    BOLAP_TRY(launch_check("q8_k"));
    int64_t h[4];
    BOLAP_TRY(fetch(ctx, h, out, sizeof(h)));
    total[0] = h[0];
    total[1] = h[1];
    nation[0] = h[2];
    nation[1] = h[3];
    return BOLAP_OK;
}

extern "C" int bolap_q9(bolap_ctx* ctx, const bolap_q9_args* a, int64_t* totals, uint8_t* used,
                        int32_t* min_year, int32_t* year_span) {
    BOLAP_ARG_CHECK(ctx && a && totals && used && min_year && year_span,
                    "bolap_q9: null argument");
    BOLAP_ARG_CHECK(a->part_rows >= 0 && a->ps_rows >= 0 && a->order_rows >= 0 && a->li_rows >= 0 &&
                        a->ps_rows < (int64_t(1) << 32),
                    "bolap_q9: rows out of range");
    BOLAP_ARG_CHECK(a->part_rows == 0 || (a->p_partkey && a->p_ok), "bolap_q9: null part column");
    BOLAP_ARG_CHECK(a->ps_rows == 0 || (a->ps_partkey && a->ps_suppkey && a->ps_supplycost),
                    "bolap_q9: null partsupp column");
    BOLAP_ARG_CHECK(a->order_rows == 0 || (a->o_orderdate && a->order_map),
                    "bolap_q9: null order column");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (a->l_partkey && a->l_suppkey && a->l_orderkey &&
                                        a->l_quantity && a->l_disc_price && a->l_supp_nation &&
                                        aligned(a->l_partkey, 16)),
                    "bolap_q9: null lineitem column (l_partkey 16-byte aligned)");
    BOLAP_ARG_CHECK(a->n_nation >= 1 && a->n_part_keys >= 0, "bolap_q9: n_nation / n_part_keys");
    *min_year = 0;
    *year_span = 0;
    if (a->order_rows == 0) return BOLAP_OK;
    // earliest and latest orderdate: the first and last rows (sorted by date)
    int16_t ends[2];
    BOLAP_TRY(fetch(ctx, &ends[0], a->o_orderdate, 2));
    BOLAP_TRY(fetch(ctx, &ends[1], a->o_orderdate + a->order_rows - 1, 2));
    auto year_of = [](int32_t off) {
        int64_t z = int64_t(off) + 8035 + 719468;
        const int64_t era = (z >= 0 ? z : z - 146096) / 146097;
        const int64_t doe = z - era * 146097;
        const int64_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
        const int64_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
        const int64_t mp = (5 * doy + 2) / 153;
        const int64_t m = mp < 10 ? mp + 3 : mp - 9;
        return int32_t(yoe + era * 400 + (m <= 2 ? 1 : 0));
    };
    const int32_t y0 = year_of(ends[0]), span = year_of(ends[1]) - y0 + 1;
    BOLAP_ARG_CHECK(span >= 1 && int64_t(span) * a->n_nation <= kQ9Groups,
                    "bolap_q9: %d nations x %d years exceed %d groups", a->n_nation, span,
                    kQ9Groups);
    *min_year = y0;
    *year_span = span;
    const int64_t np = a->n_part_keys;
    const int64_t tiles = (np + 1 + kTile - 1) / kTile;  // np + 1 offsets
    WsArena ws;
    const size_t o_mark = ws.plan(size_t(np) + 1);
    const size_t o_off = ws.plan(size_t(np + 1) * 4);
    const size_t o_tot = ws.plan(size_t(kQ9Groups) * 8);
    const size_t o_used = ws.plan(size_t(kQ9Groups));
    const size_t o_flag = ws.plan(4);
    const size_t o_zero_end = ws.off;
    const size_t o_cur = ws.plan(size_t(np + 1) * 4);
    const size_t o_tiles = ws.plan(size_t(tiles + 1) * 4);
    const size_t o_idx = ws.plan(size_t(a->ps_rows > 0 ? a->ps_rows : 1) * 4);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    BOLAP_TRY(zero_dev(ctx, ws.base, o_zero_end));
    auto* off = ws.at<uint32_t>(o_off);
    auto* cur = ws.at<uint32_t>(o_cur);
    auto* tl = ws.at<uint32_t>(o_tiles);
    uint32_t unsorted = 0;
    {
        TraceScope ts(ctx, "q9_build");
        if (a->part_rows > 0 && np > 0)
            hipLaunchKernelGGL(mark_keys_k, dim3(grid_of(ctx, a->part_rows)), dim3(kT), 0,
                               ctx->stream, a->p_partkey, static_cast<const void*>(a->p_ok),
                               int32_t(BOLAP_U8), a->part_rows, np, 1, 1,
                               static_cast<const uint8_t*>(nullptr), 0, ws.at<uint8_t>(o_mark));
        if (a->ps_rows > 0 && np > 0) {
            hipLaunchKernelGGL(ps_offsets_k, dim3(grid_of(ctx, a->ps_rows + 1)), dim3(kT), 0,
                               ctx->stream, a->ps_partkey, a->ps_rows, np, off,
                               ws.at<uint32_t>(o_flag));
            BOLAP_TRY(launch_check("ps_offsets_k"));
            BOLAP_TRY(fetch(ctx, &unsorted, ws.at<uint32_t>(o_flag), 4));
        }
        if (a->ps_rows > 0 && np > 0 && unsorted) {
            BOLAP_TRY(zero_dev(ctx, off, size_t(np + 1) * 4));
            hipLaunchKernelGGL(ps_count_k, dim3(grid_of(ctx, a->ps_rows)), dim3(kT), 0, ctx->stream,
                               a->ps_partkey, a->ps_rows, np, off);
            hipLaunchKernelGGL(tile_sum_k, dim3(unsigned(tiles)), dim3(kT), 0, ctx->stream, off,
                               np + 1, tl);
            hipLaunchKernelGGL(scan_tiles_k, dim3(1), dim3(1024), 0, ctx->stream, tl, tiles);
            hipLaunchKernelGGL(tile_scan_k, dim3(unsigned(tiles)), dim3(kT), 0, ctx->stream, off,
                               np + 1, tl);
            BOLAP_HIP_TRY(hipMemcpyAsync(cur, off, size_t(np + 1) * 4, hipMemcpyDeviceToDevice,
                                         ctx->stream));
            hipLaunchKernelGGL(ps_fill_k, dim3(grid_of(ctx, a->ps_rows)), dim3(kT), 0, ctx->stream,
                               a->ps_partkey, a->ps_rows, np, cur, ws.at<uint32_t>(o_idx));
        }
        BOLAP_TRY(launch_check("ps_fill_k"));
    }
    auto* tot = ws.at<unsigned long long>(o_tot);
    auto* usd = ws.at<uint8_t>(o_used);
    if (a->li_rows > 0 && np > 0) {
        TraceScope ts(ctx, "q9_lineitem");
        const Q9K k{a->li_rows, a->l_partkey, a->l_suppkey, a->l_orderkey, a->l_quantity,
                    a->l_disc_price, a->l_supp_nation, ws.at<uint8_t>(o_mark), np, off,
                    unsorted ? ws.at<uint32_t>(o_idx) : nullptr, a->ps_suppkey, a->ps_supplycost, a->order_map,
                    a->map_size, a->o_orderdate, a->n_nation, y0, span};
        hipLaunchKernelGGL(q9_k, dim3(grid_of(ctx, (a->li_rows + 15) / 16, 4)), dim3(kT), 0,
                           ctx->stream, k, tot, usd);
        BOLAP_TRY(launch_check("q9_k"));
    }
    const size_t g = size_t(a->n_nation) * size_t(span);
    BOLAP_TRY(fetch(ctx, totals, tot, g * 8));
    return fetch(ctx, used, usd, g);
}

// ---------------------------------------------------------------------------
// Q13 (string prefilters and substring search) and Q16 (count distinct).
// ---------------------------------------------------------------------------
namespace bolap {
namespace {

// The builder's per-string prefilter masks (builder_impl.cpp:124-150): a bit
// per letter a-z (either case) and 64 bits of (c0 * 131 + c1) & 63 over the
// byte pairs.
BOLAP_DEV uint32_t alpha_bits(uint8_t c) {
    if (c >= 'a' && c <= 'z') return 1u << (c - 'a');
    if (c >= 'A' && c <= 'Z') return 1u << (c - 'A');
    return 0;
}

__global__ __launch_bounds__(kT) void str_masks_k(const uint8_t* __restrict__ chars,
                                                  const uint32_t* __restrict__ off, int64_t n,
                                                  uint32_t* __restrict__ alpha,
                                                  uint64_t* __restrict__ bigram) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const uint32_t s = off[i], e = off[i + 1];
        uint32_t a = 0;
        uint64_t b = 0;
        uint32_t prev = 0;
        for (uint32_t p = s; p < e; ++p) {
            const uint32_t c = chars[p];
            a |= alpha_bits(uint8_t(c));
            if (p > s) b |= uint64_t(1) << ((prev * 131u + c) & 63u);
            prev = c;
        }
        alpha[i] = a;
        bigram[i] = b;
    }
}

struct Q13W {
    uint8_t w1[64], w2[64];
    int32_t n1, n2;
    uint32_t req_alpha;
    uint64_t req_bigram;
    // each word's first min(4, n) bytes packed little-endian (pre), their
    // count (m) and the right shift that aligns a 4-byte window to them (sh)
    uint32_t pre1, pre2;
    int32_t m1, m2, sh1, sh2;
};

// First position p in [s, e) where w occurs (s itself for an empty word),
// e when none (find_substring, query_q13.cpp:27-49).
BOLAP_DEV uint32_t find_word(const uint8_t* __restrict__ chars, uint32_t s, uint32_t e,
                             const uint8_t* w, int32_t n) {
    if (n == 0) return s;
    for (uint32_t p = s; p + uint32_t(n) <= e; ++p) {
        if (chars[p] != w[0]) continue;
        int32_t j = 1;
        while (j < n && chars[p + j] == w[j]) ++j;
        if (j == n) return p;
    }
    return e;
}

BOLAP_DEV bool tail_eq(const uint8_t* __restrict__ c, const uint8_t* w, int32_t from, int32_t n) {
    for (int32_t j = from; j < n; ++j)
        if (c[j] != w[j]) return false;
    return true;
}

// The two searches of the reference (query_q13.cpp:178-186): p1 = first
// occurrence of word1 in [s, e) (s for an empty word), then an occurrence of
// word2 in [p1 + n1, e); a match needs p1 != e and the second result != e.
// Strings of up to kRegBytes bytes (with their alignment lead) are loaded
// into registers once and scanned with a rolling 4-byte window against each
// word's first bytes, so no byte load waits in the loop; a window hit checks
// the word's remaining bytes in memory.  Longer strings take find_word.
constexpr int kRegWords = 24;
constexpr int kRegBytes = kRegWords * 4;

BOLAP_DEV bool comment_match(const Q13W& w, const uint8_t* __restrict__ chars, uint32_t s,
                             uint32_t e) {
    const int32_t len = int32_t(e - s);
    const uintptr_t at = reinterpret_cast<uintptr_t>(chars + s);
    const int32_t lead = int32_t(at & 3u);
    if (len + lead > kRegBytes) {
        const uint32_t p1 = find_word(chars, s, e, w.w1, w.n1);
        return p1 != e && find_word(chars, p1 + uint32_t(w.n1), e, w.w2, w.n2) != e;
    }
    if (len == 0) return false;  // p1 == e whatever the words
    // aligned words holding at least one byte of the string (never past its page)
    const uint32_t* wp = reinterpret_cast<const uint32_t*>(at - uintptr_t(lead));
    const int32_t nw = (lead + len + 3) >> 2;
    uint32_t W[kRegWords];
#pragma unroll
    for (int k = 0; k < kRegWords; ++k) W[k] = k < nw ? wp[k] : 0u;
    const uint8_t* str = chars + s;
    int32_t first1 = w.n1 == 0 ? 0 : -1;
    bool found = false;
    uint32_t win = 0;
#pragma unroll
    for (int k = 0; k < kRegWords; ++k) {
        if (k >= nw || found) break;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            win = (win >> 8) | ((W[k] >> (8 * b)) << 24);
            const int32_t r = 4 * k + b - lead;  // window ends at str[r]
            if (r < 0 || r >= len) continue;
            if (first1 < 0 && (win >> w.sh1) == w.pre1) {
                const int32_t st = r - w.m1 + 1;
                if (st >= 0 && st + w.n1 <= len && tail_eq(str + st, w.w1, w.m1, w.n1))
                    first1 = st;
            }
            if (w.n2 > 0 && first1 >= 0 && (win >> w.sh2) == w.pre2) {
                const int32_t st = r - w.m2 + 1;
                if (st >= first1 + w.n1 && st + w.n2 <= len &&
                    tail_eq(str + st, w.w2, w.m2, w.n2))
                    found = true;
            }
        }
    }
    if (w.n2 == 0) return first1 >= 0 && first1 + w.n1 < len;
    return found;
}

// Per-customer order counts of the orders whose comment does not match.  A
// workgroup takes tiles of kQ13Tile orders: orders the prefilter masks rule
// out are counted at once; the candidates are gathered into LDS and then
// searched one per lane, so the string scans run with full waves.
constexpr int kQ13Tile = kT * 8;

__global__ __launch_bounds__(kT) void q13_k(Q13W w, int64_t rows, const int32_t* __restrict__ ock,
                                            const uint8_t* __restrict__ chars,
                                            const uint32_t* __restrict__ off,
                                            const uint32_t* __restrict__ alpha,
                                            const uint64_t* __restrict__ bigram, int64_t size,
                                            int32_t* __restrict__ cnt) {
    __shared__ uint32_t s_c[kQ13Tile];
    __shared__ unsigned long long s_n;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int64_t tiles = (rows + kQ13Tile - 1) / kQ13Tile;
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t base = t * kQ13Tile;
#pragma unroll 2
        for (int j = 0; j < kQ13Tile / kT; ++j) {
            const int64_t i = base + int64_t(j) * kT + threadIdx.x;
            bool cand = false;
            if (i < rows) {
                cand = (!alpha || (alpha[i] & w.req_alpha) == w.req_alpha) &&
                       (!bigram || (bigram[i] & w.req_bigram) == w.req_bigram);
                if (!cand) {
                    const int32_t k = ock[i];
                    if (k >= 0 && k < size) atomicAdd(cnt + k, 1);
                }
            }
            const unsigned long long slot = wave_append(cand, &s_n);
            if (cand) s_c[slot] = uint32_t(i - base);
        }
        __syncthreads();
        const uint32_t n = uint32_t(s_n);
        for (uint32_t x = threadIdx.x; x < n; x += kT) {
            const int64_t i = base + s_c[x];
            if (comment_match(w, chars, off[i], off[i + 1])) continue;
            const int32_t k = ock[i];
            if (k >= 0 && k < size) atomicAdd(cnt + k, 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kT) void max_i32_arr_k(const int32_t* __restrict__ v, int64_t n,
                                                    int32_t* __restrict__ out) {
    int32_t m = 0;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        m = v[i] > m ? v[i] : m;
    m = block_max_i32(m);
    if (threadIdx.x == 0 && m) atomicMax(out, m);
}

constexpr int kDistLds = 4096;

__global__ __launch_bounds__(kT) void q13_dist_k(const int32_t* __restrict__ ck, int64_t rows,
                                                 const int32_t* __restrict__ cnt, int64_t size,
                                                 int32_t n_dist,
                                                 unsigned long long* __restrict__ dist) {
    __shared__ unsigned s_d[kDistLds];
    const bool lds = n_dist <= kDistLds;
    if (lds)
        for (int i = threadIdx.x; i < n_dist; i += kT) s_d[i] = 0;
    __syncthreads();
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = ck[i];
        const int32_t c = (k >= 0 && k < size) ? cnt[k] : 0;
        if (c >= n_dist) continue;
        if (lds) atomicAdd(&s_d[c], 1u);
        else atomicAdd(dist + c, 1ull);
    }
    __syncthreads();
    if (lds)
        for (int i = threadIdx.x; i < n_dist; i += kT)
            if (s_d[i]) atomicAdd(dist + i, (unsigned long long)s_d[i]);
}

// Q16: part key -> (row + 1) << 32 | group of the last eligible part row.
struct Q16P {
    int64_t rows;
    const int32_t* pk;
    const uint8_t* brand;
    const uint8_t* type;
    const int32_t* size;
    int32_t n_types;
    int32_t banned_brand;
    uint64_t size_mask;
    const uint8_t* type_banned;  // device, n_types bytes
    int64_t n;
};

__global__ __launch_bounds__(kT) void q16_part_k(Q16P a, unsigned long long* __restrict__ owner) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < a.rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t b = a.brand[i];
        if (b == a.banned_brand) continue;
        const int32_t sz = a.size[i];
        if (sz < 0 || sz > 50 || !((a.size_mask >> sz) & 1)) continue;
        const int32_t t = a.type[i];
        if (t >= a.n_types || a.type_banned[t]) continue;
        const int32_t k = a.pk[i];
        if (k < 0 || k >= a.n) continue;
        const uint64_t g = (uint64_t(b) * uint64_t(a.n_types) + uint64_t(t)) * 51u + uint64_t(sz);
        atomicMax(owner + k, (unsigned long long)((uint64_t(i + 1) << 32) | g));
    }
}

// The (group, suppkey) keys of the qualifying partsupp rows, in any order
// (they are sorted next).  Keys are staged in LDS and appended a stage at a
// time, so the output counter takes one atomic per flush.
constexpr int kKeyStage = 2048;

__global__ __launch_bounds__(kT) void q16_ps_k(int64_t rows, const int32_t* __restrict__ pk,
                                               const int32_t* __restrict__ sk, int64_t np,
                                               const unsigned long long* __restrict__ owner,
                                               const uint8_t* __restrict__ banned, int64_t ns,
                                               uint64_t* __restrict__ keys,
                                               unsigned long long* __restrict__ n_out) {
    __shared__ uint64_t s_keys[kKeyStage];
    __shared__ unsigned long long s_n, s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int64_t stride_g = int64_t(gridDim.x) * kT;
    const int64_t rounds = (rows + stride_g - 1) / stride_g;
    for (int64_t r = 0; r < rounds; ++r) {
        const int64_t i = r * stride_g + int64_t(blockIdx.x) * kT + threadIdx.x;
        bool hit = false;
        uint64_t key = 0;
        if (i < rows) {
            const int32_t k = pk[i];
            const unsigned long long w = (k >= 0 && k < np) ? owner[k] : 0ull;
            if (w != 0) {
                const int32_t s = sk[i];
                hit = !(s >= 0 && s < ns && banned[s]);
                key = ((w & 0xffffffffull) << 32) | uint32_t(s);
            }
        }
        const unsigned long long slot = wave_append(hit, &s_n);
        if (hit) s_keys[slot] = key;
        __syncthreads();
        const uint32_t cnt = uint32_t(s_n);
        __syncthreads();
        if (cnt > uint32_t(kKeyStage - kT))
            flush_stage<kT>(s_keys, &s_n, &s_base, cnt, keys, n_out);
    }
    const uint32_t cnt = uint32_t(s_n);
    __syncthreads();
    flush_stage<kT>(s_keys, &s_n, &s_base, cnt, keys, n_out);
}

// Distinct (group, suppkey) keys counted per group: each key is inserted
// into an open-addressing set (slots hold key + 1, zero is empty; at most
// half full, so every probe ends) and the insert that places it counts it.
__global__ __launch_bounds__(kT) void q16_distinct_k(const uint64_t* __restrict__ keys, int64_t n,
                                                     unsigned long long* __restrict__ set,
                                                     uint64_t mask,
                                                     unsigned long long* __restrict__ counts) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const uint64_t key = keys[i];
        const unsigned long long tag = key + 1;
        uint64_t slot = mix64(key) & mask;
        for (;;) {
            const unsigned long long prev = atomicCAS(set + slot, 0ull, tag);
            if (prev == 0ull) {
                atomicAdd(counts + (key >> 32), 1ull);
                break;
            }
            if (prev == tag) break;
            slot = (slot + 1) & mask;
        }
    }
}

}  // namespace
}  // namespace bolap

extern "C" int bolap_string_masks(bolap_ctx* ctx, const uint8_t* chars, const uint32_t* offsets,
                                  int64_t n, uint32_t* alpha_out, uint64_t* bigram_out) {
    BOLAP_ARG_CHECK(ctx && n >= 0 && (n == 0 || (chars && offsets && alpha_out && bigram_out)),
                    "bolap_string_masks: null argument");
    if (n == 0) return BOLAP_OK;
    hipLaunchKernelGGL(str_masks_k, dim3(grid_of(ctx, n, 8)), dim3(kT), 0, ctx->stream, chars,
                       offsets, n, alpha_out, bigram_out);
    return launch_check("str_masks_k");
}

extern "C" int bolap_q13(bolap_ctx* ctx, const bolap_q13_args* a, int64_t* dist, int32_t max_dist,
                         int32_t* max_count) {
    BOLAP_ARG_CHECK(ctx && a && max_count && (dist || max_dist == 0), "bolap_q13: null argument");
    BOLAP_ARG_CHECK(a->order_rows >= 0 && a->cust_rows >= 0, "bolap_q13: negative rows");
    BOLAP_ARG_CHECK(a->order_rows == 0 || (a->o_custkey && a->chars && a->offsets),
                    "bolap_q13: null order column");
    BOLAP_ARG_CHECK(a->cust_rows == 0 || a->c_custkey, "bolap_q13: null c_custkey");
    BOLAP_ARG_CHECK(a->word1 && a->word2 && std::strlen(a->word1) < 64 && std::strlen(a->word2) < 64,
                    "bolap_q13: words must be shorter than 64 bytes");
    Q13W w{};
    w.n1 = int32_t(std::strlen(a->word1));
    w.n2 = int32_t(std::strlen(a->word2));
    std::memcpy(w.w1, a->word1, size_t(w.n1));
    std::memcpy(w.w2, a->word2, size_t(w.n2));
    // required prefilter bits of the two words (query_q13.cpp:166-176)
    auto host_masks = [](const uint8_t* s, int32_t n, uint32_t& al, uint64_t& bg) {
        for (int32_t i = 0; i < n; ++i) {
            const uint8_t c = s[i];
            if (c >= 'a' && c <= 'z') al |= 1u << (c - 'a');
            else if (c >= 'A' && c <= 'Z') al |= 1u << (c - 'A');
            if (i > 0) bg |= uint64_t(1) << ((uint32_t(s[i - 1]) * 131u + c) & 63u);
        }
    };
    host_masks(w.w1, w.n1, w.req_alpha, w.req_bigram);
    host_masks(w.w2, w.n2, w.req_alpha, w.req_bigram);
    auto prefix = [](const uint8_t* s, int32_t n, uint32_t& pre, int32_t& m, int32_t& sh) {
        m = n < 4 ? n : 4;
        pre = 0;
        for (int32_t i = 0; i < m; ++i) pre |= uint32_t(s[i]) << (8 * i);
        sh = m == 0 ? 0 : 8 * (4 - m);  // an empty word is never looked for
    };
    prefix(w.w1, w.n1, w.pre1, w.m1, w.sh1);
    prefix(w.w2, w.n2, w.pre2, w.m2, w.sh2);
    *max_count = 0;
    for (int32_t i = 0; i < max_dist; ++i) dist[i] = 0;
    int32_t* slot = static_cast<int32_t*>(ctx->result_dev);
    int32_t kmax = 0;
    if (a->cust_rows > 0) BOLAP_TRY(max_i32(ctx, a->c_custkey, a->cust_rows, slot, &kmax));
    const int64_t size = kmax >= 0 ? int64_t(kmax) + 1 : 0;
    WsArena ws;
    const size_t o_cnt = ws.plan(size_t(size > 0 ? size : 1) * 4);
    const size_t o_max = ws.plan(8);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    auto* cnt = ws.at<int32_t>(o_cnt);
    auto* mx = ws.at<int32_t>(o_max);
    BOLAP_TRY(zero_dev(ctx, cnt, ws.off));
    {
        TraceScope ts(ctx, "q13_orders");
        if (a->order_rows > 0 && size > 0) {
            hipLaunchKernelGGL(q13_k, dim3(grid_of(ctx, (a->order_rows + 7) / 8, 8)), dim3(kT), 0,
                               ctx->stream, w, a->order_rows, a->o_custkey, a->chars, a->offsets,
                               a->alpha, a->bigram, size, cnt);
            hipLaunchKernelGGL(max_i32_arr_k, dim3(grid_of(ctx, size, 1)), dim3(kT), 0,
                               ctx->stream, cnt, size, mx);
            BOLAP_TRY(launch_check("q13_k"));
        }
    }
    int32_t m = 0;
    BOLAP_TRY(fetch(ctx, &m, mx, 4));
    *max_count = m;
    if (m + 1 > max_dist) {
        set_error("bolap_q13: counts up to %d need %d distribution slots", m, m + 1);
        return BOLAP_ERR_CAPACITY;
    }
    WsArena ws2 = ws;
    const size_t o_dist = ws2.plan(size_t(m + 1) * 8);
    BOLAP_TRY(ws_reserve(ctx, ws2.off, ws.off));
    ws2.base = static_cast<char*>(ctx->ws);
    auto* dd = ws2.at<unsigned long long>(o_dist);
    BOLAP_TRY(zero_dev(ctx, dd, size_t(m + 1) * 8));
    if (a->cust_rows > 0) {
        TraceScope ts(ctx, "q13_dist");
        hipLaunchKernelGGL(q13_dist_k, dim3(grid_of(ctx, a->cust_rows)), dim3(kT), 0, ctx->stream,
                           a->c_custkey, a->cust_rows, ws2.at<int32_t>(o_cnt), size, m + 1, dd);
        BOLAP_TRY(launch_check("q13_dist_k"));
    }
    return fetch(ctx, dist, dd, size_t(m + 1) * 8);
}

extern "C" int bolap_q16(bolap_ctx* ctx, const bolap_q16_args* a, int64_t* counts,
                         int64_t n_counts) {
    BOLAP_ARG_CHECK(ctx && a && (counts || n_counts == 0), "bolap_q16: null argument");
    BOLAP_ARG_CHECK(a->part_rows >= 0 && a->part_rows < (int64_t(1) << 31) && a->ps_rows >= 0 &&
                        a->supp_rows >= 0,
                    "bolap_q16: rows out of range");
    BOLAP_ARG_CHECK(a->n_types > 0 && a->n_types <= 256 && a->n_brands > 0 && a->n_brands <= 256,
                    "bolap_q16: %d brands / %d types", a->n_brands, a->n_types);
    BOLAP_ARG_CHECK(n_counts == int64_t(a->n_brands) * a->n_types * 51,
                    "bolap_q16: counts must hold n_brands * n_types * 51 groups");
    BOLAP_ARG_CHECK(a->part_rows == 0 || (a->p_partkey && a->p_brand && a->p_type && a->p_size),
                    "bolap_q16: null part column");
    BOLAP_ARG_CHECK(a->ps_rows == 0 || (a->ps_partkey && a->ps_suppkey),
                    "bolap_q16: null partsupp column");
    BOLAP_ARG_CHECK(a->supp_rows == 0 || (a->s_suppkey && a->s_banned),
                    "bolap_q16: null supplier column");
    for (int64_t i = 0; i < n_counts; ++i) counts[i] = 0;
    if (a->part_rows == 0 || a->ps_rows == 0) return BOLAP_OK;
    const int64_t np = a->n_part_keys, ns = a->n_supp_keys;
    WsArena ws;
    const size_t o_owner = ws.plan(size_t(np > 0 ? np : 1) * 8);
    const size_t o_ban = ws.plan(size_t(ns > 0 ? ns : 1));
    const size_t o_cnt = ws.plan(size_t(n_counts) * 8);
    const size_t o_n = ws.plan(8);
    const size_t o_zero_end = ws.off;
    const size_t o_tb = ws.plan(256);
    const size_t o_keys = ws.plan(size_t(a->ps_rows) * 8);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    BOLAP_TRY(zero_dev(ctx, ws.base, o_zero_end));
    BOLAP_HIP_TRY(hipMemcpyAsync(ws.at<uint8_t>(o_tb), a->type_banned, size_t(a->n_types),
                                 hipMemcpyHostToDevice, ctx->stream));
    auto* owner = ws.at<unsigned long long>(o_owner);
    auto* nout = ws.at<unsigned long long>(o_n);
    {
        TraceScope ts(ctx, "q16_part_supp");
        const Q16P p{a->part_rows, a->p_partkey, a->p_brand, a->p_type, a->p_size, a->n_types,
                     a->banned_brand, a->size_mask, ws.at<uint8_t>(o_tb), np};
        hipLaunchKernelGGL(q16_part_k, dim3(grid_of(ctx, a->part_rows)), dim3(kT), 0, ctx->stream,
                           p, owner);
        if (a->supp_rows > 0 && ns > 0)
            hipLaunchKernelGGL(mark_keys_k, dim3(grid_of(ctx, a->supp_rows)), dim3(kT), 0,
                               ctx->stream, a->s_suppkey, static_cast<const void*>(a->s_banned),
                               int32_t(BOLAP_U8), a->supp_rows, ns, 1, 1,
                               static_cast<const uint8_t*>(nullptr), 0, ws.at<uint8_t>(o_ban));
        hipLaunchKernelGGL(q16_ps_k, dim3(grid_of(ctx, a->ps_rows, 8)), dim3(kT), 0, ctx->stream,
                           a->ps_rows, a->ps_partkey, a->ps_suppkey, np, owner,
                           ws.at<uint8_t>(o_ban), ns, ws.at<uint64_t>(o_keys), nout);
        BOLAP_TRY(launch_check("q16_ps_k"));
    }
    unsigned long long n = 0;
    BOLAP_TRY(fetch(ctx, &n, nout, 8));
    auto* cnt = ws.at<unsigned long long>(o_cnt);
    if (n > 0) {
        TraceScope ts(ctx, "q16_distinct");
        int cap_log2 = 10;
        while ((uint64_t(1) << cap_log2) < 2 * n) ++cap_log2;
        WsArena ws2 = ws;
        const size_t o_set = ws2.plan(size_t(8) << cap_log2);
        BOLAP_TRY(ws_reserve(ctx, ws2.off, ws.off));
        ws2.base = static_cast<char*>(ctx->ws);
        cnt = ws2.at<unsigned long long>(o_cnt);
        auto* set = ws2.at<unsigned long long>(o_set);
        BOLAP_TRY(zero_dev(ctx, set, size_t(8) << cap_log2));
        hipLaunchKernelGGL(q16_distinct_k, dim3(grid_of(ctx, int64_t(n))), dim3(kT), 0,
                           ctx->stream, ws2.at<uint64_t>(o_keys), int64_t(n), set,
                           (uint64_t(1) << cap_log2) - 1, cnt);
        BOLAP_TRY(launch_check("q16_distinct_k"));
    }
    return fetch(ctx, counts, cnt, size_t(n_counts) * 8);
}
