// tpch.hip — TPC-H Q6 and Q18 pipelines and the build-time derivations of the
// reference Database layout, on the device.
//
// Q6 (bespoke_tpch/query_q6.cpp:155-397): a three-predicate range scan with
// SUM(extendedprice * discount).  The flat kernel streams the four reference
// columns (ship i16, discount u8, quantity i16, price i32: 9 B/row) with
// 8 rows per lane per unit (one 16-byte load of ship and of quantity, two of
// price, one 8-byte load of discount).  The zoned kernel does the work of the
// reference's shard pruning (query_q6.cpp:224-379) the HBM way: the build
// clusters the columns by the reference's shard key (ship month, discount,
// quantity / 100; builder_impl.cpp:855-911) with an LSD radix sort, and keeps
// one zone (min/max of each predicate column) per 2048 consecutive rows.  A
// workgroup reads a zone, skips the block when no row can pass, reads only
// the columns whose predicate the block does not satisfy wholly, and streams
// them coalesced; the skipped blocks cost one 16-byte zone read.
//
// Q18 (bespoke_tpch/query_q18.cpp:183-455): SUM(quantity) per order key into
// a dense int32 array (the reference's sum_qty), then the orders whose sum
// exceeds the threshold.  Lineitem sorted by order key (the reference's
// orderkey_sorted, builder_impl.cpp:569-577) takes a pass without atomics:
// the lane holding the first row of a run of equal keys sums the whole run
// (reading past its 8 rows when the run continues) and stores it.  A pass
// that finds the keys out of order raises a flag and the host reruns with
// atomics (in-lane run merging, one atomic per run).  Selection, customer
// lookup and the three-key order (price desc, date asc, key asc) follow.
//
// Build: truncated discounted_price and the returnflag/linestatus code
// (builder_impl.cpp:523-567), the stable date order of orders (:374-405),
// orderkey -> row and per-order lineitem ranges (:610-630, 929-991).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "radix.hpp"

namespace bolap {
namespace {

constexpr int kT = 256;
constexpr int kRowsPerLane = 8;
constexpr int64_t kZoneRows = BOLAP_ZONE_ROWS;
static_assert(kZoneRows == int64_t(kT) * kRowsPerLane, "one zone per workgroup pass");

// ---------------------------------------------------------------- dates ---
// Howard Hinnant's days-from-civil / civil-from-days algorithms (the ones the
// reference uses, builder_impl.cpp:22-40).
__host__ __device__ inline int64_t days_from_civil(int y, int m, int d) {
    y -= m <= 2;
    const int era = (y >= 0 ? y : y - 399) / 400;
    const unsigned yoe = unsigned(y - era * 400);
    const unsigned doy = (153u * unsigned(m + (m > 2 ? -3 : 9)) + 2u) / 5u + unsigned(d) - 1u;
    const unsigned doe = yoe * 365u + yoe / 4u - yoe / 100u + doy;
    return int64_t(era) * 146097 + int64_t(doe) - 719468;
}
__host__ __device__ inline void civil_from_days(int z, int& y, int& m, int& d) {
    z += 719468;
    const int era = (z >= 0 ? z : z - 146096) / 146097;
    const unsigned doe = unsigned(z - era * 146097);
    const unsigned yoe = (doe - doe / 1460u + doe / 36524u - doe / 146096u) / 365u;
    const unsigned doy = doe - (365u * yoe + yoe / 4u - yoe / 100u);
    const unsigned mp = (5u * doy + 2u) / 153u;
    d = int(doy - (153u * mp + 2u) / 5u + 1u);
    m = int(mp < 10u ? mp + 3u : mp - 9u);
    y = int(yoe) + era * 400 + (m <= 2);
}
// 1992-01-01, the base of every date offset (builder_impl.cpp:454).
constexpr int kBaseDays = 8035;

inline unsigned grid_of(bolap_ctx* ctx, int64_t items, int per_cu) {
    int64_t b = (items + kT - 1) / kT;
    const int64_t cap = int64_t(ctx->num_cus) * per_cu;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

// ------------------------------------------------------------------ Q6 ---
struct Q6K {
    const int16_t* ship;
    const uint8_t* disc;
    const int16_t* qty;
    const int32_t* price;
    int64_t rows;
    int32_t ship_lo;
    uint32_t date_span;  // ship passes when (u32)(ship - ship_lo) < date_span
    int32_t disc_lo;
    uint32_t disc_span;  // disc passes when (u32)(disc - disc_lo) <= disc_span
    int32_t qty_lt;
};

// The eight rows of unit u (rows 8u .. 8u+7, all in bounds).
struct Q6Unit {
    VecT<int16_t, 8> ship, qty;
    VecT<int32_t, 4> p0, p1;
    VecT<uint8_t, 8> disc;
};

template <bool kShip, bool kQty>
BOLAP_DEV void q6_load(const Q6K& a, int64_t u, Q6Unit& x) {
    x.p0 = load_unit<int32_t, 4>(a.price, 2 * u);
    x.p1 = load_unit<int32_t, 4>(a.price, 2 * u + 1);
    x.disc = load_unit<uint8_t, 8>(a.disc, u);
    if (kShip) x.ship = load_unit<int16_t, 8>(a.ship, u);
    if (kQty) x.qty = load_unit<int16_t, 8>(a.qty, u);
}

template <bool kShip, bool kQty>
BOLAP_DEV void q6_use(const Q6K& a, const Q6Unit& x, uint64_t& rev, uint32_t& cnt) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int32_t d = int32_t(x.disc.v[r]);
        bool pass = uint32_t(d - a.disc_lo) <= a.disc_span;
        if (kShip) pass = pass && uint32_t(int32_t(x.ship.v[r]) - a.ship_lo) < a.date_span;
        if (kQty) pass = pass && int32_t(x.qty.v[r]) < a.qty_lt;
        const int32_t p = r < 4 ? x.p0.v[r] : x.p1.v[r - 4];
        rev += pass ? uint64_t(int64_t(p) * int64_t(d)) : 0ull;
        cnt += pass ? 1u : 0u;
    }
}

BOLAP_DEV void q6_row(const Q6K& a, int64_t i, uint64_t& rev, uint32_t& cnt) {
    const int32_t d = int32_t(a.disc[i]);
    if (uint32_t(int32_t(a.ship[i]) - a.ship_lo) >= a.date_span) return;
    if (uint32_t(d - a.disc_lo) > a.disc_span) return;
    if (int32_t(a.qty[i]) >= a.qty_lt) return;
    rev += uint64_t(int64_t(a.price[i]) * int64_t(d));
    cnt += 1;
}

// Block partial (revenue, rows emitted, rows scanned) -> partials[3 * b].
BOLAP_DEV void q6_block_out(uint64_t rev, uint32_t cnt, uint64_t scanned, int64_t* partials) {
    __shared__ uint64_t s_r[kT / kWave], s_c[kT / kWave];
    rev = wave_sum_u64(rev);
    const uint64_t c = wave_sum_u64(uint64_t(cnt));
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s_r[threadIdx.x / kWave] = rev;
        s_c[threadIdx.x / kWave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t tr = 0, tc = 0;
        for (int w = 0; w < kT / kWave; ++w) {
            tr += s_r[w];
            tc += s_c[w];
        }
        partials[3 * int64_t(blockIdx.x)] = int64_t(tr);
        partials[3 * int64_t(blockIdx.x) + 1] = int64_t(tc);
        partials[3 * int64_t(blockIdx.x) + 2] = int64_t(scanned);
    }
}

// Flat scan: two units per lane in flight (the next pair is loaded while the
// current one is evaluated).
__global__ __launch_bounds__(kT) void q6_flat_k(Q6K a, int64_t* partials) {
    const int64_t units = a.rows / 8;
    const int64_t stride = int64_t(gridDim.x) * kT * 2;
    uint64_t rev = 0;
    uint32_t cnt = 0;
    int64_t u = int64_t(blockIdx.x) * kT * 2 + threadIdx.x;
    auto ld = [&](Q6Unit (&x)[2], int64_t at) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            int64_t v = at + int64_t(k) * kT;
            v = v < units ? v : units - 1;  // in bounds; dropped when used
            q6_load<true, true>(a, v, x[k]);
        }
    };
    auto use = [&](const Q6Unit (&x)[2], int64_t at) {
#pragma unroll
        for (int k = 0; k < 2; ++k)
            if (at + int64_t(k) * kT < units) q6_use<true, true>(a, x[k], rev, cnt);
    };
    if (units > 0 && u < units) {
        Q6Unit A[2], B[2];
        ld(A, u);
        for (;;) {
            ld(B, u + stride);
            use(A, u);
            u += stride;
            if (u >= units) break;
            ld(A, u + stride);
            use(B, u);
            u += stride;
            if (u >= units) break;
        }
    }
    if (blockIdx.x == 0 && int64_t(threadIdx.x) < a.rows - units * 8)
        q6_row(a, units * 8 + threadIdx.x, rev, cnt);
    q6_block_out(rev, cnt, 0, partials);
}

// Zoned scan: workgroup b takes zones b, b + grid, ...; each zone is 2048
// rows, 8 per lane.  Zone decisions are workgroup-uniform.
__global__ __launch_bounds__(kT) void q6_zoned_k(Q6K a, const bolap_zone* __restrict__ zones,
                                                 int64_t n_zones, int64_t* partials) {
    uint64_t rev = 0, scanned = 0;
    uint32_t cnt = 0;
    for (int64_t z = blockIdx.x; z < n_zones; z += gridDim.x) {
        const bolap_zone zn = zones[z];
        const int32_t smin = zn.ship_min, smax = zn.ship_max;
        const int32_t qmin = zn.qty_min, qmax = zn.qty_max;
        const int32_t dmin = zn.disc_min, dmax = zn.disc_max;
        const int32_t ship_hi = a.ship_lo + int32_t(a.date_span);
        const int32_t disc_hi = a.disc_lo + int32_t(a.disc_span);
        if (smax < a.ship_lo || smin >= ship_hi || dmax < a.disc_lo || dmin > disc_hi ||
            qmin >= a.qty_lt)
            continue;
        const bool full_ship = smin >= a.ship_lo && smax < ship_hi;
        const bool full_qty = qmax < a.qty_lt;
        const int64_t first = z * kZoneRows;
        const int64_t n = a.rows - first < kZoneRows ? a.rows - first : kZoneRows;
        scanned += uint64_t(n);
        if (n == kZoneRows) {
            const int64_t u = z * (kZoneRows / 8) + threadIdx.x;
            Q6Unit x;
            if (full_ship && full_qty) {
                q6_load<false, false>(a, u, x);
                q6_use<false, false>(a, x, rev, cnt);
            } else if (full_ship) {
                q6_load<false, true>(a, u, x);
                q6_use<false, true>(a, x, rev, cnt);
            } else if (full_qty) {
                q6_load<true, false>(a, u, x);
                q6_use<true, false>(a, x, rev, cnt);
            } else {
                q6_load<true, true>(a, u, x);
                q6_use<true, true>(a, x, rev, cnt);
            }
        } else {
            for (int64_t i = threadIdx.x; i < n; i += kT) q6_row(a, first + i, rev, cnt);
        }
    }
    // Only thread 0's count of scanned rows is kept (every lane counted the
    // same zones).
    q6_block_out(rev, cnt, threadIdx.x == 0 ? scanned : 0, partials);
}

// Sums the block partials into res[0..2] and the pinned host row.
// Sums three u64 fields of the per-workgroup partials into res[0..2] and
// the pinned host row; `flag` (optional) is copied to host[3].
__global__ __launch_bounds__(kT) void q6_merge_k(const int64_t* __restrict__ partials, int nb,
                                                  int64_t* res, int64_t* host, int64_t seq,
                                                  const int32_t* flag = nullptr) {
    uint64_t r = 0, c = 0, s = 0;
    for (int b = threadIdx.x; b < nb; b += kT) {
        r += uint64_t(partials[3 * b]);
        c += uint64_t(partials[3 * b + 1]);
        s += uint64_t(partials[3 * b + 2]);
    }
    __shared__ uint64_t sh[3][kT / kWave];
    r = wave_sum_u64(r);
    c = wave_sum_u64(c);
    s = wave_sum_u64(s);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sh[0][threadIdx.x / kWave] = r;
        sh[1][threadIdx.x / kWave] = c;
        sh[2][threadIdx.x / kWave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // One thread writes every field, then the sequence number (host slot
        // 4) with release semantics: the host waits for that value.
        for (int f = 0; f < 3; ++f) {
            uint64_t t = 0;
            for (int w = 0; w < kT / kWave; ++w) t += sh[f][w];
            res[f] = int64_t(t);
            __hip_atomic_store(host + f, int64_t(t), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        if (flag)
            __hip_atomic_store(host + 3, int64_t(*flag), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(host + 4, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// One zone per workgroup: min/max of ship, quantity and discount.
template <typename T>
BOLAP_DEV T wave_min(T v) {
    for (int o = 32; o > 0; o >>= 1) {
        const T w = T(__shfl_xor(int(v), o, kWave));
        v = w < v ? w : v;
    }
    return v;
}
template <typename T>
BOLAP_DEV T wave_max(T v) {
    for (int o = 32; o > 0; o >>= 1) {
        const T w = T(__shfl_xor(int(v), o, kWave));
        v = w > v ? w : v;
    }
    return v;
}

__global__ __launch_bounds__(kT) void zones_k(const int16_t* __restrict__ ship,
                                              const uint8_t* __restrict__ disc,
                                              const int16_t* __restrict__ qty, int64_t rows,
                                              int64_t n_zones, bolap_zone* zones) {
    __shared__ int32_t s[6][kT / kWave];
    for (int64_t z = blockIdx.x; z < n_zones; z += gridDim.x) {
        int32_t smin = INT32_MAX, smax = INT32_MIN, qmin = INT32_MAX, qmax = INT32_MIN;
        int32_t dmin = INT32_MAX, dmax = INT32_MIN;
        const int64_t first = z * kZoneRows;
        const int64_t n = rows - first < kZoneRows ? rows - first : kZoneRows;
        for (int64_t i = threadIdx.x; i < n; i += kT) {
            const int32_t sv = ship[first + i], qv = qty[first + i], dv = disc[first + i];
            smin = sv < smin ? sv : smin;
            smax = sv > smax ? sv : smax;
            qmin = qv < qmin ? qv : qmin;
            qmax = qv > qmax ? qv : qmax;
            dmin = dv < dmin ? dv : dmin;
            dmax = dv > dmax ? dv : dmax;
        }
        smin = wave_min(smin);
        smax = wave_max(smax);
        qmin = wave_min(qmin);
        qmax = wave_max(qmax);
        dmin = wave_min(dmin);
        dmax = wave_max(dmax);
        const int w = threadIdx.x / kWave;
        if ((threadIdx.x & (kWave - 1)) == 0) {
            s[0][w] = smin; s[1][w] = smax; s[2][w] = qmin;
            s[3][w] = qmax; s[4][w] = dmin; s[5][w] = dmax;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 1; k < kT / kWave; ++k) {
                s[0][0] = s[0][k] < s[0][0] ? s[0][k] : s[0][0];
                s[1][0] = s[1][k] > s[1][0] ? s[1][k] : s[1][0];
                s[2][0] = s[2][k] < s[2][0] ? s[2][k] : s[2][0];
                s[3][0] = s[3][k] > s[3][0] ? s[3][k] : s[3][0];
                s[4][0] = s[4][k] < s[4][0] ? s[4][k] : s[4][0];
                s[5][0] = s[5][k] > s[5][0] ? s[5][k] : s[5][0];
            }
            bolap_zone out{};
            out.ship_min = int16_t(s[0][0]);
            out.ship_max = int16_t(s[1][0]);
            out.qty_min = int16_t(s[2][0]);
            out.qty_max = int16_t(s[3][0]);
            out.disc_min = uint8_t(s[4][0]);
            out.disc_max = uint8_t(s[5][0]);
            zones[z] = out;
        }
        __syncthreads();
    }
}

// Shard key of the clustering sort (ascending = the reference's shard key
// order): month since 1900 (12 bits), discount (8 bits), quantity / 100
// biased by 512 (10 bits); the row index rides along as the value.
__global__ __launch_bounds__(kT) void shard_keys_k(const int16_t* __restrict__ ship,
                                                   const uint8_t* __restrict__ disc,
                                                   const int16_t* __restrict__ qty, int64_t rows,
                                                   uint64_t* keys, uint32_t* vals) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        int y, m, d;
        civil_from_days(kBaseDays + int(ship[i]), y, m, d);
        const uint64_t month = uint64_t((y - 1900) * 12 + (m - 1)) & 0xfff;
        const uint64_t qb = uint64_t(int(qty[i]) / 100 + 512) & 0x3ff;
        keys[i] = (month << 18) | (uint64_t(disc[i]) << 10) | qb;
        vals[i] = uint32_t(i);
    }
}

__global__ __launch_bounds__(kT) void gather_q6_k(const uint32_t* __restrict__ perm, int64_t rows,
                                                  const int16_t* ship, const uint8_t* disc,
                                                  const int16_t* qty, const int32_t* price,
                                                  int16_t* ship_o, uint8_t* disc_o,
                                                  int16_t* qty_o, int32_t* price_o) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const uint32_t r = perm[i];
        ship_o[i] = ship[r];
        disc_o[i] = disc[r];
        qty_o[i] = qty[r];
        price_o[i] = price[r];
    }
}

// ----------------------------------------------------------------- Q14 ---
// partkey -> promo flag (query_q14.cpp:202-223): the largest part row of a
// key wins (the reference's sequential assignment lets the last row win),
// packed as (row + 1) << 1 | promo under an atomic maximum, then narrowed to
// int8 (-1 absent).  The array covers keys 0..size-1; a larger key sets
// *ovf so that the host reruns with the exact size.
__global__ __launch_bounds__(kT) void q14_clear_k(int32_t* packed, int64_t size, int32_t* ovf) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *ovf = 0;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < size;
         i += int64_t(gridDim.x) * kT)
        packed[i] = 0;
}

__global__ __launch_bounds__(kT) void q14_part_k(const int32_t* __restrict__ pk,
                                                 const uint8_t* __restrict__ promo, int64_t n,
                                                 int32_t* packed, int64_t size, int32_t* ovf) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = pk[i];
        if (k < 0) continue;
        if (k < size)
            atomicMax(packed + k, int32_t(((i + 1) << 1) | (promo[i] ? 1 : 0)));
        else
            *ovf = 1;
    }
}

__global__ __launch_bounds__(kT) void q14_narrow_k(const int32_t* __restrict__ packed,
                                                   int64_t size, int8_t* flag) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < size;
         i += int64_t(gridDim.x) * kT)
        flag[i] = packed[i] == 0 ? int8_t(-1) : int8_t(packed[i] & 1);
}

struct Q14K {
    const int16_t* ship;
    const int32_t* pk;
    const int32_t* dp;
    int64_t rows;
    int32_t ship_lo, ship_hi;
    const int8_t* flag;
    int64_t size;
};

BOLAP_DEV void q14_row(const Q14K& a, int64_t i, uint64_t& promo, uint64_t& total,
                       uint64_t& cnt) {
    const int32_t k = a.pk[i];
    if (k < 0 || k >= a.size) return;
    const int8_t f = a.flag[k];
    if (f < 0) return;
    const int64_t v = a.dp[i];
    total += uint64_t(v);
    promo += f ? uint64_t(v) : 0ull;
    cnt += 1;
}

// Block sums of three u64 values -> partials[3 * b].
BOLAP_DEV void block_out3(uint64_t x, uint64_t y, uint64_t z, int64_t* partials) {
    __shared__ uint64_t s[3][kT / kWave];
    x = wave_sum_u64(x);
    y = wave_sum_u64(y);
    z = wave_sum_u64(z);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s[0][threadIdx.x / kWave] = x;
        s[1][threadIdx.x / kWave] = y;
        s[2][threadIdx.x / kWave] = z;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint64_t t = 0;
        for (int w = 0; w < kT / kWave; ++w) t += s[threadIdx.x][w];
        partials[3 * int64_t(blockIdx.x) + threadIdx.x] = int64_t(t);
    }
}

// Lineitem pass over a zone map: one zone (2048 rows = 256 units of 8) per
// workgroup iteration; zones outside the window are skipped, the ship test
// is skipped for zones wholly inside it, and the part key and discounted
// price of a unit are loaded only when one of its rows is in the window.
__global__ __launch_bounds__(kT) void q14_zoned_k(Q14K a, const bolap_zone* __restrict__ zones,
                                                  int64_t* partials) {
    uint64_t promo = 0, total = 0, cnt = 0;
    const int64_t units = a.rows / 8;
    const int64_t n_zones = (a.rows + kZoneRows - 1) / kZoneRows;
    for (int64_t z = blockIdx.x; z < n_zones; z += gridDim.x) {
        const bolap_zone zn = zones[z];
        if (zn.ship_max < a.ship_lo || zn.ship_min >= a.ship_hi) continue;
        const bool full = zn.ship_min >= a.ship_lo && zn.ship_max < a.ship_hi;
        const int64_t u = z * kT + threadIdx.x;  // zone z = units z*256 .. z*256+255
        if (u >= units) continue;
        uint32_t m = 0xffu;
        if (!full) {
            const VecT<int16_t, 8> sv = load_unit<int16_t, 8>(a.ship, u);
            m = 0;
#pragma unroll
            for (int r = 0; r < 8; ++r)
                m |= (sv.v[r] >= a.ship_lo && sv.v[r] < a.ship_hi ? 1u : 0u) << r;
        }
        if (!m) continue;
        const VecT<int32_t, 4> k0 = load_unit<int32_t, 4>(a.pk, 2 * u);
        const VecT<int32_t, 4> k1 = load_unit<int32_t, 4>(a.pk, 2 * u + 1);
        const VecT<int32_t, 4> d0 = load_unit<int32_t, 4>(a.dp, 2 * u);
        const VecT<int32_t, 4> d1 = load_unit<int32_t, 4>(a.dp, 2 * u + 1);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            if (!((m >> r) & 1u)) continue;
            const int32_t k = r < 4 ? k0.v[r] : k1.v[r - 4];
            if (k < 0 || k >= a.size) continue;
            const int8_t f = a.flag[k];
            if (f < 0) continue;
            const int64_t v = r < 4 ? d0.v[r] : d1.v[r - 4];
            total += uint64_t(v);
            promo += f ? uint64_t(v) : 0ull;
            cnt += 1;
        }
    }
    // Rows past the last whole unit.
    if (blockIdx.x == 0 && int64_t(threadIdx.x) < a.rows - units * 8) {
        const int64_t i = units * 8 + threadIdx.x;
        if (a.ship[i] >= a.ship_lo && a.ship[i] < a.ship_hi) q14_row(a, i, promo, total, cnt);
    }
    block_out3(promo, total, cnt, partials);
}

// Flat lineitem pass, kU units of 8 rows per lane in flight: all ship
// loads, then the part-key and price loads of the units with a row in the
// window, then one flag gather per row in the window (the flag array is
// small enough to stay in L2), so the three dependent reads of a unit
// overlap across kU units.
template <int kU>
__global__ __launch_bounds__(kT) void q14_flat_k(Q14K a, int64_t* partials) {
    uint64_t promo = 0, total = 0, cnt = 0;
    const int64_t units = a.rows / 8;
    const int64_t stride = int64_t(gridDim.x) * kT * kU;
    for (int64_t base = int64_t(blockIdx.x) * kT * kU + threadIdx.x; base < units;
         base += stride) {
        uint32_t m[kU];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const int64_t u = base + int64_t(k) * kT;
            m[k] = 0;
            if (u < units) {
                const VecT<int16_t, 8> sv = load_unit<int16_t, 8>(a.ship, u);
#pragma unroll
                for (int r = 0; r < 8; ++r)
                    m[k] |= (sv.v[r] >= a.ship_lo && sv.v[r] < a.ship_hi ? 1u : 0u) << r;
            }
        }
        int32_t key[kU][8], dp[kU][8];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const int64_t u = base + int64_t(k) * kT;
            if (m[k]) {
                const VecT<int32_t, 4> k0 = load_unit<int32_t, 4>(a.pk, 2 * u);
                const VecT<int32_t, 4> k1 = load_unit<int32_t, 4>(a.pk, 2 * u + 1);
                const VecT<int32_t, 4> d0 = load_unit<int32_t, 4>(a.dp, 2 * u);
                const VecT<int32_t, 4> d1 = load_unit<int32_t, 4>(a.dp, 2 * u + 1);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    key[k][r] = k0.v[r];
                    key[k][r + 4] = k1.v[r];
                    dp[k][r] = d0.v[r];
                    dp[k][r + 4] = d1.v[r];
                }
            }
        }
        int8_t f[kU][8];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                f[k][r] = -1;
                if (((m[k] >> r) & 1u) && key[k][r] >= 0 && key[k][r] < a.size)
                    f[k][r] = a.flag[key[k][r]];
            }
        }
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            if (!m[k]) continue;
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                if (!((m[k] >> r) & 1u) || f[k][r] < 0) continue;
                const int64_t v = dp[k][r];
                total += uint64_t(v);
                promo += f[k][r] ? uint64_t(v) : 0ull;
                cnt += 1;
            }
        }
    }
    // Rows past the last whole unit.
    if (blockIdx.x == 0 && int64_t(threadIdx.x) < a.rows - units * 8) {
        const int64_t i = units * 8 + threadIdx.x;
        if (a.ship[i] >= a.ship_lo && a.ship[i] < a.ship_hi) q14_row(a, i, promo, total, cnt);
    }
    block_out3(promo, total, cnt, partials);
}

// ----------------------------------------------------------------- Q18 ---
enum { kQ18MaxL = 0, kQ18MaxO = 1, kQ18MaxC = 2, kQ18Unsorted = 3, kQ18Count = 4,
       kQ18ResN = 8 };

// First and last order and customer keys (the density checks' inputs) into
// pinned host memory in one launch; zeros for an empty table.
__global__ void q18_ends_k(const int32_t* __restrict__ okey, int64_t n_o,
                           const int32_t* __restrict__ ckey, int64_t n_c, int32_t* out) {
    if (threadIdx.x == 0) out[0] = n_o > 0 ? okey[0] : 0;
    if (threadIdx.x == 1) out[1] = n_o > 0 ? okey[n_o - 1] : 0;
    if (threadIdx.x == 2) out[2] = n_c > 0 ? ckey[0] : 0;
    if (threadIdx.x == 3) out[3] = n_c > 0 ? ckey[n_c - 1] : 0;
}

__global__ __launch_bounds__(kT) void q18_init_k(int64_t* res) {
    if (threadIdx.x < kQ18ResN) res[threadIdx.x] = threadIdx.x < 3 ? -1 : 0;
}

// Largest key of up to three int32 columns (res[0..2], -1 when empty or all
// negative), one atomic per workgroup and column.  Each thread keeps four
// independent coalesced loads in flight per round.
BOLAP_DEV int32_t col_max(const int32_t* __restrict__ v, int64_t n) {
    const int64_t step = int64_t(gridDim.x) * kT;
    int32_t m = -1;
    int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    for (; i + 3 * step < n; i += 4 * step) {
        const int32_t a = v[i], b = v[i + step], c = v[i + 2 * step], d = v[i + 3 * step];
        const int32_t x = a > b ? a : b, y = c > d ? c : d;
        const int32_t z = x > y ? x : y;
        m = z > m ? z : m;
    }
    for (; i < n; i += step) m = v[i] > m ? v[i] : m;
    return m;
}

__global__ __launch_bounds__(kT) void q18_max_k(const int32_t* __restrict__ l, int64_t nl,
                                                const int32_t* __restrict__ o, int64_t no,
                                                const int32_t* __restrict__ c, int64_t nc,
                                                int64_t* res) {
    int32_t m[3] = {col_max(l, nl), col_max(o, no), col_max(c, nc)};
    __shared__ int32_t s[3][kT / kWave];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int32_t w = wave_max(m[k]);
        if ((threadIdx.x & (kWave - 1)) == 0) s[k][threadIdx.x / kWave] = w;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int32_t t = -1;
        for (int w = 0; w < kT / kWave; ++w) t = s[threadIdx.x][w] > t ? s[threadIdx.x][w] : t;
        // res holds -1 initially: a signed maximum through a CAS loop.
        unsigned long long* p = reinterpret_cast<unsigned long long*>(res + threadIdx.x);
        unsigned long long cur = *p;
        while (int64_t(cur) < int64_t(t)) {
            const unsigned long long prev = atomicCAS(p, cur, (unsigned long long)(int64_t(t)));
            if (prev == cur) break;
            cur = prev;
        }
    }
}

__global__ __launch_bounds__(kT) void fill_i32_k(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        p[i] = v;
}

// custkey -> customer row; the largest row wins, as the reference's
// sequential assignment lets the last row win (query_q18.cpp:226-236).
__global__ __launch_bounds__(kT) void cust_map_k(const int32_t* __restrict__ c, int64_t nc,
                                                 int32_t* map, int64_t map_size) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < nc;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = c[i];
        if (k >= 0 && k < map_size) atomicMax(map + k, int32_t(i));
    }
}

// Sorted lineitem: the lane holding a run's first row sums the run and
// stores it.  Each lane loads its own 8 rows and takes the next 8 from the
// next lane (the wave's last lane loads them), so a run that crosses into
// the next unit is finished in registers; only a run longer than that reads
// on, row by row.  Any pair of adjacent rows out of order raises the flag.
BOLAP_DEV int q18_unit(const int32_t* __restrict__ key, const int16_t* __restrict__ qty,
                       int64_t n, int64_t u, int32_t (&k)[8], int32_t (&q)[8]) {
    const int64_t r0 = u * 8;
    if (r0 + 8 <= n) {
        const VecT<int32_t, 4> k0 = load_unit<int32_t, 4>(key, 2 * u);
        const VecT<int32_t, 4> k1 = load_unit<int32_t, 4>(key, 2 * u + 1);
        const VecT<int16_t, 8> qv = load_unit<int16_t, 8>(qty, u);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            k[j] = k0.v[j];
            k[j + 4] = k1.v[j];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) q[j] = qv.v[j];
        return 8;
    }
    const int m = r0 < n ? int(n - r0) : 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        k[j] = j < m ? key[r0 + j] : 0;
        q[j] = j < m ? qty[r0 + j] : 0;
    }
    return m;
}

// One wave tile of q18_sum_sorted_k: lane `lane` holds unit u's rows; the
// last lane also holds the next unit (the next wave's first), the first
// lane the previous row's key, all loaded together.
struct Q18Tile {
    int32_t k[8], q[8];
    int32_t nk[8], nq[8];
    int m;
    int32_t prev;
};

BOLAP_DEV void q18_load_tile(const int32_t* __restrict__ key, const int16_t* __restrict__ qty,
                             int64_t n, int64_t units, int64_t u, int lane, Q18Tile& t) {
    if (u < units) {
        t.m = q18_unit(key, qty, n, u, t.k, t.q);
    } else {
        t.m = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) t.k[j] = t.q[j] = 0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) t.nk[j] = t.nq[j] = 0;
    if (lane == kWave - 1 && u + 1 < units) (void)q18_unit(key, qty, n, u + 1, t.nk, t.nq);
    t.prev = 0;
    if (lane == 0 && u < units && u > 0) t.prev = key[u * 8 - 1];
}

// Sums the runs that start in the lane's rows of tile t and stores them
// through the wave's LDS rows.
BOLAP_DEV void q18_tile_runs(const int32_t* __restrict__ key, const int16_t* __restrict__ qty,
                             int64_t n, int64_t units, int64_t u, int lane, const Q18Tile& t,
                             int2* s_out, int32_t* sum, int64_t size, bool& bad) {
    const int64_t r0 = u * 8;
    const int m = t.m;
    int32_t k[16], q[16];
    const int64_t rb = r0 + 8;
    const int mb = rb >= n ? 0 : (n - rb >= 8 ? 8 : int(n - rb));
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        k[j] = t.k[j];
        q[j] = t.q[j];
        k[j + 8] = __shfl_down(t.k[j], 1, kWave);
        q[j + 8] = __shfl_down(t.q[j], 1, kWave);
    }
    const int32_t up = __shfl_up(t.k[7], 1, kWave);
    if (lane == kWave - 1) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            k[j + 8] = t.nk[j];
            q[j + 8] = t.nq[j];
        }
    }
    const bool has_prev = r0 > 0 && m > 0;
    const int32_t prev = lane == 0 ? t.prev : up;
    if (has_prev && k[0] < prev) bad = true;
#pragma unroll
    for (int j = 1; j < 8; ++j)
        if (j < m && k[j] < k[j - 1]) bad = true;
    // One backward pass over the 16 rows: srun[x] = sum of the run of
    // equal keys from x to its end inside the window, reach bit x = that
    // run reaches row 15 of a full window (it may continue past it).
    const int lim = m == 8 ? 8 + mb : m;
    uint32_t srun[16];
    uint32_t reach = 0;
#pragma unroll
    for (int x = 15; x >= 0; --x) {
        const bool v = x < lim;
        const bool cont = x + 1 < lim && k[x + 1] == k[x];
        srun[x] = v ? uint32_t(q[x]) + (cont ? srun[x + 1 < 16 ? x + 1 : 15] : 0u) : 0u;
        const bool r = v && (x == 15 || (cont && ((reach >> (x + 1)) & 1u)));
        reach |= (r ? 1u : 0u) << x;
    }
    // Run starts among the lane's 8 rows: their sums, then the wave's rows
    // go through LDS so that the stores are issued with full waves.
    uint32_t tot = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const bool start = j < m && (j == 0 ? (!has_prev || k[0] != prev) : k[j] != k[j - 1]);
        uint32_t sj = srun[j];
        if (start && ((reach >> j) & 1u)) {  // the run continues past the window
            int64_t r = r0 + 16;
            while (r < n && key[r] == k[j]) sj += uint32_t(int32_t(qty[r++]));
        }
        const bool keep = start && k[j] >= 0 && k[j] < size;
        const uint64_t mj = __ballot(keep);
        if (keep) {
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                uint32_t(mj >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mj), 0u));
            s_out[tot + below] = make_int2(k[j], int32_t(sj));
        }
        tot += uint32_t(__popcll(mj));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    for (uint32_t i = uint32_t(lane); i < tot; i += kWave) {
        const int2 e = s_out[i];
        sum[e.x] = e.y;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Wave-uniform trip count: lanes past the end take part with no rows, so
// the lane shuffles always see the whole wave.
__global__ __launch_bounds__(kT) void q18_sum_sorted_k(const int32_t* __restrict__ key,
                                                       const int16_t* __restrict__ qty,
                                                       int64_t n, int32_t* sum, int64_t size,
                                                       int64_t* res) {
    constexpr int kW = kT / kWave;
    __shared__ int2 s_out[kW][kWave * 8];  // a wave's (key, sum) rows before the stores
    const int lane = int(threadIdx.x) & (kWave - 1);
    const int wave = int(threadIdx.x) / kWave;
    const int64_t units = (n + 7) / 8;
    const int64_t stride = int64_t(gridDim.x) * kT;
    bool bad = false;
    for (int64_t wb = int64_t(blockIdx.x) * kT + (threadIdx.x & ~(kWave - 1)); wb < units;
         wb += stride) {
        const int64_t u = wb + lane;
        Q18Tile t;
        q18_load_tile(key, qty, n, units, u, lane, t);
        q18_tile_runs(key, qty, n, units, u, lane, t, s_out[wave], sum, size, bad);
    }
    if (__ballot(bad) && (threadIdx.x & (kWave - 1)) == 0)
        atomicOr(reinterpret_cast<unsigned long long*>(res + kQ18Unsorted), 1ull);
}

// Any order: one atomic per run of equal keys within a lane's 8 rows.
__global__ __launch_bounds__(kT) void q18_sum_atomic_k(const int32_t* __restrict__ key,
                                                       const int16_t* __restrict__ qty,
                                                       int64_t n, int32_t* sum, int64_t size) {
    const int64_t units = (n + 7) / 8;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < units;
         u += int64_t(gridDim.x) * kT) {
        const int64_t r0 = u * 8;
        const int m = r0 + 8 <= n ? 8 : int(n - r0);
        int32_t cur = key[r0];
        uint32_t s = uint32_t(int32_t(qty[r0]));
        for (int j = 1; j <= m; ++j) {
            const bool end = j == m || key[r0 + j] != cur;
            if (end) {
                if (cur >= 0 && cur < size) atomicAdd(reinterpret_cast<uint32_t*>(sum + cur), s);
                if (j < m) {
                    cur = key[r0 + j];
                    s = uint32_t(int32_t(qty[r0 + j]));
                }
            } else {
                s += uint32_t(int32_t(qty[r0 + j]));
            }
        }
    }
}

struct Q18Sel {
    const int32_t* sum;
    int64_t size;
    int32_t threshold;
    const int32_t* o_orderkey;
    const int32_t* o_custkey;
    const int16_t* o_orderdate;
    const int32_t* o_totalprice;
    int64_t o_rows;
    int64_t c_rows;
    int32_t dense_cust;
    const int32_t* cust_map;
    int64_t map_size;
    bolap_q18_row* out;
    int64_t cap;
};

BOLAP_DEV void q18_emit(const Q18Sel& s, bool ok, const bolap_q18_row& row, int64_t* res) {
    const uint64_t bal = __ballot(ok);
    if (!bal) return;
    const int lane = threadIdx.x & (kWave - 1);
    uint64_t at = 0;
    if (lane == 0)
        at = atomicAdd(reinterpret_cast<unsigned long long*>(res + kQ18Count),
                       (unsigned long long)__popcll(bal));
    at = readlane_u64(at, 0);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int64_t idx = int64_t(at) + __popcll(bal & lt);
    if (ok && idx < s.cap) s.out[idx] = row;
}

// Customer row of an order's customer key, or -1.
BOLAP_DEV int64_t q18_cust(const Q18Sel& s, int32_t ck) {
    if (s.dense_cust) {
        const int64_t r = int64_t(ck) - 1;
        return (r >= 0 && r < s.c_rows) ? r : -1;
    }
    return (ck >= 0 && ck < s.map_size) ? int64_t(s.cust_map[ck]) : -1;
}

// Dense path: order key k qualifies when sum[k] > threshold; its order row is
// k - 1 (query_q18.cpp:258-300).
__global__ __launch_bounds__(kT) void q18_select_dense_k(Q18Sel s, int64_t* res) {
    for (int64_t base = int64_t(blockIdx.x) * kT; base < s.size; base += int64_t(gridDim.x) * kT) {
        const int64_t k = base + threadIdx.x;
        bool ok = k >= 1 && k < s.size && s.sum[k] > s.threshold;
        bolap_q18_row row{};
        if (ok) {
            const int64_t orow = k - 1;
            ok = orow < s.o_rows;
            if (ok) {
                const int32_t ck = s.o_custkey[orow];
                const int64_t crow = q18_cust(s, ck);
                ok = crow >= 0;
                row = bolap_q18_row{ck, int32_t(k), int32_t(s.o_orderdate[orow]),
                                    s.o_totalprice[orow], int64_t(s.sum[k]), int32_t(crow),
                                    int32_t(orow)};
            }
        }
        q18_emit(s, ok, row, res);
    }
}

// Any order keys: every order row looks up its key's sum (query_q18.cpp:
// 402-437).
__global__ __launch_bounds__(kT) void q18_select_orders_k(Q18Sel s, int64_t* res) {
    for (int64_t base = int64_t(blockIdx.x) * kT; base < s.o_rows;
         base += int64_t(gridDim.x) * kT) {
        const int64_t o = base + threadIdx.x;
        bool ok = false;
        bolap_q18_row row{};
        if (o < s.o_rows) {
            const int32_t k = s.o_orderkey[o];
            if (k >= 0 && k < s.size) {
                const int32_t total = s.sum[k];
                if (total > s.threshold) {
                    const int32_t ck = s.o_custkey[o];
                    const int64_t crow = q18_cust(s, ck);
                    ok = crow >= 0;
                    row = bolap_q18_row{ck, k, int32_t(s.o_orderdate[o]), s.o_totalprice[o],
                                        int64_t(total), int32_t(crow), int32_t(o)};
                }
            }
        }
        q18_emit(s, ok, row, res);
    }
}

// Sort keys of the output rows: pass 1 (order key asc, then order row),
// pass 2 (price desc, date asc); stable LSD sorting by pass 1 then pass 2
// gives the reference's order.
__global__ __launch_bounds__(kT) void q18_keys_k(const bolap_q18_row* __restrict__ rows,
                                                 int64_t n, int pass, uint64_t* keys,
                                                 uint32_t* vals) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        if (pass == 1) {
            const bolap_q18_row& r = rows[i];
            keys[i] = (uint64_t(uint32_t(r.orderkey) ^ 0x80000000u) << 32) | uint32_t(r.order_row);
            vals[i] = uint32_t(i);
        } else {
            const bolap_q18_row& r = rows[vals[i]];
            const uint32_t p = ~(uint32_t(r.totalprice) ^ 0x80000000u);
            const uint32_t d = uint32_t(r.orderdate) ^ 0x80000000u;
            keys[i] = (uint64_t(p) << 32) | d;
        }
    }
}

// Up to kRankMax rows are ordered by rank: a row's place is the number of
// rows before it in (price desc, date asc, key asc, order row asc), a strict
// total order because order rows are unique.  Workgroup b ranks rows
// b * 64 .. b * 64 + 63 (one per lane); its four waves split the rows they
// are compared with, so the n^2 comparisons spread over n / 64 CUs.
constexpr int kRankT = 256;
constexpr int kRankMax = 2048;

BOLAP_DEV bool q18_before(const int4& a, const int4& b) {
    if (a.x != b.x) return a.x > b.x;  // totalprice desc
    if (a.y != b.y) return a.y < b.y;  // orderdate asc
    if (a.z != b.z) return a.z < b.z;  // orderkey asc
    return a.w < b.w;                  // order row asc
}

__global__ __launch_bounds__(kRankT) void q18_rank_sort_k(const bolap_q18_row* __restrict__ rows,
                                                          int n, bolap_q18_row* out) {
    constexpr int kW = kRankT / kWave;
    __shared__ int4 s[kRankMax];
    __shared__ int s_rank[kW][kWave];
    for (int i = threadIdx.x; i < n; i += kRankT)
        s[i] = make_int4(rows[i].totalprice, rows[i].orderdate, rows[i].orderkey,
                         rows[i].order_row);
    __syncthreads();
    const int lane = int(threadIdx.x) & (kWave - 1);
    const int wave = int(threadIdx.x) / kWave;
    const int i = int(blockIdx.x) * kWave + lane;
    const int4 me = s[i < n ? i : 0];
    const int per = (n + kW - 1) / kW;
    const int j0 = wave * per;
    const int j1 = j0 + per < n ? j0 + per : n;
    int rank = 0;
    int j = j0;
    // Eight broadcast LDS reads in flight, then their comparisons.
    for (; j + 8 <= j1; j += 8) {
        int4 o[8];
#pragma unroll
        for (int x = 0; x < 8; ++x) o[x] = s[j + x];
#pragma unroll
        for (int x = 0; x < 8; ++x) rank += q18_before(o[x], me) ? 1 : 0;
    }
    for (; j < j1; ++j) rank += q18_before(s[j], me) ? 1 : 0;
    s_rank[wave][lane] = rank;
    __syncthreads();
    if (wave == 0 && i < n) {
        int r = 0;
#pragma unroll
        for (int w = 0; w < kW; ++w) r += s_rank[w][lane];
        out[r] = rows[i];
    }
}

__global__ __launch_bounds__(kT) void q18_gather_k(const bolap_q18_row* __restrict__ rows,
                                                   const uint32_t* __restrict__ order, int64_t n,
                                                   bolap_q18_row* out) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        out[i] = rows[order[i]];
}

// --------------------------------------------------------------- build ---
__global__ __launch_bounds__(kT) void derive_k(const int32_t* __restrict__ price,
                                               const uint8_t* __restrict__ disc,
                                               const uint8_t* __restrict__ rf,
                                               const uint8_t* __restrict__ ls, int32_t n_ls,
                                               const int32_t* __restrict__ okey, int64_t rows,
                                               int32_t* dp, uint8_t* code, int64_t* flag) {
    bool bad = false;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        dp[i] = int32_t(int64_t(price[i]) * (100 - int32_t(disc[i])) / 100);
        code[i] = uint8_t(uint32_t(rf[i]) * uint32_t(n_ls) + uint32_t(ls[i]));
        if (okey && i > 0 && okey[i] < okey[i - 1]) bad = true;
    }
    if (__ballot(bad) && (threadIdx.x & (kWave - 1)) == 0)
        atomicOr(reinterpret_cast<unsigned long long*>(flag), 1ull);
}

__global__ __launch_bounds__(kT) void date_keys_k(const int16_t* __restrict__ date, int64_t rows,
                                                  uint64_t* keys, uint32_t* vals) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        keys[i] = uint64_t(uint16_t(date[i]) ^ 0x8000u);
        vals[i] = uint32_t(i);
    }
}

__global__ __launch_bounds__(kT) void order_map_k(const int32_t* __restrict__ okey, int64_t rows,
                                                  int32_t* map, int64_t map_size) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = okey[i];
        if (k >= 0 && k < map_size) atomicMax(map + k, int32_t(i));
    }
}

// start = first, end = last + 1 lineitem row of each order (one atomic pair
// per run of equal keys within a lane's 8 rows).
__global__ __launch_bounds__(kT) void ranges_k(const int32_t* __restrict__ lkey, int64_t n,
                                               const int32_t* __restrict__ map, int64_t map_size,
                                               int64_t o_rows, uint32_t* start, uint32_t* end) {
    const int64_t units = (n + 7) / 8;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < units;
         u += int64_t(gridDim.x) * kT) {
        const int64_t r0 = u * 8;
        const int m = r0 + 8 <= n ? 8 : int(n - r0);
        int64_t b = r0;
        for (int j = 1; j <= m; ++j) {
            if (j < m && lkey[r0 + j] == lkey[b]) continue;
            const int32_t k = lkey[b];
            if (k >= 0 && k < map_size) {
                const int32_t row = map[k];
                if (row >= 0 && row < o_rows) {
                    atomicMin(start + row, uint32_t(b));
                    atomicMax(end + row, uint32_t(r0 + j));
                }
            }
            b = r0 + j;
        }
    }
}

// Orders without lineitem rows keep the reference's empty range {0, 0}.
__global__ __launch_bounds__(kT) void ranges_fix_k(int64_t o_rows, uint32_t* start,
                                                   const uint32_t* end) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < o_rows;
         i += int64_t(gridDim.x) * kT)
        if (end[i] == 0) start[i] = 0;
}

// ------------------------------------------------------------ host side ---
int read_result(bolap_ctx* ctx, const int64_t* res_dev, int n, int64_t* host) {
    BOLAP_TRY(pinned_reserve(ctx, size_t(n) * sizeof(int64_t)));
    BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, res_dev, size_t(n) * sizeof(int64_t),
                                 hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::memcpy(host, ctx->pinned, size_t(n) * sizeof(int64_t));
    return BOLAP_OK;
}

// "YYYY-MM-DD" -> (year, month, day); false when malformed.
bool parse_ymd(const char* s, int& y, int& m, int& d) {
    if (!s || std::strlen(s) < 10 || s[4] != '-' || s[7] != '-') return false;
    for (int i : {0, 1, 2, 3, 5, 6, 8, 9})
        if (s[i] < '0' || s[i] > '9') return false;
    y = (s[0] - '0') * 1000 + (s[1] - '0') * 100 + (s[2] - '0') * 10 + (s[3] - '0');
    m = (s[5] - '0') * 10 + (s[6] - '0');
    d = (s[8] - '0') * 10 + (s[9] - '0');
    return m >= 1 && m <= 12 && d >= 1 && d <= 31;
}

int days_in_month(int y, int m) {
    static const int kDays[] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    const bool leap = (y % 4 == 0 && y % 100 != 0) || y % 400 == 0;
    return (m == 2 && leap) ? 29 : kDays[m - 1];
}

// std::stod then llround(value * scale) (query_q6.cpp:92-95).
bool parse_scaled(const char* s, int scale, int32_t& out) {
    if (!s) return false;
    char* end = nullptr;
    const double v = std::strtod(s, &end);
    if (end == s) return false;
    out = int32_t(std::llround(v * double(scale)));
    return true;
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" {

int bolap_q6_parse(const char* date, const char* discount, const char* quantity,
                   bolap_q6_params* out) {
    BOLAP_ARG_CHECK(out, "bolap_q6_parse: null output");
    int y, m, d;
    BOLAP_ARG_CHECK(parse_ymd(date, y, m, d), "bolap_q6_parse: bad DATE '%s'", date ? date : "");
    int32_t disc_raw = 0, qty = 0;
    BOLAP_ARG_CHECK(parse_scaled(discount, 100, disc_raw), "bolap_q6_parse: bad DISCOUNT");
    BOLAP_ARG_CHECK(parse_scaled(quantity, 100, qty), "bolap_q6_parse: bad QUANTITY");
    // query_q6.cpp:171-190: [DATE, DATE + 1 year) with the day clamped to the
    // month's length; discount +-0.01; quantity below the limit.
    const int ey = y + 1;
    const int ed = d > days_in_month(ey, m) ? days_in_month(ey, m) : d;
    out->ship_lo = int32_t(int16_t(days_from_civil(y, m, d) - kBaseDays));
    out->ship_hi = int32_t(int16_t(days_from_civil(ey, m, ed) - kBaseDays));
    out->disc_lo = disc_raw - 1;
    out->disc_hi = disc_raw + 1;
    out->qty_lt = qty;
    out->month_lo = y * 12 + (m - 1);
    int32_t mhi = ey * 12 + (m - 1);
    if (ed == 1) mhi -= 1;
    out->month_hi = mhi < out->month_lo ? out->month_lo : mhi;
    return BOLAP_OK;
}

int bolap_q14_parse(const char* date, int32_t* ship_lo, int32_t* ship_hi) {
    BOLAP_ARG_CHECK(ship_lo && ship_hi, "bolap_q14_parse: null output");
    int y, m, d;
    BOLAP_ARG_CHECK(parse_ymd(date, y, m, d), "bolap_q14_parse: bad DATE '%s'", date ? date : "");
    // query_q14.cpp:54-74: one month later, the day clamped to that month.
    int mi = m;  // month index of DATE + 1 month, 0-based
    int ey = y + mi / 12;
    mi %= 12;
    const int em = mi + 1;
    const int ed = d > days_in_month(ey, em) ? days_in_month(ey, em) : d;
    *ship_lo = int32_t(int16_t(days_from_civil(y, m, d) - kBaseDays));
    *ship_hi = int32_t(int16_t(days_from_civil(ey, em, ed) - kBaseDays));
    return BOLAP_OK;
}

namespace {

constexpr int kQ14U = 4;  // units of 8 rows per lane in flight (flat pass)

// One Q14 pass with a partkey array of `size` entries; *overflow is set
// when a part key is >= size (the result is then not valid).
int q14_pass(bolap_ctx* ctx, const bolap_q14_args* a, int64_t size, bolap_q14_result* out,
             bool* overflow) {
    const int64_t units = a->li_rows / 8;
    const int64_t n_zones = (a->li_rows + kZoneRows - 1) / kZoneRows;
    const int64_t groups = a->zones ? n_zones : (units + kT * kQ14U - 1) / (kT * kQ14U);
    int64_t grid = groups < int64_t(ctx->num_cus) * 8 ? groups : int64_t(ctx->num_cus) * 8;
    if (grid < 1) grid = 1;
    WsArena ws;
    const size_t o_packed = ws.plan(size_t(size) * 4);
    const size_t o_flag = ws.plan(size_t(size));
    const size_t o_part = ws.plan(size_t(grid) * 3 * sizeof(int64_t));
    const size_t o_ovf = ws.plan(sizeof(int32_t));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    BOLAP_TRY(pinned_reserve(ctx, 5 * sizeof(int64_t)));
    ws.base = static_cast<char*>(ctx->ws);
    int32_t* packed = ws.at<int32_t>(o_packed);
    int8_t* flag = ws.at<int8_t>(o_flag);
    int32_t* ovf = ws.at<int32_t>(o_ovf);
    int64_t* res = static_cast<int64_t*>(ctx->result_dev);
    {
        TraceScope ts(ctx, "q14_part_map");
        hipLaunchKernelGGL(q14_clear_k, dim3(grid_of(ctx, size, 4)), dim3(kT), 0, ctx->stream,
                           packed, size, ovf);
        BOLAP_TRY(launch_check("q14_clear_k"));
        hipLaunchKernelGGL(q14_part_k, dim3(grid_of(ctx, a->part_rows, 4)), dim3(kT), 0,
                           ctx->stream, a->p_partkey, a->p_promo, a->part_rows, packed, size,
                           ovf);
        BOLAP_TRY(launch_check("q14_part_k"));
        hipLaunchKernelGGL(q14_narrow_k, dim3(grid_of(ctx, size, 4)), dim3(kT), 0, ctx->stream,
                           packed, size, flag);
        BOLAP_TRY(launch_check("q14_narrow_k"));
    }
    Q14K k{a->l_ship, a->l_partkey, a->l_disc_price, a->li_rows, a->ship_lo, a->ship_hi, flag,
           size};
    {
        TraceScope ts(ctx, "q14_lineitem_scan");
        if (a->zones)
            hipLaunchKernelGGL(q14_zoned_k, dim3(unsigned(grid)), dim3(kT), 0, ctx->stream,
                               k, a->zones, ws.at<int64_t>(o_part));
        else
            hipLaunchKernelGGL(q14_flat_k<kQ14U>, dim3(unsigned(grid)), dim3(kT), 0,
                               ctx->stream, k, ws.at<int64_t>(o_part));
        BOLAP_TRY(launch_check("q14_scan_k"));
    }
    const int64_t seq = ++ctx->pinned_seq;
    hipLaunchKernelGGL(q6_merge_k, dim3(1), dim3(kT), 0, ctx->stream, ws.at<int64_t>(o_part),
                       int(grid), res, static_cast<int64_t*>(ctx->pinned_dev), seq, ovf);
    BOLAP_TRY(launch_check("q6_merge_k"));
    BOLAP_TRY(wait_pinned_seq(ctx, 4, seq));
    const int64_t* hp = static_cast<const int64_t*>(ctx->pinned);
    *overflow = hp[3] != 0;
    out->promo_sum = hp[0];
    out->total_sum = hp[1];
    out->rows_emitted = hp[2];
    return BOLAP_OK;
}

}  // namespace

int bolap_q14(bolap_ctx* ctx, const bolap_q14_args* a, bolap_q14_result* out) {
    BOLAP_ARG_CHECK(ctx && a && out, "bolap_q14: null argument");
    BOLAP_ARG_CHECK(a->li_rows >= 0 && a->part_rows >= 0 && a->part_rows < (int64_t(1) << 30),
                    "bolap_q14: bad row counts");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (a->l_ship && a->l_partkey && a->l_disc_price),
                    "bolap_q14: null lineitem column");
    BOLAP_ARG_CHECK(a->part_rows == 0 || (a->p_partkey && a->p_promo),
                    "bolap_q14: null part column");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (aligned(a->l_ship, 16) && aligned(a->l_partkey, 16) &&
                                        aligned(a->l_disc_price, 16)),
                    "bolap_q14: lineitem columns must be 16-byte aligned");
    std::memset(out, 0, sizeof(*out));
    // The reference sizes its partkey array by the largest key
    // (query_q14.cpp:193-223).  Part keys are 1..part_rows in TPC-H, so the
    // first pass covers keys below part_rows * 1.5 + 1024 without a host
    // round trip; a larger key reruns with the exact size.  Keys between the
    // largest part key and the array size hold no part, so both sizes give
    // the same sums.
    int64_t size = a->part_rows + a->part_rows / 2 + 1024;
    const int64_t forced = ctx->knob("q14.array_size", 0);
    if (forced > 0) size = forced;
    bool overflow = false;
    BOLAP_TRY(q14_pass(ctx, a, size, out, &overflow));
    if (overflow) {
        int64_t* res = static_cast<int64_t*>(ctx->result_dev);
        int64_t h[kQ18ResN];
        {
            TraceScope ts(ctx, "q14_part_scan");
            hipLaunchKernelGGL(q18_init_k, dim3(1), dim3(kT), 0, ctx->stream, res);
            BOLAP_TRY(launch_check("q18_init_k"));
            hipLaunchKernelGGL(q18_max_k, dim3(grid_of(ctx, a->part_rows / 4, 4)), dim3(kT), 0,
                               ctx->stream, nullptr, 0, nullptr, 0, a->p_partkey, a->part_rows,
                               res);
            BOLAP_TRY(launch_check("q18_max_k"));
            BOLAP_TRY(read_result(ctx, res, kQ18ResN, h));
        }
        BOLAP_ARG_CHECK(h[kQ18MaxC] >= size, "bolap_q14: part key maximum %lld below %lld",
                        (long long)h[kQ18MaxC], (long long)size);
        BOLAP_TRY(q14_pass(ctx, a, h[kQ18MaxC] + 1, out, &overflow));
        BOLAP_ARG_CHECK(!overflow, "bolap_q14: overflow at the exact size");
        trace_count(ctx, "q14_fallbacks", 1);
    }
    out->promo_revenue =
        out->total_sum > 0 ? 100.0 * double(out->promo_sum) / double(out->total_sum) : 0.0;
    trace_count(ctx, "q14_lineitem_scan_rows_emitted", out->rows_emitted);
    return BOLAP_OK;
}

int bolap_q6(bolap_ctx* ctx, const bolap_q6_args* a, bolap_q6_result* out) {
    BOLAP_ARG_CHECK(ctx && a && out, "bolap_q6: null argument");
    BOLAP_ARG_CHECK(a->rows >= 0, "bolap_q6: negative rows");
    BOLAP_ARG_CHECK(a->rows == 0 || (a->ship && a->disc && a->qty && a->price),
                    "bolap_q6: null column");
    BOLAP_ARG_CHECK(a->rows == 0 || (aligned(a->ship, 16) && aligned(a->qty, 16) &&
                                     aligned(a->price, 16) && aligned(a->disc, 8)),
                    "bolap_q6: columns must be 16-byte aligned (discount 8-byte)");
    Q6K k{a->ship, a->disc, a->qty, a->price, a->rows, a->ship_lo,
          uint32_t(a->ship_hi - a->ship_lo), a->disc_lo, uint32_t(a->disc_hi - a->disc_lo),
          a->qty_lt};
    const int64_t n_zones = (a->rows + kZoneRows - 1) / kZoneRows;
    unsigned grid;
    if (a->zones)
        grid = unsigned(n_zones < int64_t(ctx->num_cus) * 8 ? (n_zones < 1 ? 1 : n_zones)
                                                            : int64_t(ctx->num_cus) * 8);
    else
        grid = grid_of(ctx, (a->rows / 8 + 1) / 2, 4);
    WsArena ws;
    const size_t o_part = ws.plan(size_t(grid) * 3 * sizeof(int64_t));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    BOLAP_TRY(pinned_reserve(ctx, 5 * sizeof(int64_t)));
    ws.base = static_cast<char*>(ctx->ws);
    int64_t* res = static_cast<int64_t*>(ctx->result_dev);
    {
        TraceScope ts(ctx, a->zones ? "q6_zoned" : "q6_flat");
        if (a->zones)
            hipLaunchKernelGGL(q6_zoned_k, dim3(grid), dim3(kT), 0, ctx->stream, k, a->zones,
                               n_zones, ws.at<int64_t>(o_part));
        else
            hipLaunchKernelGGL(q6_flat_k, dim3(grid), dim3(kT), 0, ctx->stream, k,
                               ws.at<int64_t>(o_part));
        BOLAP_TRY(launch_check("q6 scan"));
    }
    const int64_t seq = ++ctx->pinned_seq;
    hipLaunchKernelGGL(q6_merge_k, dim3(1), dim3(kT), 0, ctx->stream, ws.at<int64_t>(o_part),
                       int(grid), res, static_cast<int64_t*>(ctx->pinned_dev), seq);
    BOLAP_TRY(launch_check("q6_merge_k"));
    BOLAP_TRY(wait_pinned_seq(ctx, 4, seq));
    const int64_t* h = static_cast<const int64_t*>(ctx->pinned);
    out->revenue = h[0];
    out->rows_emitted = h[1];
    out->rows_scanned = a->zones ? h[2] : a->rows;
    trace_count(ctx, "q6_lineitem_scan_rows_scanned", out->rows_scanned);
    trace_count(ctx, "q6_lineitem_scan_rows_emitted", out->rows_emitted);
    return BOLAP_OK;
}

int bolap_q6_zones(bolap_ctx* ctx, int64_t rows, const int16_t* ship, const uint8_t* disc,
                   const int16_t* qty, bolap_zone* zones_out) {
    BOLAP_ARG_CHECK(ctx && rows >= 0, "bolap_q6_zones: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(ship && disc && qty && zones_out, "bolap_q6_zones: null array");
    const int64_t n_zones = (rows + kZoneRows - 1) / kZoneRows;
    const int64_t cap = int64_t(ctx->num_cus) * 8;
    hipLaunchKernelGGL(zones_k, dim3(unsigned(n_zones < cap ? n_zones : cap)), dim3(kT), 0,
                       ctx->stream, ship, disc, qty, rows, n_zones, zones_out);
    return launch_check("zones_k");
}

int bolap_q6_cluster(bolap_ctx* ctx, int64_t rows, const int16_t* ship, const uint8_t* disc,
                     const int16_t* qty, const int32_t* price, int16_t* ship_out,
                     uint8_t* disc_out, int16_t* qty_out, int32_t* price_out,
                     uint32_t* perm_out) {
    BOLAP_ARG_CHECK(ctx && rows >= 0 && rows < (int64_t(1) << 32),
                    "bolap_q6_cluster: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(ship && disc && qty && price && ship_out && disc_out && qty_out && price_out,
                    "bolap_q6_cluster: null array");
    WsArena ws;
    const size_t o_k = ws.plan(size_t(rows) * 8);
    const size_t o_v = ws.plan(size_t(rows) * 4);
    const size_t o_s = ws.plan(radix_sort_scratch_bytes(ctx, rows));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    uint64_t* keys = ws.at<uint64_t>(o_k);
    uint32_t* vals = ws.at<uint32_t>(o_v);
    hipLaunchKernelGGL(shard_keys_k, dim3(grid_of(ctx, rows, 16)), dim3(kT), 0, ctx->stream,
                       ship, disc, qty, rows, keys, vals);
    BOLAP_TRY(launch_check("shard_keys_k"));
    BOLAP_TRY(radix_sort_pairs(ctx, keys, vals, rows, 30, ws.at<void>(o_s), nullptr));
    hipLaunchKernelGGL(gather_q6_k, dim3(grid_of(ctx, rows, 16)), dim3(kT), 0, ctx->stream, vals,
                       rows, ship, disc, qty, price, ship_out, disc_out, qty_out, price_out);
    BOLAP_TRY(launch_check("gather_q6_k"));
    if (perm_out)
        BOLAP_HIP_TRY(hipMemcpyAsync(perm_out, vals, size_t(rows) * 4, hipMemcpyDeviceToDevice,
                                     ctx->stream));
    return BOLAP_OK;
}

int bolap_q18(bolap_ctx* ctx, const bolap_q18_args* a, bolap_q18_row* rows_out,
              int64_t max_rows, bolap_q18_stats* st) {
    BOLAP_ARG_CHECK(ctx && a && st && (rows_out || max_rows == 0) && max_rows >= 0,
                    "bolap_q18: null argument");
    BOLAP_ARG_CHECK(a->li_rows >= 0 && a->order_rows >= 0 && a->cust_rows >= 0,
                    "bolap_q18: negative rows");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (a->l_orderkey && a->l_quantity),
                    "bolap_q18: null lineitem column");
    BOLAP_ARG_CHECK(a->order_rows == 0 || (a->o_orderkey && a->o_custkey && a->o_orderdate &&
                                           a->o_totalprice),
                    "bolap_q18: null orders column");
    BOLAP_ARG_CHECK(a->cust_rows == 0 || a->c_custkey, "bolap_q18: null customer column");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (aligned(a->l_orderkey, 16) && aligned(a->l_quantity, 16)),
                    "bolap_q18: lineitem columns must be 16-byte aligned");
    BOLAP_ARG_CHECK(a->order_rows < (int64_t(1) << 31), "bolap_q18: too many orders");
    std::memset(st, 0, sizeof(*st));
    int64_t* res = static_cast<int64_t*>(ctx->result_dev);
    // The reference's density checks read the first and last keys only
    // (query_q18.cpp:194-199); the key maxima are needed only off the dense
    // path.
    int64_t h[kQ18ResN];
    int32_t ends[4] = {0, 0, 0, 0};
    {
        TraceScope ts(ctx, "q18_bounds");
        BOLAP_TRY(pinned_reserve(ctx, 4 * sizeof(int32_t)));
        int32_t* pe = static_cast<int32_t*>(ctx->pinned);
        hipLaunchKernelGGL(q18_ends_k, dim3(1), dim3(kWave), 0, ctx->stream, a->o_orderkey,
                           a->order_rows, a->c_custkey, a->cust_rows, pe);
        BOLAP_TRY(launch_check("q18_ends_k"));
        BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::memcpy(ends, pe, sizeof(ends));
    }
    hipLaunchKernelGGL(q18_init_k, dim3(1), dim3(kT), 0, ctx->stream, res);
    BOLAP_TRY(launch_check("q18_init_k"));
    {
        const bool dense_o = a->order_rows > 0 && ends[0] == 1 && ends[1] == a->order_rows;
        const bool dense_c = a->cust_rows > 0 && ends[2] == 1 && ends[3] == a->cust_rows;
        h[kQ18MaxL] = h[kQ18MaxO] = h[kQ18MaxC] = -1;
        if (!dense_o || !dense_c) {
            TraceScope ts(ctx, "q18_bounds");
            const int64_t most = a->li_rows > a->order_rows ? a->li_rows : a->order_rows;
            hipLaunchKernelGGL(q18_max_k, dim3(grid_of(ctx, most / 4, 4)), dim3(kT), 0,
                               ctx->stream, dense_o ? nullptr : a->l_orderkey,
                               dense_o ? 0 : a->li_rows, dense_o ? nullptr : a->o_orderkey,
                               dense_o ? 0 : a->order_rows, a->c_custkey, a->cust_rows, res);
            BOLAP_TRY(launch_check("q18_max_k"));
            BOLAP_TRY(read_result(ctx, res, kQ18ResN, h));
        }
    }
    const bool dense_orders = a->order_rows > 0 && ends[0] == 1 && ends[1] == a->order_rows;
    const bool dense_cust = a->cust_rows > 0 && ends[2] == 1 && ends[3] == a->cust_rows;
    st->dense_path = dense_orders ? 1 : 0;
    int64_t size = 0;
    if (dense_orders) {
        if (a->cust_rows == 0) return BOLAP_OK;  // query_q18.cpp:212-218
        size = a->order_rows + 1;
    } else {
        const int64_t mk = h[kQ18MaxO] > h[kQ18MaxL] ? h[kQ18MaxO] : h[kQ18MaxL];
        if (mk < 0) return BOLAP_OK;
        size = mk + 1;
    }
    const bool need_map = !dense_cust || !dense_orders;
    const int64_t map_size = need_map ? h[kQ18MaxC] + 1 : 0;
    if (need_map && map_size <= 0) return BOLAP_OK;  // no customer key >= 0
    const int64_t cap = max_rows;
    WsArena ws;
    const size_t o_sum = ws.plan(size_t(size) * 4);
    const size_t o_map = ws.plan(size_t(map_size > 0 ? map_size : 1) * 4);
    const size_t o_rows = ws.plan(size_t(cap > 0 ? cap : 1) * sizeof(bolap_q18_row));
    const size_t o_sorted = ws.plan(size_t(cap > 0 ? cap : 1) * sizeof(bolap_q18_row));
    const size_t o_keys = ws.plan(size_t(cap > 0 ? cap : 1) * 8);
    const size_t o_vals = ws.plan(size_t(cap > 0 ? cap : 1) * 4);
    const size_t o_rs = ws.plan(radix_sort_scratch_bytes(ctx, cap > 0 ? cap : 1));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    int32_t* sum = ws.at<int32_t>(o_sum);
    int32_t* map = ws.at<int32_t>(o_map);
    bolap_q18_row* rows_dev = ws.at<bolap_q18_row>(o_rows);
    if (need_map) {
        hipLaunchKernelGGL(fill_i32_k, dim3(grid_of(ctx, map_size, 4)), dim3(kT), 0, ctx->stream,
                           map, map_size, -1);
        BOLAP_TRY(launch_check("fill_i32_k"));
        hipLaunchKernelGGL(cust_map_k, dim3(grid_of(ctx, a->cust_rows, 4)), dim3(kT), 0,
                           ctx->stream, a->c_custkey, a->cust_rows, map, map_size);
        BOLAP_TRY(launch_check("cust_map_k"));
    }
    Q18Sel sel{sum, size, a->quantity_threshold, a->o_orderkey, a->o_custkey, a->o_orderdate,
               a->o_totalprice, a->order_rows, a->cust_rows, dense_cust ? 1 : 0, map,
               map_size, rows_dev, cap};
    bool sorted_pass = a->orderkey_sorted != 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        {
            TraceScope ts(ctx, "q18_sum");
            hipLaunchKernelGGL(fill_i32_k, dim3(grid_of(ctx, size, 8)), dim3(kT), 0,
                               ctx->stream, sum, size, 0);
            BOLAP_TRY(launch_check("fill_i32_k"));
            const unsigned g = grid_of(ctx, (a->li_rows + 7) / 8, 8);
            if (sorted_pass)
                hipLaunchKernelGGL(q18_sum_sorted_k, dim3(g), dim3(kT), 0, ctx->stream,
                                   a->l_orderkey, a->l_quantity, a->li_rows, sum, size, res);
            else
                hipLaunchKernelGGL(q18_sum_atomic_k, dim3(g), dim3(kT), 0, ctx->stream,
                                   a->l_orderkey, a->l_quantity, a->li_rows, sum, size);
            BOLAP_TRY(launch_check("q18 sum"));
        }
        {
            TraceScope ts(ctx, "q18_select");
            if (dense_orders)
                hipLaunchKernelGGL(q18_select_dense_k, dim3(grid_of(ctx, size, 8)), dim3(kT), 0,
                                   ctx->stream, sel, res);
            else
                hipLaunchKernelGGL(q18_select_orders_k, dim3(grid_of(ctx, a->order_rows, 8)),
                                   dim3(kT), 0, ctx->stream, sel, res);
            BOLAP_TRY(launch_check("q18 select"));
        }
        BOLAP_TRY(read_result(ctx, res, kQ18ResN, h));
        if (sorted_pass && h[kQ18Unsorted]) {
            // Keys out of order: rerun the sums with atomics.
            sorted_pass = false;
            st->fallbacks = 1;
            hipLaunchKernelGGL(q18_init_k, dim3(1), dim3(kT), 0, ctx->stream, res);
            BOLAP_TRY(launch_check("q18_init_k"));
            continue;
        }
        break;
    }
    const int64_t n = h[kQ18Count];
    st->rows = n;
    st->sorted_pass = sorted_pass ? 1 : 0;
    trace_count(ctx, "q18_query_output_rows", n);
    if (n > cap) {
        set_error("bolap_q18: %lld result rows exceed max_rows %lld", (long long)n,
                  (long long)cap);
        return BOLAP_ERR_CAPACITY;
    }
    if (n == 0) return BOLAP_OK;
    uint64_t* keys = ws.at<uint64_t>(o_keys);
    uint32_t* vals = ws.at<uint32_t>(o_vals);
    bolap_q18_row* sorted = ws.at<bolap_q18_row>(o_sorted);
    if (n <= kRankMax) {
        TraceScope ts(ctx, "q18_sort");
        hipLaunchKernelGGL(q18_rank_sort_k, dim3(unsigned((n + kWave - 1) / kWave)),
                           dim3(kRankT), 0, ctx->stream, rows_dev,
                           int(n), sorted);
        BOLAP_TRY(launch_check("q18_rank_sort_k"));
    } else {
        TraceScope ts(ctx, "q18_sort");
        const unsigned g = grid_of(ctx, n, 4);
        hipLaunchKernelGGL(q18_keys_k, dim3(g), dim3(kT), 0, ctx->stream, rows_dev, n, 1, keys,
                           vals);
        BOLAP_TRY(launch_check("q18_keys_k"));
        BOLAP_TRY(radix_sort_pairs(ctx, keys, vals, n, 64, ws.at<void>(o_rs), nullptr));
        hipLaunchKernelGGL(q18_keys_k, dim3(g), dim3(kT), 0, ctx->stream, rows_dev, n, 2, keys,
                           vals);
        BOLAP_TRY(launch_check("q18_keys_k"));
        BOLAP_TRY(radix_sort_pairs(ctx, keys, vals, n, 64, ws.at<void>(o_rs), nullptr));
        hipLaunchKernelGGL(q18_gather_k, dim3(g), dim3(kT), 0, ctx->stream, rows_dev, vals, n,
                           sorted);
        BOLAP_TRY(launch_check("q18_gather_k"));
    }
    BOLAP_TRY(pinned_reserve(ctx, size_t(n) * sizeof(bolap_q18_row)));
    BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, sorted, size_t(n) * sizeof(bolap_q18_row),
                                 hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::memcpy(rows_out, ctx->pinned, size_t(n) * sizeof(bolap_q18_row));
    return BOLAP_OK;
}

int bolap_derive_lineitem(bolap_ctx* ctx, int64_t rows, const int32_t* price,
                          const uint8_t* disc, const uint8_t* rf, const uint8_t* ls,
                          int32_t n_ls, const int32_t* orderkey, int32_t* disc_price_out,
                          uint8_t* code_out, int32_t* orderkey_sorted) {
    BOLAP_ARG_CHECK(ctx && rows >= 0 && orderkey_sorted, "bolap_derive_lineitem: bad arguments");
    BOLAP_ARG_CHECK(n_ls >= 1 && n_ls <= 255, "bolap_derive_lineitem: n_ls %d", n_ls);
    *orderkey_sorted = 1;
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(price && disc && rf && ls && disc_price_out && code_out,
                    "bolap_derive_lineitem: null array");
    int64_t* res = static_cast<int64_t*>(ctx->result_dev);
    hipLaunchKernelGGL(q18_init_k, dim3(1), dim3(kT), 0, ctx->stream, res);
    BOLAP_TRY(launch_check("q18_init_k"));
    hipLaunchKernelGGL(derive_k, dim3(grid_of(ctx, rows, 16)), dim3(kT), 0, ctx->stream, price,
                       disc, rf, ls, n_ls, orderkey, rows, disc_price_out, code_out,
                       res + kQ18Unsorted);
    BOLAP_TRY(launch_check("derive_k"));
    int64_t h[kQ18ResN];
    BOLAP_TRY(read_result(ctx, res, kQ18ResN, h));
    *orderkey_sorted = (orderkey && h[kQ18Unsorted]) ? 0 : 1;
    return BOLAP_OK;
}

int bolap_date_order(bolap_ctx* ctx, int64_t rows, const int16_t* date, uint32_t* order_out) {
    BOLAP_ARG_CHECK(ctx && rows >= 0 && rows < (int64_t(1) << 32),
                    "bolap_date_order: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(date && order_out, "bolap_date_order: null array");
    WsArena ws;
    const size_t o_k = ws.plan(size_t(rows) * 8);
    const size_t o_s = ws.plan(radix_sort_scratch_bytes(ctx, rows));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    uint64_t* keys = ws.at<uint64_t>(o_k);
    hipLaunchKernelGGL(date_keys_k, dim3(grid_of(ctx, rows, 16)), dim3(kT), 0, ctx->stream, date,
                       rows, keys, order_out);
    BOLAP_TRY(launch_check("date_keys_k"));
    return radix_sort_pairs(ctx, keys, order_out, rows, 16, ws.at<void>(o_s), nullptr);
}

int bolap_order_ranges(bolap_ctx* ctx, int64_t o_rows, const int32_t* o_orderkey,
                       int64_t map_size, int32_t* map_out, int64_t li_rows,
                       const int32_t* l_orderkey, uint32_t* start_out, uint32_t* end_out) {
    BOLAP_ARG_CHECK(ctx && o_rows >= 0 && li_rows >= 0 && map_size >= 0 &&
                        o_rows < (int64_t(1) << 31) && li_rows < (int64_t(1) << 32),
                    "bolap_order_ranges: bad arguments");
    BOLAP_ARG_CHECK((o_rows == 0 || (o_orderkey && start_out && end_out)) &&
                        (map_size == 0 || map_out) && (li_rows == 0 || l_orderkey),
                    "bolap_order_ranges: null array");
    if (map_size > 0) {
        hipLaunchKernelGGL(fill_i32_k, dim3(grid_of(ctx, map_size, 8)), dim3(kT), 0, ctx->stream,
                           map_out, map_size, -1);
        BOLAP_TRY(launch_check("fill_i32_k"));
        if (o_rows > 0) {
            hipLaunchKernelGGL(order_map_k, dim3(grid_of(ctx, o_rows, 8)), dim3(kT), 0,
                               ctx->stream, o_orderkey, o_rows, map_out, map_size);
            BOLAP_TRY(launch_check("order_map_k"));
        }
    }
    if (o_rows == 0) return BOLAP_OK;
    hipLaunchKernelGGL(fill_i32_k, dim3(grid_of(ctx, o_rows, 8)), dim3(kT), 0, ctx->stream,
                       reinterpret_cast<int32_t*>(start_out), o_rows, -1);
    BOLAP_TRY(launch_check("fill_i32_k"));
    hipLaunchKernelGGL(fill_i32_k, dim3(grid_of(ctx, o_rows, 8)), dim3(kT), 0, ctx->stream,
                       reinterpret_cast<int32_t*>(end_out), o_rows, 0);
    BOLAP_TRY(launch_check("fill_i32_k"));
    if (li_rows > 0 && map_size > 0) {
        hipLaunchKernelGGL(ranges_k, dim3(grid_of(ctx, (li_rows + 7) / 8, 8)), dim3(kT), 0,
                           ctx->stream, l_orderkey, li_rows, map_out, map_size, o_rows,
                           start_out, end_out);
        BOLAP_TRY(launch_check("ranges_k"));
    }
    hipLaunchKernelGGL(ranges_fix_k, dim3(grid_of(ctx, o_rows, 8)), dim3(kT), 0, ctx->stream,
                       o_rows, start_out, end_out);
    return launch_check("ranges_fix_k");
}

}  // extern "C"
