// tpch_supp.hip — TPC-H Q2, Q11, Q15 and Q20, the supplier-side pipelines,
// over the reference Database layout (partsupp stably sorted by partkey,
// bespoke_tpch/builder_impl.cpp:769-777; supplier rows in table order).
//
// Q2 walks each qualifying part's partsupp run from one lane (the run is
// found from the row's stride block or a binary search, as the reference
// does) and appends the partsupp rows at the minimum cost; a radix sort of
// those row numbers restores the reference's emission order.  Q11 sums each
// run of equal partkeys from the lane holding its first row.  Q15 adds the
// window's discounted prices into a dense per-supplier array with L2
// atomics.  Q20 sums quantities per (partkey, suppkey) in an open-addressing
// HBM table sized from an estimate of the group count, and reruns with the
// exact size when the estimate was short (a flag in the insert kernel).
//
// Wave-aggregated appends (one atomic per wave) compact every output list.
#include <algorithm>
#include <cstring>
#include <vector>

#include "radix.hpp"

namespace bolap {
namespace {

constexpr int kT = 256;

inline unsigned grid_of(bolap_ctx* ctx, int64_t items, int per_cu = 4) {
    int64_t b = (items + kT - 1) / kT;
    const int64_t cap = int64_t(ctx->num_cus) * per_cu;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

__global__ __launch_bounds__(kT) void zero_words_k(uint32_t* __restrict__ p, int64_t words) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < words;
         i += int64_t(gridDim.x) * kT)
        p[i] = 0;
}

// Zeroes `bytes` (rounded up to whole words) with a kernel, so a captured
// call holds no memset node.
int zero_dev(bolap_ctx* ctx, void* p, size_t bytes) {
    const int64_t words = int64_t((bytes + 3) / 4);
    hipLaunchKernelGGL(zero_words_k, dim3(grid_of(ctx, words)), dim3(kT), 0, ctx->stream,
                       static_cast<uint32_t*>(p), words);
    return launch_check("zero_words_k");
}

int fetch(bolap_ctx* ctx, void* host, const void* dev, size_t bytes) {
    if (bytes == 0) return BOLAP_OK;
    BOLAP_HIP_TRY(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BOLAP_OK;
}

// ---------------------------------------------------------------------------
// Q2: per qualifying part, the partsupp rows of in-region suppliers at the
// minimum supply cost.
// ---------------------------------------------------------------------------
struct Q2K {
    int64_t p_rows;
    const int32_t* size;
    const uint8_t* type;
    const uint8_t* type_ok;  // n_types flags, device copy
    int32_t n_types;
    int32_t target;
    int64_t ps_rows;
    const int32_t* pk;
    const int32_t* sk;
    const int32_t* cost;
    int64_t stride;           // ps_rows / p_rows when it divides, else 0
    int64_t s_rows;
    const int32_t* s_nation;
    const uint8_t* nation_ok;  // n_nation flags, device copy
    int32_t n_nation;
};

// Supplier row suppkey - 1 when it is a row of a supplier in the region,
// else -1 (query_q2.cpp:338-342; nations outside the nation rows are out).
BOLAP_DEV int64_t q2_supplier(const Q2K& a, int32_t suppkey) {
    const uint32_t srow = uint32_t(suppkey - 1);
    if (srow >= uint64_t(a.s_rows)) return -1;
    const uint32_t nk = uint32_t(a.s_nation[srow]);
    return (nk < uint32_t(a.n_nation) && a.nation_ok[nk]) ? int64_t(srow) : -1;
}

__global__ __launch_bounds__(kT) void q2_k(Q2K a, uint64_t* __restrict__ rows_out,
                                           unsigned long long* __restrict__ n_out) {
    const int64_t stride_g = int64_t(gridDim.x) * kT;
    // every lane of a wave runs the same number of rounds (wave_append)
    const int64_t rounds = (a.p_rows + stride_g - 1) / stride_g;
    for (int64_t r = 0; r < rounds; ++r) {
        const int64_t prow = r * stride_g + int64_t(blockIdx.x) * kT + threadIdx.x;
        bool ok = false;
        if (prow < a.p_rows) {
            const uint32_t t = a.type[prow];
            ok = a.size[prow] == a.target && t < uint32_t(a.n_types) && a.type_ok[t];
        }
        int64_t s = 0, e = 0;
        int32_t best = INT32_MAX;
        if (ok) {
            const int32_t key = int32_t(prow + 1);
            bool found = false;
            if (a.stride > 0) {  // the builder's fixed stride (:288-312)
                const int64_t base = prow * a.stride;
                if (base + a.stride <= a.ps_rows && a.pk[base] == key &&
                    a.pk[base + a.stride - 1] == key && (base == 0 || a.pk[base - 1] != key) &&
                    (base + a.stride == a.ps_rows || a.pk[base + a.stride] != key)) {
                    s = base;
                    e = base + a.stride;
                    found = true;
                }
            }
            if (!found) {  // lower_bound, then the run (:313-324)
                int64_t lo = 0, hi = a.ps_rows;
                while (lo < hi) {
                    const int64_t mid = lo + (hi - lo) / 2;
                    if (a.pk[mid] < key) lo = mid + 1;
                    else hi = mid;
                }
                s = lo;
                e = lo;
                while (e < a.ps_rows && a.pk[e] == key) ++e;
            }
            for (int64_t i = s; i < e; ++i)
                if (q2_supplier(a, a.sk[i]) >= 0 && a.cost[i] < best) best = a.cost[i];
        }
        // rows at the minimum, one round per run position (runs are short;
        // a run holds fewer than 2^32 rows)
        const uint32_t len = (ok && best != INT32_MAX) ? uint32_t(e - s) : 0u;
        uint32_t maxlen = len;
        for (int d = kWave / 2; d > 0; d >>= 1) {
            const uint32_t o = uint32_t(__shfl_xor(int(maxlen), d, kWave));
            maxlen = o > maxlen ? o : maxlen;
        }
        for (uint32_t j = 0; j < maxlen; ++j) {
            bool emit = false;
            if (j < len) {
                const int64_t i = s + int64_t(j);
                emit = a.cost[i] == best && q2_supplier(a, a.sk[i]) >= 0;
            }
            const unsigned long long slot = wave_append(emit, n_out);
            if (emit) rows_out[slot] = uint64_t(s + int64_t(j));
        }
    }
}

// (part row, supplier row) of each sorted partsupp row: the part key is its
// row + 1 and the supplier row its suppkey - 1 by the pipeline's own rule.
__global__ __launch_bounds__(kT) void q2_pairs_k(const uint64_t* __restrict__ rows, int64_t n,
                                                 const int32_t* __restrict__ pk,
                                                 const int32_t* __restrict__ sk,
                                                 int32_t* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const int64_t r = int64_t(rows[i]);
        out[2 * i] = pk[r] - 1;
        out[2 * i + 1] = sk[r] - 1;
    }
}

// ---------------------------------------------------------------------------
// Q11: suppliers of one nation, per-partkey sums over partsupp runs.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void mark_nation_k(const int32_t* __restrict__ key,
                                                    const int32_t* __restrict__ nation,
                                                    int64_t rows, int32_t target, int64_t n,
                                                    uint8_t* __restrict__ mark) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = key[i];
        if (nation[i] == target && k >= 0 && k < n) mark[k] = 1;
    }
}

struct Q11K {
    int64_t rows;
    const int32_t* pk;
    const int32_t* sk;
    const int32_t* value;
    const uint8_t* mark;
    int64_t n_mark;
};

// Each workgroup owns the contiguous row range [b * chunk, (b + 1) * chunk)
// and writes the groups whose first row lies there into its own region of
// the group arrays (capacity `chunk`: a range holds at most that many run
// heads), so the appends meet in LDS and never on one global counter.  The
// lane holding a run's first row sums the run (:163-186), reading past the
// range end when the run continues there.
__global__ __launch_bounds__(kT) void q11_runs_k(Q11K a, int64_t chunk, int32_t* __restrict__ g_pk,
                                                 int64_t* __restrict__ g_sum,
                                                 uint32_t* __restrict__ blk_n,
                                                 unsigned long long* __restrict__ n_groups,
                                                 unsigned long long* __restrict__ total) {
    __shared__ unsigned long long s_n;
    __shared__ uint64_t s_tot[kT / kWave];
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int64_t lo = int64_t(blockIdx.x) * chunk;
    const int64_t hi = lo + chunk < a.rows ? lo + chunk : a.rows;
    int32_t* r_pk = g_pk + lo;
    int64_t* r_sum = g_sum + lo;
    int64_t my_total = 0;
    for (int64_t t = lo; t < hi; t += kT) {
        const int64_t i = t + threadIdx.x;
        bool head = false;
        int32_t key = 0;
        if (i < hi) {
            key = a.pk[i];
            head = i == 0 || a.pk[i - 1] != key;
        }
        int64_t sum = 0;
        if (head) {
            for (int64_t j = i; j < a.rows && a.pk[j] == key; ++j) {
                const int32_t s = a.sk[j];
                if (s >= 0 && s < a.n_mark && a.mark[s]) sum += a.value[j];
            }
        }
        my_total += sum;
        const bool emit = head && sum != 0;
        const unsigned long long slot = wave_append(emit, &s_n);
        if (emit) {
            r_pk[slot] = key;
            r_sum[slot] = sum;
        }
    }
    const uint64_t wt = wave_sum_u64(uint64_t(my_total));
    if ((threadIdx.x & (kWave - 1)) == 0) s_tot[threadIdx.x / kWave] = wt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t bt = 0;
        for (int w = 0; w < kT / kWave; ++w) bt += s_tot[w];
        blk_n[blockIdx.x] = uint32_t(s_n);
        if (s_n) atomicAdd(n_groups, s_n);
        if (bt) atomicAdd(total, (unsigned long long)bt);
    }
}

// Order-preserving unsigned images of the signed values and keys.
BOLAP_DEV uint64_t ob64(int64_t v) { return uint64_t(v) ^ (uint64_t(1) << 63); }
BOLAP_DEV uint32_t ob32(int32_t v) { return uint32_t(v) ^ 0x80000000u; }

BOLAP_DEV uint64_t wave_max_u64(uint64_t x) {
    for (int d = kWave / 2; d > 0; d >>= 1) {
        const uint64_t o = __shfl_xor(x, d, kWave);
        x = o > x ? o : x;
    }
    return x;
}

// Groups with (double)sum > (double)total * fraction (:194-203).  Workgroup b
// filters region b: it counts its kept groups, takes their output range with
// one atomic, then writes them there.  The kept groups' value and key ranges
// go to range[4b..4b+3] (max of ~value, max of value, max of ~key, max of
// key, over the order-preserving unsigned images; 0 for none).
__global__ __launch_bounds__(kT) void q11_filter_k(const int32_t* __restrict__ g_pk,
                                                   const int64_t* __restrict__ g_sum,
                                                   const uint32_t* __restrict__ blk_n,
                                                   int64_t chunk,
                                                   const unsigned long long* __restrict__ total,
                                                   double fraction, int32_t* __restrict__ o_pk,
                                                   int64_t* __restrict__ o_sum,
                                                   unsigned long long* __restrict__ n_out,
                                                   uint64_t* __restrict__ range) {
    __shared__ uint32_t s_cnt[kT / kWave];
    __shared__ uint64_t s_r[4][kT / kWave];
    __shared__ unsigned long long s_base, s_pos;
    const int64_t n = blk_n[blockIdx.x];
    const int32_t* r_pk = g_pk + int64_t(blockIdx.x) * chunk;
    const int64_t* r_sum = g_sum + int64_t(blockIdx.x) * chunk;
    const double threshold = double(int64_t(*total)) * fraction;
    uint32_t kept = 0;
    for (int64_t i = threadIdx.x; i < n; i += kT) kept += double(r_sum[i]) > threshold ? 1u : 0u;
    const uint32_t wk = uint32_t(wave_sum_u64(kept));
    if ((threadIdx.x & (kWave - 1)) == 0) s_cnt[threadIdx.x / kWave] = wk;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long bk = 0;
        for (int w = 0; w < kT / kWave; ++w) bk += s_cnt[w];
        s_base = bk ? atomicAdd(n_out, bk) : 0;
        s_pos = 0;
    }
    __syncthreads();
    uint64_t r[4] = {0, 0, 0, 0};
    for (int64_t t = 0; t < n; t += kT) {
        const int64_t i = t + threadIdx.x;
        const bool keep = i < n && double(r_sum[i]) > threshold;
        const unsigned long long slot = wave_append(keep, &s_pos);
        if (keep) {
            const int32_t pk = r_pk[i];
            const int64_t v = r_sum[i];
            o_pk[s_base + slot] = pk;
            o_sum[s_base + slot] = v;
            const uint64_t vi = ob64(v), ki = ob32(pk);
            r[0] = ~vi > r[0] ? ~vi : r[0];
            r[1] = vi > r[1] ? vi : r[1];
            r[2] = (~ki & 0xffffffffu) > r[2] ? (~ki & 0xffffffffu) : r[2];
            r[3] = ki > r[3] ? ki : r[3];
        }
    }
    for (int j = 0; j < 4; ++j) {
        const uint64_t m = wave_max_u64(r[j]);
        if ((threadIdx.x & (kWave - 1)) == 0) s_r[j][threadIdx.x / kWave] = m;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        uint64_t m = 0;
        for (int w = 0; w < kT / kWave; ++w) m = s_r[threadIdx.x][w] > m ? s_r[threadIdx.x][w] : m;
        range[int64_t(blockIdx.x) * 4 + threadIdx.x] = m;
    }
}

// Folds the workgroups' range words into out[0..3] (one workgroup).
__global__ __launch_bounds__(kT) void q11_range_k(const uint64_t* __restrict__ range,
                                                  int64_t blocks, uint64_t* __restrict__ out) {
    __shared__ uint64_t s_r[4][kT / kWave];
    uint64_t r[4] = {0, 0, 0, 0};
    for (int64_t b = threadIdx.x; b < blocks; b += kT)
        for (int j = 0; j < 4; ++j) r[j] = range[b * 4 + j] > r[j] ? range[b * 4 + j] : r[j];
    for (int j = 0; j < 4; ++j) {
        const uint64_t m = wave_max_u64(r[j]);
        if ((threadIdx.x & (kWave - 1)) == 0) s_r[j][threadIdx.x / kWave] = m;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        uint64_t m = 0;
        for (int w = 0; w < kT / kWave; ++w) m = s_r[threadIdx.x][w] > m ? s_r[threadIdx.x][w] : m;
        out[threadIdx.x] = m;
    }
}

// Sort key of a kept group: (vmax - value) above (key - kmin), so ascending
// keys give value descending, then key ascending.
__global__ __launch_bounds__(kT) void q11_keys_k(const int32_t* __restrict__ o_pk,
                                                 const int64_t* __restrict__ o_sum, int64_t n,
                                                 uint64_t vmax, uint32_t kmin, int kbits,
                                                 uint64_t* __restrict__ keys) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        keys[i] = ((vmax - ob64(o_sum[i])) << kbits) | uint64_t(ob32(o_pk[i]) - kmin);
}

__global__ __launch_bounds__(kT) void q11_values_k(const uint64_t* __restrict__ keys, int64_t n,
                                                   uint64_t vmax, int kbits,
                                                   int64_t* __restrict__ o_sum) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        o_sum[i] = int64_t((vmax - (keys[i] >> kbits)) ^ (uint64_t(1) << 63));
}

// ---------------------------------------------------------------------------
// Q15: revenue per supplier key over a ship-date window.
// ---------------------------------------------------------------------------
struct Q15K {
    int64_t rows;
    const int16_t* ship;
    const int32_t* sk;
    const int32_t* price;
    int32_t lo;
    uint32_t span;  // (uint32)(hi - lo), the reference's unsigned window test
    int64_t n_keys;
};

BOLAP_DEV void q15_row(const Q15K& a, int64_t i, int32_t ship, unsigned long long* rev,
                       uint64_t& emitted) {
    if (uint32_t(ship - a.lo) >= a.span) return;
    ++emitted;
    const uint32_t k = uint32_t(a.sk[i]);
    if (k < uint64_t(a.n_keys)) atomicAdd(&rev[k], (unsigned long long)(int64_t(a.price[i])));
}

__global__ __launch_bounds__(kT) void q15_scan_k(Q15K a, unsigned long long* __restrict__ rev,
                                                 unsigned long long* __restrict__ emitted_out) {
    uint64_t emitted = 0;
    const int64_t units = a.rows / 8;  // eight ship days per 16-byte load
    const int64_t stride = int64_t(gridDim.x) * kT;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < units; u += stride) {
        const VecT<int16_t, 8> s8 = load_unit<int16_t, 8>(a.ship, u);
        uint32_t in = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) in |= uint32_t(uint32_t(s8.v[j] - a.lo) < a.span) << j;
        if (!in) continue;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if ((in >> j) & 1) q15_row(a, u * 8 + j, s8.v[j], rev, emitted);
    }
    for (int64_t i = units * 8 + int64_t(blockIdx.x) * kT + threadIdx.x; i < a.rows; i += stride)
        q15_row(a, i, a.ship[i], rev, emitted);
    const uint64_t t = wave_sum_u64(emitted);
    if ((threadIdx.x & (kWave - 1)) == 0 && t) atomicAdd(emitted_out, (unsigned long long)t);
}

// Largest positive revenue (0 when none) into *mx.
__global__ __launch_bounds__(kT) void q15_max_k(const unsigned long long* __restrict__ rev,
                                                int64_t n, unsigned long long* __restrict__ mx) {
    int64_t m = 0;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const int64_t v = int64_t(rev[i]);
        m = v > m ? v : m;
    }
    for (int d = kWave / 2; d > 0; d >>= 1) {
        const int64_t o = __shfl_xor(m, d, kWave);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & (kWave - 1)) == 0 && m > 0) atomicMax(mx, (unsigned long long)m);
}

// Last supplier row of every key (atomic maximum over row numbers, -1 absent).
__global__ __launch_bounds__(kT) void last_row_k(const int32_t* __restrict__ key, int64_t rows,
                                                 int64_t n, int32_t* __restrict__ row_of) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = key[i];
        if (k >= 0 && k < n) atomicMax(&row_of[k], int32_t(i));
    }
}

__global__ __launch_bounds__(kT) void fill_i32_k(int32_t* __restrict__ p, int64_t n, int32_t v) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        p[i] = v;
}

// Keys whose revenue equals the maximum and that have a supplier row.
__global__ __launch_bounds__(kT) void q15_collect_k(const unsigned long long* __restrict__ rev,
                                                    int64_t n,
                                                    const unsigned long long* __restrict__ mx,
                                                    const int32_t* __restrict__ row_of,
                                                    int32_t* __restrict__ out,
                                                    unsigned long long* __restrict__ n_out) {
    const unsigned long long m = *mx;
    const int64_t stride_g = int64_t(gridDim.x) * kT;
    const int64_t rounds = (n + stride_g - 1) / stride_g;
    for (int64_t r = 0; r < rounds; ++r) {
        const int64_t i = r * stride_g + int64_t(blockIdx.x) * kT + threadIdx.x;
        const bool hit = m > 0 && i < n && rev[i] == m && row_of[i] >= 0;
        const unsigned long long slot = wave_append(hit, n_out);
        if (hit) {
            out[2 * slot] = int32_t(i);
            out[2 * slot + 1] = row_of[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Q20: quantity per (partkey, suppkey) of coloured parts and one nation's
// suppliers, against the partsupp rows' available quantity.
// ---------------------------------------------------------------------------
constexpr unsigned long long kEmpty = ~0ull;

__global__ __launch_bounds__(kT) void color_bits_k(const int32_t* __restrict__ pk,
                                                   const uint8_t* __restrict__ ok, int64_t rows,
                                                   int32_t max_key, uint32_t* __restrict__ bits) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < rows;
         i += int64_t(gridDim.x) * kT) {
        const int32_t k = pk[i];
        if (ok[i] && k >= 0 && k <= max_key) atomicOr(&bits[k >> 5], 1u << (k & 31));
    }
}

struct Q20K {
    int64_t rows;
    const int16_t* ship;
    const int32_t* pk;
    const int32_t* sk;
    const int16_t* nation;
    const int16_t* qty;
    int32_t lo;
    uint32_t span;
    int16_t target;
    const uint32_t* bits;
    int32_t max_partkey;
    // the group table (absent in the counting pass)
    unsigned long long* keys;
    unsigned long long* sums;
    uint64_t mask;
    unsigned long long max_fill;  // groups allowed before the table counts as full
    unsigned long long* fill;
    unsigned* overflow;
};

BOLAP_DEV bool colored(const uint32_t* bits, int32_t max_key, uint32_t k) {
    return k <= uint32_t(max_key) && ((bits[k >> 5] >> (k & 31)) & 1);
}

BOLAP_DEV uint64_t slot_of(unsigned long long key, uint64_t mask) {
    return mix64(key) & mask;
}

template <bool kCount>
BOLAP_DEV void q20_row(const Q20K& a, int64_t i, uint64_t& passing) {
    if (a.nation[i] != a.target) return;
    const uint32_t p = uint32_t(a.pk[i]);
    if (!colored(a.bits, a.max_partkey, p)) return;
    ++passing;
    if (kCount) return;
    const unsigned long long key = (uint64_t(p) << 32) | uint32_t(a.sk[i]);
    const unsigned long long q = (unsigned long long)(int64_t(a.qty[i]));
    uint64_t s = slot_of(key, a.mask);
    for (uint64_t probe = 0; probe <= a.mask; ++probe) {
        unsigned long long cur = a.keys[s];
        if (cur == kEmpty) {
            cur = atomicCAS(&a.keys[s], kEmpty, key);
            if (cur == kEmpty) {  // this lane opened the group
                if (atomicAdd(a.fill, 1ull) >= a.max_fill) atomicOr(a.overflow, 1u);
                cur = key;
            }
        }
        if (cur == key) {
            atomicAdd(&a.sums[s], q);
            return;
        }
        s = (s + 1) & a.mask;
    }
    atomicOr(a.overflow, 1u);  // no free slot
}

template <bool kCount>
__global__ __launch_bounds__(kT) void q20_li_k(Q20K a, unsigned long long* __restrict__ n_pass) {
    uint64_t passing = 0;
    const int64_t units = a.rows / 8;
    const int64_t stride = int64_t(gridDim.x) * kT;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < units; u += stride) {
        const VecT<int16_t, 8> s8 = load_unit<int16_t, 8>(a.ship, u);
        uint32_t in = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) in |= uint32_t(uint32_t(s8.v[j] - a.lo) < a.span) << j;
        if (!in) continue;
        const VecT<int16_t, 8> n8 = load_unit<int16_t, 8>(a.nation, u);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (((in >> j) & 1) && n8.v[j] == a.target) q20_row<kCount>(a, u * 8 + j, passing);
    }
    for (int64_t i = units * 8 + int64_t(blockIdx.x) * kT + threadIdx.x; i < a.rows; i += stride)
        if (uint32_t(a.ship[i] - a.lo) < a.span) q20_row<kCount>(a, i, passing);
    const uint64_t t = wave_sum_u64(passing);
    if ((threadIdx.x & (kWave - 1)) == 0 && t) atomicAdd(n_pass, (unsigned long long)t);
}

__global__ __launch_bounds__(kT) void fill_u64_k(unsigned long long* __restrict__ p, int64_t n,
                                                 unsigned long long v) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        p[i] = v;
}

struct Q20P {
    int64_t rows;
    const int32_t* pk;
    const int32_t* sk;
    const int32_t* avail;
    const uint32_t* bits;
    int32_t max_partkey;
    const int16_t* supp_nation;
    int64_t n_map;
    int16_t target;
    int32_t max_suppkey;
    const unsigned long long* keys;
    const unsigned long long* sums;
    uint64_t mask;
};

__global__ __launch_bounds__(kT) void q20_ps_k(Q20P a, uint8_t* __restrict__ qualified) {
    for (int64_t j = int64_t(blockIdx.x) * kT + threadIdx.x; j < a.rows;
         j += int64_t(gridDim.x) * kT) {
        const uint32_t p = uint32_t(a.pk[j]);
        if (!colored(a.bits, a.max_partkey, p)) continue;
        const uint32_t s = uint32_t(a.sk[j]);
        if (s >= uint64_t(a.n_map) || a.supp_nation[s] != a.target) continue;
        const unsigned long long key = (uint64_t(p) << 32) | s;
        uint64_t slot = slot_of(key, a.mask);
        for (uint64_t probe = 0; probe <= a.mask; ++probe) {
            const unsigned long long cur = a.keys[slot];
            if (cur == kEmpty) break;
            if (cur == key) {
                const int64_t sum = int64_t(a.sums[slot]);
                if (int64_t(a.avail[j]) * 2 * 100 > sum && a.max_suppkey >= 0 &&
                    s <= uint32_t(a.max_suppkey))
                    qualified[s] = 1;
                break;
            }
            slot = (slot + 1) & a.mask;
        }
    }
}

uint64_t pow2_at_least(uint64_t v) {
    uint64_t c = 1;
    while (c < v) c <<= 1;
    return c;
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" int bolap_q2(bolap_ctx* ctx, const bolap_q2_args* a, int32_t* pairs_out,
                        int64_t max_rows, int64_t* rows) {
    BOLAP_ARG_CHECK(ctx && a && rows && (pairs_out || max_rows == 0), "bolap_q2: null argument");
    BOLAP_ARG_CHECK(a->part_rows >= 0 && a->part_rows < (int64_t(1) << 31) && a->ps_rows >= 0 &&
                        a->ps_rows < (int64_t(1) << 32) && a->supp_rows >= 0,
                    "bolap_q2: rows out of range");
    BOLAP_ARG_CHECK(a->n_types >= 0 && a->n_types <= 256 && a->n_nation >= 0 &&
                        a->n_nation <= 65536 && (a->type_ok || a->n_types == 0) &&
                        (a->nation_ok || a->n_nation == 0),
                    "bolap_q2: %d types / %d nations", a->n_types, a->n_nation);
    BOLAP_ARG_CHECK(a->part_rows == 0 || (a->p_size && a->p_type), "bolap_q2: null part column");
    BOLAP_ARG_CHECK(a->ps_rows == 0 || (a->ps_partkey && a->ps_suppkey && a->ps_supplycost),
                    "bolap_q2: null partsupp column");
    BOLAP_ARG_CHECK(a->supp_rows == 0 || a->s_nationkey, "bolap_q2: null supplier column");
    *rows = 0;
    if (a->part_rows == 0 || a->ps_rows == 0 || a->supp_rows == 0) return BOLAP_OK;
    WsArena ws;
    const size_t o_n = ws.plan(8);
    const size_t o_tok = ws.plan(size_t(std::max(a->n_types, 1)));
    const size_t o_nok = ws.plan(size_t(std::max(a->n_nation, 1)));
    const size_t o_keys = ws.plan(size_t(a->ps_rows) * 8);
    const size_t o_vals = ws.plan(size_t(a->ps_rows) * 4);
    const size_t o_pairs = ws.plan(size_t(a->ps_rows) * 8);
    const size_t o_sort = ws.plan(radix_sort_scratch_bytes(ctx, a->ps_rows));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    BOLAP_TRY(zero_dev(ctx, ws.at<void>(o_n), 8));
    if (a->n_types)
        BOLAP_HIP_TRY(hipMemcpyAsync(ws.at<uint8_t>(o_tok), a->type_ok, size_t(a->n_types),
                                     hipMemcpyHostToDevice, ctx->stream));
    if (a->n_nation)
        BOLAP_HIP_TRY(hipMemcpyAsync(ws.at<uint8_t>(o_nok), a->nation_ok, size_t(a->n_nation),
                                     hipMemcpyHostToDevice, ctx->stream));
    auto* n_dev = ws.at<unsigned long long>(o_n);
    {
        TraceScope ts(ctx, "q2_part_partsupp");
        const Q2K k{a->part_rows, a->p_size, a->p_type, ws.at<uint8_t>(o_tok), a->n_types,
                    a->size, a->ps_rows, a->ps_partkey, a->ps_suppkey, a->ps_supplycost,
                    a->ps_rows % a->part_rows == 0 ? a->ps_rows / a->part_rows : 0, a->supp_rows,
                    a->s_nationkey, ws.at<uint8_t>(o_nok), a->n_nation};
        hipLaunchKernelGGL(q2_k, dim3(grid_of(ctx, a->part_rows)), dim3(kT), 0, ctx->stream, k,
                           ws.at<uint64_t>(o_keys), n_dev);
        BOLAP_TRY(launch_check("q2_k"));
    }
    unsigned long long n = 0;
    BOLAP_TRY(fetch(ctx, &n, n_dev, 8));
    *rows = int64_t(n);
    if (int64_t(n) > max_rows) {
        set_error("bolap_q2: %lld result rows, room for %lld", (long long)n, (long long)max_rows);
        return BOLAP_ERR_CAPACITY;
    }
    if (n == 0) return BOLAP_OK;
    {
        TraceScope ts(ctx, "q2_order");
        int bits = 1;
        while (bits < 63 && (int64_t(1) << bits) < a->ps_rows) ++bits;
        int passes = 0;
        BOLAP_TRY(radix_sort_pairs(ctx, ws.at<uint64_t>(o_keys), ws.at<uint32_t>(o_vals),
                                   int64_t(n), bits, ws.at<void>(o_sort), &passes));
        hipLaunchKernelGGL(q2_pairs_k, dim3(grid_of(ctx, int64_t(n))), dim3(kT), 0, ctx->stream,
                           ws.at<uint64_t>(o_keys), int64_t(n), a->ps_partkey, a->ps_suppkey,
                           ws.at<int32_t>(o_pairs));
        BOLAP_TRY(launch_check("q2_pairs_k"));
    }
    return fetch(ctx, pairs_out, ws.at<int32_t>(o_pairs), size_t(n) * 8);
}

extern "C" int bolap_q11(bolap_ctx* ctx, const bolap_q11_args* a, int32_t* partkey_out,
                         int64_t* value_out, int64_t max_rows, int64_t* rows, int64_t* total) {
    BOLAP_ARG_CHECK(ctx && a && rows && total && ((partkey_out && value_out) || max_rows == 0),
                    "bolap_q11: null argument");
    BOLAP_ARG_CHECK(a->supp_rows >= 0 && a->ps_rows >= 0 && a->n_supp_keys >= 0,
                    "bolap_q11: rows out of range");
    BOLAP_ARG_CHECK(a->supp_rows == 0 || (a->s_suppkey && a->s_nationkey),
                    "bolap_q11: null supplier column");
    BOLAP_ARG_CHECK(a->ps_rows == 0 || (a->ps_partkey && a->ps_suppkey && a->ps_value),
                    "bolap_q11: null partsupp column");
    *rows = 0;
    *total = 0;
    if (a->ps_rows == 0) return BOLAP_OK;
    const int64_t nk = a->n_supp_keys > 0 ? a->n_supp_keys : 1;
    WsArena ws;
    const size_t o_cnt = ws.plan(64);  // groups, total, kept, four range words
    const size_t o_mark = ws.plan(size_t(nk));
    const size_t o_zero_end = ws.off;
    // contiguous row ranges of whole tiles, one per workgroup
    const int64_t g0 = grid_of(ctx, a->ps_rows);
    const int64_t chunk = ((a->ps_rows + g0 - 1) / g0 + kT - 1) / kT * kT;
    const unsigned blocks = unsigned((a->ps_rows + chunk - 1) / chunk);
    const size_t o_blk = ws.plan(size_t(blocks) * 4);
    const size_t o_rng = ws.plan(size_t(blocks) * 32);
    const size_t o_gpk = ws.plan(size_t(a->ps_rows) * 4);
    const size_t o_gsum = ws.plan(size_t(a->ps_rows) * 8);
    const size_t o_opk = ws.plan(size_t(a->ps_rows) * 4);
    const size_t o_osum = ws.plan(size_t(a->ps_rows) * 8);
    const size_t o_keys = ws.plan(size_t(a->ps_rows) * 8);
    const size_t o_sort = ws.plan(radix_sort_scratch_bytes(ctx, a->ps_rows));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    BOLAP_TRY(zero_dev(ctx, ws.base, o_zero_end));
    auto* cnt = ws.at<unsigned long long>(o_cnt);
    {
        TraceScope ts(ctx, "q11_partsupp");
        if (a->supp_rows > 0)
            hipLaunchKernelGGL(mark_nation_k, dim3(grid_of(ctx, a->supp_rows)), dim3(kT), 0,
                               ctx->stream, a->s_suppkey, a->s_nationkey, a->supp_rows, a->nation,
                               nk, ws.at<uint8_t>(o_mark));
        const Q11K k{a->ps_rows, a->ps_partkey, a->ps_suppkey, a->ps_value, ws.at<uint8_t>(o_mark),
                     nk};
        hipLaunchKernelGGL(q11_runs_k, dim3(blocks), dim3(kT), 0, ctx->stream, k, chunk,
                           ws.at<int32_t>(o_gpk), ws.at<int64_t>(o_gsum), ws.at<uint32_t>(o_blk),
                           cnt, cnt + 1);
        hipLaunchKernelGGL(q11_filter_k, dim3(blocks), dim3(kT), 0, ctx->stream,
                           ws.at<int32_t>(o_gpk), ws.at<int64_t>(o_gsum), ws.at<uint32_t>(o_blk),
                           chunk, cnt + 1, a->fraction, ws.at<int32_t>(o_opk),
                           ws.at<int64_t>(o_osum), cnt + 2, ws.at<uint64_t>(o_rng));
        hipLaunchKernelGGL(q11_range_k, dim3(1), dim3(kT), 0, ctx->stream,
                           ws.at<uint64_t>(o_rng), int64_t(blocks),
                           reinterpret_cast<uint64_t*>(cnt + 3));
        BOLAP_TRY(launch_check("q11_range_k"));
    }
    unsigned long long c[7];
    BOLAP_TRY(fetch(ctx, c, cnt, sizeof(c)));
    *total = int64_t(c[1]);
    *rows = int64_t(c[2]);
    trace_count(ctx, "q11_groups", int64_t(c[0]));
    if (int64_t(c[2]) > max_rows) {
        set_error("bolap_q11: %lld result rows, room for %lld", (long long)c[2],
                  (long long)max_rows);
        return BOLAP_ERR_CAPACITY;
    }
    const int64_t n = int64_t(c[2]);
    if (n == 0) return BOLAP_OK;
    // the kept groups in the reference's order: value descending, then key.
    // On the device when (vmax - value, key - kmin) packs into 64 bits.
    const uint64_t vmin = ~c[3], vmax = c[4];
    const uint32_t kmin = uint32_t(~c[5]), kmax = uint32_t(c[6]);
    auto bit_len = [](uint64_t x) {
        int b = 0;
        while (b < 64 && (x >> b) != 0) ++b;
        return b;
    };
    const int kbits = bit_len(uint64_t(kmax - kmin)), vbits = bit_len(vmax - vmin);
    if (kbits + vbits <= 64) {
        TraceScope ts(ctx, "q11_order");
        auto* keys = ws.at<uint64_t>(o_keys);
        hipLaunchKernelGGL(q11_keys_k, dim3(grid_of(ctx, n)), dim3(kT), 0, ctx->stream,
                           ws.at<int32_t>(o_opk), ws.at<int64_t>(o_osum), n, vmax, kmin, kbits,
                           keys);
        int passes = 0;
        BOLAP_TRY(radix_sort_pairs(ctx, keys, ws.at<uint32_t>(o_opk), n, kbits + vbits,
                                   ws.at<void>(o_sort), &passes));
        hipLaunchKernelGGL(q11_values_k, dim3(grid_of(ctx, n)), dim3(kT), 0, ctx->stream, keys, n,
                           vmax, kbits, ws.at<int64_t>(o_osum));
        BOLAP_TRY(launch_check("q11_values_k"));
        BOLAP_HIP_TRY(hipMemcpyAsync(partkey_out, ws.at<int32_t>(o_opk), size_t(n) * 4,
                                     hipMemcpyDeviceToHost, ctx->stream));
        return fetch(ctx, value_out, ws.at<int64_t>(o_osum), size_t(n) * 8);
    }
    std::vector<int32_t> pk(static_cast<size_t>(n));
    std::vector<int64_t> v(static_cast<size_t>(n));
    BOLAP_HIP_TRY(hipMemcpyAsync(pk.data(), ws.at<int32_t>(o_opk), size_t(n) * 4,
                                 hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_TRY(fetch(ctx, v.data(), ws.at<int64_t>(o_osum), size_t(n) * 8));
    std::vector<int64_t> idx(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) idx[size_t(i)] = i;
    std::sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) {
        if (v[size_t(x)] != v[size_t(y)]) return v[size_t(x)] > v[size_t(y)];
        return pk[size_t(x)] < pk[size_t(y)];
    });
    for (int64_t i = 0; i < n; ++i) {
        partkey_out[i] = pk[size_t(idx[size_t(i)])];
        value_out[i] = v[size_t(idx[size_t(i)])];
    }
    return BOLAP_OK;
}

extern "C" int bolap_q15(bolap_ctx* ctx, const bolap_q15_args* a, int32_t* suppkey_out,
                         int32_t* row_out, int64_t* revenue_out, int64_t max_rows, int64_t* rows,
                         int64_t* emitted) {
    BOLAP_ARG_CHECK(ctx && a && rows && emitted &&
                        ((suppkey_out && row_out && revenue_out) || max_rows == 0),
                    "bolap_q15: null argument");
    BOLAP_ARG_CHECK(a->li_rows >= 0 && a->supp_rows >= 0 && a->n_supp_keys >= 1,
                    "bolap_q15: rows out of range (n_supp_keys = max suppkey + 1, at least 1)");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (a->l_shipdate && a->l_suppkey && a->l_disc_price),
                    "bolap_q15: null lineitem column");
    BOLAP_ARG_CHECK(a->li_rows == 0 || aligned(a->l_shipdate, 16),
                    "bolap_q15: l_shipdate must be 16-byte aligned");
    BOLAP_ARG_CHECK(a->supp_rows == 0 || a->s_suppkey, "bolap_q15: null supplier column");
    *rows = 0;
    *emitted = 0;
    const int64_t nk = a->n_supp_keys;
    WsArena ws;
    const size_t o_cnt = ws.plan(32);  // emitted, max, results
    const size_t o_rev = ws.plan(size_t(nk) * 8);
    const size_t o_zero_end = ws.off;
    const size_t o_row = ws.plan(size_t(nk) * 4);
    const size_t o_out = ws.plan(size_t(nk) * 8);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    BOLAP_TRY(zero_dev(ctx, ws.base, o_zero_end));
    auto* cnt = ws.at<unsigned long long>(o_cnt);
    auto* rev = ws.at<unsigned long long>(o_rev);
    {
        TraceScope ts(ctx, "q15_lineitem");
        const Q15K k{a->li_rows, a->l_shipdate, a->l_suppkey, a->l_disc_price, a->date_lo,
                     uint32_t(a->date_hi - a->date_lo), nk};
        if (a->li_rows > 0)
            hipLaunchKernelGGL(q15_scan_k, dim3(grid_of(ctx, (a->li_rows + 7) / 8, 8)), dim3(kT),
                               0, ctx->stream, k, rev, cnt);
        BOLAP_TRY(launch_check("q15_scan_k"));
    }
    {
        TraceScope ts(ctx, "q15_top_suppliers");
        hipLaunchKernelGGL(q15_max_k, dim3(grid_of(ctx, nk)), dim3(kT), 0, ctx->stream, rev, nk,
                           cnt + 1);
        hipLaunchKernelGGL(fill_i32_k, dim3(grid_of(ctx, nk)), dim3(kT), 0, ctx->stream,
                           ws.at<int32_t>(o_row), nk, -1);
        if (a->supp_rows > 0)
            hipLaunchKernelGGL(last_row_k, dim3(grid_of(ctx, a->supp_rows)), dim3(kT), 0,
                               ctx->stream, a->s_suppkey, a->supp_rows, nk, ws.at<int32_t>(o_row));
        hipLaunchKernelGGL(q15_collect_k, dim3(grid_of(ctx, nk)), dim3(kT), 0, ctx->stream, rev,
                           nk, cnt + 1, ws.at<int32_t>(o_row), ws.at<int32_t>(o_out), cnt + 2);
        BOLAP_TRY(launch_check("q15_collect_k"));
    }
    unsigned long long c[3];
    BOLAP_TRY(fetch(ctx, c, cnt, sizeof(c)));
    *emitted = int64_t(c[0]);
    // no row in the window, or no positive revenue: no result (:454-476)
    if (c[0] == 0 || c[1] == 0) return BOLAP_OK;
    const int64_t n = int64_t(c[2]);
    *rows = n;
    if (n > max_rows) {
        set_error("bolap_q15: %lld result rows, room for %lld", (long long)n, (long long)max_rows);
        return BOLAP_ERR_CAPACITY;
    }
    std::vector<int32_t> out(static_cast<size_t>(n) * 2);
    BOLAP_TRY(fetch(ctx, out.data(), ws.at<int32_t>(o_out), size_t(n) * 8));
    std::vector<int64_t> order(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) order[size_t(i)] = i;
    std::sort(order.begin(), order.end(),
              [&](int64_t x, int64_t y) { return out[size_t(2 * x)] < out[size_t(2 * y)]; });
    for (int64_t i = 0; i < n; ++i) {
        suppkey_out[i] = out[size_t(2 * order[size_t(i)])];
        row_out[i] = out[size_t(2 * order[size_t(i)] + 1)];
        revenue_out[i] = int64_t(c[1]);
    }
    return BOLAP_OK;
}

extern "C" int bolap_q20(bolap_ctx* ctx, const bolap_q20_args* a, uint8_t* qualified,
                         bolap_q20_stats* st) {
    BOLAP_ARG_CHECK(ctx && a && st, "bolap_q20: null argument");
    BOLAP_ARG_CHECK(a->part_rows >= 0 && a->li_rows >= 0 && a->ps_rows >= 0 &&
                        a->n_supp_map >= 0 && a->capacity_hint >= 0,
                    "bolap_q20: rows out of range");
    BOLAP_ARG_CHECK(a->max_suppkey < 0 || qualified, "bolap_q20: null qualified array");
    BOLAP_ARG_CHECK(a->part_rows == 0 || (a->p_partkey && a->p_ok_row),
                    "bolap_q20: null part column");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (a->l_shipdate && a->l_partkey && a->l_suppkey &&
                                        a->l_supp_nation && a->l_quantity),
                    "bolap_q20: null lineitem column");
    BOLAP_ARG_CHECK(a->li_rows == 0 || (aligned(a->l_shipdate, 16) && aligned(a->l_supp_nation, 16)),
                    "bolap_q20: l_shipdate / l_supp_nation must be 16-byte aligned");
    BOLAP_ARG_CHECK(a->ps_rows == 0 || (a->ps_partkey && a->ps_suppkey && a->ps_availqty),
                    "bolap_q20: null partsupp column");
    BOLAP_ARG_CHECK(a->n_supp_map == 0 || a->supp_nation_by_key,
                    "bolap_q20: null supplier nation map");
    *st = bolap_q20_stats{};
    const int64_t n_q = a->max_suppkey >= 0 ? int64_t(a->max_suppkey) + 1 : 0;
    if (n_q) std::memset(qualified, 0, size_t(n_q));
    if (a->max_partkey < 0 || a->li_rows == 0) return BOLAP_OK;
    const int64_t words = int64_t(a->max_partkey) / 32 + 1;
    const Q20K base{a->li_rows, a->l_shipdate, a->l_partkey, a->l_suppkey, a->l_supp_nation,
                    a->l_quantity, a->date_lo, uint32_t(a->date_hi - a->date_lo),
                    int16_t(a->nation), nullptr, a->max_partkey, nullptr, nullptr, 0, 0, nullptr,
                    nullptr};
    const unsigned li_grid = grid_of(ctx, (a->li_rows + 7) / 8, 8);
    // group capacity: the estimate, or the passing rows counted first
    uint64_t expected = uint64_t(a->capacity_hint);
    bool exact = a->capacity_hint == 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        WsArena ws;
        const size_t o_cnt = ws.plan(32);  // passing rows, filled groups, overflow
        const size_t o_bits = ws.plan(size_t(words) * 4);
        const size_t o_q = ws.plan(size_t(std::max<int64_t>(n_q, 1)));
        const size_t o_zero_end = ws.off;
        if (exact) {  // count the rows that reach the table
            BOLAP_TRY(ws_reserve(ctx, ws.off));
            ws.base = static_cast<char*>(ctx->ws);
            BOLAP_TRY(zero_dev(ctx, ws.base, o_zero_end));
            TraceScope ts(ctx, "q20_count");
            hipLaunchKernelGGL(color_bits_k, dim3(grid_of(ctx, a->part_rows)), dim3(kT), 0,
                               ctx->stream, a->p_partkey, a->p_ok_row, a->part_rows,
                               a->max_partkey, ws.at<uint32_t>(o_bits));
            Q20K k = base;
            k.bits = ws.at<uint32_t>(o_bits);
            hipLaunchKernelGGL(q20_li_k<true>, dim3(li_grid), dim3(kT), 0, ctx->stream, k,
                               ws.at<unsigned long long>(o_cnt));
            BOLAP_TRY(launch_check("q20_li_k<count>"));
            unsigned long long passing = 0;
            BOLAP_TRY(fetch(ctx, &passing, ws.at<void>(o_cnt), 8));
            expected = passing;
        }
        // the reference's reserve: a power of two >= 2 x expected, at least 8
        // (query_q20.cpp:145-152); full beyond a 0.7 load (:158)
        const uint64_t cap = pow2_at_least(std::max<uint64_t>(expected * 2, 8));
        const size_t o_keys = ws.plan(size_t(cap) * 8);
        const size_t o_sums = ws.plan(size_t(cap) * 8);
        BOLAP_TRY(ws_reserve(ctx, ws.off, exact ? o_zero_end : 0));
        ws.base = static_cast<char*>(ctx->ws);
        auto* cnt = ws.at<unsigned long long>(o_cnt);
        if (!exact) BOLAP_TRY(zero_dev(ctx, ws.base, o_zero_end));
        else BOLAP_TRY(zero_dev(ctx, cnt, 32));
        {
            TraceScope ts(ctx, "q20_lineitem");
            if (!exact)
                hipLaunchKernelGGL(color_bits_k, dim3(grid_of(ctx, a->part_rows)), dim3(kT), 0,
                                   ctx->stream, a->p_partkey, a->p_ok_row, a->part_rows,
                                   a->max_partkey, ws.at<uint32_t>(o_bits));
            hipLaunchKernelGGL(fill_u64_k, dim3(grid_of(ctx, int64_t(cap))), dim3(kT), 0,
                               ctx->stream, ws.at<unsigned long long>(o_keys), int64_t(cap),
                               kEmpty);
            hipLaunchKernelGGL(fill_u64_k, dim3(grid_of(ctx, int64_t(cap))), dim3(kT), 0,
                               ctx->stream, ws.at<unsigned long long>(o_sums), int64_t(cap), 0ull);
            Q20K k = base;
            k.bits = ws.at<uint32_t>(o_bits);
            k.keys = ws.at<unsigned long long>(o_keys);
            k.sums = ws.at<unsigned long long>(o_sums);
            k.mask = cap - 1;
            k.max_fill = (unsigned long long)(cap * 7 / 10);
            k.fill = cnt + 1;
            k.overflow = reinterpret_cast<unsigned*>(cnt + 2);
            hipLaunchKernelGGL(q20_li_k<false>, dim3(li_grid), dim3(kT), 0, ctx->stream, k, cnt);
            BOLAP_TRY(launch_check("q20_li_k"));
        }
        unsigned long long c[3];
        BOLAP_TRY(fetch(ctx, c, cnt, sizeof(c)));
        st->capacity = int64_t(cap);
        st->pairs = int64_t(c[1]);
        if (c[2] != 0) {  // the estimate was short: size exactly and run again
            if (exact) {
                set_error("bolap_q20: exact-sized table overflowed");
                return BOLAP_ERR_HIP;
            }
            st->fallbacks += 1;
            exact = true;
            continue;
        }
        {
            TraceScope ts(ctx, "q20_partsupp");
            if (a->ps_rows > 0) {
                const Q20P p{a->ps_rows, a->ps_partkey, a->ps_suppkey, a->ps_availqty,
                             ws.at<uint32_t>(o_bits), a->max_partkey, a->supp_nation_by_key,
                             a->n_supp_map, int16_t(a->nation), a->max_suppkey,
                             ws.at<unsigned long long>(o_keys), ws.at<unsigned long long>(o_sums),
                             cap - 1};
                hipLaunchKernelGGL(q20_ps_k, dim3(grid_of(ctx, a->ps_rows)), dim3(kT), 0,
                                   ctx->stream, p, ws.at<uint8_t>(o_q));
                BOLAP_TRY(launch_check("q20_ps_k"));
            }
        }
        BOLAP_TRY(fetch(ctx, qualified, ws.at<uint8_t>(o_q), size_t(n_q)));
        for (int64_t i = 0; i < n_q; ++i) st->qualified += qualified[i];
        return BOLAP_OK;
    }
    set_error("bolap_q20: no attempt completed");
    return BOLAP_ERR_HIP;
}
