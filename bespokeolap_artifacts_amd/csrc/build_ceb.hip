// build_ceb.hip — the CEB builder's derived data: a string pool that maps
// strings to int32 ids, and per-table sort orders applied on the device.
//
// Reference: bespoke_ceb/builder_impl.hpp:14-89 (StringPool: intern, finalize
// with lower-cased copies) and builder_impl.cpp:364-466 (every string column
// becomes ids into one global pool, nulls stay -1; tables are sorted by
// (movie_id, role_id), (movie_id, info_type_id), movie_id, keyword_id,
// info_type_id or id, :432-452, with std::sort on the key columns,
// :265-340).
//
// The pool here is host code (the reference's is too); its ids follow first
// appearance over the intern calls, which is the reference's order when one
// thread interns the tables in turn (its concurrent interning leaves the ids
// open).  The sort order is an LSD radix sort on the device of the packed
// (primary, secondary) key with the row index as payload: ascending by the
// signed keys, ties in input order (std::sort leaves them open).  The key is
// packed as offsets from the columns' minima, (p - pmin) << sbits | (s -
// smin), so the sort runs only as many 8-bit passes as the key spans need
// (cast_info's (movie_id, role_id): 26 bits, four passes instead of eight).
#include <algorithm>
#include <cctype>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "radix.hpp"

struct bolap_string_pool {
    std::unordered_map<std::string, int32_t> index;
    std::vector<const std::string*> values;  // by id
};

namespace bolap {
namespace {

constexpr int kT = 256;

unsigned grid_of(bolap_ctx* ctx, int64_t n) {
    int64_t b = (n + kT - 1) / kT;
    const int64_t cap = int64_t(ctx->num_cus) * 8;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

// Signed int32 -> order-preserving unsigned.
BOLAP_DEV uint32_t bias(int32_t v) { return uint32_t(v) ^ 0x80000000u; }

// Per-workgroup minima and maxima of both columns folded with one atomic
// per workgroup and value (a grid of at most a few hundred workgroups).
__global__ __launch_bounds__(kT) void so_range_k(const int32_t* __restrict__ p,
                                                 const int32_t* __restrict__ s, int64_t n,
                                                 bool vec, unsigned* __restrict__ mm) {
    __shared__ unsigned s_v[4][kT / kWave];
    unsigned v[4] = {~0u, 0u, ~0u, 0u};  // biased: min p, max p, min s, max s
    auto fold_p = [&](int32_t x) {
        const unsigned a = bias(x);
        v[0] = a < v[0] ? a : v[0];
        v[1] = a > v[1] ? a : v[1];
    };
    auto fold_s = [&](int32_t x) {
        const unsigned b = bias(x);
        v[2] = b < v[2] ? b : v[2];
        v[3] = b > v[3] ? b : v[3];
    };
    // four rows per lane per 16-byte load (vec: both columns 16-byte
    // aligned), two units in flight; the tail (or unaligned columns) row by row
    const int64_t units = vec ? n / 4 : 0;
    const int64_t stride = int64_t(gridDim.x) * kT;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < units; u += 2 * stride) {
        const int64_t u2 = u + stride < units ? u + stride : u;
        const VecT<int32_t, 4> a1 = load_unit<int32_t, 4>(p, u), a2 = load_unit<int32_t, 4>(p, u2);
        VecT<int32_t, 4> b1{}, b2{};
        if (s) {
            b1 = load_unit<int32_t, 4>(s, u);
            b2 = load_unit<int32_t, 4>(s, u2);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            fold_p(a1.v[j]);
            fold_p(a2.v[j]);
            if (s) {
                fold_s(b1.v[j]);
                fold_s(b2.v[j]);
            }
        }
    }
    for (int64_t i = units * 4 + int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += stride) {
        fold_p(p[i]);
        if (s) fold_s(s[i]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        for (int off = kWave / 2; off > 0; off >>= 1) {
            const unsigned o = unsigned(__shfl_xor(int(v[k]), off, kWave));
            v[k] = (k & 1) ? (o > v[k] ? o : v[k]) : (o < v[k] ? o : v[k]);
        }
        if ((threadIdx.x & (kWave - 1)) == 0) s_v[k][threadIdx.x / kWave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        unsigned r = s_v[k][0];
        for (int w = 1; w < kT / kWave; ++w)
            r = (k & 1) ? (s_v[k][w] > r ? s_v[k][w] : r) : (s_v[k][w] < r ? s_v[k][w] : r);
        if (k & 1) atomicMax(mm + k, r);
        else atomicMin(mm + k, r);
    }
}

__global__ void so_range_init_k(unsigned* mm) {
    if (threadIdx.x < 4) mm[threadIdx.x] = (threadIdx.x & 1) ? 0u : ~0u;
}

__global__ __launch_bounds__(kT) void so_keys_k(const int32_t* __restrict__ p,
                                                const int32_t* __restrict__ s, int64_t n,
                                                unsigned pmin, unsigned smin, int sbits,
                                                uint64_t* __restrict__ keys,
                                                uint32_t* __restrict__ rows) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const uint64_t hi = uint64_t(bias(p[i]) - pmin);
        keys[i] = s ? (hi << sbits) | uint64_t(bias(s[i]) - smin) : hi;
        rows[i] = uint32_t(i);
    }
}

template <typename T>
__global__ __launch_bounds__(kT) void gather_k(const T* __restrict__ src,
                                               const uint32_t* __restrict__ order, int64_t n,
                                               int64_t src_rows, T* __restrict__ dst,
                                               unsigned* __restrict__ bad) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const uint32_t r = order[i];
        if (int64_t(r) < src_rows) {
            dst[i] = src[r];
        } else {
            atomicOr(bad, 1u);
        }
    }
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" {

int bolap_string_pool_create(bolap_string_pool** out) {
    BOLAP_ARG_CHECK(out != nullptr, "bolap_string_pool_create: null argument");
    *out = new bolap_string_pool();
    return BOLAP_OK;
}

void bolap_string_pool_destroy(bolap_string_pool* pool) { delete pool; }

int bolap_string_pool_intern(bolap_string_pool* pool, const char* chars, const int64_t* offsets,
                             int64_t n, const uint8_t* valid, int32_t* ids_out) {
    BOLAP_ARG_CHECK(pool && offsets && ids_out && n >= 0 && (chars || n == 0),
                    "bolap_string_pool_intern: bad arguments");
    for (int64_t i = 0; i < n; ++i) {
        if (valid && !valid[i]) {
            ids_out[i] = -1;  // a null string keeps id -1 (builder_impl.cpp:42-45, 60-66)
            continue;
        }
        BOLAP_ARG_CHECK(offsets[i] >= 0 && offsets[i + 1] >= offsets[i],
                        "bolap_string_pool_intern: offsets not ascending at %lld", (long long)i);
        std::string s(chars + offsets[i], size_t(offsets[i + 1] - offsets[i]));
        auto it = pool->index.find(s);
        if (it == pool->index.end()) {
            BOLAP_ARG_CHECK(pool->values.size() < size_t(INT32_MAX),
                            "bolap_string_pool_intern: more than 2^31 - 1 strings");
            const int32_t id = int32_t(pool->values.size());
            it = pool->index.emplace(std::move(s), id).first;
            pool->values.push_back(&it->first);
        }
        ids_out[i] = it->second;
    }
    return BOLAP_OK;
}

int bolap_string_pool_size(const bolap_string_pool* pool, int64_t* n_strings, int64_t* n_chars) {
    BOLAP_ARG_CHECK(pool && n_strings && n_chars, "bolap_string_pool_size: null argument");
    int64_t c = 0;
    for (const std::string* s : pool->values) c += int64_t(s->size());
    *n_strings = int64_t(pool->values.size());
    *n_chars = c;
    return BOLAP_OK;
}

int bolap_string_pool_export(const bolap_string_pool* pool, int32_t lower, char* chars,
                             int64_t* offsets) {
    BOLAP_ARG_CHECK(pool && offsets && (chars || pool->values.empty()),
                    "bolap_string_pool_export: null argument");
    int64_t at = 0;
    offsets[0] = 0;
    for (size_t i = 0; i < pool->values.size(); ++i) {
        const std::string& s = *pool->values[i];
        for (size_t k = 0; k < s.size(); ++k)
            chars[at + int64_t(k)] =
                lower ? char(std::tolower(static_cast<unsigned char>(s[k]))) : s[k];
        at += int64_t(s.size());
        offsets[i + 1] = at;
    }
    return BOLAP_OK;
}

int bolap_sort_order_i32(bolap_ctx* ctx, const int32_t* primary, const int32_t* secondary,
                         int64_t rows, uint32_t* order_out) {
    BOLAP_ARG_CHECK(ctx && (primary || rows == 0) && (order_out || rows == 0) && rows >= 0,
                    "bolap_sort_order_i32: bad arguments");
    BOLAP_ARG_CHECK(rows <= int64_t(UINT32_MAX), "bolap_sort_order_i32: %lld rows",
                    (long long)rows);
    if (rows == 0) return BOLAP_OK;
    TraceScope ts(ctx, "sort_order");
    WsArena ws;
    const size_t o_keys = ws.plan(size_t(rows) * 8);
    const size_t o_scr = ws.plan(radix_sort_scratch_bytes(ctx, rows));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    uint64_t* keys = ws.at<uint64_t>(o_keys);
    unsigned* mm = static_cast<unsigned*>(ctx->result_dev);
    hipLaunchKernelGGL(so_range_init_k, dim3(1), dim3(kWave), 0, ctx->stream, mm);
    hipLaunchKernelGGL(so_range_k, dim3(std::min<unsigned>(grid_of(ctx, rows / 4 + 1),
                                                            unsigned(ctx->num_cus) * 2)),
                       dim3(kT), 0, ctx->stream, primary, secondary, rows,
                       aligned(primary, 16) && (!secondary || aligned(secondary, 16)), mm);
    BOLAP_TRY(launch_check("so_range_k"));
    unsigned h[4];
    BOLAP_HIP_TRY(hipMemcpyAsync(h, mm, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    auto bits_of = [](uint64_t span) {  // bits to hold 0 .. span
        int b = 0;
        while (b < 64 && (span >> b) != 0) ++b;
        return b;
    };
    const int pbits = bits_of(uint64_t(h[1] - h[0]));
    const int sbits = secondary ? bits_of(uint64_t(h[3] - h[2])) : 0;
    hipLaunchKernelGGL(so_keys_k, dim3(grid_of(ctx, rows)), dim3(kT), 0, ctx->stream, primary,
                       secondary, rows, h[0], h[2], sbits, keys, order_out);
    BOLAP_TRY(launch_check("so_keys_k"));
    int passes = 0;
    BOLAP_TRY(radix_sort_pairs(ctx, keys, order_out, rows, std::max(1, pbits + sbits),
                               ws.at<void>(o_scr), &passes));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BOLAP_OK;
}

int bolap_gather_rows(bolap_ctx* ctx, const void* src, int64_t src_rows, int32_t elem_bytes,
                      const uint32_t* order, int64_t rows, void* dst) {
    BOLAP_ARG_CHECK(ctx && rows >= 0 && src_rows >= 0 && (rows == 0 || (src && order && dst)),
                    "bolap_gather_rows: bad arguments");
    BOLAP_ARG_CHECK(elem_bytes == 1 || elem_bytes == 2 || elem_bytes == 4 || elem_bytes == 8,
                    "bolap_gather_rows: element size %d", elem_bytes);
    if (rows == 0) return BOLAP_OK;
    unsigned* bad = static_cast<unsigned*>(ctx->result_dev);
    BOLAP_HIP_TRY(hipMemsetAsync(bad, 0, 4, ctx->stream));
    const unsigned g = grid_of(ctx, rows);
    switch (elem_bytes) {
        case 1:
            hipLaunchKernelGGL(gather_k<uint8_t>, dim3(g), dim3(kT), 0, ctx->stream,
                               static_cast<const uint8_t*>(src), order, rows, src_rows,
                               static_cast<uint8_t*>(dst), bad);
            break;
        case 2:
            hipLaunchKernelGGL(gather_k<uint16_t>, dim3(g), dim3(kT), 0, ctx->stream,
                               static_cast<const uint16_t*>(src), order, rows, src_rows,
                               static_cast<uint16_t*>(dst), bad);
            break;
        case 4:
            hipLaunchKernelGGL(gather_k<uint32_t>, dim3(g), dim3(kT), 0, ctx->stream,
                               static_cast<const uint32_t*>(src), order, rows, src_rows,
                               static_cast<uint32_t*>(dst), bad);
            break;
        default:
            hipLaunchKernelGGL(gather_k<uint64_t>, dim3(g), dim3(kT), 0, ctx->stream,
                               static_cast<const uint64_t*>(src), order, rows, src_rows,
                               static_cast<uint64_t*>(dst), bad);
    }
    BOLAP_TRY(launch_check("gather_k"));
    unsigned h = 0;
    BOLAP_HIP_TRY(hipMemcpyAsync(&h, bad, 4, hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    BOLAP_ARG_CHECK(h == 0, "bolap_gather_rows: an order entry is >= src_rows %lld",
                    (long long)src_rows);
    return BOLAP_OK;
}

}  // extern "C"
