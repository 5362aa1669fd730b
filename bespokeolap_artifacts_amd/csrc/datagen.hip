// datagen.hip — seeded synthetic tables generated in HBM.
//
// Every value is a pure function of (seed, table/column stream, row), so the
// tables are identical however the rows are split across workgroups.  Value
// distributions follow the benchmark definitions: ship day U{0..2525},
// quantity U{1..50}, price = qty * U{90000..200000} / 100 cents, discount
// U{0..10}, tax U{0..8}; TPC-H style return flag / line status.
#include <numeric>

#include "prims.hpp"

namespace bolap {
namespace {

constexpr int kGenThreads = 256;

BOLAP_DEV uint32_t stream_id(uint32_t table, uint32_t col) { return (table << 8) | col; }

inline unsigned gen_blocks(int64_t rows) {
    int64_t b = (rows + kGenThreads - 1) / kGenThreads;
    if (b > 65536) b = 65536;
    return unsigned(b < 1 ? 1 : b);
}

__global__ __launch_bounds__(kGenThreads) void gen_lineitem_k(
    uint64_t seed, int64_t rows, int32_t* ship, int64_t* qty, int64_t* price, int64_t* disc,
    int64_t* tax, uint8_t* flag_a, uint8_t* flag_b) {
    for (int64_t r = int64_t(blockIdx.x) * kGenThreads + threadIdx.x; r < rows;
         r += int64_t(gridDim.x) * kGenThreads) {
        const uint64_t u = uint64_t(r);
        const int32_t s = int32_t(gen_uniform(gen_hash(seed, stream_id(1, 0), u), 0, 2525));
        const int64_t q = gen_uniform(gen_hash(seed, stream_id(1, 1), u), 1, 50);
        const int64_t unit = gen_uniform(gen_hash(seed, stream_id(1, 2), u), 90000, 200000);
        const int32_t receipt =
            s + int32_t(gen_uniform(gen_hash(seed, stream_id(1, 5), u), 1, 30));
        ship[r] = s;
        qty[r] = q;
        price[r] = q * unit / 100;
        disc[r] = gen_uniform(gen_hash(seed, stream_id(1, 3), u), 0, 10);
        tax[r] = gen_uniform(gen_hash(seed, stream_id(1, 4), u), 0, 8);
        flag_b[r] = uint8_t(s > 1263 ? 1 : 0);
        flag_a[r] = uint8_t(
            receipt <= 1263 ? (gen_uniform(gen_hash(seed, stream_id(1, 6), u), 0, 1) ? 2 : 0)
                            : 1);
    }
}

__global__ __launch_bounds__(kGenThreads) void gen_dim_k(uint64_t seed, int64_t rows,
                                                          uint64_t mul, uint64_t off,
                                                          int32_t* key, uint8_t* cat,
                                                          uint8_t* cont) {
    for (int64_t r = int64_t(blockIdx.x) * kGenThreads + threadIdx.x; r < rows;
         r += int64_t(gridDim.x) * kGenThreads) {
        const uint64_t u = uint64_t(r);
        key[r] = int32_t((u * mul + off) % uint64_t(rows)) + 1;
        cat[r] = uint8_t(gen_uniform(gen_hash(seed, stream_id(2, 1), u), 0, 149));
        cont[r] = uint8_t(gen_uniform(gen_hash(seed, stream_id(2, 2), u), 0, 39));
    }
}

__global__ __launch_bounds__(kGenThreads) void gen_fact_k(uint64_t seed, int64_t rows,
                                                           int64_t dim_rows, int32_t* fk,
                                                           int64_t* price, int64_t* disc) {
    for (int64_t r = int64_t(blockIdx.x) * kGenThreads + threadIdx.x; r < rows;
         r += int64_t(gridDim.x) * kGenThreads) {
        const uint64_t u = uint64_t(r);
        fk[r] = int32_t(gen_uniform(gen_hash(seed, stream_id(3, 0), u), 1, dim_rows));
        const int64_t q = gen_uniform(gen_hash(seed, stream_id(3, 1), u), 1, 50);
        const int64_t unit = gen_uniform(gen_hash(seed, stream_id(3, 2), u), 90000, 200000);
        price[r] = q * unit / 100;
        disc[r] = gen_uniform(gen_hash(seed, stream_id(3, 3), u), 0, 10);
    }
}

__global__ __launch_bounds__(kGenThreads) void gen_customers_k(uint64_t seed, int64_t rows,
                                                                int32_t* key, uint8_t* seg) {
    for (int64_t r = int64_t(blockIdx.x) * kGenThreads + threadIdx.x; r < rows;
         r += int64_t(gridDim.x) * kGenThreads) {
        key[r] = int32_t(r + 1);
        seg[r] = uint8_t(gen_uniform(gen_hash(seed, stream_id(4, 1), uint64_t(r)), 0, 4));
    }
}

__global__ __launch_bounds__(kGenThreads) void gen_orders_k(uint64_t seed, int64_t rows,
                                                             int64_t cust_rows, int32_t* key,
                                                             int32_t* day, int32_t* cust,
                                                             int64_t* counts) {
    for (int64_t r = int64_t(blockIdx.x) * kGenThreads + threadIdx.x; r <= rows;
         r += int64_t(gridDim.x) * kGenThreads) {
        if (r == rows) {
            counts[r] = 0;
            continue;
        }
        const uint64_t u = uint64_t(r);
        key[r] = int32_t(r + 1);
        day[r] = int32_t(gen_uniform(gen_hash(seed, stream_id(5, 1), u), 0, 2405));
        cust[r] = int32_t(gen_uniform(gen_hash(seed, stream_id(5, 2), u), 1, cust_rows));
        counts[r] = gen_uniform(gen_hash(seed, stream_id(5, 3), u), 1, 7);
    }
}

__global__ __launch_bounds__(kGenThreads) void gen_order_facts_k(
    uint64_t seed, int64_t order_rows, const int64_t* __restrict__ offsets, int64_t fact_rows,
    int32_t* fk, int64_t* price, int64_t* disc) {
    for (int64_t o = int64_t(blockIdx.x) * kGenThreads + threadIdx.x; o < order_rows;
         o += int64_t(gridDim.x) * kGenThreads) {
        const int64_t b = offsets[o];
        int64_t e = offsets[o + 1];
        if (e > fact_rows) e = fact_rows;
        for (int64_t j = b; j < e; ++j) {
            const uint64_t u = uint64_t(j);
            fk[j] = int32_t(o + 1);
            const int64_t q = gen_uniform(gen_hash(seed, stream_id(6, 1), u), 1, 50);
            const int64_t unit =
                gen_uniform(gen_hash(seed, stream_id(6, 2), u), 90000, 200000);
            price[j] = q * unit / 100;
            disc[j] = gen_uniform(gen_hash(seed, stream_id(6, 3), u), 0, 10);
        }
    }
}


__global__ __launch_bounds__(kGenThreads) void gen_iota_k(int64_t rows, int32_t first,
                                                           int32_t* out) {
    for (int64_t r = int64_t(blockIdx.x) * kGenThreads + threadIdx.x; r < rows;
         r += int64_t(gridDim.x) * kGenThreads)
        out[r] = int32_t(int64_t(first) + r);
}

template <typename T>
__global__ __launch_bounds__(kGenThreads) void gen_uniform_k(uint64_t seed, uint32_t stream,
                                                              int64_t rows, int32_t lo,
                                                              int32_t hi, T* out) {
    for (int64_t r = int64_t(blockIdx.x) * kGenThreads + threadIdx.x; r < rows;
         r += int64_t(gridDim.x) * kGenThreads)
        out[r] = T(gen_uniform(gen_hash(seed, stream, uint64_t(r)), lo, hi));
}

// Zipf-distributed keys: the top 32 hash bits pick a point u of [0, 2^32);
// its top 16 bits index the inverse-CDF table and the low 16 bits
// interpolate linearly between neighbouring entries (integer arithmetic, so
// the host restatement is bit-identical).
__global__ __launch_bounds__(kGenThreads) void gen_zipf_k(uint64_t seed, uint32_t stream,
                                                           int64_t rows,
                                                           const int32_t* __restrict__ q,
                                                           int64_t n_keys, uint64_t mul,
                                                           int32_t* out) {
    for (int64_t r = int64_t(blockIdx.x) * kGenThreads + threadIdx.x; r < rows;
         r += int64_t(gridDim.x) * kGenThreads) {
        const uint32_t u = uint32_t(gen_hash(seed, stream, uint64_t(r)) >> 32);
        const uint32_t i = u >> 16, f = u & 0xFFFFu;
        const int64_t a = q[i], b = q[i + 1];
        const int64_t rank = a + (((b - a) * int64_t(f)) >> 16);
        out[r] = int32_t(uint64_t(rank - 1) * mul % uint64_t(n_keys)) + 1;
    }
}

__global__ __launch_bounds__(kGenThreads) void gen_price_disc_k(uint64_t seed, uint32_t stream,
                                                                 int64_t rows, int64_t* price,
                                                                 int64_t* disc) {
    for (int64_t r = int64_t(blockIdx.x) * kGenThreads + threadIdx.x; r < rows;
         r += int64_t(gridDim.x) * kGenThreads) {
        const uint64_t u = uint64_t(r);
        const int64_t q = gen_uniform(gen_hash(seed, stream, u), 1, 50);
        const int64_t unit = gen_uniform(gen_hash(seed, stream + 1, u), 90000, 200000);
        price[r] = q * unit / 100;
        disc[r] = gen_uniform(gen_hash(seed, stream + 2, u), 0, 10);
    }
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" {

int bolap_gen_lineitem(bolap_ctx* ctx, uint64_t seed, int64_t rows,
                       const bolap_gen_lineitem_cols* c) {
    BOLAP_ARG_CHECK(ctx && c && rows >= 0, "bolap_gen_lineitem: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(c->ship && c->qty && c->price && c->disc && c->tax && c->flag_a &&
                        c->flag_b,
                    "bolap_gen_lineitem: null column");
    hipLaunchKernelGGL(gen_lineitem_k, dim3(gen_blocks(rows)), dim3(kGenThreads), 0,
                       ctx->stream, seed, rows, c->ship, c->qty, c->price, c->disc, c->tax,
                       c->flag_a, c->flag_b);
    return launch_check("gen_lineitem");
}

int bolap_gen_dim(bolap_ctx* ctx, uint64_t seed, int64_t rows, int32_t* key, uint8_t* cat,
                  uint8_t* cont) {
    BOLAP_ARG_CHECK(ctx && rows >= 0 && rows < (int64_t(1) << 31),
                    "bolap_gen_dim: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(key && cat && cont, "bolap_gen_dim: null column");
    const uint64_t n = uint64_t(rows);
    uint64_t a = 1000003ULL;
    while (std::gcd(a, n) != 1) a += 2;
    hipLaunchKernelGGL(gen_dim_k, dim3(gen_blocks(rows)), dim3(kGenThreads), 0, ctx->stream,
                       seed, rows, a % n, seed % n, key, cat, cont);
    return launch_check("gen_dim");
}

int bolap_gen_fact(bolap_ctx* ctx, uint64_t seed, int64_t rows, int64_t dim_rows, int32_t* fk,
                   int64_t* price, int64_t* disc) {
    BOLAP_ARG_CHECK(ctx && rows >= 0 && dim_rows >= 1, "bolap_gen_fact: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(fk && price && disc, "bolap_gen_fact: null column");
    hipLaunchKernelGGL(gen_fact_k, dim3(gen_blocks(rows)), dim3(kGenThreads), 0, ctx->stream,
                       seed, rows, dim_rows, fk, price, disc);
    return launch_check("gen_fact");
}

int bolap_gen_customers(bolap_ctx* ctx, uint64_t seed, int64_t rows, int32_t* key,
                        uint8_t* segment) {
    BOLAP_ARG_CHECK(ctx && rows >= 0, "bolap_gen_customers: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(key && segment, "bolap_gen_customers: null column");
    hipLaunchKernelGGL(gen_customers_k, dim3(gen_blocks(rows)), dim3(kGenThreads), 0,
                       ctx->stream, seed, rows, key, segment);
    return launch_check("gen_customers");
}

int bolap_gen_orders(bolap_ctx* ctx, uint64_t seed, int64_t rows, int64_t cust_rows,
                     int32_t* key, int32_t* day, int32_t* cust, int64_t* fact_offsets) {
    BOLAP_ARG_CHECK(ctx && rows >= 0 && cust_rows >= 1 && fact_offsets,
                    "bolap_gen_orders: bad arguments");
    BOLAP_ARG_CHECK(rows == 0 || (key && day && cust), "bolap_gen_orders: null column");
    WsArena plan;
    const size_t o_tmp = plan.plan(scan_scratch_bytes(rows + 1));
    BOLAP_TRY(ws_reserve(ctx, plan.off));
    plan.base = static_cast<char*>(ctx->ws);
    hipLaunchKernelGGL(gen_orders_k, dim3(gen_blocks(rows + 1)), dim3(kGenThreads), 0,
                       ctx->stream, seed, rows, cust_rows, key, day, cust, fact_offsets);
    BOLAP_TRY(launch_check("gen_orders"));
    return exclusive_scan_i64(ctx, fact_offsets, rows + 1, plan.at<void>(o_tmp));
}

int bolap_gen_order_facts(bolap_ctx* ctx, uint64_t seed, int64_t order_rows,
                          const int64_t* fact_offsets, int64_t fact_rows, int32_t* fk_order,
                          int64_t* price, int64_t* disc) {
    BOLAP_ARG_CHECK(ctx && order_rows >= 0 && fact_offsets && fact_rows >= 0,
                    "bolap_gen_order_facts: bad arguments");
    if (order_rows == 0 || fact_rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(fk_order && price && disc, "bolap_gen_order_facts: null column");
    hipLaunchKernelGGL(gen_order_facts_k, dim3(gen_blocks(order_rows)), dim3(kGenThreads), 0,
                       ctx->stream, seed, order_rows, fact_offsets, fact_rows, fk_order, price,
                       disc);
    return launch_check("gen_order_facts");
}

int bolap_gen_iota_i32(bolap_ctx* ctx, int64_t rows, int32_t first, int32_t* out) {
    BOLAP_ARG_CHECK(ctx && rows >= 0, "bolap_gen_iota_i32: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(out && int64_t(first) + rows - 1 <= INT32_MAX,
                    "bolap_gen_iota_i32: null column or int32 overflow");
    hipLaunchKernelGGL(gen_iota_k, dim3(gen_blocks(rows)), dim3(kGenThreads), 0, ctx->stream,
                       rows, first, out);
    return launch_check("gen_iota");
}

int bolap_gen_uniform_i32(bolap_ctx* ctx, uint64_t seed, uint32_t stream, int64_t rows,
                          int32_t lo, int32_t hi, int32_t* out) {
    BOLAP_ARG_CHECK(ctx && rows >= 0 && lo <= hi, "bolap_gen_uniform_i32: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(out, "bolap_gen_uniform_i32: null column");
    hipLaunchKernelGGL(gen_uniform_k<int32_t>, dim3(gen_blocks(rows)), dim3(kGenThreads), 0,
                       ctx->stream, seed, stream, rows, lo, hi, out);
    return launch_check("gen_uniform_i32");
}

int bolap_gen_uniform_u8(bolap_ctx* ctx, uint64_t seed, uint32_t stream, int64_t rows,
                         int32_t lo, int32_t hi, uint8_t* out) {
    BOLAP_ARG_CHECK(ctx && rows >= 0 && 0 <= lo && lo <= hi && hi <= 255,
                    "bolap_gen_uniform_u8: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(out, "bolap_gen_uniform_u8: null column");
    hipLaunchKernelGGL(gen_uniform_k<uint8_t>, dim3(gen_blocks(rows)), dim3(kGenThreads), 0,
                       ctx->stream, seed, stream, rows, lo, hi, out);
    return launch_check("gen_uniform_u8");
}

int bolap_gen_zipf_i32(bolap_ctx* ctx, uint64_t seed, uint32_t stream, int64_t rows,
                       const int32_t* quantiles_dev, int64_t n_keys, uint64_t mul,
                       int32_t* out) {
    BOLAP_ARG_CHECK(ctx && rows >= 0 && n_keys >= 1 && n_keys <= INT32_MAX && mul >= 1,
                    "bolap_gen_zipf_i32: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(out && quantiles_dev, "bolap_gen_zipf_i32: null array");
    BOLAP_ARG_CHECK(std::gcd(mul % uint64_t(n_keys), uint64_t(n_keys)) == 1 || n_keys == 1,
                    "bolap_gen_zipf_i32: mul %llu is not coprime to %lld",
                    (unsigned long long)mul, (long long)n_keys);
    hipLaunchKernelGGL(gen_zipf_k, dim3(gen_blocks(rows)), dim3(kGenThreads), 0, ctx->stream,
                       seed, stream, rows, quantiles_dev, n_keys, mul, out);
    return launch_check("gen_zipf_i32");
}

int bolap_gen_price_disc(bolap_ctx* ctx, uint64_t seed, uint32_t stream, int64_t rows,
                         int64_t* price, int64_t* disc) {
    BOLAP_ARG_CHECK(ctx && rows >= 0, "bolap_gen_price_disc: bad arguments");
    if (rows == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(price && disc, "bolap_gen_price_disc: null column");
    hipLaunchKernelGGL(gen_price_disc_k, dim3(gen_blocks(rows)), dim3(kGenThreads), 0,
                       ctx->stream, seed, stream, rows, price, disc);
    return launch_check("gen_price_disc");
}

}  // extern "C"
