// stats.hip — table statistics read by the learned cardinality estimator:
// HyperLogLog registers of a key column and sample counts of star plans.
//
// No reference counterpart: the reference sizes its tables from hard-coded
// selectivities (bespoke_tpch/query_q20.cpp:16, 405-417).  The sketches
// here are computed once per column (at load, like the reference builder's
// per-column min/max, bespoke_ceb/builder_impl.hpp:88-118), and the sample
// counts once per batch of plans, all plans in one launch.
#include <cstring>
#include <vector>

#include "preds.hpp"
#include "prims.hpp"

namespace bolap {
namespace {

constexpr int kT = 256;
constexpr int kRegs = BOLAP_HLL_REGS;
constexpr int kSample = BOLAP_SAMPLE_ROWS;
constexpr int kMaxDims = BOLAP_STAR_MAX_DIMS;
constexpr uint32_t kDimStream = 0x53440000u;   // sample row streams
constexpr uint32_t kFactStream = 0x53460000u;
constexpr int kDistLog2 = 13;                  // 8192-slot LDS table for sampled group keys
constexpr int kDistCap = 1 << kDistLog2;

// Register index and rank of one value (p = 12).
BOLAP_DEV void hll_point(int32_t v, uint32_t& idx, uint32_t& rank) {
    const uint64_t h = mix64(uint64_t(uint32_t(v)) + 0x9E3779B97F4A7C15ULL);
    idx = uint32_t(h >> 52);
    const uint64_t w = h << 12;
    rank = w ? uint32_t(__clzll(w)) + 1u : 53u;
}

// One partial register array per workgroup (LDS max), reduced by hll_merge_k.
__global__ __launch_bounds__(kT) void hll_partial_k(const int32_t* __restrict__ col, int64_t n,
                                                    uint8_t* __restrict__ partial) {
    __shared__ uint32_t regs[kRegs];
    for (int i = threadIdx.x; i < kRegs; i += kT) regs[i] = 0u;
    __syncthreads();
    const int64_t stride = int64_t(gridDim.x) * kT;
    constexpr int U = 4;
    for (int64_t base = int64_t(blockIdx.x) * kT + threadIdx.x; base < n; base += U * stride) {
        int32_t v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t r = base + u * stride;
            v[u] = r < n ? col[r] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (base + u * stride >= n) continue;
            uint32_t idx, rank;
            hll_point(v[u], idx, rank);
            atomicMax(regs + idx, rank);
        }
    }
    __syncthreads();
    uint8_t* out = partial + size_t(blockIdx.x) * kRegs;
    for (int i = threadIdx.x; i < kRegs; i += kT) out[i] = uint8_t(regs[i]);
}

__global__ __launch_bounds__(kT) void hll_merge_k(const uint8_t* __restrict__ partial, int parts,
                                                  uint8_t* __restrict__ regs) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= kRegs) return;
    uint32_t m = 0;
    for (int p = 0; p < parts; ++p) m = max(m, uint32_t(partial[size_t(p) * kRegs + i]));
    regs[i] = uint8_t(m);
}

struct SDim {
    const int32_t* fk;
    int64_t rows;
    int32_t kmin;
    PredSet ps;
};

struct SPlan {
    int32_t n;
    SDim d[kMaxDims];
    const int32_t* gkey;
    int64_t fact_rows;
};

BOLAP_DEV int64_t sample_row(uint64_t seed, uint32_t stream, int i, int64_t rows) {
    return int64_t(((gen_hash(seed, stream, uint64_t(i)) >> 32) * uint64_t(rows)) >> 32);
}

BOLAP_DEV uint32_t block_sum(uint32_t v, uint32_t* s_tmp) {
    const uint32_t w = uint32_t(wave_sum(uint64_t(v)));
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) s_tmp[threadIdx.x / kWave] = w;
    __syncthreads();
    uint32_t t = 0;
    for (int i = 0; i < kT / kWave; ++i) t += s_tmp[i];
    return t;
}

// One workgroup per plan.
__global__ __launch_bounds__(kT) void sample_counts_k(const SPlan* __restrict__ plans,
                                                      uint64_t seed, int64_t* __restrict__ out) {
    extern __shared__ uint32_t dyn[];  // hk[kDistCap] keys, hc[kDistCap] counts
    uint32_t* hk = dyn;
    uint32_t* hc = dyn + kDistCap;
    __shared__ uint32_t s_tmp[kT / kWave];
    __shared__ uint32_t s_neg;  // sampled group key 0xFFFFFFFF (stored key 0 is "empty")
    const SPlan& p = plans[blockIdx.x];
    for (int i = threadIdx.x; i < kDistCap; i += kT) {
        hk[i] = 0u;
        hc[i] = 0u;
    }
    if (threadIdx.x == 0) s_neg = 0u;
    __syncthreads();
    uint32_t dpass[kMaxDims] = {0, 0, 0}, jpass[kMaxDims] = {0, 0, 0}, all = 0;
    for (int i = threadIdx.x; i < kSample; i += kT) {
        for (int j = 0; j < p.n; ++j) {
            const SDim& d = p.d[j];
            if (d.rows > 0 && eval_preds(d.ps, sample_row(seed, kDimStream, i, d.rows)))
                ++dpass[j];
        }
        if (p.fact_rows <= 0) continue;
        const int64_t fr = sample_row(seed, kFactStream, i, p.fact_rows);
        bool ok_all = true;
        for (int j = 0; j < p.n; ++j) {
            const SDim& d = p.d[j];
            const int64_t row = int64_t(d.fk[fr]) - d.kmin;
            const bool ok = row >= 0 && row < d.rows && eval_preds(d.ps, row);
            jpass[j] += ok;
            ok_all = ok_all && ok;
        }
        if (!ok_all) continue;
        ++all;
        const uint32_t key1 = uint32_t(p.gkey[fr]) + 1u;
        if (key1 == 0u) {
            atomicAdd(&s_neg, 1u);
            continue;
        }
        uint32_t s = (key1 * 0x9E3779B1u) >> (32 - kDistLog2);
        for (int probe = 0; probe < kDistCap; ++probe) {  // 4096 keys at most: always a slot
            const uint32_t prev = atomicCAS(hk + s, 0u, key1);
            if (prev == 0u || prev == key1) {
                atomicAdd(hc + s, 1u);
                break;
            }
            s = (s + 1) & (kDistCap - 1);
        }
    }
    int64_t* o = out + size_t(blockIdx.x) * BOLAP_SAMPLE_COUNTS;
    for (int j = 0; j < kMaxDims; ++j) {
        const uint32_t a = block_sum(dpass[j], s_tmp);
        const uint32_t b = block_sum(jpass[j], s_tmp);
        if (threadIdx.x == 0) {
            o[j] = a;
            o[3 + j] = b;
        }
    }
    const uint32_t a_all = block_sum(all, s_tmp);
    uint32_t distinct = 0, f1 = 0, f2 = 0;
    for (int i = threadIdx.x; i < kDistCap; i += kT) {
        const uint32_t c = hc[i];
        distinct += c > 0;
        f1 += c == 1;
        f2 += c == 2;
    }
    const uint32_t nd = block_sum(distinct, s_tmp);
    const uint32_t n1 = block_sum(f1, s_tmp);
    const uint32_t n2 = block_sum(f2, s_tmp);
    if (threadIdx.x == 0) {
        const uint32_t neg = s_neg;
        o[6] = a_all;
        o[7] = nd + (neg > 0);
        o[8] = n1 + (neg == 1);
        o[9] = n2 + (neg == 2);
        for (int k = 10; k < BOLAP_SAMPLE_COUNTS; ++k) o[k] = 0;
    }
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" {

int bolap_hll_i32(bolap_ctx* ctx, const int32_t* col, int64_t n, uint8_t* regs_host) {
    BOLAP_ARG_CHECK(ctx && regs_host && n >= 0, "bolap_hll_i32: bad arguments");
    if (n == 0) {
        std::memset(regs_host, 0, kRegs);
        return BOLAP_OK;
    }
    BOLAP_ARG_CHECK(col, "bolap_hll_i32: null column");
    TraceScope ts(ctx, "hll");
    int64_t parts = (n + kT * 4 - 1) / (kT * 4);
    parts = std::min<int64_t>(parts, int64_t(ctx->num_cus) * 2);
    parts = std::max<int64_t>(parts, 1);
    WsArena ws;
    const size_t o_part = ws.plan(size_t(parts) * kRegs);
    const size_t o_regs = ws.plan(kRegs);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    hipLaunchKernelGGL(hll_partial_k, dim3(unsigned(parts)), dim3(kT), 0, ctx->stream, col, n,
                       ws.at<uint8_t>(o_part));
    hipLaunchKernelGGL(hll_merge_k, dim3(kRegs / kT), dim3(kT), 0, ctx->stream,
                       ws.at<uint8_t>(o_part), int(parts), ws.at<uint8_t>(o_regs));
    BOLAP_TRY(launch_check("hll"));
    BOLAP_TRY(pinned_reserve(ctx, kRegs));
    BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, ws.at<uint8_t>(o_regs), kRegs,
                                 hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::memcpy(regs_host, ctx->pinned, kRegs);
    return BOLAP_OK;
}

int bolap_star_sample_counts(bolap_ctx* ctx, const bolap_star_args* plans, int32_t n_plans,
                             uint64_t seed, int64_t* counts_host) {
    BOLAP_ARG_CHECK(ctx && n_plans >= 0, "bolap_star_sample_counts: bad arguments");
    if (n_plans == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(plans && counts_host, "bolap_star_sample_counts: null array");
    std::vector<SPlan> host(static_cast<size_t>(n_plans));
    for (int32_t i = 0; i < n_plans; ++i) {
        const bolap_star_args& a = plans[i];
        SPlan& s = host[size_t(i)];
        std::memset(&s, 0, sizeof(s));
        BOLAP_ARG_CHECK(a.n_dims >= 1 && a.n_dims <= kMaxDims,
                        "bolap_star_sample_counts: plan %d has %d dimensions", i, a.n_dims);
        BOLAP_ARG_CHECK(a.fact_rows >= 0 && (a.fact_rows == 0 || a.group_key),
                        "bolap_star_sample_counts: plan %d fact columns", i);
        s.n = a.n_dims;
        s.gkey = a.group_key;
        s.fact_rows = a.fact_rows;
        for (int j = 0; j < a.n_dims; ++j) {
            const bolap_star_dim& d = a.dims[j];
            BOLAP_ARG_CHECK(d.rows >= 0 && (d.rows == 0 || d.key) &&
                                (a.fact_rows == 0 || d.fact_fk),
                            "bolap_star_sample_counts: plan %d dimension %d", i, j);
            s.d[j].fk = d.fact_fk;
            s.d[j].rows = d.rows;
            s.d[j].kmin = d.key_min;
            BOLAP_TRY(make_predset(s.d[j].ps, d.n_preds, d.preds, d.rows,
                                   "bolap_star_sample_counts"));
        }
    }
    TraceScope ts(ctx, "sample_counts");
    const size_t out_bytes = size_t(n_plans) * BOLAP_SAMPLE_COUNTS * 8;
    WsArena ws;
    const size_t o_plans = ws.plan(host.size() * sizeof(SPlan));
    const size_t o_out = ws.plan(out_bytes);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    BOLAP_HIP_TRY(hipMemcpyAsync(ws.at<SPlan>(o_plans), host.data(), host.size() * sizeof(SPlan),
                                 hipMemcpyHostToDevice, ctx->stream));
    BOLAP_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(sample_counts_k),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      2 * kDistCap * 4));
    hipLaunchKernelGGL(sample_counts_k, dim3(unsigned(n_plans)), dim3(kT),
                       size_t(2 * kDistCap * 4), ctx->stream,
                       ws.at<SPlan>(o_plans), seed, ws.at<int64_t>(o_out));
    BOLAP_TRY(launch_check("sample_counts_k"));
    BOLAP_HIP_TRY(hipMemcpyAsync(counts_host, ws.at<int64_t>(o_out), out_bytes,
                                 hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BOLAP_OK;
}

}  // extern "C"
