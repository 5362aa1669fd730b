// autotune.hip — the knob table's persistence and two autotuners that time
// the caller's own workload under each candidate configuration and keep the
// fastest in the context's knobs.
//
// No reference counterpart: the reference hard-codes its constants (the
// 16-bit radix of bespoke_tpch/query_q10.cpp:139, its loop unrolls and
// block sizes).  Both tuners check that every candidate gives the same
// result as the first one before it may win, so tuning never changes an
// answer.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "radix.hpp"

namespace bolap {
namespace {

constexpr int kT = 256;

// Mismatches between two (key, value) arrays, added into *bad.
__global__ __launch_bounds__(kT) void pairs_diff_k(const uint64_t* __restrict__ ka,
                                                   const uint32_t* __restrict__ va,
                                                   const uint64_t* __restrict__ kb,
                                                   const uint32_t* __restrict__ vb, int64_t n,
                                                   unsigned long long* __restrict__ bad) {
    uint64_t d = 0;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        d += (ka[i] != kb[i] || va[i] != vb[i]) ? 1u : 0u;
    d = wave_sum(d);
    if ((threadIdx.x & (kWave - 1)) == 0 && d) atomicAdd(bad, (unsigned long long)d);
}

__global__ void zero_u64_one_k(unsigned long long* p) { *p = 0; }

double median_of(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    return v[v.size() / 2];
}

bool knob_name_ok(const std::string& s) {
    if (s.empty() || s.size() > 63) return false;
    for (char c : s)
        if (!(std::isalnum(static_cast<unsigned char>(c)) || c == '.' || c == '_')) return false;
    return true;
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" {

int bolap_tune_save(bolap_ctx* ctx, const char* path) {
    BOLAP_ARG_CHECK(ctx && path && path[0], "bolap_tune_save: null argument");
    FILE* f = std::fopen(path, "w");
    BOLAP_ARG_CHECK(f != nullptr, "bolap_tune_save: cannot open %s: %s", path,
                    std::strerror(errno));
    std::fprintf(f, "# libbolap knobs: name value\n");
    for (const auto& kv : ctx->knobs)
        std::fprintf(f, "%s %lld\n", kv.first.c_str(), (long long)kv.second);
    const bool ok = std::fclose(f) == 0;
    BOLAP_ARG_CHECK(ok, "bolap_tune_save: cannot write %s", path);
    return BOLAP_OK;
}

int bolap_tune_load(bolap_ctx* ctx, const char* path) {
    BOLAP_ARG_CHECK(ctx && path && path[0], "bolap_tune_load: null argument");
    FILE* f = std::fopen(path, "r");
    BOLAP_ARG_CHECK(f != nullptr, "bolap_tune_load: cannot open %s: %s", path,
                    std::strerror(errno));
    // parse everything first: a malformed file leaves the knobs unchanged
    std::vector<std::pair<std::string, int64_t>> got;
    char line[256];
    int ln = 0;
    int status = BOLAP_OK;
    while (std::fgets(line, sizeof(line), f)) {
        ++ln;
        if (line[0] == '#' || line[0] == '\n' || line[0] == '\0') continue;
        char name[128];
        long long v = 0;
        char extra = 0;
        if (std::sscanf(line, "%127s %lld %c", name, &v, &extra) != 2 || !knob_name_ok(name)) {
            set_error("bolap_tune_load: %s line %d is not 'name value'", path, ln);
            status = BOLAP_ERR_ARG;
            break;
        }
        got.emplace_back(name, int64_t(v));
    }
    std::fclose(f);
    if (status != BOLAP_OK) return status;
    for (const auto& kv : got) ctx->knobs[kv.first] = kv.second;
    return BOLAP_OK;
}

// Times bolap_join_sum on the caller's arguments under every combination of
// probe loop (join.probe_variant 1..5), LDS filter bits per key
// (join.lds_bloom_bits 6..8) and probe workgroups per CU
// (join.blocks_per_cu 1, 2); keeps the fastest (median of `reps` calls).
int bolap_join_autotune(bolap_ctx* ctx, const bolap_join_sum_args* args, int32_t reps,
                        double* best_ms) {
    BOLAP_ARG_CHECK(ctx && args && reps >= 1 && reps <= 100,
                    "bolap_join_autotune: bad arguments");
    const char* kn[3] = {"join.probe_variant", "join.lds_bloom_bits", "join.blocks_per_cu"};
    int64_t saved[3];
    for (int i = 0; i < 3; ++i) saved[i] = ctx->knob(kn[i], 0);
    auto restore = [&] {
        for (int i = 0; i < 3; ++i) ctx->knobs[kn[i]] = saved[i];
    };
    bolap_join_sum_result first{}, r{};
    int status = bolap_join_sum(ctx, args, &first);
    if (status != BOLAP_OK) return status;
    double best = 1e30;
    int64_t bv = 3, bb = 7, bp = 2;
    for (int64_t v = 1; v <= 5; ++v)
        for (int64_t b = 6; b <= 8; ++b)
            for (int64_t p = 1; p <= 2; ++p) {
                ctx->knobs[kn[0]] = v;
                ctx->knobs[kn[1]] = b;
                ctx->knobs[kn[2]] = p;
                status = bolap_join_sum(ctx, args, &r);  // warm
                std::vector<double> t;
                for (int k = 0; k < reps && status == BOLAP_OK; ++k) {
                    const auto t0 = std::chrono::steady_clock::now();
                    status = bolap_join_sum(ctx, args, &r);
                    const auto t1 = std::chrono::steady_clock::now();
                    t.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
                }
                if (status != BOLAP_OK) {
                    restore();
                    return status;
                }
                if (r.revenue != first.revenue || r.matches != first.matches ||
                    r.build_rows != first.build_rows) {
                    restore();
                    set_error("bolap_join_autotune: variant %lld/%lld/%lld changed the result",
                              (long long)v, (long long)b, (long long)p);
                    return BOLAP_ERR_INTERNAL;
                }
                const double ms = median_of(t);
                if (ms < best) {
                    best = ms;
                    bv = v;
                    bb = b;
                    bp = p;
                }
            }
    ctx->knobs[kn[0]] = bv;
    ctx->knobs[kn[1]] = bb;
    ctx->knobs[kn[2]] = bp;
    if (best_ms) *best_ms = best;
    return BOLAP_OK;
}

// Times the radix sort of a copy of the caller's pairs under each digit
// width (radix.digit_bits 5..8, i.e. ceil(key_bits / width) passes); keeps
// the fastest.  The caller's arrays are not modified.
int bolap_sort_autotune(bolap_ctx* ctx, const uint64_t* keys, const uint32_t* vals, int64_t n,
                        int32_t key_bits, int32_t reps, double* best_ms) {
    BOLAP_ARG_CHECK(ctx && (n == 0 || (keys && vals)) && n >= 0 && reps >= 1 && reps <= 100 &&
                        key_bits >= 1 && key_bits <= 64,
                    "bolap_sort_autotune: bad arguments");
    BOLAP_ARG_CHECK(n >= 2, "bolap_sort_autotune: nothing to sort (n = %lld)", (long long)n);
    WsArena ws;
    const size_t o_k = ws.plan(size_t(n) * 8);
    const size_t o_v = ws.plan(size_t(n) * 4);
    const size_t o_rk = ws.plan(size_t(n) * 8);  // the first width's result
    const size_t o_rv = ws.plan(size_t(n) * 4);
    const size_t o_bad = ws.plan(8);
    const size_t o_scr = ws.plan(radix_sort_scratch_bytes(ctx, n));
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    uint64_t* k = ws.at<uint64_t>(o_k);
    uint32_t* v = ws.at<uint32_t>(o_v);
    const int64_t saved = ctx->knob("radix.digit_bits", 0);
    hipEvent_t e0, e1;
    BOLAP_HIP_TRY(hipEventCreate(&e0));
    BOLAP_HIP_TRY(hipEventCreate(&e1));
    int status = BOLAP_OK;
    double best = 1e30;
    int64_t bw = 8;
    for (int64_t w = 8; w >= 5 && status == BOLAP_OK; --w) {
        ctx->knobs["radix.digit_bits"] = w;
        std::vector<double> t;
        for (int rep = 0; rep <= reps && status == BOLAP_OK; ++rep) {  // rep 0 warms
            if (hipMemcpyAsync(k, keys, size_t(n) * 8, hipMemcpyDeviceToDevice, ctx->stream) !=
                    hipSuccess ||
                hipMemcpyAsync(v, vals, size_t(n) * 4, hipMemcpyDeviceToDevice, ctx->stream) !=
                    hipSuccess) {
                set_error("bolap_sort_autotune: copy failed");
                status = BOLAP_ERR_HIP;
                break;
            }
            (void)hipEventRecord(e0, ctx->stream);
            status = radix_sort_pairs(ctx, k, v, n, key_bits, ws.at<void>(o_scr), nullptr);
            (void)hipEventRecord(e1, ctx->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) t.push_back(ms);
        }
        if (status != BOLAP_OK) break;
        if (w == 8) {
            (void)hipMemcpyAsync(ws.at<uint64_t>(o_rk), k, size_t(n) * 8,
                                 hipMemcpyDeviceToDevice, ctx->stream);
            (void)hipMemcpyAsync(ws.at<uint32_t>(o_rv), v, size_t(n) * 4,
                                 hipMemcpyDeviceToDevice, ctx->stream);
        } else {
            auto* bad = ws.at<unsigned long long>(o_bad);
            hipLaunchKernelGGL(zero_u64_one_k, dim3(1), dim3(1), 0, ctx->stream, bad);
            const int64_t blocks = std::min<int64_t>((n + kT - 1) / kT, int64_t(ctx->num_cus) * 4);
            hipLaunchKernelGGL(pairs_diff_k, dim3(unsigned(blocks)), dim3(kT), 0, ctx->stream, k,
                               v, ws.at<uint64_t>(o_rk), ws.at<uint32_t>(o_rv), n, bad);
            unsigned long long h = 0;
            if (hipMemcpyAsync(&h, bad, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess) {
                set_error("bolap_sort_autotune: comparison failed");
                status = BOLAP_ERR_HIP;
                break;
            }
            if (h != 0) {
                set_error("bolap_sort_autotune: %lld-bit digits changed %llu entries",
                          (long long)w, h);
                status = BOLAP_ERR_INTERNAL;
                break;
            }
        }
        const double ms = median_of(t);
        if (ms < best) {
            best = ms;
            bw = w;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->knobs["radix.digit_bits"] = status == BOLAP_OK ? bw : saved;
    if (status == BOLAP_OK && best_ms) *best_ms = best;
    return status;
}

}  // extern "C"
