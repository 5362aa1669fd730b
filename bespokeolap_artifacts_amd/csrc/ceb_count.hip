// ceb_count.hip — factorised COUNT(*) of the CEB star queries over the
// reference's nullable column layout (bespoke_ceb/builder_impl.hpp:97-110:
// i32 values plus one is_null byte per row).
//
// run_q1a / run_q2a / run_q2b / run_q2c / run_q8a (bespoke_ceb/query_q1a.cpp:
// 241-621, query_q2a.cpp:240-503, ...) all follow one plan: title rows pass
// a production-year window and dimension masks; every fact table ("term")
// counts, per movie, its rows whose columns are non-null and pass their
// masks; COUNT(*) is the sum, over the passing title rows, of the product of
// their movie's term counts.  Three kinds of launch:
//
//   title_k   one pass over title: a pass flag per title row and a pass byte
//             per movie id (plain byte stores), then pack_k folds the bytes
//             into one bit per movie (the reference's title_match /
//             info_both_mask arrays, here a bitmap so that the fact scans'
//             random probes hit a 2.5 M-movie set that stays in L2: 312 KB);
//   term_k    one pass per fact table over its columns, four rows per lane
//             with 16-byte loads of the movie column; a row probes the movie
//             bit first and reads its predicate columns only when the movie
//             passed; for each of the table's terms (up to two share one
//             pass: 1a / 2a read movie_info twice) a passing row adds the
//             term's weight to the movie's u32 counter with a device atomic;
//   product_k one pass over the title flags: the product of the movie's
//             counters (u64, wrapping as the reference's int64 products do)
//             summed by DPP wave sums and one atomic per wave.
#include <cstring>

#include "common.hpp"

namespace bolap {
namespace {

constexpr int kT = 256;

struct Pred {
    const int32_t* col;
    const uint8_t* nul;
    const uint8_t* mask;
    int64_t mask_n;
};

struct PredSet {
    int32_t n;
    Pred p[BOLAP_CEB_MAX_TERM_PREDS];
};

BOLAP_DEV bool pred_pass(const Pred& p, int64_t row) {
    if (p.nul && p.nul[row]) return false;
    const int32_t v = p.col[row];
    return v >= 0 && int64_t(v) < p.mask_n && p.mask[v] != 0;
}

BOLAP_DEV bool preds_pass(const PredSet& s, int64_t row) {
    for (int i = 0; i < s.n; ++i)
        if (!pred_pass(s.p[i], row)) return false;
    return true;
}

BOLAP_DEV bool movie_bit(const uint32_t* __restrict__ bits, int32_t m) {
    return (bits[uint32_t(m) >> 5] >> (uint32_t(m) & 31)) & 1u;
}

__global__ __launch_bounds__(kT) void title_k(int64_t rows, const int32_t* __restrict__ id,
                                              const uint8_t* __restrict__ id_null,
                                              const int32_t* __restrict__ year,
                                              const uint8_t* __restrict__ year_null,
                                              int32_t year_lo, int32_t year_hi, PredSet preds,
                                              int32_t cap, uint8_t* __restrict__ flags,
                                              uint8_t* __restrict__ movie_pass,
                                              unsigned long long* __restrict__ n_pass) {
    uint64_t mine = 0;
    for (int64_t r = int64_t(blockIdx.x) * kT + threadIdx.x; r < rows;
         r += int64_t(gridDim.x) * kT) {
        bool pass = !(id_null && id_null[r]);
        const int32_t m = id[r];
        pass = pass && m >= 0 && m < cap;
        if (pass && year) {
            pass = !(year_null && year_null[r]);
            const int32_t y = year[r];
            pass = pass && y >= year_lo && y <= year_hi;
        }
        pass = pass && preds_pass(preds, r);
        flags[r] = pass ? 1 : 0;
        if (pass) {
            movie_pass[m] = 1;   // plain byte store: every writer stores the same value
            ++mine;
        }
    }
    const uint64_t s = wave_sum_u64(mine);
    if ((threadIdx.x & (kWave - 1)) == 0 && s) atomicAdd(n_pass, (unsigned long long)s);
}

// One bit per movie from the pass bytes: word w covers movies 32w .. 32w+31.
__global__ __launch_bounds__(kT) void pack_k(const uint8_t* __restrict__ movie_pass,
                                             int64_t words, uint32_t* __restrict__ bits) {
    for (int64_t w = int64_t(blockIdx.x) * kT + threadIdx.x; w < words;
         w += int64_t(gridDim.x) * kT) {
        const auto q = load_unit<uint32_t, 8>(reinterpret_cast<const uint32_t*>(movie_pass), w);
        uint32_t b = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t v = q.v[k];
            b |= ((v & 0xffu) ? 1u : 0u) << (4 * k) | ((v & 0xff00u) ? 1u : 0u) << (4 * k + 1) |
                 ((v & 0xff0000u) ? 1u : 0u) << (4 * k + 2) |
                 ((v & 0xff000000u) ? 1u : 0u) << (4 * k + 3);
        }
        bits[w] = b;
    }
}

// Up to two terms over the same fact rows (the same movie column): the
// movie column is read and its bit probed once for both.
struct TermGroup {
    int32_t n;
    PredSet ps[2];
    uint32_t weight[2];
    uint32_t* counts[2];
    unsigned long long* n_rows[2];
};

__global__ __launch_bounds__(kT) void term_k(int64_t rows, const int32_t* __restrict__ movie,
                                             const uint8_t* __restrict__ movie_null,
                                             TermGroup g, int32_t cap,
                                             const uint32_t* __restrict__ bits) {
    uint64_t mine0 = 0, mine1 = 0;
    const int64_t units = (rows + 3) / 4;
    const bool vec = (reinterpret_cast<uintptr_t>(movie) & 15) == 0;
    for (int64_t u = int64_t(blockIdx.x) * kT + threadIdx.x; u < units;
         u += int64_t(gridDim.x) * kT) {
        const int64_t r0 = u * 4;
        int32_t mv[4];
        if (vec && r0 + 3 < rows) {
            const auto q = load_unit<int32_t, 4>(movie, u);
            mv[0] = q.v[0], mv[1] = q.v[1], mv[2] = q.v[2], mv[3] = q.v[3];
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) mv[k] = r0 + k < rows ? movie[r0 + k] : -1;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t r = r0 + k;
            const int32_t m = mv[k];
            if (r >= rows || m < 0 || m >= cap || !movie_bit(bits, m)) continue;
            if (movie_null && movie_null[r]) continue;
            if (preds_pass(g.ps[0], r)) {
                atomicAdd(g.counts[0] + m, g.weight[0]);
                ++mine0;
            }
            if (g.n > 1 && preds_pass(g.ps[1], r)) {
                atomicAdd(g.counts[1] + m, g.weight[1]);
                ++mine1;
            }
        }
    }
    const uint64_t s0 = wave_sum_u64(mine0), s1 = wave_sum_u64(mine1);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (s0) atomicAdd(g.n_rows[0], (unsigned long long)s0);
        if (s1) atomicAdd(g.n_rows[1], (unsigned long long)s1);
    }
}

struct Terms {
    int32_t n;
    const uint32_t* counts[BOLAP_CEB_MAX_TERMS];
};

__global__ __launch_bounds__(kT) void product_k(int64_t rows, const int32_t* __restrict__ id,
                                                const uint8_t* __restrict__ flags, Terms t,
                                                unsigned long long* __restrict__ out) {
    uint64_t sum = 0, joined = 0;
    for (int64_t r = int64_t(blockIdx.x) * kT + threadIdx.x; r < rows;
         r += int64_t(gridDim.x) * kT) {
        if (!flags[r]) continue;
        const int32_t m = id[r];
        uint64_t p = 1;
        for (int i = 0; i < t.n && p; ++i) p *= uint64_t(t.counts[i][m]);
        if (p) {
            sum += p;
            ++joined;
        }
    }
    const uint64_t s = wave_sum_u64(sum), j = wave_sum_u64(joined);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (s) atomicAdd(out, (unsigned long long)s);
        if (j) atomicAdd(out + 1, (unsigned long long)j);
    }
}

__global__ void fill_u32_k(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        p[i] = 0;
}

unsigned grid_for(bolap_ctx* ctx, int64_t items, int per_cu = 8) {
    int64_t b = (items + kT - 1) / kT;
    const int64_t cap = int64_t(ctx->num_cus) * per_cu;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

int zero_words(bolap_ctx* ctx, void* p, int64_t words) {
    if (words <= 0) return BOLAP_OK;
    hipLaunchKernelGGL(fill_u32_k, dim3(grid_for(ctx, words)), dim3(kT), 0, ctx->stream,
                       static_cast<uint32_t*>(p), words);
    return launch_check("fill_u32_k");
}

int check_pred(const bolap_ceb_pred& p, const char* who) {
    BOLAP_ARG_CHECK(p.col && p.mask && p.mask_n > 0, "%s: predicate without column or mask",
                    who);
    BOLAP_ARG_CHECK(aligned(p.col, 4), "%s: misaligned predicate column", who);
    return BOLAP_OK;
}

PredSet pred_set(const bolap_ceb_pred* p, int32_t n) {
    PredSet s{};
    s.n = n;
    for (int i = 0; i < n; ++i) s.p[i] = Pred{p[i].col, p[i].is_null, p[i].mask, p[i].mask_n};
    return s;
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" int bolap_ceb_count(bolap_ctx* ctx, const bolap_ceb_count_args* a,
                               bolap_ceb_count_result* out) {
    BOLAP_ARG_CHECK(ctx && a && out, "bolap_ceb_count: null argument");
    BOLAP_ARG_CHECK(a->title_rows >= 0 && a->title_rows < (int64_t(1) << 31),
                    "bolap_ceb_count: %lld title rows", (long long)a->title_rows);
    BOLAP_ARG_CHECK(a->movie_cap >= 0, "bolap_ceb_count: movie_cap %d", a->movie_cap);
    BOLAP_ARG_CHECK(a->n_title_preds >= 0 && a->n_title_preds <= BOLAP_CEB_MAX_TITLE_PREDS,
                    "bolap_ceb_count: %d title predicates", a->n_title_preds);
    BOLAP_ARG_CHECK(a->n_terms >= 1 && a->n_terms <= BOLAP_CEB_MAX_TERMS,
                    "bolap_ceb_count: %d terms", a->n_terms);
    BOLAP_ARG_CHECK(a->title_rows == 0 || a->t_id, "bolap_ceb_count: null title id column");
    for (int i = 0; i < a->n_title_preds; ++i)
        BOLAP_TRY(check_pred(a->title_preds[i], "bolap_ceb_count"));
    for (int t = 0; t < a->n_terms; ++t) {
        const bolap_ceb_term& tm = a->terms[t];
        BOLAP_ARG_CHECK(tm.rows >= 0 && tm.rows < (int64_t(1) << 40),
                        "bolap_ceb_count: term %d has %lld rows", t, (long long)tm.rows);
        BOLAP_ARG_CHECK(tm.rows == 0 || tm.movie, "bolap_ceb_count: term %d without movie column",
                        t);
        BOLAP_ARG_CHECK(tm.n_preds >= 0 && tm.n_preds <= BOLAP_CEB_MAX_TERM_PREDS,
                        "bolap_ceb_count: term %d has %d predicates", t, tm.n_preds);
        for (int i = 0; i < tm.n_preds; ++i)
            BOLAP_TRY(check_pred(tm.preds[i], "bolap_ceb_count"));
    }
    std::memset(out, 0, sizeof(*out));
    const int32_t cap = a->movie_cap;
    const int64_t words = (int64_t(cap) + 31) / 32 + 1;
    WsArena ws;
    const size_t o_bits = ws.plan(size_t(words) * 4);
    const size_t o_pass = ws.plan(size_t(words) * 32);
    const size_t o_flags = ws.plan(size_t(a->title_rows) + 4);
    const size_t o_counts = ws.plan(size_t(a->n_terms) * size_t(cap + 1) * 4);
    BOLAP_TRY(ws_reserve(ctx, ws.off));
    ws.base = static_cast<char*>(ctx->ws);
    uint32_t* bits = ws.at<uint32_t>(o_bits);
    uint8_t* movie_pass = ws.at<uint8_t>(o_pass);
    uint8_t* flags = ws.at<uint8_t>(o_flags);
    uint32_t* counts = ws.at<uint32_t>(o_counts);
    // result slots: [0] title passes, [1] sum, [2] joined, [3..] term rows
    auto* res = static_cast<unsigned long long*>(ctx->result_dev);
    const int n_res = 3 + BOLAP_CEB_MAX_TERMS;
    BOLAP_TRY(zero_words(ctx, res, n_res * 2));
    BOLAP_TRY(zero_words(ctx, movie_pass, words * 8));
    BOLAP_TRY(zero_words(ctx, counts, int64_t(a->n_terms) * (cap + 1)));
    if (a->title_rows > 0) {
        TraceScope ts(ctx, "ceb_title");
        hipLaunchKernelGGL(title_k, dim3(grid_for(ctx, a->title_rows)), dim3(kT), 0, ctx->stream,
                           a->title_rows, a->t_id, a->t_id_null, a->t_year, a->t_year_null,
                           a->year_lo, a->year_hi, pred_set(a->title_preds, a->n_title_preds),
                           cap, flags, movie_pass, res);
        BOLAP_TRY(launch_check("title_k"));
    }
    hipLaunchKernelGGL(pack_k, dim3(grid_for(ctx, words)), dim3(kT), 0, ctx->stream, movie_pass,
                       words, bits);
    BOLAP_TRY(launch_check("pack_k"));
    Terms terms{};
    terms.n = a->n_terms;
    for (int t = 0; t < a->n_terms; ++t) terms.counts[t] = counts + size_t(t) * size_t(cap + 1);
    // terms over the same movie column share one scan, two at a time
    for (int t = 0; t < a->n_terms;) {
        const bolap_ceb_term& tm = a->terms[t];
        TermGroup g{};
        g.n = 1;
        if (t + 1 < a->n_terms && a->terms[t + 1].movie == tm.movie &&
            a->terms[t + 1].rows == tm.rows && a->terms[t + 1].movie_null == tm.movie_null)
            g.n = 2;
        for (int k = 0; k < g.n; ++k) {
            const bolap_ceb_term& x = a->terms[t + k];
            g.ps[k] = pred_set(x.preds, x.n_preds);
            g.weight[k] = x.weight;
            g.counts[k] = const_cast<uint32_t*>(terms.counts[t + k]);
            g.n_rows[k] = res + 3 + t + k;
        }
        if (tm.rows > 0 && a->title_rows > 0) {
            TraceScope ts(ctx, "ceb_term");
            hipLaunchKernelGGL(term_k, dim3(grid_for(ctx, (tm.rows + 3) / 4)), dim3(kT), 0,
                               ctx->stream, tm.rows, tm.movie, tm.movie_null, g, cap, bits);
            BOLAP_TRY(launch_check("term_k"));
        }
        t += g.n;
    }
    if (a->title_rows > 0) {
        TraceScope ts(ctx, "ceb_product");
        hipLaunchKernelGGL(product_k, dim3(grid_for(ctx, a->title_rows)), dim3(kT), 0,
                           ctx->stream, a->title_rows, a->t_id, flags, terms, res + 1);
        BOLAP_TRY(launch_check("product_k"));
    }
    unsigned long long host[3 + BOLAP_CEB_MAX_TERMS];
    BOLAP_HIP_TRY(hipMemcpyAsync(host, res, sizeof(host), hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    out->title_pass = int64_t(host[0]);
    out->count = int64_t(host[1]);
    out->joined = int64_t(host[2]);
    for (int t = 0; t < a->n_terms; ++t) out->term_rows[t] = int64_t(host[3 + t]);
    trace_count(ctx, "ceb_title_pass", out->title_pass);
    trace_count(ctx, "ceb_joined", out->joined);
    return BOLAP_OK;
}
