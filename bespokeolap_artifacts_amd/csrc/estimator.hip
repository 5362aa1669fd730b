// estimator.hip — learned sizing as a library call: persisted MLP weights,
// plan featurisation and the hint policy, so that a C or C++ caller sizes a
// star pipeline (pass_hint per dimension, group_hint) without Python.
//
// No reference counterpart: the reference sizes its tables from hard-coded
// selectivities (bespoke_tpch/query_q20.cpp:16, 405-417) and grows them on
// overflow (:158-160).  Hints only size tables: one that proves too small
// costs an exact rerun inside bolap_star_groupby, never a wrong result.
//
// The features and the policy are those the estimator was trained with
// (the package's estimator.py: features(), gee_groups(), from_targets(),
// sizing_hints()), computed here in the same order of operations so that
// both sides give the same hints for the same plan:
//   per plan with k dimensions, k feature rows (the dimension list rotated
//   so that each dimension takes slot 0 once) from the columns' HyperLogLog
//   registers, the plan's sample counts (bolap_star_sample_counts) and its
//   predicate constants; one batched bolap_mlp_predict; outputs
//   de-standardised (y * sd + mu in fp32) and mapped back to passing rows,
//   join rows and groups; join rows and groups averaged over the rotations
//   in log1p space; hints = estimate x margin, the group hint capped by the
//   key span and the join estimate and never below the distinct keys the
//   sample saw.
//
// File (little-endian): "BOLAPEST", u32 version 1, u32 parameter count,
// fp32 parameters (bolap_mlp_get_params order), fp32 mu[3], fp32 sd[3],
// f64 margin_dim, f64 margin_group.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.hpp"

struct bolap_estimator {
    bolap_ctx* ctx = nullptr;
    bolap_mlp* mlp = nullptr;
    std::vector<float> params;
    float mu[BOLAP_MLP_OUT] = {}, sd[BOLAP_MLP_OUT] = {};
    double margin_dim = 2.0, margin_group = 2.0;
};

namespace bolap {
namespace {

constexpr int kFeat = BOLAP_MLP_IN;
constexpr int kSample = BOLAP_SAMPLE_ROWS;
constexpr char kMagic[8] = {'B', 'O', 'L', 'A', 'P', 'E', 'S', 'T'};

double l_(double x, double scale) { return std::log10(std::max(x, 1e-12)) / scale; }

// HyperLogLog cardinality (Flajolet et al. 2007) with linear counting for
// small ranges, m = 4096 registers.  The harmonic sum is taken per register
// value in ascending order (each term count x 2^-value is exact), so that
// any implementation that sums in that order gets the same bits.
double hll_estimate(const uint8_t* regs) {
    const double m = BOLAP_HLL_REGS;
    const double alpha = 0.7213 / (1 + 1.079 / m);
    int64_t cnt[256] = {};
    for (int i = 0; i < BOLAP_HLL_REGS; ++i) ++cnt[regs[i]];
    double sum = 0;
    for (int v = 0; v < 256; ++v)
        if (cnt[v]) sum += std::ldexp(double(cnt[v]), -v);
    double est = alpha * m * m / sum;
    if (est <= 2.5 * m && cnt[0]) est = m * std::log(m / double(cnt[0]));
    return est;
}

// Column statistics: 64-bin histogram of register values (fractions of the
// registers) and the distinct estimate.
struct ColStats {
    const bolap_est_column* c = nullptr;
    float hist[64] = {};
    double distinct = 0;
};

const ColStats* find_col(const std::vector<ColStats>& cs, const void* col) {
    for (const auto& s : cs)
        if (s.c->col == col) return &s;
    return nullptr;
}

// Guaranteed-error estimator of distinct values (Charikar et al. 2000).
double gee_groups(const int64_t* counts, double joined) {
    const int64_t n = counts[6];
    if (n <= 0) return 0.0;
    const double d = double(counts[7]), f1 = double(counts[8]);
    return std::sqrt(std::max(joined, double(n)) / double(n)) * f1 + (d - f1);
}

int popcount_set(bolap_ctx* ctx, const bolap_pred& p, int64_t* n) {
    int64_t c = 0;
    if (!p.set) {
        for (int w = 0; w < 4; ++w) c += __builtin_popcountll(p.inline_set[w]);
    } else {
        std::vector<uint32_t> bits(size_t((p.set_bits + 31) / 32));
        BOLAP_HIP_TRY(hipMemcpyAsync(bits.data(), p.set, bits.size() * 4, hipMemcpyDeviceToHost,
                                     ctx->stream));
        BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (uint32_t b : bits) c += __builtin_popcount(b);
    }
    *n = c;
    return BOLAP_OK;
}

// Features of one rotation: dims[r..] + dims[..r] with the counts rotated
// alike; `group` the plan's group column statistics.
int features(bolap_ctx* ctx, const std::vector<ColStats>& cs, const bolap_star_args& a, int r,
             const int64_t* c0, float* f) {
    const int k = a.n_dims;
    int idx[BOLAP_STAR_MAX_DIMS];
    int64_t c[BOLAP_SAMPLE_COUNTS];
    std::memcpy(c, c0, sizeof(c));
    for (int i = 0; i < k; ++i) {
        idx[i] = (i + r) % k;
        c[i] = c0[idx[i]];
        c[3 + i] = c0[3 + idx[i]];
    }
    std::fill(f, f + kFeat, 0.0f);
    const ColStats* fk[BOLAP_STAR_MAX_DIMS] = {};
    for (int i = 0; i < k; ++i) {
        fk[i] = find_col(cs, a.dims[idx[i]].fact_fk);
        BOLAP_ARG_CHECK(fk[i] && fk[i]->c->kind == BOLAP_EST_FK,
                        "bolap_estimator_hints: a dimension's fact key has no FK statistics");
    }
    const ColStats* g = find_col(cs, a.group_key);
    BOLAP_ARG_CHECK(g && (g->c->kind == BOLAP_EST_FK || g->c->kind == BOLAP_EST_GROUP),
                    "bolap_estimator_hints: the group column has no statistics");
    std::copy(fk[0]->hist, fk[0]->hist + 64, f);
    std::copy(g->hist, g->hist + 64, f + 64);
    if (k > 1) std::copy(fk[1]->hist, fk[1]->hist + 64, f + 128);
    int o = 192;
    for (int j = 0; j < BOLAP_STAR_MAX_DIMS; ++j, o += 12) {
        if (j >= k) continue;
        const bolap_star_dim& d = a.dims[idx[j]];
        f[o + 0] = 1.0f;
        f[o + 1] = float(l_(double(d.rows), 8));
        for (int q = 0; q < d.n_preds; ++q) {
            const bolap_pred& p = d.preds[q];
            const ColStats* pc = find_col(cs, p.col);
            BOLAP_ARG_CHECK(pc && (pc->c->kind == BOLAP_EST_VALUE ||
                                   pc->c->kind == BOLAP_EST_CATEGORY),
                            "bolap_estimator_hints: a predicate column has no statistics");
            const double dom = double(pc->c->span);
            if (pc->c->kind == BOLAP_EST_VALUE) {
                const double w = double(p.hi - p.lo + 1) / dom;
                f[o + 2] = 1.0f;
                f[o + 3] = float(double(p.lo) / dom);
                f[o + 4] = float(w);
                f[o + 5] = float(l_(w, 6));
            } else {
                int64_t n = p.hi - p.lo + 1;
                if (p.op == BOLAP_OP_IN_SET) BOLAP_TRY(popcount_set(ctx, p, &n));
                f[o + 6] = 1.0f;
                f[o + 7] = float(double(n) / dom);
            }
        }
        f[o + 8] = float(double(c[j]) / kSample);
        f[o + 9] = float(l_((double(c[j]) + 0.5) / kSample, 4));
        f[o + 10] = float(double(c[3 + j]) / kSample);
        f[o + 11] = float(l_((double(c[3 + j]) + 0.5) / kSample, 4));
    }
    const int64_t nf = a.fact_rows;
    const double gspan = double(g->c->span);
    const double joined = double(nf * c[6]) / kSample;
    f[228] = float(double(c[6]) / kSample);
    f[229] = float(l_((double(c[6]) + 0.5) / kSample, 4));
    f[230] = float(l_(double(nf), 8));
    f[231] = float(l_(g->distinct + 1, 8));
    f[232] = float(l_(gspan, 8));
    if (g->c->kind == BOLAP_EST_FK) {
        f[233] = 1.0f;
        for (int j = 0; j < k; ++j)
            if (fk[j] == g) {
                f[234 + j] = 1.0f;
                break;
            }
    }
    f[237] = float(double(c[7]) / kSample);
    f[238] = float(double(c[8]) / kSample);
    f[239] = float(double(c[9]) / kSample);
    f[240] = float(l_(double(c[7] + 1), 4));
    f[241] = float(l_(gee_groups(c, joined) + 1, 8));
    f[242] = float(l_(joined + 1, 8));
    f[243] = float(double(k) / 3);
    for (int j = 0; j < std::min(3, k); ++j) f[244 + j] = float(l_(fk[j]->distinct + 1, 8));
    double sel = 1.0;
    for (int j = 0; j < k; ++j) sel *= (double(c[j]) + 0.5) / kSample;
    f[247] = float(l_(sel, 8));
    f[255] = 1.0f;
    return BOLAP_OK;
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" int bolap_estimator_create(bolap_ctx* ctx, const float* params, const float* mu_sd,
                                      const double* margins, bolap_estimator** out) {
    BOLAP_ARG_CHECK(ctx && params && mu_sd && margins && out,
                    "bolap_estimator_create: null argument");
    *out = nullptr;
    auto* e = new bolap_estimator();
    e->ctx = ctx;
    e->params.assign(params, params + bolap_mlp_num_params());
    for (int i = 0; i < BOLAP_MLP_OUT; ++i) {
        e->mu[i] = mu_sd[i];
        e->sd[i] = mu_sd[BOLAP_MLP_OUT + i];
    }
    e->margin_dim = margins[0];
    e->margin_group = margins[1];
    int st = bolap_mlp_create(ctx, 0, &e->mlp);
    if (st == BOLAP_OK) st = bolap_mlp_set_params(e->mlp, e->params.data());
    if (st != BOLAP_OK) {
        if (e->mlp) bolap_mlp_destroy(e->mlp);
        delete e;
        return st;
    }
    *out = e;
    return BOLAP_OK;
}

extern "C" int bolap_estimator_destroy(bolap_estimator* est) {
    if (!est) return BOLAP_OK;
    if (est->mlp) bolap_mlp_destroy(est->mlp);
    delete est;
    return BOLAP_OK;
}

extern "C" int bolap_estimator_save(const bolap_estimator* est, const char* path) {
    BOLAP_ARG_CHECK(est && path, "bolap_estimator_save: null argument");
    FILE* fp = std::fopen(path, "wb");
    BOLAP_ARG_CHECK(fp, "bolap_estimator_save: cannot open %s", path);
    const uint32_t head[2] = {1u, uint32_t(est->params.size())};
    const double m[2] = {est->margin_dim, est->margin_group};
    bool ok = std::fwrite(kMagic, 1, 8, fp) == 8 && std::fwrite(head, 4, 2, fp) == 2 &&
              std::fwrite(est->params.data(), 4, est->params.size(), fp) == est->params.size() &&
              std::fwrite(est->mu, 4, BOLAP_MLP_OUT, fp) == BOLAP_MLP_OUT &&
              std::fwrite(est->sd, 4, BOLAP_MLP_OUT, fp) == BOLAP_MLP_OUT &&
              std::fwrite(m, 8, 2, fp) == 2;
    ok = (std::fclose(fp) == 0) && ok;
    BOLAP_ARG_CHECK(ok, "bolap_estimator_save: write to %s failed", path);
    return BOLAP_OK;
}

extern "C" int bolap_estimator_load(bolap_ctx* ctx, const char* path, bolap_estimator** out) {
    BOLAP_ARG_CHECK(ctx && path && out, "bolap_estimator_load: null argument");
    *out = nullptr;
    FILE* fp = std::fopen(path, "rb");
    BOLAP_ARG_CHECK(fp, "bolap_estimator_load: cannot open %s", path);
    char magic[8];
    uint32_t head[2] = {0, 0};
    std::vector<float> p(static_cast<size_t>(bolap_mlp_num_params()));
    float musd[2 * BOLAP_MLP_OUT];
    double m[2];
    bool ok = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
              std::fread(head, 4, 2, fp) == 2 && head[0] == 1u && head[1] == p.size() &&
              std::fread(p.data(), 4, p.size(), fp) == p.size() &&
              std::fread(musd, 4, 2 * BOLAP_MLP_OUT, fp) == 2 * BOLAP_MLP_OUT &&
              std::fread(m, 8, 2, fp) == 2;
    std::fclose(fp);
    BOLAP_ARG_CHECK(ok, "bolap_estimator_load: %s is not an estimator file of version 1 with "
                    "%zu parameters", path, p.size());
    return bolap_estimator_create(ctx, p.data(), musd, m, out);
}

extern "C" int bolap_estimator_hints(bolap_estimator* est, const bolap_est_column* cols,
                                     int32_t n_cols, bolap_star_args* plans, int32_t n_plans) {
    BOLAP_ARG_CHECK(est && n_cols >= 0 && n_plans >= 0 && (cols || n_cols == 0) &&
                        (plans || n_plans == 0),
                    "bolap_estimator_hints: bad arguments");
    if (n_plans == 0) return BOLAP_OK;
    bolap_ctx* ctx = est->ctx;
    std::vector<ColStats> cs(static_cast<size_t>(n_cols));
    for (int32_t i = 0; i < n_cols; ++i) {
        ColStats& s = cs[size_t(i)];
        s.c = &cols[i];
        if (cols[i].kind == BOLAP_EST_FK || cols[i].kind == BOLAP_EST_GROUP) {
            int64_t h[64] = {};
            for (int r = 0; r < BOLAP_HLL_REGS; ++r) ++h[std::min<int>(cols[i].hll[r], 63)];
            for (int b = 0; b < 64; ++b) s.hist[b] = float(h[b]) / float(BOLAP_HLL_REGS);
            s.distinct = hll_estimate(cols[i].hll);
        }
    }
    for (int32_t i = 0; i < n_plans; ++i)
        BOLAP_ARG_CHECK(plans[i].n_dims >= 1 && plans[i].n_dims <= BOLAP_STAR_MAX_DIMS,
                        "bolap_estimator_hints: plan %d has %d dimensions", i, plans[i].n_dims);
    std::vector<int64_t> counts(size_t(n_plans) * BOLAP_SAMPLE_COUNTS);
    BOLAP_TRY(bolap_star_sample_counts(ctx, plans, n_plans, BOLAP_EST_SAMPLE_SEED,
                                       counts.data()));
    std::vector<int64_t> first(static_cast<size_t>(n_plans) + 1, 0);
    for (int32_t i = 0; i < n_plans; ++i) first[size_t(i) + 1] = first[size_t(i)] + plans[i].n_dims;
    const int64_t rows = first.back();
    std::vector<float> x(size_t(rows) * kFeat), y(size_t(rows) * BOLAP_MLP_OUT);
    for (int32_t i = 0; i < n_plans; ++i)
        for (int r = 0; r < plans[i].n_dims; ++r)
            BOLAP_TRY(features(ctx, cs, plans[i], r, &counts[size_t(i) * BOLAP_SAMPLE_COUNTS],
                               &x[size_t(first[size_t(i)] + r) * kFeat]));
    float* xd = nullptr;
    float* yd = nullptr;
    BOLAP_HIP_TRY(hipMalloc(&xd, x.size() * 4));
    if (hipMalloc(&yd, y.size() * 4) != hipSuccess) {
        (void)hipFree(xd);
        set_error("bolap_estimator_hints: out of device memory");
        return BOLAP_ERR_HIP;
    }
    int st = hipMemcpyAsync(xd, x.data(), x.size() * 4, hipMemcpyHostToDevice, ctx->stream) ==
                     hipSuccess
                 ? BOLAP_OK
                 : BOLAP_ERR_HIP;
    if (st == BOLAP_OK) st = bolap_mlp_predict(est->mlp, xd, rows, yd);
    if (st == BOLAP_OK &&
        (hipMemcpyAsync(y.data(), yd, y.size() * 4, hipMemcpyDeviceToHost, ctx->stream) !=
             hipSuccess ||
         hipStreamSynchronize(ctx->stream) != hipSuccess))
        st = BOLAP_ERR_HIP;
    (void)hipFree(xd);
    (void)hipFree(yd);
    if (st != BOLAP_OK) {
        if (st == BOLAP_ERR_HIP) set_error("bolap_estimator_hints: HIP error in the prediction");
        return st;
    }
    for (int32_t i = 0; i < n_plans; ++i) {
        bolap_star_args& a = plans[i];
        const int k = a.n_dims;
        const int64_t* c = &counts[size_t(i) * BOLAP_SAMPLE_COUNTS];
        const ColStats* g = find_col(cs, a.group_key);
        const double bound = double(std::min<int64_t>(a.fact_rows, g->c->span));
        const double nf = double(a.fact_rows);
        double lj = 0, lg = 0;
        for (int r = 0; r < k; ++r) {
            float yy[BOLAP_MLP_OUT];
            for (int o = 0; o < BOLAP_MLP_OUT; ++o) {
                const float t = y[size_t(first[size_t(i)] + r) * BOLAP_MLP_OUT + o] * est->sd[o];
                yy[o] = t + est->mu[o];
            }
            // slot 0 of rotation r is dimension r
            const double rows0 = double(a.dims[r].rows);
            const double pass = std::max(0.0, std::pow(10.0, double(yy[0])) * (rows0 + 1) - 1);
            const double join = std::max(0.0, std::pow(10.0, double(yy[1])) * (nf + 1) - 1);
            const double grp = std::max(0.0, std::pow(10.0, double(yy[2])) * (bound + 1) - 1);
            a.dims[r].pass_hint = std::max<int64_t>(1, int64_t(std::ceil(pass * est->margin_dim)));
            lj += std::log1p(join);
            lg += std::log1p(grp);
        }
        const double matches = std::exp(lj / k) - 1, groups = std::exp(lg / k) - 1;
        double gh = groups * est->margin_group;
        gh = std::min({gh, double(g->c->span), std::max(matches * est->margin_group, 1.0)});
        gh = std::max({gh, double(c[7]), 1.0});
        a.group_hint = int64_t(std::ceil(gh));
    }
    return BOLAP_OK;
}
