/*
 * star.c — CPU restatement of the star join + group-by and of the
 * estimator's statistics (HyperLogLog registers, plan sample counts), plus
 * the generic seeded column generators.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py as the checker, never by the product.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* --- generic columns (mirror of the device generators) ------------------ */
void og_uniform_i32(uint64_t seed, uint32_t stream, int64_t rows, int32_t lo, int32_t hi,
                    int32_t* out) {
    for (int64_t r = 0; r < rows; ++r)
        out[r] = (int32_t)og_uniform(og_hash(seed, stream, (uint64_t)r), lo, hi);
}

void og_uniform_u8(uint64_t seed, uint32_t stream, int64_t rows, int32_t lo, int32_t hi,
                   uint8_t* out) {
    for (int64_t r = 0; r < rows; ++r)
        out[r] = (uint8_t)og_uniform(og_hash(seed, stream, (uint64_t)r), lo, hi);
}

/* Inverse-CDF table lookup with 16-bit linear interpolation, then the
   multiplicative scatter of ranks over 1..n_keys. */
void og_zipf_i32(uint64_t seed, uint32_t stream, int64_t rows, const int32_t* q,
                 int64_t n_keys, uint64_t mul, int32_t* out) {
    for (int64_t r = 0; r < rows; ++r) {
        const uint32_t u = (uint32_t)(og_hash(seed, stream, (uint64_t)r) >> 32);
        const uint32_t i = u >> 16, f = u & 0xFFFFu;
        const int64_t a = q[i], b = q[i + 1];
        const int64_t rank = a + (((b - a) * (int64_t)f) >> 16);
        out[r] = (int32_t)((uint64_t)(rank - 1) * mul % (uint64_t)n_keys) + 1;
    }
}

void og_price_disc(uint64_t seed, uint32_t stream, int64_t rows, int64_t* price,
                   int64_t* disc) {
    for (int64_t r = 0; r < rows; ++r) {
        const int64_t qn = og_uniform(og_hash(seed, stream, (uint64_t)r), 1, 50);
        const int64_t unit = og_uniform(og_hash(seed, stream + 1, (uint64_t)r), 90000, 200000);
        price[r] = qn * unit / 100;
        disc[r] = og_uniform(og_hash(seed, stream + 2, (uint64_t)r), 0, 10);
    }
}

/* --- HyperLogLog (p = 12) ------------------------------------------------- */
void oracle_hll_i32(const int32_t* col, int64_t n, uint8_t* regs) {
    memset(regs, 0, 4096);
    for (int64_t r = 0; r < n; ++r) {
        const uint64_t h = og_mix64((uint64_t)(uint32_t)col[r] + 0x9E3779B97F4A7C15ULL);
        const uint32_t idx = (uint32_t)(h >> 52);
        const uint64_t w = h << 12;
        const uint8_t rank = w ? (uint8_t)(__builtin_clzll(w) + 1) : 53;
        if (rank > regs[idx]) regs[idx] = rank;
    }
}

/* --- star join + group-by -------------------------------------------------
 * Restates the CEB star pipelines: each dimension filter marks its passing
 * keys in a byte mask over the key domain (bespoke_ceb/query_q1a.cpp:69-152),
 * one scan of the fact table tests every foreign key against the masks
 * (query_q1a.cpp:241-306, 346-467), and the passing rows accumulate into a
 * dense array indexed by the group key, the reference's count-array layout
 * (query_q1a.cpp:505-596; query_q10a.cpp:843-856 keeps a dense matrix for
 * small domains).  Groups come out in ascending key order.
 */
static int64_t load_any(const void* col, int32_t dtype, int64_t row) {
    switch (dtype) {
        case 0: return ((const uint8_t*)col)[row];
        case 1: return ((const int16_t*)col)[row];
        case 2: return ((const int32_t*)col)[row];
        default: return ((const int64_t*)col)[row];
    }
}

int64_t oracle_star_groupby(int32_t n_dims, const oracle_star_dim* dims, int64_t fact_rows,
                            const int32_t* group_key, int32_t group_min, int32_t group_max,
                            const void* price, int32_t price_dtype, const void* disc,
                            int32_t disc_dtype, int32_t* keys_out, int64_t* rev_out,
                            int64_t max_out, oracle_star_stats* stats) {
    memset(stats, 0, sizeof(*stats));
    uint8_t* mask[3] = {0, 0, 0};
    for (int j = 0; j < n_dims; ++j) {
        const oracle_star_dim* d = &dims[j];
        const int64_t span = (int64_t)d->key_max - d->key_min + 1;
        mask[j] = (uint8_t*)calloc((size_t)span, 1);
        for (int64_t r = 0; r < d->rows; ++r) {
            int ok = 1;
            for (int p = 0; p < d->n_preds && ok; ++p) ok = oracle_pred_row(&d->preds[p], r);
            if (!ok) continue;
            stats->dim_pass[j] += 1;
            mask[j][d->key[r] - d->key_min] = 1;
        }
    }
    const int64_t gspan = (int64_t)group_max - group_min + 1;
    uint64_t* rev = (uint64_t*)calloc((size_t)gspan, 8);
    uint8_t* hit = (uint8_t*)calloc((size_t)gspan, 1);
    for (int64_t r = 0; r < fact_rows; ++r) {
        int ok = 1;
        for (int j = 0; j < n_dims && ok; ++j) {
            const oracle_star_dim* d = &dims[j];
            const int64_t v = (int64_t)d->fact_fk[r] - d->key_min;
            /* planted mistake 7: the last dimension's filter is ignored */
            if (oracle_mutation() == 7 && j == n_dims - 1) continue;
            ok = v >= 0 && v <= (int64_t)d->key_max - d->key_min && mask[j][v];
        }
        if (!ok) continue;
        stats->matches += 1;
        const int64_t g = (int64_t)group_key[r] - group_min;
        const int64_t dv = load_any(disc, disc_dtype, r);
        /* planted mistake 6: discount added instead of subtracted */
        const int64_t rv = load_any(price, price_dtype, r) * (oracle_mutation() == 6 ? 100 + dv
                                                                                   : 100 - dv);
        rev[g] += (uint64_t)rv;
        hit[g] = 1;
    }
    int64_t n = 0;
    uint64_t total = 0;
    for (int64_t g = 0; g < gspan; ++g) {
        if (!hit[g]) continue;
        if (n < max_out) {
            keys_out[n] = (int32_t)(g + group_min);
            rev_out[n] = (int64_t)rev[g];
        }
        total += rev[g];
        ++n;
    }
    stats->groups = n;
    stats->revenue = (int64_t)total;
    for (int j = 0; j < n_dims; ++j) free(mask[j]);
    free(rev);
    free(hit);
    return n;
}

/* --- sample counts of star plans (same sample rows as the device) ------- */
static int64_t sample_row(uint64_t seed, uint32_t stream, int i, int64_t rows) {
    return (int64_t)(((og_hash(seed, stream, (uint64_t)i) >> 32) * (uint64_t)rows) >> 32);
}

static int cmp_u32(const void* a, const void* b) {
    const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : x > y;
}

void oracle_star_sample_counts(int32_t n_dims, const oracle_star_dim* dims, int64_t fact_rows,
                               const int32_t* group_key, uint64_t seed, int64_t* counts) {
    memset(counts, 0, 16 * sizeof(int64_t));
    uint32_t keys[4096];
    int nk = 0;
    for (int i = 0; i < 4096; ++i) {
        for (int j = 0; j < n_dims; ++j) {
            const oracle_star_dim* d = &dims[j];
            if (d->rows <= 0) continue;
            const int64_t r = sample_row(seed, 0x53440000u, i, d->rows);
            int ok = 1;
            for (int p = 0; p < d->n_preds && ok; ++p) ok = oracle_pred_row(&d->preds[p], r);
            counts[j] += ok;
        }
        if (fact_rows <= 0) continue;
        const int64_t fr = sample_row(seed, 0x53460000u, i, fact_rows);
        int all = 1;
        for (int j = 0; j < n_dims; ++j) {
            const oracle_star_dim* d = &dims[j];
            const int64_t row = (int64_t)d->fact_fk[fr] - d->key_min;
            int ok = row >= 0 && row < d->rows;
            for (int p = 0; p < d->n_preds && ok; ++p) ok = oracle_pred_row(&d->preds[p], row);
            counts[3 + j] += ok;
            all = all && ok;
        }
        if (!all) continue;
        counts[6] += 1;
        keys[nk++] = (uint32_t)group_key[fr];
    }
    qsort(keys, (size_t)nk, sizeof(uint32_t), cmp_u32);
    for (int i = 0; i < nk;) {
        int j = i;
        while (j < nk && keys[j] == keys[i]) ++j;
        counts[7] += 1;
        counts[8] += (j - i) == 1;
        counts[9] += (j - i) == 2;
        i = j;
    }
}
