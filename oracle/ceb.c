/*
 * ceb.c — CPU restatement of the CEB star queries (factorised COUNT and
 * MIN(production_year)).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows bespoke_ceb/query_q1a.cpp: byte masks from the dimension
 * predicates (:69-152), a title scan into candidate movies (:241-306),
 * dense per-movie count arrays filled by one scan of each fact table
 * (:346-467, :505-596), and the sum over candidate movies of the product of
 * their counts (:602-621); MIN over the same joined movies as
 * query_q10a.cpp:334 ff.  Counts are summed in unsigned 128-bit arithmetic.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

static int mask_has(const uint8_t* m, int64_t n, int64_t v) {
    if (!m) return 1;
    return v >= 0 && v < n && m[v];
}

void oracle_ceb_run(const oracle_ceb_db* db, const oracle_ceb_query* q, uint64_t* out) {
    const int64_t nt = db->title_rows;
    uint8_t* tmask = (uint8_t*)calloc((size_t)nt + 1, 1);
    int64_t title_pass = 0;
    for (int64_t r = 0; r < nt; ++r) {
        const int32_t y = db->t_year[r], k = db->t_kind[r];
        const int ok = y >= q->year_lo && y <= q->year_hi && k >= 0 && k < 32 &&
                       ((q->kind_mask >> k) & 1u) &&
                       mask_has(q->title_mask, q->title_n, db->t_code[r]);
        tmask[r + 1] = (uint8_t)ok;  /* indexed by movie id = row + 1 */
        title_pass += ok;
    }
    uint32_t* cnt[3] = {0, 0, 0};
    for (int t = 0; t < 3; ++t) {
        if (!((q->tables >> t) & 1)) continue;
        cnt[t] = (uint32_t*)calloc((size_t)nt + 1, sizeof(uint32_t));
        const int64_t rows = t == 0 ? db->ci_rows : t == 1 ? db->mi_rows : db->mk_rows;
        const int32_t* mv = t == 0 ? db->ci_movie : t == 1 ? db->mi_movie : db->mk_movie;
        for (int64_t r = 0; r < rows; ++r) {
            const int32_t m = mv[r];
            if (m < 1 || m > nt || !tmask[m]) continue;
            int ok;
            if (t == 0) {
                const int32_t role = db->ci_role[r];
                ok = role >= 0 && role < 32 && ((q->role_mask >> role) & 1u);
            } else if (t == 1) {
                ok = mask_has(q->type_mask, q->type_n, db->mi_type[r]) &&
                     mask_has(q->info_mask, q->info_n, db->mi_info[r]);
            } else {
                ok = mask_has(q->kw_mask, q->kw_n, db->mk_keyword[r]);
            }
            if (ok) cnt[t][m] += 1;
        }
    }
    unsigned __int128 total = 0;
    int64_t min_year = INT64_MAX, joined = 0;
    for (int64_t m = 1; m <= nt; ++m) {
        if (!tmask[m]) continue;
        unsigned __int128 prod = 1;
        int all = 1;
        for (int t = 0; t < 3 && all; ++t) {
            if (!cnt[t]) continue;
            if (cnt[t][m] == 0) all = 0;
            prod *= cnt[t][m];
        }
        if (!all) continue;
        /* planted mistake 8: the last table's count is left out of the product */
        if (oracle_mutation() == 8 && cnt[2]) prod /= cnt[2][m];
        total += prod;
        ++joined;
        if (db->t_year[m - 1] < min_year) min_year = db->t_year[m - 1];
    }
    out[0] = (uint64_t)total;
    out[1] = (uint64_t)(total >> 64);
    out[2] = (uint64_t)min_year;
    out[3] = (uint64_t)title_pass;
    out[4] = (uint64_t)joined;
    for (int t = 0; t < 3; ++t) free(cnt[t]);
    free(tmask);
}

/* ---- grouped MIN/MAX (bespoke_ceb/query_q10a.cpp:838-1107) -------------- */

typedef struct {
    uint64_t key;
    int32_t has, val;
} mm_row;

static int mm_cmp(const void* a, const void* b) {
    const uint64_t x = ((const mm_row*)a)->key, y = ((const mm_row*)b)->key;
    return x < y ? -1 : x > y;
}

/* One aggregate slot, updated as query_q10a.cpp:880-900: created {has, v, v}
   or {0, 0, 0} by its first row; a later row with a value sets or widens it. */
static void mm_fold(int32_t* used, int32_t* has, int32_t* mn, int32_t* mx, int32_t h, int32_t v) {
    if (oracle_mutation() == 14 && !h) { h = 1; v = 0; }  /* planted: null year taken as 0 */
    if (!*used) {
        *used = 1;
        *has = h;
        *mn = h ? v : 0;
        *mx = h ? v : 0;
    } else if (h) {
        if (*has) {
            if (v < *mn) *mn = v;
            if (v > *mx) *mx = v;
        } else {
            *has = 1;
            *mn = v;
            *mx = v;
        }
    }
}

int64_t oracle_group_minmax(int64_t pairs, const int32_t* pg, const int32_t* pi, int64_t items,
                            const int32_t* off, const int32_t* cnt, const uint8_t* hasv,
                            const int32_t* val, int64_t values, const int32_t* rv,
                            int32_t groups, int32_t keys, int64_t dense_max, int32_t* out,
                            int64_t max_rows) {
    if (dense_max <= 0) dense_max = 5000000;  /* query_q10a.cpp:843 */
    const int64_t slots = (int64_t)groups * keys;
    /* bounds (the reference indexes without checks) */
    int64_t expanded = 0;
    for (int64_t p = 0; p < pairs; ++p) {
        const int32_t g = pg[p];
        if (g < 0) continue;                 /* no name group: dropped (:1078-1080) */
        const int32_t it = pi[p];
        if (g >= groups || it < 0 || it >= items) return -1;
        if (off[it] < 0 || cnt[it] < 0 || (int64_t)off[it] + cnt[it] > values) return -1;
        for (int64_t i = off[it]; i < (int64_t)off[it] + cnt[it]; ++i)
            if (rv[i] < 0 || rv[i] >= keys) return -1;
        expanded += cnt[it];
    }
    int64_t n = 0;
    if (slots <= dense_max) {                /* dense matrix (:847-853) */
        int32_t* used = (int32_t*)calloc((size_t)slots, 4);
        int32_t* has = (int32_t*)calloc((size_t)slots, 4);
        int32_t* mn = (int32_t*)calloc((size_t)slots, 4);
        int32_t* mx = (int32_t*)calloc((size_t)slots, 4);
        for (int64_t p = 0; p < pairs; ++p) {
            const int32_t g = pg[p];
            if (g < 0) continue;
            const int32_t it = pi[p];
            const int32_t h = hasv ? (hasv[it] != 0) : 1;
            for (int64_t i = off[it]; i < (int64_t)off[it] + cnt[it]; ++i) {
                const int64_t s = (int64_t)g * keys + rv[i];
                mm_fold(used + s, has + s, mn + s, mx + s, h, val[it]);
            }
        }
        for (int64_t s = 0; s < slots; ++s) {
            if (!used[s]) continue;
            if (n >= max_rows) { n = -2; break; }
            int32_t* o = out + 5 * n++;
            o[0] = (int32_t)(s / keys);
            o[1] = (int32_t)(s % keys);
            o[2] = has[s];
            o[3] = mn[s];
            o[4] = mx[s];
        }
        free(used); free(has); free(mn); free(mx);
        return n;
    }
    /* hashed in the reference (:905-921); here the expanded rows sorted by key */
    mm_row* rows = (mm_row*)malloc((size_t)(expanded > 0 ? expanded : 1) * sizeof(mm_row));
    int64_t m = 0;
    for (int64_t p = 0; p < pairs; ++p) {
        const int32_t g = pg[p];
        if (g < 0) continue;
        const int32_t it = pi[p];
        const int32_t h = hasv ? (hasv[it] != 0) : 1;
        for (int64_t i = off[it]; i < (int64_t)off[it] + cnt[it]; ++i) {
            rows[m].key = ((uint64_t)(uint32_t)g << 32) | (uint32_t)rv[i];
            rows[m].has = h;
            rows[m].val = val[it];
            ++m;
        }
    }
    qsort(rows, (size_t)m, sizeof(mm_row), mm_cmp);
    for (int64_t i = 0; i < m;) {
        int32_t used = 0, has = 0, mn = 0, mx = 0;
        int64_t j = i;
        for (; j < m && rows[j].key == rows[i].key; ++j)
            mm_fold(&used, &has, &mn, &mx, rows[j].has, rows[j].val);
        if (n >= max_rows) { n = -2; break; }
        int32_t* o = out + 5 * n++;
        o[0] = (int32_t)(rows[i].key >> 32);
        o[1] = (int32_t)(uint32_t)rows[i].key;
        o[2] = has;
        o[3] = mn;
        o[4] = mx;
        i = j;
    }
    free(rows);
    return n;
}

/* ---- table sort order (bespoke_ceb/builder_impl.cpp:265-340) ------------ */

static const int32_t* so_p;
static const int32_t* so_s;

static int so_cmp(const void* a, const void* b) {
    const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    if (so_p[x] != so_p[y]) return so_p[x] < so_p[y] ? -1 : 1;
    if (so_s && so_s[x] != so_s[y]) return so_s[x] < so_s[y] ? -1 : 1;
    return x < y ? -1 : x > y;  /* input order on ties */
}

void oracle_sort_order_i32(const int32_t* primary, const int32_t* secondary, int64_t rows,
                           uint32_t* order) {
    for (int64_t i = 0; i < rows; ++i) order[i] = (uint32_t)i;
    so_p = primary;
    so_s = secondary;
    qsort(order, (size_t)rows, 4, so_cmp);
}
