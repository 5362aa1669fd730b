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
