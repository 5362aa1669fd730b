// ceb_q10a.hip — CEB query 10a on the GPU, and ILIKE over a dictionary in HBM.
//
//   SELECT n.name, mi1.info, MIN(t.production_year), MAX(t.production_year)
//   ... GROUP BY mi1.info, n.name        (bespoke_ceb/queries.txt:451-471)
//
// The reference (bespoke_ceb/query_q10a.cpp:334-1120) runs five passes over
// the builder's tables; each is a kernel (or two) here, with the reference's
// row checks (nulls, ids inside the arrays it sizes from row counts, the
// host-built masks):
//   title       one lane per row: the last row of a title id in row order
//               decides its year (:460-500); present / has-year / year pack
//               into one u64 per id, taken with atomicMax over the row number.
//   movie_info  passing rows append their dense info index to one list in
//               row order (:533-590); a movie's range starts at its first
//               appended value and spans its count, as there.  The order
//               comes from a chunked compaction: per-workgroup counts, a
//               scan, then each workgroup writes its chunk from its offset.
//   cast_info   passing rows become (movie, person) join rows in row order
//               (:612-740, both of the reference's paths visit rows in row
//               order) and mark their person as a candidate.
//   name        candidate persons whose name matches take the name's string
//               id; the last such row of a person id wins (:756-808).  The
//               name groups are the distinct string ids so taken, numbered
//               by rank in a bitmap over the pool (any numbering gives the
//               same rows, which carry string ids).
//   result      the join rows expand through their movie's info range into
//               group_minmax.hip's aggregation: a dense matrix when
//               groups x infos <= 5,000,000 slots (:843-856), rows in the
//               order the reference first reaches them (:866-872); else
//               the expanded rows sorted and reduced run by run, rows in
//               (name group, info) order (the reference's unordered_map
//               leaves its order open).
#include <algorithm>
#include <cstring>

#include "common.hpp"
#include "prims.hpp"

namespace bolap {
namespace {

constexpr int kT = 256;
constexpr int kLikeMax = 256;

unsigned grid_of(bolap_ctx* ctx, int64_t n, int per_cu = 8) {
    int64_t b = (n + kT - 1) / kT;
    const int64_t cap = int64_t(ctx->num_cus) * per_cu;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

__global__ __launch_bounds__(kT) void fill_u32_k(uint32_t* __restrict__ p, int64_t n,
                                                 uint32_t v) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        p[i] = v;
}

// Fills `bytes` (rounded up to whole words) with a kernel.
int fill_dev(bolap_ctx* ctx, void* p, size_t bytes, uint32_t v = 0) {
    const int64_t words = int64_t((bytes + 3) / 4);
    if (words == 0) return BOLAP_OK;
    hipLaunchKernelGGL(fill_u32_k, dim3(grid_of(ctx, words)), dim3(kT), 0, ctx->stream,
                       static_cast<uint32_t*>(p), words, v);
    return launch_check("fill_u32_k");
}

// ---------------------------------------------------------------------------
// ILIKE over a dictionary: one lane per string, the reference's forms
// (query_q10a.cpp:135-209) on ASCII-lower-cased text.
// ---------------------------------------------------------------------------
struct LikeK {
    int32_t kind;  // LikeKind
    int32_t n;     // bytes of s
    char s[kLikeMax];  // the token (literal forms) or the lower-cased pattern (general)
};

BOLAP_DEV char lower_c(char c) { return (c >= 'A' && c <= 'Z') ? char(c + 32) : c; }

BOLAP_DEV bool eq_at(const char* t, int64_t at, const LikeK& P) {
    for (int32_t j = 0; j < P.n; ++j)
        if (lower_c(t[at + j]) != P.s[j]) return false;
    return true;
}

// like_match (:43-71): '_' or an equal character advances both sides; '%'
// records a restart point; a mismatch returns to the last '%' one text
// character further.
BOLAP_DEV bool like_general(const LikeK& P, const char* t, int64_t tn) {
    int64_t i = 0, j = 0, star = -1, resume = 0;
    while (j < tn) {
        const char c = lower_c(t[j]);
        if (i < P.n && (P.s[i] == '_' || P.s[i] == c)) {
            ++i;
            ++j;
        } else if (i < P.n && P.s[i] == '%') {
            star = i++;
            resume = j;
        } else if (star >= 0) {
            i = star + 1;
            j = ++resume;
        } else {
            return false;
        }
    }
    while (i < P.n && P.s[i] == '%') ++i;
    return i == P.n;
}

BOLAP_DEV bool like_one(const LikeK& P, const char* t, int64_t tn) {
    switch (LikeKind(P.kind)) {
        case LikeKind::kAll: return true;
        case LikeKind::kExact: return tn == P.n && eq_at(t, 0, P);
        case LikeKind::kPrefix: return tn >= P.n && eq_at(t, 0, P);
        case LikeKind::kSuffix: return tn >= P.n && eq_at(t, tn - P.n, P);
        case LikeKind::kContains:
            for (int64_t at = 0; at + P.n <= tn; ++at)
                if (eq_at(t, at, P)) return true;
            return P.n == 0;
        case LikeKind::kGeneral: return like_general(P, t, tn);
    }
    return false;
}

__global__ __launch_bounds__(kT) void ilike_k(LikeK P, const char* __restrict__ chars,
                                              const int64_t* __restrict__ off, int64_t n,
                                              uint8_t* __restrict__ mask) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const int64_t a = off[i], b = off[i + 1];
        mask[i] = (a <= b && like_one(P, chars + a, b - a)) ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// 10a passes
// ---------------------------------------------------------------------------
BOLAP_DEV bool is_null(const bolap_ceb_col& c, int64_t i) { return c.null && c.null[i]; }

// A value inside [0, n) and set in mask.
BOLAP_DEV bool in_mask(int32_t v, const uint8_t* mask, int64_t n) {
    return v >= 0 && int64_t(v) < n && mask[v];
}

constexpr uint64_t kHasYear = uint64_t(1) << 32;

// title (:460-500): tw[id] = (row + 1) << 33 | has_year << 32 | year of the
// last passing row of the id.
__global__ __launch_bounds__(kT) void q10_title_k(bolap_ceb_q10a_args a, int64_t t1,
                                                  unsigned long long* __restrict__ tw) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < a.title_rows;
         i += int64_t(gridDim.x) * kT) {
        if (is_null(a.t_id, i) || is_null(a.t_kind_id, i)) continue;
        if (!in_mask(a.t_kind_id.v[i], a.kind_ok, a.kind_n)) continue;
        const int32_t id = a.t_id.v[i];
        if (id < 0 || int64_t(id) >= t1) continue;
        const bool has = !is_null(a.t_year, i);
        const uint64_t w = (uint64_t(i + 1) << 33) | (has ? kHasYear : 0) |
                           (has ? uint64_t(uint32_t(a.t_year.v[i])) : 0);
        atomicMax(tw + id, (unsigned long long)w);
    }
}

// movie_info row filter (:533-590): the movie and the row's dense info.
struct MiPass {
    bolap_ceb_q10a_args a;
    int64_t t1;
    const unsigned long long* tw;
    BOLAP_DEV bool operator()(int64_t i, int32_t& movie, int32_t& dense) const {
        if (is_null(a.mi_type, i) || is_null(a.mi_info, i)) return false;
        if (!in_mask(a.mi_type.v[i], a.type_ok, a.type_n)) return false;
        const int32_t info = a.mi_info.v[i];
        if (info < 0 || int64_t(info) >= a.pool_n) return false;
        dense = a.info_dense[info];
        if (dense < 0 || is_null(a.mi_movie, i)) return false;
        movie = a.mi_movie.v[i];
        return movie >= 0 && int64_t(movie) < t1 && tw[movie] != 0;
    }
};

// cast_info row filter (:612-740): the movie has a title and info values,
// the role passes, the person id lies inside the name arrays.
struct CiPass {
    bolap_ceb_q10a_args a;
    int64_t t1, n1;
    const unsigned long long* tw;
    const uint32_t* mi_first;
    BOLAP_DEV bool operator()(int64_t i, int32_t& movie, int32_t& person) const {
        if (is_null(a.ci_movie, i)) return false;
        movie = a.ci_movie.v[i];
        if (movie < 0 || int64_t(movie) >= t1 || tw[movie] == 0 || mi_first[movie] == ~0u)
            return false;
        if (is_null(a.ci_role, i) || !in_mask(a.ci_role.v[i], a.role_ok, a.role_n)) return false;
        if (is_null(a.ci_person, i)) return false;
        person = a.ci_person.v[i];
        return person >= 0 && int64_t(person) < n1;
    }
};

// Chunked compaction in row order: workgroup b owns rows [b * chunk, ...).
template <class P>
__global__ __launch_bounds__(kT) void oc_count_k(P pass, int64_t rows, int64_t chunk,
                                                 int64_t* __restrict__ cnt) {
    __shared__ uint64_t s_w[kT / kWave];
    const int64_t c0 = int64_t(blockIdx.x) * chunk;
    const int64_t c1 = c0 + chunk < rows ? c0 + chunk : rows;
    uint64_t n = 0;
    for (int64_t i = c0 + threadIdx.x; i < c1; i += kT) {
        int32_t x, y;
        n += pass(i, x, y) ? 1u : 0u;
    }
    n = wave_sum(n);
    if ((threadIdx.x & (kWave - 1)) == 0) s_w[threadIdx.x / kWave] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < kT / kWave; ++w) t += s_w[w];
        cnt[blockIdx.x] = int64_t(t);
    }
}

// The passing rows of each chunk in row order from the chunk's offset;
// emit(row, position, x, y) with the filter's two outputs.
template <class P, class E>
__global__ __launch_bounds__(kT) void oc_write_k(P pass, E emit, int64_t rows, int64_t chunk,
                                                 const int64_t* __restrict__ off) {
    __shared__ uint32_t s_w[kT / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t c0 = int64_t(blockIdx.x) * chunk;
    const int64_t c1 = c0 + chunk < rows ? c0 + chunk : rows;
    int64_t base = off[blockIdx.x];
    // the trip count is the same for every thread of the workgroup
    for (int64_t b0 = c0; b0 < c1; b0 += kT) {
        const int64_t i = b0 + threadIdx.x;
        int32_t x = 0, y = 0;
        const bool ok = i < c1 && pass(i, x, y);
        const uint64_t bal = __ballot(ok);
        if (lane == 0) s_w[wave] = uint32_t(__popcll(bal));
        __syncthreads();
        uint32_t before = 0, tile = 0;
#pragma unroll
        for (int w = 0; w < kT / kWave; ++w) {
            before += w < wave ? s_w[w] : 0u;
            tile += s_w[w];
        }
        __syncthreads();  // s_w is rewritten by the next tile
        if (ok) emit(i, base + before + __popcll(bal & ((1ull << lane) - 1ull)), x, y);
        base += tile;
    }
}

struct MiEmit {
    int32_t* values;
    uint32_t* first;
    uint32_t* count;
    BOLAP_DEV void operator()(int64_t, int64_t pos, int32_t movie, int32_t dense) const {
        values[pos] = dense;
        atomicMin(first + movie, uint32_t(pos));
        atomicAdd(count + movie, 1u);
    }
};

struct CiEmit {
    int32_t* movie;
    int32_t* person;
    uint8_t* cand;
    BOLAP_DEV void operator()(int64_t, int64_t pos, int32_t m, int32_t p) const {
        movie[pos] = m;
        person[pos] = p;
        cand[p] = 1;
    }
};

// name (:756-808): pw[person] = (row + 1) << 32 | name string id of the last
// passing row; mark[string] = 1.
__global__ __launch_bounds__(kT) void q10_name_k(bolap_ceb_q10a_args a, int64_t n1,
                                                 const uint8_t* __restrict__ cand,
                                                 unsigned long long* __restrict__ pw,
                                                 uint8_t* __restrict__ mark) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < a.name_rows;
         i += int64_t(gridDim.x) * kT) {
        if (is_null(a.n_id, i)) continue;
        const int32_t person = a.n_id.v[i];
        if (person < 0 || int64_t(person) >= n1 || !cand[person]) continue;
        if (is_null(a.n_name, i)) continue;
        const int32_t s = a.n_name.v[i];
        if (s < 0 || int64_t(s) >= a.pool_n || (a.name_ok && !a.name_ok[s])) continue;
        mark[s] = 1;
        atomicMax(pw + person, (unsigned long long)((uint64_t(i + 1) << 32) | uint32_t(s)));
    }
}

// Mark bytes -> bitmap words and per-word popcounts.
__global__ __launch_bounds__(kT) void q10_bits_k(const uint8_t* __restrict__ mark, int64_t n,
                                                 uint32_t* __restrict__ bits,
                                                 int64_t* __restrict__ cnt) {
    const int64_t words = (n + 31) / 32;
    for (int64_t w = int64_t(blockIdx.x) * kT + threadIdx.x; w < words;
         w += int64_t(gridDim.x) * kT) {
        uint32_t b = 0;
        for (int j = 0; j < 32; ++j) {
            const int64_t s = w * 32 + j;
            if (s < n && mark[s]) b |= 1u << j;
        }
        bits[w] = b;
        cnt[w] = __popc(b);
    }
}

BOLAP_DEV int32_t rank_of(const uint32_t* bits, const int64_t* cnt, int64_t s) {
    return int32_t(cnt[s >> 5] + __popc(bits[s >> 5] & ((1u << (s & 31)) - 1u)));
}

// gstr[rank] = string id; sizes[0] = groups (one thread, after the scan).
__global__ __launch_bounds__(kT) void q10_gstr_k(const uint32_t* __restrict__ bits,
                                                 const int64_t* __restrict__ cnt, int64_t n,
                                                 int32_t* __restrict__ gstr) {
    for (int64_t s = int64_t(blockIdx.x) * kT + threadIdx.x; s < n;
         s += int64_t(gridDim.x) * kT)
        if ((bits[s >> 5] >> (s & 31)) & 1u) gstr[rank_of(bits, cnt, s)] = int32_t(s);
}

__global__ void q10_sizes_k(const uint32_t* __restrict__ bits, const int64_t* __restrict__ wcnt,
                            int64_t words, const int64_t* __restrict__ mi_off, int64_t mi_nb,
                            const int64_t* __restrict__ mi_cnt,
                            const int64_t* __restrict__ ci_off, int64_t ci_nb,
                            const int64_t* __restrict__ ci_cnt, int64_t* __restrict__ sizes) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    sizes[0] = words ? wcnt[words - 1] + __popc(bits[words - 1]) : 0;  // name groups
    sizes[1] = mi_nb ? mi_off[mi_nb - 1] + mi_cnt[mi_nb - 1] : 0;      // info values
    sizes[2] = ci_nb ? ci_off[ci_nb - 1] + ci_cnt[ci_nb - 1] : 0;      // join rows
}

// Per join row: its person's name group (rank of the name string) or -1.
__global__ __launch_bounds__(kT) void q10_pairs_k(const int32_t* __restrict__ person, int64_t n,
                                                  const unsigned long long* __restrict__ pw,
                                                  const uint32_t* __restrict__ bits,
                                                  const int64_t* __restrict__ cnt,
                                                  int32_t* __restrict__ group) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        const uint64_t w = pw[person[i]];
        group[i] = w ? rank_of(bits, cnt, int64_t(uint32_t(w))) : -1;
    }
}

// Per movie id: the info range, has-year and year.
__global__ __launch_bounds__(kT) void q10_items_k(const unsigned long long* __restrict__ tw,
                                                  uint32_t* __restrict__ first, int64_t t1,
                                                  uint8_t* __restrict__ has,
                                                  int32_t* __restrict__ year) {
    for (int64_t m = int64_t(blockIdx.x) * kT + threadIdx.x; m < t1;
         m += int64_t(gridDim.x) * kT) {
        const uint64_t w = tw[m];
        if (first[m] == ~0u) first[m] = 0;  // no values (count 0)
        has[m] = (w & kHasYear) ? 1 : 0;
        year[m] = int32_t(uint32_t(w));
    }
}

// Result rows: name group -> name string id, dense info -> info string id.
__global__ __launch_bounds__(kT) void q10_rows_k(bolap_minmax_row* __restrict__ rows, int64_t n,
                                                 const int32_t* __restrict__ gstr,
                                                 const int32_t* __restrict__ info_ids) {
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT) {
        rows[i].group = gstr[rows[i].group];
        rows[i].key = info_ids[rows[i].key];
    }
}

int chunks_of(bolap_ctx* ctx, int64_t rows, int64_t& chunk) {
    const int64_t nb = std::max<int64_t>(1, grid_of(ctx, rows));
    chunk = ((rows + nb - 1) / nb + kT - 1) / kT * kT;
    if (chunk < kT) chunk = kT;
    return int((rows + chunk - 1) / chunk);
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" int bolap_ilike_mask_device(bolap_ctx* ctx, const char* chars, const int64_t* offsets,
                                       int64_t n, const char* pattern, uint8_t* mask) {
    BOLAP_ARG_CHECK(ctx && pattern && n >= 0 && (n == 0 || (chars && offsets && mask)),
                    "bolap_ilike_mask_device: bad arguments");
    const LikePattern pt = like_classify(pattern);
    const std::string& s = pt.kind == LikeKind::kGeneral ? pt.p : pt.token;
    BOLAP_ARG_CHECK(s.size() < size_t(kLikeMax), "bolap_ilike_mask_device: pattern of %zu bytes "
                    "(at most %d)", s.size(), kLikeMax - 1);
    if (n == 0) return BOLAP_OK;
    LikeK P{};
    P.kind = int32_t(pt.kind);
    P.n = int32_t(s.size());
    std::memcpy(P.s, s.data(), s.size());
    hipLaunchKernelGGL(ilike_k, dim3(grid_of(ctx, n)), dim3(kT), 0, ctx->stream, P, chars,
                       offsets, n, mask);
    return launch_check("ilike_k");
}

extern "C" int bolap_ceb_q10a(bolap_ctx* ctx, const bolap_ceb_q10a_args* a,
                              bolap_minmax_row* rows_out, int64_t max_rows, int64_t* n_out,
                              bolap_minmax_stats* stats) {
    BOLAP_ARG_CHECK(ctx && a && n_out && (rows_out || max_rows == 0),
                    "bolap_ceb_q10a: null argument");
    BOLAP_ARG_CHECK(a->title_rows >= 0 && a->mi_rows >= 0 && a->ci_rows >= 0 &&
                        a->name_rows >= 0 && a->pool_n >= 0 && a->n_infos >= 0,
                    "bolap_ceb_q10a: negative size");
    BOLAP_ARG_CHECK(a->title_rows < (int64_t(1) << 31) && a->mi_rows < (int64_t(1) << 31) &&
                        a->ci_rows < (int64_t(1) << 31) && a->name_rows < (int64_t(1) << 31),
                    "bolap_ceb_q10a: table too large");
    BOLAP_ARG_CHECK(a->title_rows == 0 || (a->t_id.v && a->t_kind_id.v && a->t_year.v),
                    "bolap_ceb_q10a: null title column");
    BOLAP_ARG_CHECK(a->mi_rows == 0 || (a->mi_movie.v && a->mi_type.v && a->mi_info.v),
                    "bolap_ceb_q10a: null movie_info column");
    BOLAP_ARG_CHECK(a->ci_rows == 0 || (a->ci_movie.v && a->ci_person.v && a->ci_role.v),
                    "bolap_ceb_q10a: null cast_info column");
    BOLAP_ARG_CHECK(a->name_rows == 0 || (a->n_id.v && a->n_name.v),
                    "bolap_ceb_q10a: null name column");
    BOLAP_ARG_CHECK((a->kind_n == 0 || a->kind_ok) && (a->role_n == 0 || a->role_ok) &&
                        (a->type_n == 0 || a->type_ok),
                    "bolap_ceb_q10a: null mask");
    BOLAP_ARG_CHECK(a->pool_n == 0 || a->info_dense, "bolap_ceb_q10a: null info_dense");
    BOLAP_ARG_CHECK(a->n_infos == 0 || a->info_ids, "bolap_ceb_q10a: null info_ids");
    *n_out = 0;
    if (stats) *stats = bolap_minmax_stats{};
    if (a->n_infos == 0 || a->title_rows == 0 || a->mi_rows == 0 || a->ci_rows == 0 ||
        a->name_rows == 0)
        return BOLAP_OK;  // an empty stage ends the reference's run with no rows
    TraceScope ts_all(ctx, "q10a");
    const int64_t t1 = a->title_rows + 1, n1 = a->name_rows + 1, pool = a->pool_n;
    const int64_t words = (pool + 31) / 32;
    int64_t mi_chunk = 0, ci_chunk = 0;
    const int mi_nb = chunks_of(ctx, a->mi_rows, mi_chunk);
    const int ci_nb = chunks_of(ctx, a->ci_rows, ci_chunk);
    WsArena ws;
    const size_t o_tw = ws.plan(size_t(t1) * 8);
    const size_t o_cnt = ws.plan(size_t(t1) * 4);
    const size_t o_cand = ws.plan(size_t(n1));
    const size_t o_pw = ws.plan(size_t(n1) * 8);
    const size_t o_mark = ws.plan(size_t(std::max<int64_t>(pool, 1)));
    const size_t o_zero_end = ws.off;
    const size_t o_first = ws.plan(size_t(t1) * 4);
    const size_t o_has = ws.plan(size_t(t1));
    const size_t o_year = ws.plan(size_t(t1) * 4);
    const size_t o_vals = ws.plan(size_t(a->mi_rows) * 4);
    const size_t o_pm = ws.plan(size_t(a->ci_rows) * 4);
    const size_t o_pp = ws.plan(size_t(a->ci_rows) * 4);
    const size_t o_pg = ws.plan(size_t(a->ci_rows) * 4);
    const size_t o_mic = ws.plan(size_t(mi_nb) * 8), o_mio = ws.plan(size_t(mi_nb) * 8);
    const size_t o_cic = ws.plan(size_t(ci_nb) * 8), o_cio = ws.plan(size_t(ci_nb) * 8);
    const size_t o_bits = ws.plan(size_t(std::max<int64_t>(words, 1)) * 4);
    const size_t o_wcnt = ws.plan(size_t(std::max<int64_t>(words, 1)) * 8);
    const size_t o_gstr = ws.plan(size_t(std::max<int64_t>(pool, 1)) * 4);
    const size_t o_sizes = ws.plan(64);
    const size_t o_scan = ws.plan(scan_scratch_bytes(std::max<int64_t>(
        std::max<int64_t>(mi_nb, ci_nb), std::max<int64_t>(words, 1))));
    const size_t prep_end = ws.off;
    BOLAP_TRY(ws_reserve(ctx, prep_end));
    ws.base = static_cast<char*>(ctx->ws);
    BOLAP_TRY(fill_dev(ctx, ws.base, o_zero_end));
    BOLAP_TRY(fill_dev(ctx, ws.at<void>(o_first), size_t(t1) * 4, ~0u));
    auto* tw = ws.at<unsigned long long>(o_tw);
    auto* first = ws.at<uint32_t>(o_first);
    auto* count = ws.at<uint32_t>(o_cnt);
    auto* cand = ws.at<uint8_t>(o_cand);
    auto* pw = ws.at<unsigned long long>(o_pw);
    auto* mark = ws.at<uint8_t>(o_mark);
    auto* bits = ws.at<uint32_t>(o_bits);
    auto* wcnt = ws.at<int64_t>(o_wcnt);
    void* scan = ws.at<void>(o_scan);
    {
        TraceScope ts(ctx, "q10a_scans");
        hipLaunchKernelGGL(q10_title_k, dim3(grid_of(ctx, a->title_rows)), dim3(kT), 0,
                           ctx->stream, *a, t1, tw);
        const MiPass mp{*a, t1, tw};
        auto* mic = ws.at<int64_t>(o_mic);
        auto* mio = ws.at<int64_t>(o_mio);
        hipLaunchKernelGGL(oc_count_k<MiPass>, dim3(unsigned(mi_nb)), dim3(kT), 0, ctx->stream,
                           mp, a->mi_rows, mi_chunk, mic);
        BOLAP_HIP_TRY(hipMemcpyAsync(mio, mic, size_t(mi_nb) * 8, hipMemcpyDeviceToDevice,
                                     ctx->stream));
        BOLAP_TRY(exclusive_scan_i64(ctx, mio, mi_nb, scan));
        hipLaunchKernelGGL((oc_write_k<MiPass, MiEmit>), dim3(unsigned(mi_nb)), dim3(kT), 0,
                           ctx->stream, mp, MiEmit{ws.at<int32_t>(o_vals), first, count},
                           a->mi_rows, mi_chunk, mio);
        const CiPass cp{*a, t1, n1, tw, first};
        auto* cic = ws.at<int64_t>(o_cic);
        auto* cio = ws.at<int64_t>(o_cio);
        hipLaunchKernelGGL(oc_count_k<CiPass>, dim3(unsigned(ci_nb)), dim3(kT), 0, ctx->stream,
                           cp, a->ci_rows, ci_chunk, cic);
        BOLAP_HIP_TRY(hipMemcpyAsync(cio, cic, size_t(ci_nb) * 8, hipMemcpyDeviceToDevice,
                                     ctx->stream));
        BOLAP_TRY(exclusive_scan_i64(ctx, cio, ci_nb, scan));
        hipLaunchKernelGGL((oc_write_k<CiPass, CiEmit>), dim3(unsigned(ci_nb)), dim3(kT), 0,
                           ctx->stream, cp,
                           CiEmit{ws.at<int32_t>(o_pm), ws.at<int32_t>(o_pp), cand}, a->ci_rows,
                           ci_chunk, cio);
        hipLaunchKernelGGL(q10_name_k, dim3(grid_of(ctx, a->name_rows)), dim3(kT), 0,
                           ctx->stream, *a, n1, cand, pw, mark);
        hipLaunchKernelGGL(q10_bits_k, dim3(grid_of(ctx, words)), dim3(kT), 0, ctx->stream, mark,
                           pool, bits, wcnt);
        BOLAP_TRY(exclusive_scan_i64(ctx, wcnt, words, scan));
        hipLaunchKernelGGL(q10_sizes_k, dim3(1), dim3(64), 0, ctx->stream, bits, wcnt, words,
                           mio, int64_t(mi_nb), mic, cio, int64_t(ci_nb), cic,
                           ws.at<int64_t>(o_sizes));
        BOLAP_TRY(launch_check("q10_sizes_k"));
    }
    int64_t sz[3];
    BOLAP_HIP_TRY(hipMemcpyAsync(sz, ws.at<int64_t>(o_sizes), sizeof(sz), hipMemcpyDeviceToHost,
                                 ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const int64_t groups = sz[0], values = sz[1], pairs = sz[2];
    trace_count(ctx, "q10a_name_groups", groups);
    trace_count(ctx, "q10a_info_values", values);
    trace_count(ctx, "q10a_join_rows", pairs);
    if (groups == 0 || values == 0 || pairs == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(groups < (int64_t(1) << 31), "bolap_ceb_q10a: %lld name groups",
                    (long long)groups);
    {
        TraceScope ts(ctx, "q10a_join");
        hipLaunchKernelGGL(q10_pairs_k, dim3(grid_of(ctx, pairs)), dim3(kT), 0, ctx->stream,
                           ws.at<int32_t>(o_pp), pairs, pw, bits, wcnt, ws.at<int32_t>(o_pg));
        hipLaunchKernelGGL(q10_gstr_k, dim3(grid_of(ctx, pool)), dim3(kT), 0, ctx->stream, bits,
                           wcnt, pool, ws.at<int32_t>(o_gstr));
        hipLaunchKernelGGL(q10_items_k, dim3(grid_of(ctx, t1)), dim3(kT), 0, ctx->stream, tw,
                           first, t1, ws.at<uint8_t>(o_has), ws.at<int32_t>(o_year));
        BOLAP_TRY(launch_check("q10_items_k"));
    }
    bolap_minmax_args m{};
    m.items = t1;
    // room for the packed items at prep_end first (minmax_expanded writes
    // them there), so that the addresses taken below stay valid
    BOLAP_TRY(ws_reserve(ctx, minmax_pack_end(&m, prep_end), prep_end));
    ws.base = static_cast<char*>(ctx->ws);
    m.pairs = pairs;
    m.pair_group = ws.at<int32_t>(o_pg);
    m.pair_item = ws.at<int32_t>(o_pm);
    m.item_offset = ws.at<int32_t>(o_first);
    m.item_count = ws.at<int32_t>(o_cnt);
    m.item_has = ws.at<uint8_t>(o_has);
    m.item_value = ws.at<int32_t>(o_year);
    m.values = values;
    m.range_value = ws.at<int32_t>(o_vals);
    m.groups = int32_t(groups);
    m.keys = a->n_infos;
    m.dense_max_slots = a->dense_max_slots;
    int64_t expanded = 0;
    BOLAP_TRY(minmax_expanded(ctx, &m, prep_end, &expanded));
    const int64_t slots = groups * int64_t(a->n_infos);
    const int64_t row_cap = std::max<int64_t>(1, std::min(slots, expanded));
    const size_t mm_end = minmax_ws_end(ctx, &m, expanded, true, prep_end);
    const size_t o_rows = WsArena::align(mm_end);
    const size_t all_end = o_rows + size_t(row_cap) * sizeof(bolap_minmax_row);
    // one reservation for the aggregation and the rows (keeping the join
    // rows and items above and the packed items), then the pointers into
    // the moved workspace
    BOLAP_TRY(ws_reserve(ctx, all_end, minmax_pack_end(&m, prep_end)));
    ws.base = static_cast<char*>(ctx->ws);
    m.pair_group = ws.at<int32_t>(o_pg);
    m.pair_item = ws.at<int32_t>(o_pm);
    m.item_offset = ws.at<int32_t>(o_first);
    m.item_count = ws.at<int32_t>(o_cnt);
    m.item_has = ws.at<uint8_t>(o_has);
    m.item_value = ws.at<int32_t>(o_year);
    m.range_value = ws.at<int32_t>(o_vals);
    auto* rows = ws.at<bolap_minmax_row>(o_rows);
    int64_t n = 0;
    BOLAP_TRY(minmax_run(ctx, &m, expanded, true, prep_end, rows, row_cap, &n, stats));
    *n_out = n;
    if (n > max_rows) {
        set_error("bolap_ceb_q10a: %lld result rows, room for %lld", (long long)n,
                  (long long)max_rows);
        return BOLAP_ERR_CAPACITY;
    }
    if (n == 0) return BOLAP_OK;
    hipLaunchKernelGGL(q10_rows_k, dim3(grid_of(ctx, n)), dim3(kT), 0, ctx->stream, rows, n,
                       ws.at<int32_t>(o_gstr), a->info_ids);
    BOLAP_TRY(launch_check("q10_rows_k"));
    BOLAP_HIP_TRY(hipMemcpyAsync(rows_out, rows, size_t(n) * sizeof(bolap_minmax_row),
                                 hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BOLAP_OK;
}
