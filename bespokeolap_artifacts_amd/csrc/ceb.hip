// ceb.hip — CEB-style star queries over movie ids: COUNT(*) as the sum over
// qualifying movies of the product of per-table match counts, and
// MIN(production_year), one query at a time or as a batch.
//
// Reference: bespoke_ceb/query_q1a.cpp builds per-dimension masks (:69-152),
// filters title into candidate movies (:241-306), counts matching rows per
// movie in dense arrays per fact table (:346-467, :505-596) and sums the
// per-movie products (:602-621); query_q10a.cpp:334 ff. keeps MIN aggregates.
//
// Device stages (the same kernels serve one query and a batch of n):
//   1. title: per query, a bitmap over movies passing year / kind / title-
//      string predicates; one wave evaluates 64 movies for every query of
//      the batch, so the title columns are read once.
//   2. rank: per query, the exclusive prefix of the bitmap's word popcounts,
//      so qualifying movie m gets a dense index (its rank) and the per-table
//      count arrays hold only qualifying movies.
//   3. layout: per query and table, the offset of its count array in one
//      arena (an exclusive scan over the batch).
//   4. fact: one persistent kernel per fact table reads the table once and
//      tests each row against the queries that use the table and that its
//      movie passes (a movie-major copy of the title bits lists them), then
//      the row predicates and dimension bitmaps, and counts it.  Rows are
//      sorted by movie, so one popcount per movie run leaves one atomic.
//   5. final: per query, the 128-bit sum of count products and the minimum
//      year over movies with a match in every table; partials, then a merge.
// A batch is captured once into a hipGraph and replayed; the fact stage's
// workgroups per CU and query chunk come from the autotuner.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace bolap {
namespace {

constexpr int kT = 256;
constexpr int kRankT = 1024;
constexpr int kMaxBatch = 8192;
constexpr int kTables = 3;  // 0 cast_info, 1 movie_info, 2 movie_keyword

struct QK {
    int32_t tables, agg, year_lo, year_hi;
    uint32_t kind_mask, role_mask;
    const uint32_t* title_bits;
    int64_t title_bits_n;
    const uint32_t* info_bits;
    int64_t info_bits_n;
    const uint32_t* type_bits;
    int64_t type_bits_n;
    const uint32_t* kw_bits;
    int64_t kw_bits_n;
};

struct QState {
    uint64_t n_tf;              // movies passing the title predicates
    uint64_t off[kTables];      // count-array offsets in the arena (u32 units)
};

struct Part {
    uint64_t lo;
    int64_t hi;
    int64_t min_year;
    int64_t joined;
};

struct DbK {
    int64_t n_title, words;
    const int32_t *year, *kind, *code;
    int64_t rows[kTables];
    const int32_t* movie[kTables];
    const int32_t* a1[kTables];  // role / info_type / keyword
    const int32_t* a2;           // movie_info.info
};

BOLAP_DEV bool bit_of(const uint32_t* bits, int64_t n, int64_t v) {
    if (!bits) return true;
    if (v < 0 || v >= n) return false;
    return (bits[v >> 5] >> (v & 31)) & 1u;
}

// 1. Title bitmaps: wave g covers movies [64 g, 64 g + 64) = words 2g, 2g+1.
// Two layouts of the same bits: query-major `tf` (one bitmap over movies per
// query, for ranks and the final stage) and movie-major `qm` (word w of movie
// m = which of queries 32 w .. 32 w + 31 the movie passes, stored at
// qm[w * n_title + m]), which lets the fact stage visit only the queries a
// row's movie passes.
// The predicates of up to kStage queries are staged in LDS per workgroup, so
// the per-query loop reads broadcast LDS words instead of scalar loads.
constexpr int kStage = 512;  // a multiple of 32: a staged chunk holds whole query words

struct TitleQ {
    int32_t lo, hi;
    uint32_t kinds, pad;
    const uint32_t* bits;
    int64_t n;
};

// A wave takes 256 movies as four 64-movie chunks, so one LDS read of a
// query's predicate serves four chunks and the query's eight bitmap words
// for those movies go out as one eight-lane store.
constexpr int kTitleU = 4;
static_assert(kTitleU == 4, "the bitmap-word store below selects among four chunks");

__global__ __launch_bounds__(kT) void ceb_title_k(DbK db, const QK* __restrict__ qs, int32_t nq,
                                                  uint32_t* __restrict__ tf,
                                                  uint32_t* __restrict__ qm) {
    __shared__ TitleQ s_q[kStage];
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wstride = int64_t(gridDim.x) * (kT / kWave);
    constexpr int64_t kSpan = int64_t(kWave) * kTitleU;  // movies per wave step
    for (int c0 = 0; c0 < nq; c0 += kStage) {
        const int cn = nq - c0 < kStage ? nq - c0 : kStage;
        __syncthreads();  // the previous chunk's readers are done
        for (int i = threadIdx.x; i < cn; i += kT) {
            const QK& Q = qs[c0 + i];
            s_q[i] = TitleQ{Q.year_lo, Q.year_hi, Q.kind_mask, 0u, Q.title_bits, Q.title_bits_n};
        }
        __syncthreads();
        for (int64_t g = int64_t(blockIdx.x) * (kT / kWave) + threadIdx.x / kWave;
             g * kSpan < db.n_title; g += wstride) {
            int32_t y[kTitleU], k[kTitleU], c[kTitleU];
            bool in[kTitleU];
            uint32_t acc[kTitleU];
#pragma unroll
            for (int j = 0; j < kTitleU; ++j) {
                const int64_t m = g * kSpan + j * kWave + lane;
                in[j] = m < db.n_title;
                y[j] = in[j] ? db.year[m] : 0;
                k[j] = in[j] ? db.kind[m] : 0;
                c[j] = in[j] ? db.code[m] : 0;
                acc[j] = 0;
            }
            // lane l < 8 writes bitmap word 8 g + l: chunk l / 2, half l % 2
            const int64_t w_out = g * (2 * kTitleU) + lane;
            const bool writer = lane < 2 * kTitleU && w_out < db.words;
            for (int i = 0; i < cn; ++i) {
                const int q = c0 + i;
                const TitleQ Q = s_q[i];
                uint64_t bl[kTitleU];
#pragma unroll
                for (int j = 0; j < kTitleU; ++j) {
                    const bool p = in[j] && y[j] >= Q.lo && y[j] <= Q.hi &&
                                   uint32_t(k[j]) < 32u && ((Q.kinds >> k[j]) & 1u) &&
                                   bit_of(Q.bits, Q.n, c[j]);
                    acc[j] |= uint32_t(p) << (q & 31);
                    bl[j] = __ballot(p);
                }
                if ((q & 31) == 31 || q == nq - 1) {
#pragma unroll
                    for (int j = 0; j < kTitleU; ++j) {
                        if (in[j]) qm[int64_t(q >> 5) * db.n_title + g * kSpan + j * kWave + lane] =
                            acc[j];
                        acc[j] = 0;
                    }
                }
                if (writer) {
                    const int j = lane >> 1;
                    const uint64_t b = j == 0 ? bl[0] : j == 1 ? bl[1] : j == 2 ? bl[2] : bl[3];
                    tf[int64_t(q) * db.words + w_out] = (lane & 1) ? uint32_t(b >> 32) : uint32_t(b);
                }
            }
        }
    }
}

// 2. Rank prefix per query: one workgroup per query walks its bitmap in
// tiles of 4 words per thread (16-byte coalesced loads and stores), a
// workgroup scan of the per-thread popcounts per tile, and a running carry.
__global__ __launch_bounds__(kRankT) void ceb_rank_k(const uint32_t* __restrict__ tf, int64_t words,
                                                     uint32_t* __restrict__ rank,
                                                     QState* __restrict__ st) {
    constexpr int kW = kRankT / kWave;
    __shared__ uint32_t s_w[kW];
    const int q = blockIdx.x;
    const uint32_t* row = tf + int64_t(q) * words;
    uint32_t* rr = rank + int64_t(q) * words;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    // rows start at q * words: 16-byte accesses only when that stays aligned
    const bool vec = (words & 3) == 0;
    uint32_t carry = 0;
    for (int64_t t0 = 0; t0 < words; t0 += 4 * kRankT) {
        const int64_t i = t0 + 4 * int64_t(threadIdx.x);
        uint32_t v[4] = {0u, 0u, 0u, 0u};
        if (vec && i + 3 < words) {
            const uint4 u = *reinterpret_cast<const uint4*>(row + i);
            v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = i + k < words ? row[i + k] : 0u;
        }
        const uint32_t mine = __popc(v[0]) + __popc(v[1]) + __popc(v[2]) + __popc(v[3]);
        uint32_t incl = mine;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, kWave);
            if (lane >= off) incl += y;
        }
        if (lane == kWave - 1) s_w[wave] = incl;
        __syncthreads();
        uint32_t before = 0, tile = 0;
#pragma unroll
        for (int w = 0; w < kW; ++w) {
            before += w < wave ? s_w[w] : 0u;
            tile += s_w[w];
        }
        __syncthreads();  // s_w is rewritten by the next tile
        uint32_t run = carry + before + incl - mine;
        uint32_t o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            o[k] = run;
            run += __popc(v[k]);
        }
        if (vec && i + 3 < words) {
            *reinterpret_cast<uint4*>(rr + i) = make_uint4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (i + k < words) rr[i + k] = o[k];
        }
        carry += tile;
    }
    if (threadIdx.x == 0) st[q].n_tf = carry;
}

// 3. Count-array offsets: exclusive scan over the batch of n_tf x (tables).
__global__ __launch_bounds__(kRankT) void ceb_layout_k(const QK* __restrict__ qs, int32_t nq,
                                                       QState* __restrict__ st,
                                                       unsigned long long* __restrict__ total) {
    __shared__ unsigned long long s_w[kRankT / kWave];
    __shared__ unsigned long long s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    for (int base = 0; base < nq; base += kRankT) {
        const int q = base + threadIdx.x;
        unsigned long long need = 0;
        if (q < nq) need = st[q].n_tf * unsigned(__popc(uint32_t(qs[q].tables) & 7u));
        unsigned long long incl = need;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const unsigned long long y = __shfl_up(incl, off, kWave);
            if (lane >= off) incl += y;
        }
        if (lane == kWave - 1) s_w[wave] = incl;
        __syncthreads();
        unsigned long long before = s_carry, tile = 0;
        for (int w = 0; w < kRankT / kWave; ++w) {
            before += w < wave ? s_w[w] : 0ull;
            tile += s_w[w];
        }
        if (q < nq) {
            unsigned long long off = before + incl - need;
            for (int t = 0; t < kTables; ++t) {
                if ((qs[q].tables >> t) & 1) {
                    st[q].off[t] = off;
                    off += st[q].n_tf;
                } else {
                    st[q].off[t] = ~0ull;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) s_carry += tile;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = s_carry;
}

// Zeroes the used part of the arena (its length is on the device).
__global__ __launch_bounds__(kT) void ceb_zero_k(uint32_t* __restrict__ arena,
                                                 const unsigned long long* __restrict__ total) {
    const int64_t n = int64_t(*total);
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kT)
        arena[i] = 0u;
}

// A query's row predicate on fact table TB and its count-array offset, as the
// fact stage stages them in LDS: cast_info keeps the role mask, movie_info the
// info_type and info bitmaps, movie_keyword the keyword bitmap.
template <int TB>
struct FactQ;
template <>
struct FactQ<0> {
    uint64_t off;
    uint32_t role, pad;
    BOLAP_DEV void set(const QK& Q, uint64_t o) { off = o; role = Q.role_mask; pad = 0; }
    BOLAP_DEV bool pass(int32_t x, int32_t) const { return uint32_t(x) < 32u && ((role >> x) & 1u); }
};
template <>
struct FactQ<1> {
    uint64_t off;
    const uint32_t *b1, *b2;
    int64_t n1, n2;
    BOLAP_DEV void set(const QK& Q, uint64_t o) {
        off = o;
        b1 = Q.type_bits;
        n1 = Q.type_bits_n;
        b2 = Q.info_bits;
        n2 = Q.info_bits_n;
    }
    BOLAP_DEV bool pass(int32_t x, int32_t z) const { return bit_of(b1, n1, x) && bit_of(b2, n2, z); }
};
template <>
struct FactQ<2> {
    uint64_t off;
    const uint32_t* b1;
    int64_t n1;
    BOLAP_DEV void set(const QK& Q, uint64_t o) {
        off = o;
        b1 = Q.kw_bits;
        n1 = Q.kw_bits_n;
    }
    BOLAP_DEV bool pass(int32_t x, int32_t) const { return bit_of(b1, n1, x); }
};

// 4. One fact table for the queries that use it (bit q of use[q >> 5]).
// Each wave owns a contiguous range of rows and walks it in groups of 256
// rows, as four chunks of 64 (one row per lane per chunk, so each load is one
// coalesced 256-byte access).  For each 32-query word the wave loads its
// rows' movie-major title words and visits only the queries some row of the
// group passes, instead of testing every row against every query.
// Within a chunk the movie ids are non-decreasing across lanes; the last lane
// of each run of one movie holds the run's lanes as a mask, so a run's matches
// are one popcount of the query's ballot.  A run that reaches the chunk's last
// lane and continues in the next chunk of the wave's range hands its count on
// (a scalar within a group, a per-wave LDS slot per query between groups), so
// a movie with many rows costs one atomic per wave range and query rather
// than one per chunk.  The lane that ends a run adds its count to the query's
// count array at the movie's rank.
template <int TB>
__global__ __launch_bounds__(kT) void ceb_fact_k(DbK db, const QK* __restrict__ qs,
                                                 const uint32_t* __restrict__ use, int32_t qwords,
                                                 int32_t wchunk, const uint32_t* __restrict__ qm,
                                                 const uint32_t* __restrict__ tf,
                                                 const uint32_t* __restrict__ rank,
                                                 const QState* __restrict__ st,
                                                 uint32_t* __restrict__ arena) {
    constexpr int U = 4;
    constexpr int64_t kGroup = int64_t(U) * kWave;
    __shared__ FactQ<TB> s_q[kStage];
    __shared__ uint32_t s_carry[kT / kWave][kStage];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int64_t rows = db.rows[TB];
    const int32_t* __restrict__ movie = db.movie[TB];
    const int32_t* __restrict__ a1 = db.a1[TB];
    const int32_t* __restrict__ a2 = db.a2;
    const int64_t groups = (rows + kGroup - 1) / kGroup;
    const int64_t nwaves = int64_t(gridDim.x) * (kT / kWave);
    const int64_t per = (groups + nwaves - 1) / nwaves;
    const int64_t g0 = (int64_t(blockIdx.x) * (kT / kWave) + wv) * per;
    const int64_t g1 = g0 + per < groups ? g0 + per : groups;
    // lanes 0 .. lane
    const uint64_t upto = lane == kWave - 1 ? ~0ull : (2ull << lane) - 1ull;
    if (wchunk > kStage / 32) wchunk = kStage / 32;
    for (int c0 = 0; c0 < qwords; c0 += wchunk) {
        const int c1 = c0 + wchunk < qwords ? c0 + wchunk : qwords;
        const int q0 = c0 * 32;
        __syncthreads();  // the previous chunk's readers are done
        for (int i = threadIdx.x; i < (c1 - c0) * 32; i += kT) {
            const int q = q0 + i;
            if ((use[q >> 5] >> (q & 31)) & 1u) s_q[i].set(qs[q], st[q].off[TB]);
        }
        for (int i = lane; i < kStage; i += kWave) s_carry[wv][i] = 0u;
        __syncthreads();
        bool cin0 = false;  // the first run of the group continues the previous group's last
        for (int64_t g = g0; g < g1; ++g) {
            int32_t m[U], x[U], z[U];
            uint64_t run[U];  // the run's lanes, held by its last lane (else 0)
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int64_t r = (g * U + j) * kWave + lane;
                const bool in = r < rows;
                m[j] = in ? movie[r] - 1 : -1;
                x[j] = in ? a1[r] : 0;
                z[j] = (TB == 1 && in) ? a2[r] : 0;
                if (m[j] < 0 || m[j] >= db.n_title) m[j] = -1;
            }
            // first movie of the next group in this wave's range (-1: none)
            int32_t mnext = -1;
            if (g + 1 < g1) {
                mnext = movie[(g + 1) * kGroup] - 1;
                if (mnext < 0 || mnext >= db.n_title) mnext = -1;
            }
            bool cin[U], cout[U], single[U];
            uint64_t lastrun[U];
            int firsttail[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int32_t prev = __shfl_up(m[j], 1, kWave);
                const int32_t next = __shfl_down(m[j], 1, kWave);
                const uint64_t heads = __ballot(lane == 0 || prev != m[j]);
                const bool tail = m[j] >= 0 && (lane == kWave - 1 || next != m[j]);
                const int h = 63 - __clzll(heads & upto);  // lane 0 is always a head
                run[j] = tail ? (upto & ~((1ull << h) - 1ull)) : 0ull;
                const uint64_t tails = __ballot(tail);
                firsttail[j] = tails ? __builtin_ctzll(tails) : kWave;
                lastrun[j] = readlane_u64(run[j], kWave - 1);
                // every lane holds the same (valid) movie
                single[j] = heads == 1ull && tails == (1ull << (kWave - 1));
            }
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int32_t mlast = __builtin_amdgcn_readlane(m[j], kWave - 1);
                const int32_t mfollow =
                    j + 1 < U ? __builtin_amdgcn_readlane(m[j < U - 1 ? j + 1 : j], 0) : mnext;
                cout[j] = mlast >= 0 && mfollow == mlast;
                cin[j] = j == 0 ? cin0 : cout[j - 1];
            }
            for (int w = c0; w < c1; ++w) {
                const uint32_t uw = use[w];
                if (uw == 0) continue;
                uint32_t M[U], any = 0;
#pragma unroll
                for (int j = 0; j < U; ++j) {
                    M[j] = m[j] >= 0 ? qm[int64_t(w) * db.n_title + m[j]] : 0u;
                    any |= M[j];
                }
                uint32_t Ms[U];  // the chunk's movie's word when the chunk is one movie
#pragma unroll
                for (int j = 0; j < U; ++j) Ms[j] = __builtin_amdgcn_readfirstlane(M[j]);
                uint32_t todo = wave_or(any) & uw;
                while (todo) {
                    const int b = __builtin_ctz(todo);
                    todo &= todo - 1u;
                    const int q = w * 32 + b;
                    const FactQ<TB> F = s_q[q - q0];
                    uint32_t carry = cin[0] ? s_carry[wv][q - q0] : 0u;
#pragma unroll
                    for (int j = 0; j < U; ++j) {
                        const uint32_t cj = cin[j] ? carry : 0u;
                        if (single[j]) {  // wave-uniform: one movie, one run
                            uint32_t c = cj;
                            if ((Ms[j] >> b) & 1u) c += uint32_t(__popcll(__ballot(F.pass(x[j], z[j]))));
                            carry = cout[j] ? c : 0u;
                            if (!cout[j] && c && lane == 0) {
                                const int64_t wi = int64_t(q) * db.words + (m[j] >> 5);
                                const uint32_t r =
                                    rank[wi] + __popc(tf[wi] & ((1u << (m[j] & 31)) - 1u));
                                atomicAdd(arena + F.off + r, c);
                            }
                            continue;
                        }
                        const bool ok = ((M[j] >> b) & 1u) && F.pass(x[j], z[j]);
                        const uint64_t bal = __ballot(ok);
                        uint32_t v = uint32_t(__popcll(bal & run[j]));
                        if (lane == firsttail[j]) v += cj;
                        if (cout[j]) {
                            carry = uint32_t(__popcll(bal & lastrun[j])) +
                                    (firsttail[j] == kWave - 1 ? cj : 0u);
                            if (lane == kWave - 1) v = 0;
                        } else {
                            carry = 0u;
                        }
                        if (v) {
                            const int64_t wi = int64_t(q) * db.words + (m[j] >> 5);
                            const uint32_t r =
                                rank[wi] + __popc(tf[wi] & ((1u << (m[j] & 31)) - 1u));
                            atomicAdd(arena + F.off + r, v);
                        }
                    }
                    if (lane == 0) s_carry[wv][q - q0] = cout[U - 1] ? carry : 0u;
                }
            }
            cin0 = cout[U - 1];
        }
    }
}

// 5. Per (query, slice of bitmap words): sum of count products and min year
// over movies with a match in every table of the query.  A wave takes 64
// words (2048 movies) at a time; a wave scan of their popcounts numbers the
// passing movies, whose ranks are consecutive from the tile's first rank, so
// lane k of a round takes the tile's k-th passing movie and the count loads
// of a round are coalesced.  The lane finds its movie (needed for the year)
// by a binary search of the scan across lanes and a select within the word.
__global__ __launch_bounds__(kT) void ceb_final_k(DbK db, const QK* __restrict__ qs,
                                                  const uint32_t* __restrict__ tf,
                                                  const uint32_t* __restrict__ rank,
                                                  const QState* __restrict__ st,
                                                  const uint32_t* __restrict__ arena,
                                                  int32_t slices, Part* __restrict__ parts) {
    __shared__ uint64_t s_lo[kT / kWave];
    __shared__ int64_t s_hi[kT / kWave], s_min[kT / kWave], s_j[kT / kWave];
    const int q = blockIdx.y;
    const int b = blockIdx.x;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int tables = qs[q].tables;
    const QState S = st[q];
    const int64_t per = (db.words + slices - 1) / slices;
    const int64_t w0 = int64_t(b) * per;
    const int64_t w1 = w0 + per < db.words ? w0 + per : db.words;
    const uint32_t* row = tf + int64_t(q) * db.words;
    const uint32_t* rk = rank + int64_t(q) * db.words;
    Acc128 acc{0, 0};
    int64_t mn = INT64_MAX, joined = 0;
    for (int64_t t = w0 + int64_t(wave) * kWave; t < w1; t += int64_t(kT)) {
        const uint32_t word = t + lane < w1 ? row[t + lane] : 0u;
        const uint32_t cnt = __popc(word);
        uint32_t incl = cnt;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, kWave);
            if (lane >= off) incl += y;
        }
        const uint32_t excl = incl - cnt;
        const uint32_t total = uint32_t(__builtin_amdgcn_readlane(int(incl), kWave - 1));
        if (total == 0) continue;
        const uint32_t rbase = rk[t];
        for (uint32_t k0 = 0; k0 < total; k0 += kWave) {
            const uint32_t k = k0 + uint32_t(lane);
            const bool valid = k < total;
            const uint32_t r = rbase + k;
            uint64_t lo = 1, hi = 0;
            bool all = valid;
            for (int tb = 0; tb < kTables; ++tb) {
                if (!((tables >> tb) & 1)) continue;
                const uint32_t c = valid ? arena[S.off[tb] + r] : 0u;
                all = all && c != 0;
                // (hi, lo) *= c, exact while the product stays below 2^128
                const uint64_t nlo = lo * uint64_t(c);
                hi = hi * uint64_t(c) + __umul64hi(lo, uint64_t(c));
                lo = nlo;
            }
            // largest lane s with excl[s] <= k: the word holding passing movie k
            int src = 0;
#pragma unroll
            for (int step = kWave / 2; step > 0; step >>= 1) {
                const uint32_t e = uint32_t(__shfl(int(excl), src + step, kWave));
                if (e <= k) src += step;
            }
            const uint32_t wd = uint32_t(__shfl(int(word), src, kWave));
            uint32_t jj = k - uint32_t(__shfl(int(excl), src, kWave));
            if (!all) continue;
            // position of the jj-th set bit of wd
            uint32_t v = wd;
            int pos = 0;
#pragma unroll
            for (int sh = 16; sh > 0; sh >>= 1) {
                const uint32_t low = __popc(v & ((1u << sh) - 1u));
                if (jj >= low) {
                    jj -= low;
                    v >>= sh;
                    pos += sh;
                }
            }
            ++joined;
            const uint64_t nlo = acc.lo + lo;
            acc.hi += int64_t(hi) + (nlo < acc.lo ? 1 : 0);
            acc.lo = nlo;
            const int64_t y = db.year[(t + src) * 32 + pos];
            mn = y < mn ? y : mn;
        }
    }
    const Acc128 ws = wave_sum128(acc);
    int64_t wm = mn;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int64_t o = __shfl_xor(wm, off, kWave);
        wm = o < wm ? o : wm;
    }
    const int64_t wj = int64_t(wave_sum(uint64_t(joined)));
    if (lane == 0) {
        s_lo[wave] = ws.lo;
        s_hi[wave] = ws.hi;
        s_min[wave] = wm;
        s_j[wave] = wj;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Acc128 t{0, 0};
        int64_t m2 = INT64_MAX, j2 = 0;
        for (int w = 0; w < kT / kWave; ++w) {
            add128(t, Acc128{s_lo[w], s_hi[w]});
            m2 = s_min[w] < m2 ? s_min[w] : m2;
            j2 += s_j[w];
        }
        parts[int64_t(q) * slices + b] = Part{t.lo, t.hi, m2, j2};
    }
}

__global__ __launch_bounds__(kWave) void ceb_merge_k(const Part* __restrict__ parts, int32_t slices,
                                                     const QState* __restrict__ st,
                                                     bolap_ceb_result* __restrict__ out) {
    const int q = blockIdx.x;
    Acc128 acc{0, 0};
    int64_t mn = INT64_MAX, joined = 0;
    for (int b = threadIdx.x; b < slices; b += kWave) {
        const Part p = parts[int64_t(q) * slices + b];
        add128(acc, Acc128{p.lo, p.hi});
        mn = p.min_year < mn ? p.min_year : mn;
        joined += p.joined;
    }
    const Acc128 t = wave_sum128(acc);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int64_t o = __shfl_xor(mn, off, kWave);
        mn = o < mn ? o : mn;
    }
    const int64_t j = int64_t(wave_sum(uint64_t(joined)));
    if (threadIdx.x == 0)
        out[q] = bolap_ceb_result{t.lo, uint64_t(t.hi), mn, int64_t(st[q].n_tf), j};
}

// ------------------------------------------------------------------ host
struct Layout {
    int32_t nq = 0, slices = 0, qwords = 0;
    int64_t words = 0;
    size_t o_qs, o_use, o_st, o_tf, o_qm, o_rank, o_total, o_parts, o_res, o_arena, bytes;
    int64_t arena_words = 0;
};

int check_db(const bolap_ceb_db* d) {
    BOLAP_ARG_CHECK(d->title_rows > 0 && d->title_rows < (int64_t(1) << 31),
                    "bolap_ceb: title_rows %lld", (long long)d->title_rows);
    BOLAP_ARG_CHECK(d->t_year && d->t_kind && d->t_code, "bolap_ceb: null title column");
    BOLAP_ARG_CHECK(d->ci_rows >= 0 && d->mi_rows >= 0 && d->mk_rows >= 0,
                    "bolap_ceb: negative row count");
    BOLAP_ARG_CHECK(d->ci_rows == 0 || (d->ci_movie && d->ci_role), "bolap_ceb: cast_info");
    BOLAP_ARG_CHECK(d->mi_rows == 0 || (d->mi_movie && d->mi_type && d->mi_info),
                    "bolap_ceb: movie_info");
    BOLAP_ARG_CHECK(d->mk_rows == 0 || (d->mk_movie && d->mk_keyword),
                    "bolap_ceb: movie_keyword");
    return BOLAP_OK;
}

int check_query(const bolap_ceb_query* q, int i) {
    BOLAP_ARG_CHECK((q->tables & 7) != 0 && (q->tables & ~7) == 0,
                    "bolap_ceb: query %d: tables 0x%x", i, q->tables);
    BOLAP_ARG_CHECK(q->agg == BOLAP_CEB_COUNT || q->agg == BOLAP_CEB_MIN_YEAR,
                    "bolap_ceb: query %d: agg %d", i, q->agg);
    BOLAP_ARG_CHECK((!q->title_bits || q->title_bits_n >= 0) && (!q->info_bits || q->info_bits_n >= 0) &&
                        (!q->type_bits || q->type_bits_n >= 0) && (!q->kw_bits || q->kw_bits_n >= 0),
                    "bolap_ceb: query %d: negative bitmap size", i);
    return BOLAP_OK;
}

DbK make_db(const bolap_ceb_db* d) {
    DbK k;
    std::memset(&k, 0, sizeof(k));
    k.n_title = d->title_rows;
    k.words = (d->title_rows + 31) / 32;
    k.year = d->t_year;
    k.kind = d->t_kind;
    k.code = d->t_code;
    k.rows[0] = d->ci_rows;
    k.movie[0] = d->ci_movie;
    k.a1[0] = d->ci_role;
    k.rows[1] = d->mi_rows;
    k.movie[1] = d->mi_movie;
    k.a1[1] = d->mi_type;
    k.a2 = d->mi_info;
    k.rows[2] = d->mk_rows;
    k.movie[2] = d->mk_movie;
    k.a1[2] = d->mk_keyword;
    return k;
}

QK make_q(const bolap_ceb_query& q) {
    QK k;
    k.tables = q.tables;
    k.agg = q.agg;
    k.year_lo = q.year_lo;
    k.year_hi = q.year_hi;
    k.kind_mask = q.kind_mask;
    k.role_mask = q.role_mask;
    k.title_bits = q.title_bits;
    k.title_bits_n = q.title_bits_n;
    k.info_bits = q.info_bits;
    k.info_bits_n = q.info_bits_n;
    k.type_bits = q.type_bits;
    k.type_bits_n = q.type_bits_n;
    k.kw_bits = q.kw_bits;
    k.kw_bits_n = q.kw_bits_n;
    return k;
}

Layout plan_layout(bolap_ctx* ctx, const DbK& db, const bolap_ceb_query* qs, int32_t nq) {
    Layout L;
    L.nq = nq;
    L.words = db.words;
    L.qwords = (nq + 31) / 32;
    L.slices = int32_t(std::max<int64_t>(1, std::min<int64_t>(64, (db.words + 2047) / 2048)));
    (void)ctx;
    int64_t arena = 0;
    for (int i = 0; i < nq; ++i) arena += int64_t(__builtin_popcount(qs[i].tables & 7)) * db.n_title;
    L.arena_words = arena;
    WsArena a;
    L.o_qs = a.plan(size_t(nq) * sizeof(QK));
    L.o_use = a.plan(size_t(kTables) * L.qwords * 4);
    L.o_st = a.plan(size_t(nq) * sizeof(QState));
    L.o_tf = a.plan(size_t(nq) * db.words * 4);
    L.o_qm = a.plan(size_t(L.qwords) * size_t(db.n_title) * 4);
    L.o_rank = a.plan(size_t(nq) * db.words * 4);
    L.o_total = a.plan(64);
    L.o_parts = a.plan(size_t(nq) * L.slices * sizeof(Part));
    L.o_res = a.plan(size_t(nq) * sizeof(bolap_ceb_result));
    L.o_arena = a.plan(size_t(arena) * 4);
    L.bytes = a.off;
    return L;
}

int ceb_reserve(bolap_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->ceb_size) return BOLAP_OK;
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->ceb_exec) {
        (void)hipGraphExecDestroy(ctx->ceb_exec);
        ctx->ceb_exec = nullptr;
    }
    if (ctx->ceb_graph) {
        (void)hipGraphDestroy(ctx->ceb_graph);
        ctx->ceb_graph = nullptr;
    }
    ctx->ceb_key.clear();
    if (ctx->ceb_buf) BOLAP_HIP_TRY(hipFree(ctx->ceb_buf));
    ctx->ceb_buf = nullptr;
    ctx->ceb_size = 0;
    const size_t want = bytes + bytes / 8;
    BOLAP_HIP_TRY(hipMalloc(&ctx->ceb_buf, want));
    ctx->ceb_size = want;
    return BOLAP_OK;
}

unsigned grid_for(bolap_ctx* ctx, int64_t n, int per_cu) {
    int64_t b = (n + kT - 1) / kT;
    const int64_t cap = int64_t(ctx->num_cus) * per_cu;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

// Enqueues every stage of a planned batch on the context stream.
int enqueue(bolap_ctx* ctx, const DbK& db, const Layout& L, const int32_t* n_per_table) {
    char* base = static_cast<char*>(ctx->ceb_buf);
    auto* qs = reinterpret_cast<const QK*>(base + L.o_qs);
    auto* use = reinterpret_cast<const uint32_t*>(base + L.o_use);
    auto* st = reinterpret_cast<QState*>(base + L.o_st);
    auto* tf = reinterpret_cast<uint32_t*>(base + L.o_tf);
    auto* qm = reinterpret_cast<uint32_t*>(base + L.o_qm);
    auto* rank = reinterpret_cast<uint32_t*>(base + L.o_rank);
    auto* total = reinterpret_cast<unsigned long long*>(base + L.o_total);
    auto* parts = reinterpret_cast<Part*>(base + L.o_parts);
    auto* res = reinterpret_cast<bolap_ceb_result*>(base + L.o_res);
    auto* arena = reinterpret_cast<uint32_t*>(base + L.o_arena);
    const int per_cu = int(ctx->knob("ceb.blocks_per_cu", 4));
    // queries per pass over a fact table, in whole 32-query words
    const int wchunk = int(std::max<int64_t>(1, (ctx->knob("ceb.qchunk", L.nq) + 31) / 32));
    hipLaunchKernelGGL(ceb_title_k, dim3(grid_for(ctx, (db.n_title + 3) / 4, 8)),
                       dim3(kT), 0, ctx->stream, db, qs, L.nq, tf, qm);
    hipLaunchKernelGGL(ceb_rank_k, dim3(unsigned(L.nq)), dim3(kRankT), 0, ctx->stream, tf,
                       L.words, rank, st);
    hipLaunchKernelGGL(ceb_layout_k, dim3(1), dim3(kRankT), 0, ctx->stream, qs, L.nq, st, total);
    hipLaunchKernelGGL(ceb_zero_k, dim3(unsigned(ctx->num_cus * 8)), dim3(kT), 0, ctx->stream,
                       arena, total);
    BOLAP_TRY(launch_check("ceb title/rank/layout"));
    for (int t = 0; t < kTables; ++t) {
        const int32_t nql = n_per_table[t];
        if (nql == 0 || db.rows[t] == 0) continue;
        const unsigned grid = grid_for(ctx, (db.rows[t] + 3) / 4, per_cu);
        const uint32_t* tuse = use + size_t(t) * L.qwords;
        if (t == 0)
            hipLaunchKernelGGL(ceb_fact_k<0>, dim3(grid), dim3(kT), 0, ctx->stream, db, qs, tuse,
                               L.qwords, wchunk, qm, tf, rank, st, arena);
        else if (t == 1)
            hipLaunchKernelGGL(ceb_fact_k<1>, dim3(grid), dim3(kT), 0, ctx->stream, db, qs, tuse,
                               L.qwords, wchunk, qm, tf, rank, st, arena);
        else
            hipLaunchKernelGGL(ceb_fact_k<2>, dim3(grid), dim3(kT), 0, ctx->stream, db, qs, tuse,
                               L.qwords, wchunk, qm, tf, rank, st, arena);
        BOLAP_TRY(launch_check("ceb_fact_k"));
    }
    hipLaunchKernelGGL(ceb_final_k, dim3(unsigned(L.slices), unsigned(L.nq)), dim3(kT), 0,
                       ctx->stream, db, qs, tf, rank, st, arena, L.slices, parts);
    hipLaunchKernelGGL(ceb_merge_k, dim3(unsigned(L.nq)), dim3(kWave), 0, ctx->stream, parts,
                       L.slices, st, res);
    return launch_check("ceb final");
}

// Uploads the queries and, per table, a bitmap over the batch of the queries
// that use it; returns how many queries use each table.
int upload(bolap_ctx* ctx, const Layout& L, const bolap_ceb_query* qs, int32_t* n_per_table) {
    std::vector<QK> hq(static_cast<size_t>(L.nq));
    std::vector<uint32_t> lists(static_cast<size_t>(kTables) * L.qwords, 0u);
    for (int t = 0; t < kTables; ++t) n_per_table[t] = 0;
    for (int i = 0; i < L.nq; ++i) {
        hq[size_t(i)] = make_q(qs[i]);
        for (int t = 0; t < kTables; ++t)
            if ((qs[i].tables >> t) & 1) {
                lists[size_t(t) * L.qwords + size_t(i >> 5)] |= 1u << (i & 31);
                ++n_per_table[t];
            }
    }
    char* base = static_cast<char*>(ctx->ceb_buf);
    const size_t qb = hq.size() * sizeof(QK), lb = lists.size() * 4;
    BOLAP_TRY(pinned_reserve(ctx, qb + lb));
    std::memcpy(ctx->pinned, hq.data(), qb);
    std::memcpy(static_cast<char*>(ctx->pinned) + qb, lists.data(), lb);
    BOLAP_HIP_TRY(hipMemcpyAsync(base + L.o_qs, ctx->pinned, qb, hipMemcpyHostToDevice,
                                 ctx->stream));
    BOLAP_HIP_TRY(hipMemcpyAsync(base + L.o_use, static_cast<char*>(ctx->pinned) + qb, lb,
                                 hipMemcpyHostToDevice, ctx->stream));
    return BOLAP_OK;
}

int read_results(bolap_ctx* ctx, const Layout& L, bolap_ceb_result* out) {
    const size_t rb = size_t(L.nq) * sizeof(bolap_ceb_result);
    BOLAP_TRY(pinned_reserve(ctx, rb));
    BOLAP_HIP_TRY(hipMemcpyAsync(ctx->pinned, static_cast<char*>(ctx->ceb_buf) + L.o_res, rb,
                                 hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::memcpy(out, ctx->pinned, rb);
    return BOLAP_OK;
}

std::string graph_key(bolap_ctx* ctx, const bolap_ceb_db* d, const Layout& L,
                      const int32_t* npt) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%p|%lld|%p|%p|%p|%p|%p|%p|%p|%p|%p|%p|%p|%lld|%lld|%lld|"
                  "%d|%d|%d|%d|%lld|%lld",
                  ctx->ceb_buf, (long long)d->title_rows, (const void*)d->t_year,
                  (const void*)d->t_kind, (const void*)d->t_code, (const void*)d->ci_movie,
                  (const void*)d->ci_role, (const void*)d->mi_movie, (const void*)d->mi_type,
                  (const void*)d->mi_info, (const void*)d->mk_movie,
                  (const void*)d->mk_keyword, (const void*)nullptr, (long long)d->ci_rows,
                  (long long)d->mi_rows, (long long)d->mk_rows, L.nq, npt[0], npt[1], npt[2],
                  (long long)ctx->knob("ceb.blocks_per_cu", 4),
                  (long long)ctx->knob("ceb.qchunk", L.nq));
    return buf;
}

int run_batch(bolap_ctx* ctx, const bolap_ceb_db* d, const bolap_ceb_query* qs, int32_t n,
              bolap_ceb_result* out, bool use_graph) {
    BOLAP_TRY(check_db(d));
    BOLAP_ARG_CHECK(n >= 1 && n <= kMaxBatch, "bolap_ceb: batch of %d queries (1..%d)", n,
                    kMaxBatch);
    for (int i = 0; i < n; ++i) BOLAP_TRY(check_query(&qs[i], i));
    const DbK db = make_db(d);
    const Layout L = plan_layout(ctx, db, qs, n);
    BOLAP_TRY(ceb_reserve(ctx, L.bytes));
    int32_t npt[kTables];
    BOLAP_TRY(upload(ctx, L, qs, npt));
    if (!use_graph) {
        BOLAP_TRY(enqueue(ctx, db, L, npt));
        return read_results(ctx, L, out);
    }
    const std::string key = graph_key(ctx, d, L, npt);
    if (key != ctx->ceb_key || !ctx->ceb_exec) {
        if (ctx->ceb_exec) (void)hipGraphExecDestroy(ctx->ceb_exec);
        if (ctx->ceb_graph) (void)hipGraphDestroy(ctx->ceb_graph);
        ctx->ceb_exec = nullptr;
        ctx->ceb_graph = nullptr;
        ctx->ceb_key.clear();
        // Record on a private stream (the context stream may be the null
        // stream, which cannot capture); the graph then runs on ctx->stream.
        if (!ctx->capture_stream)
            BOLAP_HIP_TRY(hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
        hipStream_t run_stream = ctx->stream;
        BOLAP_HIP_TRY(hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal));
        ctx->stream = ctx->capture_stream;
        const int s = enqueue(ctx, db, L, npt);
        ctx->stream = run_stream;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->capture_stream, &g);
        BOLAP_TRY(s);
        BOLAP_HIP_TRY(e);
        ctx->ceb_graph = g;
        BOLAP_HIP_TRY(hipGraphInstantiate(&ctx->ceb_exec, g, nullptr, nullptr, 0));
        ctx->ceb_key = key;
    }
    BOLAP_HIP_TRY(hipGraphLaunch(ctx->ceb_exec, ctx->stream));
    return read_results(ctx, L, out);
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" {

int bolap_ceb_run(bolap_ctx* ctx, const bolap_ceb_db* db, const bolap_ceb_query* q,
                  bolap_ceb_result* out) {
    BOLAP_ARG_CHECK(ctx && db && q && out, "bolap_ceb_run: null argument");
    TraceScope ts(ctx, "ceb_query");
    return run_batch(ctx, db, q, 1, out, false);
}

int bolap_ceb_run_batch(bolap_ctx* ctx, const bolap_ceb_db* db, const bolap_ceb_query* qs,
                        int32_t n, bolap_ceb_result* out) {
    BOLAP_ARG_CHECK(ctx && db && qs && out, "bolap_ceb_run_batch: null argument");
    TraceScope ts(ctx, "ceb_batch");
    return run_batch(ctx, db, qs, n, out, ctx->knob("ceb.graph", 1) == 1);
}

int bolap_ceb_autotune(bolap_ctx* ctx, const bolap_ceb_db* db, const bolap_ceb_query* qs,
                       int32_t n, double* best_ms) {
    BOLAP_ARG_CHECK(ctx && db && qs && n >= 1, "bolap_ceb_autotune: bad arguments");
    std::vector<bolap_ceb_result> out(static_cast<size_t>(n));
    const int64_t per_cu[] = {2, 4, 8, 16};
    std::vector<int64_t> chunks = {64, 256, int64_t(n)};
    double best = 1e30;
    int64_t bp = 4, bc = n;
    hipEvent_t e0, e1;
    BOLAP_HIP_TRY(hipEventCreate(&e0));
    BOLAP_HIP_TRY(hipEventCreate(&e1));
    int status = BOLAP_OK;
    for (int64_t p : per_cu) {
        for (int64_t c : chunks) {
            if (c > n && c != int64_t(n)) continue;
            ctx->knobs["ceb.blocks_per_cu"] = p;
            ctx->knobs["ceb.qchunk"] = c;
            status = run_batch(ctx, db, qs, n, out.data(), true);  // capture + warm
            if (status != BOLAP_OK) break;
            float ms = 0.f;
            (void)hipEventRecord(e0, ctx->stream);
            status = run_batch(ctx, db, qs, n, out.data(), true);
            (void)hipEventRecord(e1, ctx->stream);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (status != BOLAP_OK) break;
            if (ms < best) {
                best = ms;
                bp = p;
                bc = c;
            }
        }
        if (status != BOLAP_OK) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->knobs["ceb.blocks_per_cu"] = bp;
    ctx->knobs["ceb.qchunk"] = bc;
    if (best_ms) *best_ms = best;
    return status;
}

}  // extern "C"
