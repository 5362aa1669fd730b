// radix.hip — stable LSD radix sort of (u64 key, u32 value) pairs and an
// LDS-histogram radix select for top-k.
//
// Replaces the 16-bit-digit counting sorts of the reference
// (bespoke_tpch/query_q10.cpp:134-166 radix_sort_rows_by_orderkey and
// :318-373 radix_sort_groups, where descending order comes from sorting ~key)
// with a device-wide sort:
//   per pass (8-bit digit): histogram kernel (LDS counters per workgroup over
//   a contiguous chunk) -> exclusive scan of the (digit, workgroup) counts ->
//   scatter kernel that ranks keys stably inside each 256-key tile with wave
//   ballots (peer masks over the 8 digit bits) and per-wave LDS counts.
// Top-k: rounds of 12-bit LDS histograms over the remaining candidates find
// the digit bucket holding the k-th key; keys below it are certain, keys in
// it stay candidates; once at most kSelMax remain, one workgroup sorts them by
// (key, input index) with a bitonic sort in LDS.
#include <algorithm>
#include <cstring>

#include "prims.hpp"
#include "radix.hpp"

namespace bolap {
namespace {

constexpr int kRT = 256;        // threads per workgroup
constexpr int kBins = 256;      // the widest digit (8 bits)
constexpr int kKPT = 16;        // keys per thread per scatter tile
constexpr int kTileK = kRT * kKPT;  // keys per scatter tile (4096)
constexpr int kRW = kRT / kWave;    // waves per workgroup

// Digits are B bits wide (5 <= B <= 8): the host spreads the key bits over
// the fewest 8-bit passes and then narrows the digit as far as that number
// of passes allows, so a tile's digit runs are as long as possible.
template <int B>
BOLAP_DEV uint32_t digit_of(uint64_t key, uint64_t kmask, int shift) {
    return uint32_t(((key & kmask) >> shift) & ((1u << B) - 1u));
}

// Lanes of the wave holding the same digit as this lane (among `live`),
// from B ballots over the digit bits.
template <int B>
BOLAP_DEV uint64_t digit_peers(uint32_t d, uint64_t live) {
    uint64_t peers = live;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

// Digit counts of one scatter workgroup's contiguous chunk, counted by a
// workgroup of kHT threads (more loads in flight than the scatter's grid
// gives).  Lanes with the same digit are matched by ballots and their
// lowest lane adds the group's size, so a skewed digit costs one LDS atomic
// per wave instead of 64 conflicting ones.
constexpr int kHT = 1024;

template <int B>
__global__ __launch_bounds__(kHT) void rs_hist_k(const uint64_t* __restrict__ keys, int64_t n,
                                                 int64_t chunk, uint64_t kmask, int shift,
                                                 int nblocks, int64_t* __restrict__ hist) {
    constexpr int NB = 1 << B;
    __shared__ uint32_t h[NB];
    if (threadIdx.x < NB) h[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t lt_mask = (uint64_t(1) << lane) - 1;
    const int64_t b0 = int64_t(blockIdx.x) * chunk;
    const int64_t b1 = b0 + chunk < n ? b0 + chunk : n;
    constexpr int kU = 4;
    // the trip count is uniform across the workgroup (ballots below)
    for (int64_t base = b0; base < b1; base += int64_t(kHT) * kU) {
        uint64_t k[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int64_t i = base + int64_t(u) * kHT + threadIdx.x;
            k[u] = i < b1 ? keys[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int64_t i = base + int64_t(u) * kHT + threadIdx.x;
            const uint32_t d = digit_of<B>(k[u], kmask, shift);
            const uint64_t peers = digit_peers<B>(d, __ballot(i < b1));
            if (i < b1 && (peers & lt_mask) == 0) atomicAdd(&h[d], uint32_t(__popcll(peers)));
        }
    }
    __syncthreads();
    if (threadIdx.x < NB) hist[int64_t(threadIdx.x) * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one workgroup's chunk in tiles of kTileK keys.  Wave w
// of a tile owns keys [w * kKPT * 64, (w + 1) * kKPT * 64) of it, loaded as
// kKPT coalesced rows of 64, so (wave, row, lane) order is input order.
// Ranking: per row, ballot-matched digit peers; the wave's running count per
// digit (LDS) plus the lane's rank among its peers.  The tile is then
// reordered by digit in LDS (digit start in the tile + the counts of earlier
// waves + the rank) and written out with consecutive threads on consecutive
// slots, so each digit's run of the tile leaves as contiguous stores.
template <int B>
__global__ __launch_bounds__(kRT) void rs_scatter_k(const uint64_t* __restrict__ kin,
                                                    const uint32_t* __restrict__ vin, int64_t n,
                                                    int64_t chunk, uint64_t kmask, int shift,
                                                    int nblocks,
                                                    const int64_t* __restrict__ offsets,
                                                    uint64_t* __restrict__ kout,
                                                    uint32_t* __restrict__ vout) {
    constexpr int NB = 1 << B;
    __shared__ uint64_t s_key[kTileK];
    __shared__ uint32_t s_val[kTileK];
    __shared__ uint32_t s_wcnt[kRW][NB];  // per-wave counts, then exclusive wave offsets
    __shared__ uint32_t s_start[NB];      // digit start inside the tile
    __shared__ uint32_t s_wsum[kRW];
    __shared__ int64_t s_run[NB];         // global cursor of each digit
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wave = t / kWave;
    const bool owner = t < NB;            // thread t owns digit t
    const uint64_t lt_mask = (uint64_t(1) << lane) - 1;
    if (owner) s_run[t] = offsets[int64_t(t) * nblocks + blockIdx.x];
    const int64_t b0 = int64_t(blockIdx.x) * chunk;
    const int64_t b1 = b0 + chunk < n ? b0 + chunk : n;
    for (int64_t base = b0; base < b1; base += kTileK) {
        if (owner) {
#pragma unroll
            for (int w = 0; w < kRW; ++w) s_wcnt[w][t] = 0;
        }
        const int64_t wbase = base + int64_t(wave) * (kKPT * kWave);
        uint64_t key[kKPT];
        uint32_t val[kKPT];
#pragma unroll
        for (int j = 0; j < kKPT; ++j) {
            const int64_t i = wbase + j * kWave + lane;
            key[j] = i < b1 ? kin[i] : 0;
            val[j] = i < b1 ? vin[i] : 0;
        }
        __syncthreads();  // counts cleared; the previous tile's reads are done
        uint32_t rank[kKPT];
#pragma unroll
        for (int j = 0; j < kKPT; ++j) {
            const bool valid = wbase + j * kWave + lane < b1;
            const uint32_t d = digit_of<B>(key[j], kmask, shift);
            const uint64_t peers = digit_peers<B>(d, __ballot(valid));
            const uint32_t before = s_wcnt[wave][d];
            rank[j] = before + uint32_t(__popcll(peers & lt_mask));
            // one lane per digit group bumps the wave's count; the reads
            // above precede this store in the wave's instruction order
            if (valid && (peers & lt_mask) == 0) s_wcnt[wave][d] = before + uint32_t(__popcll(peers));
        }
        __syncthreads();
        // thread t owns digit t: exclusive offsets over the waves, the tile
        // total, and the digits' exclusive scan across the workgroup
        uint32_t tot = 0;
        if (owner) {
#pragma unroll
            for (int w = 0; w < kRW; ++w) {
                const uint32_t c = s_wcnt[w][t];
                s_wcnt[w][t] = tot;
                tot += c;
            }
        }
        uint32_t inc = tot;
#pragma unroll
        for (int dlt = 1; dlt < kWave; dlt <<= 1) {
            const uint32_t y = __shfl_up(inc, dlt, kWave);
            if (lane >= dlt) inc += y;
        }
        if (lane == kWave - 1) s_wsum[wave] = inc;
        __syncthreads();
        uint32_t wpre = 0;
#pragma unroll
        for (int w = 0; w < kRW; ++w) wpre += w < wave ? s_wsum[w] : 0u;
        const uint32_t my_start = wpre + inc - tot;
        if (owner) s_start[t] = my_start;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kKPT; ++j) {
            const bool valid = wbase + j * kWave + lane < b1;
            if (valid) {
                const uint32_t d = digit_of<B>(key[j], kmask, shift);
                const uint32_t at = s_start[d] + s_wcnt[wave][d] + rank[j];
                s_key[at] = key[j];
                s_val[at] = val[j];
            }
        }
        __syncthreads();
        const int64_t left = b1 - base;
        const int cnt = left < kTileK ? int(left) : kTileK;
#pragma unroll
        for (int j = 0; j < kKPT; ++j) {
            const int at = j * kRT + t;
            if (at < cnt) {
                const uint64_t k = s_key[at];
                const uint32_t d = digit_of<B>(k, kmask, shift);
                const int64_t pos = s_run[d] + (at - int64_t(s_start[d]));
                kout[pos] = k;
                vout[pos] = s_val[at];
            }
        }
        __syncthreads();  // s_run, s_start and the staged tile are read above
        if (owner) s_run[t] += tot;
    }
}

// ---------------------------------------------------------------- top-k
constexpr int kSelBits = 12;
constexpr int kSelBins = 1 << kSelBits;
constexpr int kSelMax = 8192;   // candidates sorted by one workgroup in LDS

struct KeyIdx {
    uint64_t key;
    uint64_t idx;  // input index (stable tie order)
};

// The selection digit of a candidate: kSelBits bits of the key (field 0)
// or of the input index (field 1, used once every key bit is fixed).
BOLAP_DEV uint32_t sel_digit(const KeyIdx& e, int field, int shift) {
    return uint32_t((field ? e.idx : e.key) >> shift) & (kSelBins - 1);
}

// Histogram of one digit among the candidates, added into the global
// histogram `hist` (zero on entry) with one atomic per non-empty LDS bin.
// `cand` may be null (then the candidates are the whole input `keys`,
// tagged by position).
__global__ __launch_bounds__(kRT) void sel_hist_k(const uint64_t* __restrict__ keys,
                                                  const KeyIdx* __restrict__ cand, int64_t n,
                                                  int field, int shift,
                                                  uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kSelBins];
    for (int b = threadIdx.x; b < kSelBins; b += kRT) h[b] = 0;
    __syncthreads();
    for (int64_t i = int64_t(blockIdx.x) * kRT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kRT) {
        const KeyIdx e = cand ? cand[i] : KeyIdx{keys[i], uint64_t(i)};
        atomicAdd(&h[sel_digit(e, field, shift)], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kSelBins; b += kRT)
        if (h[b]) atomicAdd(hist + b, h[b]);
}

// Finds the bucket b* of the global histogram holding the need-th smallest
// candidate: info = {b*, count below b*, count in b*}; then clears the
// histogram for the next digit.
__global__ __launch_bounds__(1024) void sel_pick_k(uint32_t* __restrict__ hist, int64_t need,
                                                   int64_t* __restrict__ info) {
    __shared__ uint32_t tot[kSelBins];
    for (int b = threadIdx.x; b < kSelBins; b += 1024) {
        tot[b] = hist[b];
        hist[b] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = 0;
        int b = 0;
        for (; b < kSelBins; ++b) {
            if (run + tot[b] >= need) break;
            run += tot[b];
        }
        if (b == kSelBins) b = kSelBins - 1;
        info[0] = b;
        info[1] = run;
        info[2] = tot[b];
    }
}

// Appends candidates below b* to `sure` and those in b* to `next`.
__global__ __launch_bounds__(kRT) void sel_split_k(const uint64_t* __restrict__ keys,
                                                   const KeyIdx* __restrict__ cand, int64_t n,
                                                   int field, int shift,
                                                   const int64_t* __restrict__ info,
                                                   KeyIdx* __restrict__ sure,
                                                   unsigned long long* __restrict__ n_sure,
                                                   KeyIdx* __restrict__ next,
                                                   unsigned long long* __restrict__ n_next) {
    const uint32_t bstar = uint32_t(info[0]);
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t lt_mask = (uint64_t(1) << lane) - 1;
    for (int64_t base = int64_t(blockIdx.x) * kRT; base < n; base += int64_t(gridDim.x) * kRT) {
        const int64_t i = base + threadIdx.x;
        KeyIdx e{0, 0};
        if (i < n) e = cand ? cand[i] : KeyIdx{keys[i], uint64_t(i)};
        const uint32_t d = sel_digit(e, field, shift);
        const bool lo = i < n && d < bstar;
        const bool eq = i < n && d == bstar;
        // wave-aggregated appends: one atomic per wave and target
        const uint64_t mlo = __ballot(lo), meq = __ballot(eq);
        unsigned long long blo = 0, beq = 0;
        if (lane == 0) {
            if (mlo) blo = atomicAdd(n_sure, (unsigned long long)__popcll(mlo));
            if (meq) beq = atomicAdd(n_next, (unsigned long long)__popcll(meq));
        }
        blo = __shfl(blo, 0, kWave);
        beq = __shfl(beq, 0, kWave);
        if (lo) sure[blo + __popcll(mlo & lt_mask)] = e;
        if (eq) next[beq + __popcll(meq & lt_mask)] = e;
    }
}

// One workgroup: bitonic sort of m <= kSelMax (key, idx) pairs in LDS,
// writes the first k keys and their values (vals[idx]).
__global__ __launch_bounds__(1024) void sel_sort_k(const KeyIdx* __restrict__ a, int64_t m,
                                                   int32_t k, const uint32_t* __restrict__ vals,
                                                   uint64_t* __restrict__ kout,
                                                   uint32_t* __restrict__ vout) {
    extern __shared__ KeyIdx s[];
    int p = 1;
    while (p < m) p <<= 1;
    for (int i = threadIdx.x; i < p; i += 1024)
        s[i] = i < m ? a[i] : KeyIdx{~uint64_t(0), ~uint64_t(0)};
    __syncthreads();
    for (int size = 2; size <= p; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < p / 2; t += 1024) {
                const int lo = 2 * stride * (t / stride) + (t % stride);
                const int hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const KeyIdx x = s[lo], y = s[hi];
                const bool gt = x.key > y.key || (x.key == y.key && x.idx > y.idx);
                if (gt == up) {
                    s[lo] = y;
                    s[hi] = x;
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < k && i < m; i += 1024) {
        kout[i] = s[i].key;
        if (vout) vout[i] = vals ? vals[s[i].idx] : uint32_t(s[i].idx);
    }
}

__global__ __launch_bounds__(kRT) void sel_tag_k(const uint64_t* __restrict__ keys, int64_t n,
                                                 KeyIdx* __restrict__ dst) {
    for (int64_t i = int64_t(blockIdx.x) * kRT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kRT)
        dst[i] = KeyIdx{keys[i], uint64_t(i)};
}

__global__ __launch_bounds__(kRT) void sel_concat_k(const KeyIdx* __restrict__ src, int64_t n,
                                                    KeyIdx* __restrict__ dst) {
    for (int64_t i = int64_t(blockIdx.x) * kRT + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kRT)
        dst[i] = src[i];
}

__global__ void zero_u64_k(unsigned long long* p, int n) {
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < n;
         i += int(gridDim.x * blockDim.x))
        p[i] = 0;
}

// One LSD pass on B-bit digits: per-workgroup digit counts, their
// exclusive scan in (digit, workgroup) order, then the stable scatter.
template <int B>
int radix_pass(bolap_ctx* ctx, const uint64_t* sk, const uint32_t* sv, int64_t n, int64_t chunk,
               uint64_t kmask, int shift, int64_t nblocks, int64_t* hist, uint64_t* dk,
               uint32_t* dv, void* scan_tmp) {
    hipLaunchKernelGGL(rs_hist_k<B>, dim3(unsigned(nblocks)), dim3(kHT), 0, ctx->stream, sk, n,
                       chunk, kmask, shift, int(nblocks), hist);
    BOLAP_TRY(launch_check("rs_hist_k"));
    BOLAP_TRY(exclusive_scan_i64(ctx, hist, int64_t(1 << B) * nblocks, scan_tmp));
    hipLaunchKernelGGL(rs_scatter_k<B>, dim3(unsigned(nblocks)), dim3(kRT), 0, ctx->stream, sk, sv,
                       n, chunk, kmask, shift, int(nblocks), hist, dk, dv);
    return launch_check("rs_scatter_k");
}

}  // namespace

size_t radix_sort_scratch_bytes(bolap_ctx* ctx, int64_t n) {
    (void)ctx;
    const int64_t nb = 1024;
    WsArena w;
    w.plan(size_t(n) * 8);
    w.plan(size_t(n) * 4);
    w.plan(size_t(kBins) * nb * 8);
    w.plan(scan_scratch_bytes(int64_t(kBins) * nb));
    return w.off;
}

int radix_sort_pairs(bolap_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int key_bits,
                     void* scratch, int* passes_out) {
    if (passes_out) *passes_out = 0;
    if (n < 2 || key_bits <= 0) return BOLAP_OK;
    if (key_bits > 64) key_bits = 64;
    // Contiguous chunks of whole tiles per workgroup keep the scatter
    // stable; two workgroups per CU fit the scatter's 55 KB of LDS, so the
    // grid is at most that many (one round of workgroups).
    int64_t nblocks = (n + kTileK - 1) / kTileK;
    const int64_t max_blocks = std::min<int64_t>(1024, int64_t(ctx->num_cus) * 2);
    if (nblocks > max_blocks) nblocks = max_blocks;
    if (nblocks < 1) nblocks = 1;
    int64_t chunk = (n + nblocks - 1) / nblocks;
    chunk = (chunk + kTileK - 1) / kTileK * kTileK;
    nblocks = (n + chunk - 1) / chunk;
    WsArena w;
    const size_t o_k = w.plan(size_t(n) * 8);
    const size_t o_v = w.plan(size_t(n) * 4);
    const size_t o_h = w.plan(size_t(kBins) * 1024 * 8);
    const size_t o_s = w.plan(scan_scratch_bytes(int64_t(kBins) * 1024));
    w.base = static_cast<char*>(scratch);
    uint64_t* tk = w.at<uint64_t>(o_k);
    uint32_t* tv = w.at<uint32_t>(o_v);
    int64_t* hist = w.at<int64_t>(o_h);
    const uint64_t kmask = key_bits >= 64 ? ~uint64_t(0) : ((uint64_t(1) << key_bits) - 1);
    uint64_t* sk = keys;
    uint32_t* sv = vals;
    uint64_t* dk = tk;
    uint32_t* dv = tv;
    // the fewest 8-bit passes, then the narrowest digit that still covers
    // the key in that many passes (5 bits at least); knob radix.digit_bits
    // (5..8, set by the autotuner) fixes the width instead
    const int npass = (key_bits + 7) / 8;
    int dbits = std::max(5, (key_bits + npass - 1) / npass);
    const int64_t forced = ctx->knob("radix.digit_bits", 0);
    if (forced >= 5 && forced <= 8) dbits = int(forced);
    int passes = 0;
    for (int shift = 0; shift < key_bits; shift += dbits, ++passes) {
        switch (dbits) {
            case 5: BOLAP_TRY(radix_pass<5>(ctx, sk, sv, n, chunk, kmask, shift, nblocks, hist, dk, dv, w.at<void>(o_s))); break;
            case 6: BOLAP_TRY(radix_pass<6>(ctx, sk, sv, n, chunk, kmask, shift, nblocks, hist, dk, dv, w.at<void>(o_s))); break;
            case 7: BOLAP_TRY(radix_pass<7>(ctx, sk, sv, n, chunk, kmask, shift, nblocks, hist, dk, dv, w.at<void>(o_s))); break;
            default: BOLAP_TRY(radix_pass<8>(ctx, sk, sv, n, chunk, kmask, shift, nblocks, hist, dk, dv, w.at<void>(o_s)));
        }
        std::swap(sk, dk);
        std::swap(sv, dv);
    }
    if (sk != keys) {
        BOLAP_HIP_TRY(hipMemcpyAsync(keys, sk, size_t(n) * 8, hipMemcpyDeviceToDevice, ctx->stream));
        BOLAP_HIP_TRY(hipMemcpyAsync(vals, sv, size_t(n) * 4, hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (passes_out) *passes_out = passes;
    return BOLAP_OK;
}

size_t topk_scratch_bytes(int64_t n) {
    WsArena w;
    w.plan(size_t(kSelBins) * 4);
    w.plan(64);
    w.plan(64);
    w.plan(size_t(n) * sizeof(KeyIdx));  // sure
    w.plan(size_t(n) * sizeof(KeyIdx));  // cand A
    w.plan(size_t(n) * sizeof(KeyIdx));  // cand B
    return w.off;
}

int topk_pairs(bolap_ctx* ctx, const uint64_t* keys, const uint32_t* vals, int64_t n, int32_t k,
               uint64_t* kout, uint32_t* vout, void* scratch) {
    if (n <= 0 || k <= 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(k <= kSelMax, "topk: k = %d exceeds %d", k, kSelMax);
    WsArena w;
    const size_t o_h = w.plan(size_t(kSelBins) * 4);
    const size_t o_info = w.plan(64);
    const size_t o_cnt = w.plan(64);
    const size_t o_sure = w.plan(size_t(n) * sizeof(KeyIdx));
    const size_t o_ca = w.plan(size_t(n) * sizeof(KeyIdx));
    const size_t o_cb = w.plan(size_t(n) * sizeof(KeyIdx));
    w.base = static_cast<char*>(scratch);
    auto* hist = w.at<uint32_t>(o_h);
    auto* info = w.at<int64_t>(o_info);
    auto* cnt = w.at<unsigned long long>(o_cnt);  // [0] sure, [1] next
    KeyIdx* sure = w.at<KeyIdx>(o_sure);
    KeyIdx* ca = w.at<KeyIdx>(o_ca);
    KeyIdx* cb = w.at<KeyIdx>(o_cb);
    BOLAP_TRY(pinned_reserve(ctx, 64));
    auto* host = static_cast<int64_t*>(ctx->pinned);

    const KeyIdx* cand = nullptr;  // null: the whole input
    int64_t m = n;                 // candidates still undecided
    int64_t n_sure = 0;            // candidates certainly in the top k
    int64_t need = k;              // how many of the undecided are in
    // Digit schedule: key bits from the top (the last digit overlaps bits
    // already fixed, which is harmless), then index bits for ties.
    const int sched_field[8] = {0, 0, 0, 0, 0, 0, 1, 1};
    const int sched_shift[8] = {52, 40, 28, 16, 4, 0, 20, 8};
    hipLaunchKernelGGL(zero_u64_k, dim3(1), dim3(64), 0, ctx->stream, cnt, 2);
    hipLaunchKernelGGL(zero_u64_k, dim3(1), dim3(1024), 0, ctx->stream,
                       reinterpret_cast<unsigned long long*>(hist), kSelBins / 2);
    for (int r = 0; r < 9 && n_sure + m > kSelMax && m > need; ++r) {
        const int field = r < 8 ? sched_field[r] : 1;
        const int sh = r < 8 ? sched_shift[r] : 0;
        int64_t hb = (m + kRT * 8 - 1) / (kRT * 8);
        if (hb > int64_t(ctx->num_cus) * 2) hb = int64_t(ctx->num_cus) * 2;
        if (hb < 1) hb = 1;
        hipLaunchKernelGGL(sel_hist_k, dim3(unsigned(hb)), dim3(kRT), 0, ctx->stream, keys, cand,
                           m, field, sh, hist);
        hipLaunchKernelGGL(sel_pick_k, dim3(1), dim3(1024), 0, ctx->stream, hist, need, info);
        KeyIdx* next = (cand == ca) ? cb : ca;
        hipLaunchKernelGGL(zero_u64_k, dim3(1), dim3(64), 0, ctx->stream, cnt + 1, 1);
        int64_t sb = (m + kRT - 1) / kRT;
        if (sb > int64_t(ctx->num_cus) * 4) sb = int64_t(ctx->num_cus) * 4;
        hipLaunchKernelGGL(sel_split_k, dim3(unsigned(sb)), dim3(kRT), 0, ctx->stream, keys, cand,
                           m, field, sh, info, sure, cnt, next, cnt + 1);
        BOLAP_TRY(launch_check("topk rounds"));
        BOLAP_HIP_TRY(hipMemcpyAsync(host, cnt, 16, hipMemcpyDeviceToHost, ctx->stream));
        BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
        n_sure = host[0];
        m = host[1];
        need = k - n_sure;
        cand = next;
    }
    // Final candidates: the certain ones followed by the undecided ones
    // (all of them are needed, or few enough to sort).
    const KeyIdx* fin = sure;
    int64_t nf = n_sure + m;
    if (cand == nullptr) {
        hipLaunchKernelGGL(sel_tag_k, dim3(unsigned((n + kRT - 1) / kRT)), dim3(kRT), 0,
                           ctx->stream, keys, n, ca);
        BOLAP_TRY(launch_check("sel_tag_k"));
        fin = ca;
        nf = n;
    } else if (m > 0) {
        hipLaunchKernelGGL(sel_concat_k, dim3(unsigned((m + kRT - 1) / kRT)), dim3(kRT), 0,
                           ctx->stream, cand, m, sure + n_sure);
        BOLAP_TRY(launch_check("sel_concat_k"));
    }
    if (nf > kSelMax) {
        set_error("topk: %lld candidates left after every digit", (long long)nf);
        return BOLAP_ERR_CAPACITY;
    }
    int p = 1;
    while (p < nf) p <<= 1;
    const size_t lds = size_t(p) * sizeof(KeyIdx);
    BOLAP_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(sel_sort_k),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      int(size_t(kSelMax) * sizeof(KeyIdx))));
    hipLaunchKernelGGL(sel_sort_k, dim3(1), dim3(1024), lds, ctx->stream, fin, nf, k, vals, kout,
                       vout);
    return launch_check("sel_sort_k");
}

}  // namespace bolap

using namespace bolap;

extern "C" {

int bolap_radix_sort_u64(bolap_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n,
                         int32_t key_bits) {
    BOLAP_ARG_CHECK(ctx && n >= 0, "bolap_radix_sort_u64: bad arguments");
    BOLAP_ARG_CHECK(n < (int64_t(1) << 32), "bolap_radix_sort_u64: n >= 2^32");
    if (n < 2) return BOLAP_OK;
    BOLAP_ARG_CHECK(keys && vals, "bolap_radix_sort_u64: null array");
    BOLAP_ARG_CHECK(key_bits >= 0 && key_bits <= 64, "bolap_radix_sort_u64: key_bits %d",
                    key_bits);
    BOLAP_TRY(ws_reserve(ctx, radix_sort_scratch_bytes(ctx, n)));
    TraceScope ts(ctx, "radix_sort");
    return radix_sort_pairs(ctx, keys, vals, n, key_bits, ctx->ws, nullptr);
}

int bolap_topk_u64(bolap_ctx* ctx, const uint64_t* keys, const uint32_t* vals, int64_t n,
                   int32_t k, uint64_t* keys_out, uint32_t* vals_out) {
    BOLAP_ARG_CHECK(ctx && n >= 0 && k >= 0, "bolap_topk_u64: bad arguments");
    BOLAP_ARG_CHECK(n < (int64_t(1) << 32), "bolap_topk_u64: n >= 2^32");
    if (n == 0 || k == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(keys && keys_out, "bolap_topk_u64: null array");
    BOLAP_TRY(ws_reserve(ctx, topk_scratch_bytes(n)));
    TraceScope ts(ctx, "topk");
    return topk_pairs(ctx, keys, vals, n, k, keys_out, vals_out, ctx->ws);
}

}  // extern "C"
