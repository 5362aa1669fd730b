// q1_scan_agg.hip — fused filter-scan + low-cardinality group-by aggregate.
//
// Replaces the scan loop of run_q1 (bespoke_tpch/query_q1.cpp:143-228).  Per
// row that passes ship_lo <= ship <= ship_hi it adds, into the accumulators
// of group (a * n_b + b):
//   quantity, price, disc_price, disc_price * (100 + tax) (128-bit), discount,
//   and a count
// with disc_price = price * (100 - disc) (declared-width layout) or the
// truncated discounted_price column of the reference layout
// (builder_impl.cpp:523-532).
//
// Design (HBM-bound streaming scan, no MFMA):
//  * Each lane owns "units" of R consecutive rows, R = 16 B / widest column,
//    so the widest column is read with one 128-bit load per unit and every
//    other column with one narrower load of the same R rows; each load
//    instruction of a wave covers a contiguous span (fully coalesced).
//  * Tiles of U units per lane are double-buffered in registers: the next
//    tile's loads are in flight while the current one is aggregated.
//  * The predicate and the group select stay in registers: G group slots
//    (compile-time) of per-lane accumulators updated with selects, so there
//    are no atomics and no LDS traffic in the loop.
//  * DPP wave sums + LDS reduce the block into one partial row; a one-
//    workgroup merge kernel sums the rows.  Integer sums are exact, so
//    results are bit-identical to the sequential reference.
#include <cstring>

#include "common.hpp"

namespace bolap {
namespace {

constexpr int kFields = 8;  // qty, price, disc_price, charge lo, charge hi, disc, count, pad

struct Q1K {
    const void* ship;
    const void* qty;
    const void* price;
    const void* disc_price;
    const void* disc;
    const void* tax;
    const uint8_t* fa;
    const uint8_t* fb;
    int64_t rows;
    int64_t lo, hi;
    int32_t n_b;
    int32_t n_groups;
};

template <int G>
struct Q1Lane {
    int64_t qty[G], price[G], dp[G], disc[G];
    Acc128 ch[G];
    uint32_t cnt[G];
};

// Column types of the two layouts.
struct WideLayout {
    using S = int32_t; using Q = int64_t; using P = int64_t; using DP = int64_t;
    using D = int64_t; using T = int64_t;
    static constexpr bool kHasDP = false;
    static constexpr bool kTwoFlags = true;
    static constexpr int kNativeR = 2;  // 16 B of an int64 column
};
struct RefLayout {
    using S = int16_t; using Q = int16_t; using P = int32_t; using DP = int32_t;
    using D = uint8_t; using T = uint8_t;
    static constexpr bool kHasDP = true;
    static constexpr bool kTwoFlags = false;
    static constexpr int kNativeR = 4;  // 16 B of an int32 column
};

template <class L, int G>
BOLAP_DEV void q1_row(Q1Lane<G>& acc, const Q1K& a, bool valid, int64_t s, int64_t q,
                      int64_t p, int64_t dpc, int64_t d, int64_t t, uint32_t fa, uint32_t fb) {
    const uint32_t g = L::kTwoFlags ? fa * uint32_t(a.n_b) + fb : fa;
    // Codes outside the dictionaries (a >= n_a, b >= n_b) belong to no group.
    const bool in_dict = (L::kTwoFlags ? fb < uint32_t(a.n_b) : true) &&
                         g < uint32_t(a.n_groups);
    const bool pass = valid && s >= a.lo && s <= a.hi && in_dict;
    const int64_t dp = L::kHasDP ? dpc : int64_t(uint64_t(p) * uint64_t(100 - d));
    const int64_t ch = int64_t(uint64_t(dp) * uint64_t(100 + t));
    const int64_t ch_sign = ch >> 63;
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        // A real branch per group (the empty volatile asm keeps the compiler
        // from flattening it into selects): the body runs once under the mask
        // of the lanes in group gi and is skipped when no lane is.
        if (pass && g == uint32_t(gi)) {
            asm volatile("");
            acc.qty[gi] += q;
            acc.price[gi] += p;
            acc.dp[gi] += dp;
            acc.disc[gi] += d;
            acc.cnt[gi] += 1u;
            const uint64_t nlo = acc.ch[gi].lo + uint64_t(ch);
            acc.ch[gi].hi += ch_sign + (nlo < acc.ch[gi].lo ? 1 : 0);
            acc.ch[gi].lo = nlo;
        }
    }
}

// Merge of the per-block partial rows (one workgroup).  Thread t owns rows
// t, t + 256, ... and loads each row whole with independent 16-byte loads,
// accumulating every (group, field) in registers (the charge pair in 128
// bits); DPP wave sums and LDS then combine the threads.  Launched after
// the scan on the same stream, so the partials are visible without fences.
template <int G, int NT>
BOLAP_DEV void q1_merge_rows(const int64_t* __restrict__ partials, int nblocks, int n_groups,
                             bolap_q1_group* out) {
    constexpr int kP = G * kFields;
    constexpr int kWaves256 = NT / kWave;
    int64_t acc[G][kFields];
    Acc128 ch[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int i = 0; i < kFields; ++i) acc[g][i] = 0;
        ch[g] = Acc128{0, 0};
    }
    for (int b = threadIdx.x; b < nblocks; b += NT) {
        const VecT<int64_t, 2>* row = reinterpret_cast<const VecT<int64_t, 2>*>(partials) +
                                      int64_t(b) * (kP / 2);
        VecT<int64_t, 2> v[kP / 2];
#pragma unroll
        for (int j = 0; j < kP / 2; ++j) v[j] = row[j];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int64_t* x = &v[g * kFields / 2].v[0];
            acc[g][0] += x[0];
            acc[g][1] += x[1];
            acc[g][2] += x[2];
            add128(ch[g], Acc128{uint64_t(x[3]), x[4]});
            acc[g][5] += x[5];
            acc[g][6] += x[6];
        }
    }
    __shared__ int64_t s_w[kWaves256][G][kFields];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int64_t q = wave_sum(acc[g][0]);
        const int64_t p = wave_sum(acc[g][1]);
        const int64_t dp = wave_sum(acc[g][2]);
        const Acc128 c = wave_sum128(ch[g]);
        const int64_t d = wave_sum(acc[g][5]);
        const int64_t n = wave_sum(acc[g][6]);
        if (lane == 0) {
            s_w[wave][g][0] = q;
            s_w[wave][g][1] = p;
            s_w[wave][g][2] = dp;
            s_w[wave][g][3] = int64_t(c.lo);
            s_w[wave][g][4] = c.hi;
            s_w[wave][g][5] = d;
            s_w[wave][g][6] = n;
        }
    }
    __syncthreads();
    if (threadIdx.x < G && int(threadIdx.x) < n_groups) {
        const int g = threadIdx.x;
        int64_t t[kFields] = {0, 0, 0, 0, 0, 0, 0, 0};
        Acc128 tc{0, 0};
        for (int w = 0; w < kWaves256; ++w) {
            for (int i = 0; i < kFields; ++i)
                if (i != 3 && i != 4 && i != 7) t[i] += s_w[w][g][i];
            add128(tc, Acc128{uint64_t(s_w[w][g][3]), s_w[w][g][4]});
        }
        bolap_q1_group r;
        r.sum_qty = t[0];
        r.sum_base_price = t[1];
        r.sum_disc_price = t[2];
        r.sum_charge_lo = tc.lo;
        r.sum_charge_hi = tc.hi;
        r.sum_disc = t[5];
        r.count = t[6];
        r.reserved = 0;
        out[g] = r;
    }
}

template <int G>
__global__ __launch_bounds__(256) void q1_merge_k(const int64_t* __restrict__ partials,
                                                  int nblocks, int n_groups,
                                                  bolap_q1_group* out) {
    q1_merge_rows<G, 256>(partials, nblocks, n_groups, out);
}

template <class L, int R, int U, int G, bool kPipe, int kQ1Threads, bool kLastBlock>
__global__ __launch_bounds__(kQ1Threads) void q1_scan_agg_kernel(Q1K a, int64_t* partials,
                                                                  unsigned* counter,
                                                                  bolap_q1_group* out) {
    constexpr int kQ1Waves = kQ1Threads / kWave;
    using S = typename L::S; using Q = typename L::Q; using P = typename L::P;
    using DP = typename L::DP; using D = typename L::D; using T = typename L::T;
    const S* __restrict__ ship = static_cast<const S*>(a.ship);
    const Q* __restrict__ qty = static_cast<const Q*>(a.qty);
    const P* __restrict__ price = static_cast<const P*>(a.price);
    const DP* __restrict__ dpc = static_cast<const DP*>(a.disc_price);
    const D* __restrict__ disc = static_cast<const D*>(a.disc);
    const T* __restrict__ tax = static_cast<const T*>(a.tax);
    const uint8_t* __restrict__ fa = a.fa;
    const uint8_t* __restrict__ fb = a.fb;

    Q1Lane<G> acc;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        acc.qty[g] = acc.price[g] = acc.dp[g] = acc.disc[g] = 0;
        acc.ch[g].lo = 0;
        acc.ch[g].hi = 0;
        acc.cnt[g] = 0;
    }

    // Software-pipelined loop over tiles of U units per lane: the loads of
    // tile i+1 are issued before tile i is consumed (two register buffers),
    // so every wave keeps a tile in flight while it computes.
    const int64_t n_units = a.rows / R;
    const int64_t stride = int64_t(gridDim.x) * kQ1Threads * U;
    struct Tile {
        VecT<S, R> s[U];
        VecT<Q, R> q[U];
        VecT<P, R> p[U];
        VecT<DP, R> dp[U];
        VecT<D, R> d[U];
        VecT<T, R> t[U];
        VecT<uint8_t, R> fa[U], fb[U];
    };
    auto load_tile = [&](Tile& x, int64_t at) {
#pragma unroll
        for (int k = 0; k < U; ++k) {
            int64_t u = at + int64_t(k) * kQ1Threads;
            u = u < n_units ? u : n_units - 1;  // in bounds; masked when consumed
            x.s[k] = load_unit<S, R>(ship, u);
            x.q[k] = load_unit<Q, R>(qty, u);
            x.p[k] = load_unit<P, R>(price, u);
            if constexpr (L::kHasDP) x.dp[k] = load_unit<DP, R>(dpc, u);
            x.d[k] = load_unit<D, R>(disc, u);
            x.t[k] = load_unit<T, R>(tax, u);
            x.fa[k] = load_unit<uint8_t, R>(fa, u);
            if constexpr (L::kTwoFlags) x.fb[k] = load_unit<uint8_t, R>(fb, u);
        }
        asm volatile("" ::: "memory");  // keep the loads ahead of the compute
    };
    auto use_tile = [&](const Tile& x, int64_t at) {
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const bool valid = at + int64_t(k) * kQ1Threads < n_units;
#pragma unroll
            for (int r = 0; r < R; ++r)
                q1_row<L, G>(acc, a, valid, x.s[k].v[r], x.q[k].v[r], x.p[k].v[r],
                             L::kHasDP ? int64_t(x.dp[k].v[r]) : 0, x.d[k].v[r], x.t[k].v[r],
                             x.fa[k].v[r], L::kTwoFlags ? uint32_t(x.fb[k].v[r]) : 0u);
        }
    };
    int64_t base = int64_t(blockIdx.x) * kQ1Threads * U + threadIdx.x;
    if constexpr (!kPipe) {
        for (; base < n_units; base += stride) {
            Tile A;
            load_tile(A, base);
            use_tile(A, base);
        }
    } else if (n_units > 0 && base < n_units) {
        Tile A, B;
        load_tile(A, base);
        for (;;) {
            load_tile(B, base + stride);
            use_tile(A, base);
            base += stride;
            if (base >= n_units) break;
            load_tile(A, base + stride);
            use_tile(B, base);
            base += stride;
            if (base >= n_units) break;
        }
    }
    // Tail rows (rows % R) go to block 0.
    if (blockIdx.x == 0) {
        const int64_t row = n_units * R + threadIdx.x;
        if (threadIdx.x < R && row < a.rows) {
            q1_row<L, G>(acc, a, true, ship[row], qty[row], price[row],
                         L::kHasDP ? int64_t(dpc[row]) : 0, disc[row], tax[row], fa[row],
                         L::kTwoFlags ? uint32_t(fb[row]) : 0u);
        }
    }

    // Block reduction: DPP within each wave, then LDS across waves.
    __shared__ int64_t s_part[kQ1Waves][G][kFields];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int64_t q = wave_sum(acc.qty[g]);
        const int64_t p = wave_sum(acc.price[g]);
        const int64_t dp = wave_sum(acc.dp[g]);
        const int64_t d = wave_sum(acc.disc[g]);
        const int64_t c = int64_t(wave_sum(uint64_t(acc.cnt[g])));
        const Acc128 ch = wave_sum128(acc.ch[g]);
        if (lane == 0) {
            s_part[wave][g][0] = q;
            s_part[wave][g][1] = p;
            s_part[wave][g][2] = dp;
            s_part[wave][g][3] = int64_t(ch.lo);
            s_part[wave][g][4] = ch.hi;
            s_part[wave][g][5] = d;
            s_part[wave][g][6] = c;
            s_part[wave][g][7] = 0;
        }
    }
    __syncthreads();
    // Block partial: row of G*kFields int64 per block.
    constexpr int kP = G * kFields;
    if (threadIdx.x < kP) {
        const int g = threadIdx.x / kFields;
        const int f = threadIdx.x % kFields;
        int64_t v = 0;
        if (f == 3) {
            Acc128 c{0, 0};
            for (int w = 0; w < kQ1Waves; ++w)
                add128(c, Acc128{uint64_t(s_part[w][g][3]), s_part[w][g][4]});
            v = int64_t(c.lo);
            partials[int64_t(blockIdx.x) * kP + threadIdx.x + 1] = c.hi;
        } else if (f != 4) {
            for (int w = 0; w < kQ1Waves; ++w) v += s_part[w][g][f];
        }
        if (f != 4) partials[int64_t(blockIdx.x) * kP + threadIdx.x] = v;
    }
    if constexpr (kLastBlock) {
        if (!last_block_done(counter)) return;
        q1_merge_rows<G, kQ1Threads>(partials, int(gridDim.x), a.n_groups, out);
        if (threadIdx.x == 0) reset_counter(counter);
    }

}

// Kernel variants (knob "q1.variant"): 1..3 plain loop with U = 1, 2, 4
// units per lane; 4..5 software-pipelined loop with U = 1, 2.  All use
// 256-thread workgroups; "q1.blocks_per_cu" caps the persistent grid.
constexpr int kQ1Threads = 256;

template <class L, int R, int G, int U, bool kPipe, bool kLast>
int launch_q1_f(bolap_ctx* ctx, const Q1K& k, bolap_q1_group* out_dev) {
    auto kern = q1_scan_agg_kernel<L, R, U, G, kPipe, kQ1Threads, kLast>;
    const int64_t n_units = k.rows / R;
    const int64_t tile = int64_t(kQ1Threads) * U;
    int64_t blocks = (n_units + tile - 1) / tile;
    int per_cu = 0;
    BOLAP_TRY(kernel_occupancy(ctx, reinterpret_cast<const void*>(kern), kQ1Threads, 0,
                               &per_cu));
    // One workgroup per CU measured fastest (fewer partial rows to merge).
    int64_t want = ctx->knob("q1.blocks_per_cu", 1);
    if (per_cu > 0 && want > per_cu) want = per_cu;
    const int64_t max_blocks = int64_t(ctx->num_cus) * want;
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks < 1) blocks = 1;
    WsArena plan;
    const size_t o_part = plan.plan(size_t(blocks) * G * kFields * sizeof(int64_t));
    BOLAP_TRY(ws_reserve(ctx, plan.off));
    plan.base = static_cast<char*>(ctx->ws);
    {
        TraceScope ts(ctx, "q1_scan");
        hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(kQ1Threads), 0, ctx->stream, k,
                           plan.at<int64_t>(o_part), ctx->counters + 0, out_dev);
    }
    BOLAP_TRY(launch_check("q1_scan_agg_kernel"));
    if (kLast) return BOLAP_OK;
    TraceScope ts(ctx, "q1_merge");
    hipLaunchKernelGGL(q1_merge_k<G>, dim3(1), dim3(256), 0, ctx->stream,
                       plan.at<int64_t>(o_part), int(blocks), k.n_groups, out_dev);
    return launch_check("q1_merge_k");
}

// Finish (knob "q1.finish"): 1 = the last workgroup merges the partials
// (release/acquire on a counter), 2 (default) = a separate one-workgroup
// merge kernel; both measured ~10 us at 256 workgroups on MI355X.
template <class L, int R, int G, int U, bool kPipe>
int launch_q1_v(bolap_ctx* ctx, const Q1K& k, bolap_q1_group* out_dev) {
    if (ctx->knob("q1.finish", 2) == 1) return launch_q1_f<L, R, G, U, kPipe, true>(ctx, k, out_dev);
    return launch_q1_f<L, R, G, U, kPipe, false>(ctx, k, out_dev);
}

template <class L, int R, int G>
int launch_q1_g(bolap_ctx* ctx, const Q1K& k, bolap_q1_group* out_dev) {
    if (R == 1) return launch_q1_v<L, R, G, 2, false>(ctx, k, out_dev);  // unaligned columns
    switch (ctx->knob("q1.variant", 4)) {
        case 1: return launch_q1_v<L, R, G, 1, false>(ctx, k, out_dev);
        case 2: return launch_q1_v<L, R, G, 2, false>(ctx, k, out_dev);
        case 3: return launch_q1_v<L, R, G, 4, false>(ctx, k, out_dev);
        case 4: return launch_q1_v<L, R, G, 1, true>(ctx, k, out_dev);
        default: return launch_q1_v<L, R, G, 2, true>(ctx, k, out_dev);
    }
}

template <class L, int R>
int launch_q1_r(bolap_ctx* ctx, const Q1K& k, bolap_q1_group* out_dev) {
    const int g = k.n_groups;
    if (g <= 2) return launch_q1_g<L, R, 2>(ctx, k, out_dev);
    if (g <= 4) return launch_q1_g<L, R, 4>(ctx, k, out_dev);
    if (g <= 6) return launch_q1_g<L, R, 6>(ctx, k, out_dev);
    if (g <= 8) return launch_q1_g<L, R, 8>(ctx, k, out_dev);
    return launch_q1_g<L, R, 16>(ctx, k, out_dev);
}

int q1_launch(bolap_ctx* ctx, const bolap_q1_args* a, bolap_q1_group* out_dev) {
    BOLAP_ARG_CHECK(ctx && a && out_dev, "bolap_q1_scan_agg: null argument");
    BOLAP_ARG_CHECK(a->rows >= 0, "bolap_q1_scan_agg: rows < 0");
    BOLAP_ARG_CHECK(a->n_a >= 1 && a->n_b >= 1 && a->n_a * a->n_b <= BOLAP_Q1_MAX_GROUPS,
                    "bolap_q1_scan_agg: n_a*n_b must be in 1..%d", BOLAP_Q1_MAX_GROUPS);
    const bool wide = a->layout == BOLAP_Q1_LAYOUT_WIDE;
    BOLAP_ARG_CHECK(wide || a->layout == BOLAP_Q1_LAYOUT_REF,
                    "bolap_q1_scan_agg: unknown layout %d", a->layout);
    Q1K k;
    k.ship = a->ship;
    k.qty = a->qty;
    k.price = a->price;
    k.disc_price = a->disc_price;
    k.disc = a->disc;
    k.tax = a->tax;
    k.fa = a->flag_a;
    k.fb = a->flag_b;
    k.rows = a->rows;
    k.lo = a->ship_lo;
    k.hi = a->ship_hi;
    k.n_b = a->n_b;
    k.n_groups = a->n_a * a->n_b;
    if (a->rows > 0) {
        BOLAP_ARG_CHECK(a->ship && a->qty && a->price && a->disc && a->tax && a->flag_a,
                        "bolap_q1_scan_agg: null column");
        BOLAP_ARG_CHECK(!wide || a->flag_b, "bolap_q1_scan_agg: WIDE layout needs flag_b");
        BOLAP_ARG_CHECK(wide || a->disc_price,
                        "bolap_q1_scan_agg: REF layout needs disc_price");
    }
    if (wide) {
        // 128-bit path needs every column aligned to its R=2 vector width.
        const bool vec = aligned(k.ship, 8) && aligned(k.qty, 16) && aligned(k.price, 16) &&
                         aligned(k.disc, 16) && aligned(k.tax, 16) && aligned(k.fa, 2) &&
                         aligned(k.fb, 2);
        return vec ? launch_q1_r<WideLayout, 2>(ctx, k, out_dev)
                   : launch_q1_r<WideLayout, 1>(ctx, k, out_dev);
    }
    if (!k.fb) k.fb = k.fa;  // unused by the reference layout
    const bool vec = aligned(k.ship, 8) && aligned(k.qty, 8) && aligned(k.price, 16) &&
                     aligned(k.disc_price, 16) && aligned(k.disc, 4) && aligned(k.tax, 4) &&
                     aligned(k.fa, 4);
    return vec ? launch_q1_r<RefLayout, 4>(ctx, k, out_dev)
               : launch_q1_r<RefLayout, 1>(ctx, k, out_dev);
}

}  // namespace
}  // namespace bolap

using namespace bolap;

extern "C" {

int bolap_q1_scan_agg_async(bolap_ctx* ctx, const bolap_q1_args* args,
                            bolap_q1_group* out_dev) {
    return q1_launch(ctx, args, out_dev);
}

int bolap_q1_scan_agg(bolap_ctx* ctx, const bolap_q1_args* args, bolap_q1_group* out) {
    BOLAP_ARG_CHECK(ctx && args && out, "bolap_q1_scan_agg: null argument");
    BOLAP_ARG_CHECK(args->n_a >= 1 && args->n_b >= 1 &&
                        args->n_a * args->n_b <= BOLAP_Q1_MAX_GROUPS,
                    "bolap_q1_scan_agg: n_a*n_b must be in 1..%d", BOLAP_Q1_MAX_GROUPS);
    const size_t bytes = sizeof(bolap_q1_group) * size_t(args->n_a * args->n_b);
    BOLAP_TRY(pinned_reserve(ctx, bytes));
    auto* dev_out = static_cast<bolap_q1_group*>(ctx->result_dev);
    BOLAP_TRY(q1_launch(ctx, args, dev_out));
    BOLAP_HIP_TRY(
        hipMemcpyAsync(ctx->pinned, dev_out, bytes, hipMemcpyDeviceToHost, ctx->stream));
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::memcpy(out, ctx->pinned, bytes);
    return BOLAP_OK;
}

}  // extern "C"
