// fold.hpp — workgroup folds into replicated device statistics.
//
// A statistic is an int64 with kRep replicas: workgroup b folds into replica
// b % kRep, so a replica's workgroups sit on one XCD (workgroups are dealt to
// the XCDs in turn) and no address sees more than (#workgroups / kRep)
// atomics.  Each replica holds its kStatLine statistics in one 128-byte line
// of its own: atomics on one line serialise (roughly 5-10 ns each), so the
// replicas must not share lines.  The host folds the replicas after the copy
// back; the area is kStatWords int64 slots.
#pragma once

#include "common.hpp"

namespace bolap {

constexpr int kRep = 16;
constexpr int kStatLine = 16;                   // statistics per replica (one line)
constexpr int kStatWords = kRep * kStatLine;    // int64 slots of the statistics area

// Slot of statistic idx (< kStatLine) in replica rep.
__host__ __device__ constexpr int stat_slot(int idx, int rep) { return rep * kStatLine + idx; }

BOLAP_DEV int64_t wave_min(int64_t v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int64_t o = __shfl_xor(v, off, kWave);
        v = o < v ? o : v;
    }
    return v;
}

BOLAP_DEV int64_t wave_max(int64_t v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int64_t o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// Workgroup-wide fold of a wave-uniform value into statistic `idx`
// (op 0 add, 1 min, 2 max).  Every thread of the workgroup must call it.
template <int NT>
BOLAP_DEV void block_fold(int64_t* st, int idx, int64_t wave_value, int op) {
    __shared__ int64_t s_v[NT / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    if (lane == 0) s_v[threadIdx.x / kWave] = wave_value;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t v = s_v[0];
        for (int w = 1; w < NT / kWave; ++w) {
            const int64_t x = s_v[w];
            v = op == 0 ? int64_t(uint64_t(v) + uint64_t(x)) : op == 1 ? (x < v ? x : v)
                                                                       : (x > v ? x : v);
        }
        long long* dst = reinterpret_cast<long long*>(st + stat_slot(idx, blockIdx.x % kRep));
        if (op == 0) {
            if (v) atomicAdd(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)v);
        } else if (op == 1) {
            if (v != INT64_MAX) atomicMin(dst, (long long)v);
        } else if (v != INT64_MIN) {
            atomicMax(dst, (long long)v);
        }
    }
    __syncthreads();
}

// Workgroup min/max folded into statistics imin / imax.
template <int NT>
BOLAP_DEV void fold_minmax(int64_t lo, int64_t hi, int64_t* st, int imin, int imax) {
    block_fold<NT>(st, imin, wave_min(lo), 1);
    block_fold<NT>(st, imax, wave_max(hi), 2);
}

}  // namespace bolap
