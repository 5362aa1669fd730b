// prims.hpp — device-wide primitives shared by the pipelines.
#pragma once

#include "common.hpp"

namespace bolap {

// Exclusive prefix sum of n int64 values in place (three launches on the
// context stream).  Scratch: scan_scratch_bytes(n) bytes at `tmp`.
size_t scan_scratch_bytes(int64_t n);
int exclusive_scan_i64(bolap_ctx* ctx, int64_t* data, int64_t n, void* tmp);

// Counter-based hash of (seed, stream, row): SplitMix64 output function
// (Steele, Lea & Flood 2014) applied twice.
BOLAP_DEV uint64_t gen_hash(uint64_t seed, uint32_t stream, uint64_t row) {
    const uint64_t k = mix64(seed + 0x9e3779b97f4a7c15ULL * (uint64_t(stream) + 1));
    return mix64(k + 0x9e3779b97f4a7c15ULL * (row + 1));
}

// Uniform integer in [lo, hi] from the top 32 hash bits (hi - lo < 2^32).
BOLAP_DEV int64_t gen_uniform(uint64_t h, int64_t lo, int64_t hi) {
    const uint64_t range = uint64_t(hi - lo) + 1;
    return lo + int64_t(((h >> 32) * range) >> 32);
}

}  // namespace bolap
