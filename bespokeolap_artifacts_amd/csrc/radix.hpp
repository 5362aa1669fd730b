// radix.hpp — device radix sort and top-k used by the pipelines.
#pragma once

#include "common.hpp"

namespace bolap {

// Stable LSD sort of n (key, value) pairs on the low key_bits key bits, in
// place; scratch must hold radix_sort_scratch_bytes(n).
size_t radix_sort_scratch_bytes(bolap_ctx* ctx, int64_t n);
int radix_sort_pairs(bolap_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int key_bits,
                     void* scratch, int* passes_out);

// k smallest keys ascending (ties in input order) with vals[index] (or the
// input index when vals is null); scratch holds topk_scratch_bytes(n).
size_t topk_scratch_bytes(int64_t n);
int topk_pairs(bolap_ctx* ctx, const uint64_t* keys, const uint32_t* vals, int64_t n, int32_t k,
               uint64_t* kout, uint32_t* vout, void* scratch);

}  // namespace bolap
