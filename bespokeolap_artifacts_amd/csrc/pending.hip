// pending.hip — entry points declared in bolap.h whose kernels are not built yet.
#include "common.hpp"

using namespace bolap;

extern "C" {

int bolap_chain_topk(bolap_ctx*, const bolap_chain_args*, int32_t, bolap_group_row*, int32_t*,
                     bolap_chain_stats*) {
    set_error("bolap_chain_topk: not built yet");
    return BOLAP_ERR_UNSUPPORTED;
}
int bolap_chain_groups_sorted(bolap_ctx*, const bolap_chain_args*, bolap_group_row*, int64_t,
                              int64_t*, bolap_chain_stats*) {
    set_error("bolap_chain_groups_sorted: not built yet");
    return BOLAP_ERR_UNSUPPORTED;
}
int bolap_radix_sort_u64(bolap_ctx*, uint64_t*, uint32_t*, int64_t, int32_t) {
    set_error("bolap_radix_sort_u64: not built yet");
    return BOLAP_ERR_UNSUPPORTED;
}
int bolap_topk_u64(bolap_ctx*, const uint64_t*, const uint32_t*, int64_t, int32_t, uint64_t*,
                   uint32_t*) {
    set_error("bolap_topk_u64: not built yet");
    return BOLAP_ERR_UNSUPPORTED;
}

}  // extern "C"
