// runtime.hip — context, workspace and error plumbing of libbolap.
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <tuple>

#include "common.hpp"

namespace bolap {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int ws_reserve(bolap_ctx* ctx, size_t bytes, size_t keep) {
    if (bytes <= ctx->ws_size) return BOLAP_OK;
    size_t want = ctx->ws_size ? ctx->ws_size : (size_t(64) << 20);
    while (want < bytes) want *= 2;
    void* fresh = nullptr;
    BOLAP_HIP_TRY(hipMalloc(&fresh, want));
    if (ctx->ws && keep > 0) {
        const size_t n = keep < ctx->ws_size ? keep : ctx->ws_size;
        BOLAP_HIP_TRY(hipMemcpyAsync(fresh, ctx->ws, n, hipMemcpyDeviceToDevice, ctx->stream));
    }
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->ws) BOLAP_HIP_TRY(hipFree(ctx->ws));
    ctx->ws = fresh;
    ctx->ws_size = want;
    return BOLAP_OK;
}

int kernel_occupancy(bolap_ctx* ctx, const void* kern, int threads, size_t lds, int* per_cu) {
    static std::mutex mu;
    static std::map<std::tuple<int, const void*, int, size_t>, int> occ;
    static std::map<std::pair<int, const void*>, size_t> lds_limit;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_tuple(ctx->device, kern, threads, lds);
    const auto it = occ.find(key);
    if (it != occ.end()) {
        *per_cu = it->second;
        return BOLAP_OK;
    }
    // the limit only grows, so a launch with less LDS stays valid
    size_t& cur = lds_limit[std::make_pair(ctx->device, kern)];
    if (lds > cur) {
        BOLAP_HIP_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          int(lds)));
        cur = lds;
    }
    int n = 0;
    BOLAP_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, threads, lds));
    occ.emplace(key, n);
    *per_cu = n;
    return BOLAP_OK;
}

int pinned_reserve(bolap_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->pinned_size) return BOLAP_OK;
    size_t want = ctx->pinned_size ? ctx->pinned_size : (size_t(1) << 20);
    while (want < bytes) want *= 2;
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->pinned) BOLAP_HIP_TRY(hipHostFree(ctx->pinned));
    ctx->pinned = nullptr;
    ctx->pinned_size = 0;
    ctx->pinned_dev = nullptr;
    BOLAP_HIP_TRY(hipHostMalloc(&ctx->pinned, want, hipHostMallocMapped | hipHostMallocCoherent));
    ctx->pinned_size = want;
    BOLAP_HIP_TRY(hipHostGetDevicePointer(&ctx->pinned_dev, ctx->pinned, 0));
    return BOLAP_OK;
}

int wait_pinned_seq(bolap_ctx* ctx, size_t slot, int64_t seq) {
    const int64_t* flag = static_cast<const int64_t*>(ctx->pinned) + slot;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t spin = 0;; ++spin) {
        if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq) return BOLAP_OK;
        __builtin_ia32_pause();
        if ((spin & 1023u) == 1023u &&
            std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1))
            break;
    }
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq) return BOLAP_OK;
    set_error("result row %lld was not published", (long long)seq);
    return BOLAP_ERR_HIP;
}

int launch_check(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return BOLAP_ERR_HIP;
    }
    return BOLAP_OK;
}

static hipEvent_t take_event(bolap_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TraceScope::TraceScope(bolap_ctx* c, const char* name) : ctx(c) {
    if (!ctx || !ctx->trace) return;
    bolap_trace_rec r;
    std::snprintf(r.name, sizeof(r.name), "%s", name);
    r.start = take_event(ctx);
    r.stop = take_event(ctx);
    if (!r.start || !r.stop) return;
    (void)hipEventRecord(r.start, ctx->stream);
    idx = int(ctx->trace_recs.size());
    ctx->trace_recs.push_back(r);
}

TraceScope::~TraceScope() {
    if (idx >= 0) (void)hipEventRecord(ctx->trace_recs[size_t(idx)].stop, ctx->stream);
}

void trace_count(bolap_ctx* ctx, const char* name, int64_t n) {
    if (!ctx || !ctx->trace) return;
    bolap_trace_rec r;
    std::snprintf(r.name, sizeof(r.name), "%s", name);
    r.count = n;
    ctx->trace_recs.push_back(r);
}

}  // namespace bolap

using namespace bolap;

extern "C" {

const char* bolap_last_error(void) { return g_err; }

int bolap_abi_version(void) { return BOLAP_ABI_VERSION; }

int bolap_ctx_create(int device, void* stream, bolap_ctx** out) {
    BOLAP_ARG_CHECK(out != nullptr, "bolap_ctx_create: out is null");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        set_error("bolap_ctx_create: no HIP device visible");
        return BOLAP_ERR_NO_DEVICE;
    }
    BOLAP_ARG_CHECK(device >= 0 && device < n, "bolap_ctx_create: device %d of %d", device, n);
    BOLAP_HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    BOLAP_HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("bolap_ctx_create: device %d is %s; this library is built for gfx950",
                  device, prop.gcnArchName);
        return BOLAP_ERR_UNSUPPORTED;
    }
    auto* ctx = new bolap_ctx();
    ctx->device = device;
    ctx->stream = static_cast<hipStream_t>(stream);
    ctx->num_cus = prop.multiProcessorCount;
    ctx->lds_per_cu = static_cast<int>(prop.maxSharedMemoryPerMultiProcessor);
    ctx->hbm_bytes = static_cast<int64_t>(prop.totalGlobalMem);
    std::strncpy(ctx->arch, prop.gcnArchName, sizeof(ctx->arch) - 1);
    hipError_t e = hipMalloc(&ctx->counters, sizeof(unsigned) * bolap_ctx::kNumCounters);
    if (e == hipSuccess)
        e = hipMemset(ctx->counters, 0, sizeof(unsigned) * bolap_ctx::kNumCounters);
    if (e == hipSuccess) e = hipMalloc(&ctx->result_dev, bolap_ctx::kResultBytes);
    if (e == hipSuccess) e = hipMalloc(&ctx->join_state, bolap_ctx::kJoinStateBytes);
    if (e == hipSuccess) e = hipMemset(ctx->join_state, 0, bolap_ctx::kJoinStateBytes);
    int setup = BOLAP_OK;
    if (e == hipSuccess) setup = noalloc_setup(device);
    if (e != hipSuccess || setup != BOLAP_OK) {
        if (e != hipSuccess) set_error("bolap_ctx_create: %s", hipGetErrorString(e));
        if (ctx->counters) (void)hipFree(ctx->counters);
        if (ctx->result_dev) (void)hipFree(ctx->result_dev);
        if (ctx->join_state) (void)hipFree(ctx->join_state);
        delete ctx;
        return e != hipSuccess ? BOLAP_ERR_HIP : setup;
    }
    *out = ctx;
    return BOLAP_OK;
}

int bolap_ctx_destroy(bolap_ctx* ctx) {
    if (!ctx) return BOLAP_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ws) (void)hipFree(ctx->ws);
    if (ctx->ceb_exec) (void)hipGraphExecDestroy(ctx->ceb_exec);
    if (ctx->ceb_graph) (void)hipGraphDestroy(ctx->ceb_graph);
    if (ctx->ceb_buf) (void)hipFree(ctx->ceb_buf);
    if (ctx->capture_stream) (void)hipStreamDestroy(ctx->capture_stream);
    if (ctx->counters) (void)hipFree(ctx->counters);
    if (ctx->result_dev) (void)hipFree(ctx->result_dev);
    if (ctx->join_state) (void)hipFree(ctx->join_state);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    for (auto& r : ctx->trace_recs) {
        if (r.start) (void)hipEventDestroy(r.start);
        if (r.stop) (void)hipEventDestroy(r.stop);
    }
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    delete ctx;
    return BOLAP_OK;
}

int bolap_ctx_set_trace(bolap_ctx* ctx, int enabled) {
    BOLAP_ARG_CHECK(ctx != nullptr, "bolap_ctx_set_trace: null ctx");
    ctx->trace = enabled != 0;
    return BOLAP_OK;
}

int bolap_ctx_read_trace(bolap_ctx* ctx, bolap_trace_entry* out, int32_t max_entries,
                         int32_t* n_out) {
    BOLAP_ARG_CHECK(ctx && n_out && (out || max_entries == 0),
                    "bolap_ctx_read_trace: null argument");
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    int32_t n = 0;
    for (auto& r : ctx->trace_recs) {
        if (n < max_entries) {
            bolap_trace_entry& e = out[n];
            std::memset(&e, 0, sizeof(e));
            std::snprintf(e.name, sizeof(e.name), "%s", r.name);
            e.count = r.count;
            e.ms = -1.0;
            if (r.start) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, r.start, r.stop) == hipSuccess) e.ms = ms;
            }
        }
        ++n;
        if (r.start) ctx->event_pool.push_back(r.start);
        if (r.stop) ctx->event_pool.push_back(r.stop);
    }
    ctx->trace_recs.clear();
    *n_out = n;
    return BOLAP_OK;
}

int bolap_tune_set(bolap_ctx* ctx, const char* knob, int64_t value) {
    BOLAP_ARG_CHECK(ctx && knob && knob[0], "bolap_tune_set: null argument");
    ctx->knobs[knob] = value;
    return BOLAP_OK;
}

int bolap_tune_get(bolap_ctx* ctx, const char* knob, int64_t* value) {
    BOLAP_ARG_CHECK(ctx && knob && value, "bolap_tune_get: null argument");
    auto it = ctx->knobs.find(knob);
    *value = it == ctx->knobs.end() ? 0 : it->second;
    return BOLAP_OK;
}

int bolap_ctx_sync(bolap_ctx* ctx) {
    BOLAP_ARG_CHECK(ctx != nullptr, "bolap_ctx_sync: null ctx");
    BOLAP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return BOLAP_OK;
}

int bolap_ctx_device_info(bolap_ctx* ctx, bolap_device_info* out) {
    BOLAP_ARG_CHECK(ctx && out, "bolap_ctx_device_info: null argument");
    out->device = ctx->device;
    out->compute_units = ctx->num_cus;
    out->lds_bytes_per_cu = ctx->lds_per_cu;
    out->wavefront = 64;
    out->hbm_bytes = ctx->hbm_bytes;
    std::memset(out->arch, 0, sizeof(out->arch));
    std::strncpy(out->arch, ctx->arch, sizeof(out->arch) - 1);
    return BOLAP_OK;
}

}  // extern "C"
