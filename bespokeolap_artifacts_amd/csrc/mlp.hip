// mlp.hip — the learned cardinality estimator's network: MLP 256-128-64-3
// with ReLU, BF16 MFMA GEMMs accumulating in FP32, trained with Adam.
//
// No reference counterpart: the reference sizes hash tables from hard-coded
// selectivities (bespoke_tpch/query_q20.cpp:16, 405-417) and grows them on
// overflow (:158-160).  Here the network predicts log selectivity, log join
// size and log group count from statistics features, and the pipelines size
// their tables from the prediction (with an exact fallback on overflow).
//
// Every GEMM is written in "NT" form, C[M x N] = A[M x K] * Bt[N x K]^T, with
// both operands K-contiguous bf16, so each lane reads its MFMA fragment with
// one 8-byte load.  One wave computes one 16 x 16 tile of C with
// v_mfma_f32_16x16x16_bf16 (lane l holds A[l%16][4(l/16)+j],
// Bt[l%16][4(l/16)+j] and C[4(l/16)+r][l%16]).  The shapes are tiny (batch x
// <= 256), so the kernels favour simplicity over data reuse.
//
// Per layer (in -> out): fp32 master weights Wt[out][in], bf16 copies Wt
// (forward, dW) and W = Wt^T (backward dX), Adam moments m, v.  The output
// layer is padded from 3 to 16 columns whose weights stay zero.
#include <cmath>
#include <cstring>
#include <vector>

#include "prims.hpp"

typedef short bolap_s4 __attribute__((ext_vector_type(4)));
typedef float bolap_f4 __attribute__((ext_vector_type(4)));

struct bolap_mlp;

namespace bolap {
namespace {

constexpr int kIn = 256, kH1 = 128, kH2 = 64, kOut = 3, kOutP = 16;
constexpr int kLayers = 3;
constexpr int kDimIn[kLayers] = {kIn, kH1, kH2};
constexpr int kDimOut[kLayers] = {kH1, kH2, kOutP};
constexpr int kDimOutReal[kLayers] = {kH1, kH2, kOut};

// bf16 round-to-nearest-even of a finite fp32 value.
BOLAP_DEV uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
BOLAP_DEV float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// C = A * Bt^T (+ bias, ReLU); writes fp32 C and/or a bf16 copy.  M, N, K
// are multiples of 16; one wave per 16 x 16 tile.
__global__ __launch_bounds__(256) void gemm_nt_k(const uint16_t* __restrict__ A,
                                                 const uint16_t* __restrict__ Bt, int M, int N,
                                                 int K, const float* __restrict__ bias, int relu,
                                                 float* __restrict__ C32,
                                                 uint16_t* __restrict__ C16) {
    const int wave = (blockIdx.x * 256 + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int tiles_n = N / 16;
    const int tm = wave / tiles_n, tn = wave % tiles_n;
    if (tm >= M / 16) return;  // whole wave exits together
    const int r = lane % 16, kq = 4 * (lane / 16);
    const uint16_t* arow = A + size_t(tm * 16 + r) * K + kq;
    const uint16_t* brow = Bt + size_t(tn * 16 + r) * K + kq;
    bolap_f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bolap_s4 a = *reinterpret_cast<const bolap_s4*>(arow + k0);
        const bolap_s4 b = *reinterpret_cast<const bolap_s4*>(brow + k0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int col = tn * 16 + r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int row = tm * 16 + kq + j;
        float v = acc[j];
        if (bias) v += bias[col];
        if (relu) v = v > 0.f ? v : 0.f;
        if (C32) C32[size_t(row) * N + col] = v;
        if (C16) C16[size_t(row) * N + col] = to_bf16(v);
    }
}

// dst[c][r] = bf16(src[r][c]) for an R x C fp32 or bf16 matrix.
__global__ __launch_bounds__(256) void transpose_bf16_k(const float* __restrict__ s32,
                                                        const uint16_t* __restrict__ s16, int R,
                                                        int Ccols, uint16_t* __restrict__ dst) {
    const int64_t n = int64_t(R) * Ccols;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * 256) {
        const int r = int(i / Ccols), c = int(i % Ccols);
        const uint16_t v = s16 ? s16[i] : to_bf16(s32[i]);
        dst[size_t(c) * R + r] = v;
    }
}

__global__ __launch_bounds__(256) void to_bf16_k(const float* __restrict__ s, int64_t n,
                                                 uint16_t* __restrict__ d) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * 256)
        d[i] = to_bf16(s[i]);
}

// Input rows: fp32 features -> bf16, zero padding rows beyond n.
__global__ __launch_bounds__(256) void input_bf16_k(const float* __restrict__ x, int64_t n,
                                                    int64_t npad, uint16_t* __restrict__ x16) {
    const int64_t tot = npad * kIn;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < tot;
         i += int64_t(gridDim.x) * 256)
        x16[i] = (i / kIn) < n ? to_bf16(x[i]) : 0;
}

// Loss and output gradient: dZ3[b][o] = 2 (y - t) / n for real rows and
// outputs, 0 elsewhere; loss accumulated as sum of squares / n.
__global__ __launch_bounds__(256) void loss_grad_k(const float* __restrict__ y,
                                                   const float* __restrict__ t, int64_t n,
                                                   int64_t npad, float* __restrict__ dz,
                                                   float* __restrict__ loss) {
    float part = 0.f;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < npad * kOutP;
         i += int64_t(gridDim.x) * 256) {
        const int64_t b = i / kOutP;
        const int o = int(i % kOutP);
        float g = 0.f;
        if (b < n && o < kOut) {
            const float d = y[i] - t[b * kOut + o];
            part += d * d;
            g = 2.f * d / float(n);
        }
        dz[i] = g;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) part += __shfl_xor(part, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && part != 0.f) atomicAdd(loss, part / float(n));
}

// dZ *= relu'(activation) (activation is the post-ReLU value).
__global__ __launch_bounds__(256) void relu_back_k(float* __restrict__ dz,
                                                   const float* __restrict__ act, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * 256)
        if (act[i] <= 0.f) dz[i] = 0.f;
}

// db[o] = sum_b dz[b][o] (one thread per output, serial over the batch).
__global__ __launch_bounds__(256) void bias_grad_k(const float* __restrict__ dz, int64_t npad,
                                                   int nout, float* __restrict__ db) {
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= nout) return;
    float s = 0.f;
    for (int64_t b = 0; b < npad; ++b) s += dz[b * nout + o];
    db[o] = s;
}

// Adam on fp32 master weights; rows >= real_out of a padded layer stay zero.
// Refreshes the bf16 copies Wt (out x in) and W (in x out).
__global__ __launch_bounds__(256) void adam_k(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              int64_t n, int nin, int real_out, float lr,
                                              float b1, float b2, float eps, float c1, float c2,
                                              uint16_t* __restrict__ wt16,
                                              uint16_t* __restrict__ w16, int nout) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * 256) {
        const int row = nin > 0 ? int(i / nin) : int(i);
        if (row >= real_out) continue;
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wt16) {
            const uint16_t h = to_bf16(wi);
            wt16[i] = h;
            w16[size_t(i % nin) * nout + row] = h;
        }
    }
}

__global__ __launch_bounds__(256) void init_weights_k(float* __restrict__ w, int64_t n, int nin,
                                                      int real_out, float scale, uint64_t seed,
                                                      uint32_t stream) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * 256) {
        const int row = int(i / nin);
        const uint64_t h = gen_hash(seed, stream, uint64_t(i));
        const float u = float(h >> 40) * (1.0f / 16777216.0f);  // [0, 1)
        w[i] = row < real_out ? (2.f * u - 1.f) * scale : 0.f;
    }
}

__global__ __launch_bounds__(256) void refresh_bf16_k(const float* __restrict__ w, int64_t n,
                                                      int nin, int nout,
                                                      uint16_t* __restrict__ wt16,
                                                      uint16_t* __restrict__ w16) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * 256) {
        const uint16_t h = to_bf16(w[i]);
        wt16[i] = h;
        w16[size_t(i % nin) * nout + i / nin] = h;
    }
}

__global__ void zero_f32_k(float* p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * 256)
        p[i] = 0.f;
}

__global__ __launch_bounds__(256) void take_outputs_k(const float* __restrict__ y, int64_t n,
                                                      float* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n * kOut;
         i += int64_t(gridDim.x) * 256)
        out[i] = y[(i / kOut) * kOutP + i % kOut];
}

inline unsigned g256(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    return unsigned(b < 1 ? 1 : b);
}

}  // namespace
}  // namespace bolap

using namespace bolap;

struct bolap_mlp {
    bolap_ctx* ctx = nullptr;
    // per layer: master fp32 Wt, bias, Adam moments, bf16 Wt and W, grads
    float* w[kLayers] = {};
    float* b[kLayers] = {};
    float* mw[kLayers] = {};
    float* vw[kLayers] = {};
    float* mb[kLayers] = {};
    float* vb[kLayers] = {};
    float* gw[kLayers] = {};
    float* gb[kLayers] = {};
    uint16_t* wt16[kLayers] = {};
    uint16_t* w16[kLayers] = {};
    int64_t step = 0;
    // activations for the current batch capacity
    int64_t cap = 0;
    uint16_t* x16 = nullptr;             // npad x 256
    float* a32[kLayers] = {};            // npad x out (post-activation)
    uint16_t* a16[kLayers] = {};         // npad x out
    float* dz[kLayers] = {};             // npad x out
    uint16_t* dz16[kLayers] = {};        // npad x out
    uint16_t* dzT[kLayers] = {};         // out x npad
    uint16_t* actT[kLayers] = {};        // in x npad (input of each layer, transposed)
    float* dx32[kLayers] = {};           // npad x in
    float* loss = nullptr;
    std::vector<void*> allocs;
};

namespace {

int mlp_alloc(bolap_mlp* m, void** p, size_t bytes) {
    BOLAP_HIP_TRY(hipMalloc(p, bytes));
    m->allocs.push_back(*p);
    return BOLAP_OK;
}

int64_t layer_w(int l) { return int64_t(kDimOut[l]) * kDimIn[l]; }

int ensure_batch(bolap_mlp* m, int64_t npad) {
    if (npad <= m->cap) return BOLAP_OK;
    BOLAP_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    auto free_act = [&](void* p) {
        if (!p) return;
        for (auto& q : m->allocs)
            if (q == p) q = nullptr;
        (void)hipFree(p);
    };
    free_act(m->x16);
    for (int l = 0; l < kLayers; ++l) {
        free_act(m->a32[l]); free_act(m->a16[l]); free_act(m->dz[l]); free_act(m->dz16[l]);
        free_act(m->dzT[l]); free_act(m->actT[l]); free_act(m->dx32[l]);
    }
    int64_t cap = 256;
    while (cap < npad) cap *= 2;
    BOLAP_TRY(mlp_alloc(m, reinterpret_cast<void**>(&m->x16), size_t(cap) * kIn * 2));
    for (int l = 0; l < kLayers; ++l) {
        const size_t o = kDimOut[l], in = kDimIn[l];
        BOLAP_TRY(mlp_alloc(m, reinterpret_cast<void**>(&m->a32[l]), cap * o * 4));
        BOLAP_TRY(mlp_alloc(m, reinterpret_cast<void**>(&m->a16[l]), cap * o * 2));
        BOLAP_TRY(mlp_alloc(m, reinterpret_cast<void**>(&m->dz[l]), cap * o * 4));
        BOLAP_TRY(mlp_alloc(m, reinterpret_cast<void**>(&m->dz16[l]), cap * o * 2));
        BOLAP_TRY(mlp_alloc(m, reinterpret_cast<void**>(&m->dzT[l]), cap * o * 2));
        BOLAP_TRY(mlp_alloc(m, reinterpret_cast<void**>(&m->actT[l]), cap * in * 2));
        BOLAP_TRY(mlp_alloc(m, reinterpret_cast<void**>(&m->dx32[l]), cap * in * 4));
    }
    m->cap = cap;
    return BOLAP_OK;
}

int gemm(bolap_ctx* ctx, const uint16_t* A, const uint16_t* Bt, int M, int N, int K,
         const float* bias, int relu, float* C32, uint16_t* C16) {
    const int64_t waves = int64_t(M / 16) * (N / 16);
    const unsigned blocks = unsigned((waves * kWave + 255) / 256);
    hipLaunchKernelGGL(gemm_nt_k, dim3(blocks), dim3(256), 0, ctx->stream, A, Bt, M, N, K, bias,
                       relu, C32, C16);
    return launch_check("gemm_nt_k");
}

// Forward pass of npad rows (x16 filled); outputs a32/a16 per layer.
int forward(bolap_mlp* m, int64_t npad) {
    const uint16_t* in = m->x16;
    for (int l = 0; l < kLayers; ++l) {
        const int relu = l < kLayers - 1;
        BOLAP_TRY(gemm(m->ctx, in, m->wt16[l], int(npad), kDimOut[l], kDimIn[l], m->b[l], relu,
                       m->a32[l], m->a16[l]));
        in = m->a16[l];
    }
    return BOLAP_OK;
}

int load_inputs(bolap_mlp* m, const float* x, int64_t n, int64_t npad) {
    hipLaunchKernelGGL(input_bf16_k, dim3(g256(npad * kIn)), dim3(256), 0, m->ctx->stream, x, n,
                       npad, m->x16);
    return launch_check("input_bf16_k");
}

// Backward pass after forward: fills gw/gb with dLoss/dW (mean squared error).
int backward(bolap_mlp* m, const float* t, int64_t n, int64_t npad) {
    bolap_ctx* ctx = m->ctx;
    hipLaunchKernelGGL(zero_f32_k, dim3(1), dim3(256), 0, ctx->stream, m->loss, 1);
    hipLaunchKernelGGL(loss_grad_k, dim3(g256(npad * kOutP)), dim3(256), 0, ctx->stream,
                       m->a32[kLayers - 1], t, n, npad, m->dz[kLayers - 1], m->loss);
    BOLAP_TRY(launch_check("loss_grad_k"));
    for (int l = kLayers - 1; l >= 0; --l) {
        const int in = kDimIn[l], out = kDimOut[l];
        // dZ (bf16, and transposed) for the two GEMMs
        hipLaunchKernelGGL(transpose_bf16_k, dim3(g256(npad * out)), dim3(256), 0, ctx->stream,
                           m->dz[l], nullptr, int(npad), out, m->dzT[l]);
        // layer input transposed (in x npad)
        const uint16_t* x_in = l == 0 ? m->x16 : m->a16[l - 1];
        hipLaunchKernelGGL(transpose_bf16_k, dim3(g256(npad * in)), dim3(256), 0, ctx->stream,
                           nullptr, x_in, int(npad), in, m->actT[l]);
        BOLAP_TRY(launch_check("transpose_bf16_k"));
        // dWt[o][i] = sum_b dZ[b][o] X[b][i]
        BOLAP_TRY(gemm(ctx, m->dzT[l], m->actT[l], out, in, int(npad), nullptr, 0, m->gw[l],
                       nullptr));
        hipLaunchKernelGGL(bias_grad_k, dim3((out + 255) / 256), dim3(256), 0, ctx->stream,
                           m->dz[l], npad, out, m->gb[l]);
        BOLAP_TRY(launch_check("bias_grad_k"));
        if (l > 0) {
            // dX[b][i] = sum_o dZ[b][o] Wt[o][i]  (A = dZ in bf16, Bt = W, in x out)
            hipLaunchKernelGGL(to_bf16_k, dim3(g256(npad * out)), dim3(256), 0, ctx->stream,
                               m->dz[l], npad * out, m->dz16[l]);
            BOLAP_TRY(gemm(ctx, m->dz16[l], m->w16[l], int(npad), in, out, nullptr, 0,
                           m->dz[l - 1], nullptr));
            hipLaunchKernelGGL(relu_back_k, dim3(g256(npad * in)), dim3(256), 0, ctx->stream,
                               m->dz[l - 1], m->a32[l - 1], npad * in);
            BOLAP_TRY(launch_check("relu_back_k"));
        }
    }
    return BOLAP_OK;
}

}  // namespace

extern "C" {

int bolap_mlp_num_params(void) {
    return int(layer_w(0) + kH1 + layer_w(1) + kH2 + int64_t(kOut) * kH2 + kOut);
}

int bolap_mlp_create(bolap_ctx* ctx, uint64_t seed, bolap_mlp** out) {
    BOLAP_ARG_CHECK(ctx && out, "bolap_mlp_create: null argument");
    auto* m = new bolap_mlp();
    m->ctx = ctx;
    int s = BOLAP_OK;
    for (int l = 0; l < kLayers && s == BOLAP_OK; ++l) {
        const size_t nw = size_t(layer_w(l)), nb = size_t(kDimOut[l]);
        void** ptrs[] = {reinterpret_cast<void**>(&m->w[l]), reinterpret_cast<void**>(&m->mw[l]),
                         reinterpret_cast<void**>(&m->vw[l]), reinterpret_cast<void**>(&m->gw[l])};
        for (void** p : ptrs)
            if (s == BOLAP_OK) s = mlp_alloc(m, p, nw * 4);
        void** bptrs[] = {reinterpret_cast<void**>(&m->b[l]), reinterpret_cast<void**>(&m->mb[l]),
                          reinterpret_cast<void**>(&m->vb[l]), reinterpret_cast<void**>(&m->gb[l])};
        for (void** p : bptrs)
            if (s == BOLAP_OK) s = mlp_alloc(m, p, nb * 4);
        if (s == BOLAP_OK) s = mlp_alloc(m, reinterpret_cast<void**>(&m->wt16[l]), nw * 2);
        if (s == BOLAP_OK) s = mlp_alloc(m, reinterpret_cast<void**>(&m->w16[l]), nw * 2);
    }
    if (s == BOLAP_OK) s = mlp_alloc(m, reinterpret_cast<void**>(&m->loss), 64);
    if (s != BOLAP_OK) {
        bolap_mlp_destroy(m);
        return s;
    }
    for (int l = 0; l < kLayers; ++l) {
        const int64_t nw = layer_w(l);
        // He-uniform: U(-sqrt(6/fan_in), sqrt(6/fan_in))
        const float scale = std::sqrt(6.0f / float(kDimIn[l]));
        hipLaunchKernelGGL(init_weights_k, dim3(g256(nw)), dim3(256), 0, ctx->stream, m->w[l], nw,
                           kDimIn[l], kDimOutReal[l], scale, seed, uint32_t(100 + l));
        hipLaunchKernelGGL(zero_f32_k, dim3(g256(nw)), dim3(256), 0, ctx->stream, m->mw[l], nw);
        hipLaunchKernelGGL(zero_f32_k, dim3(g256(nw)), dim3(256), 0, ctx->stream, m->vw[l], nw);
        hipLaunchKernelGGL(zero_f32_k, dim3(1), dim3(256), 0, ctx->stream, m->b[l],
                           int64_t(kDimOut[l]));
        hipLaunchKernelGGL(zero_f32_k, dim3(1), dim3(256), 0, ctx->stream, m->mb[l],
                           int64_t(kDimOut[l]));
        hipLaunchKernelGGL(zero_f32_k, dim3(1), dim3(256), 0, ctx->stream, m->vb[l],
                           int64_t(kDimOut[l]));
        hipLaunchKernelGGL(refresh_bf16_k, dim3(g256(nw)), dim3(256), 0, ctx->stream, m->w[l], nw,
                           kDimIn[l], kDimOut[l], m->wt16[l], m->w16[l]);
    }
    s = launch_check("mlp init");
    if (s != BOLAP_OK) {
        bolap_mlp_destroy(m);
        return s;
    }
    *out = m;
    return BOLAP_OK;
}

int bolap_mlp_destroy(bolap_mlp* m) {
    if (!m) return BOLAP_OK;
    if (m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
    for (void* p : m->allocs)
        if (p) (void)hipFree(p);
    delete m;
    return BOLAP_OK;
}

// Parameters in the order W1t[128][256], b1, W2t[64][128], b2, W3t[3][64], b3.
int bolap_mlp_get_params(bolap_mlp* m, float* host) {
    BOLAP_ARG_CHECK(m && host, "bolap_mlp_get_params: null argument");
    BOLAP_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    float* p = host;
    for (int l = 0; l < kLayers; ++l) {
        const size_t nw = size_t(kDimOutReal[l]) * kDimIn[l];
        BOLAP_HIP_TRY(hipMemcpy(p, m->w[l], nw * 4, hipMemcpyDeviceToHost));
        p += nw;
        BOLAP_HIP_TRY(hipMemcpy(p, m->b[l], size_t(kDimOutReal[l]) * 4, hipMemcpyDeviceToHost));
        p += kDimOutReal[l];
    }
    return BOLAP_OK;
}

int bolap_mlp_set_params(bolap_mlp* m, const float* host) {
    BOLAP_ARG_CHECK(m && host, "bolap_mlp_set_params: null argument");
    BOLAP_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    const float* p = host;
    for (int l = 0; l < kLayers; ++l) {
        const size_t nw = size_t(kDimOutReal[l]) * kDimIn[l];
        BOLAP_HIP_TRY(hipMemcpy(m->w[l], p, nw * 4, hipMemcpyHostToDevice));
        p += nw;
        BOLAP_HIP_TRY(hipMemcpy(m->b[l], p, size_t(kDimOutReal[l]) * 4, hipMemcpyHostToDevice));
        p += kDimOutReal[l];
        const int64_t full = layer_w(l);
        hipLaunchKernelGGL(refresh_bf16_k, dim3(g256(full)), dim3(256), 0, m->ctx->stream,
                           m->w[l], full, kDimIn[l], kDimOut[l], m->wt16[l], m->w16[l]);
    }
    BOLAP_TRY(launch_check("refresh_bf16_k"));
    BOLAP_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return BOLAP_OK;
}

int bolap_mlp_predict(bolap_mlp* m, const float* x_dev, int64_t n, float* y_dev) {
    BOLAP_ARG_CHECK(m && n >= 0, "bolap_mlp_predict: bad arguments");
    if (n == 0) return BOLAP_OK;
    BOLAP_ARG_CHECK(x_dev && y_dev, "bolap_mlp_predict: null array");
    const int64_t npad = (n + 15) / 16 * 16;
    BOLAP_TRY(ensure_batch(m, npad));
    TraceScope ts(m->ctx, "mlp_predict");
    BOLAP_TRY(load_inputs(m, x_dev, n, npad));
    BOLAP_TRY(forward(m, npad));
    hipLaunchKernelGGL(take_outputs_k, dim3(g256(n * kOut)), dim3(256), 0, m->ctx->stream,
                       m->a32[kLayers - 1], n, y_dev);
    return launch_check("take_outputs_k");
}

int bolap_mlp_gradients(bolap_mlp* m, const float* x_dev, const float* t_dev, int64_t n,
                        float* grads_host, float* loss_out) {
    BOLAP_ARG_CHECK(m && x_dev && t_dev && n > 0 && grads_host,
                    "bolap_mlp_gradients: bad arguments");
    const int64_t npad = (n + 15) / 16 * 16;
    BOLAP_TRY(ensure_batch(m, npad));
    BOLAP_TRY(load_inputs(m, x_dev, n, npad));
    BOLAP_TRY(forward(m, npad));
    BOLAP_TRY(backward(m, t_dev, n, npad));
    BOLAP_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    float* p = grads_host;
    for (int l = 0; l < kLayers; ++l) {
        const size_t nw = size_t(kDimOutReal[l]) * kDimIn[l];
        BOLAP_HIP_TRY(hipMemcpy(p, m->gw[l], nw * 4, hipMemcpyDeviceToHost));
        p += nw;
        BOLAP_HIP_TRY(hipMemcpy(p, m->gb[l], size_t(kDimOutReal[l]) * 4, hipMemcpyDeviceToHost));
        p += kDimOutReal[l];
    }
    if (loss_out) BOLAP_HIP_TRY(hipMemcpy(loss_out, m->loss, 4, hipMemcpyDeviceToHost));
    return BOLAP_OK;
}

int bolap_mlp_train_step(bolap_mlp* m, const float* x_dev, const float* t_dev, int64_t n,
                         float lr, float* loss_out) {
    BOLAP_ARG_CHECK(m && x_dev && t_dev && n > 0, "bolap_mlp_train_step: bad arguments");
    const int64_t npad = (n + 15) / 16 * 16;
    BOLAP_TRY(ensure_batch(m, npad));
    TraceScope ts(m->ctx, "mlp_train_step");
    BOLAP_TRY(load_inputs(m, x_dev, n, npad));
    BOLAP_TRY(forward(m, npad));
    BOLAP_TRY(backward(m, t_dev, n, npad));
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, float(m->step)), c2 = 1.f - std::pow(b2, float(m->step));
    for (int l = 0; l < kLayers; ++l) {
        const int64_t nw = layer_w(l);
        hipLaunchKernelGGL(adam_k, dim3(g256(nw)), dim3(256), 0, m->ctx->stream, m->w[l],
                           m->gw[l], m->mw[l], m->vw[l], nw, kDimIn[l], kDimOutReal[l], lr, b1,
                           b2, eps, c1, c2, m->wt16[l], m->w16[l], kDimOut[l]);
        hipLaunchKernelGGL(adam_k, dim3(1), dim3(256), 0, m->ctx->stream, m->b[l], m->gb[l],
                           m->mb[l], m->vb[l], int64_t(kDimOut[l]), 0, kDimOutReal[l], lr, b1, b2,
                           eps, c1, c2, nullptr, nullptr, 0);
    }
    BOLAP_TRY(launch_check("adam_k"));
    if (loss_out) {
        BOLAP_HIP_TRY(hipMemcpyAsync(loss_out, m->loss, 4, hipMemcpyDeviceToHost,
                                     m->ctx->stream));
        BOLAP_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    }
    return BOLAP_OK;
}

}  // extern "C"
