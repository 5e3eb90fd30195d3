// mcb_mlp.hip - the vertex scorer of the learned FVS ordering on gfx950: a 4 -> 64 -> 64 -> 1 MLP
// (ReLU, logistic output) with BF16 weights and activations and FP32 accumulation.  It has no
// counterpart in the reference; it only chooses the order in which the reverse-delete FVS pass
// (include/FVS.cpp:194-213, restated in libmcb_host.so) tries the vertices, and that pass keeps the
// result a minimal feedback vertex set whatever the order.
//
// One kernel does the forward pass and, when training, the backward pass of 16 samples per wave and
// step.  The 64 x 64 layer and its two gradient products run on v_mfma_f32_16x16x16_bf16: dW2 =
// H1^T dH2 accumulates in registers across a wave's samples, and dH1 = dH2 W2^T takes dH2 through a
// per-wave LDS transpose.  The 4-wide input layer and the 1-wide output layer are plain FP32 FMAs.
// Per-lane partial gradients are summed in LDS per workgroup and added to the FP32 gradient with
// one global atomic per parameter per workgroup.  Adam keeps FP32 master weights and rewrites the
// BF16 copies the kernel reads.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <new>
#include <string>
#include <vector>

#include "mcb_mlp.h"

namespace {

thread_local std::string g_mlp_err;

int mfail(int code, const std::string &msg) {
    g_mlp_err = msg;
    return code;
}

#define MLP_TRY(call)                                                                            \
    do {                                                                                         \
        hipError_t _e = (call);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return mfail(MCB_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(_e));      \
    } while (0)

constexpr int kIn = 4, kHid = 64;
// flat parameter layout (FP32): W1 [4][64], b1 [64], W2 [64][64] (row = input unit), b2 [64], W3 [64], b3
constexpr int kW1 = 0, kB1 = kW1 + kIn * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid, kW3 = kB2 + kHid,
              kB3 = kW3 + kHid, kParams = kB3 + 1;
static_assert(kParams == MCB_MLP_PARAMS, "parameter count");
constexpr int kWaves = 4;  // waves per workgroup

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (inputs are finite).
__host__ __device__ inline uint16_t to_bf16(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

__host__ __device__ inline float from_bf16(uint16_t b) {
    const uint32_t u = static_cast<uint32_t>(b) << 16;
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}

struct MlpArgs {
    const float *x;      // [n][4]
    const float *y;      // [n] labels (training)
    const int32_t *idx;  // optional sample order: sample s is row idx[s]
    int64_t first, n;    // samples first .. first+n-1 of the order
    const uint16_t *w1, *w2, *w3;  // BF16 copies
    const float *b1, *b2, *b3;     // FP32 (inside the master parameters)
    float *logit;        // forward output [n] (inference)
    float *grad;         // [kParams + 1]: gradient of the mean loss, then the mean loss (training)
    float inv_n;         // 1 / (samples in the mean)
};

// Lane l of a wave holds, for v_mfma_f32_16x16x16_bf16 with p = l % 16 and q = l / 16: A[p][4q + t],
// B[4q + t][p] (t = 0..3) and D[4q + r][p] (r = 0..3).
template <bool TRAIN>
__global__ __launch_bounds__(kWaves * 64) void k_mlp(MlpArgs a) {
    __shared__ float s_w1[kIn * kHid], s_b1[kHid], s_b2[kHid], s_w3[kHid];
    __shared__ float s_g[TRAIN ? kParams + 1 : 1];
    __shared__ float s_t[TRAIN ? kWaves : 1][16][kHid + 1];  // per-wave transpose of dH2 (padded rows)
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int p = lane & 15, q = lane >> 4;
    for (int i = tid; i < kIn * kHid; i += blockDim.x) s_w1[i] = from_bf16(a.w1[i]);
    for (int i = tid; i < kHid; i += blockDim.x) {
        s_b1[i] = a.b1[i];
        s_b2[i] = a.b2[i];
        s_w3[i] = from_bf16(a.w3[i]);
    }
    if (TRAIN)
        for (int i = tid; i < kParams + 1; i += blockDim.x) s_g[i] = 0.0f;
    const float b3 = a.b3[0];
    // W2 as B operands (k = input unit, n = output unit) and W2^T as B operands of dH1 = dH2 W2^T
    bf16x4 bw2[4][4], bw2t[4][4];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                bw2[kt][nt][t] = static_cast<short>(a.w2[(16 * kt + 4 * q + t) * kHid + 16 * nt + p]);
                if (TRAIN) bw2t[kt][nt][t] = static_cast<short>(a.w2[(16 * nt + p) * kHid + 16 * kt + 4 * q + t]);
            }
    __syncthreads();
    f32x4 gw2[4][4];
    float gw1[kIn][4], gb1[4], gb2[4], gw3[4], gb3 = 0.0f, loss = 0.0f;
    if (TRAIN) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < 4; ++j) gw2[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            gb1[i] = gb2[i] = gw3[i] = 0.0f;
#pragma unroll
            for (int c = 0; c < kIn; ++c) gw1[c][i] = 0.0f;
        }
    }
    auto row_of = [&](int64_t s) -> int64_t { return a.idx ? a.idx[a.first + s] : a.first + s; };
    const int64_t chunks = (a.n + 15) / 16;
    const int64_t wave_g = static_cast<int64_t>(blockIdx.x) * kWaves + wv, nwaves = static_cast<int64_t>(gridDim.x) * kWaves;
    for (int64_t ch = wave_g; ch < chunks; ch += nwaves) {
        const int64_t s0 = ch * 16;
        // layer 1 in the A layout: sample p, hidden units 16kt + 4q + t
        float xs[kIn] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (s0 + p < a.n) {
            const int64_t r = row_of(s0 + p);
#pragma unroll
            for (int c = 0; c < kIn; ++c) xs[c] = a.x[r * kIn + c];
        }
        bf16x4 h1a[4];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int h = 16 * kt + 4 * q + t;
                float v = s_b1[h];
#pragma unroll
                for (int c = 0; c < kIn; ++c) v = fmaf(xs[c], s_w1[c * kHid + h], v);
                h1a[kt][t] = static_cast<short>(to_bf16(fmaxf(v, 0.0f)));
            }
        // layer 2: D[sample 4q + r][unit 16nt + p]
        f32x4 h2[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(h1a[kt], bw2[kt][nt], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) h2[nt][r] = fmaxf(acc[r] + s_b2[16 * nt + p], 0.0f);
        }
        // layer 3: logit of sample 4q + r, summed over the 16 lanes of quarter q
        float z[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = 0.0f;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) v = fmaf(h2[nt][r], s_w3[16 * nt + p], v);
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o);
            z[r] = v + b3;
        }
        if (!TRAIN) {
            if (p == 0)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (s0 + 4 * q + r < a.n) a.logit[s0 + 4 * q + r] = z[r];
            continue;
        }
        // mean binary cross-entropy: dL/dz = (sigmoid(z) - y) / n
        float dz[4], xr[4][kIn];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t s = s0 + 4 * q + r;
            dz[r] = 0.0f;
#pragma unroll
            for (int c = 0; c < kIn; ++c) xr[r][c] = 0.0f;
            if (s < a.n) {
                const int64_t row = row_of(s);
                const float yv = a.y[row];
                dz[r] = (1.0f / (1.0f + expf(-z[r])) - yv) * a.inv_n;
                if (p == 0) loss += (fmaxf(z[r], 0.0f) - z[r] * yv + log1pf(expf(-fabsf(z[r])))) * a.inv_n;
#pragma unroll
                for (int c = 0; c < kIn; ++c) xr[r][c] = a.x[row * kIn + c];
            }
            if (p == 0) gb3 += dz[r];
        }
        // dH2 (D layout), output-layer gradients
        f32x4 dh2[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const float w3 = s_w3[16 * nt + p];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                gw3[nt] = fmaf(h2[nt][r], dz[r], gw3[nt]);
                dh2[nt][r] = h2[nt][r] > 0.0f ? dz[r] * w3 : 0.0f;
                gb2[nt] += dh2[nt][r];
            }
        }
        // H1 of samples 4q + t at hidden units 16mt + p (the A operand of H1^T, the mask of dH1)
        float h1t[4][4];
        bf16x4 h1ta[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int h = 16 * mt + p;
                float v = s_b1[h];
#pragma unroll
                for (int c = 0; c < kIn; ++c) v = fmaf(xr[t][c], s_w1[c * kHid + h], v);
                h1t[mt][t] = fmaxf(v, 0.0f);
                h1ta[mt][t] = static_cast<short>(to_bf16(h1t[mt][t]));
            }
        // dW2 += H1^T dH2: A[unit 16mt + p][sample 4q + t], B[sample 4q + t][unit 16nt + p] = dH2's D layout
        bf16x4 dh2b[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int t = 0; t < 4; ++t) dh2b[nt][t] = static_cast<short>(to_bf16(dh2[nt][t]));
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                gw2[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(h1ta[mt], dh2b[nt], gw2[mt][nt], 0, 0, 0);
        // dH1 = dH2 W2^T: A[sample p][unit 16kt + 4q + t] from the transpose of dH2
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) s_t[wv][4 * q + r][16 * nt + p] = from_bf16(static_cast<uint16_t>(dh2b[nt][r]));
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        bf16x4 dh2a[4];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
#pragma unroll
            for (int t = 0; t < 4; ++t) dh2a[kt][t] = static_cast<short>(to_bf16(s_t[wv][p][16 * kt + 4 * q + t]));
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(dh2a[kt], bw2t[kt][mt], acc, 0, 0, 0);
            // acc[r]: dH1 of sample 4q + r at unit 16mt + p
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float d1 = h1t[mt][r] > 0.0f ? acc[r] : 0.0f;
                gb1[mt] += d1;
#pragma unroll
                for (int c = 0; c < kIn; ++c) gw1[c][mt] = fmaf(xr[r][c], d1, gw1[c][mt]);
            }
        }
    }
    if (!TRAIN) return;
    // workgroup sums in LDS, then one global atomic per parameter
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) atomicAdd(&s_g[kW2 + (16 * mt + 4 * q + r) * kHid + 16 * nt + p], gw2[mt][nt][r]);
        atomicAdd(&s_g[kB1 + 16 * mt + p], gb1[mt]);
        atomicAdd(&s_g[kB2 + 16 * mt + p], gb2[mt]);
        atomicAdd(&s_g[kW3 + 16 * mt + p], gw3[mt]);
#pragma unroll
        for (int c = 0; c < kIn; ++c) atomicAdd(&s_g[kW1 + c * kHid + 16 * mt + p], gw1[c][mt]);
    }
    if (p == 0) {
        atomicAdd(&s_g[kB3], gb3);
        atomicAdd(&s_g[kParams], loss);
    }
    __syncthreads();
    for (int i = tid; i < kParams + 1; i += blockDim.x) atomicAdd(&a.grad[i], s_g[i]);
}

// Adam on the FP32 masters; refreshes the BF16 copies and clears the gradient for the next step.
__global__ void k_mlp_adam(float *w, float *m, float *v, float *g, uint16_t *w1, uint16_t *w2, uint16_t *w3, float lr,
                           float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > kParams) return;
    if (i == kParams) {  // the loss slot
        g[i] = 0.0f;
        return;
    }
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    g[i] = 0.0f;
    if (i < kB1) w1[i - kW1] = to_bf16(wi);
    else if (i >= kW2 && i < kB2) w2[i - kW2] = to_bf16(wi);
    else if (i >= kW3 && i < kB3) w3[i - kW3] = to_bf16(wi);
}

__global__ void k_mlp_refresh(const float *w, uint16_t *w1, uint16_t *w2, uint16_t *w3, float *g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > kParams) return;
    g[i] = 0.0f;
    if (i == kParams) return;
    if (i < kB1) w1[i - kW1] = to_bf16(w[i]);
    else if (i >= kW2 && i < kB2) w2[i - kW2] = to_bf16(w[i]);
    else if (i >= kW3 && i < kB3) w3[i - kW3] = to_bf16(w[i]);
}

}  // namespace

struct mcb_mlp {
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // FP32 masters, moments, gradient + loss
    uint16_t *w1 = nullptr, *w2 = nullptr, *w3 = nullptr;           // BF16 copies
    float *x = nullptr, *y = nullptr, *out = nullptr;
    int32_t *idx = nullptr;
    int64_t n = 0, cap = 0, idx_cap = 0;
    int64_t steps = 0;  // Adam steps taken
    hipStream_t stream = nullptr;
};

namespace {

int mlp_grid(int64_t n) {
    const int64_t chunks = (n + 15) / 16;
    const int64_t blocks = (chunks + kWaves - 1) / kWaves;
    return static_cast<int>(blocks < 1024 ? (blocks > 0 ? blocks : 1) : 1024);
}

int ensure_rows(mcb_mlp *m, int64_t n) {
    if (n <= m->cap) return MCB_OK;
    if (m->x) (void)hipFree(m->x);
    if (m->y) (void)hipFree(m->y);
    if (m->out) (void)hipFree(m->out);
    m->x = m->y = m->out = nullptr;
    m->cap = 0;
    MLP_TRY(hipMalloc(&m->x, sizeof(float) * kIn * n));
    MLP_TRY(hipMalloc(&m->y, sizeof(float) * n));
    MLP_TRY(hipMalloc(&m->out, sizeof(float) * n));
    m->cap = n;
    return MCB_OK;
}

MlpArgs args_of(const mcb_mlp *m) {
    MlpArgs a{};
    a.x = m->x;
    a.y = m->y;
    a.w1 = m->w1;
    a.w2 = m->w2;
    a.w3 = m->w3;
    a.b1 = m->w + kB1;
    a.b2 = m->w + kB2;
    a.b3 = m->w + kB3;
    a.grad = m->g;
    return a;
}

}  // namespace

extern "C" {

const char *mcb_mlp_last_error(void) { return g_mlp_err.c_str(); }

void mcb_mlp_destroy(mcb_mlp *m) {
    if (!m) return;
    void *ptrs[] = {m->w, m->m, m->v, m->g, m->w1, m->w2, m->w3, m->x, m->y, m->out, m->idx};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int mcb_mlp_create(mcb_mlp **out) {
    if (!out) return mfail(MCB_E_INVALID, "invalid arguments");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return mfail(MCB_E_DEVICE, "no HIP device");
    mcb_mlp *m = new (std::nothrow) mcb_mlp();
    if (!m) return mfail(MCB_E_NOMEM, "out of memory");
    const size_t pb = sizeof(float) * (kParams + 1);
    if (hipMalloc(&m->w, pb) != hipSuccess || hipMalloc(&m->m, pb) != hipSuccess || hipMalloc(&m->v, pb) != hipSuccess ||
        hipMalloc(&m->g, pb) != hipSuccess || hipMalloc(&m->w1, 2 * kIn * kHid) != hipSuccess ||
        hipMalloc(&m->w2, 2 * kHid * kHid) != hipSuccess || hipMalloc(&m->w3, 2 * kHid) != hipSuccess ||
        hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        mcb_mlp_destroy(m);
        return mfail(MCB_E_DEVICE, "device allocation");
    }
    std::vector<float> zero(kParams + 1, 0.0f);
    if (hipMemcpy(m->w, zero.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->m, zero.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->v, zero.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->g, zero.data(), pb, hipMemcpyHostToDevice) != hipSuccess) {
        mcb_mlp_destroy(m);
        return mfail(MCB_E_DEVICE, "device copy");
    }
    *out = m;
    return mcb_mlp_set_params(m, zero.data());
}

int mcb_mlp_set_params(mcb_mlp *m, const float *params) {
    if (!m || !params) return mfail(MCB_E_INVALID, "invalid arguments");
    MLP_TRY(hipMemcpyAsync(m->w, params, sizeof(float) * kParams, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_mlp_refresh, dim3((kParams + 256) / 256), dim3(256), 0, m->stream, m->w, m->w1, m->w2, m->w3,
                       m->g);
    MLP_TRY(hipGetLastError());
    std::vector<float> zero(kParams + 1, 0.0f);
    MLP_TRY(hipMemcpyAsync(m->m, zero.data(), sizeof(float) * (kParams + 1), hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipMemcpyAsync(m->v, zero.data(), sizeof(float) * (kParams + 1), hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->steps = 0;
    return MCB_OK;
}

int mcb_mlp_get_params(mcb_mlp *m, float *params) {
    if (!m || !params) return mfail(MCB_E_INVALID, "invalid arguments");
    MLP_TRY(hipStreamSynchronize(m->stream));
    MLP_TRY(hipMemcpy(params, m->w, sizeof(float) * kParams, hipMemcpyDeviceToHost));
    return MCB_OK;
}

int mcb_mlp_forward(mcb_mlp *m, const float *x, int64_t n, float *logits) {
    if (!m || !x || !logits || n < 0) return mfail(MCB_E_INVALID, "invalid arguments");
    if (n == 0) return MCB_OK;
    int rc = ensure_rows(m, n);
    if (rc) return rc;
    MLP_TRY(hipMemcpyAsync(m->x, x, sizeof(float) * kIn * n, hipMemcpyHostToDevice, m->stream));
    MlpArgs a = args_of(m);
    a.n = n;
    a.logit = m->out;
    hipLaunchKernelGGL(k_mlp<false>, dim3(mlp_grid(n)), dim3(kWaves * 64), 0, m->stream, a);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipMemcpyAsync(logits, m->out, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->n = 0;  // the device rows no longer hold a training set
    return MCB_OK;
}

int mcb_mlp_gradient(mcb_mlp *m, const float *x, const float *y, int64_t n, float *grad, float *loss) {
    if (!m || !x || !y || !grad || n <= 0) return mfail(MCB_E_INVALID, "invalid arguments");
    int rc = ensure_rows(m, n);
    if (rc) return rc;
    MLP_TRY(hipMemcpyAsync(m->x, x, sizeof(float) * kIn * n, hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipMemcpyAsync(m->y, y, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_mlp_refresh, dim3((kParams + 256) / 256), dim3(256), 0, m->stream, m->w, m->w1, m->w2, m->w3,
                       m->g);
    MLP_TRY(hipGetLastError());
    MlpArgs a = args_of(m);
    a.n = n;
    a.inv_n = 1.0f / static_cast<float>(n);
    hipLaunchKernelGGL(k_mlp<true>, dim3(mlp_grid(n)), dim3(kWaves * 64), 0, m->stream, a);
    MLP_TRY(hipGetLastError());
    std::vector<float> gl(kParams + 1);
    MLP_TRY(hipMemcpyAsync(gl.data(), m->g, sizeof(float) * (kParams + 1), hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    for (int i = 0; i < kParams; ++i) grad[i] = gl[i];
    if (loss) *loss = gl[kParams];
    hipLaunchKernelGGL(k_mlp_refresh, dim3((kParams + 256) / 256), dim3(256), 0, m->stream, m->w, m->w1, m->w2, m->w3,
                       m->g);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->n = 0;
    return MCB_OK;
}

int mcb_mlp_set_data(mcb_mlp *m, const float *x, const float *y, int64_t n) {
    if (!m || !x || !y || n <= 0) return mfail(MCB_E_INVALID, "invalid arguments");
    int rc = ensure_rows(m, n);
    if (rc) return rc;
    MLP_TRY(hipMemcpyAsync(m->x, x, sizeof(float) * kIn * n, hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipMemcpyAsync(m->y, y, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->n = n;
    return MCB_OK;
}

int mcb_mlp_train_epoch(mcb_mlp *m, const int32_t *order, int64_t batch, float lr, float *mean_loss) {
    if (!m || !order || batch <= 0 || !(lr > 0.0f)) return mfail(MCB_E_INVALID, "invalid arguments");
    if (m->n <= 0) return mfail(MCB_E_STATE, "no training set (mcb_mlp_set_data)");
    for (int64_t i = 0; i < m->n; ++i)
        if (order[i] < 0 || order[i] >= m->n) return mfail(MCB_E_INVALID, "order is not a permutation of the rows");
    if (m->idx_cap < m->n) {
        if (m->idx) (void)hipFree(m->idx);
        m->idx = nullptr;
        m->idx_cap = 0;
        MLP_TRY(hipMalloc(&m->idx, sizeof(int32_t) * m->n));
        m->idx_cap = m->n;
    }
    MLP_TRY(hipMemcpyAsync(m->idx, order, sizeof(int32_t) * m->n, hipMemcpyHostToDevice, m->stream));
    const int64_t nb = (m->n + batch - 1) / batch;
    std::vector<float> losses(static_cast<size_t>(nb), 0.0f);
    float *dloss = nullptr;
    MLP_TRY(hipMalloc(&dloss, sizeof(float) * nb));
    int rc = MCB_OK;
    for (int64_t b = 0; b < nb && rc == MCB_OK; ++b) {
        MlpArgs a = args_of(m);
        a.idx = m->idx;
        a.first = b * batch;
        a.n = std::min<int64_t>(batch, m->n - a.first);
        a.inv_n = 1.0f / static_cast<float>(a.n);
        hipLaunchKernelGGL(k_mlp<true>, dim3(mlp_grid(a.n)), dim3(kWaves * 64), 0, m->stream, a);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(dloss + b, m->g + kParams, sizeof(float), hipMemcpyDeviceToDevice, m->stream) != hipSuccess) {
            rc = mfail(MCB_E_DEVICE, "training step");
            break;
        }
        ++m->steps;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f / (1.0f - std::pow(b1, static_cast<float>(m->steps)));
        const float c2 = 1.0f / (1.0f - std::pow(b2, static_cast<float>(m->steps)));
        hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + 256) / 256), dim3(256), 0, m->stream, m->w, m->m, m->v, m->g,
                           m->w1, m->w2, m->w3, lr, b1, b2, 1e-8f, c1, c2);
        if (hipGetLastError() != hipSuccess) rc = mfail(MCB_E_DEVICE, "k_mlp_adam");
    }
    if (rc == MCB_OK &&
        (hipMemcpyAsync(losses.data(), dloss, sizeof(float) * nb, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
         hipStreamSynchronize(m->stream) != hipSuccess))
        rc = mfail(MCB_E_DEVICE, "training epoch");
    (void)hipFree(dloss);
    if (rc) return rc;
    double sum = 0.0;
    for (int64_t b = 0; b < nb; ++b) sum += static_cast<double>(losses[b]) * std::min<int64_t>(batch, m->n - b * batch);
    if (mean_loss) *mean_loss = static_cast<float>(sum / static_cast<double>(m->n));
    return MCB_OK;
}

}  // extern "C"
