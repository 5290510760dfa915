// colo_mlp.hip -- FP32 MLP cost model (16 -> 64 -> 64 -> 64 -> 2, ReLU)
// with hand-written forward, backward and Adam kernels for gfx950.
//
// From 16 features logged by the miner (pattern size, instance counts,
// parent rows, filter counts and bound, cell occupancy; see colo_miner.cpp
// cand_features) it predicts two costs of a candidate: log1p(rows) of its
// instance table and log1p(refinement tests) of its count pass (instances
// of the last feature tested against its parent rows).  The schedule uses
// them to order and partition the persistent work queue (tests and rows)
// and to pre-size the level's table arena (rows); it never decides a
// pattern, so mining results do not depend on it.
//
// Parameter layout (row-major, out x in): W1[64][16] b1[64] W2[64][64] b2[64]
// W3[64][64] b3[64] W4[2][64] b4[2].  Inputs are standardised with the
// model's mean/std before the first layer.  Loss: squared error summed over
// the two outputs, averaged over the samples.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "colo_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kIn = COLO_MLP_IN, kH = COLO_MLP_HIDDEN, kOut = COLO_MLP_OUT;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kOut * kH, kNP = oB4 + kOut;
static_assert(kNP == COLO_MLP_NPARAMS, "parameter count");
constexpr int kTile = 32;  // samples per workgroup
// the output layer maps one thread to each (sample, output) of a tile
static_assert(kTile * kOut == kH, "output layer: one thread per sample and output");

// Forward for a tile of samples; one thread per hidden neuron (64 threads).
// Writes the three hidden activations (post-ReLU) and the outputs
// (Y[sample * kOut + o]).
__global__ __launch_bounds__(kH) void k_mlp_forward(const float *P, const float *X, const float *mean,
                                                    const float *stdv, uint32_t n, float *H1, float *H2, float *H3,
                                                    float *Y) {
    __shared__ float a0[kTile][kIn];
    __shared__ float a1[kTile][kH + 1];
    __shared__ float a2[kTile][kH + 1];
    __shared__ float a3[kTile][kH + 1];
    const int j = threadIdx.x;
    const uint32_t s0 = blockIdx.x * kTile;
    for (int e = j; e < kTile * kIn; e += kH) {
        int s = e / kIn, i = e - s * kIn;
        uint32_t g = s0 + s;
        a0[s][i] = g < n ? (X[(uint64_t)g * kIn + i] - mean[i]) / stdv[i] : 0.f;
    }
    __syncthreads();
    for (int s = 0; s < kTile; s++) {
        float acc = P[oB1 + j];
        for (int i = 0; i < kIn; i++) acc += P[oW1 + j * kIn + i] * a0[s][i];
        a1[s][j] = acc > 0.f ? acc : 0.f;
    }
    __syncthreads();
    for (int s = 0; s < kTile; s++) {
        float acc = P[oB2 + j];
        for (int i = 0; i < kH; i++) acc += P[oW2 + j * kH + i] * a1[s][i];
        a2[s][j] = acc > 0.f ? acc : 0.f;
    }
    __syncthreads();
    for (int s = 0; s < kTile; s++) {
        float acc = P[oB3 + j];
        for (int i = 0; i < kH; i++) acc += P[oW3 + j * kH + i] * a2[s][i];
        a3[s][j] = acc > 0.f ? acc : 0.f;
    }
    __syncthreads();
    for (int s = 0; s < kTile; s++) {
        uint32_t g = s0 + s;
        if (g >= n) break;
        if (H1) {
            H1[(uint64_t)g * kH + j] = a1[s][j];
            H2[(uint64_t)g * kH + j] = a2[s][j];
            H3[(uint64_t)g * kH + j] = a3[s][j];
        }
    }
    {
        const int s = j / kOut, o = j - s * kOut;
        uint32_t g = s0 + s;
        if (g < n) {
            float acc = P[oB4 + o];
            for (int i = 0; i < kH; i++) acc += P[oW4 + o * kH + i] * a3[s][i];
            Y[(uint64_t)g * kOut + o] = acc;
        }
    }
}

// Backward of the mean squared error for a tile of samples; gradients are
// summed into G with atomics (one add per parameter per workgroup).
__global__ __launch_bounds__(kH) void k_mlp_backward(const float *P, const float *X, const float *mean,
                                                     const float *stdv, const float *T, uint32_t n,
                                                     const float *H1, const float *H2, const float *H3,
                                                     const float *Y, float *G, float *loss) {
    __shared__ float a0[kTile][kIn];
    __shared__ float d1[kTile][kH + 1];
    __shared__ float d2[kTile][kH + 1];
    __shared__ float d3[kTile][kH + 1];
    __shared__ float h1[kTile][kH + 1];
    __shared__ float h2[kTile][kH + 1];
    __shared__ float h3[kTile][kH + 1];
    __shared__ float dy[kTile][kOut];
    const int j = threadIdx.x;
    const uint32_t s0 = blockIdx.x * kTile;
    const float inv = 1.0f / (float)n;
    for (int e = j; e < kTile * kIn; e += kH) {
        int s = e / kIn, i = e - s * kIn;
        uint32_t g = s0 + s;
        a0[s][i] = g < n ? (X[(uint64_t)g * kIn + i] - mean[i]) / stdv[i] : 0.f;
    }
    for (int s = 0; s < kTile; s++) {
        uint32_t g = s0 + s;
        bool v = g < n;
        h1[s][j] = v ? H1[(uint64_t)g * kH + j] : 0.f;
        h2[s][j] = v ? H2[(uint64_t)g * kH + j] : 0.f;
        h3[s][j] = v ? H3[(uint64_t)g * kH + j] : 0.f;
    }
    {
        const int s = j / kOut, o = j - s * kOut;
        uint32_t g = s0 + s;
        float e = g < n ? Y[(uint64_t)g * kOut + o] - T[(uint64_t)g * kOut + o] : 0.f;
        dy[s][o] = 2.f * e * inv;
        if (g < n) atomicAdd(loss, e * e * inv);
    }
    __syncthreads();
    // output layer: dW4[o][j] = sum_s dy[s][o] h3[s][j];
    // d3[s][j] = relu'(h3) sum_o dy[s][o] W4[o][j]
    {
        float w4[kOut];
        for (int o = 0; o < kOut; o++) {
            w4[o] = P[oW4 + o * kH + j];
            float gw = 0.f;
            for (int s = 0; s < kTile; s++) gw += dy[s][o] * h3[s][j];
            atomicAdd(&G[oW4 + o * kH + j], gw);
        }
        for (int s = 0; s < kTile; s++) {
            float acc = 0.f;
            for (int o = 0; o < kOut; o++) acc += dy[s][o] * w4[o];
            d3[s][j] = h3[s][j] > 0.f ? acc : 0.f;
        }
        if (j < kOut) {
            float gb = 0.f;
            for (int s = 0; s < kTile; s++) gb += dy[s][j];
            atomicAdd(&G[oB4 + j], gb);
        }
    }
    __syncthreads();
    // layer 3: dW3[j][i] = sum_s d3[s][j] * h2[s][i]; d2[s][i] = sum_j d3[s][j] W3[j][i] * relu'(h2)
    {
        float gb = 0.f;
        for (int s = 0; s < kTile; s++) gb += d3[s][j];
        atomicAdd(&G[oB3 + j], gb);
        for (int i = 0; i < kH; i++) {
            float gw = 0.f;
            for (int s = 0; s < kTile; s++) gw += d3[s][j] * h2[s][i];
            atomicAdd(&G[oW3 + j * kH + i], gw);
        }
        // thread j now computes d2[.][j] (input index j of layer 3)
        for (int s = 0; s < kTile; s++) {
            float acc = 0.f;
            for (int o = 0; o < kH; o++) acc += d3[s][o] * P[oW3 + o * kH + j];
            d2[s][j] = h2[s][j] > 0.f ? acc : 0.f;
        }
    }
    __syncthreads();
    {
        float gb = 0.f;
        for (int s = 0; s < kTile; s++) gb += d2[s][j];
        atomicAdd(&G[oB2 + j], gb);
        for (int i = 0; i < kH; i++) {
            float gw = 0.f;
            for (int s = 0; s < kTile; s++) gw += d2[s][j] * h1[s][i];
            atomicAdd(&G[oW2 + j * kH + i], gw);
        }
        for (int s = 0; s < kTile; s++) {
            float acc = 0.f;
            for (int o = 0; o < kH; o++) acc += d2[s][o] * P[oW2 + o * kH + j];
            d1[s][j] = h1[s][j] > 0.f ? acc : 0.f;
        }
    }
    __syncthreads();
    {
        float gb = 0.f;
        for (int s = 0; s < kTile; s++) gb += d1[s][j];
        atomicAdd(&G[oB1 + j], gb);
        for (int i = 0; i < kIn; i++) {
            float gw = 0.f;
            for (int s = 0; s < kTile; s++) gw += d1[s][j] * a0[s][i];
            atomicAdd(&G[oW1 + j * kIn + i], gw);
        }
    }
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) with bias correction.
__global__ void k_adam(float *P, const float *G, float *M, float *V, int np, float lr, float b1t, float b2t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float g = G[i];
    float m = 0.9f * M[i] + 0.1f * g;
    float v = 0.999f * V[i] + 0.001f * g * g;
    M[i] = m;
    V[i] = v;
    float mh = m / (1.f - b1t), vh = v / (1.f - b2t);
    P[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

__global__ void k_zero(float *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

}  // namespace

struct colo_mlp {
    hipStream_t s = nullptr;
    float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr, *mean = nullptr, *stdv = nullptr;
    float *loss = nullptr;
    float *X = nullptr, *T = nullptr, *Y = nullptr, *H = nullptr;
    uint32_t cap = 0;
    uint64_t t = 0;  // Adam step
    std::vector<float> h_mean, h_std;
};

static int mlp_cap(colo_mlp *m, uint32_t n) {
    if (n <= m->cap) return COLO_OK;
    void *ptrs[] = {m->X, m->T, m->Y, m->H};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    uint32_t c = std::max<uint32_t>(n, 256);
    if (hipMalloc(&m->X, (size_t)c * kIn * 4) != hipSuccess || hipMalloc(&m->T, (size_t)c * kOut * 4) != hipSuccess ||
        hipMalloc(&m->Y, (size_t)c * kOut * 4) != hipSuccess ||
        hipMalloc(&m->H, (size_t)c * kH * 3 * 4) != hipSuccess) {
        m->cap = 0;
        return COLO_ENOMEM;
    }
    m->cap = c;
    return COLO_OK;
}

static int mlp_forward(colo_mlp *m, uint32_t n, bool keep) {
    if (!n) return COLO_OK;
    float *H1 = m->H, *H2 = m->H + (size_t)m->cap * kH, *H3 = m->H + (size_t)m->cap * kH * 2;
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + kTile - 1) / kTile), dim3(kH), 0, m->s, m->P, m->X, m->mean, m->stdv,
                       n, keep ? H1 : nullptr, H2, H3, m->Y);
    return hipGetLastError() == hipSuccess ? COLO_OK : COLO_EHIP;
}

static int mlp_backward(colo_mlp *m, uint32_t n) {
    float *H1 = m->H, *H2 = m->H + (size_t)m->cap * kH, *H3 = m->H + (size_t)m->cap * kH * 2;
    hipLaunchKernelGGL(k_zero, dim3((kNP + 1 + 255) / 256), dim3(256), 0, m->s, m->G, kNP);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(32), 0, m->s, m->loss, 1);
    if (!n) return COLO_OK;
    hipLaunchKernelGGL(k_mlp_backward, dim3((n + kTile - 1) / kTile), dim3(kH), 0, m->s, m->P, m->X, m->mean, m->stdv,
                       m->T, n, H1, H2, H3, m->Y, m->G, m->loss);
    return hipGetLastError() == hipSuccess ? COLO_OK : COLO_EHIP;
}

namespace colo {
// Predictions for n feature rows held on the host (used by the miner).
int mlp_predict_host(colo_mlp *m, const float *X, uint32_t n, float *y) {
    if (!m || !n) return COLO_OK;
    int rc = mlp_cap(m, n);
    if (rc) return rc;
    if (hipMemcpyAsync(m->X, X, (size_t)n * kIn * 4, hipMemcpyHostToDevice, m->s) != hipSuccess) return COLO_EHIP;
    if ((rc = mlp_forward(m, n, false))) return rc;
    if (hipMemcpyAsync(y, m->Y, (size_t)n * kOut * 4, hipMemcpyDeviceToHost, m->s) != hipSuccess) return COLO_EHIP;
    return hipStreamSynchronize(m->s) == hipSuccess ? COLO_OK : COLO_EHIP;
}
}  // namespace colo

extern "C" {

int colo_mlp_create(uint32_t seed, colo_mlp **out) {
    if (!out) return COLO_EARG;
    colo_mlp *m = new colo_mlp();
    if (hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return COLO_EHIP;
    }
    if (hipMalloc(&m->P, kNP * 4) != hipSuccess || hipMalloc(&m->G, kNP * 4) != hipSuccess ||
        hipMalloc(&m->M, kNP * 4) != hipSuccess || hipMalloc(&m->V, kNP * 4) != hipSuccess ||
        hipMalloc(&m->mean, kIn * 4) != hipSuccess || hipMalloc(&m->stdv, kIn * 4) != hipSuccess ||
        hipMalloc(&m->loss, 4) != hipSuccess) {
        colo_mlp_destroy(m);
        return COLO_ENOMEM;
    }
    // He-uniform initialisation from a seeded generator; zero biases
    std::mt19937 rng(seed);
    std::vector<float> p(kNP, 0.f);
    auto fill = [&](int off, int fan_in, int cnt) {
        std::uniform_real_distribution<float> d(-std::sqrt(6.f / fan_in), std::sqrt(6.f / fan_in));
        for (int i = 0; i < cnt; i++) p[off + i] = d(rng);
    };
    fill(oW1, kIn, kH * kIn);
    fill(oW2, kH, kH * kH);
    fill(oW3, kH, kH * kH);
    fill(oW4, kH, kOut * kH);
    m->h_mean.assign(kIn, 0.f);
    m->h_std.assign(kIn, 1.f);
    std::vector<float> z(kNP, 0.f);
    if (hipMemcpy(m->P, p.data(), kNP * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->M, z.data(), kNP * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->V, z.data(), kNP * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->mean, m->h_mean.data(), kIn * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->stdv, m->h_std.data(), kIn * 4, hipMemcpyHostToDevice) != hipSuccess) {
        colo_mlp_destroy(m);
        return COLO_EHIP;
    }
    *out = m;
    return COLO_OK;
}

int colo_mlp_destroy(colo_mlp *m) {
    if (!m) return COLO_OK;
    if (m->s) (void)hipStreamSynchronize(m->s);
    void *ptrs[] = {m->P, m->G, m->M, m->V, m->mean, m->stdv, m->loss, m->X, m->T, m->Y, m->H};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
    return COLO_OK;
}

int colo_mlp_set_params(colo_mlp *m, const float *p, size_t n) {
    if (!m || n != (size_t)kNP) return COLO_EARG;
    return hipMemcpy(m->P, p, kNP * 4, hipMemcpyHostToDevice) == hipSuccess ? COLO_OK : COLO_EHIP;
}

int colo_mlp_get_params(colo_mlp *m, float *p, size_t n) {
    if (!m || n != (size_t)kNP) return COLO_EARG;
    if (hipStreamSynchronize(m->s) != hipSuccess) return COLO_EHIP;
    return hipMemcpy(p, m->P, kNP * 4, hipMemcpyDeviceToHost) == hipSuccess ? COLO_OK : COLO_EHIP;
}

int colo_mlp_set_norm(colo_mlp *m, const float *mean, const float *stdv) {
    if (!m) return COLO_EARG;
    m->h_mean.assign(mean, mean + kIn);
    m->h_std.assign(stdv, stdv + kIn);
    for (auto &v : m->h_std)
        if (!(v > 0.f)) v = 1.f;
    if (hipMemcpy(m->mean, m->h_mean.data(), kIn * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->stdv, m->h_std.data(), kIn * 4, hipMemcpyHostToDevice) != hipSuccess)
        return COLO_EHIP;
    return COLO_OK;
}

int colo_mlp_predict(colo_mlp *m, const float *X, uint32_t n, float *y) {
    if (!m || (!X && n) || (!y && n)) return COLO_EARG;
    return colo::mlp_predict_host(m, X, n, y);
}

int colo_mlp_grad(colo_mlp *m, const float *X, const float *T, uint32_t n, float *grad, float *loss) {
    if (!m || !n) return COLO_EARG;
    int rc = mlp_cap(m, n);
    if (rc) return rc;
    if (hipMemcpyAsync(m->X, X, (size_t)n * kIn * 4, hipMemcpyHostToDevice, m->s) != hipSuccess ||
        hipMemcpyAsync(m->T, T, (size_t)n * kOut * 4, hipMemcpyHostToDevice, m->s) != hipSuccess)
        return COLO_EHIP;
    if ((rc = mlp_forward(m, n, true)) || (rc = mlp_backward(m, n))) return rc;
    if (grad && hipMemcpyAsync(grad, m->G, kNP * 4, hipMemcpyDeviceToHost, m->s) != hipSuccess) return COLO_EHIP;
    if (loss && hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, m->s) != hipSuccess) return COLO_EHIP;
    return hipStreamSynchronize(m->s) == hipSuccess ? COLO_OK : COLO_EHIP;
}

int colo_mlp_train(colo_mlp *m, const float *X, const float *T, uint32_t n, uint32_t steps, uint32_t batch, float lr,
                   uint32_t seed, float *final_loss) {
    if (!m || !n || !batch) return COLO_EARG;
    if (batch > n) batch = n;
    int rc = mlp_cap(m, batch);
    if (rc) return rc;
    std::mt19937 rng(seed);
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    std::vector<float> bx((size_t)batch * kIn), bt((size_t)batch * kOut);
    uint32_t pos = n;
    float last = 0.f;
    for (uint32_t st = 0; st < steps; st++) {
        if (pos + batch > n) {
            std::shuffle(perm.begin(), perm.end(), rng);
            pos = 0;
        }
        for (uint32_t b = 0; b < batch; b++) {
            uint32_t i = perm[pos + b];
            memcpy(&bx[(size_t)b * kIn], X + (size_t)i * kIn, kIn * 4);
            memcpy(&bt[(size_t)b * kOut], T + (size_t)i * kOut, kOut * 4);
        }
        pos += batch;
        // the host batch arrays are rewritten next step: copy synchronously
        if (hipMemcpyAsync(m->X, bx.data(), bx.size() * 4, hipMemcpyHostToDevice, m->s) != hipSuccess ||
            hipMemcpyAsync(m->T, bt.data(), bt.size() * 4, hipMemcpyHostToDevice, m->s) != hipSuccess)
            return COLO_EHIP;
        if ((rc = mlp_forward(m, batch, true)) || (rc = mlp_backward(m, batch))) return rc;
        m->t++;
        float b1t = std::pow(0.9f, (float)m->t), b2t = std::pow(0.999f, (float)m->t);
        hipLaunchKernelGGL(k_adam, dim3((kNP + 255) / 256), dim3(256), 0, m->s, m->P, m->G, m->M, m->V, kNP, lr, b1t,
                           b2t);
        if (hipGetLastError() != hipSuccess) return COLO_EHIP;
        if (st + 1 == steps || (st & 63) == 63) {
            if (hipMemcpyAsync(&last, m->loss, 4, hipMemcpyDeviceToHost, m->s) != hipSuccess) return COLO_EHIP;
        }
        if (hipStreamSynchronize(m->s) != hipSuccess) return COLO_EHIP;
    }
    if (final_loss) *final_loss = last;
    return COLO_OK;
}

int colo_mlp_save(colo_mlp *m, const char *path) {
    if (!m || !path) return COLO_EARG;
    std::vector<float> p(kNP);
    int rc = colo_mlp_get_params(m, p.data(), kNP);
    if (rc) return rc;
    FILE *fp = fopen(path, "wb");
    if (!fp) return COLO_EFILE;
    const char magic[8] = {'c', 'o', 'l', 'o', 'm', 'l', 'p', '2'};
    bool ok = fwrite(magic, 1, 8, fp) == 8 && fwrite(m->h_mean.data(), 4, kIn, fp) == (size_t)kIn &&
              fwrite(m->h_std.data(), 4, kIn, fp) == (size_t)kIn && fwrite(p.data(), 4, kNP, fp) == (size_t)kNP;
    fclose(fp);
    return ok ? COLO_OK : COLO_EFILE;
}

int colo_mlp_load(const char *path, colo_mlp **out) {
    if (!path || !out) return COLO_EARG;
    FILE *fp = fopen(path, "rb");
    if (!fp) return COLO_EFILE;
    char magic[8];
    std::vector<float> mean(kIn), stdv(kIn), p(kNP);
    bool ok = fread(magic, 1, 8, fp) == 8 && !memcmp(magic, "colomlp2", 8) &&
              fread(mean.data(), 4, kIn, fp) == (size_t)kIn && fread(stdv.data(), 4, kIn, fp) == (size_t)kIn &&
              fread(p.data(), 4, kNP, fp) == (size_t)kNP;
    fclose(fp);
    if (!ok) return COLO_EPARSE;
    int rc = colo_mlp_create(0, out);
    if (rc) return rc;
    if ((rc = colo_mlp_set_norm(*out, mean.data(), stdv.data())) || (rc = colo_mlp_set_params(*out, p.data(), kNP))) {
        colo_mlp_destroy(*out);
        *out = nullptr;
    }
    return rc;
}

}  // extern "C"
