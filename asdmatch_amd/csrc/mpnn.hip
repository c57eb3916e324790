// Learned cardinality estimator: a 3-layer message-passing network over a (sub-)query graph that
// predicts log2(1 + #embeddings).  It stands in for the sampling estimator that fills
// subsetToCard (plan/estimator.cpp:7, :173-291); the planner falls back to its own estimate for
// subsets it is not asked about.  Counts never depend on it (any plan gives the exact count).
//
// Network (hidden H = 64, node features F = 8, per graph at most 32 nodes):
//   h0 = relu(x W_in + b_in)
//   h_{l+1} = relu(h_l Ws_l + (A h_l) Wn_l + b_l)        l = 0, 1, 2   (A = adjacency)
//   y = (sum_n h3[n]) . w_out + b_out
// Weights are stored in bf16 (rounded to nearest even from fp32 master copies); every product is
// accumulated in fp32.  Training: mean squared error on y, Adam on the fp32 masters.
//
// Kernels: one 256-thread workgroup per graph; node states live in LDS (7 x 32 x 64 fp32 for the
// backward pass).  Weight gradients are summed across graphs with fp32 atomics.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "engine.h"
#include "kernels_common.h"

namespace asdm {

constexpr int kF = 8, kH = 64, kL = 3, kMaxN = 32;

// Parameter layout (floats) in the master vector and the bf16 copy.
struct MpnnLayout {
    static constexpr int w_in = 0;                      // [F][H]
    static constexpr int b_in = w_in + kF * kH;         // [H]
    static constexpr int layer0 = b_in + kH;            // per layer: Ws [H][H], Wn [H][H], b [H]
    static constexpr int layer_sz = 2 * kH * kH + kH;
    static constexpr int w_out = layer0 + kL * layer_sz;  // [H]
    static constexpr int b_out = w_out + kH;              // [1]
    static constexpr int total = b_out + 1;
};

__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan unchanged
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

struct GraphBatch {  // device pointers
    const uint32_t *node_off;  // [B + 1]
    const float *x;            // [N][F]
    const uint32_t *adj_off;   // [N + 1]
    const uint32_t *adj;       // neighbour node ids (global)
};

// Forward pass of one graph into LDS; returns y in every thread.  hs[0..3], ag[0..2] are LDS
// arrays [kMaxN][kH].
__device__ inline float mpnn_forward_graph(const uint16_t *__restrict__ w, const GraphBatch &gb, uint32_t g,
                                           float (*hs)[kMaxN][kH], float (*ag)[kMaxN][kH], float *red) {
    const int tid = threadIdx.x;
    const uint32_t n0 = gb.node_off[g], nn = gb.node_off[g + 1] - n0;
    for (int o = tid; o < (int)nn * kH; o += blockDim.x) {
        const int n = o / kH, j = o % kH;
        float z = bf2f(w[MpnnLayout::b_in + j]);
        for (int f = 0; f < kF; ++f) z += gb.x[(n0 + n) * kF + f] * bf2f(w[MpnnLayout::w_in + f * kH + j]);
        hs[0][n][j] = z > 0.f ? z : 0.f;
    }
    __syncthreads();
    for (int l = 0; l < kL; ++l) {
        const uint16_t *Ws = w + MpnnLayout::layer0 + l * MpnnLayout::layer_sz;
        const uint16_t *Wn = Ws + kH * kH;
        const uint16_t *b = Wn + kH * kH;
        for (int o = tid; o < (int)nn * kH; o += blockDim.x) {
            const int n = o / kH, j = o % kH;
            float a = 0.f;
            for (uint32_t e = gb.adj_off[n0 + n]; e < gb.adj_off[n0 + n + 1]; ++e) a += hs[l][gb.adj[e] - n0][j];
            ag[l][n][j] = a;
        }
        __syncthreads();
        for (int o = tid; o < (int)nn * kH; o += blockDim.x) {
            const int n = o / kH, j = o % kH;
            float z = bf2f(b[j]);
            for (int k = 0; k < kH; ++k) z += hs[l][n][k] * bf2f(Ws[k * kH + j]) + ag[l][n][k] * bf2f(Wn[k * kH + j]);
            hs[l + 1][n][j] = z > 0.f ? z : 0.f;
        }
        __syncthreads();
    }
    // readout: y = sum_j (sum_n h3[n][j]) w_out[j] + b_out
    float part = 0.f;
    for (int j = tid; j < kH; j += blockDim.x) {
        float r = 0.f;
        for (uint32_t n = 0; n < nn; ++n) r += hs[kL][n][j];
        part += r * bf2f(w[MpnnLayout::w_out + j]);
    }
    red[tid] = part;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const float y = red[0] + bf2f(w[MpnnLayout::b_out]);
    __syncthreads();
    return y;
}

__global__ void __launch_bounds__(256) k_mpnn_forward(const uint16_t *__restrict__ w, GraphBatch gb, uint32_t B,
                                                      float *__restrict__ y) {
    __shared__ float hs[kL + 1][kMaxN][kH];
    __shared__ float ag[kL][kMaxN][kH];
    __shared__ float red[256];
    for (uint32_t g = blockIdx.x; g < B; g += gridDim.x) {
        const float v = mpnn_forward_graph(w, gb, g, hs, ag, red);
        if (threadIdx.x == 0) y[g] = v;
    }
}

// Forward + backward of the loss 0.5 * mean_g (y_g - t_g)^2; adds d loss / d w (fp32) into grad
// and y_g into y (for the loss).
__global__ void __launch_bounds__(256) k_mpnn_backward(const uint16_t *__restrict__ w, GraphBatch gb, uint32_t B,
                                                       const float *__restrict__ target, float *__restrict__ y,
                                                       float *__restrict__ grad) {
    __shared__ float hs[kL + 1][kMaxN][kH];
    __shared__ float ag[kL][kMaxN][kH];
    __shared__ float dz[kMaxN][kH];   // gradient w.r.t. the current layer's pre-activation
    __shared__ float dh[kMaxN][kH];   // gradient w.r.t. the current layer's input state
    __shared__ float red[256];
    const int tid = threadIdx.x;
    for (uint32_t g = blockIdx.x; g < B; g += gridDim.x) {
        const float yv = mpnn_forward_graph(w, gb, g, hs, ag, red);
        if (tid == 0) y[g] = yv;
        const uint32_t n0 = gb.node_off[g], nn = gb.node_off[g + 1] - n0;
        const float dy = (yv - target[g]) / (float)B;
        // readout
        for (int j = tid; j < kH; j += blockDim.x) {
            float r = 0.f;
            for (uint32_t n = 0; n < nn; ++n) r += hs[kL][n][j];
            atomicAdd(&grad[MpnnLayout::w_out + j], dy * r);
        }
        if (tid == 0) atomicAdd(&grad[MpnnLayout::b_out], dy);
        for (int o = tid; o < (int)nn * kH; o += blockDim.x) {
            const int n = o / kH, j = o % kH;
            dh[n][j] = dy * bf2f(w[MpnnLayout::w_out + j]);
        }
        __syncthreads();
        for (int l = kL - 1; l >= 0; --l) {
            const int base = MpnnLayout::layer0 + l * MpnnLayout::layer_sz;
            const uint16_t *Ws = w + base;
            const uint16_t *Wn = Ws + kH * kH;
            for (int o = tid; o < (int)nn * kH; o += blockDim.x) {
                const int n = o / kH, j = o % kH;
                dz[n][j] = hs[l + 1][n][j] > 0.f ? dh[n][j] : 0.f;
            }
            __syncthreads();
            // weight gradients: dWs[k][j] = sum_n h_l[n][k] dz[n][j], dWn with the aggregates
            for (int o = tid; o < kH * kH; o += blockDim.x) {
                const int k = o / kH, j = o % kH;
                float gs = 0.f, gn = 0.f;
                for (uint32_t n = 0; n < nn; ++n) {
                    gs += hs[l][n][k] * dz[n][j];
                    gn += ag[l][n][k] * dz[n][j];
                }
                atomicAdd(&grad[base + k * kH + j], gs);
                atomicAdd(&grad[base + kH * kH + k * kH + j], gn);
            }
            for (int j = tid; j < kH; j += blockDim.x) {
                float gb2 = 0.f;
                for (uint32_t n = 0; n < nn; ++n) gb2 += dz[n][j];
                atomicAdd(&grad[base + 2 * kH * kH + j], gb2);
            }
            __syncthreads();
            // dh_l[n][k] = sum_j dz[n][j] Ws[k][j] + sum_{m ~ n} sum_j dz[m][j] Wn[k][j]
            for (int o = tid; o < (int)nn * kH; o += blockDim.x) {
                const int n = o / kH, k = o % kH;
                float v = 0.f;
                for (int j = 0; j < kH; ++j) v += dz[n][j] * bf2f(Ws[k * kH + j]);
                for (uint32_t e = gb.adj_off[n0 + n]; e < gb.adj_off[n0 + n + 1]; ++e) {
                    const uint32_t m = gb.adj[e] - n0;  // adjacency is symmetric
                    for (int j = 0; j < kH; ++j) v += dz[m][j] * bf2f(Wn[k * kH + j]);
                }
                ag[l][n][k] = v;  // ag[l] is no longer needed: reuse it for the new dh
            }
            __syncthreads();
            for (int o = tid; o < (int)nn * kH; o += blockDim.x) {
                const int n = o / kH, k = o % kH;
                dh[n][k] = ag[l][n][k];
            }
            __syncthreads();
        }
        // input layer
        for (int o = tid; o < (int)nn * kH; o += blockDim.x) {
            const int n = o / kH, j = o % kH;
            dz[n][j] = hs[0][n][j] > 0.f ? dh[n][j] : 0.f;
        }
        __syncthreads();
        for (int o = tid; o < kF * kH; o += blockDim.x) {
            const int f = o / kH, j = o % kH;
            float gw = 0.f;
            for (uint32_t n = 0; n < nn; ++n) gw += gb.x[(n0 + n) * kF + f] * dz[n][j];
            atomicAdd(&grad[MpnnLayout::w_in + f * kH + j], gw);
        }
        for (int j = tid; j < kH; j += blockDim.x) {
            float gw = 0.f;
            for (uint32_t n = 0; n < nn; ++n) gw += dz[n][j];
            atomicAdd(&grad[MpnnLayout::b_in + j], gw);
        }
        __syncthreads();
    }
}

// Adam on the fp32 masters, then the bf16 copy.
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       uint16_t *__restrict__ wb, int n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        p[i] = pi;
        wb[i] = f2bf(pi);
    }
}

__global__ void k_to_bf16(const float *__restrict__ p, uint16_t *__restrict__ wb, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) wb[i] = f2bf(p[i]);
}

}  // namespace asdm

using namespace asdm;

struct asdm_mpnn {
    asdm_ctx *ctx = nullptr;
    float *master = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t *wb = nullptr;
    uint64_t step = 0;
};

namespace {

struct DevBatch {
    asdm_ctx *ctx;
    uint32_t *node_off = nullptr, *adj_off = nullptr, *adj = nullptr;
    float *x = nullptr, *t = nullptr, *y = nullptr;
    ~DevBatch() {
        for (void *p : {(void *)node_off, (void *)adj_off, (void *)adj, (void *)x, (void *)t, (void *)y}) dev_free(ctx, p);
    }
};

int upload_batch(asdm_ctx *ctx, DevBatch &d, uint32_t B, const uint32_t *node_off, const float *x,
                 const uint32_t *adj_off, const uint32_t *adj, const float *targets) {
    // The kernels index LDS by node - first node of the graph: every range and neighbour id is
    // checked here so that no access leaves the graph's 32 LDS rows.
    if (node_off[0] != 0) return ctx->fail(ASDM_E_INVALID, "node_off[0] must be 0");
    for (uint32_t g = 0; g < B; ++g) {
        if (node_off[g + 1] < node_off[g]) return ctx->fail(ASDM_E_INVALID, "node_off not ascending");
        if (node_off[g + 1] - node_off[g] > (uint32_t)kMaxN) return ctx->fail(ASDM_E_UNSUPPORTED, "graph with > 32 nodes");
    }
    const uint32_t N = node_off[B];
    if (adj_off[0] != 0) return ctx->fail(ASDM_E_INVALID, "adj_off[0] must be 0");
    for (uint32_t n = 0; n < N; ++n)
        if (adj_off[n + 1] < adj_off[n]) return ctx->fail(ASDM_E_INVALID, "adj_off not ascending");
    const uint32_t E = adj_off[N];
    if (E && !adj) return ASDM_E_INVALID;
    for (uint32_t g = 0; g < B; ++g)
        for (uint32_t n = node_off[g]; n < node_off[g + 1]; ++n)
            for (uint32_t e = adj_off[n]; e < adj_off[n + 1]; ++e)
                if (adj[e] < node_off[g] || adj[e] >= node_off[g + 1])
                    return ctx->fail(ASDM_E_INVALID, "neighbour outside its graph");
    int rc = dev_alloc(ctx, (void **)&d.node_off, 4ull * (B + 1ull));
    if (!rc) rc = dev_alloc(ctx, (void **)&d.x, 4 * (size_t)std::max<uint32_t>(N, 1) * kF);
    if (!rc) rc = dev_alloc(ctx, (void **)&d.adj_off, 4ull * (N + 1ull));
    if (!rc) rc = dev_alloc(ctx, (void **)&d.adj, 4ull * std::max<uint32_t>(E, 1));
    if (!rc) rc = dev_alloc(ctx, (void **)&d.y, 4ull * std::max<uint32_t>(B, 1));
    if (!rc && targets) rc = dev_alloc(ctx, (void **)&d.t, 4ull * std::max<uint32_t>(B, 1));
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return bytes ? check_hip(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "H2D") : 0;
    };
    if (!rc) rc = up(d.node_off, node_off, 4ull * (B + 1ull));
    if (!rc) rc = up(d.x, x, 4 * (size_t)N * kF);
    if (!rc) rc = up(d.adj_off, adj_off, 4ull * (N + 1ull));
    if (!rc) rc = up(d.adj, adj, 4 * (size_t)E);
    if (!rc && targets) rc = up(d.t, targets, 4ull * B);
    return rc;
}

}  // namespace

extern "C" {

int asdm_mpnn_create(asdm_ctx *ctx, uint64_t seed, asdm_mpnn **out) try {
    if (!ctx || !out) return ASDM_E_INVALID;
    (void)hipSetDevice(ctx->device);
    auto m = std::make_unique<asdm_mpnn>();
    m->ctx = ctx;
    const int n = MpnnLayout::total;
    std::vector<float> p(n, 0.f);
    asdm::Rng r(seed);
    auto fill = [&](int off, int fan_in, int cnt) {  // uniform(-a, a), a = sqrt(6 / (fan_in + H))
        const float a = std::sqrt(6.0f / (float)(fan_in + kH));
        for (int i = 0; i < cnt; ++i) p[off + i] = (float)((2.0 * r.uniform() - 1.0) * a);
    };
    fill(MpnnLayout::w_in, kF, kF * kH);
    for (int l = 0; l < kL; ++l) {
        fill(MpnnLayout::layer0 + l * MpnnLayout::layer_sz, kH, kH * kH);
        fill(MpnnLayout::layer0 + l * MpnnLayout::layer_sz + kH * kH, kH, kH * kH);
    }
    fill(MpnnLayout::w_out, kH, kH);
    int rc = dev_alloc(ctx, (void **)&m->master, 4 * n);
    if (!rc) rc = dev_alloc(ctx, (void **)&m->grad, 4 * n);
    if (!rc) rc = dev_alloc(ctx, (void **)&m->m1, 4 * n);
    if (!rc) rc = dev_alloc(ctx, (void **)&m->m2, 4 * n);
    if (!rc) rc = dev_alloc(ctx, (void **)&m->wb, 2 * n);
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(m->master, p.data(), 4 * n, hipMemcpyHostToDevice, ctx->stream), "H2D");
    if (!rc) rc = fill_u32(ctx, (uint32_t *)m->m1, n, 0);
    if (!rc) rc = fill_u32(ctx, (uint32_t *)m->m2, n, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, m->master, m->wb, n);
    rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "mpnn init");
    if (rc) return rc;
    *out = m.release();
    return ASDM_OK;
} ASDM_CTX_CATCH(ctx)

void asdm_mpnn_free(asdm_mpnn *m) {
    if (!m) return;
    for (void *p : {(void *)m->master, (void *)m->grad, (void *)m->m1, (void *)m->m2, (void *)m->wb}) dev_free(m->ctx, p);
    delete m;
}

uint32_t asdm_mpnn_num_params(void) { return (uint32_t)MpnnLayout::total; }

// fp32 master parameters (n = asdm_mpnn_num_params()) in and out; set() refreshes the bf16 copy.
int asdm_mpnn_get_params(asdm_mpnn *m, float *p) try {
    if (!m || !p) return ASDM_E_INVALID;
    int rc = check_hip(m->ctx, hipMemcpyAsync(p, m->master, 4 * MpnnLayout::total, hipMemcpyDeviceToHost, m->ctx->stream), "D2H");
    return rc ? rc : check_hip(m->ctx, hipStreamSynchronize(m->ctx->stream), "sync");
} ASDM_HOST_CATCH

int asdm_mpnn_set_params(asdm_mpnn *m, const float *p) try {
    if (!m || !p) return ASDM_E_INVALID;
    const int n = MpnnLayout::total;
    int rc = check_hip(m->ctx, hipMemcpyAsync(m->master, p, 4 * n, hipMemcpyHostToDevice, m->ctx->stream), "H2D");
    if (rc) return rc;
    hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, m->ctx->stream, m->master, m->wb, n);
    return check_hip(m->ctx, hipStreamSynchronize(m->ctx->stream), "sync");
} ASDM_HOST_CATCH

int asdm_mpnn_forward(asdm_mpnn *m, uint32_t B, const uint32_t *node_off, const float *x, const uint32_t *adj_off,
                      const uint32_t *adj, float *y) try {
    if (!m || !node_off || !x || !adj_off || !y) return ASDM_E_INVALID;
    if (B == 0) return ASDM_OK;
    asdm_ctx *ctx = m->ctx;
    DevBatch d{ctx};
    int rc = upload_batch(ctx, d, B, node_off, x, adj_off, adj, nullptr);
    if (rc) return rc;
    GraphBatch gb{d.node_off, d.x, d.adj_off, d.adj};
    KTimer kt(ctx, "mpnn_forward");
    hipLaunchKernelGGL(k_mpnn_forward, dim3(std::min<uint32_t>(B, 4096)), dim3(256), 0, ctx->stream, m->wb, gb, B, d.y);
    kt.stop();
    rc = check_hip(ctx, hipGetLastError(), "mpnn forward");
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(y, d.y, 4ull * B, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return rc;
} ASDM_HOST_CATCH

// Gradient of 0.5 * mean (y - t)^2 w.r.t. the bf16 weights (fp32), without an update.
int asdm_mpnn_grad(asdm_mpnn *m, uint32_t B, const uint32_t *node_off, const float *x, const uint32_t *adj_off,
                   const uint32_t *adj, const float *targets, float *grad, float *y) try {
    if (!m || !node_off || !x || !adj_off || !targets || !grad) return ASDM_E_INVALID;
    asdm_ctx *ctx = m->ctx;
    const int n = MpnnLayout::total;
    DevBatch d{ctx};
    int rc = upload_batch(ctx, d, B, node_off, x, adj_off, adj, targets);
    if (!rc) rc = fill_u32(ctx, (uint32_t *)m->grad, n, 0);
    if (rc) return rc;
    GraphBatch gb{d.node_off, d.x, d.adj_off, d.adj};
    hipLaunchKernelGGL(k_mpnn_backward, dim3(std::min<uint32_t>(std::max<uint32_t>(B, 1), 4096)), dim3(256), 0,
                       ctx->stream, m->wb, gb, B, d.t, d.y, m->grad);
    rc = check_hip(ctx, hipGetLastError(), "mpnn backward");
    if (!rc) rc = check_hip(ctx, hipMemcpyAsync(grad, m->grad, 4 * n, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    if (!rc && y) rc = check_hip(ctx, hipMemcpyAsync(y, d.y, 4ull * B, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return rc;
} ASDM_HOST_CATCH

// One Adam step on a batch; *loss = 0.5 * mean (y - t)^2 before the step.
int asdm_mpnn_train_step(asdm_mpnn *m, uint32_t B, const uint32_t *node_off, const float *x, const uint32_t *adj_off,
                         const uint32_t *adj, const float *targets, float lr, float *loss) try {
    if (!m || !node_off || !x || !adj_off || !targets || B == 0) return ASDM_E_INVALID;
    asdm_ctx *ctx = m->ctx;
    const int n = MpnnLayout::total;
    DevBatch d{ctx};
    int rc = upload_batch(ctx, d, B, node_off, x, adj_off, adj, targets);
    if (!rc) rc = fill_u32(ctx, (uint32_t *)m->grad, n, 0);
    if (rc) return rc;
    GraphBatch gb{d.node_off, d.x, d.adj_off, d.adj};
    {
        KTimer kt(ctx, "mpnn_backward");
        hipLaunchKernelGGL(k_mpnn_backward, dim3(std::min<uint32_t>(B, 4096)), dim3(256), 0, ctx->stream, m->wb, gb, B,
                           d.t, d.y, m->grad);
        kt.stop();
    }
    if ((rc = check_hip(ctx, hipGetLastError(), "mpnn backward"))) return rc;
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)m->step), bc2 = 1.f - std::pow(b2, (float)m->step);
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, m->master, m->grad, m->m1, m->m2, m->wb,
                       n, lr, b1, b2, 1e-8f, bc1, bc2);
    if ((rc = check_hip(ctx, hipGetLastError(), "adam"))) return rc;
    if (loss) {
        std::vector<float> y(B), t(targets, targets + B);
        rc = check_hip(ctx, hipMemcpyAsync(y.data(), d.y, 4ull * B, hipMemcpyDeviceToHost, ctx->stream), "D2H");
        if (!rc) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
        double s = 0;
        for (uint32_t i = 0; i < B; ++i) s += 0.5 * (y[i] - t[i]) * (y[i] - t[i]);
        *loss = (float)(s / B);
    }
    return rc;
} ASDM_HOST_CATCH

int asdm_mpnn_save(asdm_mpnn *m, const char *path) try {
    if (!m || !path) return ASDM_E_INVALID;
    std::vector<float> p(MpnnLayout::total);
    int rc = asdm_mpnn_get_params(m, p.data());
    if (rc) return rc;
    FILE *f = std::fopen(path, "wb");
    if (!f) return ASDM_E_IO;
    const uint32_t hdr[4] = {0x4d504e4eu, (uint32_t)kF, (uint32_t)kH, (uint32_t)MpnnLayout::total};
    bool ok = std::fwrite(hdr, 4, 4, f) == 4 && std::fwrite(p.data(), 4, p.size(), f) == p.size();
    return std::fclose(f) == 0 && ok ? ASDM_OK : ASDM_E_IO;
} ASDM_HOST_CATCH

int asdm_mpnn_load(asdm_mpnn *m, const char *path) try {
    if (!m || !path) return ASDM_E_INVALID;
    FILE *f = std::fopen(path, "rb");
    if (!f) return ASDM_E_IO;
    uint32_t hdr[4];
    std::vector<float> p(MpnnLayout::total);
    bool ok = std::fread(hdr, 4, 4, f) == 4 && hdr[0] == 0x4d504e4eu && hdr[1] == (uint32_t)kF && hdr[2] == (uint32_t)kH &&
              hdr[3] == (uint32_t)MpnnLayout::total && std::fread(p.data(), 4, p.size(), f) == p.size();
    std::fclose(f);
    if (!ok) return ASDM_E_IO;
    return asdm_mpnn_set_params(m, p.data());
} ASDM_HOST_CATCH

}  // extern "C"
