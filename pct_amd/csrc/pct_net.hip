/*
 * pct_net.hip -- the path network of the learned-path form: training
 * (forward + backward on the MFMA units, deterministic gradient reduction,
 * Adam), installation of a trained network for the surrogate walk, and the
 * inference entry point. The surrogate walk and the calibration check are in
 * pct_mlp.hip, beside the exact walk they fall back to.
 *
 * Training kernel: one wavefront per workgroup handles 64 samples. Every
 * product is a v_mfma_f32_16x16x{16,32}_bf16 over operands staged in LDS as
 * bfloat16 (weights rounded from the FP32 master copy, activations and
 * back-propagated errors rounded after FP32 accumulation); the layout is
 * "samples are rows": Z1 = X W1, Z2 = A1 W2, Z3 = A2 W3, and the backward
 * products dW = A^T dZ reduce over the 64 samples as the MFMA K dimension.
 * Each workgroup writes its partial gradient (and squared error) to its own
 * row; k_net_reduce sums the rows in a fixed order, so the gradient does not
 * depend on scheduling. The weights enter every step as a bfloat16 image in
 * the kernel's LDS layout, written by the previous optimizer step.
 */
#include <algorithm>
#include <cmath>
#include <vector>
#include "pct_internal.h"
#include "pct_plan.h"
#include "pct_net_dev.h"

namespace {

constexpr int NP = PCT_NET_PARAMS;
constexpr int PS = PCT_NET_PARAMS + 1;     /* partial row: gradient + squared error */

__device__ inline net_u32x4 gather8(const uint16_t* base, int stride) {
    net_u32x4 r;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        r[j] = (uint32_t)base[(2 * j) * stride] | ((uint32_t)base[(2 * j + 1) * stride] << 16);
    return r;
}

/* bfloat16 image of the weights in the training kernel's LDS layout (W1 with
 * rows 12..15 zero, W3 with columns 8..15 zero) and the FP32 biases; kept up
 * to date by the optimizer step, copied to LDS with 16-byte loads */
struct net_stage {
    uint16_t W1s[16 * 64], W2s[64 * 64], W3s[64 * 16];
    float bias[136];                     /* b1[64], b2[64], b3[8] */
};
static_assert(sizeof(net_stage) % 16 == 0, "stage copied in 16-byte units");

/* position of parameter p in the stage: bf16 index (>= 0) or -(1 + f32 index) */
__host__ __device__ inline int stage_pos(int p) {
    if (p < PCT_NET_OFF_B1) return p;                                              /* W1[f][h] -> f*64 + h */
    if (p < PCT_NET_OFF_W2) return -(1 + (p - PCT_NET_OFF_B1));                    /* b1 */
    if (p < PCT_NET_OFF_B2) return 16 * 64 + (p - PCT_NET_OFF_W2);                 /* W2 */
    if (p < PCT_NET_OFF_W3) return -(1 + 64 + (p - PCT_NET_OFF_B2));              /* b2 */
    if (p < PCT_NET_OFF_B3) {                                                      /* W3[h][o] -> h*16 + o */
        int q = p - PCT_NET_OFF_W3;
        return 16 * 64 + 64 * 64 + (q >> 3) * 16 + (q & 7);
    }
    return -(1 + 128 + (p - PCT_NET_OFF_B3));                                      /* b3 */
}

__device__ inline void stage_write(net_stage* st, int p, float v) {
    int q = stage_pos(p);
    if (q >= 0) ((uint16_t*)st)[q] = (uint16_t)pct_f2bf(v);
    else st->bias[-q - 1] = v;
}

/* the whole stage from the parameters (padding zeroed) */
__global__ void __launch_bounds__(256) k_net_stage(const float* __restrict__ params, net_stage* __restrict__ st) {
    for (int i = threadIdx.x; i < 16 * 64; i += 256) st->W1s[i] = 0;
    for (int i = threadIdx.x; i < 64 * 16; i += 256) st->W3s[i] = 0;
    __syncthreads();
    for (int p = threadIdx.x; p < NP; p += 256) stage_write(st, p, params[p]);
}

/* Forward + backward of 64 samples. Samples are positions j0 + 64 blockIdx.x
 * + s (s < 64) of the pass, position j < j0 + nsamp being sample
 * (pa j + pb) mod n (pa = 0: sample j). scale = 2 / (8 N) for a batch of N. */
__global__ void __launch_bounds__(64)
k_net_train(const net_stage* __restrict__ stage, const float* __restrict__ X, const float* __restrict__ Y, int64_t n,
            int64_t j0, int64_t nsamp, uint64_t pa, uint64_t pb, float scale, float* __restrict__ partial) {
    __shared__ __align__(16) net_stage S;
    __shared__ __align__(16) uint16_t Xs[64 * 16], A1s[64 * 64], A2s[64 * 64], D3s[64 * 16];
    __shared__ int64_t idx_s[64];
    __shared__ int valid_s[64];
    uint16_t* W1s = S.W1s;
    uint16_t* W2s = S.W2s;
    uint16_t* W3s = S.W3s;
    const float* b1s = S.bias;
    const float* b2s = S.bias + 64;
    const float* b3s = S.bias + 128;
    uint16_t* D2s = A2s;   /* dZ2 replaces A2 once dW3 is done */
    uint16_t* D1s = A1s;   /* dZ1 replaces A1 once dW2 is done */
    const int lane = threadIdx.x, g = lane >> 4, c16 = lane & 15;
    float* out = partial + (int64_t)blockIdx.x * PS;

    {
        const net_u32x4* src = (const net_u32x4*)stage;
        net_u32x4* dst = (net_u32x4*)&S;
        constexpr int NV = (int)(sizeof(net_stage) / 16);
#pragma unroll
        for (int i = lane; i < NV; i += 64) dst[i] = src[i];
    }
    {
        const int64_t j = j0 + (int64_t)blockIdx.x * 64 + lane;
        const bool valid = j < j0 + nsamp && j < n;
        const int64_t idx = valid ? (pa ? (int64_t)((pa * (uint64_t)j + pb) % (uint64_t)n) : j) : 0;
        idx_s[lane] = idx;
        valid_s[lane] = valid;
#pragma unroll
        for (int f = 0; f < 16; ++f)
            Xs[lane * 16 + f] = (valid && f < PCT_NET_IN) ? (uint16_t)pct_f2bf(X[idx * PCT_NET_IN + f]) : (uint16_t)0;
    }
    __syncthreads();
    const net_f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};

    /* ---- forward ---- */
    float a1r[4][4][4], a2r[4][4][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            net_u32x2 A = *(const net_u32x2*)(Xs + (16 * mt + c16) * 16 + 4 * g);
            net_u32x2 B;
            B[0] = (uint32_t)W1s[(4 * g) * 64 + 16 * nt + c16] | ((uint32_t)W1s[(4 * g + 1) * 64 + 16 * nt + c16] << 16);
            B[1] = (uint32_t)W1s[(4 * g + 2) * 64 + 16 * nt + c16] | ((uint32_t)W1s[(4 * g + 3) * 64 + 16 * nt + c16] << 16);
            net_f32x4 z = pct_mfma16(A, B, zero);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float a = pct_net_tanh(z[i] + b1s[16 * nt + c16]);
                a1r[mt][nt][i] = a;
                A1s[(16 * mt + 4 * g + i) * 64 + 16 * nt + c16] = (uint16_t)pct_f2bf(a);
            }
        }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            net_f32x4 z = zero;
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) {
                net_u32x4 A = *(const net_u32x4*)(A1s + (16 * mt + c16) * 64 + 32 * kc + 8 * g);
                net_u32x4 B = gather8(W2s + (32 * kc + 8 * g) * 64 + 16 * nt + c16, 64);
                z = pct_mfma32(A, B, z);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float a = pct_net_tanh(z[i] + b2s[16 * nt + c16]);
                a2r[mt][nt][i] = a;
                A2s[(16 * mt + 4 * g + i) * 64 + 16 * nt + c16] = (uint16_t)pct_f2bf(a);
            }
        }
    __syncthreads();
    float loss = 0.0f, db3 = 0.0f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        net_f32x4 z = zero;
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
            net_u32x4 A = *(const net_u32x4*)(A2s + (16 * mt + c16) * 64 + 32 * kc + 8 * g);
            net_u32x4 B = gather8(W3s + (32 * kc + 8 * g) * 16 + c16, 16);
            z = pct_mfma32(A, B, z);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int s = 16 * mt + 4 * g + i;
            float d = 0.0f;
            if (c16 < PCT_NET_OUT && valid_s[s]) {
                float r = (z[i] + b3s[c16]) - Y[idx_s[s] * PCT_NET_OUT + c16];
                loss += r * r;
                d = r * scale;
            }
            D3s[s * 16 + c16] = (uint16_t)pct_f2bf(d);
            db3 += d;
        }
    }
    __syncthreads();

    /* ---- backward ---- */
    /* dW3[h][o] = sum_s A2[s][h] dZ3[s][o] */
#pragma unroll
    for (int mh = 0; mh < 4; ++mh) {
        net_f32x4 z = zero;
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
            net_u32x4 A = gather8(A2s + (32 * kc + 8 * g) * 64 + 16 * mh + c16, 64);
            net_u32x4 B = gather8(D3s + (32 * kc + 8 * g) * 16 + c16, 16);
            z = pct_mfma32(A, B, z);
        }
        if (c16 < PCT_NET_OUT)
#pragma unroll
            for (int i = 0; i < 4; ++i) out[PCT_NET_OFF_W3 + (16 * mh + 4 * g + i) * 8 + c16] = z[i];
    }
    __syncthreads();
    /* dZ2 = (dZ3 W3^T) * (1 - a2^2) */
    float db2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            net_u32x2 A = *(const net_u32x2*)(D3s + (16 * mt + c16) * 16 + 4 * g);
            net_u32x2 B = *(const net_u32x2*)(W3s + (16 * nt + c16) * 16 + 4 * g);
            net_f32x4 z = pct_mfma16(A, B, zero);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float a = a2r[mt][nt][i];
                float d = z[i] * (1.0f - a * a);
                D2s[(16 * mt + 4 * g + i) * 64 + 16 * nt + c16] = (uint16_t)pct_f2bf(d);
                db2[nt] += d;
            }
        }
    __syncthreads();
    /* dW2[h][h'] = sum_s A1[s][h] dZ2[s][h'] */
#pragma unroll
    for (int mh = 0; mh < 4; ++mh)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            net_f32x4 z = zero;
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) {
                net_u32x4 A = gather8(A1s + (32 * kc + 8 * g) * 64 + 16 * mh + c16, 64);
                net_u32x4 B = gather8(D2s + (32 * kc + 8 * g) * 64 + 16 * nt + c16, 64);
                z = pct_mfma32(A, B, z);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) out[PCT_NET_OFF_W2 + (16 * mh + 4 * g + i) * 64 + 16 * nt + c16] = z[i];
        }
    __syncthreads();
    /* dZ1 = (dZ2 W2^T) * (1 - a1^2) */
    float db1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            net_f32x4 z = zero;
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) {
                net_u32x4 A = *(const net_u32x4*)(D2s + (16 * mt + c16) * 64 + 32 * kc + 8 * g);
                net_u32x4 B = *(const net_u32x4*)(W2s + (16 * nt + c16) * 64 + 32 * kc + 8 * g);
                z = pct_mfma32(A, B, z);
            }
            /* D1s aliases A1s: every read of A1s (dW2) finished at the barrier above */
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float a = a1r[mt][nt][i];
                float d = z[i] * (1.0f - a * a);
                D1s[(16 * mt + 4 * g + i) * 64 + 16 * nt + c16] = (uint16_t)pct_f2bf(d);
                db1[nt] += d;
            }
        }
    __syncthreads();
    /* dW1[f][h] = sum_s X[s][f] dZ1[s][h] */
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        net_f32x4 z = zero;
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
            net_u32x4 A = gather8(Xs + (32 * kc + 8 * g) * 16 + c16, 16);
            net_u32x4 B = gather8(D1s + (32 * kc + 8 * g) * 64 + 16 * nt + c16, 64);
            z = pct_mfma32(A, B, z);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (4 * g + i < PCT_NET_IN) out[PCT_NET_OFF_W1 + (4 * g + i) * 64 + 16 * nt + c16] = z[i];
    }
    /* biases: column sums, lanes of one column reduced over g in a fixed tree */
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        float v1 = db1[nt], v2 = db2[nt];
        v1 += __shfl_xor(v1, 16, 64);
        v1 += __shfl_xor(v1, 32, 64);
        v2 += __shfl_xor(v2, 16, 64);
        v2 += __shfl_xor(v2, 32, 64);
        if (g == 0) {
            out[PCT_NET_OFF_B1 + 16 * nt + c16] = v1;
            out[PCT_NET_OFF_B2 + 16 * nt + c16] = v2;
        }
    }
    db3 += __shfl_xor(db3, 16, 64);
    db3 += __shfl_xor(db3, 32, 64);
    if (g == 0 && c16 < PCT_NET_OUT) out[PCT_NET_OFF_B3 + c16] = db3;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) loss += __shfl_xor(loss, d, 64);
    if (lane == 0) out[NP] = loss;
}

/* Sum of the partial rows; then either the gradient is written (grad !=
 * null) or an Adam step is taken (params != null) and the bfloat16 stage
 * updated. A workgroup takes 32 parameters; its 8 row groups (rows q, q + 8,
 * ...) are summed in row order and then combined in group order, so the
 * result does not depend on scheduling. The squared-error sum goes to
 * loss_out[0]. */
__global__ void __launch_bounds__(256)
k_net_reduce(const float* __restrict__ partial, int nrows, float* __restrict__ grad, float* __restrict__ params,
             float* __restrict__ m, float* __restrict__ v, float lr, float beta1, float beta2, float eps, float c1,
             float c2, float* __restrict__ loss_out, net_stage* __restrict__ stage) {
    __shared__ float part[8][32];
    const int pl = threadIdx.x & 31, q = threadIdx.x >> 5;
    const int p = blockIdx.x * 32 + pl;
    float s = 0.0f;
    if (p <= NP) {
        int w = q;
        for (; w + 24 < nrows; w += 32) {
            float a0 = partial[(int64_t)w * PS + p], a1 = partial[(int64_t)(w + 8) * PS + p];
            float a2 = partial[(int64_t)(w + 16) * PS + p], a3 = partial[(int64_t)(w + 24) * PS + p];
            s += a0;
            s += a1;
            s += a2;
            s += a3;
        }
        for (; w < nrows; w += 8) s += partial[(int64_t)w * PS + p];
    }
    part[q][pl] = s;
    __syncthreads();
    if (q != 0 || p > NP) return;
    s = part[0][pl];
#pragma unroll
    for (int k = 1; k < 8; ++k) s += part[k][pl];
    if (p == NP) {
        loss_out[0] = s;
        return;
    }
    if (grad) grad[p] = s;
    if (params) {
        float mm = beta1 * m[p] + (1.0f - beta1) * s;
        float vv = beta2 * v[p] + (1.0f - beta2) * s * s;
        m[p] = mm;
        v[p] = vv;
        float np = params[p] - lr * (mm * c1) / (__builtin_sqrtf(vv * c2) + eps);
        params[p] = np;
        stage_write(stage, p, np);
    }
}

/* network outputs of n feature rows: each wavefront takes 64 rows */
__global__ void __launch_bounds__(256)
k_net_forward(const pct_net_frag* __restrict__ nf, const float* __restrict__ X, int64_t n, float* __restrict__ out) {
    __shared__ __align__(16) uint16_t xs[4][64 * 16];
    __shared__ float os[4][64 * 8];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t base = ((int64_t)blockIdx.x * 4 + w) * 64;
    const int64_t row = base + lane;
#pragma unroll
    for (int f = 0; f < 16; ++f)
        xs[w][lane * 16 + f] = (row < n && f < PCT_NET_IN) ? (uint16_t)pct_f2bf(X[row * PCT_NET_IN + f]) : (uint16_t)0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    pct_net_infer_wave(nf, xs[w], os[w]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (row < n)
#pragma unroll
        for (int o = 0; o < PCT_NET_OUT; ++o) out[row * PCT_NET_OUT + o] = os[w][lane * 8 + o];
}

/* inference operands from the FP32 parameters (host) */
void pack_frag(const float* p, pct_net_frag& f) {
    auto bf = [](float x) { return (uint32_t)pct_f2bf(x); };
    auto pi = [](int c, int g, int j) { return j < 4 ? 32 * c + 4 * g + j : 32 * c + 16 + 4 * g + (j - 4); };
    for (int l = 0; l < 64; ++l) {
        const int g = l >> 4, c16 = l & 15;
        for (int r = 0; r < 4; ++r) {
            uint32_t e[4];
            for (int i = 0; i < 4; ++i) {
                int fi = 4 * g + i;
                e[i] = fi < PCT_NET_IN ? bf(p[PCT_NET_OFF_W1 + fi * 64 + 16 * r + c16]) : 0u;
            }
            f.w1[r][l][0] = e[0] | (e[1] << 16);
            f.w1[r][l][1] = e[2] | (e[3] << 16);
            for (int c = 0; c < 2; ++c) {
                uint32_t q[8];
                for (int j = 0; j < 8; ++j) q[j] = bf(p[PCT_NET_OFF_W2 + pi(c, g, j) * 64 + 16 * r + c16]);
                for (int k = 0; k < 4; ++k) f.w2[r][c][l][k] = q[2 * k] | (q[2 * k + 1] << 16);
            }
        }
        for (int c = 0; c < 2; ++c) {
            uint32_t q[8];
            for (int j = 0; j < 8; ++j) q[j] = c16 < PCT_NET_OUT ? bf(p[PCT_NET_OFF_W3 + pi(c, g, j) * 8 + c16]) : 0u;
            for (int k = 0; k < 4; ++k) f.w3[c][l][k] = q[2 * k] | (q[2 * k + 1] << 16);
        }
    }
    for (int h = 0; h < 64; ++h) {
        f.b1[h] = p[PCT_NET_OFF_B1 + h];
        f.b2[h] = p[PCT_NET_OFF_B2 + h];
    }
    for (int o = 0; o < 8; ++o) f.b3[o] = p[PCT_NET_OFF_B3 + o];
}

int net_alloc(pct_ctx* ctx) {
    if (ctx->d_net) return PCT_OK;
    if (hipMalloc(&ctx->d_net, sizeof(pct_net_frag)) != hipSuccess) {
        ctx->d_net = nullptr;
        return pct_fail(ctx, PCT_E_NOMEM, "network alloc");
    }
    if (hipMalloc(&ctx->d_net_range, sizeof(float) * 2 * PCT_NET_IN) != hipSuccess ||
        hipMalloc(&ctx->d_net_cells, PCT_NET_CELLS) != hipSuccess ||
        hipMalloc(&ctx->d_net_counts, 2 * sizeof(unsigned long long)) != hipSuccess)
        return pct_fail(ctx, PCT_E_NOMEM, "network state alloc");
    PCT_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_net_counts, 0, 2 * sizeof(unsigned long long), ctx->stream));
    return PCT_OK;
}

int net_install(pct_ctx* ctx, const std::vector<float>& params, const float* in_lo, const float* in_hi) {
    int rc = net_alloc(ctx);
    if (rc) return rc;
    pct_net_frag f;
    pack_frag(params.data(), f);
    float range[2 * PCT_NET_IN];
    for (int k = 0; k < PCT_NET_IN; ++k) {
        range[k] = in_lo ? in_lo[k] : -INFINITY;
        range[PCT_NET_IN + k] = in_hi ? in_hi[k] : INFINITY;
    }
    uint8_t cells[PCT_NET_CELLS] = {0};
    /* the staging copies are synchronous: the host structures go out of scope */
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    PCT_HIP_CHECK(ctx, hipMemcpy(ctx->d_net, &f, sizeof f, hipMemcpyHostToDevice));
    PCT_HIP_CHECK(ctx, hipMemcpy(ctx->d_net_range, range, sizeof range, hipMemcpyHostToDevice));
    PCT_HIP_CHECK(ctx, hipMemcpy(ctx->d_net_cells, cells, sizeof cells, hipMemcpyHostToDevice));
    static int version = 0;
    ctx->net_version = ++version;
    return PCT_OK;
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

}  // namespace

bool pct_mlp_mode_ok(const pct_ctx* ctx, int mlp) {
    return mlp == PCT_MLP_FP64 || mlp == PCT_MLP_TABLE || (mlp == PCT_MLP_SURROGATE && ctx && ctx->net_version > 0);
}

int pct_net_release(pct_ctx* ctx) {
    if (!ctx) return PCT_OK;
    void* ptrs[] = {ctx->d_net, ctx->d_net_range, ctx->d_net_cells, ctx->d_net_counts, ctx->d_u1f};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    pct_buffer* bufs[] = {&ctx->net_partial, &ctx->net_state, &ctx->net_loss, &ctx->net_calib, &ctx->net_stage_buf,
                          &ctx->net_fallback};
    for (pct_buffer* b : bufs)
        if (b->ptr) (void)hipFree(b->ptr);
    ctx->d_net = nullptr;
    ctx->d_net_range = nullptr;
    ctx->d_net_cells = nullptr;
    ctx->d_net_counts = nullptr;
    ctx->d_u1f = nullptr;
    return PCT_OK;
}

extern "C" int pct_net_gradient(pct_ctx* ctx, const float* params, const float* features, const float* targets,
                                int64_t n, float* grad, double* loss) {
    if (!ctx || !params || !features || !targets || !grad || n <= 0)
        return pct_fail(ctx, PCT_E_INVALID, "pct_net_gradient: bad argument");
    const int64_t rows = (n + 63) / 64;
    if (rows > (1 << 24)) return pct_fail(ctx, PCT_E_INVALID, "pct_net_gradient: n too large");
    int rc = pct_ensure(ctx, ctx->net_partial, sizeof(float) * rows * PS);
    if (!rc) rc = pct_ensure(ctx, ctx->net_loss, sizeof(float) * 8);
    if (!rc) rc = pct_ensure(ctx, ctx->net_stage_buf, sizeof(net_stage));
    if (rc) return rc;
    float* partial = (float*)ctx->net_partial.ptr;
    float* dloss = (float*)ctx->net_loss.ptr;
    net_stage* stage = (net_stage*)ctx->net_stage_buf.ptr;
    hipLaunchKernelGGL(k_net_stage, dim3(1), dim3(256), 0, ctx->stream, params, stage);
    hipLaunchKernelGGL(k_net_train, dim3((unsigned)rows), dim3(64), 0, ctx->stream, (const net_stage*)stage, features,
                       targets, n, (int64_t)0, n, (uint64_t)0, (uint64_t)0, (float)(2.0 / (8.0 * (double)n)), partial);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_net_reduce, dim3(pct_grid(PS, 32)), dim3(256), 0, ctx->stream, partial, (int)rows, grad,
                       (float*)nullptr, (float*)nullptr, (float*)nullptr, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, dloss,
                       (net_stage*)nullptr);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    if (loss) {
        float l = 0.0f;
        PCT_HIP_CHECK(ctx, hipMemcpyAsync(&l, dloss, sizeof l, hipMemcpyDeviceToHost, ctx->stream));
        PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        *loss = (double)l / (8.0 * (double)n);
    }
    return PCT_OK;
}

extern "C" int pct_net_train(pct_ctx* ctx, float* params, const float* features, const float* targets, int64_t n,
                             const pct_net_train_cfg* cfg, double* epoch_loss) {
    if (!ctx || !params || !features || !targets || !cfg) return pct_fail(ctx, PCT_E_INVALID, "pct_net_train: null argument");
    if (cfg->batch <= 0 || cfg->batch % 64 || cfg->epochs <= 0 || n < cfg->batch)
        return pct_fail(ctx, PCT_E_INVALID, "pct_net_train: batch %d (a multiple of 64, <= n = %lld), epochs %d",
                        cfg->batch, (long long)n, cfg->epochs);
    const int rows = cfg->batch / 64;
    const int64_t steps = n / cfg->batch;
    const int64_t total = steps * cfg->epochs;
    int rc = pct_ensure(ctx, ctx->net_partial, sizeof(float) * (size_t)rows * PS);
    if (!rc) rc = pct_ensure(ctx, ctx->net_state, sizeof(float) * 2 * NP + sizeof(net_stage));
    if (!rc) rc = pct_ensure(ctx, ctx->net_loss, sizeof(float) * (size_t)total);
    if (rc) return rc;
    float* partial = (float*)ctx->net_partial.ptr;
    float* m = (float*)ctx->net_state.ptr;
    float* v = m + NP;
    net_stage* stage = (net_stage*)(v + NP);
    float* dloss = (float*)ctx->net_loss.ptr;
    PCT_HIP_CHECK(ctx, hipMemsetAsync(m, 0, sizeof(float) * 2 * NP, ctx->stream));
    hipLaunchKernelGGL(k_net_stage, dim3(1), dim3(256), 0, ctx->stream, (const float*)params, stage);
    uint64_t seed = cfg->seed;
    int64_t t = 0;
    const float scale = (float)(2.0 / (8.0 * (double)cfg->batch));
    for (int e = 0; e < cfg->epochs; ++e) {
        uint64_t pa = 1 + splitmix(seed) % (uint64_t)(n > 1 ? n - 1 : 1);
        while (gcd64(pa, (uint64_t)n) != 1) ++pa;
        const uint64_t pb = splitmix(seed) % (uint64_t)n;
        for (int64_t s = 0; s < steps; ++s, ++t) {
            const double lr = cfg->lr_min + 0.5 * (cfg->lr - cfg->lr_min) * (1.0 + std::cos(M_PI * (double)t / (double)total));
            const double c1 = 1.0 / (1.0 - std::pow((double)cfg->beta1, (double)(t + 1)));
            const double c2 = 1.0 / (1.0 - std::pow((double)cfg->beta2, (double)(t + 1)));
            hipLaunchKernelGGL(k_net_train, dim3(rows), dim3(64), 0, ctx->stream, (const net_stage*)stage, features,
                               targets, n, s * cfg->batch, (int64_t)cfg->batch, pa, pb, scale, partial);
            hipLaunchKernelGGL(k_net_reduce, dim3(pct_grid(PS, 32)), dim3(256), 0, ctx->stream, (const float*)partial,
                               rows, (float*)nullptr, params, m, v, (float)lr, cfg->beta1, cfg->beta2, cfg->eps,
                               (float)c1, (float)c2, dloss + t, stage);
        }
        PCT_HIP_CHECK(ctx, hipGetLastError());
    }
    if (epoch_loss) {
        std::vector<float> l((size_t)total);
        PCT_HIP_CHECK(ctx, hipMemcpyAsync(l.data(), dloss, sizeof(float) * total, hipMemcpyDeviceToHost, ctx->stream));
        PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        for (int e = 0; e < cfg->epochs; ++e) {
            double s = 0.0;
            for (int64_t k = 0; k < steps; ++k) s += l[(size_t)(e * steps + k)];
            epoch_loss[e] = s / (8.0 * (double)steps * (double)cfg->batch);
        }
    }
    return PCT_OK;
}

extern "C" int pct_net_set(pct_ctx* ctx, const float* params, const float* in_lo, const float* in_hi) {
    if (!ctx || !params) return pct_fail(ctx, PCT_E_INVALID, "pct_net_set: null argument");
    std::vector<float> p(NP);
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    PCT_HIP_CHECK(ctx, hipMemcpy(p.data(), params, sizeof(float) * NP, hipMemcpyDeviceToHost));
    for (float x : p)
        if (!std::isfinite(x)) return pct_fail(ctx, PCT_E_INVALID, "pct_net_set: non-finite parameter");
    return net_install(ctx, p, in_lo, in_hi);
}

extern "C" int pct_net_set_hermite(pct_ctx* ctx, const float* in_lo, const float* in_hi) {
    if (!ctx) return PCT_E_INVALID;
    std::vector<float> p(NP, 0.0f);
    return net_install(ctx, p, in_lo, in_hi);
}

extern "C" int pct_net_forward(pct_ctx* ctx, const float* features, int64_t n, float* out) {
    if (!ctx || !features || !out || n < 0) return pct_fail(ctx, PCT_E_INVALID, "pct_net_forward: bad argument");
    if (!ctx->net_version) return pct_fail(ctx, PCT_E_INVALID, "pct_net_forward: no network installed");
    if (n == 0) return PCT_OK;
    hipLaunchKernelGGL(k_net_forward, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, (const pct_net_frag*)ctx->d_net,
                       features, n, out);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_net_fallback_stats(pct_ctx* ctx, int64_t* fallbacks, int64_t* histories) {
    if (!ctx || !fallbacks || !histories) return pct_fail(ctx, PCT_E_INVALID, "pct_net_fallback_stats: null argument");
    *fallbacks = *histories = 0;
    if (!ctx->d_net_counts) return PCT_OK;
    unsigned long long c[2];
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(c, ctx->d_net_counts, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_net_counts, 0, sizeof c, ctx->stream));
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *fallbacks = (int64_t)c[0];
    *histories = (int64_t)c[1];
    return PCT_OK;
}
