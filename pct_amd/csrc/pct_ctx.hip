/*
 * pct_ctx.hip -- context, parameters, workspace management, device scan and
 * stable compaction of the history SoA.
 *
 * The reference keeps all state in globals (data.h:474-615), allocates and
 * frees device buffers inside every stage (e.g. .cu:6159-6194, 6229-6242) and
 * compacts surviving histories on the host with push_back loops
 * (.cu:2443-2477, 2777-2809, 6277-6298). Here the history SoA stays resident
 * in HBM; workspaces live in the context and only grow; compaction is a
 * device scan + scatter that preserves storage order.
 */
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include "pct_internal.h"
#include "pct_plan.h"
#include "pct_rdiv.h"

/* ------------------------------------------------------------------------- */
/* parameters                                                                 */
/* ------------------------------------------------------------------------- */
extern "C" void pct_default_params(pct_params* p) {
    memset(p, 0, sizeof(*p));
    p->columns = 200; p->rows = 200;
    p->slice_thickness = 0.25;
    p->ssd_t_size = 35.0; p->ssd_v_size = 9.0;
    p->recon_cyl_radius = 10.0;
    p->recon_cyl_height = p->ssd_v_size - 1.0;
    p->slices = (int32_t)(p->recon_cyl_height / p->slice_thickness);
    p->t_bin_size = 0.1; p->v_bin_size = 0.25; p->angular_bin_size = 4.0;
    p->t_bins = (int32_t)(p->ssd_t_size / p->t_bin_size + 0.5);
    p->v_bins = (int32_t)(p->ssd_v_size / p->v_bin_size + 0.5);
    p->angular_bins = (int32_t)(360 / p->angular_bin_size + 0.5);
    p->source_radius = 260.7;
    p->gantry_angle_interval = 4.0;
    p->sigmas_to_keep = 3;
    p->fbp_threshold = 0.6;
    p->rsp_air = 0.00113;
    p->sc_threshold = 0.0; p->msc_threshold = 0.0;
    p->hull_filter_threshold = 0.1; p->hull_filter_radius = 1;
    p->msc_diff_thresh = 50;
    p->e_0 = 13.6; p->x_0 = 36.08;
    /* A_k = c_k * (float)10^-e_k: pow((float)10,(float)-e) is a float
     * (data.h:306-311); the float literals below are that value */
    p->a[0] = 7.457 * (double)1e-6f;
    p->a[1] = 4.548 * (double)1e-7f;
    p->a[2] = -5.777 * (double)1e-8f;
    p->a[3] = 1.301 * (double)1e-8f;
    p->a[4] = -9.228 * (double)1e-10f;
    p->a[5] = 2.687 * (double)1e-11f;
    p->max_intersections = 512;
    p->drop_block_size = 3200;
    p->lambda = 0.001f;
    p->iterations = 12;
    p->eta = 2.5;
    p->psi_sign = 1;
    p->sm_lower_threshold = 6.0;
    p->sm_scale_threshold = 1.0;
}

pct_geo pct_make_geo(const pct_params& p) {
    pct_geo g;
    memset(&g, 0, sizeof(g));
    g.C = p.columns; g.R = p.rows; g.S = p.slices;
    g.V = (int64_t)p.columns * p.rows * p.slices;
    double diameter = 2 * p.recon_cyl_radius;
    g.voxel_width = diameter / p.columns;
    g.voxel_height = diameter / p.rows;
    g.voxel_thickness = (p.slices * p.slice_thickness) / p.slices;
    g.inv_width = 1.0 / g.voxel_width;
    g.inv_height = 1.0 / g.voxel_height;
    g.inv_thickness = 1.0 / g.voxel_thickness;
    g.x_zero = -p.recon_cyl_radius;
    g.y_zero = p.recon_cyl_radius;
    g.z_zero = p.recon_cyl_height / 2;
    g.mlp_u_step = g.voxel_width / 2;
    g.radius = p.recon_cyl_radius;
    g.height = p.recon_cyl_height;
    g.e_0 = p.e_0; g.x_0 = p.x_0;
    const double* A = p.a;
    g.A0 = A[0];
    g.a0_2 = A[0] / 2; g.a0_3 = A[0] / 3;
    g.a1_2 = A[1] / 2; g.a1_3 = A[1] / 3; g.a1_4 = A[1] / 4; g.a1_6 = A[1] / 6; g.a1_12 = A[1] / 12;
    g.a2_3 = A[2] / 3; g.a2_4 = A[2] / 4; g.a2_5 = A[2] / 5; g.a2_12 = A[2] / 12; g.a2_30 = A[2] / 30;
    g.a3_4 = A[3] / 4; g.a3_5 = A[3] / 5; g.a3_6 = A[3] / 6; g.a3_20 = A[3] / 20; g.a3_60 = A[3] / 60;
    g.a4_5 = A[4] / 5; g.a4_6 = A[4] / 6; g.a4_7 = A[4] / 7; g.a4_30 = A[4] / 30; g.a4_105 = A[4] / 105;
    g.a5_6 = A[5] / 6; g.a5_7 = A[5] / 7; g.a5_8 = A[5] / 8; g.a5_42 = A[5] / 42; g.a5_168 = A[5] / 168;
    g.max_intersections = p.max_intersections;
    g.hull_words = (g.V + 31) / 32;
    return g;
}

/* ------------------------------------------------------------------------- */
/* errors, workspaces                                                         */
/* ------------------------------------------------------------------------- */
int pct_fail(pct_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

int pct_ensure(pct_ctx* ctx, pct_buffer& b, size_t bytes) {
    if (bytes <= b.bytes && b.ptr) return PCT_OK;
    if (b.ptr) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return pct_fail(ctx, PCT_E_HIP, "sync before realloc: %s", hipGetErrorString(e));
        (void)hipFree(b.ptr);
        b.ptr = nullptr; b.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
        b.ptr = nullptr;
        return pct_fail(ctx, PCT_E_NOMEM, "hipMalloc(%zu) failed", want);
    }
    b.bytes = want;
    return PCT_OK;
}

int pct_take_status(pct_ctx* ctx, int* flags) {
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_status, ctx->d_status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const int h = *ctx->h_status;
    if (h) PCT_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_status, 0, sizeof(int), ctx->stream));
    *flags = h;
    return PCT_OK;
}

int pct_mlp_tables(pct_ctx* ctx);
int pct_noalloc_init(int device);

extern "C" pct_ctx* pct_create(const pct_params* params, void* stream) {
    if (!params) return nullptr;
    if (params->columns <= 0 || params->rows <= 0 || params->slices <= 0 || params->max_intersections <= 0)
        return nullptr;
    pct_ctx* ctx = new pct_ctx();
    ctx->p = *params;
    ctx->g = pct_make_geo(*params);
    if (hipGetDevice(&ctx->device) != hipSuccess) { delete ctx; return nullptr; }
    if (stream) {
        ctx->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return nullptr; }
        ctx->own_stream = true;
    }
    if (hipMalloc(&ctx->d_status, sizeof(int)) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_status, sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipMemsetAsync(ctx->d_status, 0, sizeof(int), ctx->stream) != hipSuccess) {
        pct_destroy(ctx);
        return nullptr;
    }
    /* the MLP tables now, and the device address of the allocation-free
     * entry point's scratch array (once per device, this device current) */
    if (pct_mlp_tables(ctx) != PCT_OK) {
        pct_destroy(ctx);
        return nullptr;
    }
    (void)pct_noalloc_init(ctx->device);
    return ctx;
}

extern "C" void pct_destroy(pct_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    pct_buffer* bufs[] = {&ctx->hull_tmp, &ctx->scan_tmp, &ctx->steps, &ctx->offsets, &ctx->path_store, &ctx->path_len,
                          &ctx->chord, &ctx->xupd, &ctx->counts, &ctx->blk_offsets, &ctx->blk_voxels,
                          &ctx->blk_lengths, &ctx->blk_chord, &ctx->x_saved};
    for (pct_buffer* b : bufs)
        if (b->ptr) (void)hipFree(b->ptr);
    for (pct_buffer& b : ctx->research)
        if (b.ptr) (void)hipFree(b.ptr);
    for (pct_buffer& b : ctx->walk_order)
        if (b.ptr) (void)hipFree(b.ptr);
    if (ctx->d_status) (void)hipFree(ctx->d_status);
    if (ctx->h_status) (void)hipHostFree(ctx->h_status);
    if (ctx->d_mlp_tab) (void)hipFree(ctx->d_mlp_tab);
    if (ctx->d_scat_tab) (void)hipFree(ctx->d_scat_tab);
    if (ctx->d_stats) (void)hipFree(ctx->d_stats);
    (void)pct_net_release(ctx);
    pct_plan_destroy(ctx->plan);
    pct_bin_order_destroy(ctx->bin_order);
    for (hipEvent_t& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->aux_stream) {
        (void)hipStreamSynchronize(ctx->aux_stream);
        (void)hipStreamDestroy(ctx->aux_stream);
    }
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" int pct_sync(pct_ctx* ctx) {
    if (!ctx) return PCT_E_INVALID;
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return PCT_OK;
}

extern "C" const char* pct_last_error(const pct_ctx* ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

/* ------------------------------------------------------------------------- */
/* exclusive scan int32 -> int64 (three passes, 1024 elements per block)      */
/* ------------------------------------------------------------------------- */
#define SCAN_T 256
#define SCAN_PER 4
#define SCAN_TILE (SCAN_T * SCAN_PER)

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
    /* inclusive scan over the 64 lanes of a wavefront */
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int64_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* total) {
    __shared__ int64_t wsum[SCAN_T / 64];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int64_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int64_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_T / 64; ++w) {
        if (w < wid) base += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_reduce(const int32_t* __restrict__ in, int64_t n, int64_t* __restrict__ block_sums) {
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER;
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k)
        if (base + k < n) s += in[base + k];
    int64_t tot;
    (void)block_excl_scan(s, &tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

/* single-block exclusive scan of the block sums, in place; writes the total */
__global__ void __launch_bounds__(SCAN_T) k_scan_blocks(int64_t* __restrict__ sums, int64_t nb, int64_t* __restrict__ total_out) {
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += SCAN_T) {
        int64_t i = base + threadIdx.x;
        int64_t v = i < nb ? sums[i] : 0;
        int64_t tot;
        int64_t ex = block_excl_scan(v, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_apply(const int32_t* __restrict__ in, int64_t n,
                                                       const int64_t* __restrict__ block_offs, int64_t* __restrict__ out) {
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER;
    int64_t vals[SCAN_PER];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        vals[k] = (base + k < n) ? in[base + k] : 0;
        s += vals[k];
    }
    int64_t tot;
    int64_t ex = block_excl_scan(s, &tot) + block_offs[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += vals[k];
    }
}

__global__ void k_store_total(const int64_t* __restrict__ total, int64_t* __restrict__ out, int64_t n) { out[n] = *total; }

int pct_exclusive_scan_i32(pct_ctx* ctx, const int32_t* in, int64_t n, int64_t* out, int64_t* total_host) {
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb < 1) nb = 1;
    int rc = pct_ensure(ctx, ctx->scan_tmp, sizeof(int64_t) * (nb + 1));
    if (rc) return rc;
    int64_t* sums = (int64_t*)ctx->scan_tmp.ptr;
    int64_t* total = sums + nb;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_T), 0, ctx->stream, in, n, sums);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(SCAN_T), 0, ctx->stream, sums, nb, total);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_T), 0, ctx->stream, in, n, (const int64_t*)sums, out);
    hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, ctx->stream, (const int64_t*)total, out, n);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    if (total_host) {
        PCT_HIP_CHECK(ctx, hipMemcpyAsync(total_host, total, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
        PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return PCT_OK;
}

/* ------------------------------------------------------------------------- */
/* stable compaction of the history SoA                                       */
/* ------------------------------------------------------------------------- */
__global__ void k_flags_to_i32(const uint8_t* __restrict__ keep, int64_t n, int32_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = keep[i] ? 1 : 0;
}

struct hist_ptrs {
    float* f[11];
    int32_t* bin;
    uint32_t* fv;
};

__global__ void k_compact_scatter(hist_ptrs src, hist_ptrs dst, const uint8_t* __restrict__ keep,
                                  const int64_t* __restrict__ pos, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    int64_t j = pos[i];
#pragma unroll
    for (int k = 0; k < 11; ++k)
        if (src.f[k]) dst.f[k][j] = src.f[k][i];
    if (src.bin) dst.bin[j] = src.bin[i];
    if (src.fv) dst.fv[j] = src.fv[i];
}

static hist_ptrs to_ptrs(const pct_hist_soa* h) {
    hist_ptrs p;
    p.f[0] = h->x_entry; p.f[1] = h->y_entry; p.f[2] = h->z_entry;
    p.f[3] = h->x_exit; p.f[4] = h->y_exit; p.f[5] = h->z_exit;
    p.f[6] = h->xy_entry_angle; p.f[7] = h->xz_entry_angle;
    p.f[8] = h->xy_exit_angle; p.f[9] = h->xz_exit_angle;
    p.f[10] = h->wepl;
    p.bin = h->bin_num;
    p.fv = h->first_voxel;
    return p;
}

extern "C" int pct_compact(pct_ctx* ctx, const pct_hist_soa* src, const uint8_t* keep,
                           const pct_hist_soa* dst, int64_t* n_kept) {
    if (!ctx || !src || !dst || !n_kept) return pct_fail(ctx, PCT_E_INVALID, "pct_compact: null argument");
    int64_t n = src->n;
    if (n < 0) return pct_fail(ctx, PCT_E_INVALID, "pct_compact: negative n");
    if (n == 0) { *n_kept = 0; return PCT_OK; }
    if (!keep) return pct_fail(ctx, PCT_E_INVALID, "pct_compact: null keep");
    hist_ptrs s = to_ptrs(src), d = to_ptrs(dst);
    for (int k = 0; k < 11; ++k)
        if ((s.f[k] == nullptr) != (d.f[k] == nullptr))
            return pct_fail(ctx, PCT_E_INVALID, "pct_compact: src/dst field %d presence differs", k);
    if ((s.bin == nullptr) != (d.bin == nullptr) || (s.fv == nullptr) != (d.fv == nullptr))
        return pct_fail(ctx, PCT_E_INVALID, "pct_compact: src/dst field presence differs");
    int rc = pct_ensure(ctx, ctx->steps, sizeof(int32_t) * n);
    if (rc) return rc;
    rc = pct_ensure(ctx, ctx->offsets, sizeof(int64_t) * (n + 1));
    if (rc) return rc;
    int32_t* flags = (int32_t*)ctx->steps.ptr;
    int64_t* pos = (int64_t*)ctx->offsets.ptr;
    hipLaunchKernelGGL(k_flags_to_i32, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, keep, n, flags);
    rc = pct_exclusive_scan_i32(ctx, flags, n, pos, n_kept);
    if (rc) return rc;
    hipLaunchKernelGGL(k_compact_scatter, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, s, d, keep,
                       (const int64_t*)pos, n);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

/* ------------------------------------------------------------------------- */
/* C1: cylinder hull, bit-packed (.cu:3316-3326)                              */
/* ------------------------------------------------------------------------- */
__global__ void k_hull_init(pct_geo g, uint32_t* __restrict__ bits) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= g.hull_words) return;
    uint32_t word = 0;
    for (int b = 0; b < 32; ++b) {
        int64_t voxel = w * 32 + b;
        if (voxel >= g.V) break;
        int column = (int)(voxel % g.C);
        int row = (int)((voxel / g.C) % g.R);
        double x = (column - g.C / 2 + 0.5) * g.voxel_width;
        double y = (g.R / 2 - row - 0.5) * g.voxel_height;
        if (x * x + y * y < g.radius * g.radius) word |= 1u << b;
    }
    bits[w] = word;
}

extern "C" int pct_hull_init(pct_ctx* ctx, uint32_t* hull_bits) {
    if (!ctx || !hull_bits) return pct_fail(ctx, PCT_E_INVALID, "pct_hull_init: null argument");
    hipLaunchKernelGGL(k_hull_init, dim3(pct_grid(ctx->g.hull_words, 256)), dim3(256), 0, ctx->stream, ctx->g, hull_bits);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

/* ------------------------------------------------------------------------- */
/* diagnostics: device evaluation of the shared math library                  */
/* ------------------------------------------------------------------------- */
__global__ void k_selftest_math(int func, const double* __restrict__ in, const double* __restrict__ in2,
                                double* __restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = in[i], b = in2 ? in2[i] : 0.0, r = 0.0;
    switch (func) {
        case 0: r = pct_log(a); break;
        case 1: r = pct_sin(a); break;
        case 2: r = pct_cos(a); break;
        case 3: r = pct_tan(a); break;
        case 4: r = pct_atan(a); break;
        case 5: r = pct_atan2(a, b); break;
        case 6: r = (double)pct_sinf((float)a); break;
        case 7: r = (double)pct_cosf((float)a); break;
        case 8: r = (double)pct_atan2f((float)a, (float)b); break;
        case 9: r = pct_sqrt(a); break;
        case 10: r = a / b; break;
        case 11: r = rdiv_q(rdiv_make(b), a); break;          /* = a / b where div_ok(a), div_ok(b) */
        case 12: r = pct_log_div(a, div_rcp{}); break;        /* = pct_log(a) */
        default: r = 0.0; break;
    }
    out[i] = r;
}

extern "C" int pct_selftest_math(pct_ctx* ctx, int func, const double* in, const double* in2, double* out, int64_t n) {
    if (!ctx || !in || !out || n < 0 || func < 0 || func > 12)
        return pct_fail(ctx, PCT_E_INVALID, "pct_selftest_math: bad argument");
    if ((func == 5 || func == 8 || func == 10 || func == 11) && !in2)
        return pct_fail(ctx, PCT_E_INVALID, "pct_selftest_math: function %d needs in2", func);
    if (n == 0) return PCT_OK;
    hipLaunchKernelGGL(k_selftest_math, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, func, in, in2, out, n);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}
