/*
 * pct_plan.hip -- the DROP iteration over a resident per-scan plan.
 *
 * The voxel lists of the MLP do not depend on the image, so the structure of
 * every DROP block -- which voxels it touches, how many of its histories touch
 * each, and which ones -- is the same in every iteration. The plan stores it
 * once per scan, next to the path store:
 *
 *   path store   CSR by history (u32 voxel ids, capacity offsets), lengths,
 *                effective chord lengths, a hash of every list;
 *   block CSC    per DROP block, the distinct voxels it touches sorted by
 *                voxel (u32), each with the block-local indices of the
 *                histories that touch it, in history order (u16).
 *
 * It is built with one stable radix sort by (block, voxel) over chunks of
 * blocks. An iteration is then two gather-only kernels per block, with no
 * atomics at all:
 *
 *   k_block_delta  one wavefront per history: a_i.x over the path, then
 *                  delta_i = L (b_i - L sum x) / (n L^2) lambda, in FP64
 *                  (block_update_GPU, .cu:5822-5834);
 *   k_block_apply  one thread per touched voxel: x_update = sum of the deltas
 *                  of its histories in history order, in FP64, then
 *                  x += x_update / count (DROP_blocks + DROP_update,
 *                  .cu:8023-8055) -- the oracle's accumulation order.
 *
 * Segments are sorted by voxel, so the image reads/writes of k_block_apply are
 * monotone in address. The whole block loop of an iteration is captured once
 * into a hipGraph and replayed. In the default mode each iteration first
 * recomputes every MLP voxel list (as the reference does, .cu:5822) and checks
 * it against the plan's hash.
 */
#include <rocprim/device/device_radix_sort.hpp>
#include "pct_internal.h"
#include "pct_plan.h"

int pct_mlp_capacity_impl(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, int64_t* offsets,
                          int64_t* total);
int pct_mlp_paths_impl(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, const int64_t* offsets,
                       uint32_t* voxels, uint32_t* lengths, double* chord, uint32_t* hash_out,
                       const uint32_t* hash_expect, int mode);
int pct_status_to_rc(pct_ctx* ctx, int flags);
int pct_drop_iteration_atomic(pct_ctx* ctx, const pct_hist_soa* hist, float* x, int block_size, float lambda,
                              int mode);

namespace {

/* ------------------------------------------------------------------------- */
/* plan construction kernels                                                  */
/* ------------------------------------------------------------------------- */
__global__ void k_lengths_i32(const uint32_t* __restrict__ len, int64_t n, int32_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int32_t)len[i];
}

/* block entry bases: blk_entry[b] = exact offset of history b*B */
__global__ void k_block_bases(const int64_t* __restrict__ ex_off, int64_t n, int64_t B, int64_t nb,
                              int64_t* __restrict__ blk_entry) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nb) return;
    int64_t h = b * B < n ? b * B : n;
    blk_entry[b] = ex_off[h];
}

/* one wavefront per history of the chunk: keys = (local block << vbits) | voxel,
 * values = block-local history index, at the history's exact entry offset */
__global__ void __launch_bounds__(256)
k_emit_keys(const uint32_t* __restrict__ voxels, const int64_t* __restrict__ cap_off,
            const uint32_t* __restrict__ len, const int64_t* __restrict__ ex_off, int64_t h0, int64_t h1,
            int64_t B, int64_t b0, int vbits, int64_t entry0, uint32_t* __restrict__ keys,
            uint32_t* __restrict__ vals) {
    int64_t h = h0 + (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (h >= h1) return;
    int64_t b = h / B;
    uint32_t hi = (uint32_t)(b - b0) << vbits;
    uint32_t local = (uint32_t)(h - b * B);
    int n = (int)len[h];
    int64_t src = cap_off[h], dst = ex_off[h] - entry0;
    for (int j = lane; j < n; j += 64) {
        keys[dst + j] = hi | voxels[src + j];
        vals[dst + j] = local;
    }
}

__global__ void k_head_flags(const uint32_t* __restrict__ keys, int64_t E, int32_t* __restrict__ flags) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < E) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

/* segments of the chunk: voxel and block-relative start of each, the CSC
 * values, and the first segment of each block */
__global__ void k_write_segments(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                 const int32_t* __restrict__ flags, const int64_t* __restrict__ segidx, int64_t E,
                                 int vbits, int64_t seg0, int64_t b0, int64_t entry0,
                                 const int64_t* __restrict__ blk_entry, uint32_t* __restrict__ seg_voxel,
                                 uint32_t* __restrict__ seg_start, uint16_t* __restrict__ csc,
                                 int64_t* __restrict__ blk_seg) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E) return;
    uint32_t k = keys[i];
    int64_t b = b0 + (int64_t)(k >> vbits);
    int64_t gi = entry0 + i;
    csc[gi] = (uint16_t)vals[i];
    if (flags[i]) {
        int64_t s = seg0 + segidx[i];
        seg_voxel[s] = k & ((1u << vbits) - 1u);
        seg_start[s] = (uint32_t)(gi - blk_entry[b]);
        if (gi == blk_entry[b]) blk_seg[b] = s;
    }
}

/* ------------------------------------------------------------------------- */
/* iteration kernels                                                          */
/* ------------------------------------------------------------------------- */
#define DELTA_WAVES 4
#define DELTA_SLOTS 8

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__global__ void __launch_bounds__(64 * DELTA_WAVES)
k_block_delta(const uint32_t* __restrict__ voxels, const int64_t* __restrict__ cap_off,
              const uint32_t* __restrict__ len, const double* __restrict__ chord, const float* __restrict__ wepl,
              int64_t h0, int64_t count, const float* __restrict__ x, double lambda, double* __restrict__ delta) {
    const int lane = threadIdx.x & 63;
    const int64_t k = (int64_t)blockIdx.x * DELTA_WAVES + (threadIdx.x >> 6);
    if (k >= count) return;
    const int64_t h = h0 + k;
    const int n = (int)len[h];
    const uint32_t* p = voxels + cap_off[h];
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < DELTA_SLOTS; ++r) {
        int j = lane + 64 * r;
        if (j < n) s += (double)x[p[j]];
    }
    for (int j = lane + 64 * DELTA_SLOTS; j < n; j += 64) s += (double)x[p[j]];
    s = wave_sum_d(s);
    if (lane == 0) {
        const double L = chord[h];
        const double a_i_dot_x_k = s * L;
        const double a_i_dot_a_i = (L * L) * n;
        delta[k] = L * ((wepl[h] - a_i_dot_x_k) / a_i_dot_a_i) * lambda;
    }
}

__global__ void __launch_bounds__(256)
k_block_apply(const uint32_t* __restrict__ seg_voxel, const uint32_t* __restrict__ seg_start,
              const uint16_t* __restrict__ csc, int64_t s0, int64_t ns, int64_t entry0, uint32_t n_entries,
              const double* __restrict__ delta, float* __restrict__ x) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    int64_t s = s0 + k;
    uint32_t a = seg_start[s];
    uint32_t e = (k + 1 < ns) ? seg_start[s + 1] : n_entries;
    const uint16_t* c = csc + entry0;
    double upd = 0.0;
    for (uint32_t i = a; i < e; ++i) upd += delta[c[i]];
    uint32_t v = seg_voxel[s];
    x[v] = (float)((double)x[v] + upd / (double)(e - a));
}

/* Precision form of the update: the same two kernels in FP32 with
 * Kahan-compensated sums (a_i.x per history, x_update per voxel). */
__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

__global__ void __launch_bounds__(64 * DELTA_WAVES)
k_block_delta_f32(const uint32_t* __restrict__ voxels, const int64_t* __restrict__ cap_off,
                  const uint32_t* __restrict__ len, const double* __restrict__ chord, const float* __restrict__ wepl,
                  int64_t h0, int64_t count, const float* __restrict__ x, float lambda, float* __restrict__ delta) {
    const int lane = threadIdx.x & 63;
    const int64_t k = (int64_t)blockIdx.x * DELTA_WAVES + (threadIdx.x >> 6);
    if (k >= count) return;
    const int64_t h = h0 + k;
    const int n = (int)len[h];
    const uint32_t* p = voxels + cap_off[h];
    float s = 0.0f, c = 0.0f;
#pragma unroll
    for (int r = 0; r < DELTA_SLOTS; ++r) {
        int j = lane + 64 * r;
        if (j < n) kahan_add(s, c, x[p[j]]);
    }
    for (int j = lane + 64 * DELTA_SLOTS; j < n; j += 64) kahan_add(s, c, x[p[j]]);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        float os = __shfl_xor(s, d, 64), oc = __shfl_xor(c, d, 64);
        c += oc;
        kahan_add(s, c, os);
    }
    if (lane == 0) {
        const float L = (float)chord[h];
        const float a_i_dot_x_k = s * L;
        const float a_i_dot_a_i = (L * L) * (float)n;
        delta[k] = L * ((wepl[h] - a_i_dot_x_k) / a_i_dot_a_i) * lambda;
    }
}

__global__ void __launch_bounds__(256)
k_block_apply_f32(const uint32_t* __restrict__ seg_voxel, const uint32_t* __restrict__ seg_start,
                  const uint16_t* __restrict__ csc, int64_t s0, int64_t ns, int64_t entry0, uint32_t n_entries,
                  const float* __restrict__ delta, float* __restrict__ x) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    int64_t s = s0 + k;
    uint32_t a = seg_start[s];
    uint32_t e = (k + 1 < ns) ? seg_start[s + 1] : n_entries;
    const uint16_t* c = csc + entry0;
    float upd = 0.0f, comp = 0.0f;
    for (uint32_t i = a; i < e; ++i) kahan_add(upd, comp, delta[c[i]]);
    uint32_t v = seg_voxel[s];
    x[v] = x[v] + upd / (float)(e - a);
}

}  // namespace

/* ------------------------------------------------------------------------- */
/* host                                                                       */
/* ------------------------------------------------------------------------- */
static void plan_release_graph(pct_plan* pl) {
    if (pl->graph) { (void)hipGraphExecDestroy(pl->graph); pl->graph = nullptr; }
}

void pct_plan_destroy(pct_plan* pl) {
    if (!pl) return;
    plan_release_graph(pl);
    pct_buffer* bufs[] = {&pl->hash, &pl->ex_off, &pl->blk_entry, &pl->blk_seg, &pl->seg_voxel, &pl->seg_start,
                          &pl->csc, &pl->delta};
    for (pct_buffer* b : bufs)
        if (b->ptr) (void)hipFree(b->ptr);
    delete pl;
}

static int vbits_for(int64_t V) {
    int b = 1;
    while (((int64_t)1 << b) < V) ++b;
    return b;
}

/* Build (or rebuild) the plan of `hist` with blocks of B histories. */
static int plan_build(pct_ctx* ctx, const pct_hist_soa* hist, int B, int mlp_mode) {
    const int64_t n = hist->n;
    if (B <= 0 || B > 65536) return pct_fail(ctx, PCT_E_INVALID, "DROP block size %d outside [1, 65536]", B);
    const int vbits = vbits_for(ctx->g.V);
    if (vbits > 28) return pct_fail(ctx, PCT_E_INVALID, "image too large for the plan (%lld voxels)", (long long)ctx->g.V);
    if (!ctx->plan) ctx->plan = new pct_plan();
    pct_plan* pl = ctx->plan;
    plan_release_graph(pl);
    pl->valid = false;
    int rc;
    /* path store + hash */
    if ((rc = pct_ensure(ctx, ctx->offsets, sizeof(int64_t) * (n + 1)))) return rc;
    int64_t total = 0;
    if ((rc = pct_mlp_capacity_impl(ctx, hist, 0, n, (int64_t*)ctx->offsets.ptr, &total))) return rc;
    if ((rc = pct_ensure(ctx, ctx->path_store, sizeof(uint32_t) * (total > 0 ? total : 1)))) return rc;
    if ((rc = pct_ensure(ctx, ctx->path_len, sizeof(uint32_t) * n))) return rc;
    if ((rc = pct_ensure(ctx, ctx->chord, sizeof(double) * n))) return rc;
    if ((rc = pct_ensure(ctx, pl->hash, sizeof(uint32_t) * n))) return rc;
    const int64_t* cap_off = (const int64_t*)ctx->offsets.ptr;
    uint32_t* vox = (uint32_t*)ctx->path_store.ptr;
    uint32_t* len = (uint32_t*)ctx->path_len.ptr;
    if ((rc = pct_mlp_paths_impl(ctx, hist, 0, n, cap_off, vox, len, (double*)ctx->chord.ptr, (uint32_t*)pl->hash.ptr,
                                 nullptr, mlp_mode)))
        return rc;
    int flags = 0;
    if ((rc = pct_take_status(ctx, &flags))) return rc;
    if ((rc = pct_status_to_rc(ctx, flags))) return rc;
    /* exact entry offsets */
    if ((rc = pct_ensure(ctx, pl->ex_off, sizeof(int64_t) * (n + 1)))) return rc;
    if ((rc = pct_ensure(ctx, ctx->steps, sizeof(int32_t) * n))) return rc;
    hipLaunchKernelGGL(k_lengths_i32, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, (const uint32_t*)len, n,
                       (int32_t*)ctx->steps.ptr);
    int64_t E = 0;
    if ((rc = pct_exclusive_scan_i32(ctx, (const int32_t*)ctx->steps.ptr, n, (int64_t*)pl->ex_off.ptr, &E))) return rc;
    const int64_t nb = (n + B - 1) / B;
    if ((rc = pct_ensure(ctx, pl->blk_entry, sizeof(int64_t) * (nb + 1)))) return rc;
    if ((rc = pct_ensure(ctx, pl->blk_seg, sizeof(int64_t) * (nb + 1)))) return rc;
    hipLaunchKernelGGL(k_block_bases, dim3(pct_grid(nb + 1, 256)), dim3(256), 0, ctx->stream,
                       (const int64_t*)pl->ex_off.ptr, n, (int64_t)B, nb, (int64_t*)pl->blk_entry.ptr);
    std::vector<int64_t> blk_entry_h(nb + 1);
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(blk_entry_h.data(), pl->blk_entry.ptr, sizeof(int64_t) * (nb + 1),
                                      hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    /* CSC values; segment arrays sized for the worst case (one segment per entry) */
    if ((rc = pct_ensure(ctx, pl->csc, sizeof(uint16_t) * (E > 0 ? E : 1)))) return rc;
    /* chunks of blocks: key = (local block << vbits) | voxel must fit 32 bits,
     * and a chunk holds at most ~2^27 entries */
    int64_t max_blocks = (int64_t)1 << (32 - vbits);
    int64_t avg = nb > 0 ? (E + nb - 1) / nb : 1;
    int64_t by_entries = ((int64_t)1 << 27) / (avg > 0 ? avg : 1);
    int64_t chunk = max_blocks < by_entries ? max_blocks : by_entries;
    if (chunk < 1) chunk = 1;
    int64_t max_chunk_entries = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += chunk) {
        int64_t b1 = b0 + chunk < nb ? b0 + chunk : nb;
        int64_t e = blk_entry_h[b1] - blk_entry_h[b0];
        if (e > max_chunk_entries) max_chunk_entries = e;
    }
    size_t ce = (size_t)(max_chunk_entries > 0 ? max_chunk_entries : 1);
    size_t sort_tmp = 0;
    {
        uint32_t* k0 = nullptr;
        uint32_t* v0 = nullptr;
        PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(nullptr, sort_tmp, k0, k0, v0, v0, ce, 0, 32, ctx->stream));
    }
    pct_buffer kin, kout, vin, vout, fl, sidx, tmp;
    auto free_all = [&]() {
        for (pct_buffer* b : {&kin, &kout, &vin, &vout, &fl, &sidx, &tmp})
            if (b->ptr) (void)hipFree(b->ptr);
    };
    if ((rc = pct_ensure(ctx, kin, 4 * ce)) || (rc = pct_ensure(ctx, kout, 4 * ce)) || (rc = pct_ensure(ctx, vin, 4 * ce)) ||
        (rc = pct_ensure(ctx, vout, 4 * ce)) || (rc = pct_ensure(ctx, fl, 4 * ce)) ||
        (rc = pct_ensure(ctx, sidx, 8 * (ce + 1))) || (rc = pct_ensure(ctx, tmp, sort_tmp))) {
        free_all();
        return rc;
    }
    /* first pass: count segments per chunk to size the segment arrays */
    std::vector<int64_t> chunk_segs;
    int64_t S_total = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int64_t seg0 = 0;
        int ci = 0;
        for (int64_t b0 = 0; b0 < nb; b0 += chunk, ++ci) {
            int64_t b1 = b0 + chunk < nb ? b0 + chunk : nb;
            int64_t h0 = b0 * B, h1 = b1 * B < n ? b1 * B : n;
            int64_t entry0 = blk_entry_h[b0];
            int64_t Ec = blk_entry_h[b1] - entry0;
            if (Ec == 0) { if (pass == 0) chunk_segs.push_back(0); continue; }
            int endbit = vbits + vbits_for(b1 - b0 + 1);
            if (endbit > 32) endbit = 32;
            hipLaunchKernelGGL(k_emit_keys, dim3((unsigned)((h1 - h0 + 3) / 4)), dim3(256), 0, ctx->stream,
                               (const uint32_t*)vox, cap_off, (const uint32_t*)len, (const int64_t*)pl->ex_off.ptr,
                               h0, h1, (int64_t)B, b0, vbits, entry0, (uint32_t*)kin.ptr, (uint32_t*)vin.ptr);
            size_t tb = tmp.bytes;
            PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(tmp.ptr, tb, (uint32_t*)kin.ptr, (uint32_t*)kout.ptr,
                                                         (uint32_t*)vin.ptr, (uint32_t*)vout.ptr, (size_t)Ec, 0, endbit,
                                                         ctx->stream));
            hipLaunchKernelGGL(k_head_flags, dim3(pct_grid(Ec, 256)), dim3(256), 0, ctx->stream,
                               (const uint32_t*)kout.ptr, Ec, (int32_t*)fl.ptr);
            int64_t Sc = 0;
            if ((rc = pct_exclusive_scan_i32(ctx, (const int32_t*)fl.ptr, Ec, (int64_t*)sidx.ptr, &Sc))) {
                free_all();
                return rc;
            }
            if (pass == 0) {
                chunk_segs.push_back(Sc);
                continue;
            }
            hipLaunchKernelGGL(k_write_segments, dim3(pct_grid(Ec, 256)), dim3(256), 0, ctx->stream,
                               (const uint32_t*)kout.ptr, (const uint32_t*)vout.ptr, (const int32_t*)fl.ptr,
                               (const int64_t*)sidx.ptr, Ec, vbits, seg0, b0, entry0,
                               (const int64_t*)pl->blk_entry.ptr, (uint32_t*)pl->seg_voxel.ptr,
                               (uint32_t*)pl->seg_start.ptr, (uint16_t*)pl->csc.ptr, (int64_t*)pl->blk_seg.ptr);
            PCT_HIP_CHECK(ctx, hipGetLastError());
            seg0 += Sc;
        }
        if (pass == 0) {
            S_total = 0;
            for (int64_t s : chunk_segs) S_total += s;
            if ((rc = pct_ensure(ctx, pl->seg_voxel, sizeof(uint32_t) * (S_total > 0 ? S_total : 1))) ||
                (rc = pct_ensure(ctx, pl->seg_start, sizeof(uint32_t) * (S_total > 0 ? S_total : 1)))) {
                free_all();
                return rc;
            }
        }
    }
    free_all();
    /* sentinel */
    PCT_HIP_CHECK(ctx, hipMemcpyAsync((int64_t*)pl->blk_seg.ptr + nb, &S_total, sizeof(int64_t), hipMemcpyHostToDevice,
                                      ctx->stream));
    pl->blk_entry_h = blk_entry_h;
    pl->blk_seg_h.assign(nb + 1, 0);
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(pl->blk_seg_h.data(), pl->blk_seg.ptr, sizeof(int64_t) * (nb + 1),
                                      hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if ((rc = pct_ensure(ctx, pl->delta, sizeof(double) * B))) return rc;
    pl->n = n;
    pl->B = B;
    pl->nb = nb;
    pl->E = E;
    pl->S = S_total;
    pl->key = *hist;
    pl->mlp_mode = mlp_mode;
    pl->valid = true;
    return PCT_OK;
}

static bool same_hist(const pct_hist_soa& a, const pct_hist_soa& b) {
    return a.n == b.n && a.x_entry == b.x_entry && a.wepl == b.wepl && a.first_voxel == b.first_voxel &&
           a.xy_entry_angle == b.xy_entry_angle && a.x_exit == b.x_exit;
}

/* the block loop of one iteration on the plan */
static int plan_block_loop(pct_ctx* ctx, pct_plan* pl, const pct_hist_soa* hist, float* x, double lambda, bool f32) {
    const int64_t B = pl->B;
    const uint32_t* vox = (const uint32_t*)ctx->path_store.ptr;
    const int64_t* cap_off = (const int64_t*)ctx->offsets.ptr;
    const uint32_t* len = (const uint32_t*)ctx->path_len.ptr;
    const double* chord = (const double*)ctx->chord.ptr;
    double* delta = (double*)pl->delta.ptr;
    for (int64_t b = 0; b < pl->nb; ++b) {
        int64_t h0 = b * B, cnt = (pl->n - h0) < B ? (pl->n - h0) : B;
        dim3 dgrid((unsigned)((cnt + DELTA_WAVES - 1) / DELTA_WAVES));
        if (f32)
            hipLaunchKernelGGL(k_block_delta_f32, dgrid, dim3(64 * DELTA_WAVES), 0, ctx->stream, vox, cap_off, len, chord,
                               (const float*)hist->wepl, h0, cnt, (const float*)x, (float)lambda, (float*)delta);
        else
            hipLaunchKernelGGL(k_block_delta, dgrid, dim3(64 * DELTA_WAVES), 0, ctx->stream, vox, cap_off, len, chord,
                               (const float*)hist->wepl, h0, cnt, (const float*)x, lambda, delta);
        int64_t s0 = pl->blk_seg_h[b], ns = pl->blk_seg_h[b + 1] - s0;
        uint32_t ne = (uint32_t)(pl->blk_entry_h[b + 1] - pl->blk_entry_h[b]);
        if (ns > 0 && f32)
            hipLaunchKernelGGL(k_block_apply_f32, dim3(pct_grid(ns, 256)), dim3(256), 0, ctx->stream,
                               (const uint32_t*)pl->seg_voxel.ptr, (const uint32_t*)pl->seg_start.ptr,
                               (const uint16_t*)pl->csc.ptr, s0, ns, pl->blk_entry_h[b], ne, (const float*)delta, x);
        else if (ns > 0)
            hipLaunchKernelGGL(k_block_apply, dim3(pct_grid(ns, 256)), dim3(256), 0, ctx->stream,
                               (const uint32_t*)pl->seg_voxel.ptr, (const uint32_t*)pl->seg_start.ptr,
                               (const uint16_t*)pl->csc.ptr, s0, ns, pl->blk_entry_h[b], ne, (const double*)delta, x);
    }
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

static int plan_run_graph(pct_ctx* ctx, pct_plan* pl, const pct_hist_soa* hist, float* x, double lambda, bool f32) {
    if (!pl->graph || pl->graph_x != x || pl->graph_lambda != lambda || pl->graph_f32 != f32) {
        plan_release_graph(pl);
        hipGraph_t g = nullptr;
        PCT_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = plan_block_loop(ctx, pl, hist, x, lambda, f32);
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return pct_fail(ctx, PCT_E_HIP, "graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&pl->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) { pl->graph = nullptr; return pct_fail(ctx, PCT_E_HIP, "graph instantiate: %s", hipGetErrorString(e)); }
        pl->graph_x = x;
        pl->graph_lambda = lambda;
        pl->graph_f32 = f32;
    }
    PCT_HIP_CHECK(ctx, hipGraphLaunch(pl->graph, ctx->stream));
    return PCT_OK;
}

extern "C" int pct_drop_prepare(pct_ctx* ctx, const pct_hist_soa* hist, int block_size, int mode) {
    if (!ctx || !hist || !hist->wepl || !hist->first_voxel || !hist->x_entry)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_prepare: null argument");
    const int mlp = mode & PCT_MLP_MODE_MASK;
    if (mlp != PCT_MLP_FP64 && mlp != PCT_MLP_TABLE)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_prepare: unknown MLP mode %d", mlp);
    if (hist->n <= 0) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_prepare: no histories");
    return plan_build(ctx, hist, block_size, mlp);
}

extern "C" int pct_drop_iteration(pct_ctx* ctx, const pct_hist_soa* hist, float* x, int block_size, float lambda, int mode) {
    if (!ctx || !hist || !hist->wepl || !hist->first_voxel || !hist->x_entry)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_iteration: null argument");
    const int mlp = mode & PCT_MLP_MODE_MASK;
    if (mlp != PCT_MLP_FP64 && mlp != PCT_MLP_TABLE)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_iteration: unknown MLP mode %d", mlp);
    if (!x || block_size <= 0) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_iteration: bad x or block size");
    if (hist->n <= 0) return PCT_OK;
    if (mode & PCT_DROP_ATOMIC) return pct_drop_iteration_atomic(ctx, hist, x, block_size, lambda, mlp);
    int rc;
    pct_plan* pl = ctx->plan;
    bool fresh = false;
    if (!pl || !pl->valid || pl->B != block_size || pl->mlp_mode != mlp || !same_hist(pl->key, *hist)) {
        if ((rc = plan_build(ctx, hist, block_size, mlp))) return rc;
        pl = ctx->plan;
        fresh = true;
    }
    if (ctx->profile) PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    if (ctx->profile) PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    if (!(mode & PCT_DROP_CACHED) && !fresh) {
        /* the reference recomputes every MLP path in every iteration; do the
         * same and check the lists against the plan */
        rc = pct_mlp_paths_impl(ctx, hist, 0, hist->n, (const int64_t*)ctx->offsets.ptr, (uint32_t*)ctx->path_store.ptr,
                                (uint32_t*)ctx->path_len.ptr, (double*)ctx->chord.ptr, nullptr,
                                (const uint32_t*)pl->hash.ptr, mlp);
        if (rc) return rc;
    }
    if (ctx->profile) PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    const bool f32 = (mode & PCT_DROP_F32) != 0;
    if (mode & PCT_DROP_NO_GRAPH) rc = plan_block_loop(ctx, pl, hist, x, (double)lambda, f32);
    else rc = plan_run_graph(ctx, pl, hist, x, (double)lambda, f32);
    if (rc) return rc;
    if (ctx->profile) {
        PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
        ctx->ev_recorded = true;
    }
    int flags = 0;
    if ((rc = pct_take_status(ctx, &flags))) return rc;
    return pct_status_to_rc(ctx, flags);
}

extern "C" int pct_drop_stats(pct_ctx* ctx, const pct_hist_soa* hist, int block_size, int64_t* sum_path,
                              int64_t* sum_touched) {
    if (!ctx || !hist || !sum_path || !sum_touched) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_stats: null argument");
    pct_plan* pl = ctx->plan;
    if (!pl || !pl->valid || pl->B != block_size || !same_hist(pl->key, *hist)) {
        int rc = plan_build(ctx, hist, block_size, pl ? pl->mlp_mode : PCT_MLP_FP64);
        if (rc) return rc;
        pl = ctx->plan;
    }
    *sum_path = pl->E;
    *sum_touched = pl->S;
    return PCT_OK;
}
