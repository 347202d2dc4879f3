/*
 * pct_drop.hip -- (e) the DROP block-iterative update.
 *
 * Reference: image_reconstruction_GPU (.cu:6538-6866) launches, for every
 * DROP block of 3200 histories, block_update_GPU (.cu:5774-5841: MLP + a_i.x +
 * two atomics per path voxel, 10 workgroups of 320 threads) and then
 * image_update_GPU (.cu:5900-5938: a scan of the whole volume), with a host
 * synchronisation per block.
 *
 * Here an iteration is split in two phases:
 *   1. the MLP voxel lists of ALL histories (they do not depend on x) are
 *      computed in one full-GPU launch into a resident path store (CSR, one
 *      u32 per path voxel, capacity offsets from a prefix sum of the step
 *      counts);
 *   2. the sequential block loop only does what depends on x: per block one
 *      wavefront per history gathers x over its list (coalesced list reads,
 *      FP64 wave reduction), forms delta_i = L (b_i - L sum x) / (n L^2) lambda
 *      (the reference's expression, .cu:5825-5834) and scatter-adds it with
 *      hardware float atomics; the image update then scans the volume with
 *      16-byte loads of the counts. Blocks stay in storage order and are never
 *      merged, so the algorithm is the reference's.
 */
#include <algorithm>
#include <vector>
#include "pct_internal.h"

int pct_mlp_capacity_impl(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, int64_t* offsets,
                          int64_t* total);
int pct_mlp_paths_impl(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, const int64_t* offsets,
                       uint32_t* voxels, uint32_t* lengths, double* chord, uint32_t* hash_out,
                       const uint32_t* hash_expect, int mode);
int pct_status_to_rc(pct_ctx* ctx, int flags);

namespace {

#define DROP_WAVES_PER_WG 4
#define DROP_MAX_SLOTS 16   /* register slots per lane: paths up to 1024 voxels in one read */

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

/* one wavefront per history of the block [0, count) (local indices into the
 * path store arrays) */
__global__ void __launch_bounds__(64 * DROP_WAVES_PER_WG)
k_drop_scatter(const uint32_t* __restrict__ voxels, const int64_t* __restrict__ offsets,
               const uint32_t* __restrict__ lengths, const double* __restrict__ chord, const float* __restrict__ wepl,
               int64_t count, const float* __restrict__ x, float* __restrict__ x_update, uint32_t* __restrict__ counts,
               float lambda) {
    const int lane = threadIdx.x & 63;
    const int64_t h = (int64_t)blockIdx.x * DROP_WAVES_PER_WG + (threadIdx.x >> 6);
    if (h >= count) return;
    const int n = (int)lengths[h];
    const int64_t off = offsets[h];
    uint32_t v[DROP_MAX_SLOTS];
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < DROP_MAX_SLOTS; ++r) {
        int j = lane + 64 * r;
        v[r] = 0;
        if (j < n) {
            v[r] = voxels[off + j];
            s += (double)x[v[r]];
        }
    }
    for (int j = lane + 64 * DROP_MAX_SLOTS; j < n; j += 64) s += (double)x[voxels[off + j]];
    s = wave_sum(s);
    const double L = chord[h];
    const double a_i_dot_x_k = s * L;
    const double a_i_dot_a_i = (L * L) * n;
    const double b_i = wepl[h];
    const float delta = (float)(L * ((b_i - a_i_dot_x_k) / a_i_dot_a_i) * lambda);
#pragma unroll
    for (int r = 0; r < DROP_MAX_SLOTS; ++r) {
        int j = lane + 64 * r;
        if (j < n) {
            unsafeAtomicAdd(&x_update[v[r]], delta);
            atomicAdd(&counts[v[r]], 1u);
        }
    }
    for (int j = lane + 64 * DROP_MAX_SLOTS; j < n; j += 64) {
        uint32_t vv = voxels[off + j];
        unsafeAtomicAdd(&x_update[vv], delta);
        atomicAdd(&counts[vv], 1u);
    }
}

/* image_update_GPU (.cu:5900-5938): x += x_update / count where count > 0,
 * then clear both. Four voxels per thread, 16-byte loads of the counts.
 * STATS: leave x alone, only clear the counts and count the touched voxels. */
template <bool STATS>
__global__ void __launch_bounds__(256)
k_drop_apply(float* __restrict__ x, float* __restrict__ x_update, uint32_t* __restrict__ counts, int64_t V,
             bool vec, unsigned long long* __restrict__ stats) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t v0 = q * 4;
    int touched = 0;
    if (v0 < V) {
        if (vec && v0 + 4 <= V) {
            uint4 c = *reinterpret_cast<const uint4*>(counts + v0);
            if ((c.x | c.y | c.z | c.w) != 0) {
                uint32_t cc[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (cc[k]) {
                        touched++;
                        if (!STATS) {
                            x[v0 + k] = x[v0 + k] + x_update[v0 + k] / (float)cc[k];
                            x_update[v0 + k] = 0.0f;
                        }
                    }
                }
                *reinterpret_cast<uint4*>(counts + v0) = make_uint4(0, 0, 0, 0);
            }
        } else {
            for (int64_t v = v0; v < V && v < v0 + 4; ++v) {
                uint32_t c = counts[v];
                if (c) {
                    touched++;
                    if (!STATS) {
                        x[v] = x[v] + x_update[v] / (float)c;
                        x_update[v] = 0.0f;
                    }
                    counts[v] = 0;
                }
            }
        }
    }
    if (STATS) {
        for (int d = 32; d >= 1; d >>= 1) touched += __shfl_xor(touched, d, 64);
        if ((threadIdx.x & 63) == 0 && touched) atomicAdd(&stats[1], (unsigned long long)touched);
    }
}

/* STATS variant of the scatter: counts only, and the path entries */
__global__ void __launch_bounds__(256)
k_drop_count(const uint32_t* __restrict__ voxels, const int64_t* __restrict__ offsets,
             const uint32_t* __restrict__ lengths, int64_t count, uint32_t* __restrict__ counts,
             unsigned long long* __restrict__ stats) {
    int64_t h = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (h >= count) return;
    int n = (int)lengths[h];
    int64_t off = offsets[h];
    for (int j = lane; j < n; j += 64) atomicAdd(&counts[voxels[off + j]], 1u);
    if (lane == 0) atomicAdd(&stats[0], (unsigned long long)n);
}

}  // namespace

static int launch_scatter(pct_ctx* ctx, const uint32_t* voxels, const int64_t* offsets, const uint32_t* lengths,
                          const double* chord, const float* wepl, int64_t count, const float* x, float* x_update,
                          uint32_t* counts, float lambda) {
    if (count <= 0) return PCT_OK;
    unsigned grid = (unsigned)((count + DROP_WAVES_PER_WG - 1) / DROP_WAVES_PER_WG);
    hipLaunchKernelGGL(k_drop_scatter, dim3(grid), dim3(64 * DROP_WAVES_PER_WG), 0, ctx->stream, voxels, offsets,
                       lengths, chord, wepl, count, x, x_update, counts, lambda);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

/* the scatter of pct_drop_block_update for lists already made (shared with
 * its allocation-free form, pct_noalloc.hip) */
int pct_drop_launch_scatter(pct_ctx* ctx, const uint32_t* voxels, const int64_t* offsets, const uint32_t* lengths,
                            const double* chord, const float* wepl, int64_t count, const float* x, float* x_update,
                            uint32_t* counts, float lambda) {
    return launch_scatter(ctx, voxels, offsets, lengths, chord, wepl, count, x, x_update, counts, lambda);
}

static int launch_apply(pct_ctx* ctx, float* x, float* x_update, uint32_t* counts, bool stats = false) {
    int64_t V = ctx->g.V;
    bool vec = ((uintptr_t)counts & 15) == 0;
    dim3 grid(pct_grid((V + 3) / 4, 256));
    if (stats)
        hipLaunchKernelGGL(k_drop_apply<true>, grid, dim3(256), 0, ctx->stream, x, x_update, counts, V, vec, ctx->d_stats);
    else
        hipLaunchKernelGGL(k_drop_apply<false>, grid, dim3(256), 0, ctx->stream, x, x_update, counts, V, vec, nullptr);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

static int ensure_accumulators(pct_ctx* ctx) {
    size_t bytes = sizeof(float) * (size_t)((ctx->g.V + 3) & ~3LL);
    bool fresh = ctx->xupd.ptr == nullptr;
    int rc = pct_ensure(ctx, ctx->xupd, bytes);
    if (rc) return rc;
    rc = pct_ensure(ctx, ctx->counts, bytes);
    if (rc) return rc;
    if (fresh) {
        PCT_HIP_CHECK(ctx, hipMemsetAsync(ctx->xupd.ptr, 0, ctx->xupd.bytes, ctx->stream));
        PCT_HIP_CHECK(ctx, hipMemsetAsync(ctx->counts.ptr, 0, ctx->counts.bytes, ctx->stream));
    }
    return PCT_OK;
}

static int check_drop_hist(pct_ctx* ctx, const pct_hist_soa* h, const char* who) {
    if (!ctx || !h) return pct_fail(ctx, PCT_E_INVALID, "%s: null argument", who);
    if (!h->wepl || !h->first_voxel || !h->x_entry)
        return pct_fail(ctx, PCT_E_INVALID, "%s: missing history field", who);
    return PCT_OK;
}

extern "C" int pct_drop_block_update(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count,
                                     const float* x, float* x_update, uint32_t* counts, float lambda, int mode) {
    int rc = check_drop_hist(ctx, hist, "pct_drop_block_update");
    if (rc) return rc;
    if (!pct_mlp_mode_ok(ctx, mode))
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_block_update: mode %d not available", mode);
    if (!x || !x_update || !counts) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_block_update: null image array");
    if (start < 0 || count < 0 || start + count > hist->n)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_block_update: bad range");
    if (count == 0) return PCT_OK;
    rc = pct_ensure(ctx, ctx->blk_offsets, sizeof(int64_t) * (count + 1));
    if (rc) return rc;
    int64_t total = 0;
    rc = pct_mlp_capacity_impl(ctx, hist, start, count, (int64_t*)ctx->blk_offsets.ptr, &total);
    if (rc) return rc;
    if ((rc = pct_ensure(ctx, ctx->blk_voxels, sizeof(uint32_t) * (total > 0 ? total : 1)))) return rc;
    if ((rc = pct_ensure(ctx, ctx->blk_lengths, sizeof(uint32_t) * count))) return rc;
    if ((rc = pct_ensure(ctx, ctx->blk_chord, sizeof(double) * count))) return rc;
    rc = pct_mlp_paths_impl(ctx, hist, start, count, (const int64_t*)ctx->blk_offsets.ptr, (uint32_t*)ctx->blk_voxels.ptr,
                            (uint32_t*)ctx->blk_lengths.ptr, (double*)ctx->blk_chord.ptr, nullptr, nullptr, mode);
    if (rc) return rc;
    rc = launch_scatter(ctx, (const uint32_t*)ctx->blk_voxels.ptr, (const int64_t*)ctx->blk_offsets.ptr,
                        (const uint32_t*)ctx->blk_lengths.ptr, (const double*)ctx->blk_chord.ptr, hist->wepl + start,
                        count, x, x_update, counts, lambda);
    if (rc) return rc;
    int flags = 0;
    rc = pct_take_status(ctx, &flags);
    if (rc) return rc;
    return pct_status_to_rc(ctx, flags);
}

extern "C" int pct_drop_image_update(pct_ctx* ctx, float* x, float* x_update, uint32_t* counts) {
    if (!ctx || !x || !x_update || !counts) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_image_update: null argument");
    return launch_apply(ctx, x, x_update, counts);
}

/* Reference structure of the iteration (.cu:6795-6857), chunk by chunk: the
 * MLP lists of a chunk of whole blocks into a path store of at most
 * `store_limit` bytes (capacity offsets from the step counts), then per block
 * the atomic scatter and the volume scan. Chunks never split a block, so the
 * blocks run in storage order as in the reference; a chunk holds at least one
 * block whatever the limit. Used for PCT_DROP_ATOMIC and, automatically, for
 * scans whose DROP plan would not fit the device (pct_plan.hip). */
int pct_drop_iteration_chunked(pct_ctx* ctx, const pct_hist_soa* hist, float* x, int block_size, float lambda,
                               int mode, int64_t store_limit) {
    int rc = check_drop_hist(ctx, hist, "pct_drop_iteration");
    if (rc) return rc;
    const int64_t n = hist->n;
    if (n <= 0) return PCT_OK;
    if ((rc = ensure_accumulators(ctx))) return rc;
    float* xu = (float*)ctx->xupd.ptr;
    uint32_t* cn = (uint32_t*)ctx->counts.ptr;
    /* step counts of every history, then chunks of blocks whose capacity
     * fits the limit */
    if ((rc = pct_ensure(ctx, ctx->offsets, sizeof(int64_t) * (n + 1)))) return rc;
    int64_t total = 0;
    if (ctx->profile) PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    if ((rc = pct_mlp_capacity_impl(ctx, hist, 0, n, (int64_t*)ctx->offsets.ptr, &total))) return rc;
    std::vector<int64_t> cap((size_t)((n + block_size - 1) / block_size) + 1);
    {
        std::vector<int64_t> all((size_t)n + 1);
        PCT_HIP_CHECK(ctx, hipMemcpyAsync(all.data(), ctx->offsets.ptr, sizeof(int64_t) * (n + 1),
                                          hipMemcpyDeviceToHost, ctx->stream));
        PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        for (size_t b = 0; b + 1 < cap.size(); ++b) cap[b] = all[b * (size_t)block_size];
        cap.back() = all[(size_t)n];
    }
    const int64_t nb = (int64_t)cap.size() - 1;
    const int64_t slots_limit = std::max<int64_t>(1, store_limit / (int64_t)sizeof(uint32_t));
    int64_t max_slots = 0, max_hist = 0;
    std::vector<int64_t> chunk_b(1, 0);
    while (chunk_b.back() < nb) {
        int64_t b0 = chunk_b.back(), b1 = b0 + 1;
        while (b1 < nb && cap[b1 + 1] - cap[b0] <= slots_limit) ++b1;
        chunk_b.push_back(b1);
        max_slots = std::max(max_slots, cap[b1] - cap[b0]);
        max_hist = std::max(max_hist, std::min(n, b1 * block_size) - b0 * block_size);
    }
    if ((rc = pct_ensure(ctx, ctx->path_store, sizeof(uint32_t) * (size_t)(max_slots > 0 ? max_slots : 1))) ||
        (rc = pct_ensure(ctx, ctx->blk_offsets, sizeof(int64_t) * (size_t)(max_hist + 1))) ||
        (rc = pct_ensure(ctx, ctx->path_len, sizeof(uint32_t) * (size_t)max_hist)) ||
        (rc = pct_ensure(ctx, ctx->chord, sizeof(double) * (size_t)max_hist)))
        return rc;
    const uint32_t* vox = (const uint32_t*)ctx->path_store.ptr;
    const uint32_t* len = (const uint32_t*)ctx->path_len.ptr;
    const double* chord = (const double*)ctx->chord.ptr;
    int64_t* offs = (int64_t*)ctx->blk_offsets.ptr;
    for (size_t c = 0; c + 1 < chunk_b.size(); ++c) {
        const int64_t h0 = chunk_b[c] * block_size, h1 = std::min(n, chunk_b[c + 1] * block_size);
        /* capacity offsets of the chunk, then its lists */
        if ((rc = pct_mlp_capacity_impl(ctx, hist, h0, h1 - h0, offs, &total))) return rc;
        if ((rc = pct_mlp_paths_impl(ctx, hist, h0, h1 - h0, offs, (uint32_t*)vox, (uint32_t*)len, (double*)chord,
                                     nullptr, nullptr, mode)))
            return rc;
        if (c == 0 && ctx->profile) {
            PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
            PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
        }
        for (int64_t s0 = h0; s0 < h1; s0 += block_size) {
            const int64_t cnt = (h1 - s0) < block_size ? (h1 - s0) : block_size;
            const int64_t l = s0 - h0;
            if ((rc = launch_scatter(ctx, vox, offs + l, len + l, chord + l, hist->wepl + s0, cnt, x, xu, cn, lambda)))
                return rc;
            if ((rc = launch_apply(ctx, x, xu, cn))) return rc;
        }
    }
    if (ctx->profile) {
        PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
        ctx->ev_recorded = true;
    }
    int flags = 0;
    if ((rc = pct_take_status(ctx, &flags))) return rc;
    return pct_status_to_rc(ctx, flags);
}

int pct_drop_iteration_atomic(pct_ctx* ctx, const pct_hist_soa* hist, float* x, int block_size, float lambda,
                              int mode) {
    /* the whole scan in one chunk when its lists take at most 8 GiB */
    return pct_drop_iteration_chunked(ctx, hist, x, block_size, lambda, mode, (int64_t)8 << 30);
}

extern "C" int pct_set_profiling(pct_ctx* ctx, int on) {
    if (!ctx) return PCT_E_INVALID;
    if (on && !ctx->ev[0])
        for (hipEvent_t& e : ctx->ev) PCT_HIP_CHECK(ctx, hipEventCreate(&e));
    ctx->profile = on != 0;
    ctx->ev_recorded = false;
    return PCT_OK;
}

extern "C" int pct_last_timings(pct_ctx* ctx, double* ms3) {
    if (!ctx || !ms3) return PCT_E_INVALID;
    if (!ctx->ev_recorded) return pct_fail(ctx, PCT_E_INVALID, "pct_last_timings: no profiled iteration");
    PCT_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev[3]));
    const int from[3] = {0, 0, 1}, to[3] = {1, 2, 3};
    for (int k = 0; k < 3; ++k) {
        float ms = 0.f;
        PCT_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev[from[k]], ctx->ev[to[k]]));
        ms3[k] = ms;
    }
    return PCT_OK;
}

