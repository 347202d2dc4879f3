/*
 * pct_research.hip -- the CPU research reconstruction loop over voxel lists
 * supplied by the caller (image_reconstruction, .cu:7413-7570): the lists and
 * per-intersection path errors the reference reads from its MLP_paths and
 * MLP_path_error files, one iteration per call, with the update pairs
 *
 *   PCT_LISTS_DROP          DROP_blocks            (.cu:8023-8043)
 *   PCT_LISTS_DROP_RED      DROP_blocks_red        (.cu:7998-8022)
 *   PCT_LISTS_ERROR_RED     DROP_blocks_error_red  (.cu:7816-7843)
 *   PCT_LISTS_ERROR         DROP_blocks_error      (.cu:7844-7870)
 *   PCT_LISTS_ERROR_RED_SP  DROP_blocks_error_red_sp (.cu:7768-7815)
 *
 * each followed by DROP_update (.cu:8044-8055) after every full block of
 * lists ((n + 1) % block_size == 0, .cu:7554). x_update (double) and counts
 * (uint32) are the caller's, as the reference's globals are: a partial last
 * block leaves its sums there and the next call's first update applies them.
 *
 * Results equal the reference's sequential loop bit for bit: per list one
 * wavefront forms the terms and one lane adds them in the list's order; the block's entries are
 * stably sorted by (block, voxel) once per call, so each voxel's run adds
 * its contributions onto x_update[v] in list order, as the loop does.
 * Arithmetic as written in the reference: the error variants weight x by
 * list position (x_k[i], not x_k[a_i[i]]), pow(c, 2.0) is c*c.
 */
#include "pct_internal.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace {

struct lists_args {
    const int64_t* offsets;
    const uint32_t* voxels;
    const float* errors;
    const double* chord;
    const float* wepl;
    int64_t V;
    double lambda, eta, psi_sign;
    int variant;
};

/* chord weight of intersection i of list k (entry e) */
__device__ __forceinline__ double entry_chord(const lists_args& a, int64_t k, int64_t i, int64_t e) {
    const double L = a.chord[k];
    switch (a.variant) {
    case PCT_LISTS_ERROR_RED: {
        double error = a.errors[e];
        return L + (error * a.eta);
    }
    case PCT_LISTS_ERROR: {
        float error = a.errors[e];
        return L + error * a.eta * L;
    }
    case PCT_LISTS_ERROR_RED_SP: return (i % 5 == 0) ? 1.8 : L;
    default: return L;
    }
}

/* one wavefront per list of the block: the lanes form the list's terms
 * (x_v, or c_i x_i and c_i^2) for up to 256 entries at a time into LDS, then
 * lane 0 adds them in list order -- the reference's sequential sums, bit for
 * bit. r[t] = the per-list factor of the update (the scaled residual, or the
 * residual and |a|^2 for the psi forms). */
constexpr int DOT_CHUNK = 256;
__global__ void __launch_bounds__(64)
k_lists_dot(lists_args a, int64_t k0, int64_t count, const float* __restrict__ x, double* __restrict__ r,
            double* __restrict__ aa_out, int* __restrict__ status) {
    __shared__ double s_dot[DOT_CHUNK], s_aa[DOT_CHUNK];
    const int64_t t = blockIdx.x;
    if (t >= count) return;
    const int lane = threadIdx.x;
    const int64_t k = k0 + t;
    const int64_t e0 = a.offsets[k], n = a.offsets[k + 1] - e0;
    const double L = a.chord[k];
    const bool by_voxel = a.variant == PCT_LISTS_DROP || a.variant == PCT_LISTS_DROP_RED;
    bool bad = false;
    double sum = 0.0, dot = 0.0, aa = 0.0;
    for (int64_t c0 = 0; c0 < n; c0 += DOT_CHUNK) {
        const int m = (int)(n - c0 < DOT_CHUNK ? n - c0 : DOT_CHUNK);
        for (int j = lane; j < m; j += 64) {
            const int64_t i = c0 + j;
            if (by_voxel) {
                const uint32_t v = a.voxels[e0 + i];
                s_dot[j] = v < a.V ? (double)x[v] : 0.0;  /* scalar_dot_product2 (.cu:7692-7698) */
                bad |= v >= a.V;
            } else {
                const double c = entry_chord(a, k, i, e0 + i);
                const double xi = i < a.V ? (double)x[i] : 0.0;
                bad |= i >= a.V;
                s_dot[j] = c * xi;
                s_aa[j] = c * c;
            }
        }
        __syncthreads();
        if (lane == 0) {
            if (by_voxel)
                for (int j = 0; j < m; ++j) sum += s_dot[j];
            else
                for (int j = 0; j < m; ++j) {
                    dot += s_dot[j];
                    aa += s_aa[j];
                }
        }
        __syncthreads();
    }
    if (bad) atomicOr(status, PCT_STATUS_OVERFLOW);
    if (lane != 0) return;
    if (by_voxel) {
        dot = L * sum;
        aa = (L * L) * (double)(uint32_t)n;
    }
    const double bi = a.wepl[k];
    const bool scaled = a.variant == PCT_LISTS_DROP || a.variant == PCT_LISTS_ERROR;
    r[t] = scaled ? (bi - dot) / aa : bi - dot;
    aa_out[t] = aa;
}

/* per sorted entry of the block: the first entry of each voxel's run adds
 * the run's contributions, in list order, onto x_update[v] and counts[v] */
__global__ void __launch_bounds__(256)
k_lists_accumulate(lists_args a, int64_t k0, int64_t e_begin, int64_t e_end, int vbits,
                   const uint64_t* __restrict__ keys, const uint64_t* __restrict__ entries,
                   const float* __restrict__ x, const double* __restrict__ r,
                   const double* __restrict__ aa, double* __restrict__ x_update, uint32_t* __restrict__ counts) {
    int64_t s = e_begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= e_end) return;
    const uint64_t key = keys[s];
    if (s > e_begin && keys[s - 1] == key) return;
    const uint64_t vmask = ((uint64_t)1 << vbits) - 1;
    const int64_t v = (int64_t)(key & vmask);
    if (v >= a.V) return;                               /* out-of-image entries (flagged by k_lists_dot) */
    const bool psi_form = a.variant != PCT_LISTS_DROP && a.variant != PCT_LISTS_ERROR;
    const double psi_i = ((1.0 - (double)x[v]) * a.eta) * a.psi_sign;
    double acc = x_update[v];
    uint32_t cnt = counts[v];
    for (int64_t q = s; q < e_end && keys[q] == key; ++q) {
        /* entry e, its list's index t in the block and its position i */
        const uint64_t pl = entries[q];
        const uint32_t e = (uint32_t)pl;
        const int64_t t = (int64_t)((pl >> 32) & 0xFFFFu);
        const int64_t i = (int64_t)(pl >> 48);
        const int64_t k = k0 + t;
        const double c = entry_chord(a, k, i, e);
        acc += psi_form ? c * a.lambda * (r[t] / (aa[t] + psi_i)) : c * a.lambda * r[t];
        cnt += 1;
    }
    x_update[v] = acc;
    counts[v] = cnt;
}

/* DROP_update over the voxels of the block's runs */
__global__ void __launch_bounds__(256)
k_lists_update_runs(int64_t V, int64_t e_begin, int64_t e_end, int vbits, const uint64_t* __restrict__ keys,
                    float* __restrict__ x, double* __restrict__ x_update, uint32_t* __restrict__ counts) {
    int64_t s = e_begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= e_end) return;
    const uint64_t key = keys[s];
    if (s > e_begin && keys[s - 1] == key) return;
    const int64_t v = (int64_t)(key & (((uint64_t)1 << vbits) - 1));
    if (v >= V || counts[v] == 0) return;
    x[v] = (float)(x[v] + x_update[v] / counts[v]);
    x_update[v] = 0.0;
    counts[v] = 0;
}

/* DROP_update over every voxel (the first update of a call, which also
 * applies what a previous call's partial block left) */
__global__ void __launch_bounds__(256)
k_lists_update_dense(int64_t V, float* __restrict__ x, double* __restrict__ x_update, uint32_t* __restrict__ counts) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || counts[v] == 0) return;
    x[v] = (float)(x[v] + x_update[v] / counts[v]);
    x_update[v] = 0.0;
    counts[v] = 0;
}

/* sort keys (block << vbits | voxel, out-of-image voxels as V) and the
 * payload of every entry: e | (list index in its block) << 32 | (position in
 * its list) << 48; one wavefront per list. Lists longer than 65535 are
 * flagged (the position has 16 bits). */
__global__ void __launch_bounds__(256)
k_lists_keys(const int64_t* __restrict__ offsets, const uint32_t* __restrict__ voxels, int64_t n, int64_t V,
             int block_size, int vbits, uint64_t* __restrict__ keys, uint64_t* __restrict__ entries,
             int* __restrict__ status) {
    int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int lane = threadIdx.x & 63;
    if (k >= n) return;
    const int64_t e0 = offsets[k], e1 = offsets[k + 1];
    if (lane == 0 && e1 - e0 > 0xFFFF) atomicOr(status, PCT_STATUS_CAPACITY);
    const uint64_t blk = (uint64_t)(k / block_size) << vbits;
    const uint64_t t = (uint64_t)(k % block_size) << 32;
    for (int64_t e = e0 + lane; e < e1; e += 64) {
        uint32_t v = voxels[e];
        keys[e] = blk | (uint64_t)(v < V ? v : V);
        entries[e] = (uint64_t)(uint32_t)e | t | ((uint64_t)((e - e0) & 0xFFFF) << 48);
    }
}

__global__ void k_block_offsets(const int64_t* __restrict__ offsets, int64_t n, int block_size, int64_t nblocks,
                                int64_t* __restrict__ out) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nblocks) return;
    int64_t k = b * block_size;
    out[b] = offsets[k < n ? k : n];
}

__global__ void k_offsets_ok(const int64_t* __restrict__ offsets, int64_t n, int* __restrict__ status) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && offsets[k + 1] < offsets[k]) atomicOr(status, PCT_STATUS_CAPACITY);
}

int bits_for(uint64_t v) {
    int b = 1;
    while (b < 63 && ((uint64_t)1 << b) <= v) ++b;
    return b;
}

}  // namespace

extern "C" int pct_drop_lists_iteration(pct_ctx* ctx, int64_t n, const int64_t* offsets, const uint32_t* voxels,
                                        const float* errors, const double* chord, const float* wepl, float* x,
                                        double* x_update, uint32_t* counts, int block_size, float lambda,
                                        int variant) {
    if (!ctx) return PCT_E_INVALID;
    const bool reuse = (variant & PCT_LISTS_REUSE) != 0;
    variant &= ~PCT_LISTS_REUSE;
    if (n < 0 || block_size <= 0 || block_size > 0xFFFF)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_lists_iteration: bad n, or block size outside 1..65535");
    if (variant < PCT_LISTS_DROP || variant > PCT_LISTS_ERROR_RED_SP)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_lists_iteration: unknown variant %d", variant);
    if (!x || !x_update || !counts) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_lists_iteration: null image or state");
    if (n == 0) return PCT_OK;
    if (!offsets || !chord || !wepl) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_lists_iteration: null list array");
    const int64_t V = ctx->g.V;
    int64_t ends[2];
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(&ends[0], offsets, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(&ends[1], offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ends[0] != 0 || ends[1] < 0) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_lists_iteration: offsets[0] must be 0");
    const int64_t E = ends[1];
    if (E >= (int64_t)UINT32_MAX || n >= (int64_t)UINT32_MAX)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_lists_iteration: more than 2^32 intersections or lists");
    /* the voxel and error arrays are read only when there are intersections */
    if (E > 0 && !voxels) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_lists_iteration: null voxel array");
    if (E > 0 && (variant == PCT_LISTS_ERROR_RED || variant == PCT_LISTS_ERROR) && !errors)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_lists_iteration: this variant needs path errors");
    int flags = 0;
    int rc = pct_take_status(ctx, &flags);
    if (rc) return rc;
    const bool cached = reuse && ctx->rs_offsets == (const void*)offsets && ctx->rs_voxels == (const void*)voxels &&
                        ctx->rs_n == n && ctx->rs_E == E && ctx->rs_block == block_size;
    if (!cached) {
        hipLaunchKernelGGL(k_offsets_ok, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, offsets, n, ctx->d_status);
        if ((rc = pct_take_status(ctx, &flags))) return rc;
        if (flags) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_lists_iteration: offsets decrease");
    }

    const int64_t nblocks = (n + block_size - 1) / block_size;
    const int vbits = bits_for((uint64_t)V);              /* V itself (out-of-image marker) fits */
    const int end_bit = vbits + bits_for((uint64_t)nblocks);
    if (end_bit > 64) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_lists_iteration: too many blocks");
    const size_t EE = (size_t)(E > 0 ? E : 1);
    pct_buffer* rb = ctx->research;
    if ((rc = pct_ensure(ctx, rb[0], 8 * EE)) || (rc = pct_ensure(ctx, rb[1], 8 * EE)) ||
        (rc = pct_ensure(ctx, rb[2], 8 * EE)) || (rc = pct_ensure(ctx, rb[3], 8 * EE)) ||
        (rc = pct_ensure(ctx, rb[5], 16 * (size_t)block_size)) || (rc = pct_ensure(ctx, rb[6], 8 * (size_t)(nblocks + 1))))
        return rc;
    uint64_t* keys_in = (uint64_t*)rb[0].ptr;
    uint64_t* keys = (uint64_t*)rb[1].ptr;
    uint64_t* ent_in = (uint64_t*)rb[2].ptr;
    uint64_t* entries = (uint64_t*)rb[3].ptr;
    double* r = (double*)rb[5].ptr;
    double* aa = r + block_size;
    int64_t* blk_off_d = (int64_t*)rb[6].ptr;

    if (!cached) ctx->rs_offsets = nullptr;         /* the sort below replaces the cached order */
    if (E > 0 && !cached) {
        hipLaunchKernelGGL(k_lists_keys, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, ctx->stream, offsets, voxels, n, V,
                           block_size, vbits, keys_in, ent_in, ctx->d_status);
        if ((rc = pct_take_status(ctx, &flags))) return rc;
        if (flags) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_lists_iteration: a list of more than 65535 voxels");
        size_t tb = 0;
        PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(nullptr, tb, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                                     (uint64_t*)nullptr, (uint64_t*)nullptr, (size_t)E, 0, end_bit,
                                                     ctx->stream));
        if ((rc = pct_ensure(ctx, rb[7], tb))) return rc;
        tb = rb[7].bytes;
        PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(rb[7].ptr, tb, keys_in, keys, ent_in, entries, (size_t)E, 0, end_bit,
                                                     ctx->stream));
    }
    if (!cached) {
        hipLaunchKernelGGL(k_block_offsets, dim3(pct_grid(nblocks + 1, 256)), dim3(256), 0, ctx->stream, offsets, n,
                           block_size, nblocks, blk_off_d);
        ctx->rs_blk_off.assign((size_t)nblocks + 1, 0);
        PCT_HIP_CHECK(ctx, hipMemcpyAsync(ctx->rs_blk_off.data(), blk_off_d, sizeof(int64_t) * (nblocks + 1),
                                          hipMemcpyDeviceToHost, ctx->stream));
        PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        ctx->rs_offsets = offsets;
        ctx->rs_voxels = voxels;
        ctx->rs_n = n;
        ctx->rs_E = E;
        ctx->rs_block = block_size;
    }
    const std::vector<int64_t>& blk_off = ctx->rs_blk_off;

    lists_args a{offsets, voxels, errors, chord, wepl, V, (double)lambda, ctx->p.eta, (double)ctx->p.psi_sign, variant};
    bool first_update = true;
    for (int64_t b = 0; b < nblocks; ++b) {
        const int64_t k0 = b * block_size;
        const int64_t count = (k0 + block_size <= n) ? block_size : n - k0;
        const int64_t e0 = blk_off[b], e1 = blk_off[b + 1];
        hipLaunchKernelGGL(k_lists_dot, dim3((unsigned)count), dim3(64), 0, ctx->stream, a, k0, count, x, r, aa,
                           ctx->d_status);
        if (e1 > e0)
            hipLaunchKernelGGL(k_lists_accumulate, dim3(pct_grid(e1 - e0, 256)), dim3(256), 0, ctx->stream, a, k0, e0, e1,
                               vbits, keys, entries, x, r, aa, x_update, counts);
        if (count == block_size) {                      /* (n + 1) % BLOCK_SIZE_RECON == 0 (.cu:7554) */
            if (first_update)
                hipLaunchKernelGGL(k_lists_update_dense, dim3(pct_grid(V, 256)), dim3(256), 0, ctx->stream, V, x,
                                   x_update, counts);
            else if (e1 > e0)
                hipLaunchKernelGGL(k_lists_update_runs, dim3(pct_grid(e1 - e0, 256)), dim3(256), 0, ctx->stream, V, e0,
                                   e1, vbits, keys, x, x_update, counts);
            first_update = false;
        }
    }
    PCT_HIP_CHECK(ctx, hipGetLastError());
    if ((rc = pct_take_status(ctx, &flags))) return rc;
    if (flags)
        return pct_fail(ctx, PCT_E_INVALID,
                        "pct_drop_lists_iteration: a voxel index, or (error variants) a list position, is outside the "
                        "image; those entries were skipped");
    return PCT_OK;
}
