/*
 * pct_plan.hip -- the DROP iteration over a resident per-scan plan.
 *
 * The voxel lists of the MLP do not depend on the image, so the structure of
 * every DROP block -- which voxels it touches, how many of its histories touch
 * each, and which ones -- is the same in every iteration. The plan stores it
 * once per scan:
 *
 *   path lists   block-strided: history k of block b at list_base[b] +
 *                k * stride[b], stride[b] = the longest list of the block,
 *                unused slots hold a sentinel; lengths, effective chord
 *                lengths and a hash of every list;
 *   block CSC    per DROP block the distinct voxels it touches as 8-byte
 *                records {voxel, count, up to 3 block-local history indices
 *                in history order}; longer segments point into a u16 array
 *                of all indices, come first in the block (each part sorted
 *                by voxel) and keep their first 8 indices in a side row.
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
 * The block loop is a chain of ~2 x 4.7k dependent launches per iteration at
 * 200x200x32, so its cost is launch latency plus the dependent memory round
 * trips of each kernel. The layout is chosen so that each kernel has exactly
 * two: the delta kernel finds a history's list from the launch arguments
 * (block base + k * stride) and reads the list and the history's scalars in
 * one round, then gathers x; the apply kernel reads one 8-byte record per voxel
 * (voxel + inline history indices; a long segment's side row beside it) and
 * then the deltas and the old x[v] together. The block loop is captured once into hipGraphs of chunks of
 * blocks and replayed. In the default mode every iteration recomputes every MLP
 * voxel list (as the reference does, .cu:5822) into the plan's list store and
 * checks it against the plan's hash; that FP64-bound pass runs on a second
 * stream, chunk by chunk ahead of the latency-bound block loop, and the graph
 * of a chunk waits for its own lists only. Blocks still run one after the
 * other in storage order and every block reads the lists computed in the
 * same iteration.
 */
#include <algorithm>
#include <rocprim/device/device_radix_sort.hpp>
#include "pct_internal.h"
#include "pct_plan.h"
#include "pct_plan_kernels.h"

int pct_mlp_capacity_impl(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, int64_t* offsets,
                          int64_t* total);
int pct_mlp_paths_impl(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, const int64_t* offsets,
                       uint32_t* voxels, uint32_t* lengths, double* chord, uint32_t* hash_out,
                       const uint32_t* hash_expect, int mode);
int pct_status_to_rc(pct_ctx* ctx, int flags);
int pct_mlp_fallback_reserve(pct_ctx* ctx, int64_t n);
int pct_drop_iteration_atomic(pct_ctx* ctx, const pct_hist_soa* hist, float* x, int block_size, float lambda,
                              int mode);
int pct_drop_iteration_chunked(pct_ctx* ctx, const pct_hist_soa* hist, float* x, int block_size, float lambda,
                               int mode, int64_t store_limit);
int pct_mlp_paths_on(pct_ctx* ctx, hipStream_t stream, const pct_hist_soa* hist, int64_t start, int64_t count,
                     const int64_t* offsets, uint32_t* voxels, uint32_t* lengths, double* chord, uint32_t* hash_out,
                     const uint32_t* hash_expect, int mode, bool beside);



namespace {

/* ------------------------------------------------------------------------- */
/* plan construction kernels                                                  */
/* ------------------------------------------------------------------------- */
__global__ void k_lengths_i32(const uint32_t* __restrict__ len, int64_t n, int32_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int32_t)len[i];
}

/* longest list of each block: one workgroup of 256 per block */
__global__ void __launch_bounds__(256)
k_block_max_len(const uint32_t* __restrict__ len, int64_t n, int64_t B, int32_t* __restrict__ out) {
    __shared__ uint32_t sm[256];
    const int64_t b = blockIdx.x;
    const int64_t h0 = b * B, h1 = (h0 + B) < n ? (h0 + B) : n;
    uint32_t m = 0;
    for (int64_t h = h0 + threadIdx.x; h < h1; h += blockDim.x) m = len[h] > m ? len[h] : m;
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] = sm[threadIdx.x] > sm[threadIdx.x + s] ? sm[threadIdx.x] : sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = (int32_t)sm[0];
}

/* placement key of every history: (block << 32) | the order-preserving bits
 * of its lateral coordinate t = (m x d) / |d|, m the chord midpoint and d the
 * chord direction in the xy plane (the distance of the path from the axis,
 * signed); the placement never changes results */
__global__ void k_lateral_keys(pct_hist_soa h, int64_t n, int64_t B, uint64_t* __restrict__ keys,
                               uint32_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float dx = h.x_exit[i] - h.x_entry[i], dy = h.y_exit[i] - h.y_entry[i];
    const float xm = 0.5f * (h.x_entry[i] + h.x_exit[i]), ym = 0.5f * (h.y_entry[i] + h.y_exit[i]);
    const float d = sqrtf(dx * dx + dy * dy);
    const float t = d > 0.0f ? (ym * dx - xm * dy) / d : 0.0f;
    uint32_t u = __float_as_uint(t);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    keys[i] = ((uint64_t)(i / B) << 32) | u;
    vals[i] = (uint32_t)i;
}

/* placement: the histories of each block sorted by (block, t) are split into
 * 8 contiguous lateral bands whose sizes are the numbers of positions q of the
 * block with (q / DELTA_WAVES) % 8 == g -- the positions the delta workgroups
 * of XCD g take (workgroup j runs on XCD j % 8) -- and band g fills those
 * positions in order. Writes per position the block-local index, the list
 * length, the chord length and the lateral rank r (the history's index in
 * the sorted order of its block), and per history its rank and the offset of
 * its row. */
__global__ void k_positions(const uint32_t* __restrict__ sorted, int64_t n, int64_t B, int64_t nb,
                            const int64_t* __restrict__ list_base, const int32_t* __restrict__ stride,
                            const uint32_t* __restrict__ len, const double* __restrict__ chord,
                            uint16_t* __restrict__ pos_k, uint32_t* __restrict__ pos_len, double* __restrict__ pos_chord,
                            uint16_t* __restrict__ pos_r, uint16_t* __restrict__ rank, int64_t* __restrict__ hist_off) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t h = sorted[i];
    const int64_t b = i / B;
    const int r = (int)(i - b * B);
    const int cnt = (int)((n - b * B) < B ? (n - b * B) : B);
    constexpr int W = DELTA_WAVES, CYC = 8 * DELTA_WAVES;
    const int full = cnt / CYC, rem = cnt % CYC;
    int g = 0, c0 = 0;
    for (; g < 8; ++g) {
        const int cg = W * full + min(W, max(0, rem - W * g));
        if (r < c0 + cg) break;
        c0 += cg;
    }
    const int m = r - c0;
    const int q = CYC * (m / W) + W * g + (m % W);
    const int64_t at = b * B + q;
    pos_k[at] = (uint16_t)(h - b * B);
    pos_len[at] = len[h];
    pos_chord[at] = chord[h];
    pos_r[at] = (uint16_t)r;
    rank[h] = (uint16_t)r;
    hist_off[h] = list_base[b] + (int64_t)q * stride[b];
    if (i == n - 1) hist_off[n] = list_base[nb];
}

/* grid-stride: a dispatch holds at most 2^32 - 1 work-items, and the list
 * store of a large scan has more entries than that */
__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, int64_t n, uint64_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

/* voxel -> segment maps of blocks [b0, b1) (one map per block, slot b % M):
 * map[slot][v] = (b << 32) | (segment index - blk_seg[b]); grid.y = block */
__global__ void __launch_bounds__(256)
k_seg_map(const uint64_t* __restrict__ segs, const int64_t* __restrict__ blk_seg, int64_t b0, int64_t M, int64_t V,
          uint64_t* __restrict__ map) {
    const int64_t b = b0 + blockIdx.y;
    const int64_t s0 = blk_seg[b], ns = blk_seg[b + 1] - s0;
    uint64_t* m = map + (b % M) * V;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ns; k += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)(segs[s0 + k] & (SEG_MAX_V - 1));
        m[v] = ((uint64_t)b << 32) | (uint64_t)k;
    }
}

/* one wavefront per history of [h0, h1): its voxel list (tiled indices)
 * into the block-strided list store and, when prev is given, slot for slot the
 * previous block's segment record of each voxel (0 when that block does not
 * touch it; looked up in the previous block's map); with tag, the slots whose
 * voxel the previous block touches carry LIST_TAG */
__global__ void __launch_bounds__(256)
k_build_entries(const uint32_t* __restrict__ path, const int64_t* __restrict__ cap_off,
                const uint32_t* __restrict__ len, const int64_t* __restrict__ hist_off, int64_t h0, int64_t h1,
                int64_t B, const uint64_t* __restrict__ segs, const int64_t* __restrict__ blk_seg, int64_t M, int64_t V,
                const uint64_t* __restrict__ map, uint32_t* __restrict__ list, uint64_t* __restrict__ prev,
                int tag, pct_blk::vox_tiling tl) {
    const int64_t h = h0 + (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (h >= h1) return;
    const int64_t b = h / B;
    const int n = (int)len[h];
    const uint32_t* src = path + cap_off[h];
    const int64_t dst = hist_off[h];
    const uint64_t* m = ((prev || tag) && b > 0) ? map + ((b - 1) % M) * V : nullptr;
    const int64_t sp = b > 0 ? blk_seg[b - 1] : 0;
    for (int j = lane; j < n; j += 64) {
        const uint32_t v = pct_blk::tiled_of(tl, src[j]);
        const uint64_t w = m ? m[v] : 0ull;
        const bool in_prev = m && (w >> 32) == (uint64_t)(b - 1);
        list[dst + j] = (tag && in_prev) ? (v | LIST_TAG) : v;
        if (prev) prev[dst + j] = in_prev ? segs[sp + (int64_t)(uint32_t)w] : 0ull;
    }
}

/* one wavefront per history of the chunk: keys = (local block << vbits) |
 * tiled voxel, values = the history's lateral rank in its block, at the
 * history's exact entry offset (histories in order, so a segment lists its
 * histories in history order after the stable sort) */
__global__ void __launch_bounds__(256)
k_emit_keys(const uint32_t* __restrict__ list, const int64_t* __restrict__ hist_off,
            const uint32_t* __restrict__ len, const int64_t* __restrict__ ex_off, int64_t h0, int64_t h1,
            int64_t B, int64_t b0, int vbits, int64_t entry0, const uint16_t* __restrict__ rank,
            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, pct_blk::vox_tiling tl) {
    int64_t h = h0 + (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (h >= h1) return;
    int64_t b = h / B;
    uint32_t hi = (uint32_t)(b - b0) << vbits;
    uint32_t local = rank[h];
    int n = (int)len[h];
    int64_t src = hist_off[h], dst = ex_off[h] - entry0;
    for (int j = lane; j < n; j += 64) {
        keys[dst + j] = hi | pct_blk::tiled_of(tl, list[src + j]);
        vals[dst + j] = local;
    }
}

__global__ void k_head_flags(const uint32_t* __restrict__ keys, int64_t E, int32_t* __restrict__ flags) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < E) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

/* the number of entries of the segment whose head is entry i (flags: heads) */
__device__ __forceinline__ uint32_t seg_count(const int32_t* __restrict__ flags, int64_t i, int64_t E) {
    int64_t j = i + 1;
    while (j < E && !flags[j]) ++j;               /* a new block always starts a new segment */
    return (uint32_t)(j - i);
}

/* long[i] = 1 at the head of a segment of more than SEG_INLINE entries */
__global__ void k_long_flags(const int32_t* __restrict__ flags, int64_t E, int32_t* __restrict__ lng) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < E) lng[i] = (flags[i] && seg_count(flags, i, E) > SEG_INLINE) ? 1 : 0;
}

/* segments of the chunk: the csc values of every entry, and at each segment
 * head the 8-byte record; a block's long segments (segidx / lngidx: exclusive
 * prefix counts of the heads / long heads over the chunk, nlong_chunk the
 * chunk's long heads) take its first record slots and a row of long_rows
 * each, its other segments the slots after them, each part in voxel order;
 * for a block's first entry blk_seg and blk_long */
__global__ void k_write_segments(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                 const int32_t* __restrict__ flags, const int64_t* __restrict__ segidx,
                                 const int64_t* __restrict__ lngidx, int64_t nlong_chunk, int64_t E, int vbits,
                                 int64_t seg0, int64_t long0, int64_t b0, int64_t entry0,
                                 const int64_t* __restrict__ blk_entry, uint64_t* __restrict__ segs,
                                 uint16_t* __restrict__ csc, pct_blk::long_row* __restrict__ long_rows,
                                 int64_t* __restrict__ blk_seg, int64_t* __restrict__ blk_long) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E) return;
    uint32_t k = keys[i];
    int64_t b = b0 + (int64_t)(k >> vbits);
    int64_t gi = entry0 + i;
    csc[gi] = (uint16_t)vals[i];
    if (!flags[i]) return;
    const int64_t eb = blk_entry[b] - entry0, eb1 = blk_entry[b + 1] - entry0;
    const int64_t seg_first = segidx[eb], long_first = lngidx[eb];
    const int64_t nlong = (eb1 < E ? lngidx[eb1] : nlong_chunk) - long_first;
    uint32_t c = seg_count(flags, i, E);
    uint64_t r = (uint64_t)(k & ((1u << vbits) - 1u));
    int64_t s;
    if (c <= SEG_INLINE) {
        r |= (uint64_t)c << SEG_VOX_BITS;
        for (uint32_t q = 0; q < c; ++q) r |= (uint64_t)(vals[i + q] & 0x1fffu) << (24 + SEG_IDX_BITS * q);
        s = seg0 + seg_first + nlong + ((segidx[i] - lngidx[i]) - (seg_first - long_first));
    } else {                                        /* inline count 0: indices in csc */
        r |= (uint64_t)(gi - blk_entry[b]) << 24;   /* block-relative first entry */
        r |= (uint64_t)c << 51;
        s = seg0 + seg_first + (lngidx[i] - long_first);
        pct_blk::long_row row;
#pragma unroll
        for (int t = 0; t < LONG_IDX / 2; ++t) {
            const uint32_t lo = vals[i + min(2u * t, c - 1u)];
            const uint32_t hi = vals[i + min(2u * t + 1u, c - 1u)];
            row.w[t] = (lo & 0xffffu) | (hi << 16);
        }
        long_rows[long0 + lngidx[i]] = row;
    }
    segs[s] = r;
    if (i == eb) {
        blk_seg[b] = seg0 + seg_first;
        blk_long[b] = long0 + long_first;
    }
}

/* ------------------------------------------------------------------------- */
/* iteration kernels                                                          */
/* ------------------------------------------------------------------------- */
/* the block loop's kernels: bodies in pct_plan_kernels.h. Both raise their
 * wave priority: they are latency-bound and share the CUs with the
 * long-running MLP waves of the next chunks. Their leading arguments are the
 * ones their first memory round needs, all scalars, so that the dispatch
 * preloads them into SGPRs (the library is compiled with kernel-argument
 * preloading); the rest are loaded while that round is in flight. */
/* SLOTS: list slots per lane in the first round, 4 for blocks whose rows
 * are at most 256 entries long (their launch then issues no slot past the
 * row), DELTA_SLOTS otherwise */
template <int SLOTS>
__global__ void __launch_bounds__(64 * DELTA_WAVES)
k_block_delta(const uint32_t* __restrict__ list, const uint16_t* __restrict__ pos_k,
              const uint32_t* __restrict__ pos_len, const double* __restrict__ pos_chord,
              const uint16_t* __restrict__ pos_r, int stride, int64_t count, const float* __restrict__ x,
              const float* __restrict__ wepl, int64_t h0, double lambda, double* __restrict__ delta) {
    __builtin_amdgcn_s_setprio(3);
    const int64_t q = (int64_t)blockIdx.x * DELTA_WAVES + (threadIdx.x >> 6);
    if (q >= count) return;
    pct_blk::delta_f64<SLOTS>(list, stride, {pos_k, pos_len, pos_chord, pos_r}, wepl, h0, q, x, lambda, delta,
                              threadIdx.x & 63);
}

template <int SLOTS>
__global__ void __launch_bounds__(64 * DELTA_WAVES)
k_block_delta_f32(const uint32_t* __restrict__ list, const uint16_t* __restrict__ pos_k,
                  const uint32_t* __restrict__ pos_len, const double* __restrict__ pos_chord,
                  const uint16_t* __restrict__ pos_r, int stride, int64_t count, const float* __restrict__ x,
                  const float* __restrict__ wepl, int64_t h0, float lambda, float* __restrict__ delta) {
    __builtin_amdgcn_s_setprio(3);
    const int64_t q = (int64_t)blockIdx.x * DELTA_WAVES + (threadIdx.x >> 6);
    if (q >= count) return;
    pct_blk::delta_f32<SLOTS>(list, stride, {pos_k, pos_len, pos_chord, pos_r}, wepl, h0, q, x, lambda, delta,
                              threadIdx.x & 63);
}

/* segs: the block's records (its long segments first), rows: their side rows */
__global__ void __launch_bounds__(APPLY_THREADS)
k_block_apply(const uint64_t* __restrict__ segs, const pct_blk::long_row* __restrict__ rows, int64_t nlong, int64_t ns,
              const uint16_t* __restrict__ csc, const double* __restrict__ delta, float* __restrict__ x) {
    __builtin_amdgcn_s_setprio(3);
    int64_t k = (int64_t)blockIdx.x * APPLY_THREADS + threadIdx.x;
    if (k >= ns) return;
    pct_blk::apply_seg<false>(k, nlong, segs[k], rows, csc, delta, x);
}

__global__ void __launch_bounds__(APPLY_THREADS)
k_block_apply_f32(const uint64_t* __restrict__ segs, const pct_blk::long_row* __restrict__ rows, int64_t nlong,
                  int64_t ns, const uint16_t* __restrict__ csc, const float* __restrict__ delta,
                  float* __restrict__ x) {
    __builtin_amdgcn_s_setprio(3);
    int64_t k = (int64_t)blockIdx.x * APPLY_THREADS + threadIdx.x;
    if (k >= ns) return;
    pct_blk::apply_seg<true>(k, nlong, segs[k], rows, csc, delta, x);
}

/* the research variants of the FP64 update (pct_plan_kernels.h) */
template <int V>
__global__ void __launch_bounds__(64 * DELTA_WAVES)
k_block_delta_var(const uint32_t* __restrict__ list, int stride, pct_blk::blk_pos pos,
                  const float* __restrict__ wepl, int64_t h0, int64_t count, const float* __restrict__ x,
                  double lambda, double* __restrict__ out, int64_t B) {
    __builtin_amdgcn_s_setprio(3);
    const int64_t q = (int64_t)blockIdx.x * DELTA_WAVES + (threadIdx.x >> 6);
    if (q >= count) return;
    pct_blk::delta_variant(V, list, stride, pos, wepl, h0, q, x, lambda, out, B, threadIdx.x & 63);
}

template <int V>
__global__ void __launch_bounds__(APPLY_THREADS)
k_block_apply_var(const uint64_t* __restrict__ segs, int64_t s0, int64_t ns, const uint16_t* __restrict__ csc,
                  const double* __restrict__ out, int64_t B, double lambda, double eta, int psi_sign,
                  float* __restrict__ x) {
    __builtin_amdgcn_s_setprio(3);
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    pct_blk::apply_variant<V>(segs[s0 + k], csc, out, B, lambda, eta, psi_sign, x);
}

/* one launch per block b: the voxel updates of block b-1 (workgroups [0, na),
 * one thread per segment) and the deltas of block b (the rest, one wavefront
 * per history); bodies in pct_plan_kernels.h */
#define FUSED_THREADS 256
#define FUSED_DELTA_WAVES (FUSED_THREADS / 64)
template <bool F32>
__global__ void __launch_bounds__(FUSED_THREADS)
k_block_fused(const uint32_t* __restrict__ list, const uint64_t* __restrict__ prev, int stride, pct_blk::blk_pos pos,
              const float* __restrict__ wepl, int64_t h0, int64_t count,
              double lambda, void* __restrict__ delta_out, const uint64_t* __restrict__ segs_p, int64_t ns_p,
              const uint16_t* __restrict__ csc_p, const void* __restrict__ delta_p, uint32_t stamp_p, unsigned na,
              uint64_t* __restrict__ xw) {
    __builtin_amdgcn_s_setprio(3);
    const pct_blk::fused_prev<F32> pv{csc_p, delta_p, stamp_p};
    if (blockIdx.x < na) {
        const int64_t k = (int64_t)blockIdx.x * FUSED_THREADS + threadIdx.x;
        if (k < ns_p) pct_blk::fused_apply<F32>(segs_p[k], xw, pv);
        return;
    }
    const int64_t q = (int64_t)(blockIdx.x - na) * FUSED_DELTA_WAVES + (threadIdx.x >> 6);
    if (q >= count) return;
    pct_blk::fused_delta<F32>(list, prev, stride, pos, wepl, h0, q, xw, pv, lambda, delta_out, threadIdx.x & 63);
}

/* one launch per block b, waiting form: the voxel updates of block b-1
 * (workgroups [0, na), na a multiple of 8, one thread per segment) and the
 * deltas of block b (workgroup na + j on XCD j % 8, DELTA_WAVES histories
 * each, the placement the plan's lateral bands assume); bodies in
 * pct_plan_kernels.h */
#define WAIT_THREADS (64 * DELTA_WAVES)
static_assert(WAIT_FLAG == PCT_STATUS_WAIT, "the waiting form's give-up flag is the status word's");
template <bool F32>
__global__ void __launch_bounds__(WAIT_THREADS)
k_block_wait(const uint32_t* __restrict__ list, int stride, pct_blk::blk_pos pos, const float* __restrict__ wepl,
             int64_t h0, int64_t count, double lambda, void* __restrict__ delta_out,
             const uint64_t* __restrict__ segs_p, int64_t ns_p, const uint16_t* __restrict__ csc_p,
             const void* __restrict__ delta_p, uint32_t stamp_p, uint32_t stamp_wait, int polls, unsigned na,
             uint64_t* __restrict__ xw, int* __restrict__ status) {
    if (blockIdx.x < na) {
        __builtin_amdgcn_s_setprio(3);
        const pct_blk::fused_prev<F32> pv{csc_p, delta_p, stamp_p};
        const int64_t k = (int64_t)blockIdx.x * WAIT_THREADS + threadIdx.x;
        if (k < ns_p) pct_blk::fused_apply<F32>(segs_p[k], xw, pv);
        return;
    }
    __builtin_amdgcn_s_setprio(2);
    const int64_t q = (int64_t)(blockIdx.x - na) * DELTA_WAVES + (threadIdx.x >> 6);
    if (q >= count) return;
    pct_blk::wait_delta<F32>(list, stride, pos, wepl, h0, q, xw, stamp_wait, polls, status, lambda, delta_out,
                             threadIdx.x & 63);
}

/* x (row order) into the tiled image and back: grid (ceil(C / 256), R, S) */
__global__ void k_tile_in(const float* __restrict__ x, pct_blk::vox_tiling t, float* __restrict__ xt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= t.C) return;
    const int r = blockIdx.y, z = blockIdx.z;
    xt[pct_blk::tiled_xyz(t, c, r, z)] = x[((int64_t)z * t.R + r) * t.C + c];
}

__global__ void k_tile_out(const float* __restrict__ xt, pct_blk::vox_tiling t, float* __restrict__ x) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= t.C) return;
    const int r = blockIdx.y, z = blockIdx.z;
    x[((int64_t)z * t.R + r) * t.C + c] = xt[pct_blk::tiled_xyz(t, c, r, z)];
}

/* the image as {value, stamp} words (stamp: no block yet) and back */
__global__ void k_xw_pack(const float* __restrict__ x, int64_t V, uint64_t* __restrict__ xw) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) xw[i] = pct_blk::xw_pack(x[i], 0xFFFFFFFFu);
}

/* x is left as it was when a wait of the waiting form gave up */
__global__ void k_xw_unpack(const uint64_t* __restrict__ xw, int64_t V, float* __restrict__ x,
                            const int* __restrict__ status) {
    if (*status & PCT_STATUS_WAIT) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) x[i] = pct_blk::xw_value(xw[i]);
}

}  // namespace

/* ------------------------------------------------------------------------- */
/* host                                                                       */
/* ------------------------------------------------------------------------- */
/* chunks of the block loop (one graph and one MLP launch each). By default
 * the block loop reads the plan's lists and never waits for the check of the
 * recomputed ones, which runs beside it as one launch (one chunk). With the
 * per-chunk waits of overlap 1 the first chunk is small so that the block
 * loop starts early, later ones double up to CHUNK_MAX blocks so that each
 * MLP launch fills the GPU for many rounds. */
#define CHUNK_FIRST 16
#define CHUNK_MAX 256

static void plan_release_graph(pct_plan* pl) {
    for (hipGraphExec_t& g : pl->graphs)
        if (g) (void)hipGraphExecDestroy(g);
    pl->graphs.clear();
}

static void release(pct_buffer& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

void pct_plan_destroy(pct_plan* pl) {
    if (!pl) return;
    plan_release_graph(pl);
    pct_buffer* bufs[] = {&pl->list, &pl->prev, &pl->hist_off, &pl->len, &pl->chord, &pl->hash, &pl->lane_of, &pl->vorder, &pl->vwork, &pl->segs,
                          &pl->csc, &pl->long_rows, &pl->blk_seg, &pl->blk_long, &pl->blk_entry, &pl->delta,
                          &pl->xw, &pl->xt, &pl->pos_k, &pl->pos_len, &pl->pos_chord, &pl->pos_r, &pl->rank};
    for (pct_buffer* b : bufs) release(*b);
    for (hipEvent_t e : pl->ev_chunk)
        if (e) (void)hipEventDestroy(e);
    if (pl->ev_start) (void)hipEventDestroy(pl->ev_start);
    delete pl;
}

static int vbits_for(int64_t V) {
    int b = 1;
    while (((int64_t)1 << b) < V) ++b;
    return b;
}

/* Build (or rebuild) the plan of `hist` with blocks of B histories. */
static int64_t plan_budget(pct_ctx* ctx);
static void plan_free_buffers(pct_plan* pl);

/* *capacity is set when the plan cannot be built for this scan on this
 * device (grid or block beyond the records' index fields, or not enough
 * memory); the context then holds no plan buffers */
static int plan_build(pct_ctx* ctx, const pct_hist_soa* hist, int B, int mlp_mode, bool* capacity = nullptr) {
    const int64_t n = hist->n;
    bool cap_dummy = false;
    bool& cap_fail = capacity ? *capacity : cap_dummy;
    cap_fail = false;
    if (n <= 0) return pct_fail(ctx, PCT_E_INVALID, "DROP plan: no histories");
    const pct_blk::vox_tiling tl = pct_blk::make_tiling(ctx->g.C, ctx->g.R, ctx->g.S);
    const int64_t Vt = tl.Vt;
    if (B <= 0 || B > SEG_MAX_B || Vt > SEG_MAX_V) {
        cap_fail = true;
        if (ctx->plan) plan_free_buffers(ctx->plan);
        return pct_fail(ctx, PCT_E_INVALID,
                        "no DROP plan for %lld voxels (%lld in whole tiles, at most %lld) in blocks of %d (at most %lld)",
                        (long long)ctx->g.V, (long long)Vt, (long long)SEG_MAX_V, B, (long long)SEG_MAX_B);
    }
    const int vbits = vbits_for(Vt);
    const bool with_prev = ctx->tune_block_form == 1;
    const bool with_tags = ctx->tune_block_form == 3;
    if (!ctx->plan) ctx->plan = new pct_plan();
    pct_plan* pl = ctx->plan;
    plan_release_graph(pl);
    pl->valid = false;
    ++pl->generation;
    int rc;
    const int64_t nb = (n + B - 1) / B;
    /* 1. the MLP lists into a capacity-strided build store */
    if ((rc = pct_ensure(ctx, ctx->offsets, sizeof(int64_t) * (n + 1)))) return rc;
    int64_t total = 0;
    if ((rc = pct_mlp_capacity_impl(ctx, hist, 0, n, (int64_t*)ctx->offsets.ptr, &total))) return rc;
    {
        const int64_t budget = plan_budget(ctx);
        int64_t need = 0;
        if (!pct_drop_plan_choice(n, total, 0, 0, Vt, B, budget, with_prev, &need)) {
            cap_fail = true;
            plan_free_buffers(pl);
            return pct_fail(ctx, PCT_E_NOMEM, "the DROP plan's build needs %.1f GB, %.1f GB available", need / 1e9,
                            budget / 1e9);
        }
    }
    if ((rc = pct_ensure(ctx, ctx->path_store, sizeof(uint32_t) * (total > 0 ? total : 1)))) return rc;
    if ((rc = pct_ensure(ctx, pl->len, sizeof(uint32_t) * n))) return rc;
    if ((rc = pct_ensure(ctx, pl->chord, sizeof(double) * n))) return rc;
    if ((rc = pct_ensure(ctx, pl->hash, sizeof(uint32_t) * n))) return rc;
    const int64_t* cap_off = (const int64_t*)ctx->offsets.ptr;
    uint32_t* len = (uint32_t*)pl->len.ptr;
    if ((rc = pct_mlp_paths_impl(ctx, hist, 0, n, cap_off, (uint32_t*)ctx->path_store.ptr, len, (double*)pl->chord.ptr,
                                 (uint32_t*)pl->hash.ptr, nullptr, mlp_mode)))
        return rc;
    int flags = 0;
    if ((rc = pct_take_status(ctx, &flags))) return rc;
    if ((rc = pct_status_to_rc(ctx, flags))) return rc;
    /* 2. block strides = longest list per block; slot offsets of the
     * block-strided list store */
    if ((rc = pct_ensure(ctx, ctx->steps, sizeof(int32_t) * (n > nb ? n : nb)))) return rc;
    int32_t* bmax = (int32_t*)ctx->steps.ptr;
    hipLaunchKernelGGL(k_block_max_len, dim3((unsigned)nb), dim3(256), 0, ctx->stream, (const uint32_t*)len, n,
                       (int64_t)B, bmax);
    pl->stride_h.assign(nb, 0);
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(pl->stride_h.data(), bmax, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    pl->list_base_h.assign(nb + 1, 0);
    for (int64_t b = 0; b < nb; ++b) {
        int64_t cnt = (n - b * B) < B ? (n - b * B) : B;
        pl->list_base_h[b + 1] = pl->list_base_h[b] + cnt * pl->stride_h[b];
    }
    const int64_t list_size = pl->list_base_h[nb];
    if ((rc = pct_ensure(ctx, pl->hist_off, sizeof(int64_t) * (n + 1)))) return rc;
    {
        pct_buffer tmp_base, tmp_stride;
        auto free_tmp = [&]() { release(tmp_base); release(tmp_stride); };
        if ((rc = pct_ensure(ctx, tmp_base, sizeof(int64_t) * (nb + 1))) ||
            (rc = pct_ensure(ctx, tmp_stride, sizeof(int32_t) * nb))) {
            free_tmp();
            return rc;
        }
        PCT_HIP_CHECK(ctx, hipMemcpyAsync(tmp_base.ptr, pl->list_base_h.data(), sizeof(int64_t) * (nb + 1),
                                          hipMemcpyHostToDevice, ctx->stream));
        PCT_HIP_CHECK(ctx, hipMemcpyAsync(tmp_stride.ptr, pl->stride_h.data(), sizeof(int32_t) * nb,
                                          hipMemcpyHostToDevice, ctx->stream));
        /* placement of each block's rows: lateral bands per XCD */
        if ((rc = pct_ensure(ctx, pl->pos_k, sizeof(uint16_t) * (size_t)n)) ||
            (rc = pct_ensure(ctx, pl->pos_len, sizeof(uint32_t) * (size_t)n)) ||
            (rc = pct_ensure(ctx, pl->pos_chord, sizeof(double) * (size_t)n)) ||
            (rc = pct_ensure(ctx, pl->pos_r, sizeof(uint16_t) * (size_t)n)) ||
            (rc = pct_ensure(ctx, pl->rank, sizeof(uint16_t) * (size_t)n))) {
            free_tmp();
            return rc;
        }
        pct_buffer kin, kout, vin, vout, stmp;
        auto free_sort = [&]() { for (pct_buffer* b : {&kin, &kout, &vin, &vout, &stmp}) release(*b); };
        size_t sort_bytes = 0;
        PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                                     (uint32_t*)nullptr, (uint32_t*)nullptr, (size_t)n, 0, 64,
                                                     ctx->stream));
        if ((rc = pct_ensure(ctx, kin, 8 * (size_t)n)) || (rc = pct_ensure(ctx, kout, 8 * (size_t)n)) ||
            (rc = pct_ensure(ctx, vin, 4 * (size_t)n)) || (rc = pct_ensure(ctx, vout, 4 * (size_t)n)) ||
            (rc = pct_ensure(ctx, stmp, sort_bytes))) {
            free_sort();
            free_tmp();
            return rc;
        }
        hipLaunchKernelGGL(k_lateral_keys, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, *hist, n, (int64_t)B,
                           (uint64_t*)kin.ptr, (uint32_t*)vin.ptr);
        size_t sb = stmp.bytes;
        int endbit = 32 + vbits_for(nb + 1);
        if (endbit > 64) endbit = 64;
        PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(stmp.ptr, sb, (uint64_t*)kin.ptr, (uint64_t*)kout.ptr,
                                                     (uint32_t*)vin.ptr, (uint32_t*)vout.ptr, (size_t)n, 0, endbit,
                                                     ctx->stream));
        hipLaunchKernelGGL(k_positions, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, (const uint32_t*)vout.ptr,
                           n, (int64_t)B, nb, (const int64_t*)tmp_base.ptr, (const int32_t*)tmp_stride.ptr,
                           (const uint32_t*)len, (const double*)pl->chord.ptr, (uint16_t*)pl->pos_k.ptr,
                           (uint32_t*)pl->pos_len.ptr, (double*)pl->pos_chord.ptr, (uint16_t*)pl->pos_r.ptr,
                           (uint16_t*)pl->rank.ptr, (int64_t*)pl->hist_off.ptr);
        PCT_HIP_CHECK(ctx, hipGetLastError());
        PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        free_sort();
        free_tmp();
    }
    const int64_t* hist_off = (const int64_t*)pl->hist_off.ptr;
    /* the build store holds every list at its capacity offset */
    const uint32_t* list = (const uint32_t*)ctx->path_store.ptr;
    /* 3. exact entry offsets (prefix sum of the lengths) */
    pct_buffer ex_off;
    if ((rc = pct_ensure(ctx, ex_off, sizeof(int64_t) * (n + 1)))) return rc;
    hipLaunchKernelGGL(k_lengths_i32, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, (const uint32_t*)len, n,
                       (int32_t*)ctx->steps.ptr);
    int64_t E = 0;
    if ((rc = pct_exclusive_scan_i32(ctx, (const int32_t*)ctx->steps.ptr, n, (int64_t*)ex_off.ptr, &E))) {
        release(ex_off);
        return rc;
    }
    std::vector<int64_t> ex_h(n + 1);
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(ex_h.data(), ex_off.ptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost,
                                      ctx->stream));
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    {
        const int64_t budget = plan_budget(ctx);
        int64_t need = 0;
        if (!pct_drop_plan_choice(n, total, list_size, E, Vt, B, budget, with_prev, &need)) {
            release(ex_off);
            release(ctx->path_store);
            cap_fail = true;
            plan_free_buffers(pl);
            return pct_fail(ctx, PCT_E_NOMEM, "the DROP plan needs %.1f GB, %.1f GB available", need / 1e9,
                            budget / 1e9);
        }
    }
    std::vector<int64_t> blk_entry_h(nb + 1);
    for (int64_t b = 0; b <= nb; ++b) blk_entry_h[b] = ex_h[b * B < n ? b * B : n];
    if ((rc = pct_ensure(ctx, pl->blk_entry, sizeof(int64_t) * (nb + 1))) ||
        (rc = pct_ensure(ctx, pl->blk_seg, sizeof(int64_t) * (nb + 1)))) {
        release(ex_off);
        return rc;
    }
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(pl->blk_entry.ptr, blk_entry_h.data(), sizeof(int64_t) * (nb + 1),
                                      hipMemcpyHostToDevice, ctx->stream));
    if ((rc = pct_ensure(ctx, pl->csc, sizeof(uint16_t) * (E > 0 ? E : 1)))) {
        release(ex_off);
        return rc;
    }
    /* 4. chunks of blocks: key = (local block << vbits) | voxel must fit 32
     * bits, and a chunk holds at most ~2^27 entries */
    int64_t max_blocks = (int64_t)1 << (32 - vbits);
    int64_t avg = nb > 0 ? (E + nb - 1) / nb : 1;
    int64_t by_entries = ((int64_t)1 << 27) / (avg > 0 ? avg : 1);
    int64_t chunk = max_blocks < by_entries ? max_blocks : by_entries;
    if (chunk < 1) chunk = 1;
    int64_t max_chunk_entries = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += chunk) {
        int64_t b1 = b0 + chunk < nb ? b0 + chunk : nb;
        max_chunk_entries = std::max(max_chunk_entries, blk_entry_h[b1] - blk_entry_h[b0]);
    }
    size_t ce = (size_t)(max_chunk_entries > 0 ? max_chunk_entries : 1);
    size_t sort_tmp = 0;
    {
        uint32_t* k0 = nullptr;
        uint32_t* v0 = nullptr;
        PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(nullptr, sort_tmp, k0, k0, v0, v0, ce, 0, 32, ctx->stream));
    }
    pct_buffer kin, kout, vin, vout, fl, sidx, lfl, lidx, tmp;
    auto free_all = [&]() {
        for (pct_buffer* b : {&kin, &kout, &vin, &vout, &fl, &sidx, &lfl, &lidx, &tmp, &ex_off}) release(*b);
    };
    if ((rc = pct_ensure(ctx, kin, 4 * ce)) || (rc = pct_ensure(ctx, kout, 4 * ce)) || (rc = pct_ensure(ctx, vin, 4 * ce)) ||
        (rc = pct_ensure(ctx, vout, 4 * ce)) || (rc = pct_ensure(ctx, fl, 4 * ce)) ||
        (rc = pct_ensure(ctx, sidx, 8 * (ce + 1))) || (rc = pct_ensure(ctx, lfl, 4 * ce)) ||
        (rc = pct_ensure(ctx, lidx, 8 * (ce + 1))) || (rc = pct_ensure(ctx, tmp, sort_tmp))) {
        free_all();
        return rc;
    }
    if ((rc = pct_ensure(ctx, pl->blk_long, sizeof(int64_t) * (nb + 1)))) {
        free_all();
        return rc;
    }
    /* first pass counts the segments and the long segments per chunk, the
     * second writes them */
    std::vector<int64_t> chunk_segs, chunk_long;
    int64_t S_total = 0, L_total = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int64_t seg0 = 0, long0 = 0;
        for (int64_t b0 = 0; b0 < nb; b0 += chunk) {
            int64_t b1 = b0 + chunk < nb ? b0 + chunk : nb;
            int64_t h0 = b0 * B, h1 = b1 * B < n ? b1 * B : n;
            int64_t entry0 = blk_entry_h[b0];
            int64_t Ec = blk_entry_h[b1] - entry0;
            if (Ec == 0) {
                if (pass == 0) { chunk_segs.push_back(0); chunk_long.push_back(0); }
                continue;
            }
            int endbit = vbits + vbits_for(b1 - b0 + 1);
            if (endbit > 32) endbit = 32;
            hipLaunchKernelGGL(k_emit_keys, dim3((unsigned)((h1 - h0 + 3) / 4)), dim3(256), 0, ctx->stream, list,
                               cap_off, (const uint32_t*)len, (const int64_t*)ex_off.ptr, h0, h1, (int64_t)B, b0, vbits,
                               entry0, (const uint16_t*)pl->rank.ptr, (uint32_t*)kin.ptr, (uint32_t*)vin.ptr, tl);
            size_t tb = tmp.bytes;
            PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(tmp.ptr, tb, (uint32_t*)kin.ptr, (uint32_t*)kout.ptr,
                                                         (uint32_t*)vin.ptr, (uint32_t*)vout.ptr, (size_t)Ec, 0, endbit,
                                                         ctx->stream));
            hipLaunchKernelGGL(k_head_flags, dim3(pct_grid(Ec, 256)), dim3(256), 0, ctx->stream,
                               (const uint32_t*)kout.ptr, Ec, (int32_t*)fl.ptr);
            int64_t Sc = 0, Lc = 0;
            if ((rc = pct_exclusive_scan_i32(ctx, (const int32_t*)fl.ptr, Ec, (int64_t*)sidx.ptr, &Sc))) {
                free_all();
                return rc;
            }
            hipLaunchKernelGGL(k_long_flags, dim3(pct_grid(Ec, 256)), dim3(256), 0, ctx->stream,
                               (const int32_t*)fl.ptr, Ec, (int32_t*)lfl.ptr);
            if ((rc = pct_exclusive_scan_i32(ctx, (const int32_t*)lfl.ptr, Ec, (int64_t*)lidx.ptr, &Lc))) {
                free_all();
                return rc;
            }
            if (pass == 0) {
                chunk_segs.push_back(Sc);
                chunk_long.push_back(Lc);
                continue;
            }
            hipLaunchKernelGGL(k_write_segments, dim3(pct_grid(Ec, 256)), dim3(256), 0, ctx->stream,
                               (const uint32_t*)kout.ptr, (const uint32_t*)vout.ptr, (const int32_t*)fl.ptr,
                               (const int64_t*)sidx.ptr, (const int64_t*)lidx.ptr, Lc, Ec, vbits, seg0, long0, b0,
                               entry0, (const int64_t*)pl->blk_entry.ptr, (uint64_t*)pl->segs.ptr,
                               (uint16_t*)pl->csc.ptr, (pct_blk::long_row*)pl->long_rows.ptr,
                               (int64_t*)pl->blk_seg.ptr, (int64_t*)pl->blk_long.ptr);
            PCT_HIP_CHECK(ctx, hipGetLastError());
            seg0 += Sc;
            long0 += Lc;
        }
        if (pass == 0) {
            S_total = 0;
            L_total = 0;
            for (int64_t s : chunk_segs) S_total += s;
            for (int64_t l : chunk_long) L_total += l;
            if ((rc = pct_ensure(ctx, pl->segs, sizeof(uint64_t) * (S_total > 0 ? S_total : 1))) ||
                (rc = pct_ensure(ctx, pl->long_rows, sizeof(pct_blk::long_row) * (L_total > 0 ? L_total : 1)))) {
                free_all();
                return rc;
            }
        }
    }
    free_all();
    release(pl->rank);
    /* sentinels */
    PCT_HIP_CHECK(ctx, hipMemcpyAsync((int64_t*)pl->blk_seg.ptr + nb, &S_total, sizeof(int64_t), hipMemcpyHostToDevice,
                                      ctx->stream));
    PCT_HIP_CHECK(ctx, hipMemcpyAsync((int64_t*)pl->blk_long.ptr + nb, &L_total, sizeof(int64_t),
                                      hipMemcpyHostToDevice, ctx->stream));
    pl->blk_entry_h = blk_entry_h;
    pl->blk_seg_h.assign(nb + 1, 0);
    pl->blk_long_h.assign(nb + 1, 0);
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(pl->blk_seg_h.data(), pl->blk_seg.ptr, sizeof(int64_t) * (nb + 1),
                                      hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(pl->blk_long_h.data(), pl->blk_long.ptr, sizeof(int64_t) * (nb + 1),
                                      hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    /* 5. the list store (unused slots the sentinel) and, for the
     * one-launch-per-block form, slot for slot the previous block's segment
     * record of each voxel. Maps from voxel to segment for M blocks at a time
     * (tagged with the block, so never cleared); the build store is freed
     * afterwards. */
    if ((rc = pct_ensure(ctx, pl->list, sizeof(uint32_t) * (size_t)(list_size > 0 ? list_size : 1)))) return rc;
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(pct_grid(list_size, 256), PCT_GRID_CAP)), dim3(256), 0, ctx->stream,
                       (uint32_t*)pl->list.ptr, list_size, LIST_SENTINEL);
    if (with_prev) {
        if ((rc = pct_ensure(ctx, pl->prev, sizeof(uint64_t) * (size_t)(list_size > 0 ? list_size : 1)))) return rc;
    } else {
        release(pl->prev);
    }
    {
        const int64_t V = Vt;
        const bool maps = with_prev || with_tags;
        int64_t M = maps ? std::max<int64_t>(2, std::min<int64_t>(256, ((int64_t)1 << 31) / (8 * V))) : 256;
        M = std::min<int64_t>(M, nb + 1);
        pct_buffer map;
        if (maps) {
            if ((rc = pct_ensure(ctx, map, sizeof(uint64_t) * (size_t)(M * V)))) return rc;
            hipLaunchKernelGGL(k_fill_u64, dim3(std::min(pct_grid(M * V, 256), PCT_GRID_CAP)), dim3(256), 0,
                               ctx->stream, (uint64_t*)map.ptr, M * V, ~0ull);
        }
        /* entries of blocks [c0, c1) read the maps of blocks [c0 - 1, c1 - 1) */
        const int64_t step = M - 1;
        for (int64_t c0 = 0; c0 < nb; c0 += step) {
            const int64_t c1 = std::min(nb, c0 + step);
            const int64_t m0 = c0 > 0 ? c0 - 1 : 0, m1 = c1 - 1;
            if (maps && m1 > m0) {
                int64_t maxs = 0;
                for (int64_t b = m0; b < m1; ++b) maxs = std::max(maxs, pl->blk_seg_h[b + 1] - pl->blk_seg_h[b]);
                if (maxs > 0)
                    hipLaunchKernelGGL(k_seg_map, dim3(std::min<unsigned>(pct_grid(maxs, 256), 1024u), (unsigned)(m1 - m0)),
                                       dim3(256), 0, ctx->stream, (const uint64_t*)pl->segs.ptr,
                                       (const int64_t*)pl->blk_seg.ptr, m0, M, V, (uint64_t*)map.ptr);
            }
            const int64_t h0 = c0 * B, h1 = std::min(n, c1 * B);
            hipLaunchKernelGGL(k_build_entries, dim3((unsigned)((h1 - h0 + 3) / 4)), dim3(256), 0, ctx->stream,
                               (const uint32_t*)ctx->path_store.ptr, cap_off, (const uint32_t*)len, hist_off, h0, h1,
                               (int64_t)B, (const uint64_t*)pl->segs.ptr, (const int64_t*)pl->blk_seg.ptr, M, V,
                               (const uint64_t*)map.ptr, (uint32_t*)pl->list.ptr,
                               with_prev ? (uint64_t*)pl->prev.ptr : nullptr, with_tags ? 1 : 0, tl);
        }
        PCT_HIP_CHECK(ctx, hipGetLastError());
        PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        release(map);
    }
    pl->has_prev = with_prev;
    pl->has_tags = with_tags;
    release(ctx->path_store);
    /* deltas of two blocks (the one-launch form reads block b-1's while writing
     * block b's), two doubles per history (ROBUST) */
    if ((rc = pct_ensure(ctx, pl->delta, sizeof(double) * 4 * (size_t)B))) return rc;
    /* the verification pass's lane order */
    if (mlp_mode == PCT_MLP_FP64) {
        if ((rc = pct_ensure(ctx, pl->lane_of, sizeof(uint16_t) * (size_t)n)) ||
            (rc = pct_mlp_verify_order(ctx, hist, n, (uint16_t*)pl->lane_of.ptr)))
            return rc;
        /* the one-launch check's step order; its gather space doubles as
         * the sort's workspace here */
        if ((rc = pct_ensure(ctx, pl->vorder, sizeof(uint32_t) * (size_t)n)) ||
            (rc = pct_ensure(ctx, pl->vwork, sizeof(uint32_t) * 12 * (size_t)n)) ||
            (rc = pct_mlp_step_order(ctx, hist, n, (uint32_t*)pl->vwork.ptr, (uint32_t*)pl->vwork.ptr + n,
                                     (uint32_t*)pl->vorder.ptr)))
            return rc;
    } else {
        release(pl->lane_of);
        release(pl->vorder);
        release(pl->vwork);
    }
    pl->n = n;
    pl->B = B;
    pl->nb = nb;
    pl->Vt = Vt;
    pl->E = E;
    pl->S = S_total;
    pl->key = *hist;
    pl->mlp_mode = mlp_mode;
    pl->net_version = ctx->net_version;
    pl->chunk_b.clear();
    pl->valid = true;
    return PCT_OK;
}

static bool same_hist(const pct_hist_soa& a, const pct_hist_soa& b) {
    return a.n == b.n && a.x_entry == b.x_entry && a.wepl == b.wepl && a.first_voxel == b.first_voxel &&
           a.xy_entry_angle == b.xy_entry_angle && a.x_exit == b.x_exit;
}

/* internal entry points for the batched form (pct_batch.hip) */
/* the plan holds the lists of these histories, this block size and MLP mode
 * (and, for the surrogate, the network and calibration it was built with) */
static bool plan_matches(const pct_ctx* ctx, const pct_plan* pl, const pct_hist_soa* hist, int B, int mlp) {
    return pl && pl->valid && pl->B == B && pl->mlp_mode == mlp && same_hist(pl->key, *hist) &&
           (mlp != PCT_MLP_SURROGATE || pl->net_version == ctx->net_version) &&
           (ctx->tune_block_form != 1 || pl->has_prev) && pl->has_tags == (ctx->tune_block_form == 3);
}

int pct_plan_build(pct_ctx* ctx, const pct_hist_soa* hist, int B, int mlp_mode) {
    return plan_build(ctx, hist, B, mlp_mode);
}
bool pct_plan_ready(const pct_ctx* ctx, const pct_hist_soa* hist, int B, int mlp_mode) {
    const pct_plan* pl = ctx->plan;
    return plan_matches(ctx, pl, hist, B, mlp_mode);
}

/* the block loop over blocks [b_begin, b_end) on the plan */
/* update variant of a pct_drop_iteration mode: 0 plain, 1 PCT_DROP_ROBUST,
 * 2 PCT_DROP_ROBUST2 */
static int update_variant(int mode) {
    return (mode & PCT_DROP_ROBUST) ? 1 : (mode & PCT_DROP_ROBUST2) ? 2 : 0;
}

/* block-loop form: 1 = one launch per block on {value, stamp} image words,
 * the previous block's update recomputed where it is not stored yet;
 * 2 = two launches per block on the float image (the research variants
 * always); 3 = one launch per block on image words, waiting for the previous
 * block's stored updates */
static int block_form(const pct_ctx* ctx, int variant) {
    if (variant != 0) return 2;
    return ctx->tune_block_form == 1 ? 1 : ctx->tune_block_form == 3 ? 3 : 2;
}

static int plan_block_loop(pct_ctx* ctx, pct_plan* pl, const pct_hist_soa* hist, float* x, double lambda, bool f32,
                           int variant, int64_t b_begin, int64_t b_end) {
    const int64_t B = pl->B;
    const uint32_t* list = (const uint32_t*)pl->list.ptr;
    const uint64_t* prev = (const uint64_t*)pl->prev.ptr;
    const double* chord = (const double*)pl->chord.ptr;
    auto pos_of = [&](int64_t b) {
        return pct_blk::blk_pos{(const uint16_t*)pl->pos_k.ptr + b * B, (const uint32_t*)pl->pos_len.ptr + b * B,
                                (const double*)pl->pos_chord.ptr + b * B, (const uint16_t*)pl->pos_r.ptr + b * B};
    };
    const uint64_t* segs = (const uint64_t*)pl->segs.ptr;
    const uint16_t* csc = (const uint16_t*)pl->csc.ptr;
    const int form = block_form(ctx, variant);
    if (form == 1 || form == 3) {
        /* launch b: deltas of block b and the voxel updates of block b-1; the
         * last chunk adds launch nb (updates of the last block only) */
        uint64_t* xw = (uint64_t*)pl->xw.ptr;
        const int64_t b_last = b_end == pl->nb ? pl->nb : b_end - 1;
        for (int64_t b = b_begin; b <= b_last; ++b) {
            const int64_t h0 = b * B, cnt = b < pl->nb ? std::min<int64_t>(B, pl->n - h0) : 0;
            const uint32_t* lb = b < pl->nb ? list + pl->list_base_h[b] : nullptr;
            const uint64_t* pb = b < pl->nb ? prev + pl->list_base_h[b] : nullptr;
            const int stride = b < pl->nb ? pl->stride_h[b] : 0;
            void* dout = (char*)pl->delta.ptr + (size_t)(b & 1) * 2 * B * sizeof(double);
            const void* dprev = (const char*)pl->delta.ptr + (size_t)((b + 1) & 1) * 2 * B * sizeof(double);
            const uint64_t* sp = nullptr;
            const uint16_t* cp = nullptr;
            int64_t nsp = 0;
            if (b > 0) {
                sp = segs + pl->blk_seg_h[b - 1];
                nsp = pl->blk_seg_h[b] - pl->blk_seg_h[b - 1];
                cp = csc + pl->blk_entry_h[b - 1];
            }
            const uint32_t stamp = (uint32_t)(b - 1);
            if (form == 3) {
                const unsigned na = nsp > 0 ? (pct_grid(nsp, WAIT_THREADS) + 7u) / 8u * 8u : 0u;
                const unsigned nd = (unsigned)((cnt + DELTA_WAVES - 1) / DELTA_WAVES);
                if (na + nd == 0) continue;
                /* the diagnostics knob: a stamp no update writes, two polls */
                const uint32_t stamp_wait = ctx->tune_wait_give_up ? stamp + 0x40000000u : stamp;
                const int polls = ctx->tune_wait_give_up ? 2 : WAIT_POLLS;
                if (f32)
                    hipLaunchKernelGGL(k_block_wait<true>, dim3(na + nd), dim3(WAIT_THREADS), 0, ctx->stream, lb,
                                       stride, pos_of(b), (const float*)hist->wepl, h0, cnt, lambda, dout, sp, nsp,
                                       cp, dprev, stamp, stamp_wait, polls, na, xw, ctx->d_status);
                else
                    hipLaunchKernelGGL(k_block_wait<false>, dim3(na + nd), dim3(WAIT_THREADS), 0, ctx->stream, lb,
                                       stride, pos_of(b), (const float*)hist->wepl, h0, cnt, lambda, dout, sp, nsp,
                                       cp, dprev, stamp, stamp_wait, polls, na, xw, ctx->d_status);
                continue;
            }
            const unsigned na = pct_grid(nsp, FUSED_THREADS) * (nsp > 0);
            const unsigned nd = (unsigned)((cnt + FUSED_DELTA_WAVES - 1) / FUSED_DELTA_WAVES);
            if (na + nd == 0) continue;
            if (f32)
                hipLaunchKernelGGL(k_block_fused<true>, dim3(na + nd), dim3(FUSED_THREADS), 0, ctx->stream, lb, pb, stride,
                                   pos_of(b), (const float*)hist->wepl, h0, cnt, lambda, dout, sp, nsp, cp, dprev,
                                   stamp, na, xw);
            else
                hipLaunchKernelGGL(k_block_fused<false>, dim3(na + nd), dim3(FUSED_THREADS), 0, ctx->stream, lb, pb, stride,
                                   pos_of(b), (const float*)hist->wepl, h0, cnt, lambda, dout, sp, nsp, cp, dprev,
                                   stamp, na, xw);
        }
        PCT_HIP_CHECK(ctx, hipGetLastError());
        return PCT_OK;
    }
    void* delta = pl->delta.ptr;
    for (int64_t b = b_begin; b < b_end; ++b) {
        int64_t h0 = b * B, cnt = (pl->n - h0) < B ? (pl->n - h0) : B;
        const uint32_t* lb = list + pl->list_base_h[b];
        int stride = pl->stride_h[b];
        dim3 dgrid((unsigned)((cnt + DELTA_WAVES - 1) / DELTA_WAVES));
        if (variant != 0) {
            const double eta = ctx->p.eta;
            const int psi = ctx->p.psi_sign;
            double* out = (double*)delta;
            if (variant == 1)
                hipLaunchKernelGGL(k_block_delta_var<1>, dgrid, dim3(64 * DELTA_WAVES), 0, ctx->stream, lb, stride,
                                   pos_of(b), (const float*)hist->wepl, h0, cnt, (const float*)x, lambda, out, B);
            else
                hipLaunchKernelGGL(k_block_delta_var<2>, dgrid, dim3(64 * DELTA_WAVES), 0, ctx->stream, lb, stride,
                                   pos_of(b), (const float*)hist->wepl, h0, cnt, (const float*)x, lambda, out, B);
            const int64_t s0 = pl->blk_seg_h[b], ns = pl->blk_seg_h[b + 1] - s0;
            const uint16_t* cb = csc + pl->blk_entry_h[b];
            if (ns > 0 && variant == 1)
                hipLaunchKernelGGL(k_block_apply_var<1>, dim3(pct_grid(ns, APPLY_THREADS)), dim3(APPLY_THREADS), 0, ctx->stream, segs, s0, ns,
                                   cb, (const double*)out, B, lambda, eta, psi, x);
            else if (ns > 0)
                hipLaunchKernelGGL(k_block_apply_var<2>, dim3(pct_grid(ns, APPLY_THREADS)), dim3(APPLY_THREADS), 0, ctx->stream, segs, s0, ns,
                                   cb, (const double*)out, B, lambda, eta, psi, x);
            continue;
        }
        const pct_blk::blk_pos pb = pos_of(b);
        const bool short_rows = stride <= 4 * 64;
        if (f32)
            hipLaunchKernelGGL((short_rows ? k_block_delta_f32<4> : k_block_delta_f32<DELTA_SLOTS>), dgrid,
                               dim3(64 * DELTA_WAVES), 0, ctx->stream, lb, pb.k, pb.len, pb.chord, pb.r, stride, cnt,
                               (const float*)x, (const float*)hist->wepl, h0, (float)lambda, (float*)delta);
        else
            hipLaunchKernelGGL((short_rows ? k_block_delta<4> : k_block_delta<DELTA_SLOTS>), dgrid,
                               dim3(64 * DELTA_WAVES), 0, ctx->stream, lb, pb.k, pb.len, pb.chord, pb.r, stride, cnt,
                               (const float*)x, (const float*)hist->wepl, h0, lambda, (double*)delta);
        int64_t s0 = pl->blk_seg_h[b], ns = pl->blk_seg_h[b + 1] - s0;
        const int64_t l0 = pl->blk_long_h[b], nlong = pl->blk_long_h[b + 1] - l0;
        const pct_blk::long_row* rows = (const pct_blk::long_row*)pl->long_rows.ptr + l0;
        const uint16_t* cb = csc + pl->blk_entry_h[b];
        if (ns > 0 && f32)
            hipLaunchKernelGGL(k_block_apply_f32, dim3(pct_grid(ns, APPLY_THREADS)), dim3(APPLY_THREADS), 0,
                               ctx->stream, segs + s0, rows, nlong, ns, cb, (const float*)delta, x);
        else if (ns > 0)
            hipLaunchKernelGGL(k_block_apply, dim3(pct_grid(ns, APPLY_THREADS)), dim3(APPLY_THREADS), 0, ctx->stream,
                               segs + s0, rows, nlong, ns, cb, (const double*)delta, x);
    }
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

/* chunk boundaries from the current knobs; a change drops the graphs */
static void plan_chunks(pct_ctx* ctx, pct_plan* pl) {
    const bool one = ctx->tune_overlap != 1 && ctx->tune_overlap != 2;
    const int64_t all = std::max<int64_t>(1, pl->nb);
    int64_t first = ctx->tune_chunk_first > 0 ? ctx->tune_chunk_first : (one ? all : CHUNK_FIRST);
    int64_t most = ctx->tune_chunk_max > 0 ? ctx->tune_chunk_max : (one ? all : CHUNK_MAX);
    if (most < first) most = first;
    if (!pl->chunk_b.empty() && pl->chunk_first == first && pl->chunk_max == most) return;
    plan_release_graph(pl);
    pl->chunk_first = first;
    pl->chunk_max = most;
    pl->chunk_b.assign(1, 0);
    for (int64_t size = first; pl->chunk_b.back() < pl->nb; size = std::min<int64_t>(2 * size, most))
        pl->chunk_b.push_back(std::min(pl->nb, pl->chunk_b.back() + size));
}

static int plan_ensure_graphs(pct_ctx* ctx, pct_plan* pl, const pct_hist_soa* hist, float* x, double lambda,
                              bool f32, int variant) {
    const int64_t nchunks = (int64_t)pl->chunk_b.size() - 1;
    if ((int64_t)pl->graphs.size() == nchunks && pl->graph_x == x && pl->graph_lambda == lambda &&
        pl->graph_f32 == f32 && pl->graph_wepl == hist->wepl && pl->graph_variant == variant &&
        pl->graph_form == block_form(ctx, variant) && pl->graph_wait_give_up == ctx->tune_wait_give_up &&
        (variant == 0 || (pl->graph_eta == ctx->p.eta && pl->graph_psi == ctx->p.psi_sign)))
        return PCT_OK;
    plan_release_graph(pl);
    for (int64_t c = 0; c < nchunks; ++c) {
        int64_t b0 = pl->chunk_b[c], b1 = pl->chunk_b[c + 1];
        hipGraph_t g = nullptr;
        PCT_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = plan_block_loop(ctx, pl, hist, x, lambda, f32, variant, b0, b1);
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); plan_release_graph(pl); return rc; }
        if (e != hipSuccess) { plan_release_graph(pl); return pct_fail(ctx, PCT_E_HIP, "graph capture: %s", hipGetErrorString(e)); }
        hipGraphExec_t ge = nullptr;
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) { plan_release_graph(pl); return pct_fail(ctx, PCT_E_HIP, "graph instantiate: %s", hipGetErrorString(e)); }
        pl->graphs.push_back(ge);
    }
    pl->graph_x = x;
    pl->graph_lambda = lambda;
    pl->graph_f32 = f32;
    pl->graph_wepl = hist->wepl;
    pl->graph_variant = variant;
    pl->graph_form = block_form(ctx, variant);
    pl->graph_wait_give_up = ctx->tune_wait_give_up;
    pl->graph_eta = ctx->p.eta;
    pl->graph_psi = ctx->p.psi_sign;
    return PCT_OK;
}

static int plan_ensure_events(pct_ctx* ctx, pct_plan* pl, int64_t nchunks) {
    if (!ctx->aux_stream)
        PCT_HIP_CHECK(ctx, hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
    if (!pl->ev_start) PCT_HIP_CHECK(ctx, hipEventCreateWithFlags(&pl->ev_start, hipEventDisableTiming));
    while ((int64_t)pl->ev_chunk.size() < nchunks) {
        hipEvent_t e = nullptr;
        PCT_HIP_CHECK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        pl->ev_chunk.push_back(e);
    }
    return PCT_OK;
}

/* the tiled image (Vt floats; zero on allocation, so the padding voxels,
 * which no list or record names, hold zero) */
static int plan_ensure_xt(pct_ctx* ctx, pct_plan* pl, hipStream_t s) {
    if (pl->xt.ptr && pl->xt.bytes >= sizeof(float) * (size_t)pl->Vt) return PCT_OK;
    int rc = pct_ensure(ctx, pl->xt, sizeof(float) * (size_t)pl->Vt);
    if (rc) return rc;
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(pct_grid(pl->Vt, 256), PCT_GRID_CAP)), dim3(256), 0, s,
                       (uint32_t*)pl->xt.ptr, pl->Vt, 0u);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

static dim3 tile_grid(const pct_blk::vox_tiling& t) {
    return dim3((unsigned)((t.C + 255) / 256), (unsigned)t.R, (unsigned)t.S);
}

int pct_plan_image_in(pct_ctx* ctx, hipStream_t s, const float* x, float** xt) {
    pct_plan* pl = ctx->plan;
    int rc = plan_ensure_xt(ctx, pl, s);
    if (rc) return rc;
    const pct_blk::vox_tiling tl = pct_blk::make_tiling(ctx->g.C, ctx->g.R, ctx->g.S);
    hipLaunchKernelGGL(k_tile_in, tile_grid(tl), dim3(256), 0, s, x, tl, (float*)pl->xt.ptr);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    *xt = (float*)pl->xt.ptr;
    return PCT_OK;
}

int pct_plan_image_out(pct_ctx* ctx, hipStream_t s, float* x) {
    const pct_blk::vox_tiling tl = pct_blk::make_tiling(ctx->g.C, ctx->g.R, ctx->g.S);
    hipLaunchKernelGGL(k_tile_out, tile_grid(tl), dim3(256), 0, s, (const float*)ctx->plan->xt.ptr, tl, x);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

/* One iteration on the plan. With `recompute`, the MLP lists of every chunk
 * of blocks are recomputed on the context's second stream, ahead of the block
 * loop, which takes chunk c as soon as its lists are done: the FP64-bound MLP
 * of later chunks runs beside the latency-bound block loop of earlier ones. */
static int plan_iterate(pct_ctx* ctx, pct_plan* pl, const pct_hist_soa* hist, float* x, double lambda, bool f32,
                        int variant, int mlp, bool recompute, bool eager) {
    const bool sequential = recompute && ctx->tune_overlap == 2;
    const bool words = block_form(ctx, variant) != 2;
    int rc;
    plan_chunks(ctx, pl);
    const int64_t nchunks = (int64_t)pl->chunk_b.size() - 1;
    const int64_t Vt = pl->Vt;
    if ((rc = plan_ensure_xt(ctx, pl, ctx->stream))) return rc;
    float* xt = (float*)pl->xt.ptr;
    if (words && (rc = pct_ensure(ctx, pl->xw, sizeof(uint64_t) * (size_t)Vt))) return rc;
    if (!eager && (rc = plan_ensure_graphs(ctx, pl, hist, xt, lambda, f32, variant))) return rc;
    const pct_blk::vox_tiling tl = pct_blk::make_tiling(ctx->g.C, ctx->g.R, ctx->g.S);
    const dim3 tgrid = tile_grid(tl);
    hipLaunchKernelGGL(k_tile_in, tgrid, dim3(256), 0, ctx->stream, (const float*)x, tl, xt);
    if (words)
        hipLaunchKernelGGL(k_xw_pack, dim3(pct_grid(Vt, 256)), dim3(256), 0, ctx->stream, (const float*)xt, Vt,
                           (uint64_t*)pl->xw.ptr);
    if (ctx->profile) PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    if (recompute && ctx->tune_overlap == 2) {
        /* sequential: all lists first, then the block loop */
        rc = mlp == PCT_MLP_FP64 && pl->vorder.ptr
                 ? pct_mlp_verify_ordered(ctx, ctx->stream, hist, (const uint32_t*)pl->vorder.ptr,
                                          (const uint32_t*)pl->hash.ptr, (uint32_t*)pl->vwork.ptr)
                 : mlp == PCT_MLP_FP64
                 ? pct_mlp_verify_lanes(ctx, ctx->stream, hist, 0, pl->n, (const uint16_t*)pl->lane_of.ptr,
                                        (const uint32_t*)pl->hash.ptr)
                 : pct_mlp_paths_on(ctx, ctx->stream, hist, 0, pl->n, (const int64_t*)pl->hist_off.ptr, nullptr,
                                    nullptr, nullptr, nullptr, (const uint32_t*)pl->hash.ptr, mlp, false);
        if (rc) return rc;
        recompute = false;
        if (ctx->profile) {
            PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
            PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
        }
    } else if (recompute) {
        if ((rc = plan_ensure_events(ctx, pl, nchunks))) return rc;
        /* the surrogate's fallback list sized for the whole scan before any
         * chunk is queued (growing it later would wait for the device) */
        if (mlp == PCT_MLP_SURROGATE && (rc = pct_mlp_fallback_reserve(ctx, pl->n))) return rc;
        hipStream_t aux = ctx->aux_stream;
        PCT_HIP_CHECK(ctx, hipEventRecord(pl->ev_start, ctx->stream));
        PCT_HIP_CHECK(ctx, hipStreamWaitEvent(aux, pl->ev_start, 0));
        for (int64_t c = 0; c < nchunks; ++c) {
            int64_t h0 = pl->chunk_b[c] * pl->B;
            int64_t h1 = std::min(pl->n, pl->chunk_b[c + 1] * pl->B);
            /* the exact walk's check: the lane-per-history pass */
            if (mlp == PCT_MLP_FP64 && nchunks == 1 && pl->vorder.ptr)
                rc = pct_mlp_verify_ordered(ctx, aux, hist, (const uint32_t*)pl->vorder.ptr,
                                            (const uint32_t*)pl->hash.ptr, (uint32_t*)pl->vwork.ptr);
            else if (mlp == PCT_MLP_FP64)
                rc = pct_mlp_verify_lanes(ctx, aux, hist, h0, h1 - h0, (const uint16_t*)pl->lane_of.ptr,
                                          (const uint32_t*)pl->hash.ptr + h0);
            else
                rc = pct_mlp_paths_on(ctx, aux, hist, h0, h1 - h0, (const int64_t*)pl->hist_off.ptr + h0, nullptr,
                                      nullptr, nullptr, nullptr, (const uint32_t*)pl->hash.ptr + h0, mlp, true);
            if (rc) return rc;
            PCT_HIP_CHECK(ctx, hipEventRecord(pl->ev_chunk[c], aux));
            if (ctx->profile && c == 0) PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], aux));
        }
        if (ctx->profile) PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[2], aux));
    } else if (ctx->profile && !sequential) {
        PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
        PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    }
    /* the block loop reads the plan's lists: by default it waits for the check
     * of the recomputed lists only at the end (a mismatch restores x then);
     * overlap 1 makes chunk c wait for the check of its own lists */
    const bool chunk_waits = recompute && ctx->tune_overlap == 1;
    for (int64_t c = 0; c < nchunks; ++c) {
        if (chunk_waits) PCT_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, pl->ev_chunk[c], 0));
        if (eager) {
            int64_t b0 = pl->chunk_b[c], b1 = pl->chunk_b[c + 1];
            if ((rc = plan_block_loop(ctx, pl, hist, xt, lambda, f32, variant, b0, b1))) return rc;
        } else {
            PCT_HIP_CHECK(ctx, hipGraphLaunch(pl->graphs[c], ctx->stream));
        }
    }
    if (recompute && !chunk_waits) PCT_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, pl->ev_chunk[nchunks - 1], 0));
    if (words)
        hipLaunchKernelGGL(k_xw_unpack, dim3(pct_grid(Vt, 256)), dim3(256), 0, ctx->stream,
                           (const uint64_t*)pl->xw.ptr, Vt, xt, ctx->d_status);
    hipLaunchKernelGGL(k_tile_out, tgrid, dim3(256), 0, ctx->stream, (const float*)xt, tl, x);
    if (ctx->profile) {
        PCT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
        ctx->ev_recorded = true;
    }
    return PCT_OK;
}

/* ------------------------------------------------------------------------- */
/* capacity: does the plan of a scan fit the device?                          */
/* ------------------------------------------------------------------------- */
/* Bytes a plan takes at its build's peak: n histories, `slots` MLP capacity
 * slots (the sum over histories of the step count + 1, an upper bound of the
 * path lengths: the capacity-strided build store holds 4 B per slot),
 * `list_slots` slots of the block-strided list store (4 B each, 12 B with the
 * one-launch form's previous-block records), `entries` path entries (at most
 * one 8 B segment record and one 2 B csc index each), 94 B per history kept
 * (length, chord, hash, row offset, placement, rank, the verification
 * pass's orders and gather space) and 34 B per history plus up
 * to 4 GiB of workspace while building (offsets, placement sort, the radix
 * sort of a chunk of blocks, the voxel maps). Before the MLP pass only n and
 * slots are known: list_slots = entries = 0 gives what the build needs to
 * start. */
extern "C" int64_t pct_drop_plan_bytes(int64_t n, int64_t slots, int64_t list_slots, int64_t entries,
                                       int with_prev) {
    return 4 * slots + (with_prev ? 12 : 4) * list_slots + 10 * entries + 128 * n + ((int64_t)4 << 30);
}

/* 1: build a plan; 0: run the chunked reference structure instead -- when
 * the plan's segment records cannot index the grid (V > 2^22) or the block
 * (more than 8191 histories), or its peak bytes exceed `budget` */
extern "C" int pct_drop_plan_choice(int64_t n, int64_t slots, int64_t list_slots, int64_t entries, int64_t V,
                                    int block_size, int64_t budget, int with_prev, int64_t* need) {
    const int64_t b = pct_drop_plan_bytes(n, slots, list_slots, entries, with_prev);
    if (need) *need = b;
    if (V > SEG_MAX_V || block_size <= 0 || block_size > SEG_MAX_B) return 0;
    return b <= budget ? 1 : 0;
}

/* device bytes a plan may take now: free memory plus what the context's plan
 * holds (a rebuild reuses or frees it), less 1 GiB, capped by the knob */
static int64_t plan_budget(pct_ctx* ctx) {
    size_t fr = 0, tot = 0;
    int64_t budget = hipMemGetInfo(&fr, &tot) == hipSuccess ? (int64_t)fr : 0;
    if (ctx->plan) {
        pct_plan* pl = ctx->plan;
        for (pct_buffer* b : {&pl->list, &pl->prev, &pl->segs, &pl->csc, &pl->long_rows, &pl->hist_off, &pl->pos_k,
                              &pl->pos_len, &pl->pos_chord, &pl->pos_r, &pl->len, &pl->chord, &pl->hash})
            budget += (int64_t)b->bytes;
    }
    budget -= (int64_t)1 << 30;
    if (ctx->tune_plan_limit_mb > 0) budget = std::min<int64_t>(budget, (int64_t)ctx->tune_plan_limit_mb << 20);
    return budget;
}

static void plan_free_buffers(pct_plan* pl) {
    plan_release_graph(pl);
    for (pct_buffer* b : {&pl->list, &pl->prev, &pl->hist_off, &pl->len, &pl->chord, &pl->hash, &pl->lane_of, &pl->vorder, &pl->vwork, &pl->segs, &pl->csc,
                          &pl->long_rows, &pl->blk_seg, &pl->blk_long, &pl->blk_entry, &pl->delta, &pl->xw, &pl->xt,
                          &pl->pos_k, &pl->pos_len, &pl->pos_chord, &pl->pos_r, &pl->rank})
        release(*b);
    pl->valid = false;
    ++pl->generation;
}

extern "C" int pct_drop_prepare(pct_ctx* ctx, const pct_hist_soa* hist, int block_size, int mode) {
    if (!ctx || !hist || !hist->wepl || !hist->first_voxel || !hist->x_entry)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_prepare: null argument");
    const int mlp = mode & PCT_MLP_MODE_MASK;
    if (!pct_mlp_mode_ok(ctx, mlp)) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_prepare: MLP mode %d not available", mlp);
    if (hist->n <= 0) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_prepare: no histories");
    bool capacity = false;
    int rc = plan_build(ctx, hist, block_size, mlp, &capacity);
    if (rc && capacity)
        return pct_fail(ctx, PCT_E_NOMEM, "pct_drop_prepare: %s; pct_drop_iteration runs this scan in chunks of "
                                          "blocks instead", ctx->err.c_str());
    return rc;
}

extern "C" int pct_drop_iteration(pct_ctx* ctx, const pct_hist_soa* hist, float* x, int block_size, float lambda, int mode) {
    if (!ctx || !hist) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_iteration: null argument");
    if (hist->n == 0 && x && block_size > 0) return PCT_OK;   /* no history: x is left as it is */
    if (!hist->wepl || !hist->first_voxel || !hist->x_entry)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_iteration: null argument");
    const int mlp = mode & PCT_MLP_MODE_MASK;
    if (!pct_mlp_mode_ok(ctx, mlp))
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_iteration: MLP mode %d not available", mlp);
    if (!x || block_size <= 0) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_iteration: bad x or block size");
    if (hist->n <= 0) return PCT_OK;
    const int variant = update_variant(mode);
    if (variant != 0 && (mode & (PCT_DROP_F32 | PCT_DROP_ATOMIC | (variant == 1 ? PCT_DROP_ROBUST2 : 0))))
        return pct_fail(ctx, PCT_E_INVALID,
                        "pct_drop_iteration: the robust updates are FP64 plan forms (no PCT_DROP_F32 or ATOMIC, one "
                        "variant at a time)");
    if (mode & PCT_DROP_ATOMIC) return pct_drop_iteration_atomic(ctx, hist, x, block_size, lambda, mlp);
    int rc;
    pct_plan* pl = ctx->plan;
    bool fresh = false;
    if (!plan_matches(ctx, pl, hist, block_size, mlp)) {
        /* a scan whose plan does not fit the device (or whose grid or block
         * size the plan's records cannot index) runs in chunks of blocks, the
         * reference's structure; the research variants need the plan */
        bool capacity = false;
        if ((rc = plan_build(ctx, hist, block_size, mlp, &capacity))) {
            if (!capacity) return rc;
            if (variant != 0)
                return pct_fail(ctx, PCT_E_NOMEM, "pct_drop_iteration: the research variants need a DROP plan (%s)",
                                ctx->err.c_str());
            const int64_t store = std::max<int64_t>(1, std::min<int64_t>((int64_t)8 << 30, plan_budget(ctx) / 2));
            return pct_drop_iteration_chunked(ctx, hist, x, block_size, lambda, mlp, store);
        }
        pl = ctx->plan;
        fresh = true;
    }
    /* the reference recomputes every MLP path in every iteration; so does the
     * default mode, beside the block loop, checking every list against the
     * plan's hash (the block loop reads the plan's lists) */
    const bool recompute = !(mode & PCT_DROP_CACHED) && !fresh;
    if (recompute) {
        /* a recomputed list that differs from the plan is only known once the
         * block loop has run on the plan's lists (and the changed inputs'
         * WEPL): keep x to restore it then */
        if ((rc = pct_ensure(ctx, ctx->x_saved, sizeof(float) * (size_t)ctx->g.V))) return rc;
        PCT_HIP_CHECK(ctx, hipMemcpyAsync(ctx->x_saved.ptr, x, sizeof(float) * (size_t)ctx->g.V,
                                          hipMemcpyDeviceToDevice, ctx->stream));
    }
    rc = plan_iterate(ctx, pl, hist, x, (double)lambda, (mode & PCT_DROP_F32) != 0, variant, mlp, recompute,
                      (mode & PCT_DROP_NO_GRAPH) != 0);
    if (rc) return rc;
    int flags = 0;
    if ((rc = pct_take_status(ctx, &flags))) return rc;
    /* recomputed lists that differ from the plan (or overflow a stride, or a
     * wait of the waiting form that gave up): x goes back to its value before
     * the call (the waiting form never wrote it), and the plan is rebuilt on
     * the next call */
    if (flags) {
        pl->valid = false;
        if (recompute && (flags & (PCT_STATUS_PLAN | PCT_STATUS_CAPACITY | PCT_STATUS_TABLE | PCT_STATUS_OVERFLOW))) {
            PCT_HIP_CHECK(ctx, hipMemcpyAsync(x, ctx->x_saved.ptr, sizeof(float) * (size_t)ctx->g.V,
                                              hipMemcpyDeviceToDevice, ctx->stream));
            PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        }
    }
    return pct_status_to_rc(ctx, flags);
}

extern "C" int pct_drop_invalidate(pct_ctx* ctx) {
    if (!ctx) return PCT_E_INVALID;
    if (ctx->plan) {
        plan_release_graph(ctx->plan);
        ctx->plan->valid = false;
        ++ctx->plan->generation;
    }
    return PCT_OK;
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

extern "C" int pct_set_tuning(pct_ctx* ctx, int key, int value) {
    if (!ctx) return PCT_E_INVALID;
    if (value < 0) value = 0;
    switch (key) {
    case PCT_TUNE_MLP_WAVES: ctx->tune_mlp_waves = value; break;
    case PCT_TUNE_CHUNK_FIRST: ctx->tune_chunk_first = value; break;
    case PCT_TUNE_CHUNK_MAX: ctx->tune_chunk_max = value; break;
    case PCT_TUNE_PLAN_LIMIT_MB: ctx->tune_plan_limit_mb = value; break;
    case PCT_TUNE_WAIT_GIVE_UP: ctx->tune_wait_give_up = value ? 1 : 0; break;
    case PCT_TUNE_BLOCK_FORM:
        if (value > 3) return pct_fail(ctx, PCT_E_INVALID, "pct_set_tuning: block form %d not in {1, 2, 3}", value);
        ctx->tune_block_form = value;
        break;
    case PCT_TUNE_OVERLAP:
        if (value > 2) return pct_fail(ctx, PCT_E_INVALID, "pct_set_tuning: overlap %d not in {1, 2}", value);
        ctx->tune_overlap = value;
        break;
    default: return pct_fail(ctx, PCT_E_INVALID, "pct_set_tuning: unknown key %d", key);
    }
    return PCT_OK;
}
