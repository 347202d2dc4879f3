/*
 * pct_plan_kernels.h -- the per-block DROP kernels' bodies (one history's
 * delta, one voxel's update), shared by the single-scan block loop
 * (pct_plan.hip) and the batched multi-scan block loop (pct_batch.hip).
 *
 * delta_i = L (b_i - L sum_{v in path} x_v) / (n L^2) lambda
 *     (block_update_GPU, .cu:5822-5834), a.x summed per lane over the list
 *     slots then across the wavefront by DPP exchanges;
 * x_v += (sum of the deltas of the block's histories through v, in history
 *     order) / count (DROP_blocks + DROP_update, .cu:8023-8055).
 * FP64 as the oracle, or FP32 with Kahan-compensated sums (precision form).
 */
#ifndef PCT_PLAN_KERNELS_H
#define PCT_PLAN_KERNELS_H

#include <stdint.h>
#include <hip/hip_runtime.h>

#define LIST_SENTINEL 0xFFFFFFFFu
/* bit 31 of a list slot: the voxel is also on a path of the previous block
 * (set only in plans built for the waiting one-launch form; every reader
 * masks it off) */
#define LIST_TAG 0x80000000u
#define LIST_VOX(e) ((e) & ~LIST_TAG)
/* 8-byte segment record: bits 0-21 voxel, 22-23 inline count (1-3; 0 = the
 * history indices are in csc), then either up to three 13-bit block-local
 * history indices (bits 24-36, 37-49, 50-62) or the block-relative first csc
 * entry (bits 24-50) and the count (bits 51-63). Hence V <= 2^22 voxels and
 * blocks of at most 8191 histories. */
#define SEG_VOX_BITS 22
#define SEG_IDX_BITS 13
#define SEG_INLINE 3
#define SEG_MAX_V ((int64_t)1 << SEG_VOX_BITS)
#define SEG_MAX_B (((int64_t)1 << SEG_IDX_BITS) - 1)

/* wavefronts (histories) per workgroup of the delta kernels: 2 measured
 * 71.4 ms for the cfg2 block loop against 72.4 ms with 4 and 72.8 with 8 */
#ifndef DELTA_WAVES
#define DELTA_WAVES 2
#endif
/* threads per workgroup of the apply kernels (one segment each) */
#ifndef APPLY_THREADS
#define APPLY_THREADS 256
#endif
#define DELTA_SLOTS 8
/* history indices of a long segment loaded per memory round */
#define CSC_BATCH 8

namespace pct_blk {

/* The plan's voxel index space: the image in tiles of TILE_X x TILE_Y x
 * TILE_Z voxels (32 FP32 values, one 128-byte line), tiles in row order, the
 * grid padded to whole tiles. A path's consecutive voxels share lines in
 * every in-plane direction, where row order shares them only along x: the
 * delta kernels' x gathers touch about a third of the lines. Lists, segment
 * records and the block loop's image use tiled indices; an iteration copies x
 * into the tiled image and back. */
#ifndef TILE_X
#define TILE_X 4
#define TILE_Y 8
#define TILE_Z 1
#endif
struct vox_tiling {
    int32_t C, R, S;        /* grid columns, rows, slices                  */
    int32_t tiles_x;        /* tiles per row of tiles: ceil(C / TILE_X)   */
    int64_t per_slab;       /* tiled voxels per TILE_Z slices              */
    int64_t Vt;             /* tiled voxels                                */
};

inline vox_tiling make_tiling(int C, int R, int S) {
    vox_tiling t;
    t.C = C;
    t.R = R;
    t.S = S;
    t.tiles_x = (C + TILE_X - 1) / TILE_X;
    t.per_slab = (int64_t)t.tiles_x * ((R + TILE_Y - 1) / TILE_Y) * (TILE_X * TILE_Y * TILE_Z);
    t.Vt = t.per_slab * ((S + TILE_Z - 1) / TILE_Z);
    return t;
}

__host__ __device__ __forceinline__ uint32_t tiled_xyz(const vox_tiling& t, int x, int y, int z) {
    return (uint32_t)((int64_t)(z / TILE_Z) * t.per_slab +
                      ((int64_t)(y / TILE_Y) * t.tiles_x + x / TILE_X) * (TILE_X * TILE_Y * TILE_Z) +
                      ((z % TILE_Z) * TILE_Y + y % TILE_Y) * TILE_X + x % TILE_X);
}

/* the tiled index of row-order voxel v */
__host__ __device__ __forceinline__ uint32_t tiled_of(const vox_tiling& t, uint32_t v) {
    const uint32_t x = v % (uint32_t)t.C, r = v / (uint32_t)t.C;
    return tiled_xyz(t, (int)x, (int)(r % (uint32_t)t.R), (int)(r / (uint32_t)t.R));
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

/* v of lane l's DPP source lane (both halves of the double moved alike) */
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

/* the Kahan-compensated sum (s, c) over the wavefront, valid in lane 0: the
 * stages of wave_sum_d below, each combining two compensated partial sums */
__device__ __forceinline__ void wave_kahan_sum(float& s, float& c) {
    auto stage = [&](float os, float oc) {
        c += oc;
        kahan_add(s, c, os);
    };
    stage(dpp_f<0xB1>(s), dpp_f<0xB1>(c));
    stage(dpp_f<0x4E>(s), dpp_f<0x4E>(c));
    stage(dpp_f<0x141>(s), dpp_f<0x141>(c));
    stage(dpp_f<0x140>(s), dpp_f<0x140>(c));
    const float s1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), 16));
    const float c1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c), 16));
    const float s2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), 32));
    const float c2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c), 32));
    const float s3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), 48));
    const float c3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c), 48));
    s = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), 0));
    c = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c), 0));
    stage(s1, c1);
    stage(s2, c2);
    stage(s3, c3);
}

/* The sum over the wavefront of v, valid in lane 0: within each row of 16
 * lanes by DPP exchanges (pairs, quads, half rows, rows: every lane of a
 * group adds the same two operands, so each stage leaves one value per
 * group), then the four rows' sums in order. Register moves instead of LDS
 * permutes on the delta kernels' critical path. The a.x sums it serves are
 * sums of FP32 values in FP64: every partial sum is a multiple of the
 * smallest term's ulp, so the sum is exact, and equal to the oracle's
 * sequential one, while sum |x_v| < 2^53 times that ulp (rows of at most
 * 1024 voxels with |x| <= 2: every nonzero |x_v| >= 2^-19). */
__device__ __forceinline__ double wave_sum_d(double v) {
    v += dpp_d<0xB1>(v);    /* quad_perm [1,0,3,2] */
    v += dpp_d<0x4E>(v);    /* quad_perm [2,3,0,1] */
    v += dpp_d<0x141>(v);   /* row_half_mirror     */
    v += dpp_d<0x140>(v);   /* row_mirror          */
    const double r0 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 0),
                                       __builtin_amdgcn_readlane(__double2loint(v), 0));
    const double r1 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 16),
                                       __builtin_amdgcn_readlane(__double2loint(v), 16));
    const double r2 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 32),
                                       __builtin_amdgcn_readlane(__double2loint(v), 32));
    const double r3 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 48),
                                       __builtin_amdgcn_readlane(__double2loint(v), 48));
    return ((r0 + r1) + r2) + r3;
}

/* x at a lane's list slots, every load issued before the first is used (one
 * memory round): a sentinel slot loads x[0], which its caller ignores. A
 * conditional load per slot would be compiled as a branch around each load
 * and its use, one round per slot. */
template <int SLOTS>
__device__ __forceinline__ void gather_slots(const float* __restrict__ x, const uint32_t (&v)[SLOTS],
                                             float (&xv)[SLOTS]) {
#pragma unroll
    for (int r = 0; r < SLOTS; ++r) xv[r] = x[v[r] != LIST_SENTINEL ? LIST_VOX(v[r]) : 0u];
}

/* The block's histories in placement order: position q's list is row q of the
 * block's rows (list + q * stride), and pos.k[q], pos.len[q], pos.chord[q]
 * are that history's block-local index, list length and chord length; the
 * block's first history is h0. Rows are placed so that the workgroups on one
 * XCD get one lateral band of the block's paths (the plan build); results do
 * not depend on the placement. pos.r[q] is the history's rank in lateral
 * order within the block: its delta goes to delta[r], and the segment
 * records name histories by rank (still listed in history order), so the
 * histories through a compact group of voxels have nearby deltas. */
struct blk_pos {
    const uint16_t* k;
    const uint32_t* len;
    const double* chord;
    const uint16_t* r;
};

/* the history at position q of the block (one wavefront); SLOTS list slots
 * per lane in the first round (rows up to 64 SLOTS long), the rest after */
template <int SLOTS = DELTA_SLOTS>
__device__ __forceinline__ void delta_f64(const uint32_t* __restrict__ list, int stride, blk_pos pos,
                                          const float* __restrict__ wepl, int64_t h0, int64_t q,
                                          const float* __restrict__ x, double lambda, double* __restrict__ delta,
                                          int lane) {
    const uint32_t k = pos.k[q];
    const uint32_t rk = pos.r[q];
    const uint32_t n = pos.len[q];
    const double L = pos.chord[q];
    const uint32_t* p = list + q * stride;
    uint32_t v[SLOTS];
#pragma unroll
    for (int r = 0; r < SLOTS; ++r) {
        int j = lane + 64 * r;
        v[r] = j < stride ? p[j] : LIST_SENTINEL;
    }
    float xv[SLOTS];
    gather_slots(x, v, xv);
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < SLOTS; ++r)
        if (v[r] != LIST_SENTINEL) s += (double)xv[r];
    for (int j = lane + 64 * SLOTS; j < stride; j += 64) {
        uint32_t w = p[j];
        if (w != LIST_SENTINEL) s += (double)x[LIST_VOX(w)];
    }
    const float b_i = wepl[h0 + k];
    s = wave_sum_d(s);
    if (lane == 0) {
        const double a_i_dot_x_k = s * L;
        const double a_i_dot_a_i = (L * L) * n;
        delta[rk] = L * ((b_i - a_i_dot_x_k) / a_i_dot_a_i) * lambda;
    }
}

template <int SLOTS = DELTA_SLOTS>
__device__ __forceinline__ void delta_f32(const uint32_t* __restrict__ list, int stride, blk_pos pos,
                                          const float* __restrict__ wepl, int64_t h0, int64_t q,
                                          const float* __restrict__ x, float lambda, float* __restrict__ delta,
                                          int lane) {
    const uint32_t k = pos.k[q];
    const uint32_t rk = pos.r[q];
    const uint32_t n = pos.len[q];
    const float L = (float)pos.chord[q];
    const uint32_t* p = list + q * stride;
    uint32_t v[SLOTS];
#pragma unroll
    for (int r = 0; r < SLOTS; ++r) {
        int j = lane + 64 * r;
        v[r] = j < stride ? p[j] : LIST_SENTINEL;
    }
    float xv[SLOTS];
    gather_slots(x, v, xv);
    float s = 0.0f, c = 0.0f;
#pragma unroll
    for (int r = 0; r < SLOTS; ++r)
        if (v[r] != LIST_SENTINEL) kahan_add(s, c, xv[r]);
    for (int j = lane + 64 * SLOTS; j < stride; j += 64) {
        uint32_t w = p[j];
        if (w != LIST_SENTINEL) kahan_add(s, c, x[LIST_VOX(w)]);
    }
    wave_kahan_sum(s, c);
    const float b_i = wepl[h0 + k];
    if (lane == 0) {
        const float a_i_dot_x_k = s * L;
        const float a_i_dot_a_i = (L * L) * (float)n;
        delta[rk] = L * ((b_i - a_i_dot_x_k) / a_i_dot_a_i) * lambda;
    }
}

/* The research variants of the FP64 update (the reference's CPU functions,
 * with its expressions and operation order):
 *   1 DROP_blocks_red + DROP_update (.cu:8005-8055): history k (rank rk)
 *     writes its residual r = b - L sum x to out[rk] and a.a = L^2 n to
 *     out[B + rk];
 *     voxel v adds, over the block's histories through it in history order,
 *     (L lambda) (r / (a.a + psi_v)), psi_v = ((1 - x_v) ETA) PSI_SIGN, and
 *     x_v += sum / count;
 *   2 DROP_blocks_robust2 + DROP_update_robust2 (.cu:7871-7901): history k
 *     writes (L lambda) (r / a.a) to out[rk]; voxel v sums those and
 *     N_v = sum L^2, and x_v += (N_v / (N_v + psi_v)) sum / count with
 *     psi_v = PSI_SIGN (1 - x_v) ETA, the first product in float.
 * Both write the chord L to out[2B + rk] for the voxel updates. x_v is the
 * value at the block's start in both. */
__device__ __forceinline__ void delta_variant(int variant, const uint32_t* __restrict__ list, int stride,
                                              blk_pos pos, const float* __restrict__ wepl, int64_t h0, int64_t q,
                                              const float* __restrict__ x, double lambda, double* __restrict__ out,
                                              int64_t B, int lane) {
    const uint32_t k = pos.k[q];
    const uint32_t rk = pos.r[q];
    const uint32_t n = pos.len[q];
    const double L = pos.chord[q];
    const uint32_t* p = list + q * stride;
    uint32_t v[DELTA_SLOTS];
#pragma unroll
    for (int r = 0; r < DELTA_SLOTS; ++r) {
        int j = lane + 64 * r;
        v[r] = j < stride ? p[j] : LIST_SENTINEL;
    }
    float xv[DELTA_SLOTS];
    gather_slots(x, v, xv);
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < DELTA_SLOTS; ++r)
        if (v[r] != LIST_SENTINEL) s += (double)xv[r];
    for (int j = lane + 64 * DELTA_SLOTS; j < stride; j += 64) {
        uint32_t w = p[j];
        if (w != LIST_SENTINEL) s += (double)x[LIST_VOX(w)];
    }
    const float b_i = wepl[h0 + k];
    s = wave_sum_d(s);
    if (lane == 0) {
        const double a_i_dot_x_k = L * s;
        const double a_i_dot_a_i = (L * L) * n;
        const double residual = b_i - a_i_dot_x_k;
        if (variant == 1) {
            out[rk] = residual;
            out[B + rk] = a_i_dot_a_i;
        } else {
            out[rk] = (L * lambda) * (residual / a_i_dot_a_i);
        }
        out[2 * B + rk] = L;
    }
}

/* history indices i .. i+CSC_BATCH-1 of a long segment (cnt > i), every load
 * issued before the first is used: positions past the end load the last
 * index again, which the caller ignores */
__device__ __forceinline__ void csc_batch(const uint16_t* __restrict__ q, uint32_t i, uint32_t cnt,
                                          uint32_t (&idx)[CSC_BATCH]) {
#pragma unroll
    for (int t = 0; t < CSC_BATCH; ++t) idx[t] = q[min(i + (uint32_t)t, cnt - 1u)];
}

/* the block-local history indices of segment record r, in history order */
template <class F>
__device__ __forceinline__ uint32_t for_each_history(uint64_t r, const uint16_t* __restrict__ csc, F f) {
    const uint32_t c = (uint32_t)(r >> SEG_VOX_BITS) & 3u;
    if (c != 0u) {
        f((uint32_t)((r >> 24) & 0x1fffu));
        if (c > 1u) f((uint32_t)((r >> 37) & 0x1fffu));
        if (c > 2u) f((uint32_t)((r >> 50) & 0x1fffu));
        return c;
    }
    const uint32_t cnt = (uint32_t)(r >> 51);
    const uint16_t* q = csc + ((r >> 24) & 0x7ffffffu);
    for (uint32_t i = 0; i < cnt; i += CSC_BATCH) {
        uint32_t idx[CSC_BATCH];
        csc_batch(q, i, cnt, idx);
#pragma unroll
        for (int t = 0; t < CSC_BATCH; ++t)
            if (i + t < cnt) f(idx[t]);
    }
    return cnt;
}

template <int VARIANT>
__device__ __forceinline__ void apply_variant(uint64_t r, const uint16_t* __restrict__ csc,
                                              const double* __restrict__ out, int64_t B, double lambda, double eta,
                                              int psi_sign, float* __restrict__ x) {
    const double* __restrict__ chord_b = out + 2 * B;
    const uint32_t v = (uint32_t)(r & (SEG_MAX_V - 1));
    const float xv = x[v];
    double upd = 0.0;
    if (VARIANT == 1) {
        const double psi = ((1.0 - (double)xv) * eta) * (double)psi_sign;
        const uint32_t cnt = for_each_history(r, csc, [&](uint32_t i) {
            upd += (chord_b[i] * lambda) * (out[i] / (out[B + i] + psi));
        });
        x[v] = (float)((double)xv + upd / (double)cnt);
    } else {
        double norm = 0.0;
        const uint32_t cnt = for_each_history(r, csc, [&](uint32_t i) {
            upd += out[i];
            norm += chord_b[i] * chord_b[i];
        });
        const double psi = (double)((float)psi_sign * (1.0f - xv)) * eta;
        x[v] = (float)((double)xv + ((norm / (norm + psi)) * upd) / (double)cnt);
    }
}

/* the new value of a voxel of a block from its value before the block and its
 * segment record r (csc points at the block's first entry): the deltas of the
 * block's histories through it summed in history order, x + sum / count */
__device__ __forceinline__ float applied_f64(float xv, uint64_t r, const uint16_t* __restrict__ csc,
                                             const double* __restrict__ delta) {
    const uint32_t c = (uint32_t)(r >> SEG_VOX_BITS) & 3u;
    double upd = 0.0;
    uint32_t cnt;
    if (c != 0u) {
        const double d0 = delta[(r >> 24) & 0x1fffu];
        const double d1 = c > 1u ? delta[(r >> 37) & 0x1fffu] : 0.0;
        const double d2 = c > 2u ? delta[(r >> 50) & 0x1fffu] : 0.0;
        upd += d0;
        if (c > 1u) upd += d1;
        if (c > 2u) upd += d2;
        cnt = c;
    } else {
        cnt = (uint32_t)(r >> 51);
        const uint16_t* q = csc + ((r >> 24) & 0x7ffffffu);
        /* CSC_BATCH indices, then their deltas, are loaded together before
         * they are added in history order: two memory rounds per batch */
        for (uint32_t i = 0; i < cnt; i += CSC_BATCH) {
            uint32_t idx[CSC_BATCH];
            double d[CSC_BATCH];
            csc_batch(q, i, cnt, idx);
#pragma unroll
            for (int t = 0; t < CSC_BATCH; ++t) d[t] = delta[idx[t]];
#pragma unroll
            for (int t = 0; t < CSC_BATCH; ++t)
                if (i + t < cnt) upd += d[t];
        }
    }
    return (float)((double)xv + upd / (double)cnt);
}

__device__ __forceinline__ float applied_f32(float xv, uint64_t r, const uint16_t* __restrict__ csc,
                                             const float* __restrict__ delta) {
    const uint32_t c = (uint32_t)(r >> SEG_VOX_BITS) & 3u;
    float upd = 0.0f, comp = 0.0f;
    uint32_t cnt;
    if (c != 0u) {
        const float d0 = delta[(r >> 24) & 0x1fffu];
        const float d1 = c > 1u ? delta[(r >> 37) & 0x1fffu] : 0.0f;
        const float d2 = c > 2u ? delta[(r >> 50) & 0x1fffu] : 0.0f;
        kahan_add(upd, comp, d0);
        if (c > 1u) kahan_add(upd, comp, d1);
        if (c > 2u) kahan_add(upd, comp, d2);
        cnt = c;
    } else {
        cnt = (uint32_t)(r >> 51);
        const uint16_t* q = csc + ((r >> 24) & 0x7ffffffu);
        for (uint32_t i = 0; i < cnt; i += CSC_BATCH) {
            uint32_t idx[CSC_BATCH];
            float d[CSC_BATCH];
            csc_batch(q, i, cnt, idx);
#pragma unroll
            for (int t = 0; t < CSC_BATCH; ++t) d[t] = delta[idx[t]];
#pragma unroll
            for (int t = 0; t < CSC_BATCH; ++t)
                if (i + t < cnt) kahan_add(upd, comp, d[t]);
        }
    }
    return xv + upd / (float)cnt;
}

/* ------------------------------------------------------------------------- */
/* Long segments (more than SEG_INLINE histories) come first in each block's  */
/* segment range, and the first LONG_IDX of their block-local history indices */
/* are also kept at a fixed stride in a side array (one 16-byte row per long  */
/* segment, in segment order; the unused tail of a row repeats the last       */
/* index). The two-launch apply kernels read a long segment's row beside its  */
/* record, so its deltas load in the same round as an inline segment's; only  */
/* indices past LONG_IDX go through csc.                                      */
/* ------------------------------------------------------------------------- */
#define LONG_IDX 8

struct long_row {
    uint32_t w[LONG_IDX / 2];
};

__device__ __forceinline__ uint32_t long_idx(const long_row& row, int t) {
    return (row.w[t >> 1] >> (16 * (t & 1))) & 0xffffu;
}

/* the new value of a long segment's voxel: as applied_f64, with the first
 * LONG_IDX history indices from its side row */
__device__ __forceinline__ float applied_long_f64(float xv, uint64_t r, const long_row& row,
                                                  const uint16_t* __restrict__ csc, const double* __restrict__ delta) {
    const uint32_t cnt = (uint32_t)(r >> 51);
    double d[LONG_IDX];
#pragma unroll
    for (int t = 0; t < LONG_IDX; ++t) d[t] = delta[long_idx(row, t)];
    double upd = 0.0;
#pragma unroll
    for (int t = 0; t < LONG_IDX; ++t)
        if ((uint32_t)t < cnt) upd += d[t];
    if (cnt > LONG_IDX) {
        const uint16_t* q = csc + ((r >> 24) & 0x7ffffffu);
        for (uint32_t i = LONG_IDX; i < cnt; i += CSC_BATCH) {
            uint32_t idx[CSC_BATCH];
            double e[CSC_BATCH];
            csc_batch(q, i, cnt, idx);
#pragma unroll
            for (int t = 0; t < CSC_BATCH; ++t) e[t] = delta[idx[t]];
#pragma unroll
            for (int t = 0; t < CSC_BATCH; ++t)
                if (i + t < cnt) upd += e[t];
        }
    }
    return (float)((double)xv + upd / (double)cnt);
}

__device__ __forceinline__ float applied_long_f32(float xv, uint64_t r, const long_row& row,
                                                  const uint16_t* __restrict__ csc, const float* __restrict__ delta) {
    const uint32_t cnt = (uint32_t)(r >> 51);
    float d[LONG_IDX];
#pragma unroll
    for (int t = 0; t < LONG_IDX; ++t) d[t] = delta[long_idx(row, t)];
    float upd = 0.0f, comp = 0.0f;
#pragma unroll
    for (int t = 0; t < LONG_IDX; ++t)
        if ((uint32_t)t < cnt) kahan_add(upd, comp, d[t]);
    if (cnt > LONG_IDX) {
        const uint16_t* q = csc + ((r >> 24) & 0x7ffffffu);
        for (uint32_t i = LONG_IDX; i < cnt; i += CSC_BATCH) {
            uint32_t idx[CSC_BATCH];
            float e[CSC_BATCH];
            csc_batch(q, i, cnt, idx);
#pragma unroll
            for (int t = 0; t < CSC_BATCH; ++t) e[t] = delta[idx[t]];
#pragma unroll
            for (int t = 0; t < CSC_BATCH; ++t)
                if (i + t < cnt) kahan_add(upd, comp, e[t]);
        }
    }
    return xv + upd / (float)cnt;
}

/* segment k of a block (record r): long rows for k < nlong */
template <bool F32>
__device__ __forceinline__ void apply_seg(int64_t k, int64_t nlong, uint64_t r, const long_row* __restrict__ rows,
                                          const uint16_t* __restrict__ csc, const void* __restrict__ delta,
                                          float* __restrict__ x) {
    const uint32_t v = (uint32_t)(r & (SEG_MAX_V - 1));
    if (k < nlong) {
        const long_row row = rows[k];
        const float xv = x[v];
        x[v] = F32 ? applied_long_f32(xv, r, row, csc, (const float*)delta)
                   : applied_long_f64(xv, r, row, csc, (const double*)delta);
    } else {
        const float xv = x[v];
        x[v] = F32 ? applied_f32(xv, r, csc, (const float*)delta) : applied_f64(xv, r, csc, (const double*)delta);
    }
}

/* segment record r of a block; csc points at the block's first entry */
__device__ __forceinline__ void apply_f64(uint64_t r, const uint16_t* __restrict__ csc,
                                          const double* __restrict__ delta, float* __restrict__ x) {
    const uint32_t v = (uint32_t)(r & (SEG_MAX_V - 1));
    x[v] = applied_f64(x[v], r, csc, delta);
}

__device__ __forceinline__ void apply_f32(uint64_t r, const uint16_t* __restrict__ csc,
                                          const float* __restrict__ delta, float* __restrict__ x) {
    const uint32_t v = (uint32_t)(r & (SEG_MAX_V - 1));
    x[v] = applied_f32(x[v], r, csc, delta);
}

/* ------------------------------------------------------------------------- */
/* One launch per block: the voxel updates of block b-1 and the deltas of     */
/* block b together. The image is held as 8-byte words {value, stamp}; the    */
/* update of block b-1 writes its voxels' words with stamp b-1. A history of  */
/* block b reads each path voxel's word: a voxel block b-1 does not touch has */
/* its final value already; one it touches either carries stamp b-1 (written  */
/* by this launch) or still its value before block b-1, and then the history  */
/* applies block b-1's update to it itself, from the segment record stored    */
/* beside its path slot -- the same expression as the voxel update, so the    */
/* value is the same bit for bit. Nothing waits on anything: the launch       */
/* boundary orders block b-1's deltas before both halves.                     */
/* ------------------------------------------------------------------------- */
__device__ __forceinline__ uint64_t xw_pack(float v, uint32_t stamp) {
    return ((uint64_t)stamp << 32) | (uint64_t)__float_as_uint(v);
}
__device__ __forceinline__ float xw_value(uint64_t w) { return __uint_as_float((uint32_t)w); }
__device__ __forceinline__ uint32_t xw_stamp(uint64_t w) { return (uint32_t)(w >> 32); }

__device__ __forceinline__ uint64_t xw_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xw_store(uint64_t* p, uint64_t w) {
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool F32>
struct fused_prev {
    const uint16_t* csc;          /* block b-1's first csc entry */
    const void* delta;            /* block b-1's deltas          */
    uint32_t stamp;               /* b - 1                        */
};

/* x value of path voxel v of block b after block b-1; rec = block b-1's
 * segment record of v, 0 when block b-1 does not touch v */
template <bool F32>
__device__ __forceinline__ float fused_x(uint32_t v, uint64_t rec, const uint64_t* __restrict__ xw,
                                         const fused_prev<F32>& pv) {
    const uint64_t w = xw_load(xw + v);
    if (rec == 0 || xw_stamp(w) == pv.stamp) return xw_value(w);
    if (F32) return applied_f32(xw_value(w), rec, pv.csc, (const float*)pv.delta);
    return applied_f64(xw_value(w), rec, pv.csc, (const double*)pv.delta);
}

/* list / prev: the block's voxel lists and, slot for slot, the previous
 * block's segment records of those voxels */
template <bool F32>
__device__ __forceinline__ void fused_delta(const uint32_t* __restrict__ list, const uint64_t* __restrict__ prev,
                                            int stride, blk_pos pos, const float* __restrict__ wepl, int64_t h0,
                                            int64_t q, const uint64_t* __restrict__ xw, const fused_prev<F32>& pv,
                                            double lambda, void* __restrict__ delta, int lane) {
    const uint32_t k = pos.k[q];
    const uint32_t rk = pos.r[q];
    const uint32_t n = pos.len[q];
    const double Ld = pos.chord[q];
    const uint32_t* p = list + q * stride;
    const uint64_t* pr = prev + q * stride;
    uint32_t v[DELTA_SLOTS];
    uint64_t e[DELTA_SLOTS];
#pragma unroll
    for (int r = 0; r < DELTA_SLOTS; ++r) {
        int j = lane + 64 * r;
        v[r] = j < stride ? p[j] : LIST_SENTINEL;
        e[r] = j < stride ? pr[j] : 0ull;
    }
    if (F32) {
        float s = 0.0f, c = 0.0f;
#pragma unroll
        for (int r = 0; r < DELTA_SLOTS; ++r)
            if (v[r] != LIST_SENTINEL) kahan_add(s, c, fused_x<F32>(LIST_VOX(v[r]), e[r], xw, pv));
        for (int j = lane + 64 * DELTA_SLOTS; j < stride; j += 64) {
            const uint32_t q = p[j];
            if (q != LIST_SENTINEL) kahan_add(s, c, fused_x<F32>(LIST_VOX(q), pr[j], xw, pv));
        }
        wave_kahan_sum(s, c);
        const float b_i = wepl[h0 + k];
        if (lane == 0) {
            const float L = (float)Ld;
            const float a_i_dot_x_k = s * L;
            const float a_i_dot_a_i = (L * L) * (float)n;
            ((float*)delta)[rk] = L * ((b_i - a_i_dot_x_k) / a_i_dot_a_i) * (float)lambda;
        }
    } else {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < DELTA_SLOTS; ++r)
            if (v[r] != LIST_SENTINEL) s += (double)fused_x<F32>(LIST_VOX(v[r]), e[r], xw, pv);
        for (int j = lane + 64 * DELTA_SLOTS; j < stride; j += 64) {
            const uint32_t q = p[j];
            if (q != LIST_SENTINEL) s += (double)fused_x<F32>(LIST_VOX(q), pr[j], xw, pv);
        }
        const float b_i = wepl[h0 + k];
        s = wave_sum_d(s);
        if (lane == 0) {
            const double a_i_dot_x_k = s * Ld;
            const double a_i_dot_a_i = (Ld * Ld) * n;
            ((double*)delta)[rk] = Ld * ((b_i - a_i_dot_x_k) / a_i_dot_a_i) * lambda;
        }
    }
}

/* the voxel update of block b-1's segment record r on the image words */
template <bool F32>
__device__ __forceinline__ void fused_apply(uint64_t r, uint64_t* __restrict__ xw, const fused_prev<F32>& pv) {
    const uint32_t v = (uint32_t)(r & (SEG_MAX_V - 1));
    const float xv = xw_value(xw[v]);
    const float nv = F32 ? applied_f32(xv, r, pv.csc, (const float*)pv.delta)
                         : applied_f64(xv, r, pv.csc, (const double*)pv.delta);
    xw_store(xw + v, xw_pack(nv, pv.stamp));
}

/* ------------------------------------------------------------------------- */
/* One launch per block, waiting form. The launch for block b runs block      */
/* b-1's voxel updates in its first workgroups (fused_apply above) and block  */
/* b's deltas in the rest, on the {value, stamp} words. The plan tags every   */
/* list slot of block b whose voxel block b-1 also touches (LIST_TAG). An     */
/* untagged slot's word is final when the launch starts -- its last update    */
/* ran in an earlier launch -- and is read with a plain load. A tagged slot's */
/* word is read at device scope until it carries stamp b-1, i.e. until block  */
/* b-1's update of that voxel has been stored; the history then sums exactly  */
/* the values the two-launch form reads, in the same order. The updates never */
/* wait and their workgroups come first in the grid. Every wait is bounded    */
/* all the same (WAIT_POLLS polls, or fewer once another wait has given up):  */
/* a wait that gives up sets WAIT_FLAG in the status word, the launch runs to */
/* its end, and the iteration reports the error with x unchanged.             */
/* ------------------------------------------------------------------------- */
#define WAIT_POLLS 4096
#define WAIT_FLAG 16   /* PCT_STATUS_WAIT */

__device__ __forceinline__ uint64_t wait_word(const uint64_t* p, uint64_t w, uint32_t stamp, int polls,
                                              int* status) {
    for (int i = 1; xw_stamp(w) != stamp; ++i) {
        if (i >= polls) {
            atomicOr(status, WAIT_FLAG);
            break;
        }
        if ((i & 31) == 0 && (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & WAIT_FLAG))
            break;
        __builtin_amdgcn_s_sleep(2);
        w = xw_load(p);
    }
    return w;
}

/* the word of list slot e (not the sentinel): plain for an untagged slot,
 * first device-scope read for a tagged one */
__device__ __forceinline__ uint64_t slot_word(const uint64_t* __restrict__ xw, uint32_t e) {
    const uint64_t* a = xw + LIST_VOX(e);
    return (e & LIST_TAG) ? xw_load(a) : *a;
}

template <bool F32>
__device__ __forceinline__ void wait_delta(const uint32_t* __restrict__ list, int stride, blk_pos pos,
                                           const float* __restrict__ wepl, int64_t h0, int64_t q,
                                           const uint64_t* __restrict__ xw, uint32_t stamp, int polls,
                                           int* status, double lambda, void* __restrict__ delta, int lane) {
    const uint32_t k = pos.k[q];
    const uint32_t rk = pos.r[q];
    const uint32_t n = pos.len[q];
    const double Ld = pos.chord[q];
    const uint32_t* p = list + q * stride;
    uint32_t e[DELTA_SLOTS];
    uint64_t w[DELTA_SLOTS];
#pragma unroll
    for (int r = 0; r < DELTA_SLOTS; ++r) {
        const int j = lane + 64 * r;
        e[r] = j < stride ? p[j] : LIST_SENTINEL;
    }
#pragma unroll
    for (int r = 0; r < DELTA_SLOTS; ++r) w[r] = e[r] != LIST_SENTINEL ? slot_word(xw, e[r]) : 0ull;
#pragma unroll
    for (int r = 0; r < DELTA_SLOTS; ++r)
        if (e[r] != LIST_SENTINEL && (e[r] & LIST_TAG))
            w[r] = wait_word(xw + LIST_VOX(e[r]), w[r], stamp, polls, status);
    auto tail_value = [&](uint32_t t) {
        uint64_t u = slot_word(xw, t);
        if (t & LIST_TAG) u = wait_word(xw + LIST_VOX(t), u, stamp, polls, status);
        return xw_value(u);
    };
    if (F32) {
        float s = 0.0f, c = 0.0f;
#pragma unroll
        for (int r = 0; r < DELTA_SLOTS; ++r)
            if (e[r] != LIST_SENTINEL) kahan_add(s, c, xw_value(w[r]));
        for (int j = lane + 64 * DELTA_SLOTS; j < stride; j += 64) {
            const uint32_t t = p[j];
            if (t != LIST_SENTINEL) kahan_add(s, c, tail_value(t));
        }
        wave_kahan_sum(s, c);
        const float b_i = wepl[h0 + k];
        if (lane == 0) {
            const float L = (float)Ld;
            const float a_i_dot_x_k = s * L;
            const float a_i_dot_a_i = (L * L) * (float)n;
            ((float*)delta)[rk] = L * ((b_i - a_i_dot_x_k) / a_i_dot_a_i) * (float)lambda;
        }
    } else {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < DELTA_SLOTS; ++r)
            if (e[r] != LIST_SENTINEL) s += (double)xw_value(w[r]);
        for (int j = lane + 64 * DELTA_SLOTS; j < stride; j += 64) {
            const uint32_t t = p[j];
            if (t != LIST_SENTINEL) s += (double)tail_value(t);
        }
        const float b_i = wepl[h0 + k];
        s = wave_sum_d(s);
        if (lane == 0) {
            const double a_i_dot_x_k = s * Ld;
            const double a_i_dot_a_i = (Ld * Ld) * n;
            ((double*)delta)[rk] = Ld * ((b_i - a_i_dot_x_k) / a_i_dot_a_i) * lambda;
        }
    }
}

}  // namespace pct_blk

#endif
