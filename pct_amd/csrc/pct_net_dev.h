/*
 * pct_net_dev.h -- device side of the path network (learned-path form):
 * BF16 conversion, the activation, the MFMA operand types and the inference
 * of one wavefront's 64 histories. Shared by pct_net.hip (training, forward
 * entry point) and pct_mlp.hip (the surrogate walk).
 *
 * MFMA lane layouts on gfx950 (wave64, 16x16 output tile, lane l, g = l / 16):
 *   v_mfma_f32_16x16x16_bf16 : A[l%16][4g+i], B[4g+i][l%16], i = 0..3
 *   v_mfma_f32_16x16x32_bf16 : A[l%16][8g+i], B[8g+i][l%16], i = 0..7
 *   result D (both)          : D[4g+i][l%16], i = 0..3
 */
#ifndef PCT_NET_DEV_H
#define PCT_NET_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/pct_net.h"

typedef __attribute__((ext_vector_type(4))) short net_bf16x4;
typedef __attribute__((ext_vector_type(8))) __bf16 net_bf16x8;
typedef __attribute__((ext_vector_type(4))) float net_f32x4;
typedef __attribute__((ext_vector_type(4))) uint32_t net_u32x4;
typedef __attribute__((ext_vector_type(2))) uint32_t net_u32x2;

/* Inference operands of the installed network, built on the host from the FP32
 * parameters (pct_net.hip). The network runs transposed -- hidden units are the
 * rows of every product, the wave's histories the columns -- so each layer's
 * result tiles stay in registers and serve as the next layer's B operand; the
 * K index of the 16x16x32 products is therefore permuted: slot 8g + j of K tile
 * c is hidden unit 32c + 4g + j (j < 4) or 32c + 16 + 4g + j - 4 (j >= 4). */
struct pct_net_frag {
    net_u32x2 w1[4][64];        /* A of layer 1, M tile r: W1[4g+i][16r + l%16] (0 for rows >= 12) */
    net_u32x4 w2[4][2][64];     /* A of layer 2, M tile r, K tile c                               */
    net_u32x4 w3[2][64];        /* A of layer 3, K tile c: W3[hidden][l%16] (0 for l%16 >= 8)      */
    float b1[64], b2[64], b3[8];
};

/* float -> bfloat16 bits, round to nearest even (finite inputs) */
__host__ __device__ static inline uint32_t pct_f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

/* tanh from one hardware exp2 and one reciprocal: sign(x) (1 - e) / (1 + e),
 * e = 2^(-2 |x| log2(e)); absolute error below 1e-7 */
__device__ static inline float pct_net_tanh(float x) {
    float ax = __builtin_fabsf(x);
    float e = __builtin_amdgcn_exp2f(ax * -2.8853900817779268f);
    float t = (1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e);
    return __builtin_copysignf(t, x);
}

__device__ static inline net_f32x4 pct_mfma16(net_u32x2 a, net_u32x2 b, net_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(net_bf16x4, a),
                                                      __builtin_bit_cast(net_bf16x4, b), c, 0, 0, 0);
}
__device__ static inline net_f32x4 pct_mfma32(net_u32x4 a, net_u32x4 b, net_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(net_bf16x8, a),
                                                   __builtin_bit_cast(net_bf16x8, b), c, 0, 0, 0);
}
__device__ static inline uint32_t pct_pack_bf2(float lo, float hi) { return pct_f2bf(lo) | (pct_f2bf(hi) << 16); }

/* Network outputs of the 64 histories of one wavefront. xs[64][16] holds each
 * history's 12 inputs as bfloat16 bits (columns 12..15 zero); out[64][8]
 * receives the outputs. Every lane of the wave must call it. */
__device__ static inline void pct_net_infer_wave(const pct_net_frag* __restrict__ nf, const uint16_t* xs, float* out) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, c16 = lane & 15;
    const net_f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma nounroll
    for (int nt = 0; nt < 4; ++nt) {
        /* a compiler-only barrier (no instruction): the weights are read again
         * for each group of 16 histories instead of being held in registers
         * across the groups, which would raise the register count of every
         * kernel that runs the network */
        __asm__ volatile("" ::: "memory");
        /* B of layer 1: inputs 4g..4g+3 of history 16 nt + l%16 */
        net_u32x2 xb = *(const net_u32x2*)(xs + (16 * nt + c16) * 16 + 4 * g);
        net_u32x4 b2op[2];
        for (int r = 0; r < 4; ++r) {
            net_f32x4 z = pct_mfma16(nf->w1[r][lane], xb, zero);
            float a[4];
            for (int i = 0; i < 4; ++i) a[i] = pct_net_tanh(z[i] + nf->b1[16 * r + 4 * g + i]);
            const int c = r >> 1, half = (r & 1) * 2;
            b2op[c][half] = pct_pack_bf2(a[0], a[1]);
            b2op[c][half + 1] = pct_pack_bf2(a[2], a[3]);
        }
        net_u32x4 b3op[2];
        for (int r = 0; r < 4; ++r) {
            net_f32x4 z = pct_mfma32(nf->w2[r][0][lane], b2op[0], zero);
            z = pct_mfma32(nf->w2[r][1][lane], b2op[1], z);
            float a[4];
            for (int i = 0; i < 4; ++i) a[i] = pct_net_tanh(z[i] + nf->b2[16 * r + 4 * g + i]);
            const int c = r >> 1, half = (r & 1) * 2;
            b3op[c][half] = pct_pack_bf2(a[0], a[1]);
            b3op[c][half + 1] = pct_pack_bf2(a[2], a[3]);
        }
        net_f32x4 z = pct_mfma32(nf->w3[0][lane], b3op[0], zero);
        z = pct_mfma32(nf->w3[1][lane], b3op[1], z);
        if (g < 2)
            for (int i = 0; i < 4; ++i) out[(16 * nt + c16) * 8 + 4 * g + i] = z[i] + nf->b3[4 * g + i];
    }
}

#endif
