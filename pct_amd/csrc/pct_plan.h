/*
 * pct_plan.h -- the per-scan DROP plan (see pct_plan.hip).
 */
#ifndef PCT_PLAN_H
#define PCT_PLAN_H

#include <vector>
#include "pct_internal.h"

struct pct_plan {
    bool valid = false;
    /* bumped by every build: graphs captured elsewhere (the batched form)
     * that hold this plan's pointers and offsets compare it */
    uint64_t generation = 0;
    pct_hist_soa key{};            /* the histories the plan was built from */
    int64_t n = 0;                 /* histories                             */
    int B = 0;                     /* DROP block size                       */
    int64_t nb = 0;                /* blocks                                */
    int64_t E = 0;                 /* path entries (sum of list lengths)    */
    int64_t S = 0;                 /* segments = sum over blocks of |V_B|   */
    int mlp_mode = 0;              /* MLP mode the lists were built with    */
    int net_version = 0;           /* network + calibration of a surrogate plan */
    /* path lists, block-strided: the history at position q of block b at
     * list[list_base[b] + q * stride[b]], unused slots hold the sentinel */
    pct_buffer list;               /* u32                                   */
    /* for the one-launch-per-block form only: slot for slot, the previous
     * block's segment record of the voxel (0 if it does not touch it) */
    pct_buffer prev;               /* u64                                   */
    bool has_prev = false;
    bool has_tags = false;      /* list slots tagged for the waiting one-launch form */
    pct_buffer hist_off;           /* int64 per history+1: row offsets      */
    /* rows of a block in placement order (lateral bands per XCD): per
     * position the block-local history index, list length, chord length */
    pct_buffer pos_k;              /* u16 per history                       */
    pct_buffer pos_len;            /* u32 per history                       */
    pct_buffer pos_chord;          /* f64 per history                       */
    pct_buffer pos_r;              /* u16 per history: lateral rank, the
                                      slot of its delta                     */
    pct_buffer rank;               /* u16 per history (build only): the
                                      same by history                       */
    pct_buffer len;                /* u32 per history: list length          */
    pct_buffer chord;              /* f64 per history: effective chord      */
    pct_buffer hash;               /* u32 per history: hash of its list     */
    pct_buffer lane_of;            /* u16 per history: the verification
                                      pass's lane order within each group of
                                      64 (by step count)                    */
    pct_buffer vorder;             /* u32 per history: all histories by step
                                      count (the one-launch check's order)  */
    pct_buffer vwork;              /* 12 u32 per history: the check's inputs
                                      gathered in that order each iteration */
    /* block transpose: per block the touched voxels, those with more than the
     * inline number of histories first, each part sorted by voxel, each an
     * 8-byte record {voxel, count | history indices}; csc holds all
     * block-local history indices in (block, voxel, history) order for the
     * records with more than the inline number */
    pct_buffer segs;               /* u64 per segment                       */
    pct_buffer csc;                /* u16 per entry                         */
    pct_buffer blk_seg;            /* int64 per block+1: first segment      */
    pct_buffer blk_entry;          /* int64 per block+1: first csc entry    */
    /* the segments with more than the inline number of histories come first
     * in each block's range; long_rows holds, per such segment in segment
     * order, its first eight history indices (16 bytes; at most 6 B per entry
     * with the record, under the 8 B an inline record takes) */
    pct_buffer long_rows;          /* 16 B per long segment                 */
    pct_buffer blk_long;           /* int64 per block+1: first long segment */
    pct_buffer delta;              /* 2 f64 per history of two blocks       */
    pct_buffer xw;                 /* u64 per tiled voxel: {x, stamp} words
                                      of the one-launch-per-block forms     */
    /* the block loop's image in the tiled voxel order the lists and records
     * use (Vt floats, padding voxels zero) */
    pct_buffer xt;
    int64_t Vt = 0;
    std::vector<int64_t> blk_entry_h, blk_seg_h, blk_long_h, list_base_h;
    std::vector<int32_t> stride_h;
    /* the block loop in chunks of blocks [chunk_b[c], chunk_b[c+1]), one
     * hipGraph each; chunk c may start when the MLP lists of its histories are
     * done (ev_chunk[c]) */
    std::vector<int64_t> chunk_b;
    int64_t chunk_first = 0, chunk_max = 0;
    std::vector<hipGraphExec_t> graphs;
    std::vector<hipEvent_t> ev_chunk;
    hipEvent_t ev_start = nullptr;
    float* graph_x = nullptr;
    const float* graph_wepl = nullptr;
    double graph_lambda = 0.0;
    bool graph_f32 = false;
    int graph_variant = 0;
    int graph_form = 0;
    int graph_wait_give_up = 0;
    double graph_eta = 0.0;
    int graph_psi = 0;
};

void pct_plan_destroy(pct_plan* pl);
/* (re)build the plan of ctx for hist with blocks of B and the given MLP mode */
int pct_plan_build(pct_ctx* ctx, const pct_hist_soa* hist, int B, int mlp_mode);
bool pct_plan_ready(const pct_ctx* ctx, const pct_hist_soa* hist, int B, int mlp_mode);
/* x (row order) into the plan's tiled image on stream s, *xt its address;
 * and the tiled image back into x */
int pct_plan_image_in(pct_ctx* ctx, hipStream_t s, const float* x, float** xt);
int pct_plan_image_out(pct_ctx* ctx, hipStream_t s, float* x);
/* pct_mlp.hip: the verification pass of the FP64 walk with one lane per
 * history */
int pct_mlp_verify_order(pct_ctx* ctx, const pct_hist_soa* hist, int64_t count, uint16_t* lane_of);
int pct_mlp_step_order(pct_ctx* ctx, const pct_hist_soa* hist, int64_t count, uint32_t* keys, uint32_t* vals,
                       uint32_t* order);
int pct_mlp_verify_ordered(pct_ctx* ctx, hipStream_t stream, const pct_hist_soa* hist, const uint32_t* order,
                           const uint32_t* hash_expect, uint32_t* work);
int pct_mlp_verify_lanes(pct_ctx* ctx, hipStream_t stream, const pct_hist_soa* hist, int64_t start, int64_t count,
                         const uint16_t* lane_of, const uint32_t* hash_expect);

#endif
