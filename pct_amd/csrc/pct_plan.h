/*
 * pct_plan.h -- the per-scan DROP plan (see pct_plan.hip).
 */
#ifndef PCT_PLAN_H
#define PCT_PLAN_H

#include <vector>
#include "pct_internal.h"

struct pct_plan {
    bool valid = false;
    pct_hist_soa key{};            /* the histories the plan was built from */
    int64_t n = 0;                 /* histories                             */
    int B = 0;                     /* DROP block size                       */
    int64_t nb = 0;                /* blocks                                */
    int64_t E = 0;                 /* path entries (sum of list lengths)    */
    int64_t S = 0;                 /* segments = sum over blocks of |V_B|   */
    pct_buffer hash;               /* u32 per history: hash of its list     */
    pct_buffer ex_off;             /* int64 per history+1: exact offsets    */
    pct_buffer blk_entry;          /* int64 per block+1: first entry        */
    pct_buffer blk_seg;            /* int64 per block+1: first segment      */
    pct_buffer seg_voxel;          /* u32 per segment                       */
    pct_buffer seg_start;          /* u32 per segment, block-relative entry */
    pct_buffer csc;                /* u16 per entry: block-local history    */
    pct_buffer delta;              /* f64 per history of one block          */
    std::vector<int64_t> blk_entry_h, blk_seg_h;
    hipGraphExec_t graph = nullptr;
    float* graph_x = nullptr;
    double graph_lambda = 0.0;
    bool graph_f32 = false;
    int mlp_mode = 0;              /* MLP mode the lists were built with    */
};

void pct_plan_destroy(pct_plan* pl);

#endif
