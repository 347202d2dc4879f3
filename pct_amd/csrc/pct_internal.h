/*
 * pct_internal.h -- internal definitions of libpct (context, derived geometry,
 * launch helpers). Not part of the public C ABI (include/pct.h).
 */
#ifndef PCT_INTERNAL_H
#define PCT_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/pct.h"
#include "../../include/pct_net.h"
#include "pct_math.h"

/* Derived constants of the reference's macros (data.h:264-348), computed once
 * on the host and passed to kernels by value. */
struct pct_geo {
    int C, R, S;
    int64_t V;
    double voxel_width, voxel_height, voxel_thickness;   /* data.h:268-270 */
    double inv_width, inv_height, inv_thickness;         /* 1 / size, rounded once */
    double x_zero, y_zero, z_zero;                       /* data.h:272-274 */
    double mlp_u_step;                                   /* data.h:300     */
    double radius, height;
    double e_0, x_0;
    double A0;
    /* A_i / k folds (data.h:322-348) */
    double a0_2, a0_3, a1_2, a1_3, a1_4, a1_6, a1_12, a2_3, a2_4, a2_5, a2_12, a2_30;
    double a3_4, a3_5, a3_6, a3_20, a3_60, a4_5, a4_6, a4_7, a4_30, a4_105;
    double a5_6, a5_7, a5_8, a5_42, a5_168;
    int max_intersections;
    int64_t hull_words;                                  /* ceil(V / 32)   */
};

pct_geo pct_make_geo(const pct_params& p);

/* Per-depth FP64 table of the u_1-only MLP terms (see pct_mlp.hip). Field f of
 * step k is at tab[f * K + k]. */
enum { MLP_TAB_U1 = 0, MLP_TAB_S1I0, MLP_TAB_S1I1, MLP_TAB_S1I3, MLP_TAB_C1, MLP_TAB_C2, MLP_TAB_C3, MLP_TAB_FIELDS };

struct pct_plan;
struct pct_bin_order;
struct pct_net_frag;

struct pct_buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct pct_ctx {
    pct_params p;
    pct_geo g;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    /* second stream of the context (created on first use): the MLP lists of
     * a DROP iteration run on it ahead of the block loop */
    hipStream_t aux_stream = nullptr;
    int device = 0;
    std::string err;
    /* device-side status word written by kernels (overflow flags etc.) */
    int* d_status = nullptr;
    /* pinned host copy the status word is read back into */
    int* h_status = nullptr;
    /* MLP per-depth table */
    double* d_mlp_tab = nullptr;
    int mlp_tab_k = 0;
    /* precision form: sigma_2 scattering coefficient per remaining depth */
    double* d_scat_tab = nullptr;
    int scat_q = 0;
    /* resident workgroups per CU of the MLP paths kernel per walk, and the
     * CU count (0 = not queried) */
    int mlp_occ[3] = {0, 0, 0};
    int cus = 0;
    /* learned-path form (pct_net.hip): installed network (inference operands),
     * per-input trained range lo[12], hi[12], calibration cell flags (1 = exact
     * walk), fallback/history counters, and a float copy of the u_1 column of
     * the MLP table. net_version counts installs (0 = none) so that plans built
     * with another network are rebuilt. */
    pct_net_frag* d_net = nullptr;
    float* d_net_range = nullptr;
    uint8_t* d_net_cells = nullptr;
    unsigned long long* d_net_counts = nullptr;
    float* d_u1f = nullptr;
    int net_version = 0;
    pct_buffer net_partial, net_state, net_loss, net_calib, net_stage_buf;
    /* the surrogate walk's fallback list: u64 count, indices from byte 256 */
    pct_buffer net_fallback;
    /* execution knobs (pct_set_tuning; 0 = default) */
    int tune_mlp_waves = 0, tune_chunk_first = 0, tune_chunk_max = 0, tune_overlap = 0, tune_block_form = 0;
    int tune_plan_limit_mb = 0;
    int tune_wait_give_up = 0;
    /* workspaces, grown on demand */
    /* research list iteration: sort keys, entries, lists, per-list factors,
     * block offsets, sort storage (pct_research.hip) */
    pct_buffer research[8];
    /* the lists whose sorted order research[1], research[3] hold (reused by
     * PCT_LISTS_REUSE calls on the same lists) */
    const void* rs_offsets = nullptr;
    const void* rs_voxels = nullptr;
    int64_t rs_n = -1, rs_E = -1;
    int rs_block = 0;
    std::vector<int64_t> rs_blk_off;
    /* walk order of the SM counts: keys, indices (in/out), sort storage */
    pct_buffer walk_order[5];
    /* slices of the SM count's LDS window (-1 = not decided, 0 = none) */
    int sm_window = -1;
    pct_buffer hull_tmp, scan_tmp, steps, offsets, path_store, path_len, chord, xupd, counts;
    pct_buffer blk_offsets, blk_voxels, blk_lengths, blk_chord;
    /* the image before a recomputing DROP iteration, restored when the
     * recomputed lists turn out to differ from the plan */
    pct_buffer x_saved;
    /* phase timing of pct_drop_iteration (HIP events on `stream`) */
    bool profile = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ev_recorded = false;
    /* statistics counters (device, 2 x int64) */
    unsigned long long* d_stats = nullptr;
    /* DROP plan of the last scan (pct_plan.hip) */
    pct_plan* plan = nullptr;
    /* stable order of the histories by bin (pct_bins.hip) */
    pct_bin_order* bin_order = nullptr;
};

void pct_bin_order_destroy(pct_bin_order* o);

/* error helpers */
int pct_fail(pct_ctx* ctx, int code, const char* fmt, ...);
#define PCT_HIP_CHECK(ctx, expr)                                                        \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return pct_fail((ctx), PCT_E_HIP, "%s failed: %s (%s:%d)", #expr,           \
                            hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

int pct_ensure(pct_ctx* ctx, pct_buffer& b, size_t bytes);
/* reads and clears the device status word; returns nonzero flags */
int pct_take_status(pct_ctx* ctx, int* flags);

/* status flags */
#define PCT_STATUS_OVERFLOW   1   /* path longer than max_intersections       */
#define PCT_STATUS_CAPACITY   2   /* path longer than its precomputed capacity */
#define PCT_STATUS_TABLE      4   /* chord longer than the per-depth table     */
#define PCT_STATUS_PLAN       8   /* recomputed lists differ from the DROP plan */
#define PCT_STATUS_WAIT      16   /* a block-loop wait gave up (waiting form)   */

/* MLP modes accepted by the path and DROP entry points (the surrogate only
 * once a network is installed) */
bool pct_mlp_mode_ok(const pct_ctx* ctx, int mlp);
int pct_net_release(pct_ctx* ctx);

/* exclusive prefix sum of n int32 values into n+1 int64 offsets (device) */
int pct_exclusive_scan_i32(pct_ctx* ctx, const int32_t* in, int64_t n, int64_t* out, int64_t* total_host);

/* workgroups of a grid-stride launch over a count that may exceed what one
 * dispatch can hold (2^32 - 1 work-items) */
#define PCT_GRID_CAP 65536u

static inline unsigned pct_grid(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b > 0 ? b : 1);
}

#endif
