/*
 * pct.h -- C ABI of the MI355X proton-CT reconstruction engine.
 *
 * The reference (lynch829/pCT) has no plugin or FFI layer: its "interface" is a
 * set of host stage functions that read and write global device arrays and
 * launch CUDA kernels with positional SoA pointers
 * (include/pCT_functions.h:35-70, 180-237). This header is the drop-in
 * replacement for those stage/kernel boundaries on the DROP hot path: every
 * entry point below names the reference function it replaces.
 *
 * Conventions
 *   - All array pointers are DEVICE pointers owned by the caller unless stated.
 *   - Every call is ordered on the context's HIP stream and returns an int
 *     status: 0 = ok, negative = error (PCT_E_*); pct_last_error() gives text.
 *   - A context is not thread-safe; use one context per host thread.
 *   - Units are the reference's: cm, radians, gantry angle in integer degrees.
 */
#ifndef PCT_H
#define PCT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Parameters: the reference's compile-time constants as one runtime struct.  */
/* Field names follow include/pCT_data.h and include/pCT_config.h.            */
/* ------------------------------------------------------------------------- */
typedef struct pct_params {
    int32_t columns, rows, slices;           /* COLUMNS, ROWS, SLICES (data.h:264-266)          */
    int32_t t_bins, v_bins, angular_bins;    /* T_BINS, V_BINS, ANGULAR_BINS (data.h:238-243)   */
    double  recon_cyl_radius;                /* RECON_CYL_RADIUS [cm] (data.h:255)              */
    double  recon_cyl_height;                /* RECON_CYL_HEIGHT [cm] (data.h:257)              */
    double  slice_thickness;                 /* SLICE_THICKNESS [cm] (data.h:271)               */
    double  t_bin_size, v_bin_size;          /* T_BIN_SIZE, V_BIN_SIZE [cm] (data.h:237,239)     */
    double  angular_bin_size;                /* ANGULAR_BIN_SIZE [deg] (data.h:241)             */
    double  ssd_t_size, ssd_v_size;          /* SSD_T_SIZE, SSD_V_SIZE [cm] (data.h:227-228)     */
    double  source_radius;                   /* SOURCE_RADIUS [cm] (data.h:222)                 */
    double  gantry_angle_interval;           /* GANTRY_ANGLE_INTERVAL [deg] (data.h:223)        */
    double  sigmas_to_keep;                  /* SIGMAS_TO_KEEP (data.h:244)                     */
    double  fbp_threshold;                   /* FBP_THRESHOLD (data.h:248)                      */
    double  rsp_air;                         /* RSP_AIR (data.h:297)                            */
    double  sc_threshold, msc_threshold;     /* SC_THRESHOLD, MSC_THRESHOLD (data.h:282-283)    */
    double  hull_filter_threshold;           /* HULL_FILTER_THRESHOLD (data.h:287)              */
    int32_t hull_filter_radius;              /* HULL_FILTER_RADIUS (data.h:288)                 */
    int32_t msc_diff_thresh;                 /* MSC_DIFF_THRESH (data.h:281)                    */
    double  e_0, x_0;                        /* E_0 [MeV/c], X_0 [cm] (data.h:294-295)           */
    double  a[6];                            /* A_0..A_5 (data.h:306-311), folded to doubles    */
    int32_t max_intersections;               /* MAX_INTERSECTIONS (data.h:299)                  */
    int32_t drop_block_size;                 /* DROP_BLOCK_SIZE / BLOCK_SIZE_RECON (data.h:377-378) */
    float   lambda;                          /* LAMBDA (data.h:374)                             */
    int32_t iterations;                      /* ITERATIONS (cfg.h:5)                            */
    double  eta;                             /* ETA of the robust DROP variants (data.h:369,
                                                command line .cu:329-335)                      */
    int32_t psi_sign;                        /* PSI_SIGN (data.h:371)                           */
    double  sm_lower_threshold;              /* SM_LOWER_THRESHOLD [cm] (data.h:284)            */
    double  sm_scale_threshold;              /* SM_SCALE_THRESHOLD (data.h:286)                 */
} pct_params;

/* Raw tracker data of one read chunk (data.h:497-508): 4 tracker planes in the
 * gantry frame (u along the beam, t lateral, v vertical) + WEPL. */
typedef struct pct_raw_soa {
    const int32_t* gantry_angle;                          /* [deg]   */
    const float *t_in_1, *t_in_2, *t_out_1, *t_out_2;
    const float *u_in_1, *u_in_2, *u_out_1, *u_out_2;
    const float *v_in_1, *v_in_2, *v_out_1, *v_out_2;
    const float* wepl;
    int64_t n;
} pct_raw_soa;

/* Per-history record after intersection (data.h:577-590, same field set and
 * order as the reference's *_vector SoA). */
typedef struct pct_hist_soa {
    float *x_entry, *y_entry, *z_entry;
    float *x_exit, *y_exit, *z_exit;
    float *xy_entry_angle, *xz_entry_angle;
    float *xy_exit_angle, *xz_exit_angle;
    float* wepl;
    int32_t* bin_num;
    uint32_t* first_voxel;                               /* first_MLP_voxel (data.h:590) */
    int64_t n;
} pct_hist_soa;

/* Status codes */
#define PCT_OK              0
#define PCT_E_INVALID     (-1)   /* bad argument / size                         */
#define PCT_E_HIP         (-2)   /* HIP runtime error                          */
#define PCT_E_OVERFLOW    (-3)   /* a path longer than max_intersections        */
#define PCT_E_NOMEM       (-4)   /* device allocation failed                    */

/* MLP evaluation modes for the path and DROP entry points */
#define PCT_MLP_FP64      0      /* exact FP64 Schulte MLP (.cu:5279-5501)      */
#define PCT_MLP_TABLE     1      /* precision form, after the tabulated MLP
                                    (.cu:5502-5656): FP64 per-depth sigma
                                    tables and 2x2 inversions, FP32 stepping  */
#define PCT_MLP_MODE_MASK 0xff
/* flags of pct_drop_iteration's mode, OR-ed with the MLP mode */
#define PCT_DROP_CACHED    0x100 /* reuse the plan's voxel lists instead of
                                    recomputing the MLP in this iteration      */
#define PCT_DROP_ATOMIC    0x200 /* reference structure: atomic scatter and a
                                    volume scan per block, no plan (also what
                                    the default runs for a scan whose plan
                                    does not fit the device)                   */
#define PCT_DROP_NO_GRAPH  0x400 /* launch the block loop eagerly (no hipGraph) */
#define PCT_DROP_F32       0x800 /* precision form of the update: a_i.x, the
                                    residual and the per-voxel sums in FP32,
                                    sums Kahan-compensated                      */
/* The reference's research variants of the FP64 update (CPU functions there),
 * with psi_v = ETA and PSI_SIGN applied to (1 - x_v) at the block's start: */
#define PCT_DROP_ROBUST    0x1000 /* DROP_blocks_red + DROP_update (.cu:8005-8055):
                                    every path entry adds
                                    L lambda r_i / (a_i.a_i + psi_v)          */
#define PCT_DROP_ROBUST2   0x2000 /* DROP_blocks_robust2 + DROP_update_robust2
                                    (.cu:7871-7901): the plain update scaled per
                                    voxel by N_v / (N_v + psi_v), N_v the sum of
                                    the block's L_i^2 through v                */
#define PCT_DROP_PERSISTENT 0x4000 /* pct_batch_iteration only: every block step
                                    of the iteration in one persistent launch
                                    (grid barriers between the delta and update
                                    phases); the MLP check runs before it, not
                                    beside it. Results are identical to the
                                    default two launches per step. A barrier
                                    wait is bounded: one that gives up returns
                                    PCT_E_HIP with every image unchanged.      */

typedef struct pct_ctx pct_ctx;

/* Context: owns a HIP stream (or borrows `stream` if non-NULL) and the
 * workspaces of the DROP loop. Replaces the global device state of
 * data.h:512-573 and cudaSetDevice(0) in image_reconstruction_GPU (.cu:6543). */
pct_ctx*    pct_create(const pct_params* params, void* stream);
void        pct_destroy(pct_ctx* ctx);
int         pct_sync(pct_ctx* ctx);
const char* pct_last_error(const pct_ctx* ctx);
/* Fill `out` with the reference defaults (pCT_data.h / pCT_config.h). */
void        pct_default_params(pct_params* out);

/* (a) Recon-volume intersection: replaces recon_volume_intersections_GPU
 * (.cu:2169-2300) and calculate_intercepts (.cu:2301-2387). Writes the entry
 * and exit points/angles of every history and missed[i] = 1 when the path
 * misses the reconstruction cylinder. out->n must equal raw->n; out->wepl is
 * copied from raw->wepl, out->bin_num / first_voxel are left untouched. */
int pct_recon_volume_intersections(pct_ctx* ctx, const pct_raw_soa* raw,
                                   const pct_hist_soa* out, uint8_t* missed);

/* Per-bin statistics of the (t, v, angle) binning (data.h:522-531 arrays).
 * Every array has t_bins * v_bins * angular_bins entries, bin
 * t + angle * T_BINS + v * T_BINS * ANGULAR_BINS. */
typedef struct pct_bin_stats {
    int32_t* counts;                                   /* bin_counts          */
    float *mean_wepl, *mean_rel_ut, *mean_rel_uv;      /* mean_* arrays       */
    float *stddev_wepl, *stddev_rel_ut, *stddev_rel_uv;/* stddev_* arrays     */
} pct_bin_stats;

/* (a) Binning: replaces binning_GPU (.cu:2507-2578). Writes hist->bin_num
 * (0 for a history whose bin is out of range, as the reference's calloc'd
 * array leaves it) and in_range[i]; orders the non-missed histories by bin
 * (stably) for the two calls below, keeping a snapshot of what they read
 * (WEPL, exit-minus-entry angles, in_range): the histories' WEPL and angles
 * must not change before those calls. With hist->n == 0 the per-history
 * arrays (and missed, in_range) may be null. The same holds for every stage
 * entry point of this header whose history count or range is empty, and
 * pct_drop_iteration returns at once for an empty set; pct_drop_prepare,
 * pct_drop_stats and the batched form need at least one history. */
int pct_binning(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed, uint8_t* in_range);
/* (a) Per-bin count and means of WEPL and relative angles over in-range
 * non-missed histories, squared deviations over all non-missed ones, standard
 * deviations: replaces the binning sums + calculate_means_GPU (.cu:2618) +
 * sum_squared_deviations_GPU (.cu:2671) + calculate_standard_deviations_GPU
 * (.cu:2718). Sums are formed in storage order (deterministic). */
int pct_bin_statistics(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed, const uint8_t* in_range,
                       const pct_bin_stats* stats);
/* (a) 3-sigma cuts + sinogram: replaces statistical_cuts_GPU (.cu:2835-2869)
 * and construct_sinogram_GPU (.cu:2912-2918). failed[i] = 1 for histories
 * that fail a cut or are missed; stats->counts becomes the survivors' count
 * and sinogram[bin] their mean WEPL. */
int pct_statistical_cuts(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed, const pct_bin_stats* stats,
                         uint8_t* failed, float* sinogram);

/* (c) Cylinder hull: replaces initialize_hull_GPU (.cu:3316-3326). The hull is
 * bit-packed, one bit per voxel, voxel v at word v>>5, bit v&31;
 * ceil(V/32) words. */
int pct_hull_init(pct_ctx* ctx, uint32_t* hull_bits);

/* (c) Space carving: replaces SC_GPU (.cu:3338-3444). Clears the bit of every
 * voxel on the straight entry->exit walk (take_3D/2D_step_GPU, .cu:8754-8867)
 * of each non-missed history with WEPL <= SC_THRESHOLD. */
int pct_space_carve(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed, uint32_t* hull_bits);
/* (c) Modified space carving: replaces initialize_hull_GPU<int> on the MSC
 * counts (.cu:3303) and MSC_GPU (.cu:3456-3558): counts[v] starts at 1 inside
 * the cylinder, 0 outside, and is incremented for every walk voxel of each
 * non-missed history with WEPL <= MSC_THRESHOLD. */
int pct_msc_init(pct_ctx* ctx, int32_t* counts);
int pct_msc_count(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed, int32_t* counts);
/* (c) MSC edge detection: replaces MSC_edge_detection_GPU (.cu:3569-3596),
 * out of place (out != counts): out[v] = 0 where the in-slice 3x3 count drop
 * exceeds MSC_DIFF_THRESH or the voxel is on the cylinder rim, else 1. */
int pct_msc_edge(pct_ctx* ctx, const int32_t* counts, int32_t* out);
/* (c) Space modelling: replaces SM_GPU (.cu:3598-3710): counts start as
 * pct_msc_init's cylinder (initialize_hull on the SM counts, .cu:3301) and are
 * incremented for every walk voxel of each non-missed history with WEPL >=
 * SM_LOWER_THRESHOLD (the protons that crossed the object). */
int pct_sm_count(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed, int32_t* counts);
/* (c) SM edge detection: replaces SM_edge_detection (.cu:3711-3773) with
 * carve_differences (.cu:3274-3292) and SM_edge_detection_GPU (.cu:3774-3789),
 * out of place (out != counts). Per slice, the difference of a pixel is its
 * largest in-slice 3x3 count drop (0 on the image border); the slice's
 * threshold is the count of the first pixel, in pixel order, with the
 * largest difference (0 when no difference is positive). out[v] = 1 where
 * counts[v] > SM_SCALE_THRESHOLD x threshold and the voxel is inside the
 * cylinder rim, else 0. thresholds (slices values; may be null) receives the
 * per-slice thresholds. */
int pct_sm_edge(pct_ctx* ctx, const int32_t* counts, int32_t* out, int32_t* thresholds);
/* (c) Hull selection + averaging filter: replaces hull_selection (.cu:3954)
 * and averaging_filter_GPU (.cu:4007-4029): hull = counts != 0, then a voxel
 * is kept when its in-slice (2r+1)^2 neighbourhood sum exceeds
 * HULL_FILTER_THRESHOLD times the neighbourhood size. Bit-packed output. */
int pct_hull_select(pct_ctx* ctx, const int32_t* counts, uint32_t* hull_bits);

/* (b) FBP: replaces filter_GPU (.cu:3036-3069, Shepp-Logan), backprojection_GPU
 * (.cu:3155-3230) and FBP_image_2_hull_GPU (.cu:3250-3261). Sinograms have
 * t_bins * v_bins * angular_bins entries, images columns * rows * slices. */
int pct_fbp_filter(pct_ctx* ctx, const float* sinogram, float* filtered);
int pct_backproject(pct_ctx* ctx, const float* filtered, float* image);
int pct_fbp_threshold(pct_ctx* ctx, const float* image, uint32_t* hull_bits);
/* (e) Initial iterate x0 = FBP image * hull (X_K0 = HYBRID,
 * define_initial_iterate .cu:7191-7224). */
int pct_initial_iterate(pct_ctx* ctx, const float* fbp_image, const uint32_t* hull_bits, float* x0);

/* (d) MLP endpoints: replaces collect_MLP_endpoints_GPU (.cu:5729-5770) with
 * find_MLP_endpoints_GPU (.cu:4997-5277). For histories that enter and exit
 * the hull, overwrites x/y/z entry/exit with the hull surface points and sets
 * first_voxel; intersected[i] = 1 for those, 0 otherwise. */
int pct_mlp_endpoints(pct_ctx* ctx, const pct_hist_soa* hist,
                      const uint32_t* hull_bits, uint8_t* intersected);

/* Stable, order-preserving compaction of every field of `src` whose keep[i]
 * is nonzero into `dst` (dst arrays sized >= src->n). Replaces the host
 * push_back loops of .cu:2443-2477, 2777-2809, 6277-6298. The number of kept
 * histories is written to *n_kept (host pointer) and to dst->n is NOT
 * written (the struct is const); the caller sets it. */
int pct_compact(pct_ctx* ctx, const pct_hist_soa* src, const uint8_t* keep,
                const pct_hist_soa* dst, int64_t* n_kept);

/* (d) MLP voxel lists for histories [start, start+count): replaces the path
 * part of find_MLP_path_GPU (.cu:5279-5501). Path of history h is written to
 * voxels[offsets[h-start] ...] (offsets is a device array of count+1 int64
 * prefix offsets from pct_mlp_capacity), its length to lengths[h-start].
 * Returns PCT_E_OVERFLOW if a path exceeds max_intersections. */
int pct_mlp_capacity(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start,
                     int64_t count, int64_t* offsets, int64_t* total);
int pct_mlp_paths(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start,
                  int64_t count, const int64_t* offsets, uint32_t* voxels,
                  uint32_t* lengths, int mode);

/* (e) One DROP block: replaces block_update_GPU (.cu:5774-5841): for every
 * history of [start, start+count) computes its MLP path, a_i.x, the residual
 * and scatters count[v] += 1, x_update[v] += delta_i. */
int pct_drop_block_update(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start,
                          int64_t count, const float* x, float* x_update,
                          uint32_t* counts, float lambda, int mode);

/* Allocation-free form of pct_drop_block_update (no counterpart in the
 * reference): same arguments, same kernels and results, for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc. The block's MLP lists live
 * in a fixed device array of the library (about 17 MB per GPU, held until the
 * process ends) instead of context workspaces grown on demand, so a call
 * takes at most PCT_NOALLOC_MAX_HISTORIES histories and needs the context's
 * max_intersections <= PCT_NOALLOC_MAX_SLOTS (PCT_E_INVALID otherwise). The
 * call allocates nothing (the context's tables and its pinned status word
 * are made by pct_create). Every call on one GPU shares that array: calls
 * that overlap on one GPU (other streams or threads) race. */
#define PCT_NOALLOC_MAX_HISTORIES 8192
#define PCT_NOALLOC_MAX_SLOTS 512
int pct_drop_block_update_noalloc(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start,
                                  int64_t count, const float* x, float* x_update,
                                  uint32_t* counts, float lambda, int mode);

/* (e) Image update: replaces image_update_GPU (.cu:5900-5938):
 * x[v] += x_update[v] / counts[v] where counts[v] > 0, then clears both. */
int pct_drop_image_update(pct_ctx* ctx, float* x, float* x_update, uint32_t* counts);

/* (e) Effective chord length of every history: replaces
 * EffectiveChordLength_GPU (.cu:4974-4995) with the GPU driver's double
 * average of the float angles (.cu:5315) when float_average == 0, and
 * EffectiveChordLength (.cu:7380-7400) with the CPU research loop's float
 * average (.cu:7539) when float_average != 0. out: hist->n doubles. */
int pct_chord_lengths(pct_ctx* ctx, const pct_hist_soa* hist, int float_average, double* out);

/* Update rules of pct_drop_lists_iteration (each followed by DROP_update,
 * .cu:8044-8055, after every full block) */
#define PCT_LISTS_DROP          0   /* DROP_blocks (.cu:8023-8043)                       */
#define PCT_LISTS_DROP_RED      1   /* DROP_blocks_red (.cu:7998-8022): psi = (1-x_v) ETA PSI_SIGN */
#define PCT_LISTS_ERROR_RED     2   /* DROP_blocks_error_red (.cu:7816-7843): chord L + e ETA, psi */
#define PCT_LISTS_ERROR         3   /* DROP_blocks_error (.cu:7844-7870): chord L + e ETA L        */
#define PCT_LISTS_ERROR_RED_SP  4   /* DROP_blocks_error_red_sp (.cu:7768-7815): chord 1.8 cm at
                                       every 5th intersection, psi                                 */
#define PCT_LISTS_REUSE    0x100    /* or-ed to the rule: the lists (offsets, voxels, their
                                       contents, block_size) are those of the previous call on
                                       this context, so their sorted order is reused; with other
                                       arrays or sizes the call sorts as usual                     */

/* (e) One iteration of the CPU research reconstruction loop over voxel
 * lists given by the caller: replaces the iteration body of
 * image_reconstruction (.cu:7513-7565), which reads the lists from its
 * MLP_paths file and the per-intersection path errors from its
 * MLP_path_error file. List k is voxels[offsets[k] .. offsets[k+1]) (device
 * arrays; offsets[0] = 0), with WEPL wepl[k], chord length chord[k] and, for
 * the error variants, errors[offsets[k] + j] for its j-th intersection
 * (may be null otherwise). With no intersection at all (offsets[n] = 0)
 * voxels and errors may be null. Lists of up to 65535 voxels (the MLP's
 * lists are at most MAX_INTERSECTIONS long), blocks of 1..65535 lists. After every full block of block_size lists
 * ((n+1) % BLOCK_SIZE_RECON == 0, .cu:7554) the update runs; x_update
 * (double) and counts (uint32), V values each, are the caller's and persist
 * between calls like the reference's x_update_h / block_counts_h: zero them
 * before the first iteration; a partial last block leaves its sums there for
 * the next call. ETA and PSI_SIGN come from the parameters. Results equal
 * the reference's loop bit for bit, including its error variants' weighting
 * of x by list position (x_k[i]). Voxel indices, or for the error variants
 * list positions, outside the image are skipped and reported as
 * PCT_E_INVALID. */
int pct_drop_lists_iteration(pct_ctx* ctx, int64_t n, const int64_t* offsets, const uint32_t* voxels,
                             const float* errors, const double* chord, const float* wepl, float* x,
                             double* x_update, uint32_t* counts, int block_size, float lambda, int variant);

/* (e) Per-scan DROP plan: computes every MLP voxel list of `hist` and the
 * per-block voxel->history transpose used by pct_drop_iteration. Replaces the
 * set-up part of image_reconstruction_GPU (.cu:6764-6782). Called implicitly
 * by pct_drop_iteration when the histories or the block size change. */
int pct_drop_prepare(pct_ctx* ctx, const pct_hist_soa* hist, int block_size, int mode);

/* (e) One whole DROP iteration over all histories in storage order, blocks of
 * `block_size`: replaces the body of the iteration loop of
 * image_reconstruction_GPU (.cu:6795-6857). Equivalent to the sequence of
 * pct_drop_block_update + pct_drop_image_update per block. By default the
 * MLP lists are recomputed (and checked against the plan) in every iteration,
 * as the reference does; see the PCT_DROP_* flags.
 * The plan is keyed on the history arrays' addresses, the count, the block
 * size and the MLP mode, not on their contents. By default a recomputed list
 * that differs from the plan's (the histories changed in place) is detected
 * by a 32-bit hash per list after the block loop ran: the call then restores
 * x to its value before the call, marks the plan stale and returns
 * PCT_E_INVALID; the next call rebuilds the plan. With PCT_DROP_CACHED the
 * lists are not recomputed, so the caller must call pct_drop_invalidate (or
 * pct_drop_prepare) after changing histories in place. A scan whose plan does
 * not fit the device (pct_drop_plan_choice) runs the reference's structure
 * (scatter + image update per block) over chunks of whole blocks instead;
 * the research variants need the plan and return PCT_E_NOMEM then. */
int pct_drop_iteration(pct_ctx* ctx, const pct_hist_soa* hist, float* x,
                       int block_size, float lambda, int mode);
/* Marks the context's DROP plan stale (no counterpart in the reference, which
 * recomputes every list in every block): the next pct_drop_iteration rebuilds
 * it. */
int pct_drop_invalidate(pct_ctx* ctx);

/* Measurement (counterpart of the reference's per-iteration clock() and
 * cudaEvent timing, .cu:6773-6849): when profiling is on, pct_drop_iteration
 * records HIP events and pct_last_timings returns, for the last iteration, in
 * ms: [0] from its start until the block loop may start (the MLP lists of the
 * first chunk of blocks), [1] from its start until the MLP lists of all
 * histories are done, [2] from the start of the block loop to the end of the
 * iteration. The MLP and the block loop overlap (chunk by chunk), so the
 * iteration takes [0] + [2]; without a MLP pass [0] = [1] = 0. */
int pct_set_profiling(pct_ctx* ctx, int on);
int pct_last_timings(pct_ctx* ctx, double* ms3);
/* Work statistics of one DROP iteration (x is not modified): *sum_path =
 * sum over histories of the path length n_i, *sum_touched = sum over blocks of
 * the number of distinct voxels the block touches (|V_B|). */
int pct_drop_stats(pct_ctx* ctx, const pct_hist_soa* hist, int block_size, int64_t* sum_path, int64_t* sum_touched);

/* Execution knobs of pct_drop_iteration (no counterpart in the reference,
 * whose launch shapes are compile-time constants, data.h:377-378). They change
 * scheduling only, never results. value <= 0 restores the default. */
#define PCT_TUNE_MLP_WAVES     1  /* resident MLP wavefronts per SIMD while the
                                     block loop runs beside it (default: the
                                     exact walk's full occupancy, 3 for the
                                     precision form, 2 for the surrogate)      */
#define PCT_TUNE_CHUNK_FIRST   2  /* blocks in the first chunk (default: all
                                     blocks in one chunk; 16 with overlap 1)   */
#define PCT_TUNE_CHUNK_MAX     3  /* largest chunk in blocks (default: all; 256
                                     with overlap 1)                           */
#define PCT_TUNE_OVERLAP       4  /* default: the check of the recomputed MLP
                                     lists beside the block loop, which reads
                                     the plan's lists and joins the check at
                                     the end; 1: beside, the block loop's chunk
                                     c waiting for the check of chunk c;
                                     2: MLP first, then the block loop         */
#define PCT_TUNE_BLOCK_FORM    5  /* 1: one launch per block, the voxel updates
                                     of block b-1 beside the deltas of block b,
                                     which recompute the updates not stored yet;
                                     2: two launches per block (default; the
                                     research variants always);
                                     3: one launch per block, the deltas of
                                     block b waiting (bounded) for block b-1's
                                     stored updates of their voxels; a wait
                                     that gives up returns PCT_E_HIP with x
                                     unchanged. All three give the same image
                                     bit for bit.                              */
#define PCT_TUNE_PLAN_LIMIT_MB 6  /* cap, in MiB, on the device memory a DROP
                                     plan may take (default: free memory); a
                                     scan whose plan exceeds it runs in chunks
                                     of blocks (see pct_drop_plan_choice)     */
#define PCT_TUNE_WAIT_GIVE_UP  7  /* diagnostics of block form 3 and of a
                                     persistent batch (PCT_DROP_PERSISTENT,
                                     set on the batch's first context): 1 makes
                                     their waits target a value nothing writes
                                     and give up after two polls, so that the
                                     failure path (PCT_E_HIP, images unchanged)
                                     can be exercised; 0 (default) off         */
int pct_set_tuning(pct_ctx* ctx, int key, int value);

/* Capacity (no counterpart in the reference, which walks every path inside
 * its per-block kernel, .cu:5774-5841): the peak device bytes of building a
 * DROP plan for n histories with `slots` MLP capacity slots (the sum of the
 * step counts + 1, pct_mlp_capacity's total), `list_slots` slots of its
 * block-strided list store and `entries` path entries (0, 0 before the MLP
 * pass: what the build needs to start), with the one-launch form's records or
 * not; and the choice pct_drop_iteration makes before and after its MLP pass
 * -- 1: build the plan; 0: run the reference's structure (per block an atomic
 * scatter and a volume scan) over chunks of whole blocks whose MLP lists take
 * at most half the budget -- when the plan would exceed `budget` bytes (free
 * device memory, or PCT_TUNE_PLAN_LIMIT_MB), V > 2^22 or the block size >
 * 8191 (V: the plan's voxel index space, the grid padded to whole tiles of
 * 4 x 8 voxels of a slice). *need receives the bytes. Pure functions (no
 * device). */
int64_t pct_drop_plan_bytes(int64_t n, int64_t slots, int64_t list_slots, int64_t entries, int with_prev);
int     pct_drop_plan_choice(int64_t n, int64_t slots, int64_t list_slots, int64_t entries, int64_t V,
                             int block_size, int64_t budget, int with_prev, int64_t* need);

/* (e) Batched form: the DROP iterations of several independent scans (each
 * with its own context, i.e. its own grid and parameters) run together. The
 * reference reconstructs one scan per process run (main, .cu:24-253); this
 * entry point has no counterpart and serves callers with many scans. Blocks
 * are still processed in storage order within every scan; block step j runs
 * block j of every scan that has one in one pair of launches (one kernel for
 * the deltas, one for the voxel updates of all those blocks), so the
 * latency-bound chains of the scans share each launch. Each scan's MLP lists
 * are recomputed every iteration on a second stream, one launch per scan
 * beside the block steps (with PCT_TUNE_OVERLAP 1 on the first context: chunk
 * of steps by chunk of steps, ahead of the block steps that need them). The
 * batch runs on the stream of scans[0]; results are identical to
 * pct_drop_iteration per scan.
 * At most PCT_BATCH_MAX scans. */
#define PCT_BATCH_MAX 16
typedef struct pct_batch pct_batch;
pct_batch* pct_batch_create(pct_ctx* const* scans, int n_scans);
void       pct_batch_destroy(pct_batch* b);
/* builds (or keeps) each scan's plan for its histories and block size */
int        pct_batch_prepare(pct_batch* b, const pct_hist_soa* hists, const int* block_sizes, int mode);
/* one DROP iteration of every scan: xs[s] is scan s's image, lambdas[s] its
 * relaxation; mode as pct_drop_iteration (MLP mode | PCT_DROP_F32 |
 * PCT_DROP_CACHED | PCT_DROP_NO_GRAPH | PCT_DROP_PERSISTENT), with the same contract for lists that
 * differ from a plan (every such scan's image is restored). A context may
 * appear once in a batch (pct_batch_create returns NULL otherwise); a plan
 * rebuilt through its own context is picked up by the next batch call. */
int        pct_batch_iteration(pct_batch* b, const pct_hist_soa* hists, float* const* xs, const int* block_sizes,
                               const float* lambdas, int mode);

/* Diagnostics (no counterpart in the reference): evaluates one function of
 * the shared FP64/FP32 math library (pct_math.h) on the device for n inputs,
 * so host/device bit equality can be checked directly.
 * func: 0 log, 1 sin, 2 cos, 3 tan, 4 atan, 5 atan2(in[i], in2[i]),
 *       6 sinf, 7 cosf, 8 atan2f, 9 sqrt, 10 division in[i] / in2[i].
 * Single-precision functions read and write the low 32 bits as float. */
int pct_selftest_math(pct_ctx* ctx, int func, const double* in, const double* in2, double* out, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* PCT_H */
