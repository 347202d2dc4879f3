/*
 * pct_net.h -- C ABI of the learned-path form of the MLP.
 *
 * The reference computes every proton's most likely path with the FP64 Schulte
 * MLP (find_MLP_path_GPU, .cu:5279-5501, or its tabulated twin .cu:5502-5656).
 * This form replaces that walk for most histories by a cheap FP32 surrogate: a
 * small network (12 -> 64 -> 64 -> 8, tanh, BF16 weights and activations with
 * FP32 accumulation on the MFMA units) predicts, from the entry/exit position,
 * angles, u-span and WEPL of a history, the offsets of the exact MLP from the
 * cubic Hermite path at the four interior knots sigma = 1/5 .. 4/5 of the
 * normalised depth, for t and for v. The path is then stepped in FP32 at the
 * reference's u_1 grid (.cu:5376-5489) as Hermite + clamped cubic spline of the
 * offsets (tables in pct_amd/csrc/pct_net_basis.h) and voxelised like the
 * precision form. A history falls back to the exact FP64 walk when its inputs
 * lie outside the range the network was trained on, or when its cell of
 * (u-span, path-bending magnitude) failed the calibration check of
 * pct_net_calibrate. These entry points have no counterpart in the reference.
 *
 * Mode bit: pct_mlp_paths / pct_drop_prepare / pct_drop_iteration /
 * pct_batch_* accept PCT_MLP_SURROGATE as MLP mode (after pct_net_set).
 *
 * Conventions are those of pct.h: device pointers, calls ordered on the
 * context's stream, int status.
 */
#ifndef PCT_NET_H
#define PCT_NET_H

#include <stdint.h>
#include "pct.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PCT_MLP_SURROGATE 2          /* MLP mode: learned path + exact fallback */

/* network shape and parameter layout (FP32 master copy, row-major):
 *   W1[12][64] at 0, b1[64] at 768, W2[64][64] at 832, b2[64] at 4928,
 *   W3[64][8] at 4992, b3[8] at 5504; hidden activations tanh, output linear */
#define PCT_NET_IN      12
#define PCT_NET_HIDDEN  64
#define PCT_NET_OUT     8
#define PCT_NET_OFF_W1  0
#define PCT_NET_OFF_B1  768
#define PCT_NET_OFF_W2  832
#define PCT_NET_OFF_B2  4928
#define PCT_NET_OFF_W3  4992
#define PCT_NET_OFF_B3  5504
#define PCT_NET_PARAMS  5512

/* Inputs (FP64 arithmetic on the MLP's own per-history quantities, after the
 * u_in > u_out flip of .cu:5321-5334, then rounded to float):
 *   0 (t_out - t_in) * 2      1 (v_out - v_in) * 2      [cm]
 *   2 dtheta_t * 20           3 dtheta_v * 20           [rad] exit - entry angle
 *   4 u_2 * 0.125 - 1.5       5 wepl * 0.125 - 1.5      [cm]  u-span, WEPL
 *   6 t_in * 0.1              7 t_out * 0.1             [cm]
 *   8 v_in * 0.25             9 v_out * 0.25            [cm]
 *  10 xz_entry_angle * 10    11 xz_exit_angle * 10      [rad]
 * Outputs k = 0..3: (t_MLP - t_Hermite) at sigma = (k+1)/5, k = 4..7 the same
 * for v, in units of PCT_NET_OUT_SCALE cm. */
#define PCT_NET_OUT_SCALE 0.01

/* Calibration cells: cell = 8 * min(7, (int)(u_2 / 3 cm)) + m, with m the
 * number of thresholds 0.02 * 2^k cm (k = 0..6) not above the bending
 * magnitude max(|t_out - t_in|, |v_out - v_in|, u_2 |dtheta_t|, u_2 |dtheta_v|). */
#define PCT_NET_CELLS 64

/* Training data: for histories [start, start+count) of `hist` (after
 * pct_mlp_endpoints), features[count][12] and, if targets is non-NULL, the
 * exact FP64 MLP offsets targets[count][8] (the MLP evaluated at the knot
 * depths u_1 = sigma u_2 with the expressions of .cu:5376-5428). */
int pct_net_features(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count,
                     float* features, float* targets);

/* Mean-squared-error gradient of the network at params over n samples
 * (loss = sum of squared output errors / (8 n)), computed by the training
 * kernels (BF16 MFMA, FP32 accumulation, fixed reduction order): grad[5512]
 * and, if loss is a host pointer, the loss. Deterministic. */
int pct_net_gradient(pct_ctx* ctx, const float* params, const float* features, const float* targets,
                     int64_t n, float* grad, double* loss);

/* Adam training on the device (the same kernels): `epochs` passes over the n
 * samples in minibatches of `batch` (a multiple of 64), each epoch visiting
 * the samples in an affine permutation (i * a + b) mod n seeded by `seed`;
 * learning rate cosine-annealed from lr to lr_min over all steps. params (device,
 * FP32 master copy) is updated in place; epoch_loss (host, may be NULL)
 * receives the mean training loss of every epoch. */
typedef struct pct_net_train_cfg {
    int32_t epochs;
    int32_t batch;
    float lr, lr_min;
    float beta1, beta2, eps;
    uint64_t seed;
} pct_net_train_cfg;
int pct_net_train(pct_ctx* ctx, float* params, const float* features, const float* targets, int64_t n,
                  const pct_net_train_cfg* cfg, double* epoch_loss);

/* Installs a network for PCT_MLP_SURROGATE (params: device, FP32 layout
 * above; in_lo/in_hi: host, the per-input range a history must lie in to use
 * the network). Clears the calibration (every cell allowed). */
int pct_net_set(pct_ctx* ctx, const float* params, const float* in_lo, const float* in_hi);
/* Same network with zero offsets: the Hermite path alone (ablation). */
int pct_net_set_hermite(pct_ctx* ctx, const float* in_lo, const float* in_hi);

/* Evaluates the installed network on n feature rows (device) into out[n][8]
 * with the inference kernel of the surrogate walk. */
int pct_net_forward(pct_ctx* ctx, const float* features, int64_t n, float* out);

/* Calibration check on every `stride`-th history of [start, start+count):
 * both walks run; per sampled history the largest |t| and |v| deviation of the
 * surrogate path from the exact MLP over all steps and whether the two voxel
 * lists are equal. Cells are 8 bins of u2 x 8 bins of the bending magnitude.
 * With rule PCT_NET_RULE_DEVIATION a cell whose 99th-percentile deviation
 * exceeds `threshold` cm goes to the exact walk from then on; with
 * PCT_NET_RULE_LISTS (the north star's rule: fall back wherever the voxel
 * list disagrees) a cell whose share of sampled histories with a voxel list
 * different from the exact walk's exceeds `threshold` (0: any disagreement)
 * does. Either way a cell with fewer than `min_samples` samples falls back.
 * threshold < 0 only measures (against the cell flags in force). Results are
 * in *out (host). */
#define PCT_NET_RULE_DEVIATION 0
#define PCT_NET_RULE_LISTS     1
typedef struct pct_net_calib {
    int64_t samples;             /* histories checked                          */
    int64_t list_mismatch;       /* of them, voxel list != exact voxel list     */
    int64_t in_range;            /* of them, inputs inside the trained range    */
    int64_t surrogate;           /* of them, walked on the surrogate path with
                                    the cell flags in force after the call     */
    int64_t surrogate_mismatch;  /* of those, voxel list != exact voxel list    */
    double  dev_median, dev_p99, dev_max;   /* per-history max deviation [cm]  */
    int32_t cells_fallback;      /* cells sent to the exact walk               */
    int32_t cell_samples[PCT_NET_CELLS];
    float   cell_p99[PCT_NET_CELLS];
    uint8_t cell_fallback[PCT_NET_CELLS];
    float   cell_list_mismatch[PCT_NET_CELLS];  /* share of the cell's samples
                                                   whose lists disagree     */
} pct_net_calib;
int pct_net_calibrate_rule(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, int64_t stride,
                           int rule, float threshold, int32_t min_samples, pct_net_calib* out);
/* the deviation rule with tol as its threshold (tol <= 0: measure only) */
int pct_net_calibrate(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, int64_t stride,
                      float tol, int32_t min_samples, pct_net_calib* out);

/* Histories walked by the exact fallback (and in total) in the MLP passes of
 * PCT_MLP_SURROGATE since the last call; the counters are then reset. */
int pct_net_fallback_stats(pct_ctx* ctx, int64_t* fallbacks, int64_t* histories);

#ifdef __cplusplus
}
#endif

#endif /* PCT_NET_H */
