/*
 * oracle_bins.cpp -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
 *
 * Restatement of the reference's binning, per-bin statistics, 3-sigma cuts and
 * sinogram (src/pCT_Reconstruction_Data_Segments.cu, ".cu" below):
 *   binning_GPU                       .cu:2507-2578  (+ host compaction 2443-2477)
 *   calculate_means_GPU               .cu:2618-2628
 *   sum_squared_deviations_GPU        .cu:2671-2699
 *   calculate_standard_deviations_GPU .cu:2718-2731
 *   statistical_cuts_GPU              .cu:2835-2869  (+ host compaction 2777-2809)
 *   construct_sinogram_GPU            .cu:2912-2918
 * with the reference's float/double typing. The reference accumulates the
 * per-bin sums with float atomics in whatever order the threads run; the
 * canonical order here is storage order (a sequential pass), which the GPU
 * reproduces with a stable sort by bin.
 *
 * Quirks kept on purpose: a history whose bin falls outside the sinogram keeps
 * bin 0 (the host array is calloc'd, .cu:2395/2548) and is kept, so it enters
 * the sum of squared deviations and the cuts of bin 0 but not its mean; a bin
 * with at most one member keeps its raw sum of squares as "stddev" (.cu:2722);
 * the cut is strict for the angles and inclusive for WEPL (.cu:2857-2861).
 */
#include <stdint.h>
#include <vector>
#include "oracle_common.h"

namespace {

struct BinGeo {
    int T, V, A;
    double t_size, v_size, a_size;
    int64_t bins;
};

BinGeo make_bin_geo(const pct_params& p) {
    BinGeo b;
    b.T = p.t_bins; b.V = p.v_bins; b.A = p.angular_bins;
    b.t_size = p.t_bin_size; b.v_size = p.v_bin_size; b.a_size = p.angular_bin_size;
    b.bins = (int64_t)b.T * b.V * b.A;
    return b;
}

/* relative angle wrapped to [-pi, pi] (.cu:2556-2565) */
inline double wrap(double a) {
    if (a > PCT_PI) a -= 2 * PCT_PI;
    if (a < -PCT_PI) a += 2 * PCT_PI;
    return a;
}

}  // namespace

extern "C" {

/* A2 for all histories: bin_num (0 when out of range), in_range flag, and the
 * per-bin count and float sums of WEPL and relative angles over histories that
 * are in range and not missed, accumulated in storage order. */
int oracle_binning(const pct_params* p, const pct_hist_soa* h, const uint8_t* missed, uint8_t* in_range,
                   int32_t* counts, float* sum_wepl, float* sum_ut, float* sum_uv) {
    BinGeo b = make_bin_geo(*p);
    const double RADIANS_TO_ANGLE = 180.0 / PCT_PI, ANGLE_TO_RADIANS = PCT_PI / 180.0;
    for (int64_t i = 0; i < h->n; ++i) {
        double x_mid = (h->x_entry[i] + h->x_exit[i]) / 2;      /* float arithmetic */
        double y_mid = (h->y_entry[i] + h->y_exit[i]) / 2;
        double z_mid = (h->z_entry[i] + h->z_exit[i]) / 2;
        double path_angle = pct_atan2f(h->y_exit[i] - h->y_entry[i], h->x_exit[i] - h->x_entry[i]);
        if (path_angle < 0) path_angle += 2 * PCT_PI;
        int angle_bin = (int)((path_angle * RADIANS_TO_ANGLE / b.a_size) + 0.5) % b.A;
        double angle = angle_bin * b.a_size * ANGLE_TO_RADIANS;
        double t = y_mid * pct_cos(angle) - x_mid * pct_sin(angle);
        int t_bin = (int)((t / b.t_size) + b.T / 2);
        double v = z_mid;
        int v_bin = (int)((v / b.v_size) + b.V / 2);
        bool ok = (t_bin >= 0) && (v_bin >= 0) && (t_bin < b.T) && (v_bin < b.V);
        in_range[i] = ok;
        h->bin_num[i] = ok ? t_bin + angle_bin * b.T + v_bin * b.T * b.A : 0;
        if (ok && !missed[i]) {
            int32_t bin = h->bin_num[i];
            double rel_ut = wrap(h->xy_exit_angle[i] - h->xy_entry_angle[i]);
            double rel_uv = wrap(h->xz_exit_angle[i] - h->xz_entry_angle[i]);
            counts[bin] += 1;
            sum_wepl[bin] += h->wepl[i];
            sum_ut[bin] += (float)rel_ut;
            sum_uv[bin] += (float)rel_uv;
        }
    }
    return 0;
}

/* A3 (means), A4 (sum of squared deviations over kept = !missed histories in
 * storage order), A5 (standard deviations); counts are left unchanged. */
int oracle_bin_statistics(const pct_params* p, const pct_hist_soa* h, const uint8_t* missed, const int32_t* counts,
                          float* mean_wepl, float* mean_ut, float* mean_uv, float* sd_wepl, float* sd_ut, float* sd_uv) {
    BinGeo b = make_bin_geo(*p);
    for (int64_t bin = 0; bin < b.bins; ++bin)
        if (counts[bin] > 0) {
            mean_wepl[bin] /= counts[bin];
            mean_ut[bin] /= counts[bin];
            mean_uv[bin] /= counts[bin];
        }
    for (int64_t i = 0; i < h->n; ++i) {
        if (missed[i]) continue;
        int32_t bin = h->bin_num[i];
        double rel_ut = wrap(h->xy_exit_angle[i] - h->xy_entry_angle[i]);
        double rel_uv = wrap(h->xz_exit_angle[i] - h->xz_entry_angle[i]);
        double dw = h->wepl[i] - mean_wepl[bin];                 /* float - float */
        double du = rel_ut - mean_ut[bin];
        double dv = rel_uv - mean_uv[bin];
        sd_wepl[bin] += (float)(dw * dw);
        sd_ut[bin] += (float)(du * du);
        sd_uv[bin] += (float)(dv * dv);
    }
    for (int64_t bin = 0; bin < b.bins; ++bin)
        if (counts[bin] > 1) {
            sd_wepl[bin] = pct_sqrtf(sd_wepl[bin] / (counts[bin] - 1));
            sd_ut[bin] = pct_sqrtf(sd_ut[bin] / (counts[bin] - 1));
            sd_uv[bin] = pct_sqrtf(sd_uv[bin] / (counts[bin] - 1));
        }
    return 0;
}

/* A6 + A7: cuts of kept histories in storage order; survivors re-counted and
 * their WEPL summed into the sinogram, which is then divided by the counts.
 * failed[i] = 1 for missed histories too. counts are overwritten. */
int oracle_statistical_cuts(const pct_params* p, const pct_hist_soa* h, const uint8_t* missed, int32_t* counts,
                            const float* mean_wepl, const float* mean_ut, const float* mean_uv, const float* sd_wepl,
                            const float* sd_ut, const float* sd_uv, uint8_t* failed, float* sinogram) {
    BinGeo b = make_bin_geo(*p);
    const float sig = (float)p->sigmas_to_keep;
    for (int64_t bin = 0; bin < b.bins; ++bin) { counts[bin] = 0; sinogram[bin] = 0.0f; }
    for (int64_t i = 0; i < h->n; ++i) {
        if (missed[i]) { failed[i] = 1; continue; }
        int32_t bin = h->bin_num[i];
        double rel_ut = wrap(h->xy_exit_angle[i] - h->xy_entry_angle[i]);
        double rel_uv = wrap(h->xz_exit_angle[i] - h->xz_entry_angle[i]);
        bool pu = pct_fabs(rel_ut - mean_ut[bin]) < (double)(sig * sd_ut[bin]);
        bool pv = pct_fabs(rel_uv - mean_uv[bin]) < (double)(sig * sd_uv[bin]);
        bool pw = pct_fabsf(mean_wepl[bin] - h->wepl[i]) <= sig * sd_wepl[bin];
        failed[i] = !pu || !pv || !pw;
        if (!failed[i]) {
            counts[bin] += 1;
            sinogram[bin] += h->wepl[i];
        }
    }
    for (int64_t bin = 0; bin < b.bins; ++bin)
        if (counts[bin] > 0) sinogram[bin] /= counts[bin];
    return 0;
}

}  // extern "C"
