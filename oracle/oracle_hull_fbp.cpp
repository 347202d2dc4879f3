/*
 * oracle_hull_fbp.cpp -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
 *
 * Restatement of the reference's hull detection, FBP and initial iterate
 * (src/pCT_Reconstruction_Data_Segments.cu, ".cu" below):
 *   SC_GPU / MSC_GPU                  .cu:3338-3558 with take_3D_step_GPU /
 *                                     take_2D_step_GPU .cu:8754-8867
 *   MSC_edge_detection_GPU            .cu:3569-3596 (read from the counts,
 *                                     write to a separate output: the
 *                                     reference's in-place version races)
 *   hull_selection + averaging_filter .cu:3954-3990, 4007-4029
 *   filter_GPU (Shepp-Logan)          .cu:3036-3069
 *   backprojection_GPU                .cu:3155-3230
 *   FBP_image_2_hull_GPU              .cu:3250-3261
 *   define_initial_iterate (HYBRID)   .cu:7191-7224
 * with the reference's float/double typing. Byte masks (0/1) stand for the
 * reference's bool arrays.
 *
 * Guards added where the reference has none (no-ops on valid inputs): the
 * carving walk also ends when a voxel coordinate becomes negative or after
 * 4 (C+R+S) + 16 steps; backprojection treats a sinogram neighbour outside
 * [0, NUM_BINS) as 0.
 */
#include <math.h>
#include <stdint.h>
#include <vector>
#include "oracle_common.h"

using namespace oracle;

namespace {

inline int calc_voxel(double zero, double pos, double size) {
    double q = pct_fabs(pos - zero) / size;
    return q < 1.0e9 ? (int)q : -1;
}
inline double edge_coord(double zero, int voxel_entered, double size, int inc, int step) {
    int on_edge = (step == inc) ? voxel_entered : voxel_entered + 1;
    return zero + (inc * on_edge * size);
}
inline double dist_rem(double zero, double pos, int inc, int step, double size, int cur) {
    return pct_fabs(edge_coord(zero, cur + inc * step, size, inc, step) - pos);
}
inline double corr(double m, double x, double x0, double y0) { return m * (x - x0) + y0; }

const int X_INC = 1, Y_INC = -1, Z_INC = -1;

/* the straight voxel walk of SC/MSC from entry to exit; calls visit(voxel)
 * for every voxel before the exit voxel (.cu:3385-3444) */
template <class F>
void carve_walk(const Geo& g, float xe, float ye, float ze, float xx, float yx, float zx, F visit) {
    const int C = g.C, R = g.R, S = g.S;
    int xm = (xe <= xx) - (xe >= xx);
    int ym = (ye <= yx) - (ye >= yx);
    int zm = (ze <= zx) - (ze >= zx);
    double x = xe, y = ye, z = ze;
    int vx = calc_voxel(g.x_zero, xe, g.voxel_width);
    int vy = calc_voxel(g.y_zero, ye, g.voxel_height);
    int vz = calc_voxel(g.z_zero, ze, g.voxel_thickness);
    int64_t voxel = vx + (int64_t)vy * C + (int64_t)vz * C * R;
    double x_to_go = dist_rem(g.x_zero, x, X_INC, xm, g.voxel_width, vx);
    double y_to_go = dist_rem(g.y_zero, y, Y_INC, ym, g.voxel_height, vy);
    double z_to_go = dist_rem(g.z_zero, z, Z_INC, zm, g.voxel_thickness, vz);
    double dy_dx = (yx - ye) / (xx - xe);          /* float / float */
    double dz_dx = (zx - ze) / (xx - xe);
    double dz_dy = (zx - ze) / (yx - ye);
    double dx_dy = (xx - xe) / (yx - ye);
    double dx_dz = (xx - xe) / (zx - ze);
    double dy_dz = (yx - ye) / (zx - ze);
    int vxo = calc_voxel(g.x_zero, xx, g.voxel_width);
    int vyo = calc_voxel(g.y_zero, yx, g.voxel_height);
    int vzo = calc_voxel(g.z_zero, zx, g.voxel_thickness);
    int64_t voxel_out = vxo + (int64_t)vyo * C + (int64_t)vzo * C * R;
    auto ended = [&]() {
        return (voxel == voxel_out) || (vx >= C) || (vy >= R) || (vz >= S) || vx < 0 || vy < 0 || vz < 0;
    };
    bool end_walk = ended();
    if (!end_walk) visit(voxel);
    int guard = 4 * (C + R + S) + 16;
    while (!end_walk && guard-- > 0) {
        if (zm != 0) {
            double x_extension = pct_fabs(dz_dx) * x_to_go;
            double y_extension = pct_fabs(dz_dy) * y_to_go;
            if ((z_to_go <= x_extension) && (z_to_go <= y_extension)) {
                vz -= zm;
                z = edge_coord(g.z_zero, vz, g.voxel_thickness, Z_INC, zm);
                x = corr(dx_dz, z, ze, xe);
                y = corr(dy_dz, z, ze, ye);
                x_to_go = dist_rem(g.x_zero, x, X_INC, xm, g.voxel_width, vx);
                y_to_go = dist_rem(g.y_zero, y, Y_INC, ym, g.voxel_height, vy);
                z_to_go = g.voxel_thickness;
            } else if (x_extension <= y_extension) {
                vx += xm;
                x = edge_coord(g.x_zero, vx, g.voxel_width, X_INC, xm);
                y = corr(dy_dx, x, xe, ye);
                z = corr(dz_dx, x, xe, ze);
                x_to_go = g.voxel_width;
                y_to_go = dist_rem(g.y_zero, y, Y_INC, ym, g.voxel_height, vy);
                z_to_go = dist_rem(g.z_zero, z, Z_INC, zm, g.voxel_thickness, vz);
            } else {
                vy -= ym;
                y = edge_coord(g.y_zero, vy, g.voxel_height, Y_INC, ym);
                x = corr(dx_dy, y, ye, xe);
                z = corr(dz_dy, y, ye, ze);
                x_to_go = dist_rem(g.x_zero, x, X_INC, xm, g.voxel_width, vx);
                y_to_go = g.voxel_height;
                z_to_go = dist_rem(g.z_zero, z, Z_INC, zm, g.voxel_thickness, vz);
            }
            if (x_to_go == 0) { x_to_go = g.voxel_width; vx += xm; }
            if (y_to_go == 0) { y_to_go = g.voxel_height; vy -= ym; }
            if (z_to_go == 0) { z_to_go = g.voxel_thickness; vz -= zm; }
        } else {
            double y_extension = pct_fabs(dx_dy) * y_to_go;
            if (x_to_go <= y_extension) {
                vx += xm;
                x = edge_coord(g.x_zero, vx, g.voxel_width, X_INC, xm);
                y = corr(dy_dx, x, xe, ye);
                x_to_go = g.voxel_width;
                y_to_go = dist_rem(g.y_zero, y, Y_INC, ym, g.voxel_height, vy);
            } else {
                vy -= ym;
                y = edge_coord(g.y_zero, vy, g.voxel_height, Y_INC, ym);
                x = corr(dx_dy, y, ye, xe);
                x_to_go = dist_rem(g.x_zero, x, X_INC, xm, g.voxel_width, vx);
                y_to_go = g.voxel_height;
            }
            if (x_to_go == 0) { x_to_go = g.voxel_width; vx += xm; }
            if (y_to_go == 0) { y_to_go = g.voxel_height; vy -= ym; }
        }
        vz = vz > 0 ? vz : 0;
        voxel = vx + (int64_t)vy * C + (int64_t)vz * C * R;
        end_walk = ended();
        if (!end_walk) visit(voxel);
    }
}

}  // namespace

extern "C" {

/* C2: SC_GPU over all histories (hull: byte mask, carved in place) */
int oracle_space_carve(const pct_params* p, const pct_hist_soa* h, const uint8_t* missed, uint8_t* hull) {
    Geo g = make_geo(*p);
    for (int64_t i = 0; i < h->n; ++i) {
        if (missed[i] || !(h->wepl[i] <= p->sc_threshold)) continue;
        carve_walk(g, h->x_entry[i], h->y_entry[i], h->z_entry[i], h->x_exit[i], h->y_exit[i], h->z_exit[i],
                   [&](int64_t v) { if (v >= 0 && v < g.V) hull[v] = 0; });
    }
    return 0;
}

/* C3: MSC_GPU over all histories (counts: int, incremented) */
int oracle_msc_count(const pct_params* p, const pct_hist_soa* h, const uint8_t* missed, int32_t* counts) {
    Geo g = make_geo(*p);
    for (int64_t i = 0; i < h->n; ++i) {
        if (missed[i] || !(h->wepl[i] <= p->msc_threshold)) continue;
        carve_walk(g, h->x_entry[i], h->y_entry[i], h->z_entry[i], h->x_exit[i], h->y_exit[i], h->z_exit[i],
                   [&](int64_t v) { if (v >= 0 && v < g.V) counts[v] += 1; });
    }
    return 0;
}

/* C3 (SM): SM_GPU over all histories (.cu:3608-3710): the same walk for the
 * histories with WEPL >= SM_LOWER_THRESHOLD (.cu:3615) */
int oracle_sm_count(const pct_params* p, const pct_hist_soa* h, const uint8_t* missed, int32_t* counts) {
    Geo g = make_geo(*p);
    for (int64_t i = 0; i < h->n; ++i) {
        if (missed[i] || !(h->wepl[i] >= p->sm_lower_threshold)) continue;
        carve_walk(g, h->x_entry[i], h->y_entry[i], h->z_entry[i], h->x_exit[i], h->y_exit[i], h->z_exit[i],
                   [&](int64_t v) { if (v >= 0 && v < g.V) counts[v] += 1; });
    }
    return 0;
}

/* C4 (SM): SM_edge_detection (.cu:3711-3773): carve_differences (.cu:3274-3292)
 * into a zeroed array, the per-slice threshold from an in-order scan with a
 * strict > on the counts as they were before edge detection (the host copy
 * written for WRITE_SM_COUNTS, .cu:3933-3937), then SM_edge_detection_GPU
 * (.cu:3774-3789) out of place */
int oracle_sm_edge(const pct_params* p, const int32_t* counts, int32_t* out, int32_t* thresholds) {
    Geo g = make_geo(*p);
    const int C = g.C, R = g.R;
    std::vector<int32_t> differences(g.V, 0);
    for (int slice = 0; slice < g.S; ++slice)
        for (int row = 1; row < R - 1; ++row)
            for (int column = 1; column < C - 1; ++column) {
                int64_t voxel = column + (int64_t)row * C + (int64_t)slice * C * R;
                int max_difference = 0;
                for (int cr = row - 1; cr <= row + 1; ++cr)
                    for (int cc = column - 1; cc <= column + 1; ++cc) {
                        int d = counts[voxel] - counts[cc + (int64_t)cr * C + (int64_t)slice * C * R];
                        if (d > max_difference) max_difference = d;
                    }
                differences[voxel] = max_difference;
            }
    double rr = g.radius - (g.voxel_width > g.voxel_height ? g.voxel_width : g.voxel_height) / 2;
    for (int slice = 0; slice < g.S; ++slice) {
        int threshold = 0, max_difference = 0;
        for (int64_t pixel = 0; pixel < (int64_t)C * R; ++pixel) {
            int64_t voxel = pixel + (int64_t)slice * C * R;
            if (differences[voxel] > max_difference) {
                max_difference = differences[voxel];
                threshold = counts[voxel];
            }
        }
        if (thresholds) thresholds[slice] = threshold;
        for (int row = 0; row < R; ++row)
            for (int column = 0; column < C; ++column) {
                int64_t voxel = column + (int64_t)row * C + (int64_t)slice * C * R;
                float x = (float)((column - C / 2 + 0.5) * g.voxel_width);
                float y = (float)((R / 2 - row - 0.5) * g.voxel_height);
                int v = counts[voxel] > p->sm_scale_threshold * threshold ? 1 : 0;
                float s2 = x * x + y * y;
                if ((double)s2 >= rr * rr) v = 0;
                out[voxel] = v;
            }
    }
    return 0;
}

/* C4: MSC_edge_detection_GPU, out of place */
int oracle_msc_edge(const pct_params* p, const int32_t* counts, int32_t* out) {
    Geo g = make_geo(*p);
    const int C = g.C, R = g.R;
    double rr = g.radius - (g.voxel_width > g.voxel_height ? g.voxel_width : g.voxel_height) / 2;
    for (int slice = 0; slice < g.S; ++slice)
        for (int row = 0; row < R; ++row)
            for (int column = 0; column < C; ++column) {
                int64_t voxel = column + (int64_t)row * C + (int64_t)slice * C * R;
                float x = (float)((column - C / 2 + 0.5) * g.voxel_width);
                float y = (float)((R / 2 - row - 0.5) * g.voxel_height);
                int max_difference = 0;
                if (row != 0 && row != R - 1 && column != 0 && column != C - 1)
                    for (int cr = row - 1; cr <= row + 1; ++cr)
                        for (int cc = column - 1; cc <= column + 1; ++cc) {
                            int d = counts[voxel] - counts[cc + (int64_t)cr * C + (int64_t)slice * C * R];
                            if (d > max_difference) max_difference = d;
                        }
                int v = max_difference > p->msc_diff_thresh ? 0 : 1;
                float s2 = x * x + y * y;
                if ((double)s2 >= rr * rr) v = 0;
                out[voxel] = v;
            }
    return 0;
}

/* C5: hull_selection (hull = counts != 0) + averaging_filter_GPU with
 * perform_threshold (radius, threshold) */
int oracle_hull_select(const pct_params* p, const int32_t* counts, uint8_t* hull) {
    Geo g = make_geo(*p);
    const int C = g.C, R = g.R, r = p->hull_filter_radius;
    std::vector<uint8_t> x_hull(g.V);
    for (int64_t v = 0; v < g.V; ++v) x_hull[v] = counts[v] != 0;
    for (int z = 0; z < g.S; ++z)
        for (int vy = 0; vy < R; ++vy)
            for (int vx = 0; vx < C; ++vx) {
                unsigned left = vx - r > 0 ? vx - r : 0, right = vx + r < C - 1 ? vx + r : C - 1;
                unsigned top = vy - r > 0 ? vy - r : 0, bottom = vy + r < R - 1 ? vy + r : R - 1;
                int nvox = (int)((right - left + 1) * (bottom - top + 1));
                double sum_threshold = nvox * p->hull_filter_threshold;
                double sum = 0.0;
                for (int c = (int)left; c <= (int)right; ++c)
                    for (int rr = (int)top; rr <= (int)bottom; ++rr) sum += x_hull[c + (int64_t)rr * C + (int64_t)z * C * R];
                hull[vx + (int64_t)vy * C + (int64_t)z * C * R] = sum > sum_threshold;
            }
    return 0;
}

/* B1: filter_GPU with the Shepp-Logan kernel of .cu:3064 */
int oracle_fbp_filter(const pct_params* p, const float* sino, float* out) {
    const int T = p->t_bins, V = p->v_bins, A = p->angular_bins;
    const double TS = p->t_bin_size, VS = p->v_bin_size, SR = p->source_radius;
    for (int v_bin = 0; v_bin < V; ++v_bin)
        for (int angle_bin = 0; angle_bin < A; ++angle_bin)
            for (int t_bin = 0; t_bin < T; ++t_bin) {
                double v = (v_bin - V / 2) * VS + VS / 2.0;
                int64_t strip = (int64_t)v_bin * A * T + (int64_t)angle_bin * T;
                float acc = 0.0f;
                for (int t_ref = 0; t_ref < T; ++t_ref) {
                    double t = (t_ref - T / 2) * TS + TS / 2.0;
                    int sep = t_bin - t_ref;
                    double scale = SR / pct_sqrt(SR * SR + t * t + v * v);
                    double filtered = 1 / ((TS * PCT_PI * TS * PCT_PI) * (1.0 - (2 * sep * 2 * sep)));
                    acc = (float)(acc + TS * sino[strip + t_ref] * filtered * scale);
                }
                out[strip + t_bin] = acc;
            }
    return 0;
}

/* B2: backprojection_GPU */
int oracle_backprojection(const pct_params* p, const float* sf, float* image) {
    Geo g = make_geo(*p);
    const int T = p->t_bins, V = p->v_bins, A = p->angular_bins;
    const int64_t NB = (int64_t)T * V * A;
    const double SR = p->source_radius;
    const double delta = p->gantry_angle_interval * (PCT_PI / 180.0);
    auto at = [&](int64_t b) -> double { return (b >= 0 && b < NB) ? (double)sf[b] : 0.0; };
    for (int slice = 0; slice < g.S; ++slice)
        for (int row = 0; row < g.R; ++row)
            for (int column = 0; column < g.C; ++column) {
                int64_t voxel = (int64_t)slice * g.C * g.R + (int64_t)row * g.C + column;
                double x = -g.radius + (column + 0.5) * g.voxel_width;
                double y = g.radius - (row + 0.5) * g.voxel_height;
                double z = -g.height / 2.0 + (slice + 0.5) * p->slice_thickness;
                if ((x * x + y * y) > (g.radius * g.radius)) {
                    image[voxel] = (float)p->rsp_air;
                    continue;
                }
                float acc = 0.0f;
                for (int a = 0; a < A; ++a) {
                    double s, c;
                    pct_sincos(a * delta, &s, &c);
                    double u = x * c + y * s;
                    double t = -x * s + y * c;
                    double v = z;
                    double dnt = (t - u * (t / (SR + u))) / p->t_bin_size + T / 2.0;
                    int t_bin = (int)dnt;
                    if (t_bin > dnt) t_bin -= 1;
                    double eta = dnt - t_bin;
                    double dnv = (v - u * (v / (SR + u))) / p->v_bin_size + V / 2.0;
                    int v_bin = (int)dnv;
                    if (v_bin > dnv) v_bin -= 1;
                    double eps = dnv - v_bin;
                    double q = SR / (SR + u);
                    double scale = q * q;
                    int64_t bin = t_bin + (int64_t)a * T + (int64_t)v_bin * A * T;
                    if (v_bin == V - 1 || bin < 0)
                        acc = (float)(acc + scale * (((1 - eta) * at(bin)) + (eta * at(bin + 1))));
                    else
                        acc = (float)(acc + scale * ((1 - eta) * (1 - eps) * at(bin) + eta * (1 - eps) * at(bin + 1)
                                                     + (1 - eta) * eps * at(bin + (int64_t)A * T)
                                                     + eta * eps * at(bin + (int64_t)A * T + 1)));
                }
                image[voxel] = (float)(acc * delta);
            }
    return 0;
}

/* B3: FBP_image_2_hull_GPU */
int oracle_fbp_hull(const pct_params* p, const float* image, uint8_t* hull) {
    Geo g = make_geo(*p);
    for (int slice = 0; slice < g.S; ++slice)
        for (int row = 0; row < g.R; ++row)
            for (int column = 0; column < g.C; ++column) {
                int64_t voxel = (int64_t)slice * g.C * g.R + (int64_t)row * g.C + column;
                double x = -g.radius + (column + 0.5) * g.voxel_width;
                double y = g.radius - (row + 0.5) * g.voxel_height;
                double d2 = x * x + y * y;
                hull[voxel] = (image[voxel] > p->fbp_threshold && d2 < g.radius * g.radius) ? 1 : 0;
            }
    return 0;
}

/* E5: HYBRID initial iterate x0 = FBP * hull (float multiply) */
int oracle_initial_iterate(const pct_params* p, const float* fbp, const uint8_t* hull, float* x0) {
    Geo g = make_geo(*p);
    for (int64_t v = 0; v < g.V; ++v) x0[v] = fbp[v] * (float)hull[v];
    return 0;
}

}  // extern "C"
