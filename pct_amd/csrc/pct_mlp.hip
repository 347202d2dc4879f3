/*
 * pct_mlp.hip -- (d) MLP endpoints and the exact FP64 most-likely-path walk.
 *
 * Endpoints: replaces collect_MLP_endpoints_GPU (.cu:5729-5770) and
 * find_MLP_endpoints_GPU (.cu:4997-5277): one thread per history, two
 * straight-line voxel walks on a bit-packed hull (1 bit per voxel instead of
 * the reference's bool array: 8x less hull traffic, 160 KB at 200x200x32).
 *
 * MLP: replaces the path part of find_MLP_path_GPU (.cu:5279-5501). The
 * reference runs one thread per history through ~2u2/w sequential steps and
 * appends the step's voxel when it differs from the previous one. The steps
 * are independent of each other except for that dedup, so here one wavefront
 * takes one history at a time and its 64 lanes evaluate 64 consecutive steps;
 * the dedup becomes a ballot + prefix popcount and the path is written
 * coalesced. Terms that depend only on the depth u_1 (sigma_1 and its
 * inverse, the three common sigma_2 polynomials) are read from a per-depth
 * FP64 table built on the host with the very same expressions and the same
 * accumulated u_1 sequence (u_1 += MLP_U_STEP, .cu:5490), so every step value
 * is bit-identical to the reference's inline evaluation; only the u_2
 * dependent terms (one log, the Sigma_2 inverse, the 2x2 solve, the rotation
 * and the voxel index) are computed per step.
 */
#include <algorithm>
#include <vector>
#include <rocprim/device/device_radix_sort.hpp>
#include "pct_internal.h"
#include "pct_net_basis.h"
#include "pct_net_dev.h"
#include "pct_rdiv.h"

namespace {

/* ------------------------------------------------------------------------- */
/* voxel helpers (.cu:8508-8510, 8721-8752)                                   */
/* ------------------------------------------------------------------------- */
__device__ __forceinline__ int calc_voxel(double zero, double pos, double size) {
    double q = pct_fabs(pos - zero) / size;
    return q < 1.0e9 ? (int)q : -1;
}
#define PCT_LDS_PER_CU (160u * 1024u)        /* LDS of one CDNA4 compute unit */
#define PCT_MLP_LDS_TAB_MAX (64u * 1024u)   /* largest per-depth table staged in LDS */

/* calc_voxel with the division replaced by a product with the rounded
 * reciprocal. q = |pos - zero| * inv carries two roundings (of inv and of the
 * product), the division's quotient one, so the two differ by less than
 * 2^-51 q and have the same integer part unless an integer lies that close to
 * q: then `near` is set and the caller evaluates calc_voxel itself (also for
 * q = 0). Branch-free, so that two steps in flight stay one straight-line
 * block. */
__device__ __forceinline__ int calc_voxel_r(double zero, double pos, double inv, bool& near) {
    double a = pct_fabs(pos - zero);
    double q = a * inv;
    near = near || (pct_fabs(q - __builtin_rint(q)) <= q * 0x1p-49);
    return q < 1.0e9 ? (int)q : -1;
}
__device__ __forceinline__ double edge_coord(double zero, int voxel_entered, double size, int inc, int step) {
    int on_edge = (step == inc) ? voxel_entered : voxel_entered + 1;
    return zero + (inc * on_edge * size);
}
__device__ __forceinline__ double dist_remaining(double zero, double pos, int inc, int step, double size, int cur) {
    return pct_fabs(edge_coord(zero, cur + inc * step, size, inc, step) - pos);
}
__device__ __forceinline__ double corr_coord(double m, double x, double x0, double y0) { return m * (x - x0) + y0; }
__device__ __forceinline__ bool hull_bit(const uint32_t* __restrict__ hull, int64_t v) {
    return (hull[v >> 5] >> (v & 31)) & 1u;
}

/* one endpoint walk (.cu:4997-5277); returns hit_hull */
__device__ bool endpoint_walk(const pct_geo& g, const uint32_t* __restrict__ hull, double x_start, double y_start,
                              double z_start, double xy_angle, double xz_angle, double& x_obj, double& y_obj,
                              double& z_obj, int& vx, int& vy, int& vz, bool entering) {
    const int C = g.C, R = g.R, S = g.S;
    const int X_INC = 1, Y_INC = -1, Z_INC = -1;
    double x = x_start, y = y_start, z = z_start;
    if (!entering) xy_angle += PCT_PI;
    double sxy, cxy, sxz, cxz;
    pct_sincos(xy_angle, &sxy, &cxy);
    pct_sincos(xz_angle, &sxz, &cxz);
    int xm = (cxy >= 0) - (cxy <= 0);
    int ym = (sxy >= 0) - (sxy <= 0);
    int zm = (sxz >= 0) - (sxz <= 0);
    if (xm < 0) zm *= -1;
    vx = calc_voxel(g.x_zero, x, g.voxel_width);
    vy = calc_voxel(g.y_zero, y, g.voxel_height);
    vz = calc_voxel(g.z_zero, z, g.voxel_thickness);
    double x_to_go = dist_remaining(g.x_zero, x, X_INC, xm, g.voxel_width, vx);
    double y_to_go = dist_remaining(g.y_zero, y, Y_INC, ym, g.voxel_height, vy);
    double z_to_go = dist_remaining(g.z_zero, z, Z_INC, zm, g.voxel_thickness, vz);
    double txy = sxy / cxy, txz = sxz / cxz;
    double delta_yx = pct_fabs(txy), delta_zx = pct_fabs(txz), delta_zy = pct_fabs(txz / txy);
    double dy_dx = txy, dz_dx = txz, dz_dy = txz / txy, dx_dy = 1.0 / txy, dx_dz = 1.0 / txz, dy_dz = txy / txz;
    bool hit = false;
    bool outside = (vx >= C) || (vy >= R) || (vz >= S) || (vx < 0) || (vy < 0) || (vz < 0);
    if (!outside) hit = hull_bit(hull, vx + (int64_t)vy * C + (int64_t)vz * C * R);
    bool end = outside || hit;
    int guard = 4 * (C + R + S) + 16;
    if (zm != 0) {
        while (!end && guard-- > 0) {
            double x_ext = delta_zx * x_to_go;
            double y_ext = delta_zy * y_to_go;
            if ((z_to_go <= x_ext) && (z_to_go <= y_ext)) {
                vz -= zm;
                z = edge_coord(g.z_zero, vz, g.voxel_thickness, Z_INC, zm);
                x = corr_coord(dx_dz, z, z_start, x_start);
                y = corr_coord(dy_dz, z, z_start, y_start);
                x_to_go = dist_remaining(g.x_zero, x, X_INC, xm, g.voxel_width, vx);
                y_to_go = dist_remaining(g.y_zero, y, Y_INC, ym, g.voxel_height, vy);
                z_to_go = g.voxel_thickness;
            } else if (x_ext <= y_ext) {
                vx += xm;
                x = edge_coord(g.x_zero, vx, g.voxel_width, X_INC, xm);
                y = corr_coord(dy_dx, x, x_start, y_start);
                z = corr_coord(dz_dx, x, x_start, z_start);
                x_to_go = g.voxel_width;
                y_to_go = dist_remaining(g.y_zero, y, Y_INC, ym, g.voxel_height, vy);
                z_to_go = dist_remaining(g.z_zero, z, Z_INC, zm, g.voxel_thickness, vz);
            } else {
                vy -= ym;
                y = edge_coord(g.y_zero, vy, g.voxel_height, Y_INC, ym);
                x = corr_coord(dx_dy, y, y_start, x_start);
                z = corr_coord(dz_dy, y, y_start, z_start);
                x_to_go = dist_remaining(g.x_zero, x, X_INC, xm, g.voxel_width, vx);
                y_to_go = g.voxel_height;
                z_to_go = dist_remaining(g.z_zero, z, Z_INC, zm, g.voxel_thickness, vz);
            }
            if (x_to_go == 0) { x_to_go = g.voxel_width; vx += xm; }
            if (y_to_go == 0) { y_to_go = g.voxel_height; vy -= ym; }
            if (z_to_go == 0) { z_to_go = g.voxel_thickness; vz -= zm; }
            vz = vz > 0 ? vz : 0;
            outside = (vx >= C) || (vy >= R) || (vz >= S) || (vx < 0) || (vy < 0) || (vz < 0);
            if (!outside) hit = hull_bit(hull, vx + (int64_t)vy * C + (int64_t)vz * C * R);
            end = outside || hit;
        }
    } else {
        while (!end && guard-- > 0) {
            double y_ext = y_to_go / delta_yx;
            if (x_to_go <= y_ext) {
                vx += xm;
                x = edge_coord(g.x_zero, vx, g.voxel_width, X_INC, xm);
                y = corr_coord(dy_dx, x, x_start, y_start);
                x_to_go = g.voxel_width;
                y_to_go = dist_remaining(g.y_zero, y, Y_INC, ym, g.voxel_height, vy);
            } else {
                vy -= ym;
                y = edge_coord(g.y_zero, vy, g.voxel_height, Z_INC, ym);
                x = corr_coord(dx_dy, y, y_start, x_start);
                x_to_go = dist_remaining(g.x_zero, x, X_INC, xm, g.voxel_width, vx);
                y_to_go = g.voxel_height;
            }
            if (x_to_go == 0) { x_to_go = g.voxel_width; vx += xm; }
            if (y_to_go == 0) { y_to_go = g.voxel_height; vy -= ym; }
            outside = (vx >= C) || (vy >= R) || (vz >= S) || (vx < 0) || (vy < 0) || (vz < 0);
            if (!outside) hit = hull_bit(hull, vx + (int64_t)vy * C + (int64_t)vz * C * R);
            end = outside || hit;
        }
    }
    if (!end) hit = false;
    if (hit) { x_obj = x; y_obj = y; z_obj = z; }
    return hit;
}

__global__ void __launch_bounds__(256)
k_mlp_endpoints(pct_geo g, pct_hist_soa h, const uint32_t* __restrict__ hull, uint8_t* __restrict__ intersected) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h.n) return;
    double xi = 0, yi = 0, zi = 0, xo = 0, yo = 0, zo = 0;
    int vx, vy, vz, vxo, vyo, vzo;
    bool en = endpoint_walk(g, hull, h.x_entry[i], h.y_entry[i], h.z_entry[i], h.xy_entry_angle[i], h.xz_entry_angle[i],
                            xi, yi, zi, vx, vy, vz, true);
    bool ex = endpoint_walk(g, hull, h.x_exit[i], h.y_exit[i], h.z_exit[i], h.xy_exit_angle[i], h.xz_exit_angle[i],
                            xo, yo, zo, vxo, vyo, vzo, false);
    uint32_t first = 0;
    if (en && ex) {
        first = (uint32_t)(vx + g.C * vy + g.R * g.C * vz);
        h.x_entry[i] = (float)xi; h.y_entry[i] = (float)yi; h.z_entry[i] = (float)zi;
        h.x_exit[i] = (float)xo;  h.y_exit[i] = (float)yo;  h.z_exit[i] = (float)zo;
    }
    h.first_voxel[i] = first;
    intersected[i] = (uint8_t)(en && ex);
}

/* ------------------------------------------------------------------------- */
/* D3: EffectiveChordLength_GPU (.cu:4974-4995). sinf/cosf receive a float.   */
/* ------------------------------------------------------------------------- */
__device__ __forceinline__ double chord_length(const pct_geo& g, double abs_angle_t, double abs_angle_v) {
    double eff_angle_t = abs_angle_t - ((int)(abs_angle_t / (PCT_PI / 2))) * (PCT_PI / 2);
    double eff_angle_v = pct_fabs(abs_angle_v);
    double step_fraction = g.mlp_u_step / g.voxel_width;
    float s2t = pct_sinf((float)(2 * eff_angle_t));
    float cs_t = pct_cosf((float)eff_angle_t) + pct_sinf((float)eff_angle_t);
    double chord_2d = (1 / 3.0) * ((step_fraction * step_fraction * s2t - 6) / (step_fraction * s2t - 2 * cs_t)
                                   + step_fraction * step_fraction * s2t / (2 * cs_t));
    double mean_pixel_width = g.voxel_width / cs_t;
    double height_fraction = g.voxel_thickness / mean_pixel_width;
    step_fraction = g.mlp_u_step / mean_pixel_width;
    float s2v = pct_sinf((float)(2 * eff_angle_v));
    double hcs = height_fraction * pct_cosf((float)eff_angle_v) + pct_sinf((float)eff_angle_v);
    double chord_3d = (1 / 3.0) * ((step_fraction * step_fraction * s2v - 6 * height_fraction) / (step_fraction * s2v - 2 * hcs)
                                   + step_fraction * step_fraction * s2v / (2 * hcs));
    return g.voxel_width * chord_2d * chord_3d;
}

/* chord length per history: the double average of the float angles (the
 * GPU call, .cu:5315) or their float average (the CPU research call, .cu:7539) */
__global__ void __launch_bounds__(256)
k_chord_lengths(pct_geo g, pct_hist_soa h, int float_average, double* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h.n) return;
    double t, v;
    if (float_average) {
        float st = h.xy_entry_angle[i] + h.xy_exit_angle[i];
        float sv = h.xz_entry_angle[i] + h.xz_exit_angle[i];
        t = st / 2;
        v = sv / 2;
    } else {
        t = ((double)h.xy_entry_angle[i] + (double)h.xy_exit_angle[i]) / 2.0;
        v = ((double)h.xz_entry_angle[i] + (double)h.xz_exit_angle[i]) / 2.0;
    }
    out[i] = chord_length(g, t, v);
}

/* per-history MLP constants (.cu:5306-5374) */
struct mlp_hist {
    double ca, sa, u_in, t_in, t_out, dT, v_in, v_out, dV, u_2, pre1, pre2, pre3;
    double chord;
    int S;
};

__device__ __forceinline__ int steps_below(const double* __restrict__ u1tab, int K, double lim) {
    /* number of leading k with u1tab[k] < lim (u1tab strictly increasing);
     * 0 for NaN, K when every entry is below */
    int lo = 0, hi = K;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (u1tab[mid] < lim) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ void mlp_setup(const pct_geo& g, const pct_hist_soa& h, int64_t i,
                                          const double* __restrict__ u1tab, int K, mlp_hist& m) {
    double x_in = h.x_entry[i], y_in = h.y_entry[i], z_in = h.z_entry[i];
    double x_out = h.x_exit[i], y_out = h.y_exit[i], z_out = h.z_exit[i];
    double xy_in = h.xy_entry_angle[i], xz_in = h.xz_entry_angle[i];
    double xy_out = h.xy_exit_angle[i], xz_out = h.xz_exit_angle[i];
    m.chord = chord_length(g, (xy_in + xy_out) / 2.0, (xz_in + xz_out) / 2.0);
    double sa, ca;
    pct_sincos(xy_in, &sa, &ca);
    double u_in = (ca * x_in) + (sa * y_in);
    double u_out = (ca * x_out) + (sa * y_out);
    double t_in = (ca * y_in) - (sa * x_in);
    double t_out = (ca * y_out) - (sa * x_out);
    if (u_in > u_out) {
        xy_in += PCT_PI;
        xy_out += PCT_PI;
        pct_sincos(xy_in, &sa, &ca);
        u_in = (ca * x_in) + (sa * y_in);
        u_out = (ca * x_out) + (sa * y_out);
        t_in = (ca * y_in) - (sa * x_in);
        t_out = (ca * y_out) - (sa * x_out);
    }
    m.ca = ca; m.sa = sa;
    m.u_in = u_in; m.t_in = t_in; m.t_out = t_out; m.dT = xy_out - xy_in;
    m.v_in = z_in; m.v_out = z_out; m.dV = xz_out - xz_in;
    double u_2 = pct_fabs(u_out - u_in);
    m.u_2 = u_2;
    m.pre1 = u_2 * u_2 * u_2 * (g.a0_3 + u_2 * (g.a1_12 + u_2 * (g.a2_30 + u_2 * (g.a3_60 + u_2 * (g.a4_105 + u_2 * g.a5_168)))));
    m.pre2 = u_2 * u_2 * (g.a0_2 + u_2 * (g.a1_6 + u_2 * (g.a2_12 + u_2 * (g.a3_20 + u_2 * (g.a4_30 + u_2 * g.a5_42)))));
    m.pre3 = u_2 * (g.A0 + u_2 * (g.a1_2 + u_2 * (g.a2_3 + u_2 * (g.a3_4 + u_2 * (g.a4_5 + u_2 * g.a5_6)))));
    m.S = steps_below(u1tab, K, u_2 - g.mlp_u_step);
}

/* the u_2-dependent part of one MLP step (.cu:5402-5472): (t_1, v_1) from the
 * u_1-only terms S1I* (Sigma_1^-1) and cst* (the common sigma_2 polynomials).
 * With FAST the seven quotients share their reciprocals (rdiv) and the return
 * value says whether every operand was inside div_ok's range -- when it is
 * not, the caller evaluates the step again without FAST. */
template <bool FAST = false>
__device__ __forceinline__ bool mlp_tv_terms(const pct_geo& g, const mlp_hist& m, double u_1, double S1I0, double S1I1,
                                             double S1I3, double cst1, double cst2, double cst3, double& t_out,
                                             double& v_out) {
    double d = m.u_2 - u_1;                       /* R_1[1] = R_1T[2] */
    bool ok = true;
    double l2, c2;
    if constexpr (FAST) {
        /* (u_2 - u_1) / x0 is positive and normal when d is positive and
         * in div_ok's range: the logarithm without its special cases */
        const rdiv x0 = rdiv_make(g.x_0);
        l2 = g.e_0 * (1 + 0.038 * pct_log_normal(rdiv_q(x0, m.u_2 - u_1), div_rcp{}));
        const double l2sq = l2 * l2;
        c2 = rdiv_q(x0, l2sq);
        ok = div_ok(g.x_0) & div_ok(d) & (d > 0.0) & div_ok(l2sq);
    } else {
        l2 = g.e_0 * (1 + 0.038 * pct_log((m.u_2 - u_1) / g.x_0));
        c2 = l2 * l2 / g.x_0;
    }
    double sigma_t2 = c2 * (m.pre1 - m.u_2 * m.u_2 * cst1 + 2 * m.u_2 * cst2 - cst3);
    double sigma_t2_theta2 = c2 * (m.pre2 - m.u_2 * cst1 + cst2);
    double sigma_theta2 = c2 * (m.pre3 - cst1);
    double det2 = sigma_t2 * sigma_theta2 - sigma_t2_theta2 * sigma_t2_theta2;
    double S2I0, S2I1, S2I3;
    if constexpr (FAST) {
        const rdiv r2 = rdiv_make(det2);
        S2I0 = rdiv_q(r2, sigma_theta2);
        S2I1 = rdiv_q(r2, -sigma_t2_theta2);
        S2I3 = rdiv_q(r2, sigma_t2);
        ok = ok & div_ok(det2) & div_ok(sigma_theta2) & div_ok(sigma_t2_theta2) & div_ok(sigma_t2);
    } else {
        S2I0 = sigma_theta2 / det2;
        S2I1 = -sigma_t2_theta2 / det2;          /* = Sigma_2^-1[2] */
        S2I3 = sigma_t2 / det2;
    }
    /* The reference forms these products with R_1 = {1, d, 0, 1}, R_1T =
     * {1, 0, d, 1}, R_0 = {1, u_1, 0, 1} and T_0 = {t_in, 0}. A product with
     * an exact 1 is the operand itself, and a product with an exact 0 adds a
     * signed zero, which leaves every nonzero sum as it is; both are left out
     * and the remaining terms are summed in the reference's order, so t_1 and
     * v_1 are the reference's values (up to the sign of an exactly zero
     * result). */
    double f24_1 = S2I0 * d + S2I1;
    double f24_2 = S2I1 * d + S2I3;
    double F0 = S1I0 + S2I0;
    double F1 = S1I1 + f24_1;
    double F2 = (S1I1 + d * S2I0) + S2I1;
    double F3 = (S1I3 + d * f24_1) + f24_2;
    double detF = F0 * F3 - F1 * F2;
    double I0, I1;
    if constexpr (FAST) {
        const rdiv rF = rdiv_make(detF);
        I0 = rdiv_q(rF, F3);
        I1 = rdiv_q(rF, -F1);
        ok = ok & div_ok(detF) & div_ok(F3) & div_ok(F1);
    } else {
        I0 = F3 / detF;
        I1 = -F1 / detF;
    }
    double a3 = S2I0 * m.t_out + S2I1 * m.dT;
    double a4 = S2I1 * m.t_out + S2I3 * m.dT;
    double s0 = S1I0 * m.t_in + a3;
    double s1 = (S1I1 * m.t_in + d * a3) + a4;
    double t_1 = I0 * s0 + I1 * s1;
    double b3 = S2I0 * m.v_out + S2I1 * m.dV;
    double b4 = S2I1 * m.v_out + S2I3 * m.dV;
    double r0 = S1I0 * m.v_in + b3;
    double r1 = (S1I1 * m.v_in + d * b3) + b4;
    double v_1 = I0 * r0 + I1 * r1;
    t_out = t_1;
    v_out = v_1;
    return ok;
}

/* (t_1, v_1) of the MLP at table index k (FAST: see mlp_tv_terms) */
template <bool FAST = false>
__device__ __forceinline__ bool mlp_tv(const pct_geo& g, const mlp_hist& m, const double* __restrict__ tab, int K,
                                       int k, double& u_1, double& t_1, double& v_1) {
    u_1 = tab[MLP_TAB_U1 * K + k];
    return mlp_tv_terms<FAST>(g, m, u_1, tab[MLP_TAB_S1I0 * K + k], tab[MLP_TAB_S1I1 * K + k],
                              tab[MLP_TAB_S1I3 * K + k], tab[MLP_TAB_C1 * K + k], tab[MLP_TAB_C2 * K + k],
                              tab[MLP_TAB_C3 * K + k], t_1, v_1);
}

/* (t_1, v_1) of the MLP at an arbitrary depth u_1: the u_1-only terms
 * evaluated directly with the expressions of the per-depth table */
__device__ __forceinline__ void mlp_point_at(const pct_geo& g, const mlp_hist& m, double u_1, double& t_1, double& v_1) {
    const double u_0 = 0.0;
    double l1 = g.e_0 * (1 + 0.038 * pct_log((u_1 - u_0) / g.x_0));
    double c1 = l1 * l1 / g.x_0;
    double sigma_t1 = c1 * (u_1 * u_1 * u_1 * (g.a0_3 + u_1 * (g.a1_12 + u_1 * (g.a2_30 + u_1 * (g.a3_60 + u_1 * (g.a4_105 + u_1 * g.a5_168))))));
    double sigma_t1_theta1 = c1 * (u_1 * u_1 * (g.a0_2 + u_1 * (g.a1_6 + u_1 * (g.a2_12 + u_1 * (g.a3_20 + u_1 * (g.a4_30 + u_1 * g.a5_42))))));
    double sigma_theta1 = c1 * (u_1 * (g.A0 + u_1 * (g.a1_2 + u_1 * (g.a2_3 + u_1 * (g.a3_4 + u_1 * (g.a4_5 + u_1 * g.a5_6))))));
    double det1 = sigma_t1 * sigma_theta1 - sigma_t1_theta1 * sigma_t1_theta1;
    double cst1 = u_1 * (g.A0 + u_1 * (g.a1_2 + u_1 * (g.a2_3 + u_1 * (g.a3_4 + u_1 * (g.a4_5 + u_1 * g.a5_6)))));
    double cst2 = u_1 * u_1 * (g.a0_2 + u_1 * (g.a1_3 + u_1 * (g.a2_4 + u_1 * (g.a3_5 + u_1 * (g.a4_6 + u_1 * g.a5_7)))));
    double cst3 = u_1 * u_1 * u_1 * (g.a0_3 + u_1 * (g.a1_4 + u_1 * (g.a2_5 + u_1 * (g.a3_6 + u_1 * (g.a4_7 + u_1 * g.a5_8)))));
    mlp_tv_terms(g, m, u_1, sigma_theta1 / det1, -sigma_t1_theta1 / det1, sigma_t1 / det1, cst1, cst2, cst3, t_1, v_1);
}

/* one MLP step (.cu:5376-5489) at table index k; returns the voxel index or -1
 * when it falls outside the grid */
__device__ __forceinline__ int64_t mlp_step(const pct_geo& g, const mlp_hist& m, const double* __restrict__ tab, int K, int k) {
    double u_1, t_1, v_1;
    mlp_tv(g, m, tab, K, k, u_1, t_1, v_1);
    double x_1 = (m.ca * (m.u_in + u_1)) - (m.sa * t_1);
    double y_1 = (m.sa * (m.u_in + u_1)) + (m.ca * t_1);
    int vx = calc_voxel(g.x_zero, x_1, g.voxel_width);
    int vy = calc_voxel(g.y_zero, y_1, g.voxel_height);
    int vz = calc_voxel(g.z_zero, v_1, g.voxel_thickness);
    int64_t voxel = (int64_t)vx + (int64_t)vy * g.C + (int64_t)vz * g.C * g.R;
    if (vx < 0 || vy < 0 || vz < 0 || voxel >= g.V) return -1;
    return voxel;
}

/* mlp_step with the shared-reciprocal quotients and the voxel index from the
 * reciprocal products; returns -2 where an operand is outside the quotients'
 * range or a product too close to a voxel face to decide, and mlp_step then
 * gives the step's voxel */
__device__ __forceinline__ int64_t mlp_step_r(const pct_geo& g, const mlp_hist& m, const double* __restrict__ tab, int K,
                                              int k) {
    double u_1, t_1, v_1;
    if (!mlp_tv<true>(g, m, tab, K, k, u_1, t_1, v_1)) return -2;
    double x_1 = (m.ca * (m.u_in + u_1)) - (m.sa * t_1);
    double y_1 = (m.sa * (m.u_in + u_1)) + (m.ca * t_1);
    bool near = false;
    int vx = calc_voxel_r(g.x_zero, x_1, g.inv_width, near);
    int vy = calc_voxel_r(g.y_zero, y_1, g.inv_height, near);
    int vz = calc_voxel_r(g.z_zero, v_1, g.inv_thickness, near);
    int64_t voxel = (int64_t)vx + (int64_t)vy * g.C + (int64_t)vz * g.C * g.R;
    if (vx < 0 || vy < 0 || vz < 0 || voxel >= g.V) voxel = -1;
    return near ? -2 : voxel;
}

/* mlp_step_r with step k's table row given (u_1, S1I0, S1I1, S1I3, cst1,
 * cst2, cst3): the same expressions, so the same voxel bit for bit */
__device__ __forceinline__ int64_t mlp_step_r_row(const pct_geo& g, const mlp_hist& m, const double (&row)[MLP_TAB_FIELDS]) {
    const double u_1 = row[MLP_TAB_U1];
    double t_1, v_1;
    if (!mlp_tv_terms<true>(g, m, u_1, row[MLP_TAB_S1I0], row[MLP_TAB_S1I1], row[MLP_TAB_S1I3], row[MLP_TAB_C1],
                            row[MLP_TAB_C2], row[MLP_TAB_C3], t_1, v_1))
        return -2;
    double x_1 = (m.ca * (m.u_in + u_1)) - (m.sa * t_1);
    double y_1 = (m.sa * (m.u_in + u_1)) + (m.ca * t_1);
    bool near = false;
    int vx = calc_voxel_r(g.x_zero, x_1, g.inv_width, near);
    int vy = calc_voxel_r(g.y_zero, y_1, g.inv_height, near);
    int vz = calc_voxel_r(g.z_zero, v_1, g.inv_thickness, near);
    int64_t voxel = (int64_t)vx + (int64_t)vy * g.C + (int64_t)vz * g.C * g.R;
    if (vx < 0 || vy < 0 || vz < 0 || voxel >= g.V) voxel = -1;
    return near ? -2 : voxel;
}

/* position-aware mix of one list element; a list's hash is the sum over its
 * elements plus a length term, so it can be formed by a wave reduction */
__device__ __forceinline__ uint32_t pct_list_mix(uint32_t v, uint32_t pos) {
    uint32_t h = (v ^ (pos * 0x9E3779B9u)) * 0x85EBCA6Bu;
    h ^= h >> 13;
    return h * 0xC2B2AE35u;
}

__device__ __forceinline__ double bcast_d(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ int bcast_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ int64_t bcast_l(int64_t v, int lane) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), lane);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

/* The voxel list of one history from its S steps, one wavefront: 64 lanes
 * evaluate 64 consecutive steps (step(k) returns the voxel or -1 outside the
 * grid), a step is appended when it differs from the previous valid one (the
 * dedup of .cu:5478-5485: a ballot and a prefix popcount across the wave), the
 * list is written from voxels[hoff] on (the first voxel `carry` included)
 * while it fits hcap, and its position-aware hash is accumulated. Returns the
 * list length n (uniform) and this lane's part of the hash. A step that
 * returns -2 is evaluated again by fix(k). */
template <class StepF, class FixF>
__device__ __forceinline__ void walk_list(StepF step, FixF fix, int S, uint32_t carry, int64_t hoff, int64_t hcap,
                                          uint32_t* __restrict__ voxels, int lane, uint64_t lt_mask, int64_t& n_out,
                                          uint32_t& hsh_out) {
    if (lane == 0 && hcap > 0 && voxels) voxels[hoff] = carry;
    int64_t n = 1;
    uint32_t hsh = lane == 0 ? pct_list_mix(carry, 0) : 0u;
    for (int c = 0; c < S; c += 64) {
        int k = c + lane;
        int64_t vox = -1;
        if (k < S) {
            vox = step(k);
            if (vox == -2) vox = fix(k);
        }
        bool valid = vox >= 0;
        uint64_t vmask = __ballot(valid);
        uint32_t prev;
        if (vmask == ~0ull) {
            /* every step of the chunk inside the grid: lane l's previous
             * valid step is lane l - 1's (one lane shift), lane 0's the carry */
            prev = (uint32_t)__builtin_amdgcn_update_dpp((int)carry, (int)(uint32_t)vox, 0x138 /* wave_shr:1 */, 0xf,
                                                         0xf, false);
        } else {
            uint64_t below = vmask & lt_mask;
            int src = below ? 63 - __clzll(below) : 0;
            uint32_t pv = (uint32_t)__shfl((int)(uint32_t)vox, src, 64);
            prev = below ? pv : carry;
        }
        bool is_new = valid && (uint32_t)vox != prev;
        uint64_t nmask = __ballot(is_new);
        if (is_new) {
            int64_t pos = n + __popcll(nmask & lt_mask);
            if (pos < hcap && voxels) voxels[hoff + pos] = (uint32_t)vox;
            hsh += pct_list_mix((uint32_t)vox, (uint32_t)pos);
        }
        n += __popcll(nmask);
        if (vmask) carry = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)vox, 63 - __clzll(vmask));
    }
    n_out = n;
    hsh_out = hsh;
}

/* the list length (capped at the capacity); verification passes against a
 * DROP plan's hashes store no lists and no lengths */
__device__ __forceinline__ void store_length(uint32_t* __restrict__ lengths, int64_t at, int64_t n, int64_t hcap,
                                             int lane) {
    if (lengths && lane == 0) lengths[at] = (uint32_t)(n < hcap ? n : hcap);
}

/* capacity pass: steps[k] = S + 1 (upper bound of the path length) */
__global__ void __launch_bounds__(256)
k_mlp_steps(pct_geo g, pct_hist_soa h, int64_t start, int64_t count, const double* __restrict__ tab, int K,
            int32_t* __restrict__ steps) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    mlp_hist m;
    mlp_setup(g, h, start + k, tab + MLP_TAB_U1 * K, K, m);
    steps[k] = m.S + 1;
}

__device__ __forceinline__ float bcast_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

/* ------------------------------------------------------------------------- */
/* Precision form: find_MLP_path_GPU_tabulated (.cu:5502-5656).               */
/*                                                                            */
/* The sigma integrals and both 2x2 inversions stay FP64: the u_1-only terms  */
/* come from the per-depth table above, the sigma_2 scattering coefficient    */
/* (E_0 (1 + 0.038 ln(d / X_0)))^2 / X_0 from a per-depth table over the      */
/* remaining depth d = u_2 - u_1 at SCAT_SUB entries per MLP step (5e-5 cm at */
/* 1 mm voxels, the reference's DEPTH_TABLE_STEP), indexed like the reference */
/* (round((u_2 - u_1) / step) at the first step, then minus SCAT_SUB per      */
/* step, .cu:5552 and 5651). The 2x2 solve is reduced to the three            */
/* coefficients t_1 = alpha t_in + beta t_out + gamma (theta_out - theta_in)  */
/* (v_1 likewise), which need one reciprocal per inverse instead of the five  */
/* divisions of the literal expressions. The stepping -- t_1, v_1, the        */
/* rotation back to x/y and the voxel index -- runs in FP32.                  */
/*                                                                            */
/* The depth table is stored as [r][q] with j = q SCAT_SUB + r: the 64 lanes  */
/* of a wave, on 64 consecutive steps of one history, read 64 consecutive     */
/* doubles (one row r, columns q_0 - k).                                      */
/* ------------------------------------------------------------------------- */
#define SCAT_SUB 1000

struct mlp_thist {
    double u_2, u2sq, two_u2, pre1, pre2, pre3;
    float ca, sa, u_in, t_in, t_out, dT, v_in, v_out, dV;
    int q0, r;
};

__device__ __forceinline__ int calc_voxel_f(float zero, float pos, float inv_size) {
    float q = __builtin_fabsf(pos - zero) * inv_size;
    return q < 1.0e9f ? (int)q : -1;
}

/* floats of the precision form's voxel index */
struct grid_f {
    float x_zero, y_zero, z_zero, inv_w, inv_h, inv_t;
};

__device__ __forceinline__ void mlp_table_setup(const pct_geo& g, const mlp_hist& m, mlp_thist& t) {
    t.u_2 = m.u_2;
    t.u2sq = m.u_2 * m.u_2;
    t.two_u2 = 2 * m.u_2;
    t.pre1 = m.pre1; t.pre2 = m.pre2; t.pre3 = m.pre3;
    t.ca = (float)m.ca; t.sa = (float)m.sa; t.u_in = (float)m.u_in;
    t.t_in = (float)m.t_in; t.t_out = (float)m.t_out; t.dT = (float)m.dT;
    t.v_in = (float)m.v_in; t.v_out = (float)m.v_out; t.dV = (float)m.dV;
    double j0 = (m.u_2 - g.mlp_u_step) / (g.mlp_u_step / SCAT_SUB) + 0.5;
    int j = j0 > 0 ? (j0 < 2.0e9 ? (int)j0 : 2000000000) : 0;
    t.r = j % SCAT_SUB;
    t.q0 = j / SCAT_SUB;
}

__device__ __forceinline__ int64_t mlp_step_table(const pct_geo& g, const grid_f& gf, const mlp_thist& m,
                                                  const double* __restrict__ tab, int K, const double* __restrict__ scat,
                                                  int Q, int k) {
    double u_1 = tab[MLP_TAB_U1 * K + k];
    double S1I0 = tab[MLP_TAB_S1I0 * K + k];
    double S1I1 = tab[MLP_TAB_S1I1 * K + k];
    double S1I3 = tab[MLP_TAB_S1I3 * K + k];
    double cst1 = tab[MLP_TAB_C1 * K + k];
    double cst2 = tab[MLP_TAB_C2 * K + k];
    double cst3 = tab[MLP_TAB_C3 * K + k];
    int q = m.q0 - k;
    q = q < 0 ? 0 : (q >= Q ? Q - 1 : q);
    double c2 = scat[(int64_t)m.r * Q + q];
    double d = m.u_2 - u_1;
    double st2 = c2 * (m.pre1 - m.u2sq * cst1 + m.two_u2 * cst2 - cst3);
    double stt2 = c2 * (m.pre2 - m.u_2 * cst1 + cst2);
    double sth2 = c2 * (m.pre3 - cst1);
    double r2 = 1.0 / (st2 * sth2 - stt2 * stt2);
    double A = sth2 * r2, Bq = -stt2 * r2, C = st2 * r2;    /* Sigma_2^-1 = {A, Bq; Bq, C} */
    double e = A * d + Bq;
    double f = Bq * d + C;
    double F0 = S1I0 + A;
    double F1 = S1I1 + e;
    double F3 = S1I3 + d * e + f;
    double rF = 1.0 / (F0 * F3 - F1 * F1);
    double I0 = F3 * rF, I1 = -F1 * rF;
    float alpha = (float)(I0 * S1I0 + I1 * S1I1);
    float beta = (float)(I0 * A + I1 * e);
    float gamma = (float)(I0 * Bq + I1 * f);
    float t_1 = alpha * m.t_in + beta * m.t_out + gamma * m.dT;
    float v_1 = alpha * m.v_in + beta * m.v_out + gamma * m.dV;
    float us = m.u_in + (float)u_1;
    float x_1 = m.ca * us - m.sa * t_1;
    float y_1 = m.sa * us + m.ca * t_1;
    int vx = calc_voxel_f(gf.x_zero, x_1, gf.inv_w);
    int vy = calc_voxel_f(gf.y_zero, y_1, gf.inv_h);
    int vz = calc_voxel_f(gf.z_zero, v_1, gf.inv_t);
    int64_t voxel = (int64_t)vx + (int64_t)vy * g.C + (int64_t)vz * g.C * g.R;
    if (vx < 0 || vy < 0 || vz < 0 || voxel >= g.V) return -1;
    return voxel;
}

/* The two walks as policies of one paths kernel: setup of lane j's history,
 * its broadcast to the wave, and one step. */
struct exact_walk {
    const double* tab;
    int K;
    typedef mlp_hist H;
    __device__ void setup(const pct_geo& g, const pct_hist_soa& h, int64_t i, H& m, double& chord, int& S) const {
        mlp_setup(g, h, i, tab + MLP_TAB_U1 * K, K, m);
        chord = m.chord;
        S = m.S;
    }
    __device__ H bcast(const H& m, int j) const {
        H b;
        b.ca = bcast_d(m.ca, j); b.sa = bcast_d(m.sa, j); b.u_in = bcast_d(m.u_in, j);
        b.t_in = bcast_d(m.t_in, j); b.t_out = bcast_d(m.t_out, j); b.dT = bcast_d(m.dT, j);
        b.v_in = bcast_d(m.v_in, j); b.v_out = bcast_d(m.v_out, j); b.dV = bcast_d(m.dV, j);
        b.u_2 = bcast_d(m.u_2, j); b.pre1 = bcast_d(m.pre1, j); b.pre2 = bcast_d(m.pre2, j);
        b.pre3 = bcast_d(m.pre3, j);
        return b;
    }
    __device__ int64_t step(const pct_geo& g, const H& b, int k) const { return mlp_step_r(g, b, tab, K, k); }
    __device__ int64_t fix(const pct_geo& g, const H& b, int k) const { return mlp_step(g, b, tab, K, k); }
};

struct table_walk {
    const double* tab;
    int K;
    const double* scat;
    int Q;
    grid_f gf;
    typedef mlp_thist H;
    __device__ void setup(const pct_geo& g, const pct_hist_soa& h, int64_t i, H& t, double& chord, int& S) const {
        mlp_hist m;
        mlp_setup(g, h, i, tab + MLP_TAB_U1 * K, K, m);
        mlp_table_setup(g, m, t);
        chord = m.chord;
        S = m.S;
    }
    __device__ H bcast(const H& m, int j) const {
        H b;
        b.u_2 = bcast_d(m.u_2, j); b.u2sq = bcast_d(m.u2sq, j); b.two_u2 = bcast_d(m.two_u2, j);
        b.pre1 = bcast_d(m.pre1, j); b.pre2 = bcast_d(m.pre2, j); b.pre3 = bcast_d(m.pre3, j);
        b.ca = bcast_f(m.ca, j); b.sa = bcast_f(m.sa, j); b.u_in = bcast_f(m.u_in, j);
        b.t_in = bcast_f(m.t_in, j); b.t_out = bcast_f(m.t_out, j); b.dT = bcast_f(m.dT, j);
        b.v_in = bcast_f(m.v_in, j); b.v_out = bcast_f(m.v_out, j); b.dV = bcast_f(m.dV, j);
        b.q0 = bcast_i(m.q0, j); b.r = bcast_i(m.r, j);
        return b;
    }
    __device__ int64_t step(const pct_geo& g, const H& b, int k) const {
        return mlp_step_table(g, gf, b, tab, K, scat, Q, k);
    }
    __device__ int64_t fix(const pct_geo& g, const H& b, int k) const { return step(g, b, k); }
};

/* MLP paths, compiled for at least 5 wavefronts per SIMD (102 registers: the
 * exact walk keeps a few per-history values in scratch, outside its step
 * loop; 56.6 against 58.0 ms at 4 wavefronts over the 200x200x32 scan).
 * Each wavefront owns 64 consecutive histories: lane j computes the
 * per-history constants of history j, then the wave walks the histories one
 * by one with the 64 lanes on 64 consecutive steps. With LDS_TAB the
 * per-depth table (7 FP64 fields per step, read by every step of every
 * history) is staged in LDS once per workgroup, so that the steps' table
 * reads stay off the vector memory path that the block kernels running
 * beside the walk depend on. With IDX the launch walks histories
 * idx[0 .. *idx_count) of the range instead (the surrogate form's fallbacks),
 * each written where the direct launch would write it, in groups of G <= 64
 * per wavefront chosen from the list's length so that a short list spreads
 * over the waves instead of queueing in a few. */
template <class W, bool LDS_TAB, bool IDX = false>
__global__ void __launch_bounds__(256, 5)
k_mlp_paths(pct_geo g, pct_hist_soa h, int64_t start, int64_t count, W walk,
            const int64_t* __restrict__ offsets, uint32_t* __restrict__ voxels, uint32_t* __restrict__ lengths,
            double* __restrict__ chord_out, int* __restrict__ status, uint32_t* __restrict__ hash_out,
            const uint32_t* __restrict__ hash_expect, const uint32_t* __restrict__ idx,
            const unsigned long long* __restrict__ idx_count) {
    extern __shared__ double s_mlp_tab[];
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    int G = 64;
    if (IDX) {
        /* a list launch is sized for the whole range: the group size follows
         * the list's length, and workgroups past its end leave before
         * staging the table */
        count = (int64_t)*idx_count;
        const int64_t gsz = (count + waves - 1) / waves;
        G = gsz < 1 ? 1 : (gsz > 64 ? 64 : (int)gsz);
        if ((int64_t)blockIdx.x * (blockDim.x >> 6) * G >= count) return;
    }
    if (LDS_TAB) {
        const int n = MLP_TAB_FIELDS * walk.K;
        for (int i = threadIdx.x; i < n; i += blockDim.x) s_mlp_tab[i] = walk.tab[i];
        __syncthreads();
        walk.tab = s_mlp_tab;
    }
    const int K = walk.K;
    const int lane = threadIdx.x & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int local_flags = 0;
    /* grid-stride over groups of G histories: the grid is sized to the
     * resident waves, so a launch has no partial last round */
    for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * G; base < count;
         base += waves * G) {
    /* list position base + lane; its history (relative to start) is mine */
    const bool have = lane < G && base + lane < count;
    const int64_t mine = (IDX && have) ? (int64_t)idx[base + lane] : base + lane;
    typename W::H m;
    int mS = 0;
    uint32_t first = 0;
    int64_t off = 0, cap = 0;
    if (have) {
        double chord;
        walk.setup(g, h, start + mine, m, chord, mS);
        first = h.first_voxel[start + mine];
        /* a verification pass (no list store) needs no capacity */
        off = voxels ? offsets[mine] : 0;
        cap = voxels ? offsets[mine + 1] - off : ((int64_t)1 << 40);
        if (chord_out) chord_out[mine] = chord;
        if (mS >= K) atomicOr(status, PCT_STATUS_TABLE);
    }
    const int nh = (int)((count - base) < G ? (count - base) : G);
    for (int j = 0; j < nh; ++j) {
        typename W::H b = walk.bcast(m, j);
        const int64_t hj = IDX ? bcast_l(mine, j) : base + j;
        int S = bcast_i(mS, j);
        if (S > K) S = K;
        uint32_t carry = (uint32_t)bcast_i((int)first, j);
        int64_t hoff = bcast_l(off, j);
        int64_t hcap = bcast_l(cap, j);
        int64_t n;
        uint32_t hsh;
        walk_list([&](int k) { return walk.step(g, b, k); }, [&](int k) { return walk.fix(g, b, k); }, S, carry, hoff,
                  hcap, voxels, lane, lt_mask, n, hsh);
        if (n > hcap) local_flags |= PCT_STATUS_CAPACITY;
        if (n > g.max_intersections) local_flags |= PCT_STATUS_OVERFLOW;
        store_length(lengths, hj, n, hcap, lane);
        if (hash_out || hash_expect) {
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) hsh += (uint32_t)__shfl_xor((int)hsh, d, 64);
            hsh += (uint32_t)n * 0x27d4eb2fu;
            if (lane == 0) {
                if (hash_out) hash_out[hj] = hsh;
                if (hash_expect && hash_expect[hj] != hsh) local_flags |= PCT_STATUS_PLAN;
            }
        }
    }
    }
    if (lane == 0 && local_flags) atomicOr(status, local_flags);
}

/* The verification pass of a recomputing DROP iteration with one lane per
 * history: a wavefront takes 64 consecutive histories, its lanes in the
 * order of their step counts (lane_of, from the plan: lanes whose walks have
 * ended are then contiguous; any order that spreads a wave's histories over
 * more than these 64 took the set-up reads out of L2 and slowed the block
 * loop beside it), and walks them side by side, step k of all of them
 * together. The per-depth table row of step k is then the same in every
 * lane (one LDS address per step, broadcast reads), the dedup of
 * .cu:5478-5485 is a compare with the lane's
 * previous voxel, and no history's constants are broadcast. Nothing is
 * stored: each list's hash (the position-aware mixes of its elements plus
 * the length term, as walk_list forms it) is compared with the plan's. The
 * steps are mlp_step_r's expressions (walk.fix where it defers), so the
 * voxels are the lane-per-step walk's bit for bit. Only for the exact walk. */
template <class W>
__global__ void __launch_bounds__(256, 5)
k_mlp_verify(pct_geo g, pct_hist_soa h, int64_t start, int64_t count, W walk, const uint16_t* __restrict__ lane_of,
             const uint32_t* __restrict__ hash_expect, int* __restrict__ status) {
    /* the per-depth table staged in LDS one row per step (the 7 fields of
     * step k together: one address per step, the fields at fixed offsets) */
    extern __shared__ double s_rows[];
    const int K = walk.K;
    for (int i = threadIdx.x; i < MLP_TAB_FIELDS * K; i += blockDim.x) {
        const int f = i / K, k = i - f * K;
        s_rows[k * MLP_TAB_FIELDS + f] = walk.tab[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    int flags = 0;
    for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64; base < count;
         base += waves * 64) {
        const bool have = base + lane < count;
        typename W::H m;
        int S = 0;
        uint32_t prev = 0, hix = 0;
        if (have) {
            /* the wave's 64 histories, lanes in step-count order */
            hix = (uint32_t)(base + (lane_of ? lane_of[start + base + lane] : lane));
            double chord;
            walk.setup(g, h, start + hix, m, chord, S);
            prev = h.first_voxel[start + hix];
            if (S >= K) flags |= PCT_STATUS_TABLE;
            if (S > K) S = K;
        }
        int smax = S;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) smax = max(smax, __shfl_xor(smax, d, 64));
        uint32_t n = 1, hsh = pct_list_mix(prev, 0);
        for (int k = 0; k < smax; ++k) {
            if (k < S) {
                double row[MLP_TAB_FIELDS];
#pragma unroll
                for (int f = 0; f < MLP_TAB_FIELDS; ++f) row[f] = s_rows[k * MLP_TAB_FIELDS + f];
                int64_t vox = mlp_step_r_row(g, m, row);
                if (vox == -2) vox = walk.fix(g, m, k);
                if (vox >= 0) {
                    const uint32_t v = (uint32_t)vox;
                    if (v != prev) {
                        hsh += pct_list_mix(v, n);
                        ++n;
                    }
                    prev = v;
                }
            }
        }
        if (have) {
            hsh += n * 0x27d4eb2fu;
            if (hash_expect[hix] != hsh) flags |= PCT_STATUS_PLAN;
            if ((int64_t)n > g.max_intersections) flags |= PCT_STATUS_OVERFLOW;
        }
    }
    if (flags) atomicOr(status, flags);
}


/* ------------------------------------------------------------------------- */
/* Learned-path form (include/pct_net.h): the network's inputs, the FP32      */
/* surrogate path (cubic Hermite between the MLP's own boundary values plus   */
/* the clamped cubic spline of the predicted knot offsets, per segment of the */
/* normalised depth a cubic in tau whose coefficients are formed once per     */
/* history), the guard deciding between the surrogate and the exact walk, and */
/* the walk kernels.                                                          */
/* ------------------------------------------------------------------------- */
__device__ __forceinline__ void net_features(const mlp_hist& m, double xz_in, double xz_out, double wepl, float* f) {
    f[0] = (float)((m.t_out - m.t_in) * 2.0);
    f[1] = (float)((m.v_out - m.v_in) * 2.0);
    f[2] = (float)(m.dT * 20.0);
    f[3] = (float)(m.dV * 20.0);
    f[4] = (float)(m.u_2 * 0.125 - 1.5);
    f[5] = (float)(wepl * 0.125 - 1.5);
    f[6] = (float)(m.t_in * 0.1);
    f[7] = (float)(m.t_out * 0.1);
    f[8] = (float)(m.v_in * 0.25);
    f[9] = (float)(m.v_out * 0.25);
    f[10] = (float)(xz_in * 10.0);
    f[11] = (float)(xz_out * 10.0);
}

__device__ __forceinline__ int net_cell(const mlp_hist& m) {
    int ub = (int)(m.u_2 / 3.0);
    ub = ub > 7 ? 7 : (ub < 0 ? 0 : ub);
    double mag = pct_fabs(m.t_out - m.t_in);
    double a = pct_fabs(m.v_out - m.v_in), b = m.u_2 * pct_fabs(m.dT), c = m.u_2 * pct_fabs(m.dV);
    mag = a > mag ? a : mag;
    mag = b > mag ? b : mag;
    mag = c > mag ? c : mag;
    int mb = 0;
    double thr = 0.02;
    for (int k = 0; k < 7; ++k, thr *= 2)
        if (mag >= thr) mb = k + 1;
    return 8 * ub + mb;
}

__device__ __forceinline__ bool net_in_range(const float* f, const float* __restrict__ range) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < PCT_NET_IN; ++k) ok = ok && f[k] >= range[k] && f[k] <= range[PCT_NET_IN + k];
    return ok;
}

/* FP32 constants of one history's surrogate path */
struct sur_hist {
    float ca, sa, u_in, inv_u2;          /* rotation, entry depth, 1 / u_2             */
    float t_in, dlt, wt, v_in, dlv, wv;  /* t_in, t_out - t_in, u_2 dtheta_t; v alike  */
};

__device__ __forceinline__ sur_hist sur_setup(const mlp_hist& m) {
    sur_hist s;
    s.ca = (float)m.ca; s.sa = (float)m.sa; s.u_in = (float)m.u_in; s.inv_u2 = (float)(1.0 / m.u_2);
    s.t_in = (float)m.t_in; s.dlt = (float)(m.t_out - m.t_in); s.wt = (float)(m.u_2 * m.dT);
    s.v_in = (float)m.v_in; s.dlv = (float)(m.v_out - m.v_in); s.wv = (float)(m.u_2 * m.dV);
    return s;
}

__device__ __forceinline__ sur_hist sur_bcast(const sur_hist& s, int j) {
    sur_hist b;
    b.ca = bcast_f(s.ca, j); b.sa = bcast_f(s.sa, j); b.u_in = bcast_f(s.u_in, j); b.inv_u2 = bcast_f(s.inv_u2, j);
    b.t_in = bcast_f(s.t_in, j); b.dlt = bcast_f(s.dlt, j); b.wt = bcast_f(s.wt, j);
    b.v_in = bcast_f(s.v_in, j); b.dlv = bcast_f(s.dlv, j); b.wv = bcast_f(s.wv, j);
    return b;
}

/* coefficient q (0..39) of the per-segment cubics: t for q < 20, v from 20 on;
 * segment (q % 20) / 4, power q % 4; out8 are the network outputs */
__device__ __forceinline__ float sur_coef(int q, const sur_hist& s, const float* out8) {
    const bool isv = q >= 20;
    const int r = q - (isv ? 20 : 0), i = r >> 2, k = r & 3;
    const float base = isv ? s.v_in : s.t_in, dl = isv ? s.dlv : s.dlt, w = isv ? s.wv : s.wt;
    const float* o = out8 + (isv ? 4 : 0);
    float a = (k == 0 ? base : 0.0f) + dl * NET_H01[i][k];
    a = a + w * NET_H11[i][k];
#pragma unroll
    for (int j = 0; j < 4; ++j) a = a + (o[j] * (float)PCT_NET_OUT_SCALE) * NET_PHI[j][i][k];
    return a;
}

/* surrogate (t, v) at step k from the coefficients Q[40] */
__device__ __forceinline__ void sur_pos(const sur_hist& s, const float* Q, float u1, float& t, float& v) {
    float s5 = (u1 * s.inv_u2) * 5.0f;
    int seg = (int)s5;
    seg = seg > 4 ? 4 : seg;
    float tau = s5 - (float)seg;
    const float* qt = Q + 4 * seg;
    const float* qv = Q + 20 + 4 * seg;
    t = qt[0] + tau * (qt[1] + tau * (qt[2] + tau * qt[3]));
    v = qv[0] + tau * (qv[1] + tau * (qv[2] + tau * qv[3]));
}

__device__ __forceinline__ int64_t sur_step(const pct_geo& g, const grid_f& gf, const sur_hist& s, const float* Q,
                                            const float* __restrict__ u1f, int k) {
    float u1 = u1f[k], t, v;
    sur_pos(s, Q, u1, t, v);
    float us = s.u_in + u1;
    float x = s.ca * us - s.sa * t;
    float y = s.sa * us + s.ca * t;
    int vx = calc_voxel_f(gf.x_zero, x, gf.inv_w);
    int vy = calc_voxel_f(gf.y_zero, y, gf.inv_h);
    int vz = calc_voxel_f(gf.z_zero, v, gf.inv_t);
    int64_t voxel = (int64_t)vx + (int64_t)vy * g.C + (int64_t)vz * g.C * g.R;
    if (vx < 0 || vy < 0 || vz < 0 || voxel >= g.V) return -1;
    return voxel;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* per-wavefront LDS of the surrogate kernels */
struct sur_lds {
    uint16_t xs[64 * 16];     /* network inputs (bfloat16 bits) */
    float out[64 * 8];        /* network outputs                */
    float Q[40];              /* current history's cubics       */
};

/* the network for the whole wave when any lane needs it (inputs in L.xs,
 * outputs to L.out) */
__device__ __forceinline__ void sur_wave_infer(const pct_net_frag* __restrict__ nf, sur_lds& L, bool use) {
    if (__ballot(use)) pct_net_infer_wave(nf, L.xs, L.out);
    wave_lds_sync();
}

/* setup of the 64 histories of a wavefront: lane j's MLP constants and its
 * network inputs into LDS (sur_wave_infer then runs the network) */
__device__ __forceinline__ void sur_wave_prep(const pct_geo& g, const pct_hist_soa& h, int64_t i, bool have,
                                              const double* __restrict__ tab, int K, const float* __restrict__ range,
                                              const uint8_t* __restrict__ cells, bool use_cells, sur_lds& L,
                                              mlp_hist& m, bool& use, bool& in_range, int& cell) {
    const int lane = threadIdx.x & 63;
    float f[PCT_NET_IN];
#pragma unroll
    for (int k = 0; k < PCT_NET_IN; ++k) f[k] = 0.0f;
    use = in_range = false;
    cell = 0;
    if (have) {
        mlp_setup(g, h, i, tab + MLP_TAB_U1 * K, K, m);
        net_features(m, h.xz_entry_angle[i], h.xz_exit_angle[i], h.wepl[i], f);
        in_range = net_in_range(f, range);
        cell = net_cell(m);
        use = in_range && (!use_cells || cells[cell] == 0);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) L.xs[lane * 16 + k] = k < PCT_NET_IN ? (uint16_t)pct_f2bf(f[k]) : (uint16_t)0;
    wave_lds_sync();
}

/* Surrogate MLP paths: like k_mlp_paths, one wavefront per 64 histories; a
 * history the guard accepts is walked on its FP32 surrogate path, any other
 * is appended to the fallback list fb_idx (*fb_count entries, one atomic per
 * wavefront) for the exact walk's launch that follows (k_mlp_paths with the
 * list), so that this kernel holds no FP64 walk state. counts[0] +=
 * fallbacks, counts[1] += histories. */
/* at least four wavefronts per SIMD: the register budget (128) is set by the
 * walk; the network's pass over the wave's 64 histories, once per 64
 * histories, keeps a few values in scratch to fit it */
__global__ void __launch_bounds__(256, 4)
k_mlp_paths_sur(pct_geo g, pct_hist_soa h, int64_t start, int64_t count, const double* __restrict__ tab, int K,
                const float* __restrict__ u1f, grid_f gf, const pct_net_frag* __restrict__ nf,
                const float* __restrict__ range, const uint8_t* __restrict__ cells, const int64_t* __restrict__ offsets,
                uint32_t* __restrict__ voxels, uint32_t* __restrict__ lengths, double* __restrict__ chord_out,
                int* __restrict__ status, uint32_t* __restrict__ hash_out, const uint32_t* __restrict__ hash_expect,
                unsigned long long* __restrict__ counts, uint32_t* __restrict__ fb_idx,
                unsigned long long* __restrict__ fb_count) {
    __shared__ __align__(16) sur_lds lds[4];
    extern __shared__ float u1s[];          /* the K step depths, staged once per workgroup */
    for (int k = threadIdx.x; k < K; k += blockDim.x) u1s[k] = u1f[k];
    __syncthreads();
    sur_lds& L = lds[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int local_flags = 0;
    unsigned long long fb = 0, tot = 0;
    for (int64_t base = ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63)); base < count;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t mine = base + lane;
        bool use, in_range;
        int mS = 0;
        uint32_t first = 0;
        int64_t off = 0, cap = 0;
        sur_hist sh{};
        {
            /* the FP64 constants end before the network runs */
            mlp_hist m;
            int cell;
            sur_wave_prep(g, h, start + mine, mine < count, tab, K, range, cells, true, L, m, use, in_range, cell);
            if (mine < count) {
                mS = m.S;
                first = h.first_voxel[start + mine];
                off = voxels ? offsets[mine] : 0;
                cap = voxels ? offsets[mine + 1] - off : ((int64_t)1 << 40);
                if (chord_out) chord_out[mine] = m.chord;
                if (mS >= K) atomicOr(status, PCT_STATUS_TABLE);
                sh = sur_setup(m);
            }
        }
        sur_wave_infer(nf, L, use);
        const uint64_t umask = __ballot(use);
        const int nh = (int)((count - base) < 64 ? (count - base) : 64);
        tot += nh;
        const uint64_t fmask = __ballot(mine < count && !use);
        if (fmask) {
            unsigned long long at = 0;
            if (lane == 0) at = atomicAdd(fb_count, (unsigned long long)__popcll(fmask));
            at = (unsigned long long)bcast_l((int64_t)at, 0);
            if ((fmask >> lane) & 1) fb_idx[at + __popcll(fmask & lt_mask)] = (uint32_t)mine;
            fb += __popcll(fmask);
        }
        for (uint64_t rest = umask; rest; rest &= rest - 1) {
            const int j = __ffsll((unsigned long long)rest) - 1;
            int S = bcast_i(mS, j);
            if (S > K) S = K;
            uint32_t carry = (uint32_t)bcast_i((int)first, j);
            int64_t hoff = bcast_l(off, j);
            int64_t hcap = bcast_l(cap, j);
            int64_t n;
            uint32_t hsh;
            sur_hist b = sur_bcast(sh, j);
            if (lane < 40) L.Q[lane] = sur_coef(lane, b, L.out + 8 * j);
            wave_lds_sync();
            auto st = [&](int k) { return sur_step(g, gf, b, L.Q, u1s, k); };
            walk_list(st, st, S, carry, hoff, hcap, voxels, lane, lt_mask, n, hsh);
            wave_lds_sync();
            if (n > hcap) local_flags |= PCT_STATUS_CAPACITY;
            if (n > g.max_intersections) local_flags |= PCT_STATUS_OVERFLOW;
            store_length(lengths, base + j, n, hcap, lane);
            if (hash_out || hash_expect) {
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) hsh += (uint32_t)__shfl_xor((int)hsh, d, 64);
                hsh += (uint32_t)n * 0x27d4eb2fu;
                if (lane == 0) {
                    if (hash_out) hash_out[base + j] = hsh;
                    if (hash_expect && hash_expect[base + j] != hsh) local_flags |= PCT_STATUS_PLAN;
                }
            }
        }
    }
    if (lane == 0) {
        if (local_flags) atomicOr(status, local_flags);
        if (tot) {
            atomicAdd(&counts[0], fb);
            atomicAdd(&counts[1], tot);
        }
    }
}

/* Calibration check: sampled history q (index start + q stride) -- largest
 * |t| / |v| deviation of the surrogate path from the exact MLP over its steps,
 * whether the two voxel lists are equal, its cell and whether its inputs are
 * inside the trained range. */
__global__ void __launch_bounds__(256)
k_net_calib(pct_geo g, pct_hist_soa h, int64_t start, int64_t stride, int64_t nsamp, const double* __restrict__ tab,
            int K, const float* __restrict__ u1f, grid_f gf, const pct_net_frag* __restrict__ nf,
            const float* __restrict__ range, float* __restrict__ dev_out, int32_t* __restrict__ cell_out,
            uint8_t* __restrict__ mismatch_out, uint8_t* __restrict__ in_range_out) {
    __shared__ __align__(16) sur_lds lds[4];
    extern __shared__ float u1s[];
    for (int k = threadIdx.x; k < K; k += blockDim.x) u1s[k] = u1f[k];
    __syncthreads();
    sur_lds& L = lds[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int64_t base = ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63)); base < nsamp;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = base + lane;
        mlp_hist m;
        bool use, in_range;
        int cell;
        sur_wave_prep(g, h, start + q * stride, q < nsamp, tab, K, range, nullptr, false, L, m, use, in_range, cell);
        sur_wave_infer(nf, L, use);
        int mS = 0;
        uint32_t first = 0;
        sur_hist sh{};
        if (q < nsamp) {
            mS = m.S < K ? m.S : K;
            first = h.first_voxel[start + q * stride];
            sh = sur_setup(m);
        }
        const int nh = (int)((nsamp - base) < 64 ? (nsamp - base) : 64);
        for (int j = 0; j < nh; ++j) {
            const int S = bcast_i(mS, j);
            const uint32_t carry = (uint32_t)bcast_i((int)first, j);
            sur_hist b = sur_bcast(sh, j);
            exact_walk ew{tab, K};
            mlp_hist e = ew.bcast(m, j);
            if (lane < 40) L.Q[lane] = sur_coef(lane, b, L.out + 8 * j);
            wave_lds_sync();
            double dev = 0.0;
            for (int c = 0; c < S; c += 64) {
                const int k = c + lane;
                if (k < S) {
                    double u, te, ve;
                    mlp_tv(g, e, tab, K, k, u, te, ve);
                    float ts, vs;
                    sur_pos(b, L.Q, u1s[k], ts, vs);
                    double d1 = pct_fabs((double)ts - te), d2 = pct_fabs((double)vs - ve);
                    d1 = d1 > d2 ? d1 : d2;
                    dev = d1 > dev ? d1 : dev;
                }
            }
            for (int d = 32; d >= 1; d >>= 1) {
                double o = __shfl_xor(dev, d, 64);
                dev = o > dev ? o : dev;
            }
            int64_t n1, n2;
            uint32_t h1, h2;
            auto st1 = [&](int k) { return sur_step(g, gf, b, L.Q, u1s, k); };
            auto st2 = [&](int k) { return mlp_step(g, e, tab, K, k); };
            walk_list(st1, st1, S, carry, 0, 0, nullptr, lane, lt_mask, n1, h1);
            walk_list(st2, st2, S, carry, 0, 0, nullptr, lane, lt_mask, n2, h2);
            uint32_t hd = h1 - h2;
            for (int d = 32; d >= 1; d >>= 1) hd += (uint32_t)__shfl_xor((int)hd, d, 64);
            const int ir = bcast_i((int)in_range, j), cl = bcast_i(cell, j);
            if (lane == 0) {
                dev_out[base + j] = (float)dev;
                cell_out[base + j] = cl;
                mismatch_out[base + j] = (n1 != n2 || hd != 0) ? 1 : 0;
                in_range_out[base + j] = (uint8_t)ir;
            }
            wave_lds_sync();
        }
    }
}

/* training data: network inputs and the exact FP64 MLP offsets at the knots */
__global__ void __launch_bounds__(256)
k_net_features(pct_geo g, pct_hist_soa h, int64_t start, int64_t count, const double* __restrict__ tab, int K,
               float* __restrict__ features, float* __restrict__ targets) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const int64_t i = start + q;
    mlp_hist m;
    mlp_setup(g, h, i, tab + MLP_TAB_U1 * K, K, m);
    float f[PCT_NET_IN];
    net_features(m, h.xz_entry_angle[i], h.xz_exit_angle[i], h.wepl[i], f);
#pragma unroll
    for (int k = 0; k < PCT_NET_IN; ++k) features[q * PCT_NET_IN + k] = f[k];
    if (!targets) return;
    for (int j = 1; j <= 4; ++j) {
        double s = j / 5.0;
        double u_1 = m.u_2 * j / 5.0;
        double tt, vv;
        mlp_point_at(g, m, u_1, tt, vv);
        double h01 = 3 * s * s - 2 * s * s * s, h11 = s * s * s - s * s;
        double tH = m.t_in + (m.t_out - m.t_in) * h01 + m.u_2 * m.dT * h11;
        double vH = m.v_in + (m.v_out - m.v_in) * h01 + m.u_2 * m.dV * h11;
        targets[q * PCT_NET_OUT + (j - 1)] = (float)((tt - tH) / PCT_NET_OUT_SCALE);
        targets[q * PCT_NET_OUT + 4 + (j - 1)] = (float)((vv - vH) / PCT_NET_OUT_SCALE);
    }
}

}  // namespace

/* ------------------------------------------------------------------------- */
/* host side                                                                  */
/* ------------------------------------------------------------------------- */

/* Per-depth table: the reference's u_1-only expressions (.cu:5382-5402) at the
 * accumulated depths u_1 = w/2, w/2 + w/2, ... (repeated addition as in the
 * reference loop). Built once per context with the same operation order. */
static int ensure_mlp_table(pct_ctx* ctx) {
    if (ctx->d_mlp_tab) return PCT_OK;
    const pct_geo& g = ctx->g;
    /* a chord between two points of the image square is at most 2*sqrt(2)*R */
    int K = (int)(2.0 * 1.4142135623730951 * g.radius / g.mlp_u_step) + 8;
    std::vector<double> tab((size_t)MLP_TAB_FIELDS * K);
    double u_0 = 0.0, u_1 = g.mlp_u_step;
    for (int k = 0; k < K; ++k) {
        double l1 = g.e_0 * (1 + 0.038 * pct_log((u_1 - u_0) / g.x_0));
        double c1 = l1 * l1 / g.x_0;
        double sigma_t1 = c1 * (u_1 * u_1 * u_1 * (g.a0_3 + u_1 * (g.a1_12 + u_1 * (g.a2_30 + u_1 * (g.a3_60 + u_1 * (g.a4_105 + u_1 * g.a5_168))))));
        double sigma_t1_theta1 = c1 * (u_1 * u_1 * (g.a0_2 + u_1 * (g.a1_6 + u_1 * (g.a2_12 + u_1 * (g.a3_20 + u_1 * (g.a4_30 + u_1 * g.a5_42))))));
        double sigma_theta1 = c1 * (u_1 * (g.A0 + u_1 * (g.a1_2 + u_1 * (g.a2_3 + u_1 * (g.a3_4 + u_1 * (g.a4_5 + u_1 * g.a5_6))))));
        double det1 = sigma_t1 * sigma_theta1 - sigma_t1_theta1 * sigma_t1_theta1;
        tab[MLP_TAB_U1 * K + k] = u_1;
        tab[MLP_TAB_S1I0 * K + k] = sigma_theta1 / det1;
        tab[MLP_TAB_S1I1 * K + k] = -sigma_t1_theta1 / det1;
        tab[MLP_TAB_S1I3 * K + k] = sigma_t1 / det1;
        tab[MLP_TAB_C1 * K + k] = u_1 * (g.A0 + u_1 * (g.a1_2 + u_1 * (g.a2_3 + u_1 * (g.a3_4 + u_1 * (g.a4_5 + u_1 * g.a5_6)))));
        tab[MLP_TAB_C2 * K + k] = u_1 * u_1 * (g.a0_2 + u_1 * (g.a1_3 + u_1 * (g.a2_4 + u_1 * (g.a3_5 + u_1 * (g.a4_6 + u_1 * g.a5_7)))));
        tab[MLP_TAB_C3 * K + k] = u_1 * u_1 * u_1 * (g.a0_3 + u_1 * (g.a1_4 + u_1 * (g.a2_5 + u_1 * (g.a3_6 + u_1 * (g.a4_7 + u_1 * g.a5_8)))));
        u_1 += g.mlp_u_step;
    }
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * tab.size()) != hipSuccess) return pct_fail(ctx, PCT_E_NOMEM, "MLP table alloc");
    hipError_t e = hipMemcpy(d, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(d); return pct_fail(ctx, PCT_E_HIP, "MLP table copy: %s", hipGetErrorString(e)); }
    ctx->d_mlp_tab = d;
    ctx->mlp_tab_k = K;
    /* the u_1 column rounded to float, for the FP32 surrogate path */
    std::vector<float> u1f(K);
    for (int k = 0; k < K; ++k) u1f[k] = (float)tab[MLP_TAB_U1 * K + k];
    if (hipMalloc(&ctx->d_u1f, sizeof(float) * K) != hipSuccess) {
        ctx->d_u1f = nullptr;
        return pct_fail(ctx, PCT_E_NOMEM, "u_1 table alloc");
    }
    e = hipMemcpy(ctx->d_u1f, u1f.data(), sizeof(float) * K, hipMemcpyHostToDevice);
    if (e != hipSuccess) return pct_fail(ctx, PCT_E_HIP, "u_1 table copy: %s", hipGetErrorString(e));
    return PCT_OK;
}

/* Precision form: sigma_2 scattering coefficient at the remaining depth
 * j * step / SCAT_SUB (generate_scattering_coefficient_table, .cu:4830-4861),
 * stored as [r][q] with j = q SCAT_SUB + r. Q covers the per-depth table's
 * K steps plus two. */
static int ensure_scat_table(pct_ctx* ctx) {
    if (ctx->d_scat_tab) return PCT_OK;
    const pct_geo& g = ctx->g;
    const int Q = ctx->mlp_tab_k + 2;
    const double dstep = g.mlp_u_step / SCAT_SUB;
    std::vector<double> tab((size_t)SCAT_SUB * Q);
    for (int q = 0; q < Q; ++q)
        for (int r = 0; r < SCAT_SUB; ++r) {
            int64_t j = (int64_t)q * SCAT_SUB + r;
            double depth = (j > 0 ? j : 1) * dstep;
            double l = g.e_0 * (1 + 0.038 * pct_log(depth / g.x_0));
            tab[(size_t)r * Q + q] = l * l / g.x_0;
        }
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * tab.size()) != hipSuccess) return pct_fail(ctx, PCT_E_NOMEM, "depth table alloc");
    hipError_t e = hipMemcpy(d, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(d); return pct_fail(ctx, PCT_E_HIP, "depth table copy: %s", hipGetErrorString(e)); }
    ctx->d_scat_tab = d;
    ctx->scat_q = Q;
    return PCT_OK;
}

static grid_f make_grid_f(const pct_geo& g) {
    return grid_f{(float)g.x_zero, (float)g.y_zero, (float)g.z_zero, (float)(1.0 / g.voxel_width),
                  (float)(1.0 / g.voxel_height), (float)(1.0 / g.voxel_thickness)};
}

/* resident workgroups per CU of each walk kernel (the grid-stride launch size) */
static int mlp_occupancy(pct_ctx* ctx) {
    if (ctx->cus) return PCT_OK;
    PCT_HIP_CHECK(ctx, hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    PCT_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&ctx->mlp_occ[PCT_MLP_FP64],
                                                                    k_mlp_paths<exact_walk, false>, 256, 0));
    PCT_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&ctx->mlp_occ[PCT_MLP_TABLE],
                                                                    k_mlp_paths<table_walk, false>, 256, 0));
    PCT_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&ctx->mlp_occ[PCT_MLP_SURROGATE], k_mlp_paths_sur,
                                                                    256, sizeof(float) * ctx->mlp_tab_k));
    ctx->cus = std::max(1, ctx->cus);
    return PCT_OK;
}

/* the per-depth tables of both MLP forms and the walk kernels' occupancy,
 * made when the context is created (so that no later MLP call allocates) */
int pct_mlp_tables(pct_ctx* ctx) {
    int rc;
    if ((rc = ensure_mlp_table(ctx)) || (rc = ensure_scat_table(ctx))) return rc;
    return mlp_occupancy(ctx);
}

static int check_hist(pct_ctx* ctx, const pct_hist_soa* h, int64_t start, int64_t count, const char* who) {
    if (!ctx || !h) return pct_fail(ctx, PCT_E_INVALID, "%s: null argument", who);
    if (start < 0 || count < 0 || start + count > h->n)
        return pct_fail(ctx, PCT_E_INVALID, "%s: range [%lld, %lld) outside [0, %lld)", who, (long long)start,
                        (long long)(start + count), (long long)h->n);
    if (count == 0) return PCT_OK;                  /* no history: the arrays are not read */
    if (!h->x_entry || !h->y_entry || !h->z_entry || !h->x_exit || !h->y_exit || !h->z_exit || !h->xy_entry_angle ||
        !h->xz_entry_angle || !h->xy_exit_angle || !h->xz_exit_angle || !h->first_voxel)
        return pct_fail(ctx, PCT_E_INVALID, "%s: missing history field", who);
    return PCT_OK;
}

extern "C" int pct_mlp_endpoints(pct_ctx* ctx, const pct_hist_soa* hist, const uint32_t* hull_bits, uint8_t* intersected) {
    int rc = check_hist(ctx, hist, 0, hist ? hist->n : 0, "pct_mlp_endpoints");
    if (rc) return rc;
    if (hist->n == 0) return PCT_OK;
    if (!hull_bits || !intersected) return pct_fail(ctx, PCT_E_INVALID, "pct_mlp_endpoints: null hull/intersected");
    hipLaunchKernelGGL(k_mlp_endpoints, dim3(pct_grid(hist->n, 256)), dim3(256), 0, ctx->stream, ctx->g, *hist,
                       hull_bits, intersected);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

/* internal: capacity offsets into `offsets` (count+1), total path slots */
int pct_mlp_capacity_impl(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, int64_t* offsets,
                          int64_t* total) {
    int rc = ensure_mlp_table(ctx);
    if (rc) return rc;
    rc = pct_ensure(ctx, ctx->steps, sizeof(int32_t) * (count > 0 ? count : 1));
    if (rc) return rc;
    int32_t* steps = (int32_t*)ctx->steps.ptr;
    if (count > 0)
        hipLaunchKernelGGL(k_mlp_steps, dim3(pct_grid(count, 256)), dim3(256), 0, ctx->stream, ctx->g, *hist, start, count,
                           (const double*)ctx->d_mlp_tab, ctx->mlp_tab_k, steps);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return pct_exclusive_scan_i32(ctx, steps, count, offsets, total);
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p) {
    if (threadIdx.x == 0) *p = 0ull;
}

/* internal: the MLP lists of histories [start, start+count) on `stream`;
 * offsets/lengths/chord/hash arrays are indexed from `start` (entry 0 is
 * history start). The surrogate form lists its fallback histories in fb_idx
 * (room for count entries) and their number in *fb_count, then walks them
 * with the exact walk's kernel. */
int pct_mlp_paths_fb(pct_ctx* ctx, hipStream_t stream, const pct_hist_soa* hist, int64_t start, int64_t count,
                     const int64_t* offsets, uint32_t* voxels, uint32_t* lengths, double* chord, uint32_t* hash_out,
                     const uint32_t* hash_expect, int mode, bool beside, uint32_t* fb_idx,
                     unsigned long long* fb_count) {
    int rc = ensure_mlp_table(ctx);
    if (rc) return rc;
    if (!pct_mlp_mode_ok(ctx, mode))
        return pct_fail(ctx, PCT_E_INVALID, mode == PCT_MLP_SURROGATE ? "surrogate MLP mode without a network (pct_net_set)"
                                                                      : "unknown MLP mode %d", mode);
    if (mode == PCT_MLP_TABLE && (rc = ensure_scat_table(ctx))) return rc;
    if (count == 0) return PCT_OK;
    if ((rc = mlp_occupancy(ctx))) return rc;
    /* a workgroup of 256 is one wavefront on each of the CU's 4 SIMDs. Beside
     * a single scan's block loop the table and surrogate walks leave room
     * (wave slots and registers) for the latency-bound block kernels; the
     * exact walk keeps its full occupancy (measured at 200x200x32: the exact
     * walk, compiled for 5 wavefronts per SIMD, 61.7-62.1 ms per iteration at
     * 5 beside against 62.5-62.8 at 4 and 64.7 at 3; the table walk best at
     * 3, 55.9 ms against 62.5 at 2; the surrogate at 2, 54.0 ms against 54.7
     * at 3 and 68.9 at 1; the batched form's steps do better with the full
     * occupancy). 0: the walk's own occupancy. */
    static const int beside_default[3] = {0, 3, 2};
    /* workgroups per CU of a walk launch: the walk's occupancy, capped beside
     * the block loop or by the knob; for the table walks the per-depth table
     * goes to LDS when the workgroups the launch keeps on a CU fit in its
     * 160 KB beside their registers' limit */
    auto walk_shape = [&](int walk_mode, int& per_cu, bool& lds, size_t& shm) {
        per_cu = std::max(1, ctx->mlp_occ[walk_mode]);
        const int lim = ctx->tune_mlp_waves > 0 ? ctx->tune_mlp_waves : (beside ? beside_default[walk_mode] : 0);
        if (lim > 0 && lim < per_cu) per_cu = lim;
        lds = false;
        shm = 0;
        if (walk_mode == PCT_MLP_SURROGATE) return;
        const size_t tab_bytes = sizeof(double) * MLP_TAB_FIELDS * (size_t)ctx->mlp_tab_k;
        lds = tab_bytes <= PCT_MLP_LDS_TAB_MAX && tab_bytes * (size_t)per_cu <= PCT_LDS_PER_CU;
        if (!lds && tab_bytes <= PCT_MLP_LDS_TAB_MAX && PCT_LDS_PER_CU / tab_bytes >= 3) {
            /* a larger grid's table: one wavefront per SIMD less, table in LDS */
            per_cu = (int)(PCT_LDS_PER_CU / tab_bytes);
            lds = true;
        }
        shm = lds ? tab_bytes : 0;
    };
    auto exact_launch = [&](const uint32_t* idx, const unsigned long long* idx_count) {
        int per_cu;
        bool lds;
        size_t shm;
        walk_shape(PCT_MLP_FP64, per_cu, lds, shm);
        const unsigned tgrid = std::min(pct_grid(count, 256), (unsigned)(ctx->cus * per_cu));
        exact_walk w{(const double*)ctx->d_mlp_tab, ctx->mlp_tab_k};
        if (idx)
            hipLaunchKernelGGL((lds ? k_mlp_paths<exact_walk, true, true> : k_mlp_paths<exact_walk, false, true>),
                               dim3(tgrid), dim3(256), shm, stream, ctx->g, *hist, start, count, w, offsets, voxels,
                               lengths, chord, ctx->d_status, hash_out, hash_expect, idx, idx_count);
        else
            hipLaunchKernelGGL((lds ? k_mlp_paths<exact_walk, true> : k_mlp_paths<exact_walk, false>), dim3(tgrid),
                               dim3(256), shm, stream, ctx->g, *hist, start, count, w, offsets, voxels, lengths, chord,
                               ctx->d_status, hash_out, hash_expect, nullptr, nullptr);
    };
    if (mode == PCT_MLP_SURROGATE) {
        if (!fb_idx || !fb_count) return pct_fail(ctx, PCT_E_INVALID, "surrogate MLP paths without a fallback list");
        int per_cu;
        bool lds;
        size_t shm;
        walk_shape(PCT_MLP_SURROGATE, per_cu, lds, shm);
        const unsigned grid = std::min(pct_grid(count, 256), (unsigned)(ctx->cus * per_cu));
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, fb_count);
        hipLaunchKernelGGL(k_mlp_paths_sur, dim3(grid), dim3(256), sizeof(float) * ctx->mlp_tab_k, stream, ctx->g, *hist,
                           start, count,
                           (const double*)ctx->d_mlp_tab, ctx->mlp_tab_k, (const float*)ctx->d_u1f, make_grid_f(ctx->g),
                           (const pct_net_frag*)ctx->d_net, (const float*)ctx->d_net_range,
                           (const uint8_t*)ctx->d_net_cells, offsets, voxels, lengths, chord, ctx->d_status, hash_out,
                           hash_expect, ctx->d_net_counts, fb_idx, fb_count);
        exact_launch(fb_idx, fb_count);
    } else if (mode == PCT_MLP_FP64) {
        exact_launch(nullptr, nullptr);
    } else {
        int per_cu;
        bool lds;
        size_t shm;
        walk_shape(PCT_MLP_TABLE, per_cu, lds, shm);
        const unsigned tgrid = std::min(pct_grid(count, 256), (unsigned)(ctx->cus * per_cu));
        table_walk w{(const double*)ctx->d_mlp_tab, ctx->mlp_tab_k, (const double*)ctx->d_scat_tab, ctx->scat_q,
                     make_grid_f(ctx->g)};
        hipLaunchKernelGGL((lds ? k_mlp_paths<table_walk, true> : k_mlp_paths<table_walk, false>), dim3(tgrid),
                           dim3(256), shm, stream, ctx->g, *hist, start, count, w, offsets, voxels, lengths, chord,
                           ctx->d_status, hash_out, hash_expect, nullptr, nullptr);
    }
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

/* room for the surrogate form's fallback list of n histories in the
 * context (its count in the first 8 bytes, the indices from byte 256 on) */
int pct_mlp_fallback_reserve(pct_ctx* ctx, int64_t n) {
    return pct_ensure(ctx, ctx->net_fallback, 256 + sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
}

int pct_mlp_paths_on(pct_ctx* ctx, hipStream_t stream, const pct_hist_soa* hist, int64_t start, int64_t count,
                     const int64_t* offsets, uint32_t* voxels, uint32_t* lengths, double* chord, uint32_t* hash_out,
                     const uint32_t* hash_expect, int mode, bool beside) {
    uint32_t* fb_idx = nullptr;
    unsigned long long* fb_count = nullptr;
    if (mode == PCT_MLP_SURROGATE && count > 0) {
        int rc = pct_mlp_fallback_reserve(ctx, count);
        if (rc) return rc;
        fb_count = (unsigned long long*)ctx->net_fallback.ptr;
        fb_idx = (uint32_t*)((char*)ctx->net_fallback.ptr + 256);
    }
    return pct_mlp_paths_fb(ctx, stream, hist, start, count, offsets, voxels, lengths, chord, hash_out, hash_expect,
                            mode, beside, fb_idx, fb_count);
}

int pct_mlp_paths_impl(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, const int64_t* offsets,
                       uint32_t* voxels, uint32_t* lengths, double* chord, uint32_t* hash_out,
                       const uint32_t* hash_expect, int mode) {
    return pct_mlp_paths_on(ctx, ctx->stream, hist, start, count, offsets, voxels, lengths, chord, hash_out,
                            hash_expect, mode, false);
}

/* step count of history k (the key of the check's global order) and k */
__global__ void k_verify_keys(pct_geo g, pct_hist_soa h, int64_t count, const double* __restrict__ tab, int K,
                              uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    mlp_hist m;
    mlp_setup(g, h, k, tab + MLP_TAB_U1 * K, K, m);
    keys[k] = (uint32_t)(m.S < 0 ? 0 : (m.S > 0xffff ? 0xffff : m.S));
    vals[k] = (uint32_t)k;
}

/* the histories [0, count) in the order of their step counts (order: count
 * u32; keys and vals: count u32 each of workspace) */
int pct_mlp_step_order(pct_ctx* ctx, const pct_hist_soa* hist, int64_t count, uint32_t* keys, uint32_t* vals,
                       uint32_t* order) {
    int rc = mlp_occupancy(ctx);
    if (rc) return rc;
    if (count <= 0) return PCT_OK;
    hipLaunchKernelGGL(k_verify_keys, dim3(pct_grid(count, 256)), dim3(256), 0, ctx->stream, ctx->g, *hist, count,
                       (const double*)ctx->d_mlp_tab, ctx->mlp_tab_k, keys, vals);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    size_t tb = 0;
    PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(nullptr, tb, keys, keys, vals, order, (size_t)count, 0, 16,
                                                 ctx->stream));
    pct_buffer tmp, kout;
    auto free_both = [&]() {
        for (pct_buffer* b : {&tmp, &kout})
            if (b->ptr) {
                (void)hipFree(b->ptr);
                b->ptr = nullptr;
                b->bytes = 0;
            }
    };
    if ((rc = pct_ensure(ctx, tmp, tb > 0 ? tb : 1)) || (rc = pct_ensure(ctx, kout, sizeof(uint32_t) * (size_t)count))) {
        free_both();
        return rc;
    }
    hipError_t e = rocprim::radix_sort_pairs(tmp.ptr, tb, keys, (uint32_t*)kout.ptr, vals, order, (size_t)count, 0,
                                             16, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    free_both();
    if (e != hipSuccess) return pct_fail(ctx, PCT_E_HIP, "step order: %s", hipGetErrorString(e));
    return PCT_OK;
}

/* the check's inputs in step order: the fields the MLP set-up reads, the
 * first voxel and the plan's hash of history order[i] into slot i (dst:
 * 12 arrays of count u32, fields in pct_hist_soa order, then first voxel,
 * then hash) */
__global__ void k_gather_steps(pct_hist_soa h, const uint32_t* __restrict__ hash, const uint32_t* __restrict__ order,
                               int64_t count, uint32_t* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t k = order[i];
    const float* f[10] = {h.x_entry, h.y_entry, h.z_entry, h.x_exit, h.y_exit, h.z_exit,
                          h.xy_entry_angle, h.xz_entry_angle, h.xy_exit_angle, h.xz_exit_angle};
#pragma unroll
    for (int a = 0; a < 10; ++a) dst[a * count + i] = __float_as_uint(f[a][k]);
    dst[10 * count + i] = h.first_voxel[k];
    dst[11 * count + i] = hash[k];
}

int pct_mlp_verify_lanes(pct_ctx* ctx, hipStream_t stream, const pct_hist_soa* hist, int64_t start, int64_t count,
                         const uint16_t* lane_of, const uint32_t* hash_expect);

/* the verification pass over all histories in step order: gather, then
 * k_mlp_verify on the gathered copy (the inputs are read again every call,
 * so a history changed in place still fails its hash) */
int pct_mlp_verify_ordered(pct_ctx* ctx, hipStream_t stream, const pct_hist_soa* hist, const uint32_t* order,
                           const uint32_t* hash_expect, uint32_t* work) {
    const int64_t n = hist->n;
    if (n <= 0) return PCT_OK;
    hipLaunchKernelGGL(k_gather_steps, dim3(pct_grid(n, 256)), dim3(256), 0, stream, *hist, hash_expect, order, n,
                       work);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    pct_hist_soa s = *hist;
    float** fp[10] = {&s.x_entry, &s.y_entry, &s.z_entry, &s.x_exit, &s.y_exit, &s.z_exit,
                      &s.xy_entry_angle, &s.xz_entry_angle, &s.xy_exit_angle, &s.xz_exit_angle};
    for (int a = 0; a < 10; ++a) *fp[a] = (float*)(work + a * n);
    s.first_voxel = work + 10 * n;
    s.wepl = nullptr;
    s.bin_num = nullptr;
    return pct_mlp_verify_lanes(ctx, stream, &s, 0, n, nullptr, work + 11 * n);
}

/* lane_of for k_mlp_verify: per group of 64 consecutive histories (from
 * history 0; a launch's start is a multiple of 64), lane_of[g * 64 + r] is
 * the offset within the group of the history whose step count has rank r
 * (ties by offset): one wavefront per group, each lane counting the lanes
 * ranked before it */
__global__ void k_verify_lanes_order(pct_geo g, pct_hist_soa h, int64_t count, const double* __restrict__ tab, int K,
                                     uint16_t* __restrict__ lane_of) {
    const int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64;
    const int lane = threadIdx.x & 63;
    if (base >= count) return;
    const bool have = base + lane < count;
    int S = 0x7fffffff;
    if (have) {
        mlp_hist m;
        mlp_setup(g, h, base + lane, tab + MLP_TAB_U1 * K, K, m);
        S = m.S;
    }
    int rank = 0;
    for (int j = 0; j < 64; ++j) {
        const int sj = __shfl(S, j, 64);
        rank += (sj < S) || (sj == S && j < lane);
    }
    if (have) lane_of[base + rank] = (uint16_t)lane;
}

int pct_mlp_verify_order(pct_ctx* ctx, const pct_hist_soa* hist, int64_t count, uint16_t* lane_of) {
    int rc = mlp_occupancy(ctx);
    if (rc) return rc;
    if (count <= 0) return PCT_OK;
    const int64_t groups = (count + 63) / 64;
    hipLaunchKernelGGL(k_verify_lanes_order, dim3((unsigned)((groups + 3) / 4)), dim3(256), 0, ctx->stream, ctx->g,
                       *hist, count, (const double*)ctx->d_mlp_tab, ctx->mlp_tab_k, lane_of);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

/* the verification pass over histories [start, start + count) with one lane
 * per history (FP64 exact walk), on `stream`; hash_expect[i] is history
 * start + i's; lane_of (pct_mlp_verify_order of the whole scan) is used when
 * the range is the whole scan, else the lanes keep history order */
int pct_mlp_verify_lanes(pct_ctx* ctx, hipStream_t stream, const pct_hist_soa* hist, int64_t start, int64_t count,
                         const uint16_t* lane_of, const uint32_t* hash_expect) {
    int rc = mlp_occupancy(ctx);
    if (rc) return rc;
    if (count <= 0) return PCT_OK;
    const size_t shm = sizeof(double) * MLP_TAB_FIELDS * (size_t)ctx->mlp_tab_k;
    if (shm > PCT_MLP_LDS_TAB_MAX)   /* a table beyond LDS: the lane-per-step pass (no list store) */
        return pct_mlp_paths_on(ctx, stream, hist, start, count, nullptr, nullptr, nullptr, nullptr, nullptr,
                                hash_expect, PCT_MLP_FP64, true);
    int occ = 0;
    PCT_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_mlp_verify<exact_walk>, 256, shm));
    int per_cu = std::max(1, occ);
    if (ctx->tune_mlp_waves > 0 && ctx->tune_mlp_waves < per_cu) per_cu = ctx->tune_mlp_waves;
    const unsigned grid = std::min(pct_grid((count + 63) / 64 * 64, 256), (unsigned)(ctx->cus * per_cu));
    exact_walk w{(const double*)ctx->d_mlp_tab, ctx->mlp_tab_k};
    if (start != 0 || count != hist->n) lane_of = nullptr;
    hipLaunchKernelGGL(k_mlp_verify<exact_walk>, dim3(grid), dim3(256), shm, stream, ctx->g, *hist, start, count, w,
                       lane_of, hash_expect, ctx->d_status);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

int pct_status_to_rc(pct_ctx* ctx, int flags) {
    if (flags & PCT_STATUS_PLAN)
        return pct_fail(ctx, PCT_E_INVALID, "recomputed MLP voxel lists differ from the DROP plan (inputs changed?)");
    if (flags & PCT_STATUS_WAIT)
        return pct_fail(ctx, PCT_E_HIP,
                        "DROP block loop: a bounded wait (the waiting one-launch form's wait for the previous block's "
                        "stored updates, or a persistent batch's grid barrier) gave up (x unchanged)");
    if (flags & PCT_STATUS_TABLE) return pct_fail(ctx, PCT_E_OVERFLOW, "MLP chord longer than the per-depth table");
    if (flags & PCT_STATUS_CAPACITY) return pct_fail(ctx, PCT_E_OVERFLOW, "MLP path longer than its capacity");
    if (flags & PCT_STATUS_OVERFLOW)
        return pct_fail(ctx, PCT_E_OVERFLOW, "MLP path longer than max_intersections (%d)", ctx->g.max_intersections);
    return PCT_OK;
}

extern "C" int pct_mlp_capacity(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, int64_t* offsets,
                                int64_t* total) {
    int rc = check_hist(ctx, hist, start, count, "pct_mlp_capacity");
    if (rc) return rc;
    if (!offsets) return pct_fail(ctx, PCT_E_INVALID, "pct_mlp_capacity: null offsets");
    return pct_mlp_capacity_impl(ctx, hist, start, count, offsets, total);
}

extern "C" int pct_mlp_paths(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, const int64_t* offsets,
                             uint32_t* voxels, uint32_t* lengths, int mode) {
    int rc = check_hist(ctx, hist, start, count, "pct_mlp_paths");
    if (rc) return rc;
    if (!pct_mlp_mode_ok(ctx, mode)) return pct_fail(ctx, PCT_E_INVALID, "pct_mlp_paths: mode %d not available", mode);
    if (count == 0) return PCT_OK;
    if (!offsets || !voxels || !lengths) return pct_fail(ctx, PCT_E_INVALID, "pct_mlp_paths: null output");
    rc = pct_mlp_paths_impl(ctx, hist, start, count, offsets, voxels, lengths, nullptr, nullptr, nullptr, mode);
    if (rc) return rc;
    int flags = 0;
    rc = pct_take_status(ctx, &flags);
    if (rc) return rc;
    return pct_status_to_rc(ctx, flags);
}

/* ------------------------------------------------------------------------- */
/* learned-path form: training data and the calibration check                 */
/* ------------------------------------------------------------------------- */
extern "C" int pct_net_features(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, float* features,
                                float* targets) {
    int rc = check_hist(ctx, hist, start, count, "pct_net_features");
    if (rc) return rc;
    if (count == 0) return PCT_OK;
    if (!features || !hist->wepl) return pct_fail(ctx, PCT_E_INVALID, "pct_net_features: null features/wepl");
    if ((rc = ensure_mlp_table(ctx))) return rc;
    hipLaunchKernelGGL(k_net_features, dim3(pct_grid(count, 256)), dim3(256), 0, ctx->stream, ctx->g, *hist, start,
                       count, (const double*)ctx->d_mlp_tab, ctx->mlp_tab_k, features, targets);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_net_calibrate(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count, int64_t stride,
                                 float tol, int32_t min_samples, pct_net_calib* out) {
    return pct_net_calibrate_rule(ctx, hist, start, count, stride, PCT_NET_RULE_DEVIATION, tol > 0.0f ? tol : -1.0f,
                                  min_samples, out);
}

extern "C" int pct_net_calibrate_rule(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count,
                                      int64_t stride, int rule, float threshold, int32_t min_samples,
                                      pct_net_calib* out) {
    int rc = check_hist(ctx, hist, start, count, "pct_net_calibrate");
    if (rc) return rc;
    if (!out || stride <= 0 || !hist->wepl || (rule != PCT_NET_RULE_DEVIATION && rule != PCT_NET_RULE_LISTS))
        return pct_fail(ctx, PCT_E_INVALID, "pct_net_calibrate: bad argument");
    if (!ctx->net_version) return pct_fail(ctx, PCT_E_INVALID, "pct_net_calibrate: no network installed");
    if ((rc = ensure_mlp_table(ctx)) || (rc = mlp_occupancy(ctx))) return rc;
    const int64_t nsamp = (count + stride - 1) / stride;
    *out = pct_net_calib{};
    if (nsamp == 0) return PCT_OK;
    const size_t per = sizeof(float) + sizeof(int32_t) + 2;
    if ((rc = pct_ensure(ctx, ctx->net_calib, per * (size_t)nsamp + 64))) return rc;
    char* base = (char*)ctx->net_calib.ptr;
    float* dev = (float*)base;
    int32_t* cell = (int32_t*)(base + sizeof(float) * nsamp);
    uint8_t* mis = (uint8_t*)(cell + nsamp);
    uint8_t* inr = mis + nsamp;
    const unsigned grid = std::min(pct_grid(nsamp, 256), (unsigned)(ctx->cus * std::max(1, ctx->mlp_occ[PCT_MLP_SURROGATE])));
    hipLaunchKernelGGL(k_net_calib, dim3(grid), dim3(256), sizeof(float) * ctx->mlp_tab_k, ctx->stream, ctx->g, *hist,
                       start, stride, nsamp,
                       (const double*)ctx->d_mlp_tab, ctx->mlp_tab_k, (const float*)ctx->d_u1f, make_grid_f(ctx->g),
                       (const pct_net_frag*)ctx->d_net, (const float*)ctx->d_net_range, dev, cell, mis, inr);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    std::vector<float> hdev(nsamp);
    std::vector<int32_t> hcell(nsamp);
    std::vector<uint8_t> hmis(nsamp), hinr(nsamp);
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(hdev.data(), dev, sizeof(float) * nsamp, hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(hcell.data(), cell, sizeof(int32_t) * nsamp, hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(hmis.data(), mis, nsamp, hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipMemcpyAsync(hinr.data(), inr, nsamp, hipMemcpyDeviceToHost, ctx->stream));
    PCT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    /* statistics over the samples whose inputs are inside the trained range
     * (the others take the exact walk whatever their cell) */
    std::vector<std::vector<float>> per_cell(PCT_NET_CELLS);
    std::vector<int64_t> cell_mis(PCT_NET_CELLS, 0);
    std::vector<float> all;
    out->samples = nsamp;
    for (int64_t q = 0; q < nsamp; ++q) {
        out->list_mismatch += hmis[q];
        if (!hinr[q]) continue;
        out->in_range++;
        all.push_back(hdev[q]);
        int c = hcell[q] < 0 ? 0 : (hcell[q] >= PCT_NET_CELLS ? PCT_NET_CELLS - 1 : hcell[q]);
        per_cell[c].push_back(hdev[q]);
        cell_mis[c] += hmis[q];
    }
    auto quant = [](std::vector<float>& v, double p) -> double {
        if (v.empty()) return 0.0;
        size_t k = (size_t)std::min<double>((double)v.size() - 1, std::floor(p * (double)(v.size() - 1) + 0.5));
        std::nth_element(v.begin(), v.begin() + k, v.end());
        return v[k];
    };
    out->dev_median = quant(all, 0.5);
    out->dev_p99 = quant(all, 0.99);
    out->dev_max = all.empty() ? 0.0 : *std::max_element(all.begin(), all.end());
    uint8_t flags[PCT_NET_CELLS];
    for (int c = 0; c < PCT_NET_CELLS; ++c) {
        const int64_t ns = (int64_t)per_cell[c].size();
        out->cell_samples[c] = (int32_t)ns;
        out->cell_p99[c] = (float)quant(per_cell[c], 0.99);
        out->cell_list_mismatch[c] = ns > 0 ? (float)((double)cell_mis[c] / (double)ns) : 0.0f;
        const float score = rule == PCT_NET_RULE_LISTS ? out->cell_list_mismatch[c] : out->cell_p99[c];
        bool fb = ns < min_samples || score > threshold;
        out->cell_fallback[c] = fb ? 1 : 0;
        flags[c] = out->cell_fallback[c];
        out->cells_fallback += fb ? 1 : 0;
    }
    if (threshold >= 0.0f) {
        PCT_HIP_CHECK(ctx, hipMemcpy(ctx->d_net_cells, flags, sizeof flags, hipMemcpyHostToDevice));
        static int calib_version = 1 << 20;
        ctx->net_version = ++calib_version;     /* plans built before the calibration are stale */
    } else {
        /* measurement only: report the flags already in force */
        PCT_HIP_CHECK(ctx, hipMemcpy(flags, ctx->d_net_cells, sizeof flags, hipMemcpyDeviceToHost));
        out->cells_fallback = 0;
        for (int c = 0; c < PCT_NET_CELLS; ++c) {
            out->cell_fallback[c] = flags[c];
            out->cells_fallback += flags[c] ? 1 : 0;
        }
    }
    for (int64_t q = 0; q < nsamp; ++q) {
        int c = hcell[q] < 0 ? 0 : (hcell[q] >= PCT_NET_CELLS ? PCT_NET_CELLS - 1 : hcell[q]);
        if (!hinr[q] || flags[c]) continue;
        out->surrogate++;
        out->surrogate_mismatch += hmis[q];
    }
    return PCT_OK;
}

extern "C" int pct_chord_lengths(pct_ctx* ctx, const pct_hist_soa* hist, int float_average, double* out) {
    if (!ctx || !hist) return pct_fail(ctx, PCT_E_INVALID, "pct_chord_lengths: null argument");
    if (hist->n < 0) return pct_fail(ctx, PCT_E_INVALID, "pct_chord_lengths: negative history count");
    if (hist->n == 0) return PCT_OK;
    if (!out || !hist->xy_entry_angle || !hist->xy_exit_angle || !hist->xz_entry_angle || !hist->xz_exit_angle)
        return pct_fail(ctx, PCT_E_INVALID, "pct_chord_lengths: missing angle or output array");
    hipLaunchKernelGGL(k_chord_lengths, dim3(pct_grid(hist->n, 256)), dim3(256), 0, ctx->stream, ctx->g, *hist,
                       float_average, out);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}
