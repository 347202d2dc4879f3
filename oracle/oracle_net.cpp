/*
 * oracle_net.cpp -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
 *
 * CPU restatement of the learned-path form's arithmetic (the network itself is
 * restated in numpy, oracle/net.py):
 *   - the exact FP64 MLP position at an arbitrary depth u_1, with the
 *     expressions of find_MLP_path's step (.cu:4439-4633, device twin
 *     .cu:5376-5428), as the training targets need it between the steps;
 *   - the network inputs and the training targets (offsets of the exact MLP
 *     from the cubic Hermite path at sigma = 1/5 .. 4/5), as include/pct_net.h
 *     defines them;
 *   - the FP32 surrogate walk: Hermite + clamped-spline path on the
 *     reference's u_1 grid, voxelised like the precision form, for given
 *     network outputs (tables of pct_amd/csrc/pct_net_basis.h, which a test
 *     pins against the exact rational spline);
 *   - the calibration cell and trained-range test of a history.
 * Compiled with -ffp-contract=off: the operation sequence is the source's.
 */
#include <math.h>
#include <stdint.h>
#include <vector>
#include "oracle_common.h"
#include "../include/pct_net.h"
#include "../pct_amd/csrc/pct_net_basis.h"

using namespace oracle;

namespace {

/* per-history quantities of the MLP after the u_in > u_out flip (.cu:4468-4500) */
struct Hist {
    double ca, sa, u_in, t_in, t_out, v_in, v_out, dT, dV, u_2, pre1, pre2, pre3;
    double xz_in, xz_out, wepl;
};

void setup(const Geo& g, const pct_hist_soa* h, int64_t i, Hist& m) {
    double x_in = h->x_entry[i], y_in = h->y_entry[i], z_in = h->z_entry[i];
    double x_out = h->x_exit[i], y_out = h->y_exit[i], z_out = h->z_exit[i];
    double xy_in = h->xy_entry_angle[i], xy_out = h->xy_exit_angle[i];
    double ca = pct_cos(xy_in), sa = pct_sin(xy_in);
    double u_in = (ca * x_in) + (sa * y_in), u_out = (ca * x_out) + (sa * y_out);
    double t_in = (ca * y_in) - (sa * x_in), t_out = (ca * y_out) - (sa * x_out);
    if (u_in > u_out) {
        xy_in += PCT_PI;
        xy_out += PCT_PI;
        ca = pct_cos(xy_in); sa = pct_sin(xy_in);
        u_in = (ca * x_in) + (sa * y_in); u_out = (ca * x_out) + (sa * y_out);
        t_in = (ca * y_in) - (sa * x_in); t_out = (ca * y_out) - (sa * x_out);
    }
    m.ca = ca; m.sa = sa; m.u_in = u_in;
    m.t_in = t_in; m.t_out = t_out; m.v_in = z_in; m.v_out = z_out;
    m.dT = xy_out - xy_in;
    m.dV = (double)h->xz_exit_angle[i] - (double)h->xz_entry_angle[i];
    double u_2 = pct_fabs(u_out - u_in);
    m.u_2 = u_2;
    m.pre1 = u_2 * u_2 * u_2 * (g.a0_3 + u_2 * (g.a1_12 + u_2 * (g.a2_30 + u_2 * (g.a3_60 + u_2 * (g.a4_105 + u_2 * g.a5_168)))));
    m.pre2 = u_2 * u_2 * (g.a0_2 + u_2 * (g.a1_6 + u_2 * (g.a2_12 + u_2 * (g.a3_20 + u_2 * (g.a4_30 + u_2 * g.a5_42)))));
    m.pre3 = u_2 * (g.A0 + u_2 * (g.a1_2 + u_2 * (g.a2_3 + u_2 * (g.a3_4 + u_2 * (g.a4_5 + u_2 * g.a5_6)))));
    m.xz_in = h->xz_entry_angle[i];
    m.xz_out = h->xz_exit_angle[i];
    m.wepl = h->wepl ? (double)h->wepl[i] : 0.0;
}

/* exact MLP (t_1, v_1) at depth u_1: the step body of .cu:4506-4609 */
void point(const Geo& g, const Hist& m, double u_1, double* t_out, double* v_out) {
    const double u_0 = 0.0, u_2 = m.u_2;
    double l1 = g.e_0 * (1 + 0.038 * pct_log((u_1 - u_0) / g.x_0));
    double c1 = l1 * l1 / g.x_0;
    double sigma_t1 = c1 * (u_1 * u_1 * u_1 * (g.a0_3 + u_1 * (g.a1_12 + u_1 * (g.a2_30 + u_1 * (g.a3_60 + u_1 * (g.a4_105 + u_1 * g.a5_168))))));
    double sigma_t1_theta1 = c1 * (u_1 * u_1 * (g.a0_2 + u_1 * (g.a1_6 + u_1 * (g.a2_12 + u_1 * (g.a3_20 + u_1 * (g.a4_30 + u_1 * g.a5_42))))));
    double sigma_theta1 = c1 * (u_1 * (g.A0 + u_1 * (g.a1_2 + u_1 * (g.a2_3 + u_1 * (g.a3_4 + u_1 * (g.a4_5 + u_1 * g.a5_6))))));
    double det1 = sigma_t1 * sigma_theta1 - sigma_t1_theta1 * sigma_t1_theta1;
    double S1[4] = {sigma_theta1 / det1, -sigma_t1_theta1 / det1, -sigma_t1_theta1 / det1, sigma_t1 / det1};

    double l2 = g.e_0 * (1 + 0.038 * pct_log((u_2 - u_1) / g.x_0));
    double c2 = l2 * l2 / g.x_0;
    double k1 = u_1 * (g.A0 + u_1 * (g.a1_2 + u_1 * (g.a2_3 + u_1 * (g.a3_4 + u_1 * (g.a4_5 + u_1 * g.a5_6)))));
    double k2 = u_1 * u_1 * (g.a0_2 + u_1 * (g.a1_3 + u_1 * (g.a2_4 + u_1 * (g.a3_5 + u_1 * (g.a4_6 + u_1 * g.a5_7)))));
    double k3 = u_1 * u_1 * u_1 * (g.a0_3 + u_1 * (g.a1_4 + u_1 * (g.a2_5 + u_1 * (g.a3_6 + u_1 * (g.a4_7 + u_1 * g.a5_8)))));
    double sigma_t2 = c2 * (m.pre1 - u_2 * u_2 * k1 + 2 * u_2 * k2 - k3);
    double sigma_t2_theta2 = c2 * (m.pre2 - u_2 * k1 + k2);
    double sigma_theta2 = c2 * (m.pre3 - k1);
    double det2 = sigma_t2 * sigma_theta2 - sigma_t2_theta2 * sigma_t2_theta2;
    double S2[4] = {sigma_theta2 / det2, -sigma_t2_theta2 / det2, -sigma_t2_theta2 / det2, sigma_t2 / det2};

    double R0[4] = {1.0, u_1, 0.0, 1.0};
    double R1[4] = {1.0, u_2 - u_1, 0.0, 1.0};
    double R1T[4] = {1.0, 0.0, u_2 - u_1, 1.0};
    double c13_1 = S2[0] * R1[0] + S2[1] * R1[2];
    double c13_2 = S2[2] * R1[0] + S2[3] * R1[2];
    double c24_1 = S2[0] * R1[1] + S2[1] * R1[3];
    double c24_2 = S2[2] * R1[1] + S2[3] * R1[3];
    double F[4];
    F[0] = S1[0] + R1T[0] * c13_1 + R1T[1] * c13_2;
    F[1] = S1[1] + R1T[0] * c24_1 + R1T[1] * c24_2;
    F[2] = S1[2] + R1T[2] * c13_1 + R1T[3] * c13_2;
    F[3] = S1[3] + R1T[2] * c24_1 + R1T[3] * c24_2;
    double detF = F[0] * F[3] - F[1] * F[2];
    double tmp = F[0];
    F[0] = F[3] / detF;
    F[1] = -F[1] / detF;
    F[2] = -F[2] / detF;
    F[3] = tmp / detF;
    double T0[2] = {m.t_in, 0.0}, T2[2] = {m.t_out, m.dT};
    double V0[2] = {m.v_in, 0.0}, V2[2] = {m.v_out, m.dV};
    const double* Y0[2] = {T0, V0};
    const double* Y2[2] = {T2, V2};
    double res[2];
    for (int c = 0; c < 2; ++c) {
        double s1 = R0[0] * Y0[c][0] + R0[1] * Y0[c][1];
        double s2 = R0[2] * Y0[c][0] + R0[3] * Y0[c][1];
        double s3 = S2[0] * Y2[c][0] + S2[1] * Y2[c][1];
        double s4 = S2[2] * Y2[c][0] + S2[3] * Y2[c][1];
        double st0 = S1[0] * s1 + S1[1] * s2 + R1T[0] * s3 + R1T[1] * s4;
        double st1 = S1[2] * s1 + S1[3] * s2 + R1T[2] * s3 + R1T[3] * s4;
        res[c] = F[0] * st0 + F[1] * st1;
    }
    *t_out = res[0];
    *v_out = res[1];
}

void features(const Hist& m, float* f) {
    f[0] = (float)((m.t_out - m.t_in) * 2.0);
    f[1] = (float)((m.v_out - m.v_in) * 2.0);
    f[2] = (float)(m.dT * 20.0);
    f[3] = (float)(m.dV * 20.0);
    f[4] = (float)(m.u_2 * 0.125 - 1.5);
    f[5] = (float)(m.wepl * 0.125 - 1.5);
    f[6] = (float)(m.t_in * 0.1);
    f[7] = (float)(m.t_out * 0.1);
    f[8] = (float)(m.v_in * 0.25);
    f[9] = (float)(m.v_out * 0.25);
    f[10] = (float)(m.xz_in * 10.0);
    f[11] = (float)(m.xz_out * 10.0);
}

int cell_of(const Hist& m) {
    int ub = (int)(m.u_2 / 3.0);
    if (ub > 7) ub = 7;
    if (ub < 0) ub = 0;
    double mag = pct_fabs(m.t_out - m.t_in);
    double a = pct_fabs(m.v_out - m.v_in), b = m.u_2 * pct_fabs(m.dT), c = m.u_2 * pct_fabs(m.dV);
    if (a > mag) mag = a;
    if (b > mag) mag = b;
    if (c > mag) mag = c;
    int mb = 0;
    double thr = 0.02;
    for (int k = 0; k < 7; ++k, thr *= 2)
        if (mag >= thr) mb = k + 1;
    return 8 * ub + mb;
}

inline int calculate_voxel_f(float zero, float pos, float inv_size) {
    float q = fabsf(pos - zero) * inv_size;
    return q < 1.0e9f ? (int)q : -1;
}

/* per-segment cubic coefficients of the surrogate path (t and v) */
void path_coefficients(const Hist& m, const float* out8, float Qt[5][4], float Qv[5][4]) {
    const float t_in = (float)m.t_in, dlt = (float)(m.t_out - m.t_in), wt = (float)(m.u_2 * m.dT);
    const float v_in = (float)m.v_in, dlv = (float)(m.v_out - m.v_in), wv = (float)(m.u_2 * m.dV);
    float dt[4], dv[4];
    for (int j = 0; j < 4; ++j) {
        dt[j] = out8[j] * (float)PCT_NET_OUT_SCALE;
        dv[j] = out8[4 + j] * (float)PCT_NET_OUT_SCALE;
    }
    for (int i = 0; i < 5; ++i)
        for (int k = 0; k < 4; ++k) {
            float a = (k == 0 ? t_in : 0.0f) + dlt * NET_H01[i][k];
            a = a + wt * NET_H11[i][k];
            float b = (k == 0 ? v_in : 0.0f) + dlv * NET_H01[i][k];
            b = b + wv * NET_H11[i][k];
            for (int j = 0; j < 4; ++j) {
                a = a + dt[j] * NET_PHI[j][i][k];
                b = b + dv[j] * NET_PHI[j][i][k];
            }
            Qt[i][k] = a;
            Qv[i][k] = b;
        }
}

}  // namespace

extern "C" {

/* exact MLP positions of history i at the n depths u1[] */
int oracle_mlp_points(const pct_params* p, const pct_hist_soa* h, int64_t i, int n, const double* u1, double* t,
                      double* v) {
    Geo g = make_geo(*p);
    Hist m;
    setup(g, h, i, m);
    for (int k = 0; k < n; ++k) point(g, m, u1[k], t + k, v + k);
    return 0;
}

/* features[n][12], targets[n][8] (targets may be NULL), cells[n] and
 * u_2[n] (both may be NULL) of histories [0, n) */
int oracle_net_features(const pct_params* p, const pct_hist_soa* h, int64_t n, float* feat, float* targ, int32_t* cells,
                        double* u2) {
    Geo g = make_geo(*p);
    for (int64_t i = 0; i < n; ++i) {
        Hist m;
        setup(g, h, i, m);
        features(m, feat + i * PCT_NET_IN);
        if (cells) cells[i] = cell_of(m);
        if (u2) u2[i] = m.u_2;
        if (!targ) continue;
        for (int j = 1; j <= 4; ++j) {
            double s = j / 5.0;
            double u_1 = m.u_2 * j / 5.0;
            double tt, vv;
            point(g, m, u_1, &tt, &vv);
            double h01 = 3 * s * s - 2 * s * s * s, h11 = s * s * s - s * s;
            double tH = m.t_in + (m.t_out - m.t_in) * h01 + m.u_2 * m.dT * h11;
            double vH = m.v_in + (m.v_out - m.v_in) * h01 + m.u_2 * m.dV * h11;
            targ[i * PCT_NET_OUT + (j - 1)] = (float)((tt - tH) / PCT_NET_OUT_SCALE);
            targ[i * PCT_NET_OUT + 4 + (j - 1)] = (float)((vv - vH) / PCT_NET_OUT_SCALE);
        }
    }
    return 0;
}

/* FP32 surrogate walk of history i for network outputs out8[8]; the path is
 * written like find_MLP_path's (first voxel, then each new voxel). With a
 * non-NULL t/v, the per-step surrogate positions (up to cap) are recorded.
 * Returns the path length, -1 beyond max_intersections. */
int oracle_surrogate_path(const pct_params* p, const pct_hist_soa* h, int64_t i, const float* out8, uint32_t* path,
                          float* t_rec, float* v_rec, int cap, int* n_steps) {
    Geo g = make_geo(*p);
    Hist m;
    setup(g, h, i, m);
    float Qt[5][4], Qv[5][4];
    path_coefficients(m, out8, Qt, Qv);
    const float ca = (float)m.ca, sa = (float)m.sa, u_in = (float)m.u_in, inv_u2 = (float)(1.0 / m.u_2);
    const float x_zero = (float)g.x_zero, y_zero = (float)g.y_zero, z_zero = (float)g.z_zero;
    const float inv_w = (float)(1.0 / g.voxel_width), inv_h = (float)(1.0 / g.voxel_height);
    const float inv_t = (float)(1.0 / g.voxel_thickness);
    int n = 0, k = 0;
    path[n++] = h->first_voxel[i];
    const double STEP = g.mlp_u_step;
    for (double u_1 = STEP; u_1 < m.u_2 - STEP; u_1 += STEP, ++k) {
        float u1 = (float)u_1;
        float s5 = (u1 * inv_u2) * 5.0f;
        int seg = (int)s5;
        if (seg > 4) seg = 4;
        float tau = s5 - (float)seg;
        float t = Qt[seg][0] + tau * (Qt[seg][1] + tau * (Qt[seg][2] + tau * Qt[seg][3]));
        float v = Qv[seg][0] + tau * (Qv[seg][1] + tau * (Qv[seg][2] + tau * Qv[seg][3]));
        if (t_rec && k < cap) { t_rec[k] = t; v_rec[k] = v; }
        float us = u_in + u1;
        float x = ca * us - sa * t;
        float y = sa * us + ca * t;
        int vx = calculate_voxel_f(x_zero, x, inv_w);
        int vy = calculate_voxel_f(y_zero, y, inv_h);
        int vz = calculate_voxel_f(z_zero, v, inv_t);
        int64_t voxel = (int64_t)vx + (int64_t)vy * g.C + (int64_t)vz * g.C * g.R;
        if (vx < 0 || vy < 0 || vz < 0 || voxel >= g.V) continue;
        if ((uint32_t)voxel != path[n - 1]) {
            if (n >= g.max_intersections) return -1;
            path[n++] = (uint32_t)voxel;
        }
    }
    if (n_steps) *n_steps = k;
    return n;
}

}  // extern "C"
