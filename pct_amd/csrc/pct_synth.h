/*
 * pct_synth.h -- seeded synthetic proton-CT scan generator (host and device).
 *
 * The reference reads measured PCTD projection files (.cu:1769-1933); the
 * benchmark configurations instead use a synthetic scan, so this header
 * produces the reference's raw tracker record (data.h:497-508) for history
 * index i from (seed, i) alone:
 *
 *   - Philox-4x32-10 counter-based RNG (Salmon, Moraes, Dror, Shaw, "Parallel
 *     random numbers: as easy as 1, 2, 3", SC'11), keyed by the seed and
 *     counted by (history, draw), so any subset of histories can be generated
 *     in any order on any device with identical results.
 *   - Gantry angles in equal groups (angle-major storage order, like the
 *     reference's one-file-per-angle input); the first N mod A angles get one
 *     extra history.
 *   - Cone-beam entry: hit (t, v) uniform at the inner entry plane, direction
 *     from a point source at u = -SOURCE_RADIUS.
 *   - Phantom: cylinders along z (infinite height) or rotated ellipsoids;
 *     RSP of a point = the RSP of the last listed body containing it (inserts
 *     nested in the body). Line integrals are exact chord lengths per body.
 *   - Multiple Coulomb scattering: at the object exit, per projection plane,
 *     displacement and angle from the PDG joint-Gaussian prescription
 *     (y = L theta0 (z1/sqrt(12) + z2/2), theta = theta0 z2) with the Highland
 *     theta0 = 13.6 MeV/(beta c p) sqrt(L/X0) (1 + 0.038 ln(L/X0)) at the
 *     water-equivalent thickness L, fixed beam momentum (no energy loss).
 *   - WEPL = line integral * (1 + straggle * N(0,1)); 0 exactly on a miss.
 * All arithmetic uses pct_math.h, so host and device outputs are identical.
 */
#ifndef PCT_SYNTH_H
#define PCT_SYNTH_H

#include "pct_math.h"

#include "../../include/pct_synth_api.h"

/* Philox-4x32-10 */
PCT_HD void pct_philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}
PCT_HD double pct_u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; } /* (0,1) */

/* 8 uniforms for history i, draw block d */
PCT_HD void pct_synth_uniforms(uint64_t seed, int64_t i, uint32_t d, double u[4]) {
    uint32_t c[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), d, 0x70C7u};
    pct_philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    for (int k = 0; k < 4; ++k) u[k] = pct_u01(c[k]);
}
PCT_HD void pct_box_muller(double u1, double u2, double* z1, double* z2) {
    double r = pct_sqrt(-2.0 * pct_log(u1));
    double s, c;
    pct_sincos(6.283185307179586 * u2, &s, &c);
    *z1 = r * c; *z2 = r * s;
}

/* angle index of history i (angle-major, first N mod A angles get one more) */
PCT_HD int32_t pct_synth_angle_index(const pct_synth_cfg* g, int64_t i) {
    int64_t q = g->n_histories / g->n_angles, rem = g->n_histories % g->n_angles;
    int64_t big = rem * (q + 1);
    if (i < big) return (int32_t)(i / (q + 1));
    return (int32_t)(rem + (i - big) / (q > 0 ? q : 1));
}

/* chord interval [s0, s1] (in u) of the line t = t0 + m u with the circle of
 * centre (cu, ct) and radius r in the (u, t) plane; returns 0 when disjoint */
PCT_HD int pct_synth_chord(double t0, double m, double cu, double ct, double r, double* s0, double* s1) {
    /* (u - cu)^2 + (t0 + m u - ct)^2 = r^2 */
    double d = t0 - ct;
    double a = 1.0 + m * m;
    double b = 2.0 * (m * d - cu);
    double c = cu * cu + d * d - r * r;
    double disc = b * b - 4.0 * a * c;
    if (!(disc > 0.0)) return 0;
    double sq = pct_sqrt(disc);
    *s0 = (-b - sq) / (2.0 * a);
    *s1 = (-b + sq) / (2.0 * a);
    return 1;
}

/* chord interval [s0, s1] (in u) of the line (u, t0 + m_t u, v0 + m_v u) with
 * ellipsoid k of the configuration, in the gantry frame at angle theta
 * (cos ct, sin st); returns 0 when disjoint */
PCT_HD int pct_synth_ell_chord(const pct_synth_cfg* g, int k, double ct, double st, double t0, double m_t, double v0,
                               double m_v, double* s0, double* s1) {
    double cu = ct * g->ell_x[k] + st * g->ell_y[k];
    double ctt = ct * g->ell_y[k] - st * g->ell_x[k];
    /* principal axes at angle phi - theta in the (u, t) plane */
    double sa, ca;
    pct_sincos(g->ell_phi[k] - pct_atan2(st, ct), &sa, &ca);
    double pu = -cu, pt = t0 - ctt, pv = v0 - g->ell_z[k];
    double q0x = (ca * pu + sa * pt) / g->ell_a[k], qdx = (ca + sa * m_t) / g->ell_a[k];
    double q0y = (-sa * pu + ca * pt) / g->ell_b[k], qdy = (-sa + ca * m_t) / g->ell_b[k];
    double q0z = pv / g->ell_c[k], qdz = m_v / g->ell_c[k];
    double A = qdx * qdx + qdy * qdy + qdz * qdz;
    double Bh = q0x * qdx + q0y * qdy + q0z * qdz;
    double Cc = q0x * q0x + q0y * q0y + q0z * q0z - 1.0;
    double disc = Bh * Bh - A * Cc;
    if (!(disc > 0.0)) return 0;
    double sq = pct_sqrt(disc);
    *s0 = (-Bh - sq) / A;
    *s1 = (-Bh + sq) / A;
    return 1;
}

/* 1 when the centre of ellipsoid k lies inside ellipsoid j */
PCT_HD int pct_synth_ell_contains(const pct_synth_cfg* g, int j, int k) {
    double dx = g->ell_x[k] - g->ell_x[j], dy = g->ell_y[k] - g->ell_y[j], dz = g->ell_z[k] - g->ell_z[j];
    double sp, cp;
    pct_sincos(g->ell_phi[j], &sp, &cp);
    double lx = (cp * dx + sp * dy) / g->ell_a[j], ly = (-sp * dx + cp * dy) / g->ell_b[j], lz = dz / g->ell_c[j];
    return lx * lx + ly * ly + lz * lz <= 1.0;
}

typedef struct pct_synth_record {
    int32_t gantry_angle;
    float t_in_1, t_in_2, t_out_1, t_out_2;
    float u_in_1, u_in_2, u_out_1, u_out_2;
    float v_in_1, v_in_2, v_out_1, v_out_2;
    float wepl;
} pct_synth_record;

PCT_HD void pct_synth_history(const pct_synth_cfg* g, int64_t i, pct_synth_record* out) {
    int32_t a = pct_synth_angle_index(g, i);
    double theta_deg = a * g->angle_step_deg;
    double theta = theta_deg * (PCT_PI / 180.0);
    double ct, st;
    pct_sincos(theta, &st, &ct);
    double uA[4], uB[4], uC[4];
    pct_synth_uniforms(g->seed, i, 0, uA);
    pct_synth_uniforms(g->seed, i, 1, uB);
    pct_synth_uniforms(g->seed, i, 2, uC);

    double t2 = (2.0 * uA[0] - 1.0) * g->t_half;
    double v2 = (2.0 * uA[1] - 1.0) * g->v_half;
    double m_t = t2 / (g->u_in_2 + g->source_radius);
    double m_v = v2 / (g->u_in_2 + g->source_radius);
    /* straight line in the gantry frame: t(u) = t0 + m_t u, v(u) = v0 + m_v u */
    double t0 = t2 - m_t * g->u_in_2;
    double v0 = v2 - m_v * g->u_in_2;
    double ds_du = pct_sqrt(1.0 + m_t * m_t + m_v * m_v);

    /* exact line integral: RSP of the last containing cylinder / ellipsoid */
    double wepl_true = 0.0, u_exit_obj = g->u_in_2, path_len = 0.0;
    int hit = 0;
    for (int k = 0; k < g->n_ell; ++k) {
        double s0, s1;
        if (!pct_synth_ell_chord(g, k, ct, st, t0, m_t, v0, m_v, &s0, &s1)) continue;
        double len = (s1 - s0) * ds_du;
        double parent = 0.0;
        for (int j = 0; j < k; ++j)
            if (pct_synth_ell_contains(g, j, k)) parent = g->ell_rsp[j];
        wepl_true += (g->ell_rsp[k] - parent) * len;
        if (k == 0) { hit = 1; u_exit_obj = s1; path_len = len; }
    }
    for (int k = 0; g->n_ell <= 0 && k < g->n_cyl; ++k) {
        /* cylinder centre in the gantry frame: u = c x + s y, t = c y - s x */
        double cu = ct * g->cyl_x[k] + st * g->cyl_y[k];
        double ctt = ct * g->cyl_y[k] - st * g->cyl_x[k];
        double s0, s1;
        if (!pct_synth_chord(t0, m_t, cu, ctt, g->cyl_r[k], &s0, &s1)) continue;
        double len = (s1 - s0) * ds_du;
        /* this cylinder replaces the RSP of the cylinders listed before it
         * that contain it (phantoms are nested: body first, inserts after) */
        double parent = 0.0;
        for (int j = 0; j < k; ++j) {
            double dx = g->cyl_x[k] - g->cyl_x[j], dy = g->cyl_y[k] - g->cyl_y[j];
            if (pct_sqrt(dx * dx + dy * dy) + g->cyl_r[k] <= g->cyl_r[j]) parent = g->cyl_rsp[j];
        }
        wepl_true += (g->cyl_rsp[k] - parent) * len;
        if (k == 0) { hit = 1; u_exit_obj = s1; path_len = len; }
    }
    double z1, z2, z3, z4, z5, z6;
    pct_box_muller(uB[0], uB[1], &z1, &z2);
    pct_box_muller(uB[2], uB[3], &z3, &z4);
    pct_box_muller(uC[0], uC[1], &z5, &z6);
    double dt = 0.0, dv = 0.0, th_t = 0.0, th_v = 0.0;
    if (hit && wepl_true > 0.0) {
        double L = wepl_true;
        double theta0 = 13.6 / g->beta_cp * pct_sqrt(L / g->x0) * (1.0 + 0.038 * pct_log(L / g->x0));
        (void)path_len;
        dt = L * theta0 * (z1 / 3.4641016151377544 + z2 / 2.0);
        th_t = theta0 * z2;
        dv = L * theta0 * (z3 / 3.4641016151377544 + z4 / 2.0);
        th_v = theta0 * z4;
    } else {
        u_exit_obj = g->u_in_2;
    }
    double mt_out = m_t + th_t, mv_out = m_v + th_v;
    double t_exit_obj = t0 + m_t * u_exit_obj + dt;
    double v_exit_obj = v0 + m_v * u_exit_obj + dv;
    double t_o1 = t_exit_obj + mt_out * (g->u_out_1 - u_exit_obj);
    double v_o1 = v_exit_obj + mv_out * (g->u_out_1 - u_exit_obj);
    double t_o2 = t_o1 + mt_out * (g->u_out_2 - g->u_out_1);
    double v_o2 = v_o1 + mv_out * (g->u_out_2 - g->u_out_1);

    out->gantry_angle = (int32_t)theta_deg;
    out->u_in_1 = (float)g->u_in_1; out->u_in_2 = (float)g->u_in_2;
    out->u_out_1 = (float)g->u_out_1; out->u_out_2 = (float)g->u_out_2;
    out->t_in_1 = (float)(t0 + m_t * g->u_in_1); out->t_in_2 = (float)t2;
    out->v_in_1 = (float)(v0 + m_v * g->u_in_1); out->v_in_2 = (float)v2;
    out->t_out_1 = (float)t_o1; out->t_out_2 = (float)t_o2;
    out->v_out_1 = (float)v_o1; out->v_out_2 = (float)v_o2;
    out->wepl = wepl_true > 0.0 ? (float)(wepl_true * (1.0 + g->straggle * z5)) : 0.0f;
}

#endif /* PCT_SYNTH_H */
