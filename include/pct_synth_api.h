/*
 * pct_synth_api.h -- C ABI of the seeded synthetic scan generator.
 *
 * No counterpart in the reference, which reads measured PCTD projection files
 * (read_data_chunk_v02, .cu:1769-1933). The generator produces the same raw
 * tracker record (data.h:497-508) so that the pipeline from
 * pct_recon_volume_intersections onwards runs unchanged on synthetic scans.
 */
#ifndef PCT_SYNTH_API_H
#define PCT_SYNTH_API_H

#include <stdint.h>
#include "pct.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PCT_SYNTH_MAX_CYL 16

typedef struct pct_synth_cfg {
    uint64_t seed;
    int64_t  n_histories;
    int32_t  n_angles;            /* projection angles                             */
    double   angle_step_deg;      /* gantry step [deg]                             */
    double   u_in_1, u_in_2, u_out_1, u_out_2;   /* tracker planes [cm]           */
    double   t_half, v_half;      /* entry hits uniform in |t|<t_half, |v|<v_half  */
    double   source_radius;       /* point source distance [cm]                    */
    double   beta_cp;             /* beta*c*p [MeV] of the beam                    */
    double   x0;                  /* radiation length of water [cm]                */
    double   straggle;            /* relative WEPL noise                           */
    int32_t  n_cyl;               /* cylinders along z, body first, inserts after  */
    double   cyl_x[PCT_SYNTH_MAX_CYL], cyl_y[PCT_SYNTH_MAX_CYL];
    double   cyl_r[PCT_SYNTH_MAX_CYL], cyl_rsp[PCT_SYNTH_MAX_CYL];
    /* ellipsoids (used instead of the cylinders when n_ell > 0), body first:
     * centre [cm], semi-axes [cm], rotation about z [rad], RSP; an ellipsoid
     * replaces the RSP of the last listed one containing its centre */
    int32_t  n_ell;
    double   ell_x[PCT_SYNTH_MAX_CYL], ell_y[PCT_SYNTH_MAX_CYL], ell_z[PCT_SYNTH_MAX_CYL];
    double   ell_a[PCT_SYNTH_MAX_CYL], ell_b[PCT_SYNTH_MAX_CYL], ell_c[PCT_SYNTH_MAX_CYL];
    double   ell_phi[PCT_SYNTH_MAX_CYL], ell_rsp[PCT_SYNTH_MAX_CYL];
} pct_synth_cfg;

/* Raw tracker arrays written by the generator (same fields as pct_raw_soa). */
typedef struct pct_synth_soa {
    int32_t* gantry_angle;
    float *t_in_1, *t_in_2, *t_out_1, *t_out_2;
    float *u_in_1, *u_in_2, *u_out_1, *u_out_2;
    float *v_in_1, *v_in_2, *v_out_1, *v_out_2;
    float* wepl;
} pct_synth_soa;

/* Histories [start, start+count) of the scan, on the device (out: device
 * arrays of `count` elements) ... */
int pct_synth_generate(pct_ctx* ctx, const pct_synth_cfg* cfg, int64_t start, int64_t count,
                       const pct_synth_soa* out);
/* ... or on the host (out: host arrays); bit-identical results. */
int pct_synth_generate_host(const pct_synth_cfg* cfg, int64_t start, int64_t count, const pct_synth_soa* out);

#ifdef __cplusplus
}
#endif

#endif
