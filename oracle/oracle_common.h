/*
 * oracle_common.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Shared definitions of the CPU oracle: a plain restatement of the reference
 * (lynch829/pCT, src/pCT_Reconstruction_Data_Segments.cu) used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker. The
 * product (pct_amd/) never includes, links or calls anything under oracle/.
 *
 * The oracle computes its own derived constants from pct_params exactly as the
 * reference's macros do (include/pCT_data.h:215-453) and uses the shared
 * deterministic math of pct_amd/csrc/pct_math.h (pinned separately against
 * numpy/glibc) so that voxel lists can be compared bit for bit.
 */
#ifndef ORACLE_COMMON_H
#define ORACLE_COMMON_H

#include <stdint.h>
#include "../include/pct.h"
#include "../pct_amd/csrc/pct_math.h"

namespace oracle {

/* Derived constants, following the reference's macro definitions. */
struct Geo {
    int C, R, S;
    int64_t V;
    double voxel_width;      /* RECON_CYL_DIAMETER / COLUMNS          data.h:268 */
    double voxel_height;     /* RECON_CYL_DIAMETER / ROWS             data.h:269 */
    double voxel_thickness;  /* IMAGE_THICKNESS / SLICES              data.h:270 */
    double x_zero, y_zero, z_zero;  /* data.h:272-274 */
    double mlp_u_step;       /* VOXEL_WIDTH / 2                       data.h:300 */
    double radius, height;
    double e_0, x_0;
    /* A_i and the A_i_OVER_k folds of data.h:306-348 */
    double A0, A1, A2, A3, A4, A5;
    double a0_2, a0_3, a1_2, a1_3, a1_4, a1_6, a1_12, a2_3, a2_4, a2_5, a2_12, a2_30;
    double a3_4, a3_5, a3_6, a3_20, a3_60, a4_5, a4_6, a4_7, a4_30, a4_105;
    double a5_6, a5_7, a5_8, a5_42, a5_168;
    int max_intersections;
};

inline Geo make_geo(const pct_params& p) {
    Geo g;
    g.C = p.columns; g.R = p.rows; g.S = p.slices;
    g.V = (int64_t)p.columns * p.rows * p.slices;
    double diameter = 2 * p.recon_cyl_radius;
    g.voxel_width = diameter / p.columns;
    g.voxel_height = diameter / p.rows;
    double image_thickness = p.slices * p.slice_thickness;
    g.voxel_thickness = image_thickness / p.slices;
    g.x_zero = -p.recon_cyl_radius;
    g.y_zero = p.recon_cyl_radius;
    g.z_zero = p.recon_cyl_height / 2;
    g.mlp_u_step = g.voxel_width / 2;
    g.radius = p.recon_cyl_radius;
    g.height = p.recon_cyl_height;
    g.e_0 = p.e_0; g.x_0 = p.x_0;
    g.A0 = p.a[0]; g.A1 = p.a[1]; g.A2 = p.a[2]; g.A3 = p.a[3]; g.A4 = p.a[4]; g.A5 = p.a[5];
    g.a0_2 = g.A0 / 2; g.a0_3 = g.A0 / 3;
    g.a1_2 = g.A1 / 2; g.a1_3 = g.A1 / 3; g.a1_4 = g.A1 / 4; g.a1_6 = g.A1 / 6; g.a1_12 = g.A1 / 12;
    g.a2_3 = g.A2 / 3; g.a2_4 = g.A2 / 4; g.a2_5 = g.A2 / 5; g.a2_12 = g.A2 / 12; g.a2_30 = g.A2 / 30;
    g.a3_4 = g.A3 / 4; g.a3_5 = g.A3 / 5; g.a3_6 = g.A3 / 6; g.a3_20 = g.A3 / 20; g.a3_60 = g.A3 / 60;
    g.a4_5 = g.A4 / 5; g.a4_6 = g.A4 / 6; g.a4_7 = g.A4 / 7; g.a4_30 = g.A4 / 30; g.a4_105 = g.A4 / 105;
    g.a5_6 = g.A5 / 6; g.a5_7 = g.A5 / 7; g.a5_8 = g.A5 / 8; g.a5_42 = g.A5 / 42; g.a5_168 = g.A5 / 168;
    g.max_intersections = p.max_intersections;
    return g;
}

}  // namespace oracle

#endif
