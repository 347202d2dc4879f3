/*
 * oracle_hotpath.cpp -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
 *
 * FP64 restatement of the reference's DROP hot path and the stages that feed
 * it, following src/pCT_Reconstruction_Data_Segments.cu (".cu" below) line by
 * line, including its float/double typing (values stored to float arrays are
 * rounded to float exactly where the reference stores them). Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
 * checker; the product never links it.
 *
 * Canonical choices where the reference is ambiguous or disagrees with itself
 * (each is listed with its reason in the design notes of the repository):
 *   - pow(u,2.0) -> u*u, pow(u,3.0) -> u*u*u (the GPU form of .cu:5372).
 *   - A_i_OVER_k macros folded to double constants A_i/k once.
 *   - transcendental functions from pct_math.h (own implementations).
 *   - EffectiveChordLength takes the double average of the float angles as in
 *     the GPU call (.cu:5315), not the float average of the CPU call (.cu:7531).
 *   - a path point whose voxel index falls outside [0, NUM_VOXELS) is skipped
 *     (the reference writes out of bounds there, .cu:5481 / 5833).
 *   - calculate_intercepts leaves its outputs at 0 when the line misses the
 *     cylinder (the reference leaves them uninitialised, .cu:2359-2386).
 *   - DROP: per-history arithmetic of block_update_GPU (.cu:5822-5834), x_update
 *     accumulated in double in storage order (DROP_blocks .cu:8023-8043), image
 *     update of DROP_update (.cu:8044-8055) after every block including the
 *     last partial one (the GPU driver's blocking, .cu:6803-6844).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "oracle_common.h"

using namespace oracle;

namespace {

/* calculate_voxel (.cu:8445-8448): int( |pos - zero| / size ). Returns -1 when
 * the quotient does not fit an int (the reference's behaviour is undefined). */
inline int calculate_voxel(double zero_coordinate, double current_position, double voxel_size) {
    double q = pct_fabs(current_position - zero_coordinate) / voxel_size;
    if (!(q < 1.0e9)) return -1;
    return (int)q;
}

/* edge_coordinate (.cu:8584-8589) */
inline double edge_coordinate(double zero_coordinate, int voxel_entered, double voxel_size,
                              int increasing_direction, int step_direction) {
    int on_edge = (step_direction == increasing_direction) ? voxel_entered : voxel_entered + 1;
    return zero_coordinate + (increasing_direction * on_edge * voxel_size);
}
/* distance_remaining (.cu:8571-8582) */
inline double distance_remaining(double zero_coordinate, double current_position, int increasing_direction,
                                 int step_direction, double voxel_size, int current_voxel) {
    int next_voxel = current_voxel + increasing_direction * step_direction;
    double next_edge = edge_coordinate(zero_coordinate, next_voxel, voxel_size, increasing_direction, step_direction);
    return pct_fabs(next_edge - current_position);
}
/* corresponding_coordinate (.cu:8598-8603) */
inline double corresponding_coordinate(double m, double x, double x0, double y0) { return m * (x - x0) + y0; }

const int RIGHT = 1, DOWN = -1;
const int X_INC = RIGHT, Y_INC = DOWN, Z_INC = DOWN;   /* data.h:275-277 */

/* ------------------------------------------------------------------------ */
/* A1: calculate_intercepts (.cu:2301-2387)                                  */
/* ------------------------------------------------------------------------ */
bool calculate_intercepts(const Geo& g, double u, double t, double ut_angle, double& u_intercept, double& t_intercept) {
    const double PIO4 = 0.7853981633974483;   /* atan(1.0) */
    bool entry_in_cone = ((ut_angle > PIO4) && (ut_angle < 3 * PIO4)) || ((ut_angle > 5 * PIO4) && (ut_angle < 7 * PIO4));
    double u_temp;
    if (entry_in_cone) {
        u_temp = u;
        u = -t;
        t = u_temp;
        ut_angle += 2 * PIO4;
    }
    double m = pct_tan(ut_angle);
    double b_in = t - m * u;
    double a = 1 + m * m;
    double b = 2 * m * b_in;
    double c = b_in * b_in - g.radius * g.radius;
    double entry_discriminant = b * b - (4 * a * c);
    bool intersected = (entry_discriminant > 0);
    u_intercept = 0.0;
    t_intercept = 0.0;
    if (intersected) {
        double u_intercept_1 = (pct_sqrt(entry_discriminant) - b) / (2 * a);
        double u_intercept_2 = (pct_sqrt(entry_discriminant) + b) / (2 * a);
        double t_intercept_1 = m * u_intercept_1 + b_in;
        double t_intercept_2 = m * u_intercept_2 - b_in;
        double d1u = u_intercept_1 - u, d1t = t_intercept_1 - t;
        double d2u = u_intercept_2 + u, d2t = t_intercept_2 + t;
        double squared_distance_1 = d1u * d1u + d1t * d1t;
        double squared_distance_2 = d2u * d2u + d2t * d2t;
        double s1 = (double)(squared_distance_1 <= squared_distance_2);
        double s2 = (double)(squared_distance_1 > squared_distance_2);
        u_intercept = u_intercept_1 * s1 - u_intercept_2 * s2;
        t_intercept = t_intercept_1 * s1 - t_intercept_2 * s2;
    }
    if (entry_in_cone) {
        u_temp = u_intercept;
        u_intercept = t_intercept;
        t_intercept = -u_temp;
    }
    return intersected;
}

inline double sgn_int(float z) { return (double)((z >= 0) - (z < 0)); }

/* ------------------------------------------------------------------------ */
/* D1: find_MLP_endpoints (.cu:4155-4438, identical to the device version   */
/* .cu:4997-5277) on a byte hull.                                            */
/* ------------------------------------------------------------------------ */
bool find_MLP_endpoints(const Geo& g, const uint8_t* image, double x_start, double y_start, double z_start,
                        double xy_angle, double xz_angle, double& x_object, double& y_object, double& z_object,
                        int& voxel_x, int& voxel_y, int& voxel_z, bool entering) {
    const int C = g.C, R = g.R, S = g.S;
    int x_move_direction, y_move_direction, z_move_direction;
    double x = x_start, y = y_start, z = z_start;
    double x_to_go, y_to_go, z_to_go, x_extension, y_extension;
    int voxel;
    bool hit_hull = false, end_walk, outside_image;
    if (!entering) xy_angle += PCT_PI;
    double cxy = pct_cos(xy_angle), sxy = pct_sin(xy_angle), sxz = pct_sin(xz_angle);
    x_move_direction = (cxy >= 0) - (cxy <= 0);
    y_move_direction = (sxy >= 0) - (sxy <= 0);
    z_move_direction = (sxz >= 0) - (sxz <= 0);
    if (x_move_direction < 0) z_move_direction *= -1;

    voxel_x = calculate_voxel(g.x_zero, x, g.voxel_width);
    voxel_y = calculate_voxel(g.y_zero, y, g.voxel_height);
    voxel_z = calculate_voxel(g.z_zero, z, g.voxel_thickness);

    x_to_go = distance_remaining(g.x_zero, x, X_INC, x_move_direction, g.voxel_width, voxel_x);
    y_to_go = distance_remaining(g.y_zero, y, Y_INC, y_move_direction, g.voxel_height, voxel_y);
    z_to_go = distance_remaining(g.z_zero, z, Z_INC, z_move_direction, g.voxel_thickness, voxel_z);

    voxel = voxel_x + voxel_y * C + voxel_z * C * R;
    double txy = pct_tan(xy_angle), txz = pct_tan(xz_angle);
    double delta_yx = pct_fabs(txy);
    double delta_zx = pct_fabs(txz);
    double delta_zy = pct_fabs(txz / txy);
    double dy_dx = txy;
    double dz_dx = txz;
    double dz_dy = txz / txy;
    double dx_dy = 1.0 / txy;          /* pow( tan(xy_angle), -1.0 ) */
    double dx_dz = 1.0 / txz;          /* pow( tan(xz_angle), -1.0 ) */
    double dy_dz = txy / txz;
    (void)dx_dz;

    outside_image = (voxel_x >= C) || (voxel_y >= R) || (voxel_z >= S) || (voxel_x < 0) || (voxel_y < 0) || (voxel_z < 0);
    if (!outside_image) hit_hull = (image[voxel] == 1);
    end_walk = outside_image || hit_hull;
    /* the reference loops without a bound; a walk always leaves the grid in at
     * most C+R+S steps, the guard only protects against NaN inputs */
    int guard = 4 * (C + R + S) + 16;
    if (z_move_direction != 0) {
        while (!end_walk && guard-- > 0) {
            x_extension = delta_zx * x_to_go;
            y_extension = delta_zy * y_to_go;
            if ((z_to_go <= x_extension) && (z_to_go <= y_extension)) {
                voxel_z -= z_move_direction;
                z = edge_coordinate(g.z_zero, voxel_z, g.voxel_thickness, Z_INC, z_move_direction);
                x = corresponding_coordinate(dx_dz, z, z_start, x_start);
                y = corresponding_coordinate(dy_dz, z, z_start, y_start);
                x_to_go = distance_remaining(g.x_zero, x, X_INC, x_move_direction, g.voxel_width, voxel_x);
                y_to_go = distance_remaining(g.y_zero, y, Y_INC, y_move_direction, g.voxel_height, voxel_y);
                z_to_go = g.voxel_thickness;
            } else if (x_extension <= y_extension) {
                voxel_x += x_move_direction;
                x = edge_coordinate(g.x_zero, voxel_x, g.voxel_width, X_INC, x_move_direction);
                y = corresponding_coordinate(dy_dx, x, x_start, y_start);
                z = corresponding_coordinate(dz_dx, x, x_start, z_start);
                x_to_go = g.voxel_width;
                y_to_go = distance_remaining(g.y_zero, y, Y_INC, y_move_direction, g.voxel_height, voxel_y);
                z_to_go = distance_remaining(g.z_zero, z, Z_INC, z_move_direction, g.voxel_thickness, voxel_z);
            } else {
                voxel_y -= y_move_direction;
                y = edge_coordinate(g.y_zero, voxel_y, g.voxel_height, Y_INC, y_move_direction);
                x = corresponding_coordinate(dx_dy, y, y_start, x_start);
                z = corresponding_coordinate(dz_dy, y, y_start, z_start);
                x_to_go = distance_remaining(g.x_zero, x, X_INC, x_move_direction, g.voxel_width, voxel_x);
                y_to_go = g.voxel_height;
                z_to_go = distance_remaining(g.z_zero, z, Z_INC, z_move_direction, g.voxel_thickness, voxel_z);
            }
            if (x_to_go == 0) { x_to_go = g.voxel_width; voxel_x += x_move_direction; }
            if (y_to_go == 0) { y_to_go = g.voxel_height; voxel_y -= y_move_direction; }
            if (z_to_go == 0) { z_to_go = g.voxel_thickness; voxel_z -= z_move_direction; }
            voxel_z = voxel_z > 0 ? voxel_z : 0;
            voxel = voxel_x + voxel_y * C + voxel_z * C * R;
            outside_image = (voxel_x >= C) || (voxel_y >= R) || (voxel_z >= S) || (voxel_x < 0) || (voxel_y < 0) || (voxel_z < 0);
            if (!outside_image) hit_hull = (image[voxel] == 1);
            end_walk = outside_image || hit_hull;
        }
    } else {
        while (!end_walk && guard-- > 0) {
            y_extension = y_to_go / delta_yx;
            if (x_to_go <= y_extension) {
                voxel_x += x_move_direction;
                x = edge_coordinate(g.x_zero, voxel_x, g.voxel_width, X_INC, x_move_direction);
                y = corresponding_coordinate(dy_dx, x, x_start, y_start);
                x_to_go = g.voxel_width;
                y_to_go = distance_remaining(g.y_zero, y, Y_INC, y_move_direction, g.voxel_height, voxel_y);
            } else {
                voxel_y -= y_move_direction;
                y = edge_coordinate(g.y_zero, voxel_y, g.voxel_height, Z_INC, y_move_direction);
                x = corresponding_coordinate(dx_dy, y, y_start, x_start);
                x_to_go = distance_remaining(g.x_zero, x, X_INC, x_move_direction, g.voxel_width, voxel_x);
                y_to_go = g.voxel_height;
            }
            if (x_to_go == 0) { x_to_go = g.voxel_width; voxel_x += x_move_direction; }
            if (y_to_go == 0) { y_to_go = g.voxel_height; voxel_y -= y_move_direction; }
            voxel = voxel_x + voxel_y * C + voxel_z * C * R;
            outside_image = (voxel_x >= C) || (voxel_y >= R) || (voxel_z >= S) || (voxel_x < 0) || (voxel_y < 0) || (voxel_z < 0);
            if (!outside_image) hit_hull = (image[voxel] == 1);
            end_walk = outside_image || hit_hull;
        }
    }
    if (!end_walk) hit_hull = false;   /* guard exhausted: treated as a miss */
    if (hit_hull) { x_object = x; y_object = y; z_object = z; }
    return hit_hull;
}

/* ------------------------------------------------------------------------ */
/* D3: EffectiveChordLength (.cu:4974-4995 device, 7380-7401 host)           */
/* sinf/cosf take a float argument: the double argument is rounded first.    */
/* ------------------------------------------------------------------------ */
double effective_chord_length(const Geo& g, double abs_angle_t, double abs_angle_v) {
    const double PIO2 = 2 * 0.7853981633974483;
    double eff_angle_t = abs_angle_t - ((int)(abs_angle_t / (PCT_PI / 2))) * (PCT_PI / 2);
    (void)PIO2;
    double eff_angle_v = pct_fabs(abs_angle_v);
    double step_fraction = g.mlp_u_step / g.voxel_width;
    float s2t = pct_sinf((float)(2 * eff_angle_t));
    float ct = pct_cosf((float)eff_angle_t), st = pct_sinf((float)eff_angle_t);
    float cs_t = ct + st;                       /* float + float */
    double chord_length_2D = (1 / 3.0) * ((step_fraction * step_fraction * s2t - 6) / (step_fraction * s2t - 2 * cs_t)
                                          + step_fraction * step_fraction * s2t / (2 * cs_t));
    double mean_pixel_width = g.voxel_width / cs_t;
    double height_fraction = g.voxel_thickness / mean_pixel_width;
    step_fraction = g.mlp_u_step / mean_pixel_width;
    float s2v = pct_sinf((float)(2 * eff_angle_v));
    float cv = pct_cosf((float)eff_angle_v), sv = pct_sinf((float)eff_angle_v);
    double hcs = height_fraction * cv + sv;
    double chord_length_3D = (1 / 3.0) * ((step_fraction * step_fraction * s2v - 6 * height_fraction) / (step_fraction * s2v - 2 * hcs)
                                          + step_fraction * step_fraction * s2v / (2 * hcs));
    return g.voxel_width * chord_length_2D * chord_length_3D;
}

/* ------------------------------------------------------------------------ */
/* D2: find_MLP_path (.cu:4439-4633; device twin .cu:5279-5501)              */
/* Returns the path length, or -1 when it would exceed max_intersections.   */
/* ------------------------------------------------------------------------ */
struct MlpTrace {            /* optional per-step record for the tests */
    double* u1;
    double* t1;
    double* v1;
    double* x1;
    double* y1;
    int cap;
    int n;
};

int find_MLP_path(const Geo& g, uint32_t* path, uint32_t first_voxel,
                  double x_in_object, double y_in_object, double z_in_object,
                  double x_out_object, double y_out_object, double z_out_object,
                  double xy_entry_angle, double xz_entry_angle, double xy_exit_angle, double xz_exit_angle,
                  int* skipped, MlpTrace* trace = nullptr) {
    const int C = g.C, R = g.R;
    const double STEP = g.mlp_u_step;
    int num_intersections = 0;
    double u_0 = 0.0, u_1 = STEP, u_2 = 0.0;
    double T_0[2] = {0.0, 0.0}, T_2[2] = {0.0, 0.0}, V_0[2] = {0.0, 0.0}, V_2[2] = {0.0, 0.0};
    double R_0[4] = {1.0, 0.0, 0.0, 1.0};
    double R_1[4] = {1.0, 0.0, 0.0, 1.0};
    double R_1T[4] = {1.0, 0.0, 0.0, 1.0};
    double Sigma_1I[4], Sigma_2I[4], first_term[4], second_term[2];
    if (skipped) *skipped = 0;

    path[num_intersections++] = first_voxel;

    double ca = pct_cos(xy_entry_angle), sa = pct_sin(xy_entry_angle);
    double u_in_object = (ca * x_in_object) + (sa * y_in_object);
    double u_out_object = (ca * x_out_object) + (sa * y_out_object);
    double t_in_object = (ca * y_in_object) - (sa * x_in_object);
    double t_out_object = (ca * y_out_object) - (sa * x_out_object);
    double v_in_object = z_in_object;
    double v_out_object = z_out_object;
    if (u_in_object > u_out_object) {
        xy_entry_angle += PCT_PI;
        xy_exit_angle += PCT_PI;
        ca = pct_cos(xy_entry_angle); sa = pct_sin(xy_entry_angle);
        u_in_object = (ca * x_in_object) + (sa * y_in_object);
        u_out_object = (ca * x_out_object) + (sa * y_out_object);
        t_in_object = (ca * y_in_object) - (sa * x_in_object);
        t_out_object = (ca * y_out_object) - (sa * x_out_object);
        v_in_object = z_in_object;
        v_out_object = z_out_object;
    }
    T_0[0] = t_in_object;
    T_2[0] = t_out_object;
    T_2[1] = xy_exit_angle - xy_entry_angle;
    V_0[0] = v_in_object;
    V_2[0] = v_out_object;
    V_2[1] = xz_exit_angle - xz_entry_angle;

    u_0 = 0.0;
    u_1 = STEP;
    u_2 = pct_fabs(u_out_object - u_in_object);

    double sigma_2_pre_1 = u_2 * u_2 * u_2 * (g.a0_3 + u_2 * (g.a1_12 + u_2 * (g.a2_30 + u_2 * (g.a3_60 + u_2 * (g.a4_105 + u_2 * g.a5_168)))));
    double sigma_2_pre_2 = u_2 * u_2 * (g.a0_2 + u_2 * (g.a1_6 + u_2 * (g.a2_12 + u_2 * (g.a3_20 + u_2 * (g.a4_30 + u_2 * g.a5_42)))));
    double sigma_2_pre_3 = u_2 * (g.A0 + u_2 * (g.a1_2 + u_2 * (g.a2_3 + u_2 * (g.a3_4 + u_2 * (g.a4_5 + u_2 * g.a5_6)))));

    while (u_1 < u_2 - STEP) {
        R_0[1] = u_1;
        R_1[1] = u_2 - u_1;
        R_1T[2] = u_2 - u_1;

        double l1 = g.e_0 * (1 + 0.038 * pct_log((u_1 - u_0) / g.x_0));
        double sigma_1_coefficient = l1 * l1 / g.x_0;
        double sigma_t1 = sigma_1_coefficient * (u_1 * u_1 * u_1 * (g.a0_3 + u_1 * (g.a1_12 + u_1 * (g.a2_30 + u_1 * (g.a3_60 + u_1 * (g.a4_105 + u_1 * g.a5_168))))));
        double sigma_t1_theta1 = sigma_1_coefficient * (u_1 * u_1 * (g.a0_2 + u_1 * (g.a1_6 + u_1 * (g.a2_12 + u_1 * (g.a3_20 + u_1 * (g.a4_30 + u_1 * g.a5_42))))));
        double sigma_theta1 = sigma_1_coefficient * (u_1 * (g.A0 + u_1 * (g.a1_2 + u_1 * (g.a2_3 + u_1 * (g.a3_4 + u_1 * (g.a4_5 + u_1 * g.a5_6))))));
        double determinant_Sigma_1 = sigma_t1 * sigma_theta1 - sigma_t1_theta1 * sigma_t1_theta1;
        Sigma_1I[0] = sigma_theta1 / determinant_Sigma_1;
        Sigma_1I[1] = -sigma_t1_theta1 / determinant_Sigma_1;
        Sigma_1I[2] = -sigma_t1_theta1 / determinant_Sigma_1;
        Sigma_1I[3] = sigma_t1 / determinant_Sigma_1;

        double l2 = g.e_0 * (1 + 0.038 * pct_log((u_2 - u_1) / g.x_0));
        double sigma_2_coefficient = l2 * l2 / g.x_0;
        double common_sigma_2_term_1 = u_1 * (g.A0 + u_1 * (g.a1_2 + u_1 * (g.a2_3 + u_1 * (g.a3_4 + u_1 * (g.a4_5 + u_1 * g.a5_6)))));
        double common_sigma_2_term_2 = u_1 * u_1 * (g.a0_2 + u_1 * (g.a1_3 + u_1 * (g.a2_4 + u_1 * (g.a3_5 + u_1 * (g.a4_6 + u_1 * g.a5_7)))));
        double common_sigma_2_term_3 = u_1 * u_1 * u_1 * (g.a0_3 + u_1 * (g.a1_4 + u_1 * (g.a2_5 + u_1 * (g.a3_6 + u_1 * (g.a4_7 + u_1 * g.a5_8)))));
        double sigma_t2 = sigma_2_coefficient * (sigma_2_pre_1 - u_2 * u_2 * common_sigma_2_term_1 + 2 * u_2 * common_sigma_2_term_2 - common_sigma_2_term_3);
        double sigma_t2_theta2 = sigma_2_coefficient * (sigma_2_pre_2 - u_2 * common_sigma_2_term_1 + common_sigma_2_term_2);
        double sigma_theta2 = sigma_2_coefficient * (sigma_2_pre_3 - common_sigma_2_term_1);
        double determinant_Sigma_2 = sigma_t2 * sigma_theta2 - sigma_t2_theta2 * sigma_t2_theta2;
        Sigma_2I[0] = sigma_theta2 / determinant_Sigma_2;
        Sigma_2I[1] = -sigma_t2_theta2 / determinant_Sigma_2;
        Sigma_2I[2] = -sigma_t2_theta2 / determinant_Sigma_2;
        Sigma_2I[3] = sigma_t2 / determinant_Sigma_2;

        double first_term_common_13_1 = Sigma_2I[0] * R_1[0] + Sigma_2I[1] * R_1[2];
        double first_term_common_13_2 = Sigma_2I[2] * R_1[0] + Sigma_2I[3] * R_1[2];
        double first_term_common_24_1 = Sigma_2I[0] * R_1[1] + Sigma_2I[1] * R_1[3];
        double first_term_common_24_2 = Sigma_2I[2] * R_1[1] + Sigma_2I[3] * R_1[3];
        first_term[0] = Sigma_1I[0] + R_1T[0] * first_term_common_13_1 + R_1T[1] * first_term_common_13_2;
        first_term[1] = Sigma_1I[1] + R_1T[0] * first_term_common_24_1 + R_1T[1] * first_term_common_24_2;
        first_term[2] = Sigma_1I[2] + R_1T[2] * first_term_common_13_1 + R_1T[3] * first_term_common_13_2;
        first_term[3] = Sigma_1I[3] + R_1T[2] * first_term_common_24_1 + R_1T[3] * first_term_common_24_2;

        double determinant_first_term = first_term[0] * first_term[3] - first_term[1] * first_term[2];
        double first_term_inversion_temp = first_term[0];
        first_term[0] = first_term[3] / determinant_first_term;
        first_term[1] = -first_term[1] / determinant_first_term;
        first_term[2] = -first_term[2] / determinant_first_term;
        first_term[3] = first_term_inversion_temp / determinant_first_term;

        double second_term_common_1 = R_0[0] * T_0[0] + R_0[1] * T_0[1];
        double second_term_common_2 = R_0[2] * T_0[0] + R_0[3] * T_0[1];
        double second_term_common_3 = Sigma_2I[0] * T_2[0] + Sigma_2I[1] * T_2[1];
        double second_term_common_4 = Sigma_2I[2] * T_2[0] + Sigma_2I[3] * T_2[1];
        second_term[0] = Sigma_1I[0] * second_term_common_1 + Sigma_1I[1] * second_term_common_2
                       + R_1T[0] * second_term_common_3 + R_1T[1] * second_term_common_4;
        second_term[1] = Sigma_1I[2] * second_term_common_1 + Sigma_1I[3] * second_term_common_2
                       + R_1T[2] * second_term_common_3 + R_1T[3] * second_term_common_4;
        double t_1 = first_term[0] * second_term[0] + first_term[1] * second_term[1];

        second_term_common_1 = R_0[0] * V_0[0] + R_0[1] * V_0[1];
        second_term_common_2 = R_0[2] * V_0[0] + R_0[3] * V_0[1];
        second_term_common_3 = Sigma_2I[0] * V_2[0] + Sigma_2I[1] * V_2[1];
        second_term_common_4 = Sigma_2I[2] * V_2[0] + Sigma_2I[3] * V_2[1];
        second_term[0] = Sigma_1I[0] * second_term_common_1 + Sigma_1I[1] * second_term_common_2
                       + R_1T[0] * second_term_common_3 + R_1T[1] * second_term_common_4;
        second_term[1] = Sigma_1I[2] * second_term_common_1 + Sigma_1I[3] * second_term_common_2
                       + R_1T[2] * second_term_common_3 + R_1T[3] * second_term_common_4;
        double v_1 = first_term[0] * second_term[0] + first_term[1] * second_term[1];

        double x_1 = (ca * (u_in_object + u_1)) - (sa * t_1);
        double y_1 = (sa * (u_in_object + u_1)) + (ca * t_1);
        double z_1 = v_1;
        if (trace && trace->n < trace->cap) {
            trace->u1[trace->n] = u_1; trace->t1[trace->n] = t_1; trace->v1[trace->n] = v_1;
            trace->x1[trace->n] = x_1; trace->y1[trace->n] = y_1;
            trace->n++;
        }

        int voxel_x = calculate_voxel(g.x_zero, x_1, g.voxel_width);
        int voxel_y = calculate_voxel(g.y_zero, y_1, g.voxel_height);
        int voxel_z = calculate_voxel(g.z_zero, z_1, g.voxel_thickness);
        int64_t voxel = (int64_t)voxel_x + (int64_t)voxel_y * C + (int64_t)voxel_z * C * R;
        if (voxel_x < 0 || voxel_y < 0 || voxel_z < 0 || voxel >= g.V) {
            if (skipped) (*skipped)++;
        } else if ((uint32_t)voxel != path[num_intersections - 1]) {
            if (num_intersections >= g.max_intersections) return -1;
            path[num_intersections++] = (uint32_t)voxel;
        }
        u_1 += STEP;
    }
    return num_intersections;
}

/* ------------------------------------------------------------------------ */
/* D4: the precision form, after find_MLP_path_GPU_tabulated (.cu:5502-5656) */
/*                                                                          */
/* As in the tabulated reference kernel: the sigma integrals and the 2x2    */
/* inversions are FP64; the sigma_2 scattering coefficient is read at the   */
/* remaining depth rounded to a table of step STEP/1000 (5e-5 cm at 1 mm    */
/* voxels, DEPTH_TABLE_STEP), the index being rounded once at the first     */
/* step and decremented by 1000 per step (.cu:5552, 5651). Choices of this  */
/* form that differ from the reference's tabulated kernel: sin/cos of the   */
/* entry angle and the u_2 polynomials are evaluated exactly instead of     */
/* from 0.001-degree and 5e-5 cm tables; the 2x2 solve is reduced to the    */
/* coefficients alpha, beta, gamma of t_1 = alpha t_in + beta t_out + gamma */
/* (theta_out - theta_in) (one reciprocal per inverse); t_1, v_1, the       */
/* rotation back to x/y and the voxel index are FP32 (voxel index by        */
/* multiplication with the FP32 reciprocal of the voxel size). The table    */
/* index is clamped to the depths a chord inside the image can have.        */
/* ------------------------------------------------------------------------ */
const int SCAT_SUB = 1000;

inline int calculate_voxel_f(float zero, float pos, float inv_size) {
    float q = fabsf(pos - zero) * inv_size;
    return q < 1.0e9f ? (int)q : -1;
}

inline double depth_coefficient(const Geo& g, int64_t j) {
    double depth = (j > 0 ? j : 1) * (g.mlp_u_step / SCAT_SUB);
    double l = g.e_0 * (1 + 0.038 * pct_log(depth / g.x_0));
    return l * l / g.x_0;
}

int find_MLP_path_table(const Geo& g, uint32_t* path, uint32_t first_voxel,
                        double x_in_object, double y_in_object, double z_in_object,
                        double x_out_object, double y_out_object, double z_out_object,
                        double xy_entry_angle, double xz_entry_angle, double xy_exit_angle, double xz_exit_angle,
                        int* skipped, MlpTrace* trace = nullptr) {
    const int C = g.C, R = g.R;
    const double STEP = g.mlp_u_step;
    const int K = (int)(2.0 * 1.4142135623730951 * g.radius / STEP) + 8;   /* steps a chord can take */
    const int Q = K + 2;
    const float x_zero = (float)g.x_zero, y_zero = (float)g.y_zero, z_zero = (float)g.z_zero;
    const float inv_w = (float)(1.0 / g.voxel_width), inv_h = (float)(1.0 / g.voxel_height);
    const float inv_t = (float)(1.0 / g.voxel_thickness);
    int num_intersections = 0;
    if (skipped) *skipped = 0;
    path[num_intersections++] = first_voxel;

    double ca = pct_cos(xy_entry_angle), sa = pct_sin(xy_entry_angle);
    double u_in_object = (ca * x_in_object) + (sa * y_in_object);
    double u_out_object = (ca * x_out_object) + (sa * y_out_object);
    double t_in_object = (ca * y_in_object) - (sa * x_in_object);
    double t_out_object = (ca * y_out_object) - (sa * x_out_object);
    if (u_in_object > u_out_object) {
        xy_entry_angle += PCT_PI;
        xy_exit_angle += PCT_PI;
        ca = pct_cos(xy_entry_angle); sa = pct_sin(xy_entry_angle);
        u_in_object = (ca * x_in_object) + (sa * y_in_object);
        u_out_object = (ca * x_out_object) + (sa * y_out_object);
        t_in_object = (ca * y_in_object) - (sa * x_in_object);
        t_out_object = (ca * y_out_object) - (sa * x_out_object);
    }
    const float f_ca = (float)ca, f_sa = (float)sa, f_u_in = (float)u_in_object;
    const float f_t_in = (float)t_in_object, f_t_out = (float)t_out_object;
    const float f_dT = (float)(xy_exit_angle - xy_entry_angle);
    const float f_v_in = (float)z_in_object, f_v_out = (float)z_out_object;
    const float f_dV = (float)(xz_exit_angle - xz_entry_angle);

    double u_0 = 0.0, u_1 = STEP;
    double u_2 = pct_fabs(u_out_object - u_in_object);
    double sigma_2_pre_1 = u_2 * u_2 * u_2 * (g.a0_3 + u_2 * (g.a1_12 + u_2 * (g.a2_30 + u_2 * (g.a3_60 + u_2 * (g.a4_105 + u_2 * g.a5_168)))));
    double sigma_2_pre_2 = u_2 * u_2 * (g.a0_2 + u_2 * (g.a1_6 + u_2 * (g.a2_12 + u_2 * (g.a3_20 + u_2 * (g.a4_30 + u_2 * g.a5_42)))));
    double sigma_2_pre_3 = u_2 * (g.A0 + u_2 * (g.a1_2 + u_2 * (g.a2_3 + u_2 * (g.a3_4 + u_2 * (g.a4_5 + u_2 * g.a5_6)))));
    double u_2_sq = u_2 * u_2, two_u_2 = 2 * u_2;
    /* sigma_2 table index: rounded at the first step, then -SCAT_SUB per step */
    double j_first = (u_2 - STEP) / (STEP / SCAT_SUB) + 0.5;
    int j0 = j_first > 0 ? (j_first < 2.0e9 ? (int)j_first : 2000000000) : 0;
    const int row = j0 % SCAT_SUB, q_first = j0 / SCAT_SUB;

    for (int k = 0; u_1 < u_2 - STEP; ++k) {
        double l1 = g.e_0 * (1 + 0.038 * pct_log((u_1 - u_0) / g.x_0));
        double sigma_1_coefficient = l1 * l1 / g.x_0;
        double sigma_t1 = sigma_1_coefficient * (u_1 * u_1 * u_1 * (g.a0_3 + u_1 * (g.a1_12 + u_1 * (g.a2_30 + u_1 * (g.a3_60 + u_1 * (g.a4_105 + u_1 * g.a5_168))))));
        double sigma_t1_theta1 = sigma_1_coefficient * (u_1 * u_1 * (g.a0_2 + u_1 * (g.a1_6 + u_1 * (g.a2_12 + u_1 * (g.a3_20 + u_1 * (g.a4_30 + u_1 * g.a5_42))))));
        double sigma_theta1 = sigma_1_coefficient * (u_1 * (g.A0 + u_1 * (g.a1_2 + u_1 * (g.a2_3 + u_1 * (g.a3_4 + u_1 * (g.a4_5 + u_1 * g.a5_6))))));
        double determinant_Sigma_1 = sigma_t1 * sigma_theta1 - sigma_t1_theta1 * sigma_t1_theta1;
        double S1I0 = sigma_theta1 / determinant_Sigma_1;
        double S1I1 = -sigma_t1_theta1 / determinant_Sigma_1;
        double S1I3 = sigma_t1 / determinant_Sigma_1;
        double common_1 = u_1 * (g.A0 + u_1 * (g.a1_2 + u_1 * (g.a2_3 + u_1 * (g.a3_4 + u_1 * (g.a4_5 + u_1 * g.a5_6)))));
        double common_2 = u_1 * u_1 * (g.a0_2 + u_1 * (g.a1_3 + u_1 * (g.a2_4 + u_1 * (g.a3_5 + u_1 * (g.a4_6 + u_1 * g.a5_7)))));
        double common_3 = u_1 * u_1 * u_1 * (g.a0_3 + u_1 * (g.a1_4 + u_1 * (g.a2_5 + u_1 * (g.a3_6 + u_1 * (g.a4_7 + u_1 * g.a5_8)))));

        int q = q_first - k;
        q = q < 0 ? 0 : (q >= Q ? Q - 1 : q);
        double c2 = depth_coefficient(g, (int64_t)q * SCAT_SUB + row);
        double d = u_2 - u_1;
        double sigma_t2 = c2 * (sigma_2_pre_1 - u_2_sq * common_1 + two_u_2 * common_2 - common_3);
        double sigma_t2_theta2 = c2 * (sigma_2_pre_2 - u_2 * common_1 + common_2);
        double sigma_theta2 = c2 * (sigma_2_pre_3 - common_1);
        double r2 = 1.0 / (sigma_t2 * sigma_theta2 - sigma_t2_theta2 * sigma_t2_theta2);
        double S2I0 = sigma_theta2 * r2, S2I1 = -sigma_t2_theta2 * r2, S2I3 = sigma_t2 * r2;
        /* (Sigma_1^-1 + R_1^T Sigma_2^-1 R_1), R_1 = {1, d; 0, 1}; symmetric */
        double e = S2I0 * d + S2I1;
        double f = S2I1 * d + S2I3;
        double F0 = S1I0 + S2I0;
        double F1 = S1I1 + e;
        double F3 = S1I3 + d * e + f;
        double rF = 1.0 / (F0 * F3 - F1 * F1);
        double I0 = F3 * rF, I1 = -F1 * rF;
        float alpha = (float)(I0 * S1I0 + I1 * S1I1);
        float beta = (float)(I0 * S2I0 + I1 * e);
        float gamma = (float)(I0 * S2I1 + I1 * f);
        float t_1 = alpha * f_t_in + beta * f_t_out + gamma * f_dT;
        float v_1 = alpha * f_v_in + beta * f_v_out + gamma * f_dV;
        float u_shifted = f_u_in + (float)u_1;
        float x_1 = f_ca * u_shifted - f_sa * t_1;
        float y_1 = f_sa * u_shifted + f_ca * t_1;
        if (trace && trace->n < trace->cap) {
            trace->u1[trace->n] = u_1; trace->t1[trace->n] = t_1; trace->v1[trace->n] = v_1;
            trace->x1[trace->n] = x_1; trace->y1[trace->n] = y_1;
            trace->n++;
        }

        int voxel_x = calculate_voxel_f(x_zero, x_1, inv_w);
        int voxel_y = calculate_voxel_f(y_zero, y_1, inv_h);
        int voxel_z = calculate_voxel_f(z_zero, v_1, inv_t);
        int64_t voxel = (int64_t)voxel_x + (int64_t)voxel_y * C + (int64_t)voxel_z * C * R;
        if (voxel_x < 0 || voxel_y < 0 || voxel_z < 0 || voxel >= g.V) {
            if (skipped) (*skipped)++;
        } else if ((uint32_t)voxel != path[num_intersections - 1]) {
            if (num_intersections >= g.max_intersections) return -1;
            path[num_intersections++] = (uint32_t)voxel;
        }
        u_1 += STEP;
    }
    return num_intersections;
}

}  // namespace

/* ======================================================================== */
/* extern "C" surface (ctypes)                                               */
/* ======================================================================== */
extern "C" {

/* A1 over n histories (.cu:2169-2300). Host arrays. */
int oracle_recon_volume_intersections(const pct_params* p, const pct_raw_soa* raw, const pct_hist_soa* out, uint8_t* missed) {
    Geo g = make_geo(*p);
    const double ANGLE_TO_RADIANS = PCT_PI / 180.0;
    const double H = p->recon_cyl_height;
    for (int64_t i = 0; i < raw->n; ++i) {
        double rotation_angle_radians = raw->gantry_angle[i] * ANGLE_TO_RADIANS;
        double cr = pct_cos(rotation_angle_radians), sr = pct_sin(rotation_angle_radians);
        float dt_in = raw->t_in_2[i] - raw->t_in_1[i], du_in = raw->u_in_2[i] - raw->u_in_1[i];
        /* atan2 of two floats resolves to the float overload in device code */
        double ut_entry_angle = pct_atan2f(dt_in, du_in);
        double u_entry, t_entry;
        bool entered = calculate_intercepts(g, raw->u_in_2[i], raw->t_in_2[i], ut_entry_angle, u_entry, t_entry);
        out->xy_entry_angle[i] = (float)(ut_entry_angle + rotation_angle_radians);
        out->x_entry[i] = (float)((cr * u_entry) - (sr * t_entry));
        out->y_entry[i] = (float)((sr * u_entry) + (cr * t_entry));

        float dt_out = raw->t_out_2[i] - raw->t_out_1[i], du_out = raw->u_out_2[i] - raw->u_out_1[i];
        double ut_exit_angle = pct_atan2f(dt_out, du_out);
        double u_exit, t_exit;
        bool exited = calculate_intercepts(g, raw->u_out_1[i], raw->t_out_1[i], ut_exit_angle, u_exit, t_exit);
        out->xy_exit_angle[i] = (float)(ut_exit_angle + rotation_angle_radians);
        out->x_exit[i] = (float)((cr * u_exit) - (sr * t_exit));
        out->y_exit[i] = (float)((sr * u_exit) + (cr * t_exit));

        float dv_in = raw->v_in_2[i] - raw->v_in_1[i], dv_out = raw->v_out_2[i] - raw->v_out_1[i];
        float du_out2 = raw->u_out_2[i] - raw->u_out_1[i];
        double uv_entry_slope = dv_in / du_in;          /* float / float */
        double uv_exit_slope = dv_out / du_out2;
        out->xz_entry_angle[i] = pct_atan2f(dv_in, du_in);
        out->xz_exit_angle[i] = pct_atan2f(dv_out, du_out2);

        u_entry = (cr * out->x_entry[i]) + (sr * out->y_entry[i]);
        u_exit = (cr * out->x_exit[i]) + (sr * out->y_exit[i]);
        out->z_entry[i] = (float)(raw->v_in_2[i] + uv_entry_slope * (u_entry - raw->u_in_2[i]));
        out->z_exit[i] = (float)(raw->v_out_1[i] - uv_exit_slope * (raw->u_out_1[i] - u_exit));

        if (entered && exited) {
            if ((pct_fabsf(out->z_entry[i]) < H * 0.5) && (pct_fabsf(out->z_exit[i]) > H * 0.5)) {
                float dz = out->z_exit[i] - out->z_entry[i];
                double recon_cyl_fraction = pct_fabs((sgn_int(out->z_exit[i]) * H * 0.5 - out->z_entry[i]) / dz);
                float dx = out->x_exit[i] - out->x_entry[i], dy = out->y_exit[i] - out->y_entry[i];
                out->x_exit[i] = (float)(out->x_entry[i] + recon_cyl_fraction * dx);
                out->y_exit[i] = (float)(out->y_entry[i] + recon_cyl_fraction * dy);
                out->z_exit[i] = (float)(sgn_int(out->z_exit[i]) * H * 0.5);
            } else if (pct_fabsf(out->z_entry[i]) > H * 0.5) {
                entered = false;
                exited = false;
            }
            if ((pct_fabsf(out->z_entry[i]) > H * 0.5) && (pct_fabsf(out->z_exit[i]) < H * 0.5)) {
                float dz = out->z_exit[i] - out->z_entry[i];
                double recon_cyl_fraction = pct_fabs((sgn_int(out->z_exit[i]) * H * 0.5 - out->z_exit[i]) / dz);
                float dx = out->x_exit[i] - out->x_entry[i], dy = out->y_exit[i] - out->y_entry[i];
                out->x_entry[i] = (float)(out->x_exit[i] + recon_cyl_fraction * dx);
                out->y_entry[i] = (float)(out->y_exit[i] + recon_cyl_fraction * dy);
                out->z_entry[i] = (float)(sgn_int(out->z_entry[i]) * H * 0.5);
            }
            if ((pct_fabsf(raw->t_out_1[i] - raw->t_in_2[i]) > 5) || (pct_fabsf(raw->v_out_1[i] - raw->v_in_2[i]) > 5)) {
                entered = false;
                exited = false;
            }
        }
        missed[i] = (uint8_t)(!entered || !exited);
        out->wepl[i] = raw->wepl[i];
    }
    return 0;
}

/* C1: initialize_hull_GPU (.cu:3316-3326) as a byte mask. */
int oracle_hull_init(const pct_params* p, uint8_t* hull) {
    Geo g = make_geo(*p);
    for (int slice = 0; slice < g.S; ++slice)
        for (int row = 0; row < g.R; ++row)
            for (int column = 0; column < g.C; ++column) {
                int64_t voxel = column + (int64_t)row * g.C + (int64_t)slice * g.C * g.R;
                double x = (column - g.C / 2 + 0.5) * g.voxel_width;
                double y = (g.R / 2 - row - 0.5) * g.voxel_height;
                hull[voxel] = (x * x + y * y < g.radius * g.radius) ? 1 : 0;
            }
    return 0;
}

/* D1: collect_MLP_endpoints (.cu:5729-5770 / 4634-4683) over n histories,
 * in place on the hist arrays, byte hull. */
int oracle_mlp_endpoints(const pct_params* p, const pct_hist_soa* h, const uint8_t* hull, uint8_t* intersected) {
    Geo g = make_geo(*p);
    for (int64_t i = 0; i < h->n; ++i) {
        double xi, yi, zi, xo, yo, zo;
        int vx, vy, vz, vxo, vyo, vzo;
        intersected[i] = 0;
        h->first_voxel[i] = 0;
        bool entered = find_MLP_endpoints(g, hull, h->x_entry[i], h->y_entry[i], h->z_entry[i], h->xy_entry_angle[i],
                                          h->xz_entry_angle[i], xi, yi, zi, vx, vy, vz, true);
        bool exited = find_MLP_endpoints(g, hull, h->x_exit[i], h->y_exit[i], h->z_exit[i], h->xy_exit_angle[i],
                                         h->xz_exit_angle[i], xo, yo, zo, vxo, vyo, vzo, false);
        if (entered && exited) {
            intersected[i] = 1;
            h->first_voxel[i] = (uint32_t)(vx + g.C * vy + g.R * g.C * vz);
            h->x_entry[i] = (float)xi; h->y_entry[i] = (float)yi; h->z_entry[i] = (float)zi;
            h->x_exit[i] = (float)xo;  h->y_exit[i] = (float)yo;  h->z_exit[i] = (float)zo;
        }
    }
    return 0;
}

/* D3 */
/* EffectiveChordLength called as the CPU research loop calls it (.cu:7539):
 * the float average of the float angles */
double oracle_chord_length_float_avg(const pct_params* p, float xy_in, float xy_out, float xz_in, float xz_out) {
    Geo g = make_geo(*p);
    float t = (xy_in + xy_out) / 2;
    float v = (xz_in + xz_out) / 2;
    return effective_chord_length(g, t, v);
}

double oracle_chord_length(const pct_params* p, float xy_in, float xy_out, float xz_in, float xz_out) {
    Geo g = make_geo(*p);
    return effective_chord_length(g, ((double)xy_in + (double)xy_out) / 2.0, ((double)xz_in + (double)xz_out) / 2.0);
}

/* D2 for one history; returns path length, -1 on overflow. */
int oracle_mlp_path(const pct_params* p, const pct_hist_soa* h, int64_t i, uint32_t* path, int* skipped) {
    Geo g = make_geo(*p);
    return find_MLP_path(g, path, h->first_voxel[i], h->x_entry[i], h->y_entry[i], h->z_entry[i],
                         h->x_exit[i], h->y_exit[i], h->z_exit[i], h->xy_entry_angle[i], h->xz_entry_angle[i],
                         h->xy_exit_angle[i], h->xz_exit_angle[i], skipped);
}

/* D2 for histories [start, start+count): CSR output with capacity stride
 * `stride` per history (path of local h at voxels[h*stride]). */
int oracle_mlp_paths(const pct_params* p, const pct_hist_soa* h, int64_t start, int64_t count,
                     uint32_t* voxels, int64_t stride, uint32_t* lengths) {
    Geo g = make_geo(*p);
    std::vector<uint32_t> buf(g.max_intersections + 1);
    int rc = 0;
    for (int64_t k = 0; k < count; ++k) {
        int64_t i = start + k;
        int n = find_MLP_path(g, buf.data(), h->first_voxel[i], h->x_entry[i], h->y_entry[i], h->z_entry[i],
                              h->x_exit[i], h->y_exit[i], h->z_exit[i], h->xy_entry_angle[i], h->xz_entry_angle[i],
                              h->xy_exit_angle[i], h->xz_exit_angle[i], nullptr);
        if (n < 0) { rc = -3; n = 0; }
        if (n > stride) { rc = -3; n = (int)stride; }
        memcpy(voxels + k * stride, buf.data(), sizeof(uint32_t) * n);
        lengths[k] = (uint32_t)n;
    }
    return rc;
}

/* D4 (precision form) for histories [start, start+count), same layout. */
int oracle_mlp_paths_table(const pct_params* p, const pct_hist_soa* h, int64_t start, int64_t count,
                           uint32_t* voxels, int64_t stride, uint32_t* lengths) {
    Geo g = make_geo(*p);
    std::vector<uint32_t> buf(g.max_intersections + 1);
    int rc = 0;
    for (int64_t k = 0; k < count; ++k) {
        int64_t i = start + k;
        int n = find_MLP_path_table(g, buf.data(), h->first_voxel[i], h->x_entry[i], h->y_entry[i], h->z_entry[i],
                                    h->x_exit[i], h->y_exit[i], h->z_exit[i], h->xy_entry_angle[i],
                                    h->xz_entry_angle[i], h->xy_exit_angle[i], h->xz_exit_angle[i], nullptr);
        if (n < 0) { rc = -3; n = 0; }
        if (n > stride) { rc = -3; n = (int)stride; }
        memcpy(voxels + k * stride, buf.data(), sizeof(uint32_t) * n);
        lengths[k] = (uint32_t)n;
    }
    return rc;
}

/* D2 (table=0) or D4 (table=1) with the per-step record (u_1, t_1, v_1, x_1,
 * y_1) of every loop iteration; returns the number of steps recorded (path
 * length in *n_path). */
int oracle_mlp_trace(const pct_params* p, const pct_hist_soa* h, int64_t i, double* u1, double* t1, double* v1,
                     double* x1, double* y1, int cap, int* n_path, int table) {
    Geo g = make_geo(*p);
    std::vector<uint32_t> path(g.max_intersections + 1);
    MlpTrace tr{u1, t1, v1, x1, y1, cap, 0};
    auto walk = table ? find_MLP_path_table : find_MLP_path;
    int n = walk(g, path.data(), h->first_voxel[i], h->x_entry[i], h->y_entry[i], h->z_entry[i],
                 h->x_exit[i], h->y_exit[i], h->z_exit[i], h->xy_entry_angle[i], h->xz_entry_angle[i],
                 h->xy_exit_angle[i], h->xz_exit_angle[i], nullptr, &tr);
    if (n_path) *n_path = n;
    return tr.n;
}

/* DROP on explicit voxel lists (DROP_blocks .cu:8023-8043 + DROP_update
 * .cu:8044-8055 with the per-history expression of .cu:5822-5834): list k is
 * voxels[offsets[k] .. offsets[k+1]), chord[k] its effective chord length,
 * b[k] its WEPL. Blocks of block_size in list order, `iterations` passes. */
int oracle_drop_lists(int64_t V, int64_t n_lists, const int64_t* offsets, const uint32_t* voxels, const double* chord,
                      const float* b, float* x, int block_size, float lambda, int iterations) {
    std::vector<double> x_update(V, 0.0);
    std::vector<uint32_t> counts(V, 0);
    std::vector<uint32_t> touched;
    for (int it = 0; it < iterations; ++it) {
        for (int64_t start = 0; start < n_lists; start += block_size) {
            int64_t end = start + block_size < n_lists ? start + block_size : n_lists;
            touched.clear();
            for (int64_t k = start; k < end; ++k) {
                int64_t n = offsets[k + 1] - offsets[k];
                const uint32_t* a = voxels + offsets[k];
                double L = chord[k];
                double sum = 0.0;
                for (int64_t j = 0; j < n; ++j) sum += x[a[j]];
                double a_i_dot_x_k = sum * L;
                double a_i_dot_a_i = (L * L) * n;
                double update = L * ((b[k] - a_i_dot_x_k) / a_i_dot_a_i) * lambda;
                for (int64_t j = 0; j < n; ++j) {
                    if (counts[a[j]] == 0) touched.push_back(a[j]);
                    x_update[a[j]] += update;
                    counts[a[j]] += 1;
                }
            }
            for (uint32_t v : touched) {
                x[v] = (float)(x[v] + x_update[v] / counts[v]);
                x_update[v] = 0.0;
                counts[v] = 0;
            }
        }
    }
    return 0;
}

/* One iteration of the CPU research loop over given lists (image_reconstruction
 * .cu:7513-7565): list k (history WEPL b[k], chord L = chord[k], path errors
 * errors[offsets[k] + i]) goes through
 *   variant 0 DROP_blocks (.cu:8023-8043), 1 DROP_blocks_red (.cu:7998-8022),
 *   2 DROP_blocks_error_red (.cu:7816-7843), 3 DROP_blocks_error (.cu:7844-7870),
 *   4 DROP_blocks_error_red_sp (.cu:7768-7815),
 * and DROP_update (.cu:8044-8055) runs when (k+1) % block_size == 0
 * (.cu:7554). x_update and counts persist between calls (the reference's
 * globals). The error variants form a.x with x_k[i], the list position, as
 * written; pow(c, 2.0) is c*c; LAMBDA is the float lambda. */
int oracle_drop_lists_research(int64_t V, int64_t n_lists, const int64_t* offsets, const uint32_t* voxels,
                               const float* errors, const double* chord, const float* b, float* x, double* x_update,
                               uint32_t* counts, int block_size, float lambda, double eta, int psi_sign, int variant) {
    const double LAMBDA = lambda, ETA = eta;
    const int PSI_SIGN = psi_sign;
    for (int64_t k = 0; k < n_lists; ++k) {
        const uint32_t* a_i = voxels + offsets[k];
        const float* row_error = errors ? errors + offsets[k] : nullptr;
        const unsigned int num_intersections = (unsigned int)(offsets[k + 1] - offsets[k]);
        const double mean_chord_length = chord[k];
        const double bi = b[k];
        if (variant == 0 || variant == 1) {
            double sum = 0;
            for (unsigned int i = 0; i < num_intersections; i++) sum += x[a_i[i]];
            double a_i_dot_x_k = mean_chord_length * sum;
            double a_i_dot_a_i = (mean_chord_length * mean_chord_length) * num_intersections;
            if (variant == 0) {
                double scaled_residual = (bi - a_i_dot_x_k) / a_i_dot_a_i;
                double update_value = mean_chord_length * LAMBDA * scaled_residual;
                for (unsigned int j = 0; j < num_intersections; j++) {
                    x_update[a_i[j]] += update_value;
                    counts[a_i[j]] += 1;
                }
            } else {
                double residual = bi - a_i_dot_x_k;
                for (unsigned int j = 0; j < num_intersections; j++) {
                    double psi_i = ((1.0 - x[a_i[j]]) * ETA) * PSI_SIGN;
                    x_update[a_i[j]] += mean_chord_length * LAMBDA * (residual / (a_i_dot_a_i + psi_i));
                    counts[a_i[j]] += 1;
                }
            }
        } else {
            auto weight = [&](unsigned int i) -> double {
                if (variant == 2) {
                    double error = row_error[i];
                    return mean_chord_length + (error * ETA);
                }
                if (variant == 3) {
                    float error = row_error[i];
                    return mean_chord_length + error * ETA * mean_chord_length;
                }
                return (i % 5 == 0) ? 1.8 : mean_chord_length;
            };
            double a_i_dot_x_k = 0.0, a_i_dot_a_i = 0.0;
            for (unsigned int i = 0; i < num_intersections; i++) {
                double c = weight(i);
                a_i_dot_x_k += (c * x[i]);
                a_i_dot_a_i += c * c;
            }
            if (variant == 3) {
                double scaled_residual = (bi - a_i_dot_x_k) / a_i_dot_a_i;
                for (unsigned int j = 0; j < num_intersections; j++) {
                    x_update[a_i[j]] += weight(j) * LAMBDA * scaled_residual;
                    counts[a_i[j]] += 1;
                }
            } else {
                double residual = bi - a_i_dot_x_k;
                for (unsigned int j = 0; j < num_intersections; j++) {
                    double psi_i = ((1.0 - x[a_i[j]]) * ETA) * PSI_SIGN;
                    x_update[a_i[j]] += weight(j) * LAMBDA * (residual / (a_i_dot_a_i + psi_i));
                    counts[a_i[j]] += 1;
                }
            }
        }
        if ((k + 1) % block_size == 0)
            for (int64_t v = 0; v < V; v++)
                if (counts[v] > 0) {
                    x[v] += x_update[v] / counts[v];
                    x_update[v] = 0;
                    counts[v] = 0;
                }
    }
    return 0;
}

/* E6 + E4 semantics: `iterations` DROP iterations over all histories in
 * storage order with blocks of `block_size`; x is updated in place.
 * x_update is double (data.h:558) and counts unsigned (DROP_blocks).
 * variant 1: DROP_blocks_red + DROP_update (.cu:8005-8055), variant 2:
 * DROP_blocks_robust2 + DROP_update_robust2 (.cu:7871-7901), with
 * ETA = p->eta and PSI_SIGN = p->psi_sign. */
int oracle_drop_variant_mt(const pct_params* p, const pct_hist_soa* h, float* x, int block_size, float lambda,
                           int iterations, int64_t max_histories, int threads, int variant) {
    /* Per block the histories' paths and update values are independent of
     * each other (they read x as it was at the start of the block), so they
     * may be computed by several threads; the accumulation into x_update and
     * the image update then run sequentially in storage order, exactly as in
     * the single-threaded DROP_blocks/DROP_update, so the result does not
     * depend on the thread count. */
    Geo g = make_geo(*p);
    if (threads < 1) threads = 1;
    const double ETA = p->eta;
    const int PSI_SIGN = p->psi_sign;
    const double LAMBDA = lambda;
    std::vector<double> x_update(g.V, 0.0);
    std::vector<uint32_t> counts(g.V, 0);
    std::vector<double> norm_Ai(variant == 2 ? g.V : 0, 0.0);
    const int stride = g.max_intersections + 1;
    std::vector<uint32_t> paths((size_t)block_size * stride);
    std::vector<int> lens(block_size);
    std::vector<double> upd(block_size), res(block_size), aa(block_size), chord(block_size);
    std::vector<uint32_t> touched;
    int64_t n_hist = h->n;
    if (max_histories >= 0 && max_histories < n_hist) n_hist = max_histories;
    int rc = 0;
    for (int it = 0; it < iterations; ++it) {
        for (int64_t start = 0; start < n_hist; start += block_size) {
            int64_t end = start + block_size < n_hist ? start + block_size : n_hist;
            int64_t cnt = end - start;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 16)
            for (int64_t k = 0; k < cnt; ++k) {
                int64_t i = start + k;
                uint32_t* path = paths.data() + k * stride;
                int n = find_MLP_path(g, path, h->first_voxel[i], h->x_entry[i], h->y_entry[i], h->z_entry[i],
                                      h->x_exit[i], h->y_exit[i], h->z_exit[i], h->xy_entry_angle[i],
                                      h->xz_entry_angle[i], h->xy_exit_angle[i], h->xz_exit_angle[i], nullptr);
                lens[k] = n;
                if (n < 0) continue;
                double L = effective_chord_length(g, ((double)h->xy_entry_angle[i] + (double)h->xy_exit_angle[i]) / 2.0,
                                                  ((double)h->xz_entry_angle[i] + (double)h->xz_exit_angle[i]) / 2.0);
                double sum = 0.0;
                for (int j = 0; j < n; ++j) sum += x[path[j]];
                double b_i = h->wepl[i];
                if (variant == 0) {
                    double a_i_dot_x_k = sum * L;
                    double a_i_dot_a_i = (L * L) * n;
                    upd[k] = L * ((b_i - a_i_dot_x_k) / a_i_dot_a_i) * lambda;
                } else {
                    /* scalar_dot_product2 (.cu:7686): scalar * sum; pow(L, 2.0)
                     * taken as the square rounded once, L * L */
                    double a_i_dot_x_k = L * sum;
                    double a_i_dot_a_i = (L * L) * n;
                    chord[k] = L;
                    res[k] = b_i - a_i_dot_x_k;
                    aa[k] = a_i_dot_a_i;
                    upd[k] = L * LAMBDA * (res[k] / a_i_dot_a_i);   /* robust2's update_value */
                }
            }
            touched.clear();
            for (int64_t k = 0; k < cnt; ++k) {
                int n = lens[k];
                if (n < 0) { rc = -3; continue; }
                const uint32_t* path = paths.data() + k * stride;
                for (int j = 0; j < n; ++j) {
                    uint32_t v = path[j];
                    if (counts[v] == 0) touched.push_back(v);
                    if (variant == 1) {
                        double psi_i = ((1.0 - x[v]) * ETA) * PSI_SIGN;
                        x_update[v] += chord[k] * LAMBDA * (res[k] / (aa[k] + psi_i));
                    } else {
                        x_update[v] += upd[k];
                    }
                    if (variant == 2) norm_Ai[v] += chord[k] * chord[k];
                    counts[v] += 1;
                }
            }
            /* DROP_update restricted to touched voxels (equal to the full scan) */
            for (uint32_t v : touched) {
                if (variant == 2) {
                    double psi_i = PSI_SIGN * (1 - x[v]) * ETA;
                    x[v] += (norm_Ai[v] / (norm_Ai[v] + psi_i)) * x_update[v] / counts[v];
                    norm_Ai[v] = 0.0;
                } else {
                    x[v] = (float)(x[v] + x_update[v] / counts[v]);
                }
                x_update[v] = 0.0;
                counts[v] = 0;
            }
        }
    }
    return rc;
}

int oracle_drop_mt(const pct_params* p, const pct_hist_soa* h, float* x, int block_size, float lambda,
                   int iterations, int64_t max_histories, int threads) {
    return oracle_drop_variant_mt(p, h, x, block_size, lambda, iterations, max_histories, threads, 0);
}

int oracle_drop(const pct_params* p, const pct_hist_soa* h, float* x, int block_size, float lambda,
                int iterations, int64_t max_histories) {
    return oracle_drop_mt(p, h, x, block_size, lambda, iterations, max_histories, 1);
}

}  // extern "C"
