/*
 * pct_intersect.hip -- (a) recon-volume intersection and the synthetic scan
 * generator.
 *
 * pct_recon_volume_intersections replaces recon_volume_intersections_GPU
 * (.cu:2169-2300) + calculate_intercepts (.cu:2301-2387): one thread per
 * history, coalesced SoA loads/stores. The arithmetic keeps the reference's
 * float/double typing: differences of float tracker coordinates are float
 * subtractions, atan2 of two floats is the float overload, and every value the
 * reference stores to a float array is rounded to float before it is read
 * again (e.g. x_entry before u_entry, .cu:2261-2263). This keeps the output
 * bit-identical to a CPU evaluation of the same expressions.
 */
#include "pct_internal.h"
#include "pct_synth.h"
#include "../../include/pct_synth_api.h"

namespace {

struct intercept_t { double u, t; bool hit; };

/* line through (u, t) with angle ut_angle against the cylinder x^2+y^2=R^2;
 * nearest intersection to (u, t) (.cu:2301-2387) */
__device__ __forceinline__ intercept_t intercept(double R, double u, double t, double ut_angle) {
    const double PIO4 = 0.7853981633974483;
    bool in_cone = ((ut_angle > PIO4) && (ut_angle < 3 * PIO4)) || ((ut_angle > 5 * PIO4) && (ut_angle < 7 * PIO4));
    if (in_cone) {
        double tmp = u;
        u = -t;
        t = tmp;
        ut_angle += 2 * PIO4;
    }
    double m = pct_tan(ut_angle);
    double b_in = t - m * u;
    double a = 1 + m * m;
    double b = 2 * m * b_in;
    double c = b_in * b_in - R * R;
    double disc = b * b - (4 * a * c);
    intercept_t r;
    r.hit = disc > 0;
    r.u = 0.0;
    r.t = 0.0;
    if (r.hit) {
        double sq = pct_sqrt(disc);
        double u1 = (sq - b) / (2 * a);
        double u2 = (sq + b) / (2 * a);
        double t1 = m * u1 + b_in;
        double t2 = m * u2 - b_in;
        double e1 = u1 - u, f1 = t1 - t, e2 = u2 + u, f2 = t2 + t;
        double d1 = e1 * e1 + f1 * f1;
        double d2 = e2 * e2 + f2 * f2;
        double s1 = (double)(d1 <= d2), s2 = (double)(d1 > d2);
        r.u = u1 * s1 - u2 * s2;
        r.t = t1 * s1 - t2 * s2;
    }
    if (in_cone) {
        double tmp = r.u;
        r.u = r.t;
        r.t = -tmp;
    }
    return r;
}

__device__ __forceinline__ double sgn_of(float z) { return (double)((z >= 0) - (z < 0)); }

__global__ void __launch_bounds__(256)
k_recon_volume_intersections(pct_geo g, pct_raw_soa raw, pct_hist_soa out, uint8_t* __restrict__ missed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= raw.n) return;
    const double ANGLE_TO_RADIANS = PCT_PI / 180.0;
    const double H = g.height;
    float t_in_1 = raw.t_in_1[i], t_in_2 = raw.t_in_2[i], t_out_1 = raw.t_out_1[i], t_out_2 = raw.t_out_2[i];
    float u_in_1 = raw.u_in_1[i], u_in_2 = raw.u_in_2[i], u_out_1 = raw.u_out_1[i], u_out_2 = raw.u_out_2[i];
    float v_in_1 = raw.v_in_1[i], v_in_2 = raw.v_in_2[i], v_out_1 = raw.v_out_1[i], v_out_2 = raw.v_out_2[i];

    double rot = raw.gantry_angle[i] * ANGLE_TO_RADIANS;
    double sr, cr;
    pct_sincos(rot, &sr, &cr);

    float du_in = u_in_2 - u_in_1, du_out = u_out_2 - u_out_1;
    double ut_entry_angle = pct_atan2f(t_in_2 - t_in_1, du_in);
    intercept_t en = intercept(g.radius, u_in_2, t_in_2, ut_entry_angle);
    float xy_entry = (float)(ut_entry_angle + rot);
    float x_entry = (float)((cr * en.u) - (sr * en.t));
    float y_entry = (float)((sr * en.u) + (cr * en.t));

    double ut_exit_angle = pct_atan2f(t_out_2 - t_out_1, du_out);
    intercept_t ex = intercept(g.radius, u_out_1, t_out_1, ut_exit_angle);
    float xy_exit = (float)(ut_exit_angle + rot);
    float x_exit = (float)((cr * ex.u) - (sr * ex.t));
    float y_exit = (float)((sr * ex.u) + (cr * ex.t));

    float dv_in = v_in_2 - v_in_1, dv_out = v_out_2 - v_out_1;
    double uv_entry_slope = dv_in / du_in;
    double uv_exit_slope = dv_out / du_out;
    float xz_entry = pct_atan2f(dv_in, du_in);
    float xz_exit = pct_atan2f(dv_out, du_out);

    double u_entry = (cr * x_entry) + (sr * y_entry);
    double u_exit = (cr * x_exit) + (sr * y_exit);
    float z_entry = (float)(v_in_2 + uv_entry_slope * (u_entry - u_in_2));
    float z_exit = (float)(v_out_1 - uv_exit_slope * (u_out_1 - u_exit));

    bool entered = en.hit, exited = ex.hit;
    if (entered && exited) {
        if ((pct_fabsf(z_entry) < H * 0.5) && (pct_fabsf(z_exit) > H * 0.5)) {
            float dz = z_exit - z_entry;
            double frac = pct_fabs((sgn_of(z_exit) * H * 0.5 - z_entry) / dz);
            float dx = x_exit - x_entry, dy = y_exit - y_entry;
            x_exit = (float)(x_entry + frac * dx);
            y_exit = (float)(y_entry + frac * dy);
            z_exit = (float)(sgn_of(z_exit) * H * 0.5);
        } else if (pct_fabsf(z_entry) > H * 0.5) {
            entered = false;
            exited = false;
        }
        if ((pct_fabsf(z_entry) > H * 0.5) && (pct_fabsf(z_exit) < H * 0.5)) {
            float dz = z_exit - z_entry;
            double frac = pct_fabs((sgn_of(z_exit) * H * 0.5 - z_exit) / dz);
            float dx = x_exit - x_entry, dy = y_exit - y_entry;
            x_entry = (float)(x_exit + frac * dx);
            y_entry = (float)(y_exit + frac * dy);
            z_entry = (float)(sgn_of(z_entry) * H * 0.5);
        }
        if ((pct_fabsf(t_out_1 - t_in_2) > 5) || (pct_fabsf(v_out_1 - v_in_2) > 5)) {
            entered = false;
            exited = false;
        }
    }
    missed[i] = (uint8_t)(!entered || !exited);
    out.x_entry[i] = x_entry; out.y_entry[i] = y_entry; out.z_entry[i] = z_entry;
    out.x_exit[i] = x_exit;   out.y_exit[i] = y_exit;   out.z_exit[i] = z_exit;
    out.xy_entry_angle[i] = xy_entry; out.xz_entry_angle[i] = xz_entry;
    out.xy_exit_angle[i] = xy_exit;   out.xz_exit_angle[i] = xz_exit;
    out.wepl[i] = raw.wepl[i];
}

__global__ void __launch_bounds__(256)
k_synth(pct_synth_cfg cfg, int64_t start, int64_t count, pct_synth_soa out) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    pct_synth_record r;
    pct_synth_history(&cfg, start + k, &r);
    out.gantry_angle[k] = r.gantry_angle;
    out.t_in_1[k] = r.t_in_1; out.t_in_2[k] = r.t_in_2; out.t_out_1[k] = r.t_out_1; out.t_out_2[k] = r.t_out_2;
    out.u_in_1[k] = r.u_in_1; out.u_in_2[k] = r.u_in_2; out.u_out_1[k] = r.u_out_1; out.u_out_2[k] = r.u_out_2;
    out.v_in_1[k] = r.v_in_1; out.v_in_2[k] = r.v_in_2; out.v_out_1[k] = r.v_out_1; out.v_out_2[k] = r.v_out_2;
    out.wepl[k] = r.wepl;
}

}  // namespace

extern "C" int pct_recon_volume_intersections(pct_ctx* ctx, const pct_raw_soa* raw, const pct_hist_soa* out,
                                              uint8_t* missed) {
    if (!ctx || !raw || !out) return pct_fail(ctx, PCT_E_INVALID, "pct_recon_volume_intersections: null argument");
    if (raw->n != out->n || raw->n < 0)
        return pct_fail(ctx, PCT_E_INVALID, "pct_recon_volume_intersections: raw->n (%lld) != out->n (%lld)",
                        (long long)raw->n, (long long)out->n);
    if (raw->n == 0) return PCT_OK;                 /* no history: the arrays are not read */
    if (!missed) return pct_fail(ctx, PCT_E_INVALID, "pct_recon_volume_intersections: null missed");
    hipLaunchKernelGGL(k_recon_volume_intersections, dim3(pct_grid(raw->n, 256)), dim3(256), 0, ctx->stream,
                       ctx->g, *raw, *out, missed);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

/* ------------------------------------------------------------------------- */
/* synthetic scans                                                            */
/* ------------------------------------------------------------------------- */
extern "C" int pct_synth_generate(pct_ctx* ctx, const pct_synth_cfg* cfg, int64_t start, int64_t count,
                                  const pct_synth_soa* out) {
    if (!ctx || !cfg || !out) return pct_fail(ctx, PCT_E_INVALID, "pct_synth_generate: null argument");
    if (cfg->n_cyl < 0 || cfg->n_cyl > PCT_SYNTH_MAX_CYL || cfg->n_angles <= 0 || start < 0 || count < 0 ||
        start + count > cfg->n_histories)
        return pct_fail(ctx, PCT_E_INVALID, "pct_synth_generate: bad configuration or range");
    if (count == 0) return PCT_OK;
    hipLaunchKernelGGL(k_synth, dim3(pct_grid(count, 256)), dim3(256), 0, ctx->stream, *cfg, start, count, *out);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_synth_generate_host(const pct_synth_cfg* cfg, int64_t start, int64_t count, const pct_synth_soa* out) {
    if (!cfg || !out) return PCT_E_INVALID;
    if (cfg->n_cyl < 0 || cfg->n_cyl > PCT_SYNTH_MAX_CYL || cfg->n_angles <= 0 || start < 0 || count < 0 ||
        start + count > cfg->n_histories)
        return PCT_E_INVALID;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < count; ++k) {
        pct_synth_record r;
        pct_synth_history(cfg, start + k, &r);
        out->gantry_angle[k] = r.gantry_angle;
        out->t_in_1[k] = r.t_in_1; out->t_in_2[k] = r.t_in_2; out->t_out_1[k] = r.t_out_1; out->t_out_2[k] = r.t_out_2;
        out->u_in_1[k] = r.u_in_1; out->u_in_2[k] = r.u_in_2; out->u_out_1[k] = r.u_out_1; out->u_out_2[k] = r.u_out_2;
        out->v_in_1[k] = r.v_in_1; out->v_in_2[k] = r.v_in_2; out->v_out_1[k] = r.v_out_1; out->v_out_2[k] = r.v_out_2;
        out->wepl[k] = r.wepl;
    }
    return PCT_OK;
}
