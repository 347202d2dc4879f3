/*
 * pct_fbp.hip -- (b) filtered backprojection, FBP hull and the initial iterate.
 *
 * filter_GPU (.cu:3036-3069): one workgroup per (v, angle) sinogram row, the
 * row staged in LDS together with the per-reference-bin cone-beam weights and
 * the Shepp-Logan kernel values of every bin separation (the reference form
 * 1/((dt pi)^2 (1 - 4 s^2)), .cu:3064), so the O(T^2) convolution reads only
 * LDS. Each output bin accumulates over the reference bins in increasing
 * order, rounding to float after every term as the reference's float array
 * does.
 * backprojection_GPU (.cu:3155-3230): one thread per voxel, the 90 projection
 * angles' sin/cos from LDS, bilinear interpolation in (t, v) with the
 * (D/(D+u))^2 weight, float accumulation in angle order; voxels outside the
 * reconstruction cylinder get RSP_AIR.
 * FBP_image_2_hull_GPU (.cu:3250-3261) and the HYBRID initial iterate
 * x0 = FBP * hull (define_initial_iterate, .cu:7191-7224).
 */
#include "pct_internal.h"

namespace {

struct fbp_geo {
    int T, V, A;
    double ts, vs, sr, delta;
    int64_t bins;
};

#define FBP_MAX_T 1024

__global__ void __launch_bounds__(256)
k_fbp_filter(fbp_geo f, const float* __restrict__ sino, float* __restrict__ out) {
    __shared__ float row[FBP_MAX_T];
    __shared__ double scale[FBP_MAX_T];
    __shared__ double kern[2 * FBP_MAX_T];
    const int v_bin = blockIdx.x, angle_bin = blockIdx.y;
    const int T = f.T;
    const int64_t strip = (int64_t)v_bin * f.A * T + (int64_t)angle_bin * T;
    const double v = (v_bin - f.V / 2) * f.vs + f.vs / 2.0;
    for (int k = threadIdx.x; k < T; k += blockDim.x) {
        row[k] = sino[strip + k];
        double t = (k - T / 2) * f.ts + f.ts / 2.0;
        scale[k] = f.sr / pct_sqrt(f.sr * f.sr + t * t + v * v);
    }
    for (int k = threadIdx.x; k < 2 * T - 1; k += blockDim.x) {
        int sep = k - (T - 1);
        kern[k] = 1 / ((f.ts * PCT_PI * f.ts * PCT_PI) * (1.0 - (2 * sep * 2 * sep)));
    }
    __syncthreads();
    for (int t_bin = threadIdx.x; t_bin < T; t_bin += blockDim.x) {
        float acc = 0.0f;
        for (int t_ref = 0; t_ref < T; ++t_ref)
            acc = (float)(acc + f.ts * row[t_ref] * kern[t_bin - t_ref + (T - 1)] * scale[t_ref]);
        out[strip + t_bin] = acc;
    }
}

__global__ void __launch_bounds__(256)
k_backproject(pct_geo g, fbp_geo f, float rsp_air, double slice_thickness, const float* __restrict__ sf,
              float* __restrict__ image) {
    __shared__ double s_sin[512], s_cos[512];
    for (int a = threadIdx.x; a < f.A; a += blockDim.x) pct_sincos(a * f.delta, &s_sin[a], &s_cos[a]);
    __syncthreads();
    int64_t voxel = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (voxel >= g.V) return;
    int column = (int)(voxel % g.C), row = (int)((voxel / g.C) % g.R), slice = (int)(voxel / ((int64_t)g.C * g.R));
    double x = -g.radius + (column + 0.5) * g.voxel_width;
    double y = g.radius - (row + 0.5) * g.voxel_height;
    double z = -g.height / 2.0 + (slice + 0.5) * slice_thickness;
    if ((x * x + y * y) > (g.radius * g.radius)) {
        image[voxel] = rsp_air;
        return;
    }
    const int64_t AT = (int64_t)f.A * f.T;
    float acc = 0.0f;
    for (int a = 0; a < f.A; ++a) {
        double s = s_sin[a], c = s_cos[a];
        double u = x * c + y * s;
        double t = -x * s + y * c;
        double v = z;
        double dnt = (t - u * (t / (f.sr + u))) / f.ts + f.T / 2.0;
        int t_bin = (int)dnt;
        if (t_bin > dnt) t_bin -= 1;
        double eta = dnt - t_bin;
        double dnv = (v - u * (v / (f.sr + u))) / f.vs + f.V / 2.0;
        int v_bin = (int)dnv;
        if (v_bin > dnv) v_bin -= 1;
        double eps = dnv - v_bin;
        double q = f.sr / (f.sr + u);
        double scale = q * q;
        int64_t bin = t_bin + (int64_t)a * f.T + (int64_t)v_bin * AT;
        auto at = [&](int64_t b) -> double { return (b >= 0 && b < f.bins) ? (double)sf[b] : 0.0; };
        if (v_bin == f.V - 1 || bin < 0)
            acc = (float)(acc + scale * (((1 - eta) * at(bin)) + (eta * at(bin + 1))));
        else
            acc = (float)(acc + scale * ((1 - eta) * (1 - eps) * at(bin) + eta * (1 - eps) * at(bin + 1)
                                         + (1 - eta) * eps * at(bin + AT) + eta * eps * at(bin + AT + 1)));
    }
    image[voxel] = (float)(acc * f.delta);
}

__global__ void __launch_bounds__(256)
k_fbp_threshold(pct_geo g, double threshold, const float* __restrict__ image, uint32_t* __restrict__ bits) {
    int64_t voxel = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool on = false;
    if (voxel < g.V) {
        int column = (int)(voxel % g.C), row = (int)((voxel / g.C) % g.R);
        double x = -g.radius + (column + 0.5) * g.voxel_width;
        double y = g.radius - (row + 0.5) * g.voxel_height;
        double d2 = x * x + y * y;
        on = image[voxel] > threshold && d2 < g.radius * g.radius;
    }
    uint64_t m = __ballot(on);
    int lane = threadIdx.x & 63;
    if ((lane & 31) == 0 && voxel < g.V) bits[voxel >> 5] = (uint32_t)(lane == 0 ? m : (m >> 32));
}

__global__ void k_initial_iterate(int64_t V, const float* __restrict__ fbp, const uint32_t* __restrict__ bits,
                                  float* __restrict__ x0) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    float h = ((bits[v >> 5] >> (v & 31)) & 1u) ? 1.0f : 0.0f;
    x0[v] = fbp[v] * h;
}

}  // namespace

static fbp_geo make_fbp_geo(const pct_params& p) {
    fbp_geo f;
    f.T = p.t_bins; f.V = p.v_bins; f.A = p.angular_bins;
    f.ts = p.t_bin_size; f.vs = p.v_bin_size; f.sr = p.source_radius;
    f.delta = p.gantry_angle_interval * (PCT_PI / 180.0);
    f.bins = (int64_t)f.T * f.V * f.A;
    return f;
}

extern "C" int pct_fbp_filter(pct_ctx* ctx, const float* sinogram, float* filtered) {
    if (!ctx || !sinogram || !filtered || sinogram == filtered) return pct_fail(ctx, PCT_E_INVALID, "pct_fbp_filter: bad arguments");
    fbp_geo f = make_fbp_geo(ctx->p);
    if (f.T > FBP_MAX_T) return pct_fail(ctx, PCT_E_INVALID, "pct_fbp_filter: more than %d t bins", FBP_MAX_T);
    hipLaunchKernelGGL(k_fbp_filter, dim3(f.V, f.A), dim3(256), 0, ctx->stream, f, sinogram, filtered);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_backproject(pct_ctx* ctx, const float* filtered, float* image) {
    if (!ctx || !filtered || !image) return pct_fail(ctx, PCT_E_INVALID, "pct_backproject: null argument");
    fbp_geo f = make_fbp_geo(ctx->p);
    if (f.A > 512) return pct_fail(ctx, PCT_E_INVALID, "pct_backproject: more than 512 angles");
    hipLaunchKernelGGL(k_backproject, dim3(pct_grid(ctx->g.V, 256)), dim3(256), 0, ctx->stream, ctx->g, f,
                       (float)ctx->p.rsp_air, ctx->p.slice_thickness, filtered, image);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_fbp_threshold(pct_ctx* ctx, const float* image, uint32_t* hull_bits) {
    if (!ctx || !image || !hull_bits) return pct_fail(ctx, PCT_E_INVALID, "pct_fbp_threshold: null argument");
    hipLaunchKernelGGL(k_fbp_threshold, dim3(pct_grid(ctx->g.V, 256)), dim3(256), 0, ctx->stream, ctx->g,
                       ctx->p.fbp_threshold, image, hull_bits);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_initial_iterate(pct_ctx* ctx, const float* fbp_image, const uint32_t* hull_bits, float* x0) {
    if (!ctx || !fbp_image || !hull_bits || !x0) return pct_fail(ctx, PCT_E_INVALID, "pct_initial_iterate: null argument");
    hipLaunchKernelGGL(k_initial_iterate, dim3(pct_grid(ctx->g.V, 256)), dim3(256), 0, ctx->stream, ctx->g.V, fbp_image,
                       hull_bits, x0);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}
