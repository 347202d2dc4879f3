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
 * backprojection_GPU (.cu:3155-3230): a workgroup per 16 x 16 voxel tile of a
 * slice, the 90 projection angles' sin/cos in LDS and, angle by angle, the
 * tile's footprint on the filtered sinogram staged in LDS; bilinear
 * interpolation in (t, v) with the (D/(D+u))^2 weight, float accumulation in
 * angle order; voxels outside the reconstruction cylinder get RSP_AIR.
 * FBP_image_2_hull_GPU (.cu:3250-3261) and the HYBRID initial iterate
 * x0 = FBP * hull (define_initial_iterate, .cu:7191-7224).
 */
#include <climits>
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

/* Backprojection over tiles: a workgroup of 256 threads takes 16 x 16 voxels
 * of one slice. The detector position a voxel maps to at an angle,
 * t (SR / (SR + u)) and z (SR / (SR + u)) with u, t linear in the voxel
 * centre, is a linear-fractional (hence monotone along every line) function
 * of the centre, so over a tile its extremes are at the tile's four corner
 * voxels: the footprint of the tile at each angle -- the t bins and v rows of
 * its voxels' bilinear neighbours -- is taken from the corners, widened by one
 * bin each way against rounding. The footprints of chunks of consecutive
 * angles -- as many as fit BP_STAGE floats in slots of the largest
 * footprint's size -- are staged in LDS with coalesced loads, as linear bin
 * ranges v AT + a T + t, so a t neighbour past the end of a row reads the
 * next row as the reference's index arithmetic does and a bin outside the
 * sinogram reads 0; each voxel then reads its four values from LDS. A voxel
 * whose neighbours would fall outside its angle's staged footprint reads from
 * global memory, and so does every voxel when one footprint alone exceeds the
 * staging area (coarse grids). The arithmetic and the accumulation are the
 * reference's, so the image is the same bit for bit. The kernel is bound by
 * its FP64 arithmetic (five correctly rounded divisions per voxel and angle),
 * not by the sinogram reads the staging takes off the memory path. */
#define BP_TILE 16
#define BP_STAGE 4096      /* staged floats (16 KB): several workgroups per CU */

struct bp_pos {            /* a voxel's bilinear position at one angle */
    int t_bin, v_bin;
    double eta, eps, scale;
};

__device__ __forceinline__ bp_pos bp_project(const fbp_geo& f, double x, double y, double z, double s, double c) {
    bp_pos r;
    double u = x * c + y * s;
    double t = -x * s + y * c;
    double v = z;
    double dnt = (t - u * (t / (f.sr + u))) / f.ts + f.T / 2.0;
    r.t_bin = (int)dnt;
    if (r.t_bin > dnt) r.t_bin -= 1;
    r.eta = dnt - r.t_bin;
    double dnv = (v - u * (v / (f.sr + u))) / f.vs + f.V / 2.0;
    r.v_bin = (int)dnv;
    if (r.v_bin > dnv) r.v_bin -= 1;
    r.eps = dnv - r.v_bin;
    double q = f.sr / (f.sr + u);
    r.scale = q * q;
    return r;
}

__global__ void __launch_bounds__(256)
k_backproject(pct_geo g, fbp_geo f, float rsp_air, double slice_thickness, const float* __restrict__ sf,
              float* __restrict__ image) {
    /* LDS (dynamic, sized by the angle count A): sin, cos [A] doubles;
     * corner [A][4][2] and footprint [A][5] ints; then BP_STAGE floats of
     * staged footprints */
    extern __shared__ __align__(16) unsigned char bp_lds[];
    const int A = f.A;
    double* s_sin = (double*)bp_lds;
    double* s_cos = s_sin + A;
    int (*corner)[4][2] = (int (*)[4][2])(s_cos + A);
    int (*fp)[5] = (int (*)[5])(corner + A);
    float* stage = (float*)(fp + A);
    __shared__ int max_size;
    if (threadIdx.x == 0) max_size = 0;
    for (int a = threadIdx.x; a < A; a += blockDim.x) pct_sincos(a * f.delta, &s_sin[a], &s_cos[a]);
    const int tx0 = blockIdx.x * BP_TILE, ty0 = blockIdx.y * BP_TILE, slice = blockIdx.z;
    const double z = -g.height / 2.0 + (slice + 0.5) * slice_thickness;
    __syncthreads();
    /* the tile's corner voxels at every angle */
    for (int k = threadIdx.x; k < 4 * A; k += blockDim.x) {
        const int a = k >> 2, cq = k & 3;
        const int col = (cq & 1) ? min(g.C - 1, tx0 + BP_TILE - 1) : tx0;
        const int row = (cq & 2) ? min(g.R - 1, ty0 + BP_TILE - 1) : ty0;
        const double cx = -g.radius + (col + 0.5) * g.voxel_width;
        const double cy = g.radius - (row + 0.5) * g.voxel_height;
        const bp_pos r = bp_project(f, cx, cy, z, s_sin[a], s_cos[a]);
        corner[a][cq][0] = r.t_bin;
        corner[a][cq][1] = r.v_bin;
    }
    __syncthreads();
    /* each angle's footprint: rows v_bin..v_bin+1 and bins t_bin..t_bin+1 of
     * the tile's voxels, one bin of margin each way */
    for (int a = threadIdx.x; a < A; a += blockDim.x) {
        int tmin = INT_MAX, tmax = INT_MIN, vmin = INT_MAX, vmax = INT_MIN;
        for (int cq = 0; cq < 4; ++cq) {
            tmin = min(tmin, corner[a][cq][0]); tmax = max(tmax, corner[a][cq][0]);
            vmin = min(vmin, corner[a][cq][1]); vmax = max(vmax, corner[a][cq][1]);
        }
        tmin -= 1;
        vmin -= 1;
        const int nt = tmax + 3 - tmin, nv = vmax + 3 - vmin;
        fp[a][0] = tmin; fp[a][1] = vmin; fp[a][2] = nt; fp[a][3] = nv;
        atomicMax(&max_size, nt * nv);
    }
    __syncthreads();
    /* chunks of K consecutive angles, each footprint in a slot of the largest
     * footprint's size; none staged when one footprint exceeds the area */
    const int slot = max_size;
    const int K = slot <= BP_STAGE ? BP_STAGE / slot : 0;
    const int lx = threadIdx.x % BP_TILE, ly = threadIdx.x / BP_TILE;
    const int column = tx0 + lx, row = ty0 + ly;
    const int64_t voxel = (int64_t)slice * g.C * g.R + (int64_t)row * g.C + column;
    const bool in_grid = column < g.C && row < g.R;
    const double x = -g.radius + (column + 0.5) * g.voxel_width;
    const double y = g.radius - (row + 0.5) * g.voxel_height;
    const bool active = in_grid && !((x * x + y * y) > (g.radius * g.radius));
    const int64_t AT = (int64_t)A * f.T;
    auto at = [&](int64_t b) -> double { return (b >= 0 && b < f.bins) ? (double)sf[b] : 0.0; };
    __syncthreads();
    float acc = 0.0f;
    const int step = K > 0 ? K : A;
    for (int a0 = 0; a0 < A; a0 += step) {
        const int a1 = min(A, a0 + step);
        if (K > 0) {
            if (a0 > 0) __syncthreads();       /* the previous chunk's reads are done */
            for (int a = a0; a < a1; ++a) {
                const int tmin = fp[a][0], vmin = fp[a][1], nt = fp[a][2], size = nt * fp[a][3];
                float* dst = stage + (a - a0) * slot;
                for (int k = threadIdx.x; k < size; k += blockDim.x) {
                    const int r = k / nt, cidx = k - r * nt;
                    const int64_t b = (int64_t)(vmin + r) * AT + (int64_t)a * f.T + (tmin + cidx);
                    dst[k] = (b >= 0 && b < f.bins) ? sf[b] : 0.0f;
                }
            }
            __syncthreads();
        }
        if (active) {
            for (int a = a0; a < a1; ++a) {
                const bp_pos r = bp_project(f, x, y, z, s_sin[a], s_cos[a]);
                const int64_t bin = r.t_bin + (int64_t)a * f.T + (int64_t)r.v_bin * AT;
                const int dt = r.t_bin - fp[a][0], dv = r.v_bin - fp[a][1], nt = fp[a][2];
                double s00, s01, s10, s11;
                if (K > 0 && dt >= 0 && dt <= nt - 2 && dv >= 0 && dv <= fp[a][3] - 2) {
                    const float* pr = stage + (a - a0) * slot + dv * nt + dt;
                    s00 = pr[0];
                    s01 = pr[1];
                    s10 = pr[nt];
                    s11 = pr[nt + 1];
                } else {
                    s00 = at(bin);
                    s01 = at(bin + 1);
                    s10 = at(bin + AT);
                    s11 = at(bin + AT + 1);
                }
                if (r.v_bin == f.V - 1 || bin < 0)
                    acc = (float)(acc + r.scale * (((1 - r.eta) * s00) + (r.eta * s01)));
                else
                    acc = (float)(acc + r.scale * ((1 - r.eta) * (1 - r.eps) * s00 + r.eta * (1 - r.eps) * s01
                                                   + (1 - r.eta) * r.eps * s10 + r.eta * r.eps * s11));
            }
        }
    }
    if (in_grid) image[voxel] = active ? (float)(acc * f.delta) : rsp_air;
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
    if (f.A <= 0 || f.A > 512) return pct_fail(ctx, PCT_E_INVALID, "pct_backproject: %d angles (1..512)", f.A);
    const dim3 grid((unsigned)((ctx->g.C + BP_TILE - 1) / BP_TILE), (unsigned)((ctx->g.R + BP_TILE - 1) / BP_TILE),
                    (unsigned)ctx->g.S);
    const size_t lds = (size_t)f.A * (2 * sizeof(double) + 8 * sizeof(int) + 5 * sizeof(int)) +
                       BP_STAGE * sizeof(float);
    hipLaunchKernelGGL(k_backproject, grid, dim3(BP_TILE * BP_TILE), lds, ctx->stream, ctx->g, f, (float)ctx->p.rsp_air,
                       ctx->p.slice_thickness, filtered, image);
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
