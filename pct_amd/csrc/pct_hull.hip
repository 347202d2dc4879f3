/*
 * pct_hull.hip -- (c) space carving, modified space carving, MSC edge
 * detection and hull selection with the averaging filter.
 *
 * Replaces SC_GPU / MSC_GPU (.cu:3338-3558, straight voxel walk
 * take_3D_step_GPU / take_2D_step_GPU .cu:8754-8867), MSC_edge_detection_GPU
 * (.cu:3569-3596), hull_selection (.cu:3954-3990) and averaging_filter_GPU
 * (.cu:4007-4029), and the space-modelling hull SM_GPU / SM_edge_detection
 * (.cu:3598-3789). Hulls are bit-packed (1 bit per voxel): SC clears bits with
 * atomicAnd, MSC counts with integer atomics (both order-independent, so the
 * result is deterministic), the filtered hull is packed with a wavefront
 * ballot. The reference's edge detection rewrites the counts in place while
 * other workgroups still read them (.cu:3582 vs 3589); here it reads the counts
 * and writes a separate array, the race-free meaning of that kernel.
 */
#include "pct_internal.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace {

__device__ __forceinline__ int calc_voxel(double zero, double pos, double size) {
    double q = pct_fabs(pos - zero) / size;
    return q < 1.0e9 ? (int)q : -1;
}
__device__ __forceinline__ double edge_coord(double zero, int voxel_entered, double size, int inc, int step) {
    int on_edge = (step == inc) ? voxel_entered : voxel_entered + 1;
    return zero + (inc * on_edge * size);
}
__device__ __forceinline__ double dist_rem(double zero, double pos, int inc, int step, double size, int cur) {
    return pct_fabs(edge_coord(zero, cur + inc * step, size, inc, step) - pos);
}
__device__ __forceinline__ double corr(double m, double x, double x0, double y0) { return m * (x - x0) + y0; }

/* the carving walk from entry to exit: visit() for every voxel before the exit voxel */
template <class F>
__device__ void carve_walk(const pct_geo& g, float xe, float ye, float ze, float xx, float yx, float zx, F visit) {
    const int C = g.C, R = g.R, S = g.S;
    const int X_INC = 1, Y_INC = -1, Z_INC = -1;
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
    double dy_dx = (yx - ye) / (xx - xe);
    double dz_dx = (zx - ze) / (xx - xe);
    double dz_dy = (zx - ze) / (yx - ye);
    double dx_dy = (xx - xe) / (yx - ye);
    double dx_dz = (xx - xe) / (zx - ze);
    double dy_dz = (yx - ye) / (zx - ze);
    int vxo = calc_voxel(g.x_zero, xx, g.voxel_width);
    int vyo = calc_voxel(g.y_zero, yx, g.voxel_height);
    int vzo = calc_voxel(g.z_zero, zx, g.voxel_thickness);
    int64_t voxel_out = vxo + (int64_t)vyo * C + (int64_t)vzo * C * R;
    bool end_walk = (voxel == voxel_out) || (vx >= C) || (vy >= R) || (vz >= S) || vx < 0 || vy < 0 || vz < 0;
    if (!end_walk) visit(voxel);
    int guard = 4 * (C + R + S) + 16;
    while (!end_walk && guard-- > 0) {
        if (zm != 0) {
            double x_ext = pct_fabs(dz_dx) * x_to_go;
            double y_ext = pct_fabs(dz_dy) * y_to_go;
            if ((z_to_go <= x_ext) && (z_to_go <= y_ext)) {
                vz -= zm;
                z = edge_coord(g.z_zero, vz, g.voxel_thickness, Z_INC, zm);
                x = corr(dx_dz, z, ze, xe);
                y = corr(dy_dz, z, ze, ye);
                x_to_go = dist_rem(g.x_zero, x, X_INC, xm, g.voxel_width, vx);
                y_to_go = dist_rem(g.y_zero, y, Y_INC, ym, g.voxel_height, vy);
                z_to_go = g.voxel_thickness;
            } else if (x_ext <= y_ext) {
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
            double y_ext = pct_fabs(dx_dy) * y_to_go;
            if (x_to_go <= y_ext) {
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
        end_walk = (voxel == voxel_out) || (vx >= C) || (vy >= R) || (vz >= S) || vx < 0 || vy < 0 || vz < 0;
        if (!end_walk) visit(voxel);
    }
}

__global__ void __launch_bounds__(256)
k_space_carve(pct_geo g, pct_hist_soa h, const uint8_t* __restrict__ missed, double threshold,
              uint32_t* __restrict__ bits) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h.n || missed[i] || !(h.wepl[i] <= threshold)) return;
    carve_walk(g, h.x_entry[i], h.y_entry[i], h.z_entry[i], h.x_exit[i], h.y_exit[i], h.z_exit[i], [&](int64_t v) {
        if (v >= 0 && v < g.V) atomicAnd(&bits[v >> 5], ~(1u << (v & 31)));
    });
}

/* MSC (SM = false): histories with WEPL <= threshold; SM (SM = true):
 * histories with WEPL >= threshold (.cu:3615) */
template <bool SM>
__global__ void __launch_bounds__(256)
k_walk_count(pct_geo g, pct_hist_soa h, const uint8_t* __restrict__ missed, double threshold,
             int32_t* __restrict__ counts) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h.n || missed[i]) return;
    const float w = h.wepl[i];
    if (SM ? !(w >= threshold) : !(w <= threshold)) return;
    carve_walk(g, h.x_entry[i], h.y_entry[i], h.z_entry[i], h.x_exit[i], h.y_exit[i], h.z_exit[i], [&](int64_t v) {
        if (v >= 0 && v < g.V) atomicAdd(&counts[v], 1);
    });
}

/* The SM walk visits most histories, and the per-voxel atomics of a
 * wavefront cost one cache line each: walked in storage order (angle-major,
 * random lateral position) 64 lanes touch ~64 lines per step. Ordered by
 * (height, beam direction, lateral offset) the lanes of a wave trace adjacent
 * parallel paths and touch a few neighbouring lines. The order only changes
 * the order of integer increments, so the counts are the same. Key: half-slice
 * row of the entry height, 1-degree direction, 0.5 mm lateral offset;
 * histories that do not walk sort last. */
__global__ void __launch_bounds__(256)
k_walk_keys(pct_geo g, pct_hist_soa h, const uint8_t* __restrict__ missed, double threshold, int vbits,
            uint32_t* __restrict__ keys, uint32_t* __restrict__ order) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h.n) return;
    order[i] = (uint32_t)i;
    const float w = h.wepl[i];
    if (missed[i] || !(w >= threshold)) {
        keys[i] = 0xFFFFFFFFu;
        return;
    }
    const float xe = h.x_entry[i], ye = h.y_entry[i];
    const float a = atan2f(h.y_exit[i] - ye, h.x_exit[i] - xe);
    const float t = -sinf(a) * xe + cosf(a) * ye;
    int vq = (int)(fabsf(h.z_entry[i] - (float)g.z_zero) / (float)(0.5 * g.voxel_thickness));
    int aq = (int)((a + 3.14159265f) * (360.0f / 6.2831853f));
    int tq = (int)((t + (float)g.radius) * 20.0f);
    vq = vq < 0 ? 0 : (vq > (1 << vbits) - 1 ? (1 << vbits) - 1 : vq);
    aq = aq < 0 ? 0 : (aq > 359 ? 359 : aq);
    tq = tq < 0 ? 0 : (tq > 4095 ? 4095 : tq);
    keys[i] = ((uint32_t)vq << 21) | ((uint32_t)aq << 12) | (uint32_t)tq;
}

/* the SM walk over the histories in `order`; the first n_walk of them walk */
__global__ void __launch_bounds__(256)
k_walk_count_ordered(pct_geo g, pct_hist_soa h, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ order,
                     int32_t* __restrict__ counts) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= h.n || keys[j] == 0xFFFFFFFFu) return;
    const int64_t i = order[j];
    carve_walk(g, h.x_entry[i], h.y_entry[i], h.z_entry[i], h.x_exit[i], h.y_exit[i], h.z_exit[i], [&](int64_t v) {
        if (v >= 0 && v < g.V) atomicAdd(&counts[v], 1);
    });
}

/* The same walk with the increments of a slab of `wslices` slices kept in
 * LDS: a workgroup takes SMW_CHUNK consecutive walkers of the order above
 * (they enter at the same height), counts the voxels of its window in 16-bit
 * LDS halves (two per word; a chunk has fewer than 65536 histories and adds
 * at most 1 per history to a half, so a half cannot overflow), sends the other
 * voxels to global atomics, and adds the window to the counts with one
 * coalesced pass. Window: the entry slice and its neighbour on the side of
 * the entry's half slice. */
constexpr int SMW_THREADS = 1024;
constexpr int SMW_PER_THREAD = 16;
constexpr int64_t SMW_CHUNK = (int64_t)SMW_THREADS * SMW_PER_THREAD;
static_assert(SMW_CHUNK < 65536, "16-bit window counters");

__global__ void __launch_bounds__(SMW_THREADS)
k_walk_count_window(pct_geo g, pct_hist_soa h, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ order,
                    int wslices, int32_t* __restrict__ counts) {
    extern __shared__ uint32_t s_win[];
    const int64_t base = (int64_t)blockIdx.x * SMW_CHUNK;
    const uint32_t key0 = keys[base];
    if (key0 == 0xFFFFFFFFu) return;                /* no walker in this chunk (uniform) */
    const int64_t plane = (int64_t)g.C * g.R;
    const int64_t wvox = wslices * plane;
    const int64_t wwords = (wvox + 1) / 2;
    const int vq = (int)(key0 >> 21);
    int wb = vq >> 1;
    if (wslices == 2 && (vq & 1) == 0) wb -= 1;
    if (wb > g.S - wslices) wb = g.S - wslices;
    if (wb < 0) wb = 0;
    const int64_t lo = wb * plane, hi = lo + wvox;
    for (int64_t w = threadIdx.x; w < wwords; w += SMW_THREADS) s_win[w] = 0u;
    __syncthreads();
    for (int j = 0; j < SMW_PER_THREAD; ++j) {
        const int64_t q = base + (int64_t)j * SMW_THREADS + threadIdx.x;
        if (q >= h.n || keys[q] == 0xFFFFFFFFu) break;
        const int64_t i = order[q];
        int64_t prev = -1;
        carve_walk(g, h.x_entry[i], h.y_entry[i], h.z_entry[i], h.x_exit[i], h.y_exit[i], h.z_exit[i], [&](int64_t v) {
            if (v < 0 || v >= g.V) return;
            /* a walk's column, row and slice indices are monotone (the slice
             * index stops at 0), so a voxel recurs only as the previous one;
             * recurrences go to global memory, keeping each history's LDS
             * increment per voxel at most 1 */
            const bool again = v == prev;
            prev = v;
            if (!again && v >= lo && v < hi) {
                const int64_t o = v - lo;
                atomicAdd(&s_win[o >> 1], (o & 1) ? 0x10000u : 1u);
            } else {
                atomicAdd(&counts[v], 1);
            }
        });
    }
    __syncthreads();
    for (int64_t w = threadIdx.x; w < wwords; w += SMW_THREADS) {
        const uint32_t c = s_win[w];
        if (!c) continue;
        const int64_t o = 2 * w;
        if (c & 0xFFFFu) atomicAdd(&counts[lo + o], (int)(c & 0xFFFFu));
        if ((c >> 16) && o + 1 < wvox) atomicAdd(&counts[lo + o + 1], (int)(c >> 16));
    }
}

/* initialize_hull_GPU<int> (.cu:3316) for the MSC counts */
__global__ void k_counts_init(pct_geo g, int32_t* __restrict__ counts) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.V) return;
    int column = (int)(v % g.C), row = (int)((v / g.C) % g.R);
    double x = (column - g.C / 2 + 0.5) * g.voxel_width;
    double y = (g.R / 2 - row - 0.5) * g.voxel_height;
    counts[v] = (x * x + y * y < g.radius * g.radius) ? 1 : 0;
}

__global__ void k_msc_edge(pct_geo g, int diff_thresh, const int32_t* __restrict__ counts, int32_t* __restrict__ out) {
    int64_t voxel = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (voxel >= g.V) return;
    const int C = g.C, R = g.R;
    int column = (int)(voxel % C), row = (int)((voxel / C) % R);
    int64_t slice_base = voxel - column - (int64_t)row * C;
    float x = (float)((column - C / 2 + 0.5) * g.voxel_width);
    float y = (float)((R / 2 - row - 0.5) * g.voxel_height);
    int max_difference = 0;
    if (row != 0 && row != R - 1 && column != 0 && column != C - 1) {
        int c0 = counts[voxel];
        for (int cr = row - 1; cr <= row + 1; ++cr)
            for (int cc = column - 1; cc <= column + 1; ++cc) {
                int d = c0 - counts[slice_base + cc + (int64_t)cr * C];
                if (d > max_difference) max_difference = d;
            }
    }
    int v = max_difference > diff_thresh ? 0 : 1;
    double rr = g.radius - (g.voxel_width > g.voxel_height ? g.voxel_width : g.voxel_height) / 2;
    float s2 = x * x + y * y;
    if ((double)s2 >= rr * rr) v = 0;
    out[voxel] = v;
}

/* SM edge detection, step 1: per slice the largest carve_differences value
 * (.cu:3274-3292) and the first pixel holding it, as one 64-bit key
 * (difference << 32 | ~pixel): the largest key is the largest difference at
 * the smallest pixel, which is what the reference's in-order scan with a
 * strict > keeps (.cu:3736-3747). Differences of 0 leave the key at 0. A
 * workgroup covers SM_TILE pixels of one slice (blockIdx.y) and adds one
 * atomic max. */
constexpr int SM_TILE = 1024;
__global__ void __launch_bounds__(256)
k_sm_diff_max(pct_geo g, const int32_t* __restrict__ counts, unsigned long long* __restrict__ keys) {
    const int C = g.C, R = g.R;
    const int64_t plane = (int64_t)C * R;
    const int64_t slice_base = (int64_t)blockIdx.y * plane;
    unsigned long long best = 0;
    for (int k = 0; k < SM_TILE / 256; ++k) {
        int64_t pixel = (int64_t)blockIdx.x * SM_TILE + k * 256 + threadIdx.x;
        if (pixel >= plane) break;
        int column = (int)(pixel % C), row = (int)(pixel / C);
        if (row == 0 || row == R - 1 || column == 0 || column == C - 1) continue;
        const int32_t* c = counts + slice_base;
        int c0 = c[pixel];
        int max_difference = 0;
        for (int cr = row - 1; cr <= row + 1; ++cr)
            for (int cc = column - 1; cc <= column + 1; ++cc) {
                int d = c0 - c[cc + (int64_t)cr * C];
                if (d > max_difference) max_difference = d;
            }
        if (max_difference > 0) {
            unsigned long long key = ((unsigned long long)(uint32_t)max_difference << 32) | (uint32_t)~(uint32_t)pixel;
            best = key > best ? key : best;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(best, off, 64);
        best = o > best ? o : best;
    }
    __shared__ unsigned long long s_best[4];
    if ((threadIdx.x & 63) == 0) s_best[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) best = s_best[w] > best ? s_best[w] : best;
        if (best) atomicMax(&keys[blockIdx.y], best);
    }
}

__global__ void k_sm_keys_clear(unsigned long long* keys, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = 0;
}

/* SM edge detection, step 2: SM_edge_detection_GPU (.cu:3774-3789) with the
 * slice's threshold = the counts at the key's pixel */
__global__ void __launch_bounds__(256)
k_sm_edge(pct_geo g, double scale, const int32_t* __restrict__ counts, const unsigned long long* __restrict__ keys,
          int32_t* __restrict__ out, int32_t* __restrict__ thresholds) {
    int64_t voxel = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (voxel >= g.V) return;
    const int C = g.C, R = g.R;
    const int64_t plane = (int64_t)C * R;
    const int slice = (int)(voxel / plane);
    const int64_t pixel = voxel - slice * plane;
    const int column = (int)(pixel % C), row = (int)(pixel / C);
    const unsigned long long key = keys[slice];
    const int threshold = key ? counts[slice * plane + (uint32_t)~(uint32_t)key] : 0;
    float x = (float)((column - C / 2 + 0.5) * g.voxel_width);
    float y = (float)((R / 2 - row - 0.5) * g.voxel_height);
    int v = counts[voxel] > scale * threshold ? 1 : 0;
    double rr = g.radius - (g.voxel_width > g.voxel_height ? g.voxel_width : g.voxel_height) / 2;
    float s2 = x * x + y * y;
    if ((double)s2 >= rr * rr) v = 0;
    out[voxel] = v;
    if (pixel == 0 && thresholds) thresholds[slice] = threshold;
}

/* hull = (counts != 0) then the averaging filter with threshold; one voxel per
 * lane, 64 consecutive voxels per wave packed into two words */
__global__ void __launch_bounds__(256)
k_hull_select(pct_geo g, int radius, double threshold, const int32_t* __restrict__ counts, uint32_t* __restrict__ bits) {
    int64_t voxel = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool on = false;
    if (voxel < g.V) {
        const int C = g.C, R = g.R;
        int vx = (int)(voxel % C), vy = (int)((voxel / C) % R);
        int64_t slice_base = voxel - vx - (int64_t)vy * C;
        unsigned left = vx - radius > 0 ? vx - radius : 0, right = vx + radius < C - 1 ? vx + radius : C - 1;
        unsigned top = vy - radius > 0 ? vy - radius : 0, bottom = vy + radius < R - 1 ? vy + radius : R - 1;
        int nvox = (int)((right - left + 1) * (bottom - top + 1));
        double sum_threshold = nvox * threshold;
        double sum = 0.0;
        for (int c = (int)left; c <= (int)right; ++c)
            for (int r = (int)top; r <= (int)bottom; ++r) sum += (counts[slice_base + c + (int64_t)r * C] != 0) ? 1.0 : 0.0;
        on = sum > sum_threshold;
    }
    uint64_t m = __ballot(on);
    int lane = threadIdx.x & 63;
    int64_t w0 = voxel >> 5;       /* lane 0 and lane 32 write their word */
    if ((lane & 31) == 0 && voxel < g.V) bits[w0] = (uint32_t)(lane == 0 ? m : (m >> 32));
}

}  // namespace

static int check_carve(pct_ctx* ctx, const pct_hist_soa* h, const uint8_t* missed, const void* out, const char* who) {
    if (!ctx || !h || !out) return pct_fail(ctx, PCT_E_INVALID, "%s: null argument", who);
    if (h->n < 0) return pct_fail(ctx, PCT_E_INVALID, "%s: negative history count", who);
    if (h->n == 0) return PCT_OK;                   /* no history: the arrays are not read */
    if (!missed) return pct_fail(ctx, PCT_E_INVALID, "%s: null missed", who);
    if (!h->x_entry || !h->x_exit || !h->wepl) return pct_fail(ctx, PCT_E_INVALID, "%s: missing history field", who);
    return PCT_OK;
}

extern "C" int pct_space_carve(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed, uint32_t* hull_bits) {
    int rc = check_carve(ctx, hist, missed, hull_bits, "pct_space_carve");
    if (rc) return rc;
    if (hist->n > 0)
        hipLaunchKernelGGL(k_space_carve, dim3(pct_grid(hist->n, 256)), dim3(256), 0, ctx->stream, ctx->g, *hist, missed,
                           ctx->p.sc_threshold, hull_bits);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_msc_init(pct_ctx* ctx, int32_t* counts) {
    if (!ctx || !counts) return pct_fail(ctx, PCT_E_INVALID, "pct_msc_init: null argument");
    hipLaunchKernelGGL(k_counts_init, dim3(pct_grid(ctx->g.V, 256)), dim3(256), 0, ctx->stream, ctx->g, counts);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_msc_count(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed, int32_t* counts) {
    int rc = check_carve(ctx, hist, missed, counts, "pct_msc_count");
    if (rc) return rc;
    if (hist->n > 0)
        hipLaunchKernelGGL(k_walk_count<false>, dim3(pct_grid(hist->n, 256)), dim3(256), 0, ctx->stream, ctx->g, *hist, missed,
                           ctx->p.msc_threshold, counts);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_msc_edge(pct_ctx* ctx, const int32_t* counts, int32_t* out) {
    if (!ctx || !counts || !out || counts == out) return pct_fail(ctx, PCT_E_INVALID, "pct_msc_edge: bad arguments");
    hipLaunchKernelGGL(k_msc_edge, dim3(pct_grid(ctx->g.V, 256)), dim3(256), 0, ctx->stream, ctx->g,
                       ctx->p.msc_diff_thresh, counts, out);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_sm_count(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed, int32_t* counts) {
    int rc = check_carve(ctx, hist, missed, counts, "pct_sm_count");
    if (rc) return rc;
    const int64_t n = hist->n;
    if (n == 0) return PCT_OK;
    if (n >= (int64_t)UINT32_MAX) {                 /* too many to order with 32-bit indices: storage order */
        hipLaunchKernelGGL(k_walk_count<true>, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, ctx->g, *hist, missed,
                           ctx->p.sm_lower_threshold, counts);
        PCT_HIP_CHECK(ctx, hipGetLastError());
        return PCT_OK;
    }
    int vbits = 1;
    while ((1 << vbits) < 2 * ctx->g.S && vbits < 11) ++vbits;
    pct_buffer* w = ctx->walk_order;
    const size_t nn = (size_t)n;
    if ((rc = pct_ensure(ctx, w[0], 4 * nn)) || (rc = pct_ensure(ctx, w[1], 4 * nn)) ||
        (rc = pct_ensure(ctx, w[2], 4 * nn)) || (rc = pct_ensure(ctx, w[3], 4 * nn)))
        return rc;
    uint32_t *k_in = (uint32_t*)w[0].ptr, *k_out = (uint32_t*)w[1].ptr;
    uint32_t *o_in = (uint32_t*)w[2].ptr, *o_out = (uint32_t*)w[3].ptr;
    hipLaunchKernelGGL(k_walk_keys, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, ctx->g, *hist, missed,
                       ctx->p.sm_lower_threshold, vbits, k_in, o_in);
    size_t tb = 0;
    PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(nullptr, tb, k_in, k_out, o_in, o_out, nn, 0, 32, ctx->stream));
    if ((rc = pct_ensure(ctx, w[4], tb))) return rc;
    tb = w[4].bytes;
    PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(w[4].ptr, tb, k_in, k_out, o_in, o_out, nn, 0, 32, ctx->stream));
    /* the largest slab window (2 or 1 slices of 16-bit counters) the
     * workgroup's LDS holds, decided once per context */
    if (ctx->sm_window < 0) {
        int lim = 0, optin = 0;
        (void)hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
        (void)hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, ctx->device);
        const int64_t avail = optin > lim ? optin : lim;
        const int64_t plane = (int64_t)ctx->g.C * ctx->g.R;
        ctx->sm_window = 0;
        for (int ws = 2; ws >= 1 && ctx->sm_window == 0; --ws) {
            const int64_t bytes = 4 * ((ws * plane + 1) / 2);
            if (ws > ctx->g.S || bytes > avail) continue;
            if (bytes > lim && hipFuncSetAttribute((const void*)k_walk_count_window,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) {
                (void)hipGetLastError();
                continue;
            }
            ctx->sm_window = ws;
        }
    }
    if (ctx->sm_window > 0) {
        const int ws = ctx->sm_window;
        const size_t bytes = 4 * (size_t)((ws * (int64_t)ctx->g.C * ctx->g.R + 1) / 2);
        hipLaunchKernelGGL(k_walk_count_window, dim3((unsigned)((n + SMW_CHUNK - 1) / SMW_CHUNK)), dim3(SMW_THREADS), bytes,
                           ctx->stream, ctx->g, *hist, k_out, o_out, ws, counts);
    } else {
        hipLaunchKernelGGL(k_walk_count_ordered, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, ctx->g, *hist, k_out,
                           o_out, counts);
    }
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_sm_edge(pct_ctx* ctx, const int32_t* counts, int32_t* out, int32_t* thresholds) {
    if (!ctx || !counts || !out || counts == out) return pct_fail(ctx, PCT_E_INVALID, "pct_sm_edge: bad arguments");
    const int64_t plane = (int64_t)ctx->g.C * ctx->g.R;
    if (plane > INT32_MAX) return pct_fail(ctx, PCT_E_INVALID, "pct_sm_edge: slice of more than 2^31 pixels");
    const int S = ctx->g.S;
    int rc = pct_ensure(ctx, ctx->hull_tmp, sizeof(unsigned long long) * S);
    if (rc) return rc;
    unsigned long long* keys = (unsigned long long*)ctx->hull_tmp.ptr;
    hipLaunchKernelGGL(k_sm_keys_clear, dim3((S + 255) / 256), dim3(256), 0, ctx->stream, keys, S);
    hipLaunchKernelGGL(k_sm_diff_max, dim3((unsigned)((plane + SM_TILE - 1) / SM_TILE), S), dim3(256), 0, ctx->stream,
                       ctx->g, counts, keys);
    hipLaunchKernelGGL(k_sm_edge, dim3(pct_grid(ctx->g.V, 256)), dim3(256), 0, ctx->stream, ctx->g,
                       ctx->p.sm_scale_threshold, counts, keys, out, thresholds);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_hull_select(pct_ctx* ctx, const int32_t* counts, uint32_t* hull_bits) {
    if (!ctx || !counts || !hull_bits) return pct_fail(ctx, PCT_E_INVALID, "pct_hull_select: null argument");
    if (ctx->p.hull_filter_radius < 0) return pct_fail(ctx, PCT_E_INVALID, "pct_hull_select: negative radius");
    hipLaunchKernelGGL(k_hull_select, dim3(pct_grid(ctx->g.V, 256)), dim3(256), 0, ctx->stream, ctx->g,
                       ctx->p.hull_filter_radius, ctx->p.hull_filter_threshold, counts, hull_bits);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}
