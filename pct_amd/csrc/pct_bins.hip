/*
 * pct_bins.hip -- (a) binning, per-bin statistics, 3-sigma cuts, sinogram.
 *
 * Replaces binning_GPU (.cu:2507-2578), calculate_means_GPU (.cu:2618),
 * sum_squared_deviations_GPU (.cu:2671), calculate_standard_deviations_GPU
 * (.cu:2718), statistical_cuts_GPU (.cu:2835) and construct_sinogram_GPU
 * (.cu:2912). The reference forms every per-bin sum with float atomics, so
 * its statistics -- and therefore which histories the cuts keep -- depend on
 * thread scheduling. Here the sums are formed in storage order: the binning
 * kernel writes each history's bin key and a 16-byte record of what the
 * statistics read (WEPL, the two exit-minus-entry angle differences as the
 * float subtractions give them, the in-range flag), the records are stably
 * sorted by bin once (rocPRIM radix sort), and one thread per bin walks its
 * bin's contiguous records in storage order, accumulating in float as the
 * reference's arrays do. The cut flags are then evaluated per history in
 * storage order against its bin's statistics (coalesced, no scatter). The
 * result is deterministic and bit-identical to a sequential pass; the
 * decision paths (means, deviations, cut masks) are exact. The records are a
 * snapshot taken by pct_binning: the statistics and cuts that follow use the
 * WEPL and angles the histories had then.
 *
 * Reference quirks kept: an out-of-range history keeps bin 0 and is kept
 * (it enters bin 0's squared deviations and cuts, not its mean); a bin with
 * at most one member keeps its raw sum of squares as its deviation; angle cuts
 * are strict, the WEPL cut inclusive.
 */
#include <rocprim/device/device_radix_sort.hpp>
#include "pct_internal.h"

namespace {

struct bin_geo {
    int T, V, A;
    double t_size, v_size, a_size;
    int64_t bins;
};

__device__ __forceinline__ double wrap_pi(double a) {
    if (a > PCT_PI) a -= 2 * PCT_PI;
    if (a < -PCT_PI) a += 2 * PCT_PI;
    return a;
}

/* what the statistics and the cuts read of a history, sorted by bin */
struct __align__(16) bin_rec {
    float wepl;
    float d_ut;          /* xy_exit_angle - xy_entry_angle (float subtraction) */
    float d_uv;          /* xz_exit_angle - xz_entry_angle                     */
    uint32_t in_range;
};

/* bin number of every history; sort key = bin for kept (not missed)
 * histories, `bins` (sorts last) for missed ones */
__global__ void __launch_bounds__(256)
k_binning(bin_geo b, pct_hist_soa h, const uint8_t* __restrict__ missed, uint8_t* __restrict__ in_range,
          uint32_t* __restrict__ keys, bin_rec* __restrict__ recs) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h.n) return;
    const double RADIANS_TO_ANGLE = 180.0 / PCT_PI, ANGLE_TO_RADIANS = PCT_PI / 180.0;
    float xe = h.x_entry[i], ye = h.y_entry[i], ze = h.z_entry[i];
    float xx = h.x_exit[i], yx = h.y_exit[i], zx = h.z_exit[i];
    double x_mid = (xe + xx) / 2;
    double y_mid = (ye + yx) / 2;
    double z_mid = (ze + zx) / 2;
    double path_angle = pct_atan2f(yx - ye, xx - xe);
    if (path_angle < 0) path_angle += 2 * PCT_PI;
    int angle_bin = (int)((path_angle * RADIANS_TO_ANGLE / b.a_size) + 0.5) % b.A;
    double angle = angle_bin * b.a_size * ANGLE_TO_RADIANS;
    double s, c;
    pct_sincos(angle, &s, &c);
    double t = y_mid * c - x_mid * s;
    int t_bin = (int)((t / b.t_size) + b.T / 2);
    int v_bin = (int)((z_mid / b.v_size) + b.V / 2);
    bool ok = (t_bin >= 0) && (v_bin >= 0) && (t_bin < b.T) && (v_bin < b.V);
    int32_t bin = ok ? t_bin + angle_bin * b.T + v_bin * b.T * b.A : 0;
    h.bin_num[i] = bin;
    in_range[i] = ok;
    keys[i] = missed[i] ? (uint32_t)b.bins : (uint32_t)bin;
    bin_rec r;
    r.wepl = h.wepl[i];
    r.d_ut = h.xy_exit_angle[i] - h.xy_entry_angle[i];
    r.d_uv = h.xz_exit_angle[i] - h.xz_entry_angle[i];
    r.in_range = ok;
    recs[i] = r;
}

/* [first, last) member range of every bin in the sorted order */
__global__ void k_bin_ranges(const uint32_t* __restrict__ keys, int64_t n, int64_t bins, uint32_t* __restrict__ first,
                             uint32_t* __restrict__ last) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    if (k >= bins) return;
    if (i == 0 || keys[i - 1] != k) first[k] = (uint32_t)i;
    if (i == n - 1 || keys[i + 1] != k) last[k] = (uint32_t)(i + 1);
}

__global__ void k_fill_u32(uint32_t* __restrict__ a, int64_t n, uint32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

/* f(record) for the records [a, e) in order. A thread walks its own bin, so
 * the records are read in batches of 8 independent loads before any is used:
 * one memory round trip per 8 records instead of one per record. */
template <class F>
__device__ __forceinline__ void for_each_rec(const bin_rec* __restrict__ recs, uint32_t a, uint32_t e, F f) {
    uint32_t k = a;
    for (; k + 8 <= e; k += 8) {
        bin_rec r[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) r[u] = recs[k + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) f(r[u]);
    }
    for (; k < e; ++k) f(recs[k]);
}

/* A2 sums + A3 means + A4 squared deviations + A5 deviations, one thread per bin */
__global__ void __launch_bounds__(256)
k_bin_stats(int64_t bins, const uint32_t* __restrict__ first, const uint32_t* __restrict__ last,
            const bin_rec* __restrict__ recs, pct_bin_stats st) {
    int64_t bin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (bin >= bins) return;
    uint32_t a = first[bin], e = last[bin];
    int32_t count = 0;
    float sw = 0.0f, su = 0.0f, sv = 0.0f;
    for_each_rec(recs, a, e, [&](const bin_rec& r) {
        if (!r.in_range) return;
        double rel_ut = wrap_pi(r.d_ut);
        double rel_uv = wrap_pi(r.d_uv);
        count += 1;
        sw += r.wepl;
        su += (float)rel_ut;
        sv += (float)rel_uv;
    });
    if (count > 0) {
        sw /= count;
        su /= count;
        sv /= count;
    }
    float qw = 0.0f, qu = 0.0f, qv = 0.0f;
    for_each_rec(recs, a, e, [&](const bin_rec& r) {
        double rel_ut = wrap_pi(r.d_ut);
        double rel_uv = wrap_pi(r.d_uv);
        double dw = r.wepl - sw;
        double du = rel_ut - su;
        double dv = rel_uv - sv;
        qw += (float)(dw * dw);
        qu += (float)(du * du);
        qv += (float)(dv * dv);
    });
    if (count > 1) {
        qw = pct_sqrtf(qw / (count - 1));
        qu = pct_sqrtf(qu / (count - 1));
        qv = pct_sqrtf(qv / (count - 1));
    }
    st.counts[bin] = count;
    st.mean_wepl[bin] = sw; st.mean_rel_ut[bin] = su; st.mean_rel_uv[bin] = sv;
    st.stddev_wepl[bin] = qw; st.stddev_rel_ut[bin] = qu; st.stddev_rel_uv[bin] = qv;
}

/* A6: does a history fail its bin's 3-sigma cuts (strict angle cuts,
 * inclusive WEPL cut)? */
__device__ __forceinline__ bool cut_fails(float w, float d_ut, float d_uv, float mw, float mu, float mv, float cw,
                                          float cu, float cv) {
    double rel_ut = wrap_pi(d_ut);
    double rel_uv = wrap_pi(d_uv);
    bool pu = pct_fabs(rel_ut - mu) < (double)cu;
    bool pv = pct_fabs(rel_uv - mv) < (double)cv;
    bool pw = pct_fabsf(mw - w) <= cw;
    return !pu || !pv || !pw;
}

/* a bin's cut window: means and sigmas_to_keep x deviations, one 32-byte
 * record per bin so that a history's check gathers one cache line */
struct __align__(32) bin_cut {
    float mw, mu, mv, cw, cu, cv, pad0, pad1;
};

__global__ void __launch_bounds__(256)
k_pack_cuts(int64_t bins, pct_bin_stats st, float sig, bin_cut* __restrict__ cut) {
    int64_t bin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (bin >= bins) return;
    bin_cut c;
    c.mw = st.mean_wepl[bin]; c.mu = st.mean_rel_ut[bin]; c.mv = st.mean_rel_uv[bin];
    c.cw = sig * st.stddev_wepl[bin]; c.cu = sig * st.stddev_rel_ut[bin]; c.cv = sig * st.stddev_rel_uv[bin];
    c.pad0 = c.pad1 = 0.0f;
    cut[bin] = c;
}

/* A6 per history in storage order: failed = missed, or outside its bin's
 * cuts (an out-of-range history is judged in bin 0, as it is keyed) */
__global__ void __launch_bounds__(256)
k_cut_flags(pct_hist_soa h, const uint8_t* __restrict__ missed, const bin_cut* __restrict__ cut,
            uint8_t* __restrict__ failed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h.n) return;
    bool f = true;
    if (!missed[i]) {
        bin_cut c = cut[h.bin_num[i]];
        f = cut_fails(h.wepl[i], h.xy_exit_angle[i] - h.xy_entry_angle[i], h.xz_exit_angle[i] - h.xz_entry_angle[i],
                      c.mw, c.mu, c.mv, c.cw, c.cu, c.cv);
    }
    failed[i] = f;
}

/* A7 + the survivors' count, one thread per bin over its sorted records */
__global__ void __launch_bounds__(256)
k_bin_sinogram(int64_t bins, const uint32_t* __restrict__ first, const uint32_t* __restrict__ last,
               const bin_rec* __restrict__ recs, pct_bin_stats st, float sig, float* __restrict__ sinogram) {
    int64_t bin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (bin >= bins) return;
    uint32_t a = first[bin], e = last[bin];
    float mw = st.mean_wepl[bin], mu = st.mean_rel_ut[bin], mv = st.mean_rel_uv[bin];
    float cw = sig * st.stddev_wepl[bin], cu = sig * st.stddev_rel_ut[bin], cv = sig * st.stddev_rel_uv[bin];
    int32_t count = 0;
    float sum = 0.0f;
    for_each_rec(recs, a, e, [&](const bin_rec& r) {
        if (!cut_fails(r.wepl, r.d_ut, r.d_uv, mw, mu, mv, cw, cu, cv)) {
            count += 1;
            sum += r.wepl;
        }
    });
    if (count > 0) sum /= count;
    st.counts[bin] = count;
    sinogram[bin] = sum;
}

}  // namespace

/* sorted order of the kept histories by bin, cached per history array */
struct pct_bin_order {
    pct_hist_soa key{};
    int64_t n = 0;
    pct_buffer keys_in, keys_out, recs_in, recs_out, first, last, tmp, cut;
};

static bin_geo make_bin_geo(const pct_params& p) {
    bin_geo b;
    b.T = p.t_bins; b.V = p.v_bins; b.A = p.angular_bins;
    b.t_size = p.t_bin_size; b.v_size = p.v_bin_size; b.a_size = p.angular_bin_size;
    b.bins = (int64_t)b.T * b.V * b.A;
    return b;
}

static int bins_bits(int64_t bins) {
    int b = 1;
    while (((int64_t)1 << b) <= bins) ++b;
    return b;
}

/* per-history arrays may be null when there is no history */
static int check_bin_args(pct_ctx* ctx, const pct_hist_soa* h, const uint8_t* missed, const char* who) {
    if (!ctx || !h) return pct_fail(ctx, PCT_E_INVALID, "%s: null argument", who);
    if (h->n < 0 || h->n >= ((int64_t)1 << 32)) return pct_fail(ctx, PCT_E_INVALID, "%s: bad history count", who);
    if (h->n == 0) return PCT_OK;
    if (!missed) return pct_fail(ctx, PCT_E_INVALID, "%s: null missed", who);
    if (!h->bin_num || !h->wepl || !h->x_entry || !h->xy_entry_angle || !h->xz_entry_angle)
        return pct_fail(ctx, PCT_E_INVALID, "%s: missing history field", who);
    return PCT_OK;
}

static pct_bin_order* order_of(pct_ctx* ctx) {
    if (!ctx->bin_order) ctx->bin_order = new pct_bin_order();
    return ctx->bin_order;
}

void pct_bin_order_destroy(pct_bin_order* o) {
    if (!o) return;
    for (pct_buffer* b : {&o->keys_in, &o->keys_out, &o->recs_in, &o->recs_out, &o->first, &o->last, &o->tmp, &o->cut})
        if (b->ptr) (void)hipFree(b->ptr);
    delete o;
}

/* binning + stable sort of the statistics records by bin */
static int bin_and_sort(pct_ctx* ctx, const pct_hist_soa* h, const uint8_t* missed, uint8_t* in_range) {
    bin_geo b = make_bin_geo(ctx->p);
    pct_bin_order* o = order_of(ctx);
    const int64_t n = h->n;
    int rc;
    size_t nn = (size_t)(n > 0 ? n : 1);
    if ((rc = pct_ensure(ctx, o->keys_in, 4 * nn)) || (rc = pct_ensure(ctx, o->keys_out, 4 * nn)) ||
        (rc = pct_ensure(ctx, o->recs_in, sizeof(bin_rec) * nn)) || (rc = pct_ensure(ctx, o->recs_out, sizeof(bin_rec) * nn)) ||
        (rc = pct_ensure(ctx, o->first, 4 * (size_t)b.bins)) || (rc = pct_ensure(ctx, o->last, 4 * (size_t)b.bins)))
        return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_binning, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, b, *h, missed, in_range,
                           (uint32_t*)o->keys_in.ptr, (bin_rec*)o->recs_in.ptr);
    size_t tb = 0;
    int endbit = bins_bits(b.bins);
    PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(nullptr, tb, (uint32_t*)nullptr, (uint32_t*)nullptr, (bin_rec*)nullptr,
                                                 (bin_rec*)nullptr, nn, 0, endbit, ctx->stream));
    if ((rc = pct_ensure(ctx, o->tmp, tb))) return rc;
    tb = o->tmp.bytes;
    if (n > 0)
        PCT_HIP_CHECK(ctx, rocprim::radix_sort_pairs(o->tmp.ptr, tb, (uint32_t*)o->keys_in.ptr, (uint32_t*)o->keys_out.ptr,
                                                     (bin_rec*)o->recs_in.ptr, (bin_rec*)o->recs_out.ptr, (size_t)n, 0,
                                                     endbit, ctx->stream));
    hipLaunchKernelGGL(k_fill_u32, dim3(pct_grid(b.bins, 256)), dim3(256), 0, ctx->stream, (uint32_t*)o->first.ptr,
                       b.bins, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(pct_grid(b.bins, 256)), dim3(256), 0, ctx->stream, (uint32_t*)o->last.ptr,
                       b.bins, 0u);
    if (n > 0)
        hipLaunchKernelGGL(k_bin_ranges, dim3(pct_grid(n, 256)), dim3(256), 0, ctx->stream, (const uint32_t*)o->keys_out.ptr,
                           n, b.bins, (uint32_t*)o->first.ptr, (uint32_t*)o->last.ptr);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    o->key = *h;
    o->n = n;
    return PCT_OK;
}

extern "C" int pct_binning(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed, uint8_t* in_range) {
    int rc = check_bin_args(ctx, hist, missed, "pct_binning");
    if (rc) return rc;
    if (!in_range && hist->n > 0) return pct_fail(ctx, PCT_E_INVALID, "pct_binning: null in_range");
    return bin_and_sort(ctx, hist, missed, in_range);
}

static bool same_arrays(const pct_hist_soa& a, const pct_hist_soa& b) {
    return a.n == b.n && a.bin_num == b.bin_num && a.wepl == b.wepl && a.x_entry == b.x_entry;
}

extern "C" int pct_bin_statistics(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed, const uint8_t* in_range,
                                  const pct_bin_stats* stats) {
    int rc = check_bin_args(ctx, hist, missed, "pct_bin_statistics");
    if (rc) return rc;
    if ((!in_range && hist->n > 0) || !stats || !stats->counts || !stats->mean_wepl || !stats->mean_rel_ut || !stats->mean_rel_uv ||
        !stats->stddev_wepl || !stats->stddev_rel_ut || !stats->stddev_rel_uv)
        return pct_fail(ctx, PCT_E_INVALID, "pct_bin_statistics: null output");
    pct_bin_order* o = order_of(ctx);
    if (!same_arrays(o->key, *hist))
        return pct_fail(ctx, PCT_E_INVALID, "pct_bin_statistics: call pct_binning on these histories first");
    bin_geo b = make_bin_geo(ctx->p);
    hipLaunchKernelGGL(k_bin_stats, dim3(pct_grid(b.bins, 256)), dim3(256), 0, ctx->stream, b.bins,
                       (const uint32_t*)o->first.ptr, (const uint32_t*)o->last.ptr, (const bin_rec*)o->recs_out.ptr,
                       *stats);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}

extern "C" int pct_statistical_cuts(pct_ctx* ctx, const pct_hist_soa* hist, const uint8_t* missed,
                                    const pct_bin_stats* stats, uint8_t* failed, float* sinogram) {
    int rc = check_bin_args(ctx, hist, missed, "pct_statistical_cuts");
    if (rc) return rc;
    if (!stats || (!failed && hist->n > 0) || !sinogram)
        return pct_fail(ctx, PCT_E_INVALID, "pct_statistical_cuts: null argument");
    pct_bin_order* o = order_of(ctx);
    if (!same_arrays(o->key, *hist))
        return pct_fail(ctx, PCT_E_INVALID, "pct_statistical_cuts: call pct_binning on these histories first");
    bin_geo b = make_bin_geo(ctx->p);
    const float sig = (float)ctx->p.sigmas_to_keep;
    if (hist->n > 0) {
        if ((rc = pct_ensure(ctx, o->cut, sizeof(bin_cut) * (size_t)b.bins))) return rc;
        hipLaunchKernelGGL(k_pack_cuts, dim3(pct_grid(b.bins, 256)), dim3(256), 0, ctx->stream, b.bins, *stats, sig,
                           (bin_cut*)o->cut.ptr);
        hipLaunchKernelGGL(k_cut_flags, dim3(pct_grid(hist->n, 256)), dim3(256), 0, ctx->stream, *hist, missed,
                           (const bin_cut*)o->cut.ptr, failed);
    }
    /* after the flags: this kernel overwrites the counts with the survivors' */
    hipLaunchKernelGGL(k_bin_sinogram, dim3(pct_grid(b.bins, 256)), dim3(256), 0, ctx->stream, b.bins,
                       (const uint32_t*)o->first.ptr, (const uint32_t*)o->last.ptr, (const bin_rec*)o->recs_out.ptr,
                       *stats, sig, sinogram);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    return PCT_OK;
}
