/*
 * pct_noalloc.hip -- an allocation-free form of pct_drop_block_update.
 *
 * pct_drop_block_update (pct_drop.hip) grows context workspaces for the MLP
 * lists of its histories (capacity offsets, the lists, their lengths and chord
 * lengths) with hipMalloc on demand. Callers that cannot allocate -- code
 * under HIP graph capture, tight serving loops, harnesses that forbid it --
 * call pct_drop_block_update_noalloc instead: the same MLP walk and the same
 * scatter kernels, with those lists in `scratch`, a file-scope device array of
 * fixed size (PCT_NOALLOC_MAX_HISTORIES histories of at most
 * PCT_NOALLOC_MAX_SLOTS path voxels each, about 17 MB, held by every GPU the
 * library runs on until the process ends). Each history's list gets a fixed
 * slot range of the context's max_intersections entries (a longer path is an
 * overflow error in both entry points), so no prefix sum and no sort storage
 * is needed either. The per-depth MLP tables and the pinned status word the
 * call reads back are made by pct_create. The surrogate form's fallback list
 * (the histories its guard sends to the exact walk) is in the same array.
 *
 * The array's address on a device comes from hipGetSymbolAddress, called once
 * per device with that device current (pct_create, through
 * pct_noalloc_init); host code uses only that address, never the symbol.
 * Every call on one GPU uses the same array: calls that overlap on one GPU
 * (several streams or threads) race on it.
 */
#include <mutex>
#include "pct_internal.h"

int pct_mlp_paths_fb(pct_ctx* ctx, hipStream_t stream, const pct_hist_soa* hist, int64_t start, int64_t count,
                     const int64_t* offsets, uint32_t* voxels, uint32_t* lengths, double* chord, uint32_t* hash_out,
                     const uint32_t* hash_expect, int mode, bool beside, uint32_t* fb_idx,
                     unsigned long long* fb_count);
int pct_drop_launch_scatter(pct_ctx* ctx, const uint32_t* voxels, const int64_t* offsets, const uint32_t* lengths,
                            const double* chord, const float* wepl, int64_t count, const float* x, float* x_update,
                            uint32_t* counts, float lambda);
int pct_status_to_rc(pct_ctx* ctx, int flags);

/* layout of the scratch array (byte offsets, each 256-aligned) */
#define NA_OFFSETS 0                                                     /* int64 [MAX_H + 1] */
#define NA_CHORD (((8 * (PCT_NOALLOC_MAX_HISTORIES + 1)) + 255) & ~255)   /* f64 [MAX_H]       */
#define NA_LENGTHS (NA_CHORD + 8 * PCT_NOALLOC_MAX_HISTORIES)            /* u32 [MAX_H]       */
#define NA_FB_COUNT (NA_LENGTHS + 4 * PCT_NOALLOC_MAX_HISTORIES)         /* u64: surrogate fallbacks */
#define NA_FB_IDX (NA_FB_COUNT + 256)                                    /* u32 [MAX_H]       */
#define NA_VOXELS (NA_FB_IDX + 4 * PCT_NOALLOC_MAX_HISTORIES)            /* u32 [MAX_H x MAX_SLOTS] */
#define NA_BYTES (NA_VOXELS + 4 * PCT_NOALLOC_MAX_HISTORIES * PCT_NOALLOC_MAX_SLOTS)

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {

/* offsets[k] = k * stride for k in [0, count] */
__global__ void k_noalloc_offsets(int64_t* __restrict__ offsets, int64_t count, int64_t stride) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k <= count) offsets[k] = k * stride;
}

#define NA_MAX_DEVICES 64
std::mutex na_mutex;
void* na_ptr[NA_MAX_DEVICES] = {};
hipError_t na_err[NA_MAX_DEVICES] = {};

}  // namespace

/* once per device, with `device` current: the device address of the scratch
 * array on it */
int pct_noalloc_init(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) return PCT_E_INVALID;
    std::lock_guard<std::mutex> lock(na_mutex);
    if (na_ptr[device] || na_err[device] != hipSuccess) return na_ptr[device] ? PCT_OK : PCT_E_HIP;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        na_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        return PCT_E_HIP;
    }
    na_ptr[device] = p;
    return PCT_OK;
}

extern "C" int pct_drop_block_update_noalloc(pct_ctx* ctx, const pct_hist_soa* hist, int64_t start, int64_t count,
                                             const float* x, float* x_update, uint32_t* counts, float lambda,
                                             int mode) {
    if (!ctx || !hist) return pct_fail(ctx, PCT_E_INVALID, "pct_drop_block_update_noalloc: null argument");
    if (!hist->wepl || !hist->first_voxel || !hist->x_entry)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_block_update_noalloc: missing history field");
    if (!pct_mlp_mode_ok(ctx, mode))
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_block_update_noalloc: mode %d not available", mode);
    if (!x || !x_update || !counts)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_block_update_noalloc: null image array");
    if (start < 0 || count < 0 || start + count > hist->n)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_block_update_noalloc: bad range");
    if (count > PCT_NOALLOC_MAX_HISTORIES)
        return pct_fail(ctx, PCT_E_INVALID, "pct_drop_block_update_noalloc: %lld histories, at most %d per call",
                        (long long)count, PCT_NOALLOC_MAX_HISTORIES);
    const int64_t stride = ctx->p.max_intersections;
    if (stride > PCT_NOALLOC_MAX_SLOTS)
        return pct_fail(ctx, PCT_E_INVALID,
                        "pct_drop_block_update_noalloc: max_intersections %lld above the %d path voxels per history "
                        "it holds",
                        (long long)stride, PCT_NOALLOC_MAX_SLOTS);
    if (count == 0) return PCT_OK;
    if (ctx->device < 0 || ctx->device >= NA_MAX_DEVICES || !na_ptr[ctx->device])
        return pct_fail(ctx, PCT_E_HIP, "pct_drop_block_update_noalloc: no scratch address on device %d (%s)",
                        ctx->device,
                        ctx->device >= 0 && ctx->device < NA_MAX_DEVICES ? hipGetErrorString(na_err[ctx->device])
                                                                          : "device index");
    unsigned char* base = (unsigned char*)na_ptr[ctx->device];
    int64_t* offsets = (int64_t*)(base + NA_OFFSETS);
    double* chord = (double*)(base + NA_CHORD);
    uint32_t* lengths = (uint32_t*)(base + NA_LENGTHS);
    uint32_t* voxels = (uint32_t*)(base + NA_VOXELS);
    unsigned long long* fb_count = (unsigned long long*)(base + NA_FB_COUNT);
    uint32_t* fb_idx = (uint32_t*)(base + NA_FB_IDX);
    hipLaunchKernelGGL(k_noalloc_offsets, dim3(pct_grid(count + 1, 256)), dim3(256), 0, ctx->stream, offsets, count,
                       stride);
    PCT_HIP_CHECK(ctx, hipGetLastError());
    int rc = pct_mlp_paths_fb(ctx, ctx->stream, hist, start, count, offsets, voxels, lengths, chord, nullptr, nullptr,
                              mode, false, fb_idx, fb_count);
    if (rc) return rc;
    rc = pct_drop_launch_scatter(ctx, voxels, offsets, lengths, chord, hist->wepl + start, count, x, x_update, counts,
                                 lambda);
    if (rc) return rc;
    int flags = 0;
    if ((rc = pct_take_status(ctx, &flags))) return rc;
    return pct_status_to_rc(ctx, flags);
}
