/*
 * pct_batch.hip -- (e) the batched form: DROP iterations of several scans at
 * once.
 *
 * The block loop of one scan is a chain of dependent launches whose cost is
 * launch latency and dependent memory rounds, not bandwidth: one launch of a
 * block keeps only a sliver of the GPU busy. Scans are independent, so block
 * step j launches the deltas of block j of every scan that has one in ONE
 * kernel and their voxel updates in ONE kernel. The per-scan descriptors (list
 * base, stride, histories, image, deltas) travel in the kernel arguments;
 * each workgroup finds its scan from a prefix of workgroup counts. Steps are
 * captured into hipGraphs in chunks; each scan's MLP lists of a chunk of steps
 * are recomputed on a second stream ahead of the chunk (as the single-scan
 * iteration does). The arithmetic is the single-scan kernels'
 * (pct_plan_kernels.h), so every scan's image is identical to its own
 * pct_drop_iteration.
 */
#include <algorithm>
#include <vector>
#include "pct_internal.h"
#include "pct_plan.h"
#include "pct_plan_kernels.h"

int pct_mlp_paths_on(pct_ctx* ctx, hipStream_t stream, const pct_hist_soa* hist, int64_t start, int64_t count,
                     const int64_t* offsets, uint32_t* voxels, uint32_t* lengths, double* chord, uint32_t* hash_out,
                     const uint32_t* hash_expect, int mode, bool beside);
int pct_status_to_rc(pct_ctx* ctx, int flags);
int pct_mlp_fallback_reserve(pct_ctx* ctx, int64_t n);

namespace {

struct batch_delta_args {
    int n;
    uint32_t wg0[PCT_BATCH_MAX + 1];
    const uint32_t* list[PCT_BATCH_MAX];
    int stride[PCT_BATCH_MAX];
    pct_blk::blk_pos pos[PCT_BATCH_MAX];
    const float* wepl[PCT_BATCH_MAX];
    int64_t h0[PCT_BATCH_MAX];
    int64_t count[PCT_BATCH_MAX];
    const float* x[PCT_BATCH_MAX];
    double lambda[PCT_BATCH_MAX];
    void* delta[PCT_BATCH_MAX];
};

struct batch_apply_args {
    int n;
    uint32_t wg0[PCT_BATCH_MAX + 1];
    const uint64_t* segs[PCT_BATCH_MAX];
    int64_t ns[PCT_BATCH_MAX];
    const uint16_t* csc[PCT_BATCH_MAX];
    const void* delta[PCT_BATCH_MAX];
    float* x[PCT_BATCH_MAX];
};

__device__ __forceinline__ int batch_scan_of(const uint32_t* wg0, int n, uint32_t wg) {
    int s = 0;
    while (s + 1 < n && wg >= wg0[s + 1]) ++s;
    return s;
}

template <bool F32>
__global__ void __launch_bounds__(64 * DELTA_WAVES)
k_batch_delta(batch_delta_args a) {
    __builtin_amdgcn_s_setprio(3);
    const int s = batch_scan_of(a.wg0, a.n, blockIdx.x);
    const int64_t k = (int64_t)(blockIdx.x - a.wg0[s]) * DELTA_WAVES + (threadIdx.x >> 6);
    if (k >= a.count[s]) return;
    if (F32)
        pct_blk::delta_f32(a.list[s], a.stride[s], a.pos[s], a.wepl[s], a.h0[s], k, a.x[s], (float)a.lambda[s],
                           (float*)a.delta[s], threadIdx.x & 63);
    else
        pct_blk::delta_f64(a.list[s], a.stride[s], a.pos[s], a.wepl[s], a.h0[s], k, a.x[s], a.lambda[s],
                           (double*)a.delta[s], threadIdx.x & 63);
}

template <bool F32>
__global__ void __launch_bounds__(256)
k_batch_apply(batch_apply_args a) {
    __builtin_amdgcn_s_setprio(3);
    const int s = batch_scan_of(a.wg0, a.n, blockIdx.x);
    const int64_t k = (int64_t)(blockIdx.x - a.wg0[s]) * 256 + threadIdx.x;
    if (k >= a.ns[s]) return;
    if (F32) pct_blk::apply_f32(a.segs[s][k], a.csc[s], (const float*)a.delta[s], a.x[s]);
    else pct_blk::apply_f64(a.segs[s][k], a.csc[s], (const double*)a.delta[s], a.x[s]);
}

}  // namespace

/* chunks of block steps: small first, doubling up to BATCH_CHUNK_MAX */
#define BATCH_CHUNK_FIRST 16
#define BATCH_CHUNK_MAX 256

struct pct_batch {
    std::vector<pct_ctx*> scans;
    hipStream_t aux = nullptr;
    int64_t steps = 0;
    std::vector<int64_t> chunk;            /* step boundaries */
    std::vector<hipGraphExec_t> graphs;
    std::vector<hipEvent_t> ev_chunk;
    hipEvent_t ev_start = nullptr;
    /* what the graphs were captured for */
    std::vector<float*> g_x;
    std::vector<const float*> g_wepl;
    std::vector<double> g_lambda;
    /* generation of every scan's plan when the graphs were captured: a plan
     * rebuilt by its own context (pct_drop_iteration, pct_drop_prepare) may
     * hold other buffers and offsets */
    std::vector<uint64_t> g_gen;
    bool g_f32 = false;
};

static void batch_release_graphs(pct_batch* b) {
    for (hipGraphExec_t g : b->graphs)
        if (g) (void)hipGraphExecDestroy(g);
    b->graphs.clear();
}

extern "C" pct_batch* pct_batch_create(pct_ctx* const* scans, int n_scans) {
    if (!scans || n_scans <= 0 || n_scans > PCT_BATCH_MAX) return nullptr;
    for (int s = 0; s < n_scans; ++s) {
        if (!scans[s] || scans[s]->device != scans[0]->device) return nullptr;
        for (int r = 0; r < s; ++r)
            if (scans[r] == scans[s]) return nullptr;     /* one plan per context: a scan at most once */
    }
    pct_batch* b = new pct_batch();
    b->scans.assign(scans, scans + n_scans);
    if (hipStreamCreateWithFlags(&b->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_start, hipEventDisableTiming) != hipSuccess) {
        pct_batch_destroy(b);
        return nullptr;
    }
    return b;
}

extern "C" void pct_batch_destroy(pct_batch* b) {
    if (!b) return;
    if (!b->scans.empty() && b->scans[0]->stream) (void)hipStreamSynchronize(b->scans[0]->stream);
    if (b->aux) (void)hipStreamSynchronize(b->aux);
    batch_release_graphs(b);
    for (hipEvent_t e : b->ev_chunk)
        if (e) (void)hipEventDestroy(e);
    if (b->ev_start) (void)hipEventDestroy(b->ev_start);
    if (b->aux) (void)hipStreamDestroy(b->aux);
    delete b;
}

static int check_batch_args(pct_batch* b, const pct_hist_soa* hists, const int* block_sizes, int mode, const char* who) {
    if (!b) return PCT_E_INVALID;
    pct_ctx* c0 = b->scans[0];
    if (!hists || !block_sizes) return pct_fail(c0, PCT_E_INVALID, "%s: null argument", who);
    const int mlp = mode & PCT_MLP_MODE_MASK;
    for (pct_ctx* c : b->scans)
        if (!pct_mlp_mode_ok(c, mlp)) return pct_fail(c0, PCT_E_INVALID, "%s: MLP mode %d not available", who, mlp);
    if (mode & PCT_DROP_ATOMIC) return pct_fail(c0, PCT_E_INVALID, "%s: no atomic form in a batch", who);
    if (mode & (PCT_DROP_ROBUST | PCT_DROP_ROBUST2))
        return pct_fail(c0, PCT_E_INVALID, "%s: the robust updates run per scan (pct_drop_iteration)", who);
    for (size_t s = 0; s < b->scans.size(); ++s) {
        const pct_hist_soa& h = hists[s];
        if (!h.wepl || !h.first_voxel || !h.x_entry || h.n <= 0)
            return pct_fail(c0, PCT_E_INVALID, "%s: scan %d has no histories", who, (int)s);
        if (block_sizes[s] <= 0) return pct_fail(c0, PCT_E_INVALID, "%s: scan %d block size %d", who, (int)s, block_sizes[s]);
    }
    return PCT_OK;
}

/* builds the plans that are missing or stale; returns 1 in *rebuilt if any was */
static int batch_plans(pct_batch* b, const pct_hist_soa* hists, const int* block_sizes, int mlp, bool* rebuilt) {
    *rebuilt = false;
    for (size_t s = 0; s < b->scans.size(); ++s) {
        pct_ctx* c = b->scans[s];
        if (pct_plan_ready(c, &hists[s], block_sizes[s], mlp)) continue;
        int rc = pct_plan_build(c, &hists[s], block_sizes[s], mlp);
        if (rc) {
            pct_fail(b->scans[0], rc, "scan %d: %s", (int)s, pct_last_error(c));
            return rc;
        }
        *rebuilt = true;
    }
    int64_t steps = 0;
    for (pct_ctx* c : b->scans) steps = std::max(steps, c->plan->nb);
    for (size_t s = 0; !*rebuilt && s < b->scans.size() && s < b->g_gen.size(); ++s)
        if (b->g_gen[s] != b->scans[s]->plan->generation) batch_release_graphs(b);
    if (*rebuilt || steps != b->steps || b->chunk.empty()) {
        batch_release_graphs(b);
        b->steps = steps;
        b->chunk.assign(1, 0);
        for (int64_t size = BATCH_CHUNK_FIRST; b->chunk.back() < steps; size = std::min<int64_t>(2 * size, BATCH_CHUNK_MAX))
            b->chunk.push_back(std::min(steps, b->chunk.back() + size));
    }
    return PCT_OK;
}

extern "C" int pct_batch_prepare(pct_batch* b, const pct_hist_soa* hists, const int* block_sizes, int mode) {
    int rc = check_batch_args(b, hists, block_sizes, mode, "pct_batch_prepare");
    if (rc) return rc;
    bool rebuilt;
    return batch_plans(b, hists, block_sizes, mode & PCT_MLP_MODE_MASK, &rebuilt);
}

/* block steps [j0, j1) on the batch stream */
static int batch_steps(pct_batch* b, float* const* xs, const pct_hist_soa* hists, const float* lambdas, bool f32,
                       int64_t j0, int64_t j1) {
    pct_ctx* c0 = b->scans[0];
    const int n = (int)b->scans.size();
    for (int64_t j = j0; j < j1; ++j) {
        batch_delta_args da{};
        batch_apply_args aa{};
        uint32_t dwg = 0, awg = 0;
        for (int s = 0; s < n; ++s) {
            const pct_plan* pl = b->scans[s]->plan;
            if (j >= pl->nb) continue;
            const int64_t h0 = j * pl->B, cnt = std::min<int64_t>(pl->B, pl->n - h0);
            const int64_t s0 = pl->blk_seg_h[j], ns = pl->blk_seg_h[j + 1] - s0;
            int d = da.n++;
            da.wg0[d] = dwg;
            da.list[d] = (const uint32_t*)pl->list.ptr + pl->list_base_h[j];
            da.stride[d] = pl->stride_h[j];
            da.pos[d] = pct_blk::blk_pos{(const uint16_t*)pl->pos_k.ptr + h0, (const uint32_t*)pl->pos_len.ptr + h0,
                                         (const double*)pl->pos_chord.ptr + h0};
            da.wepl[d] = hists[s].wepl;
            da.h0[d] = h0;
            da.count[d] = cnt;
            da.x[d] = xs[s];
            da.lambda[d] = (double)lambdas[s];
            da.delta[d] = pl->delta.ptr;
            dwg += (uint32_t)((cnt + DELTA_WAVES - 1) / DELTA_WAVES);
            if (ns > 0) {
                int a = aa.n++;
                aa.wg0[a] = awg;
                aa.segs[a] = (const uint64_t*)pl->segs.ptr + s0;
                aa.ns[a] = ns;
                aa.csc[a] = (const uint16_t*)pl->csc.ptr + pl->blk_entry_h[j];
                aa.delta[a] = pl->delta.ptr;
                aa.x[a] = xs[s];
                awg += (uint32_t)((ns + 255) / 256);
            }
        }
        da.wg0[da.n] = dwg;
        aa.wg0[aa.n] = awg;
        if (da.n == 0) continue;
        if (f32) hipLaunchKernelGGL(k_batch_delta<true>, dim3(dwg), dim3(64 * DELTA_WAVES), 0, c0->stream, da);
        else hipLaunchKernelGGL(k_batch_delta<false>, dim3(dwg), dim3(64 * DELTA_WAVES), 0, c0->stream, da);
        if (aa.n > 0) {
            if (f32) hipLaunchKernelGGL(k_batch_apply<true>, dim3(awg), dim3(256), 0, c0->stream, aa);
            else hipLaunchKernelGGL(k_batch_apply<false>, dim3(awg), dim3(256), 0, c0->stream, aa);
        }
    }
    PCT_HIP_CHECK(c0, hipGetLastError());
    return PCT_OK;
}

static int batch_graphs(pct_batch* b, float* const* xs, const pct_hist_soa* hists, const float* lambdas, bool f32) {
    const size_t n = b->scans.size();
    bool same = b->graphs.size() + 1 == b->chunk.size() && b->g_f32 == f32 && b->g_x.size() == n;
    for (size_t s = 0; same && s < n; ++s)
        same = b->g_x[s] == xs[s] && b->g_wepl[s] == hists[s].wepl && b->g_lambda[s] == (double)lambdas[s] &&
               b->g_gen.size() == n && b->g_gen[s] == b->scans[s]->plan->generation;
    if (same) return PCT_OK;
    batch_release_graphs(b);
    pct_ctx* c0 = b->scans[0];
    for (size_t c = 0; c + 1 < b->chunk.size(); ++c) {
        hipGraph_t g = nullptr;
        PCT_HIP_CHECK(c0, hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
        int rc = batch_steps(b, xs, hists, lambdas, f32, b->chunk[c], b->chunk[c + 1]);
        hipError_t e = hipStreamEndCapture(c0->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); batch_release_graphs(b); return rc; }
        if (e != hipSuccess) { batch_release_graphs(b); return pct_fail(c0, PCT_E_HIP, "graph capture: %s", hipGetErrorString(e)); }
        hipGraphExec_t ge = nullptr;
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) { batch_release_graphs(b); return pct_fail(c0, PCT_E_HIP, "graph instantiate: %s", hipGetErrorString(e)); }
        b->graphs.push_back(ge);
    }
    b->g_x.assign(xs, xs + n);
    b->g_gen.resize(n);
    for (size_t s = 0; s < n; ++s) b->g_gen[s] = b->scans[s]->plan->generation;
    b->g_wepl.resize(n);
    b->g_lambda.resize(n);
    for (size_t s = 0; s < n; ++s) {
        b->g_wepl[s] = hists[s].wepl;
        b->g_lambda[s] = (double)lambdas[s];
    }
    b->g_f32 = f32;
    return PCT_OK;
}

extern "C" int pct_batch_iteration(pct_batch* b, const pct_hist_soa* hists, float* const* xs, const int* block_sizes,
                                   const float* lambdas, int mode) {
    int rc = check_batch_args(b, hists, block_sizes, mode, "pct_batch_iteration");
    if (rc) return rc;
    pct_ctx* c0 = b->scans[0];
    if (!xs || !lambdas) return pct_fail(c0, PCT_E_INVALID, "pct_batch_iteration: null images or lambdas");
    const int n = (int)b->scans.size();
    for (int s = 0; s < n; ++s)
        if (!xs[s]) return pct_fail(c0, PCT_E_INVALID, "pct_batch_iteration: scan %d has no image", s);
    const int mlp = mode & PCT_MLP_MODE_MASK;
    const bool f32 = (mode & PCT_DROP_F32) != 0, eager = (mode & PCT_DROP_NO_GRAPH) != 0;
    bool rebuilt;
    if ((rc = batch_plans(b, hists, block_sizes, mlp, &rebuilt))) return rc;
    const bool recompute = !(mode & PCT_DROP_CACHED) && !rebuilt;
    if (!eager && (rc = batch_graphs(b, xs, hists, lambdas, f32))) return rc;
    const int64_t nchunks = (int64_t)b->chunk.size() - 1;
    while ((int64_t)b->ev_chunk.size() < nchunks) {
        hipEvent_t e = nullptr;
        PCT_HIP_CHECK(c0, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->ev_chunk.push_back(e);
    }
    if (recompute) {
        /* lists that differ from a plan are only known after the block steps
         * ran on them: keep every image to restore it then */
        for (int s = 0; s < n; ++s) {
            pct_ctx* cs = b->scans[s];
            if ((rc = pct_ensure(cs, cs->x_saved, sizeof(float) * (size_t)cs->g.V))) {
                pct_fail(c0, rc, "scan %d: %s", s, pct_last_error(cs));
                return rc;
            }
            PCT_HIP_CHECK(c0, hipMemcpyAsync(cs->x_saved.ptr, xs[s], sizeof(float) * (size_t)cs->g.V,
                                             hipMemcpyDeviceToDevice, c0->stream));
            /* the surrogate's fallback list sized for the whole scan before
             * any chunk is queued (growing it later would wait for the device) */
            if (mlp == PCT_MLP_SURROGATE && (rc = pct_mlp_fallback_reserve(cs, cs->plan->n))) {
                pct_fail(c0, rc, "scan %d: %s", s, pct_last_error(cs));
                return rc;
            }
        }
    }
    if (c0->profile) PCT_HIP_CHECK(c0, hipEventRecord(c0->ev[0], c0->stream));
    if (recompute) {
        /* every scan's lists for the histories of each chunk of steps, on the
         * second stream, ahead of the block steps */
        PCT_HIP_CHECK(c0, hipEventRecord(b->ev_start, c0->stream));
        PCT_HIP_CHECK(c0, hipStreamWaitEvent(b->aux, b->ev_start, 0));
        for (int64_t c = 0; c < nchunks; ++c) {
            for (int s = 0; s < n; ++s) {
                pct_ctx* cs = b->scans[s];
                pct_plan* pl = cs->plan;
                const int64_t h0 = std::min(pl->n, b->chunk[c] * pl->B);
                const int64_t h1 = std::min(pl->n, b->chunk[c + 1] * pl->B);
                if (h1 <= h0) continue;
                rc = pct_mlp_paths_on(cs, b->aux, &hists[s], h0, h1 - h0, (const int64_t*)pl->hist_off.ptr + h0,
                                      nullptr, nullptr, nullptr, nullptr, (const uint32_t*)pl->hash.ptr + h0, mlp,
                                      false);
                if (rc) {
                    pct_fail(c0, rc, "scan %d: %s", s, pct_last_error(cs));
                    return rc;
                }
            }
            PCT_HIP_CHECK(c0, hipEventRecord(b->ev_chunk[c], b->aux));
            if (c0->profile && c == 0) PCT_HIP_CHECK(c0, hipEventRecord(c0->ev[1], b->aux));
        }
        if (c0->profile) PCT_HIP_CHECK(c0, hipEventRecord(c0->ev[2], b->aux));
    } else if (c0->profile) {
        PCT_HIP_CHECK(c0, hipEventRecord(c0->ev[1], c0->stream));
        PCT_HIP_CHECK(c0, hipEventRecord(c0->ev[2], c0->stream));
    }
    /* the block steps read the plans' lists: by default they wait for the
     * check of the recomputed lists only at the end; with overlap 1 of the
     * first scan's context, chunk c waits for the check of its lists */
    const bool chunk_waits = recompute && c0->tune_overlap == 1;
    for (int64_t c = 0; c < nchunks; ++c) {
        if (chunk_waits) PCT_HIP_CHECK(c0, hipStreamWaitEvent(c0->stream, b->ev_chunk[c], 0));
        if (eager) {
            if ((rc = batch_steps(b, xs, hists, lambdas, f32, b->chunk[c], b->chunk[c + 1]))) return rc;
        } else {
            PCT_HIP_CHECK(c0, hipGraphLaunch(b->graphs[c], c0->stream));
        }
    }
    if (recompute && !chunk_waits && nchunks > 0)
        PCT_HIP_CHECK(c0, hipStreamWaitEvent(c0->stream, b->ev_chunk[nchunks - 1], 0));
    if (c0->profile) {
        PCT_HIP_CHECK(c0, hipEventRecord(c0->ev[3], c0->stream));
        c0->ev_recorded = true;
    }
    /* every kernel of the iteration has joined the batch stream */
    PCT_HIP_CHECK(c0, hipStreamSynchronize(c0->stream));
    int first_rc = PCT_OK;
    for (int s = 0; s < n; ++s) {
        pct_ctx* cs = b->scans[s];
        int flags = 0;
        if ((rc = pct_take_status(cs, &flags))) return rc;
        if (flags) {
            cs->plan->valid = false;
            if (recompute) {
                PCT_HIP_CHECK(c0, hipMemcpyAsync(xs[s], cs->x_saved.ptr, sizeof(float) * (size_t)cs->g.V,
                                                 hipMemcpyDeviceToDevice, c0->stream));
                PCT_HIP_CHECK(c0, hipStreamSynchronize(c0->stream));
            }
        }
        if ((rc = pct_status_to_rc(cs, flags)) && first_rc == PCT_OK) {
            pct_fail(c0, rc, "scan %d: %s", s, pct_last_error(cs));
            first_rc = rc;
        }
    }
    return first_rc;
}
