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
 * pct_drop_iteration. With PCT_DROP_PERSISTENT all block steps of an
 * iteration run in one launch instead (k_batch_persist: grid barriers
 * between the phases, bounded waits).
 */
#include <algorithm>
#include <cstdio>
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

/* ------------------------------------------------------------------------- */
/* Persistent form: every block step of an iteration in ONE launch. The grid  */
/* is the device's resident capacity for this kernel (every workgroup is      */
/* resident at once: the iteration runs its MLP check before the launch, so   */
/* nothing of the batch runs beside it); block step j runs the deltas of      */
/* block j of every scan, a grid barrier, the voxel updates, a grid barrier.  */
/* A step's histories (segments) are numbered across the scans from per-step  */
/* prefix counts; wavefront w takes histories w, w + W, ... (thread t         */
/* segments t, t + T, ...). The arithmetic is the two-launch kernels', so     */
/* every scan's image is bit-identical to the graph form's.                   */
/* ------------------------------------------------------------------------- */
struct batch_scan {                 /* one scan's plan and image */
    const uint32_t* list;
    const uint16_t* pos_k;
    const uint32_t* pos_len;
    const double* pos_chord;
    const uint16_t* pos_r;
    const float* wepl;
    float* x;
    double lambda;
    void* delta;
    const uint64_t* segs;
    const uint16_t* csc;
    const pct_blk::long_row* rows;
};

struct batch_step {                 /* block j of scan s (a scan without one: cnt = ns = 0) */
    int64_t list_base, h0, s0, entry, l0;
    int32_t stride, cnt, ns, nlong;
};

struct batch_persist_args {
    int n;
    int64_t steps;
    batch_scan sc[PCT_BATCH_MAX];
    const batch_step* desc;         /* steps x n                                        */
    const int32_t* pre;             /* per step: n+1 history prefix counts, n+1 segment */
    unsigned* bar;                  /* barrier words, zero at the launch                */
    int* status;
    unsigned target_add;            /* added to every barrier's target (0; the          */
    int polls;                      /*   diagnostics knob makes it unreachable)         */
};

#define PERSIST_THREADS 256
/* workgroups per CU: the kernel's occupancy less one, at most 6 (every
 * workgroup must be resident at once; the margin keeps that true when a
 * CU holds a few other waves) */
#define PERSIST_WG_PER_CU 6
/* a barrier wait gives up after this long on the 100 MHz real-time counter
 * (0.5 s; a barrier takes microseconds), or after `polls` polls */
#define PERSIST_WAIT_TICKS 50000000ull
#define PERSIST_POLLS (1 << 30)

/* Grid barrier number `round` (1, 2, ...), two levels so that no counter
 * takes more than a few dozen arrivals: workgroup w arrives (one thread, at
 * agent scope, acquire-release) at the counter of its group w % PERSIST_GROUPS;
 * the last arrival of a group arrives at the top counter; the last arrival
 * there stores the round into the generation word, which every workgroup
 * polls (a load at agent scope after a short sleep), then an agent-scope
 * acquire. Each counter and the generation word have a 256-byte line of their
 * own (bar[0] generation, bar[64] top, bar[128 + 64 g] group g). A poll count
 * over `polls`, or the give-up flag set by another workgroup, sets the flag
 * and ends the wait; the caller then returns. */
#define PERSIST_GROUPS 32
#define PERSIST_BAR_WORDS (128 + 64 * PERSIST_GROUPS)
__device__ __forceinline__ bool persist_barrier(const batch_persist_args& a, unsigned round) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned G = gridDim.x, g = blockIdx.x % PERSIST_GROUPS;
        const unsigned groups = G < PERSIST_GROUPS ? G : PERSIST_GROUPS;
        const unsigned gsize = G / PERSIST_GROUPS + (g < G % PERSIST_GROUPS ? 1u : 0u);
        unsigned* gen = a.bar;
        unsigned* top = a.bar + 64;
        unsigned* grp = a.bar + 128 + 64 * g;
        if (__hip_atomic_fetch_add(grp, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u == round * gsize &&
            __hip_atomic_fetch_add(top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u == round * groups)
            __hip_atomic_store(gen, round, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned target = round + a.target_add;
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        for (int i = 0; __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; ++i) {
            if (i >= a.polls || __builtin_amdgcn_s_memrealtime() - t0 > PERSIST_WAIT_TICKS ||
                ((i & 63) == 63 && (__hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
                                    PCT_STATUS_WAIT))) {
                atomicOr(a.status, PCT_STATUS_WAIT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    /* every thread reads the flag: no LDS, so a workgroup needs none */
    return !(__hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & PCT_STATUS_WAIT);
}

__device__ __forceinline__ int persist_scan_of(const int32_t* pre, int n, int64_t f) {
    int s = 0;
    while (s + 1 < n && f >= pre[s + 1]) ++s;
    return s;
}

template <bool F32>
__global__ void __launch_bounds__(PERSIST_THREADS)
k_batch_persist(batch_persist_args a) {
    __builtin_amdgcn_s_setprio(3);
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * (PERSIST_THREADS / 64);
    const int64_t wave = (int64_t)blockIdx.x * (PERSIST_THREADS / 64) + (threadIdx.x >> 6);
    const int64_t nthreads = (int64_t)gridDim.x * PERSIST_THREADS;
    const int64_t tid = (int64_t)blockIdx.x * PERSIST_THREADS + threadIdx.x;
    unsigned round = 0;
    for (int64_t j = 0; j < a.steps; ++j) {
        const int32_t* ph = a.pre + j * 2 * (a.n + 1);
        const int32_t* ps = ph + (a.n + 1);
        const batch_step* d = a.desc + j * a.n;
        /* deltas of block j of every scan */
        for (int64_t f = wave; f < ph[a.n]; f += nwaves) {
            const int s = persist_scan_of(ph, a.n, f);
            const batch_scan& c = a.sc[s];
            const batch_step& e = d[s];
            const pct_blk::blk_pos pos{c.pos_k + e.h0, c.pos_len + e.h0, c.pos_chord + e.h0, c.pos_r + e.h0};
            if (F32)
                pct_blk::delta_f32(c.list + e.list_base, e.stride, pos, c.wepl, e.h0, f - ph[s], c.x,
                                   (float)c.lambda, (float*)c.delta, lane);
            else
                pct_blk::delta_f64(c.list + e.list_base, e.stride, pos, c.wepl, e.h0, f - ph[s], c.x, c.lambda,
                                   (double*)c.delta, lane);
        }
        if (!persist_barrier(a, ++round)) return;
        /* their voxel updates */
        for (int64_t f = tid; f < ps[a.n]; f += nthreads) {
            const int s = persist_scan_of(ps, a.n, f);
            const batch_scan& c = a.sc[s];
            const batch_step& e = d[s];
            const int64_t k = f - ps[s];
            pct_blk::apply_seg<F32>(k, e.nlong, c.segs[e.s0 + k], c.rows + e.l0, c.csc + e.entry, c.delta, c.x);
        }
        if (!persist_barrier(a, ++round)) return;
    }
}

__global__ void k_zero_u32(unsigned* __restrict__ p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

}  // namespace

/* chunks of block steps when chunk c waits for its own MLP check (overlap 1
 * of the first scan's context): small first, doubling up to BATCH_CHUNK_MAX.
 * Otherwise one chunk: the check of every scan is one MLP launch per scan,
 * no launch a fraction of the resident grid. */
#define BATCH_CHUNK_FIRST 16
#define BATCH_CHUNK_MAX 256

struct pct_batch {
    std::vector<pct_ctx*> scans;
    hipStream_t aux = nullptr;
    int64_t steps = 0;
    std::vector<int64_t> chunk;            /* step boundaries */
    int64_t chunk_first = 0, chunk_max = 0; /* what chunk was laid out with */
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
    bool g_persist = false;
    int g_wait_give_up = 0;
    /* persistent form: step descriptors and prefix counts, barrier counter */
    pct_buffer p_desc, p_bar;
    int64_t p_pre_off = 0;          /* byte offset of the prefix counts in p_desc */
    unsigned p_grid = 0;
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
    for (pct_buffer* q : {&b->p_desc, &b->p_bar})
        if (q->ptr) (void)hipFree(q->ptr);
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
    /* chunk layout from the first context's knobs, as pct_drop_iteration's */
    const pct_ctx* c0 = b->scans[0];
    const bool one = c0->tune_overlap != 1 && c0->tune_overlap != 2;
    const int64_t all = std::max<int64_t>(1, steps);
    const int64_t first = c0->tune_chunk_first > 0 ? c0->tune_chunk_first : (one ? all : BATCH_CHUNK_FIRST);
    const int64_t most = std::max(first, c0->tune_chunk_max > 0 ? c0->tune_chunk_max : (one ? all : BATCH_CHUNK_MAX));
    if (*rebuilt || steps != b->steps || b->chunk.empty() || first != b->chunk_first || most != b->chunk_max) {
        batch_release_graphs(b);
        b->steps = steps;
        b->chunk_first = first;
        b->chunk_max = most;
        b->chunk.assign(1, 0);
        for (int64_t size = first; b->chunk.back() < steps; size = std::min<int64_t>(2 * size, most))
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
                                         (const double*)pl->pos_chord.ptr + h0, (const uint16_t*)pl->pos_r.ptr + h0};
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

/* the persistent launch's arguments: per scan its plan and image, per block
 * step its blocks' descriptors (uploaded to p_desc) */
static int batch_persist_args_of(pct_batch* b, float* const* xs, const pct_hist_soa* hists, const float* lambdas,
                                 bool f32, batch_persist_args& a) {
    pct_ctx* c0 = b->scans[0];
    const int n = (int)b->scans.size();
    const int64_t steps = b->steps;
    std::vector<batch_step> desc((size_t)(steps * n));
    std::vector<int32_t> pre((size_t)(steps * 2 * (n + 1)), 0);
    a = batch_persist_args{};
    a.n = n;
    a.steps = steps;
    for (int s = 0; s < n; ++s) {
        const pct_plan* pl = b->scans[s]->plan;
        a.sc[s] = batch_scan{(const uint32_t*)pl->list.ptr, (const uint16_t*)pl->pos_k.ptr,
                             (const uint32_t*)pl->pos_len.ptr, (const double*)pl->pos_chord.ptr,
                             (const uint16_t*)pl->pos_r.ptr, hists[s].wepl, xs[s],
                             (double)lambdas[s], pl->delta.ptr, (const uint64_t*)pl->segs.ptr,
                             (const uint16_t*)pl->csc.ptr, (const pct_blk::long_row*)pl->long_rows.ptr};
    }
    for (int64_t j = 0; j < steps; ++j) {
        int32_t* ph = pre.data() + j * 2 * (n + 1);
        int32_t* ps = ph + (n + 1);
        for (int s = 0; s < n; ++s) {
            const pct_plan* pl = b->scans[s]->plan;
            batch_step e{};
            if (j < pl->nb) {
                e.h0 = j * pl->B;
                e.cnt = (int32_t)std::min<int64_t>(pl->B, pl->n - e.h0);
                e.list_base = pl->list_base_h[j];
                e.stride = pl->stride_h[j];
                e.s0 = pl->blk_seg_h[j];
                e.ns = (int32_t)(pl->blk_seg_h[j + 1] - e.s0);
                e.entry = pl->blk_entry_h[j];
                e.l0 = pl->blk_long_h[j];
                e.nlong = (int32_t)(pl->blk_long_h[j + 1] - e.l0);
            }
            desc[(size_t)(j * n + s)] = e;
            ph[s + 1] = ph[s] + e.cnt;
            ps[s + 1] = ps[s] + e.ns;
        }
    }
    const size_t dbytes = sizeof(batch_step) * desc.size();
    b->p_pre_off = (int64_t)((dbytes + 255) / 256 * 256);
    int rc;
    if ((rc = pct_ensure(c0, b->p_desc, (size_t)b->p_pre_off + sizeof(int32_t) * pre.size())) ||
        (rc = pct_ensure(c0, b->p_bar, sizeof(unsigned) * PERSIST_BAR_WORDS)))
        return rc;
    PCT_HIP_CHECK(c0, hipMemcpy(b->p_desc.ptr, desc.data(), dbytes, hipMemcpyHostToDevice));
    PCT_HIP_CHECK(c0, hipMemcpy((char*)b->p_desc.ptr + b->p_pre_off, pre.data(), sizeof(int32_t) * pre.size(),
                                hipMemcpyHostToDevice));
    a.desc = (const batch_step*)b->p_desc.ptr;
    a.pre = (const int32_t*)((const char*)b->p_desc.ptr + b->p_pre_off);
    a.bar = (unsigned*)b->p_bar.ptr;
    a.status = c0->d_status;
    /* the diagnostics knob: a target no arrival count reaches, two polls */
    a.target_add = c0->tune_wait_give_up ? 0x40000000u : 0u;
    a.polls = c0->tune_wait_give_up ? 2 : PERSIST_POLLS;
    /* every workgroup resident at once: the kernel's occupancy on every CU */
    int occ = 0;
    PCT_HIP_CHECK(c0, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &occ, f32 ? (const void*)k_batch_persist<true> : (const void*)k_batch_persist<false>,
                          PERSIST_THREADS, 0));
    int cus = 0;
    PCT_HIP_CHECK(c0, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c0->device));
    b->p_grid = (unsigned)(std::max(1, cus) * std::max(1, std::min(occ - 1, PERSIST_WG_PER_CU)));
    return PCT_OK;
}

/* the persistent form's launch of one iteration: zero the barrier counter
 * (a kernel, so that the graph holds no memset node), then the loop */
static int batch_persist_launch(pct_batch* b, const batch_persist_args& a, bool f32) {
    pct_ctx* c0 = b->scans[0];
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, c0->stream, (unsigned*)b->p_bar.ptr, PERSIST_BAR_WORDS);
    if (f32) hipLaunchKernelGGL(k_batch_persist<true>, dim3(b->p_grid), dim3(PERSIST_THREADS), 0, c0->stream, a);
    else hipLaunchKernelGGL(k_batch_persist<false>, dim3(b->p_grid), dim3(PERSIST_THREADS), 0, c0->stream, a);
    PCT_HIP_CHECK(c0, hipGetLastError());
    return PCT_OK;
}

static int batch_graphs(pct_batch* b, float* const* xs, const pct_hist_soa* hists, const float* lambdas, bool f32,
                        bool persist) {
    const size_t n = b->scans.size();
    const size_t want = persist ? 1 : b->chunk.size() - 1;
    bool same = b->graphs.size() == want && b->g_f32 == f32 && b->g_persist == persist && b->g_x.size() == n &&
                b->g_wait_give_up == b->scans[0]->tune_wait_give_up;
    for (size_t s = 0; same && s < n; ++s)
        same = b->g_x[s] == xs[s] && b->g_wepl[s] == hists[s].wepl && b->g_lambda[s] == (double)lambdas[s] &&
               b->g_gen.size() == n && b->g_gen[s] == b->scans[s]->plan->generation;
    if (same) return PCT_OK;
    batch_release_graphs(b);
    pct_ctx* c0 = b->scans[0];
    batch_persist_args pa{};
    if (persist) {
        int rc = batch_persist_args_of(b, xs, hists, lambdas, f32, pa);
        if (rc) return rc;
    }
    for (size_t c = 0; c < want; ++c) {
        hipGraph_t g = nullptr;
        PCT_HIP_CHECK(c0, hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
        int rc = persist ? batch_persist_launch(b, pa, f32)
                         : batch_steps(b, xs, hists, lambdas, f32, b->chunk[c], b->chunk[c + 1]);
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
    b->g_persist = persist;
    b->g_wait_give_up = c0->tune_wait_give_up;
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
    const bool persist = (mode & PCT_DROP_PERSISTENT) != 0;
    bool rebuilt;
    if ((rc = batch_plans(b, hists, block_sizes, mlp, &rebuilt))) return rc;
    const bool recompute = !(mode & PCT_DROP_CACHED) && !rebuilt;
    /* the block steps run on each plan's tiled image: x copied in first,
     * back after the last step */
    std::vector<float*> xts((size_t)n, nullptr);
    for (int s = 0; s < n; ++s)
        if ((rc = pct_plan_image_in(b->scans[s], c0->stream, xs[s], &xts[s]))) {
            pct_fail(c0, rc, "scan %d: %s", s, pct_last_error(b->scans[s]));
            return rc;
        }
    float* const* xt = xts.data();
    if (!eager && (rc = batch_graphs(b, xt, hists, lambdas, f32, persist))) return rc;
    const int64_t nchunks = (int64_t)b->chunk.size() - 1;
    while ((int64_t)b->ev_chunk.size() < nchunks) {
        hipEvent_t e = nullptr;
        PCT_HIP_CHECK(c0, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->ev_chunk.push_back(e);
    }
    if (recompute || persist) {
        /* lists that differ from a plan are only known after the block steps
         * ran on them, and a persistent launch whose barrier gave up stops
         * midway: keep every image to restore it then */
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
            if (recompute && mlp == PCT_MLP_SURROGATE && (rc = pct_mlp_fallback_reserve(cs, cs->plan->n))) {
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
                rc = mlp == PCT_MLP_FP64
                         ? pct_mlp_verify_lanes(cs, b->aux, &hists[s], h0, h1 - h0, (const uint16_t*)pl->lane_of.ptr,
                                                (const uint32_t*)pl->hash.ptr + h0)
                         : pct_mlp_paths_on(cs, b->aux, &hists[s], h0, h1 - h0, (const int64_t*)pl->hist_off.ptr + h0,
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
    const bool chunk_waits = recompute && c0->tune_overlap == 1 && !persist;
    if (persist) {
        /* the persistent launch needs every workgroup resident: it follows
         * the MLP check instead of running beside it */
        if (recompute && nchunks > 0) PCT_HIP_CHECK(c0, hipStreamWaitEvent(c0->stream, b->ev_chunk[nchunks - 1], 0));
        if (eager) {
            batch_persist_args pa{};
            if ((rc = batch_persist_args_of(b, xt, hists, lambdas, f32, pa)) ||
                (rc = batch_persist_launch(b, pa, f32)))
                return rc;
        } else {
            PCT_HIP_CHECK(c0, hipGraphLaunch(b->graphs[0], c0->stream));
        }
    }
    for (int64_t c = 0; !persist && c < nchunks; ++c) {
        if (chunk_waits) PCT_HIP_CHECK(c0, hipStreamWaitEvent(c0->stream, b->ev_chunk[c], 0));
        if (eager) {
            if ((rc = batch_steps(b, xt, hists, lambdas, f32, b->chunk[c], b->chunk[c + 1]))) return rc;
        } else {
            PCT_HIP_CHECK(c0, hipGraphLaunch(b->graphs[c], c0->stream));
        }
    }
    if (recompute && !chunk_waits && !persist && nchunks > 0)
        PCT_HIP_CHECK(c0, hipStreamWaitEvent(c0->stream, b->ev_chunk[nchunks - 1], 0));
    for (int s = 0; s < n; ++s)
        if ((rc = pct_plan_image_out(b->scans[s], c0->stream, xs[s]))) {
            pct_fail(c0, rc, "scan %d: %s", s, pct_last_error(b->scans[s]));
            return rc;
        }
    if (c0->profile) {
        PCT_HIP_CHECK(c0, hipEventRecord(c0->ev[3], c0->stream));
        c0->ev_recorded = true;
    }
    /* every kernel of the iteration has joined the batch stream */
    PCT_HIP_CHECK(c0, hipStreamSynchronize(c0->stream));
    int first_rc = PCT_OK;
    /* a persistent launch reports a barrier that gave up in scans[0]'s status */
    int flags0 = 0;
    if ((rc = pct_take_status(c0, &flags0))) return rc;
    const bool gave_up = persist && (flags0 & PCT_STATUS_WAIT);
    char arrivals[96] = "";
    if (gave_up && b->p_bar.ptr) {
        /* what the barrier words held: how far the launch got */
        unsigned w[128];
        PCT_HIP_CHECK(c0, hipMemcpy(w, b->p_bar.ptr, sizeof(w), hipMemcpyDeviceToHost));
        snprintf(arrivals, sizeof(arrivals), " (grid %u: generation %u, %u group arrivals)", b->p_grid, w[0], w[64]);
    }
    for (int s = 0; s < n; ++s) {
        pct_ctx* cs = b->scans[s];
        int flags = s == 0 ? flags0 : 0;
        if (s > 0 && (rc = pct_take_status(cs, &flags))) return rc;
        if (gave_up && s > 0) flags |= PCT_STATUS_WAIT;
        /* a barrier that gave up leaves the plan intact; anything else the
         * MLP check found invalidates it */
        if (flags & ~PCT_STATUS_WAIT) cs->plan->valid = false;
        if ((flags & PCT_STATUS_WAIT) || (flags && recompute)) {
            PCT_HIP_CHECK(c0, hipMemcpyAsync(xs[s], cs->x_saved.ptr, sizeof(float) * (size_t)cs->g.V,
                                             hipMemcpyDeviceToDevice, c0->stream));
            PCT_HIP_CHECK(c0, hipStreamSynchronize(c0->stream));
        }
        if ((rc = pct_status_to_rc(cs, flags)) && first_rc == PCT_OK) {
            pct_fail(c0, rc, "scan %d: %s%s", s, pct_last_error(cs), gave_up ? arrivals : "");
            first_rc = rc;
        }
    }
    return first_rc;
}
