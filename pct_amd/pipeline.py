"""The reconstruction pipeline of the reference's main() (.cu:24-253) over the
C ABI, stage by stage, with the reference's stage names:

  recon_volume_intersections (.cu:2066)  -> missed
  binning (.cu:2388)                      -> bin numbers, kept = not missed
  hull_detection / MSC (.cu:3265, 3446)   -> MSC counts
  calculate_means .. statistical_cuts (.cu:2582-2869) -> failed, sinogram
  FBP (.cu:2919): filter + backprojection -> FBP image (+ FBP hull)
  hull_detection_finish / hull_selection (.cu:3892, 3954) -> hull
  define_initial_iterate (.cu:7191)       -> x0 = FBP * hull
  reconstruction_cuts (.cu:6142)          -> MLP endpoints + compaction
  image_reconstruction_GPU (.cu:6538)     -> DROP iterations

All device work runs in libpct's HIP kernels; this module only sequences the
calls and keeps the intermediate device tensors.
"""
from dataclasses import dataclass, field

from . import Context, Histories, RawScan, PCT_MLP_FP64


@dataclass
class Preprocessed:
    """Everything the DROP loop needs, resident on the device."""
    hist: Histories              # reconstruction histories (post cuts, inside the hull)
    x0: object                   # initial iterate (f32, V)
    hull: object                 # bit-packed MLP hull
    fbp: object                  # FBP image (f32, V)
    sinogram: object             # post-cut mean-WEPL sinogram
    counts: dict = field(default_factory=dict)


# Algorithmic HBM bytes of each stage, the roofline numerator: every array the
# reference's kernels for the stage read or write, counted once at its stored
# width. n = histories in, k = histories kept (or walked) by the stage, V =
# voxels, B = sinogram bins, Bt = bins holding at least one history. A history
# row after the cuts is bin_num, gantry angle, WEPL, 6 positions and 4 angles:
# 52 B. Per-voxel or per-bin atomics count once per array element.
STAGE_BYTES = {
    # recon_volume_intersections_GPU (.cu:2169): gantry angle and 12 tracker
    # coordinates in; missed, 6 positions and 4 angles out
    "recon_volume_intersections": lambda n, k, V, B, Bt: 52 * n + 41 * n,
    # binning_GPU (.cu:2507): missed, 6 positions, WEPL and 4 angles in,
    # bin_num out; per touched bin the count and three sums
    "binning": lambda n, k, V, B, Bt: 45 * n + 4 * n + 16 * Bt,
    # calculate_means_GPU (.cu:2618): per bin the count and three sums in,
    # three means out; sum_squared_deviations_GPU (.cu:2671): per history
    # bin_num, WEPL and 4 angles, per touched bin three means in and three
    # sums out; calculate_standard_deviations_GPU (.cu:2718): per bin the
    # count and three sums in, three deviations out
    "bin_statistics": lambda n, k, V, B, Bt: 28 * B + 24 * n + 24 * Bt + 28 * B,
    # statistical_cuts_GPU (.cu:2835): per history bin_num, WEPL, 4 angles
    # in, failed out; per touched bin count, means, deviations in and the
    # sinogram sum out; construct_sinogram_GPU (.cu:2912): per bin the count
    # and the sum in, the mean out
    "statistical_cuts": lambda n, k, V, B, Bt: 25 * n + 32 * Bt + 12 * B,
    # MSC_GPU (.cu:3456), SM_GPU (.cu:3608): missed and WEPL per history, the
    # 6 positions of the k histories that walk; the counts read and written
    # once
    "msc_counts": lambda n, k, V, B, Bt: 5 * n + 24 * k + 8 * V,
    "sm_counts": lambda n, k, V, B, Bt: 5 * n + 24 * k + 8 * V,
    # SC_GPU (.cu:3338): as MSC, the hull at its stored width (1 bit here)
    "space_carve": lambda n, k, V, B, Bt: 5 * n + 24 * k + V / 4,
    # host compaction (.cu:2777-2809): a keep flag per history, each kept row
    # read and written
    "compact_post_cut": lambda n, k, V, B, Bt: n + 104 * k,
    # filter_GPU (.cu:3036): a bin in, a bin out
    "fbp_filter": lambda n, k, V, B, Bt: 8 * B,
    # backprojection_GPU (.cu:3155): the filtered sinogram in, the image out
    "backproject": lambda n, k, V, B, Bt: 4 * B + 4 * V,
    # MSC_edge_detection_GPU (.cu:3569), SM_edge_detection (.cu:3711): the
    # counts in, the edge image out
    "msc_edge": lambda n, k, V, B, Bt: 8 * V,
    "sm_edge": lambda n, k, V, B, Bt: 8 * V,
    # averaging_filter_GPU (.cu:4007) + threshold: image in, hull out
    "hull_select": lambda n, k, V, B, Bt: 4 * V + V / 8,
    "fbp_threshold": lambda n, k, V, B, Bt: 4 * V + V / 8,
    # define_initial_iterate (.cu:7191): FBP image and hull in, x out
    "initial_iterate": lambda n, k, V, B, Bt: 8 * V + V / 8,
    # collect_MLP_endpoints_GPU (.cu:5729): 6 positions and 4 angles in, 6
    # positions, first voxel and the intersected flag out; the hull once
    "mlp_endpoints": lambda n, k, V, B, Bt: 40 * n + 29 * n + V / 8,
    "compact_reconstruction": lambda n, k, V, B, Bt: n + 104 * k,
}

# HBM bytes this implementation moves per stage (reported beside the
# algorithmic bytes): its own arrays at their stored widths, including the
# radix sort of the statistics records, the bin ranges and cut windows that
# replace the reference's per-bin float atomics, and the raw row as read.
STAGE_TRAFFIC = {
    # raw row (14 x 4 B) in; 11 f32 + missed out
    "recon_volume_intersections": lambda n, k, V, B, Bt: 56 * n + 45 * n,
    # 11 f32 + missed in; bin_num, in_range, sort key and 16-byte statistics
    # record out; one key + record pass of the sort (20 B in, 20 B out);
    # sorted keys read for the bin ranges; first/last per bin written
    "binning": lambda n, k, V, B, Bt: 45 * n + 25 * n + 40 * n + 4 * n + 8 * B,
    # cylinder counts written; missed and wepl of every history, the 6 f32
    # positions of the k histories that walk; counts read and written once
    "msc_counts": lambda n, k, V, B, Bt: 4 * V + 5 * n + 24 * k + 8 * V,
    "sm_counts": lambda n, k, V, B, Bt: 4 * V + 5 * n + 24 * k + 8 * V,
    "space_carve": lambda n, k, V, B, Bt: 5 * n + 24 * k + V / 4,
    # sorted records once; ranges in, 7 statistics out per bin
    "bin_statistics": lambda n, k, V, B, Bt: 16 * n + 36 * B,
    # per history: missed, bin_num, wepl, 4 angles in, failed out, and its
    # sorted record; per bin: ranges and 6 statistics in, the 32-byte cut
    # window written, count and sinogram out
    "statistical_cuts": lambda n, k, V, B, Bt: 26 * n + 16 * n + 72 * B,
    "compact_post_cut": lambda n, k, V, B, Bt: n + 104 * k,
    "fbp_filter": lambda n, k, V, B, Bt: 8 * B,
    "backproject": lambda n, k, V, B, Bt: 4 * B + 4 * V,
    "msc_edge": lambda n, k, V, B, Bt: 8 * V,
    # counts read for the differences and again for the threshold test; hull out
    "sm_edge": lambda n, k, V, B, Bt: 12 * V,
    "hull_select": lambda n, k, V, B, Bt: 4 * V + V / 8,
    "fbp_threshold": lambda n, k, V, B, Bt: 4 * V + V / 8,
    "initial_iterate": lambda n, k, V, B, Bt: 8 * V + V / 8,
    # 10 f32 in; 6 f32, first_voxel and the intersected flag out
    "mlp_endpoints": lambda n, k, V, B, Bt: 40 * n + 29 * n,
    "compact_reconstruction": lambda n, k, V, B, Bt: n + 104 * k,
}


class _StageTimer:
    """HIP events on the context's stream around each stage, when the caller
    asks for timings: appends (stage, ms, algorithmic bytes, implementation
    bytes) per stage once the preprocessing is done (finish()).

    A spin kernel of about a millisecond runs on the stream ahead of the
    first event, so the host has queued the stage's launches by the time the
    event is stamped and the interval holds the stage's device work, not the
    host's launch latency (a stage that reads a count back to the host still
    pays for that round trip)."""

    GATE_CYCLES = 3_000_000

    def __init__(self, ctx, out):
        self.ctx, self.out = ctx, out
        self.pending = []
        self.touched_bins = 0

    def __call__(self, name, n, fn, *args, kept=None):
        if self.out is None:
            return fn(*args)
        torch = self.ctx.torch
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        self.ctx.sync()
        if hasattr(torch.cuda, "_sleep"):
            with torch.cuda.stream(self.ctx.stream):
                torch.cuda._sleep(self.GATE_CYCLES)
        e0.record(self.ctx.stream)
        r = fn(*args)
        e1.record(self.ctx.stream)
        torch.cuda.synchronize()
        k = kept(r) if kept is not None else 0
        self.pending.append((name, e0.elapsed_time(e1), n, k))
        return r

    def finish(self):
        if self.out is None:
            return
        V, B, Bt = self.ctx.V, self.ctx.n_bins, self.touched_bins
        for name, ms, n, k in self.pending:
            self.out.append((name, ms, float(STAGE_BYTES[name](n, k, V, B, Bt)),
                             float(STAGE_TRAFFIC[name](n, k, V, B, Bt))))
        self.pending = []


def preprocess(ctx: Context, raw: RawScan, hull_kind: str = "msc", timings: list = None) -> Preprocessed:
    """Raw tracker data -> reconstruction histories, hull and x0. With a list
    as `timings`, each stage runs alone between two events and (stage, ms,
    algorithmic bytes, implementation bytes) is appended per stage."""
    torch = ctx.torch
    n_raw = raw.n
    T = _StageTimer(ctx, timings)
    hist, missed = T("recon_volume_intersections", n_raw, ctx.recon_volume_intersections, raw)
    in_range = T("binning", n_raw, ctx.binning, hist, missed)
    def walkers(thr):                     # histories whose straight path is carved
        return lambda _: int(((missed == 0) & (hist.wepl <= thr)).sum())
    def crossers(thr):                    # histories whose path is counted by SM
        return lambda _: int(((missed == 0) & (hist.wepl >= thr)).sum())
    if hull_kind == "msc":
        counts = T("msc_counts", n_raw, ctx.msc_counts, hist, missed, kept=walkers(ctx.params.msc_threshold))
    elif hull_kind == "sm":
        counts = T("sm_counts", n_raw, ctx.sm_counts, hist, missed, kept=crossers(ctx.params.sm_lower_threshold))
    elif hull_kind == "sc":
        sc_bits = T("space_carve", n_raw, ctx.space_carve, hist, missed, ctx.hull_init(),
                    kept=walkers(ctx.params.sc_threshold))
    st = T("bin_statistics", n_raw, ctx.bin_statistics, hist, missed, in_range)
    if timings is not None:
        T.touched_bins = int((st["counts"] > 0).sum())
    failed, sinogram = T("statistical_cuts", n_raw, ctx.statistical_cuts, hist, missed, st)
    post_cut = T("compact_post_cut", n_raw, ctx.compact, hist, (failed == 0).to(torch.uint8), kept=lambda h: h.n)
    fbp = T("backproject", 0, ctx.backproject, T("fbp_filter", 0, ctx.fbp_filter, sinogram))
    if hull_kind == "msc":
        hull = T("hull_select", 0, ctx.hull_select, T("msc_edge", 0, ctx.msc_edge, counts))
    elif hull_kind == "sm":
        hull = T("hull_select", 0, ctx.hull_select, T("sm_edge", 0, ctx.sm_edge, counts))
    elif hull_kind == "sc":
        hull = sc_bits
    elif hull_kind == "fbp":
        hull = T("fbp_threshold", 0, ctx.fbp_threshold, fbp)
    elif hull_kind == "cylinder":
        hull = ctx.hull_init()
    else:
        raise ValueError(f"unknown hull kind {hull_kind}")
    x0 = T("initial_iterate", 0, ctx.initial_iterate, fbp, hull)
    inter = T("mlp_endpoints", post_cut.n, ctx.mlp_endpoints, post_cut, hull)
    recon = T("compact_reconstruction", post_cut.n, ctx.compact, post_cut, inter, kept=lambda h: h.n)
    ctx.sync()
    T.finish()
    return Preprocessed(hist=recon, x0=x0, hull=hull, fbp=fbp, sinogram=sinogram,
                        counts={"histories": n_raw, "post_cut": post_cut.n, "reconstruction": recon.n})


def image_reconstruction(ctx: Context, pre: Preprocessed, iterations=None, lam=None, block_size=None,
                         mode=PCT_MLP_FP64, callback=None):
    """DROP iterations from x0 (image_reconstruction_GPU, .cu:6538-6866)."""
    iterations = ctx.params.iterations if iterations is None else iterations
    x = pre.x0.clone()
    for it in range(1, iterations + 1):
        ctx.drop_iteration(pre.hist, x, block_size, lam, mode)
        if callback is not None:
            callback(it, x)
    return x
