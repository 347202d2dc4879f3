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


def preprocess(ctx: Context, raw: RawScan, hull_kind: str = "msc") -> Preprocessed:
    """Raw tracker data -> reconstruction histories, hull and x0."""
    torch = ctx.torch
    n_raw = raw.n
    hist, missed = ctx.recon_volume_intersections(raw)
    in_range = ctx.binning(hist, missed)
    if hull_kind == "msc":
        counts = ctx.msc_counts(hist, missed)
    elif hull_kind == "sc":
        sc_bits = ctx.space_carve(hist, missed, ctx.hull_init())
    st = ctx.bin_statistics(hist, missed, in_range)
    failed, sinogram = ctx.statistical_cuts(hist, missed, st)
    post_cut = ctx.compact(hist, (failed == 0).to(torch.uint8))
    fbp = ctx.backproject(ctx.fbp_filter(sinogram))
    if hull_kind == "msc":
        hull = ctx.hull_select(ctx.msc_edge(counts))
    elif hull_kind == "sc":
        hull = sc_bits
    elif hull_kind == "fbp":
        hull = ctx.fbp_threshold(fbp)
    elif hull_kind == "cylinder":
        hull = ctx.hull_init()
    else:
        raise ValueError(f"unknown hull kind {hull_kind}")
    x0 = ctx.initial_iterate(fbp, hull)
    inter = ctx.mlp_endpoints(post_cut, hull)
    recon = ctx.compact(post_cut, inter)
    ctx.sync()
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
