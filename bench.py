"""DROP-iteration benchmark on one MI355X.

A step is one DROP iteration (MLP voxel lists of every history + the
sequential block update, blocks of 3200 in storage order) over the
reconstruction histories of a synthetic scan resident in HBM. Workload: the
200x200x32 grid of 1x1x2.5 mm voxels with 20 M histories over 90 angles
(the plain FP64-MLP GPU configuration); pass --config cfg1 for 100x100x16 /
2 M histories.

Prints one JSON line with the metric (histories/s per DROP iteration), the
roofline of the dominant phase measured with HIP events on the stream the
kernels run on, and the CPU oracle timed on a bounded sample on the host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config cfg2]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64, TFLOP/s
# FP64 operations of one MLP step of the reference loop body (.cu:5376-5491):
# +,-,*,/ and each log/pow counted once.
MLP_FLOPS_PER_STEP = 235


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def batched_form(log):
    """The batched configuration: 12 scans of uneven size (grids 64x64x8 to
    256x256x48, 0.5-30 M histories, random ellipsoid phantoms), each
    preprocessed on the GPU, block sizes from the project's tuner, then timed
    DROP iterations of the whole batch (FP64 and precision forms) and, for
    comparison, the same scans' iterations one after the other."""
    import torch
    import pct_amd
    from pct_amd import configs, pipeline, tuner
    t0 = time.time()
    scans = configs.batch_scans()
    ctxs, hists, x0s, info = [], [], [], []
    for sc in scans:
        ctx = pct_amd.Context(sc.params)
        pre = pipeline.preprocess(ctx, ctx.generate(sc.synth))
        ctxs.append(ctx)
        hists.append(pre.hist)
        x0s.append(pre.x0)
        info.append({"name": sc.name, "grid": [sc.params.columns, sc.params.rows, sc.params.slices],
                     "histories": int(sc.synth.n_histories), "reconstruction_histories": int(pre.hist.n),
                     "ellipsoids": int(sc.synth.n_ell)})
        del pre
    setup_s = time.time() - t0
    batch = pct_amd.Batch(ctxs)
    best, table = tuner.tune(batch, hists, x0s, configs.BATCH_BLOCK_SIZES, mode=0, log=log)
    B = best["block_sizes"]
    n = sum(h.n for h in hists)
    res = {"scans": info, "setup_s": round(setup_s, 2), "tuner": table, "block_sizes": B,
           "reconstruction_histories": n}

    def timed_batch(mode, steps=3):
        batch.prepare(hists, B, mode & 0xff)
        xs = [x.clone() for x in x0s]
        batch.iteration(hists, xs, B, None, mode)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(ctxs[0].stream)
        for _ in range(steps):
            batch.iteration(hists, xs, B, None, mode)
        e1.record(ctxs[0].stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    P = pct_amd.PCT_DROP_PERSISTENT
    for name, mode in (("fp64", 0), ("precision", pct_amd.PCT_PRECISION), ("fp64_persistent", P),
                       ("precision_persistent", pct_amd.PCT_PRECISION | P)):
        ms = timed_batch(mode)
        res[name] = {"ms_per_iteration": round(ms, 3), "histories_per_s": round(n / (ms / 1e3), 1)}
        log(f"batched {name}: {ms:.2f} ms/iteration, {n / (ms / 1e3) / 1e6:.1f} M histories/s")
    # the block steps alone (cached lists): two launches per step (graphs)
    # against one persistent launch per iteration
    res["block_steps_cached_ms"] = {"two_launches_per_step": round(timed_batch(pct_amd.PCT_DROP_CACHED), 3),
                                    "persistent": round(timed_batch(pct_amd.PCT_DROP_CACHED | P), 3)}
    res["persistent_note"] = ("PCT_DROP_PERSISTENT: every block step of an iteration in one launch with grid "
                              "barriers; its MLP check runs before the launch instead of beside the steps")
    log(f"batched block steps alone (cached lists): {res['block_steps_cached_ms']}")
    seq = 0.0
    for ctx, h, x0, b in zip(ctxs, hists, x0s, B):
        x = x0.clone()
        ctx.drop_iteration(h, x, b)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(ctx.stream)
        ctx.drop_iteration(h, x, b)
        e1.record(ctx.stream)
        torch.cuda.synchronize()
        seq += e0.elapsed_time(e1)
    res["scans_one_after_the_other_fp64_ms"] = round(seq, 3)
    batch.close()
    for c in ctxs:
        c.close()
    return res


def _rel_errors(x, refs, x0_np):
    inside = x0_np != 0
    out = {}
    for it, r in refs.items():
        d = x[it].astype(np.float64) - r.astype(np.float64)
        rm = float(np.sqrt(np.mean(d[inside] ** 2)))
        out[f"iteration_{it}"] = {"rmse": rm, "rmse_rel": rm / float(np.mean(np.abs(r[inside]))),
                                  "max_abs": float(np.max(np.abs(d)))}
    return out


def learned_form(ctx, hist, x0, ref, B, p, ms_fp64, args, timed, log, oracle_imgs=None):
    """The learned-path form on this scan: training set (10 M exact-MLP paths
    of 8 random-ellipsoid phantom scans of other grids), training with the
    library's kernels, calibration check on 1 in 64 histories (cells whose
    99th-percentile path deviation exceeds 3 % of a voxel take the exact
    walk), the agreement with the exact walk over every history, timed DROP
    iterations (surrogate MLP recomputed every iteration + FP32/Kahan update)
    and the RSP error against the exact FP64 form after iterations 1, 6, 12."""
    import torch
    import pct_amd
    from pct_amd import configs, learned
    res = {}
    t0 = time.time()
    scans = learned.training_scans(configs.LEARNED_TRAIN_SCANS, configs.LEARNED_TRAIN_HISTORIES_PER_SCAN)
    F, Y = learned.training_set(configs.LEARNED_TRAIN_PATHS, scans)
    torch.cuda.synchronize()
    t_data = time.time() - t0
    t0 = time.time()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(ctx.stream)
    net = learned.train_network(ctx, F, Y, epochs=20, batch=8192, lr=3e-3, seed=3)
    e1.record(ctx.stream)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    steps = 20 * (F.shape[0] // 8192)
    # MACs of one sample through the padded network (16x64 + 64x64 + 64x16),
    # forward + the two backward products per layer
    flops = 3 * 2 * (16 * 64 + 64 * 64 + 64 * 16) * 8192.0 * steps
    train_ms = e0.elapsed_time(e1)
    res["training"] = {"paths": int(F.shape[0]), "scans": [sc.name for sc in scans], "epochs": 20, "batch": 8192,
                       "optimizer_steps": steps, "data_s": round(t_data, 2), "train_s": round(t_train, 2),
                       "loss_first_epoch": net.epoch_loss[0], "loss_last_epoch": net.epoch_loss[-1],
                       "mfma": {"achieved_tflops": round(flops / (train_ms / 1e3) / 1e12, 3), "peak_tflops": 2500.0,
                                "frac": round(flops / (train_ms / 1e3) / 1e12 / 2500.0, 6),
                                "note": "two launches per optimizer step of 8192 samples (128 workgroups): "
                                        "latency-bound, not MFMA-bound"}}
    del F, Y
    log(f"learned: training set {t_data:.1f}s, training {t_train:.1f}s, loss {net.epoch_loss[0]:.3e} -> "
        f"{net.epoch_loss[-1]:.3e}")
    res.update(evaluate_learned(ctx, hist, x0, ref, B, p, net, timed, args.steps, log, oracle_imgs))
    res["speedup_vs_fp64_form"] = round(ms_fp64 / res["ms_per_iteration"], 3)
    res["lists_rule"]["speedup_vs_fp64_form"] = round(ms_fp64 / res["lists_rule"]["ms_per_iteration"], 3)
    if oracle_imgs and "error_vs_cpu_oracle" in res:
        vs_oracle = res.pop("error_vs_cpu_oracle")
        res["error_vs_cpu_oracle"] = vs_oracle
    return res, net


def cfg1_pipeline(log, threads):
    """The plain reference pipeline of the 100x100x16 / 2 M-history
    configuration as the reference runs it on the CPU -- the oracle's
    preprocessing (intersections, binning, 3-sigma cuts, FBP, MSC hull,
    endpoints) and 12 DROP iterations (MLP + DROP_blocks/DROP_update) on
    `threads` host cores -- timed, with the same pipeline on the GPU beside it
    and the two images compared."""
    import torch
    import oracle
    import pct_amd
    from pct_amd import configs, pipeline
    p = configs.params("cfg1")
    cfg = configs.synth_cfg("cfg1")
    iters = p.iterations
    t0 = time.perf_counter()
    raw = pct_amd.generate_host(cfg).numpy()
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    h, x0, _, _ = oracle.preprocess(p, raw)
    t_pre = time.perf_counter() - t0
    del raw
    t0 = time.perf_counter()
    x_cpu = oracle.drop(p, h, x0, iterations=iters, threads=threads)
    t_drop = time.perf_counter() - t0
    log(f"cfg1 on the CPU ({threads} threads): preprocess {t_pre:.1f} s, {iters} iterations {t_drop:.1f} s")
    ctx = pct_amd.Context(p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pre = pipeline.preprocess(ctx, ctx.generate(cfg))
    torch.cuda.synchronize()
    g_pre = time.perf_counter() - t0
    B = p.drop_block_size
    x = pre.x0.clone()
    t0 = time.perf_counter()
    ctx.drop_prepare(pre.hist, B)
    ctx.sync()
    g_plan = time.perf_counter() - t0
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(ctx.stream)
    for _ in range(iters):
        ctx.drop_iteration(pre.hist, x, B, p.lambda_)
    e1.record(ctx.stream)
    torch.cuda.synchronize()
    g_ms = e0.elapsed_time(e1)
    xg = x.cpu().numpy()
    inside = x0 != 0
    d = xg.astype(np.float64) - x_cpu.astype(np.float64)
    res = {"workload": f"cfg1: {p.columns}x{p.rows}x{p.slices} voxels, {cfg.n_histories} histories, "
                       f"{iters} DROP iterations, block {B}, lambda {p.lambda_:g}",
           "reconstruction_histories": int(h.n), "gpu_reconstruction_histories": int(pre.hist.n),
           "cpu": {"threads": threads, "host_nproc": os.cpu_count(), "cpu_model": _cpu_model(),
                   "generate_s": round(t_gen, 2), "preprocess_s": round(t_pre, 2), "drop_s": round(t_drop, 2),
                   "histories_per_s_per_iteration": round(h.n * iters / t_drop, 1)},
           "gpu": {"preprocess_s": round(g_pre, 3), "plan_s": round(g_plan, 3),
                   "drop_ms_per_iteration": round(g_ms / iters, 3),
                   "histories_per_s_per_iteration": round(pre.hist.n * iters / (g_ms / 1e3), 1)},
           "image_gpu_vs_cpu": {"rmse": float(np.sqrt(np.mean(d[inside] ** 2))),
                                "rmse_rel": float(np.sqrt(np.mean(d[inside] ** 2)) / np.mean(np.abs(x_cpu[inside]))),
                                "max_abs": float(np.max(np.abs(d))),
                                "identical_voxels": float(np.mean(xg == x_cpu))}}
    log(f"cfg1: GPU {g_ms / iters:.2f} ms/iteration; image vs CPU {res['image_gpu_vs_cpu']}")
    ctx.close()
    return res


def evaluate_learned(ctx, hist, x0, ref, B, p, net, timed, steps, log, oracle_imgs=None):
    """The installed network under both fallback rules of the calibration check
    (1 history in 64): "deviation" -- a cell whose 99th-percentile path
    deviation from the exact MLP exceeds 3 % of a voxel width takes the exact
    walk; "lists" (the north star's rule) -- a cell in which any sampled
    history's voxel list differs from the exact walk's does. Per rule: the
    agreement with the exact walk over every history, timed iterations and
    the RSP error against the exact FP64 form."""
    import pct_amd
    from pct_amd import learned
    res = {}
    vw = 2 * p.recon_cyl_radius / p.columns
    tol = 0.03 * vw
    for rule in ("deviation", "lists"):
        r = {}
        net.install(ctx)
        if rule == "deviation":
            cal = learned.calibrate(ctx, hist, stride=64, tol=tol)
            r["calibration"] = {"rule": f"cell p99 path deviation > {tol:.4g} cm", "samples": cal["samples"]}
        else:
            cal = learned.calibrate(ctx, hist, stride=64, rule="lists", threshold=0.0)
            r["calibration"] = {"rule": "any sampled voxel list of the cell differs from the exact walk's",
                                "samples": cal["samples"],
                                "cell_list_disagreement": [round(v, 4) for v, c in
                                                           zip(cal["cell_list_mismatch"], cal["cell_samples"]) if c]}
        r["calibration"].update({"cells_fallback": cal["cells_fallback"],
                                 "fallback_rate": round(cal["fallback_rate"], 5)})
        full = learned.calibrate(ctx, hist, stride=1)
        r["vs_exact_walk_all_histories"] = {
            "histories": full["samples"], "in_trained_range": full["in_range"],
            "surrogate_histories": full["surrogate"], "fallback_rate": round(full["fallback_rate"], 5),
            "list_mismatch_of_surrogate_histories": round(full["surrogate_mismatch"] / max(full["surrogate"], 1), 4),
            "list_mismatch_of_all_histories": round(full["surrogate_mismatch"] / max(full["samples"], 1), 4),
            "path_deviation_um": {"median": round(full["dev_median_cm"] * 1e4, 3),
                                  "p99": round(full["dev_p99_cm"] * 1e4, 3), "max": round(full["dev_max_cm"] * 1e4, 3)}}
        log(f"learned ({rule} rule): {r['vs_exact_walk_all_histories']}")
        t0 = time.time()
        ctx.drop_prepare(hist, B, pct_amd.PCT_MLP_SURROGATE)
        ctx.sync()
        r["plan_setup_once_ms"] = round((time.time() - t0) * 1e3, 1)
        xs = x0.clone()
        ctx.drop_iteration(hist, xs, B, p.lambda_, pct_amd.PCT_LEARNED)
        learned.fallback_stats(ctx)
        ms, ph = timed(pct_amd.PCT_LEARNED, steps, xs)
        ms /= steps
        fb, tot = learned.fallback_stats(ctx)
        _, mlp_ms, blk_ms = ph.mean(axis=0)
        pct_amd.set_tuning(ctx, overlap=2)
        _, ph_seq = timed(pct_amd.PCT_LEARNED, 1, xs)
        pct_amd.set_tuning(ctx, overlap=0)
        n = hist.n
        r.update({"histories_per_s": round(n / (ms / 1e3), 1), "ms_per_iteration": round(ms, 3),
                  "phases_ms": {"mlp_span": round(float(mlp_ms), 3), "block_loop_span": round(float(blk_ms), 3),
                                "surrogate_mlp_alone": round(float(ph_seq[0][0]), 3)},
                  "fallback_rate_timed_iterations": round(fb / max(tot, 1), 5)})
        if ref:
            xl = x0.clone()
            snaps = {}
            for it in range(1, max(ref) + 1):
                ctx.drop_iteration(hist, xl, B, p.lambda_, pct_amd.PCT_LEARNED)
                if it in ref:
                    snaps[it] = xl.cpu().numpy()
            r["error_vs_fp64"] = _rel_errors(snaps, ref, x0.cpu().numpy())
            if oracle_imgs:
                r["error_vs_cpu_oracle"] = _rel_errors(snaps, oracle_imgs, x0.cpu().numpy())
        log(f"learned form ({rule} rule): {ms:.2f} ms/iteration, fallback {fb / max(tot, 1):.4f}, "
            f"errors {r.get('error_vs_fp64')}")
        if rule == "deviation":
            res.update(r)        # the form as timed by default
        else:
            res["lists_rule"] = r
    res["error_region"] = "voxels of the hull (x0 != 0); reference image: exact FP64 form"
    res["rules"] = ("top level: the deviation rule; lists_rule: the same network with the list rule "
                    "(fall back wherever a sampled voxel list disagrees with the exact walk)")
    return res


def learned_large(net, steps, log):
    """The learned-path form on the 256x256x48 / 50 M-history scan (random
    ellipsoid phantom): the exact FP64 form gives the reference images after
    iterations 1, 6 and 12, then the same network as on the 200x200x32 scan."""
    import torch
    import pct_amd
    from pct_amd import configs, pipeline
    sc = configs.learned_large()
    ctx = pct_amd.Context(sc.params)
    t0 = time.time()
    pre = pipeline.preprocess(ctx, ctx.generate(sc.synth))
    hist = pre.hist
    B = sc.params.drop_block_size
    ctx.drop_prepare(hist, B)
    ctx.sync()
    setup_s = time.time() - t0
    pct_amd.set_profiling(ctx, True)

    def timed(mode, k, img):
        phases = []
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(ctx.stream)
        for _ in range(k):
            ctx.drop_iteration(hist, img, B, sc.params.lambda_, mode)
            phases.append(pct_amd.last_timings(ctx))
        ev1.record(ctx.stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1), np.array(phases)

    ref = {}
    xe = pre.x0.clone()
    t_ref = 0.0
    for it in range(1, 13):
        ms, _ = timed(0, 1, xe)
        t_ref += ms
        if it in (1, 6, 12):
            ref[it] = xe.cpu().numpy()
    res = {"scan": sc.name, "histories": int(sc.synth.n_histories), "reconstruction_histories": int(hist.n),
           "setup_s": round(setup_s, 1), "fp64_form_ms_per_iteration": round(t_ref / 12, 3),
           "fp64_form_histories_per_s": round(hist.n / (t_ref / 12 / 1e3), 1)}
    res.update(evaluate_learned(ctx, hist, pre.x0, ref, B, sc.params, net, timed, steps, log))
    res["speedup_vs_fp64_form"] = round(res["fp64_form_ms_per_iteration"] / res["ms_per_iteration"], 3)
    ctx.close()
    return res


def research_forms(ctx, hist, x0, B, p, steps, timed, log, m_lists=2_000_000):
    """The research update rules on this scan: the robust variants on the
    block loop (FP64, MLP lists recomputed as in the default form), and the
    research loop over given lists (pct_drop_lists_iteration, the live
    DROP_blocks_error_red rule with random path errors) on the MLP lists of
    the first m_lists reconstruction histories; per call the loop also sorts
    the lists' entries by (block, voxel)."""
    import torch
    import pct_amd
    res = {}
    for name, flag in (("robust_drop_blocks_red", pct_amd.PCT_DROP_ROBUST),
                       ("robust2_drop_blocks_robust2", pct_amd.PCT_DROP_ROBUST2)):
        xv = x0.clone()
        ctx.drop_iteration(hist, xv, B, p.lambda_, pct_amd.PCT_MLP_FP64 | flag)
        ms, _ = timed(pct_amd.PCT_MLP_FP64 | flag, steps, xv)
        ms /= steps
        res[name] = {"ms_per_iteration": round(ms, 3), "histories_per_s": round(hist.n / (ms / 1e3), 1),
                     "eta": p.eta, "psi_sign": p.psi_sign}
    m = min(m_lists, hist.n)
    sub = hist.head(m)
    offs, vox, lens = ctx.mlp_paths(sub)
    lens = lens.to(torch.int64)
    E = int(lens.sum())
    starts = torch.repeat_interleave(offs[:-1], lens)
    first = torch.cumsum(lens, 0) - lens
    idx = starts + torch.arange(E, device=lens.device) - torch.repeat_interleave(first, lens)
    voxels = vox[idx].contiguous()
    del starts, idx, vox
    offsets = torch.zeros(m + 1, dtype=torch.int64, device=lens.device)
    offsets[1:] = torch.cumsum(lens, 0)
    g = torch.Generator(device=lens.device)
    g.manual_seed(7)
    errors = (0.5 * torch.randn(E, generator=g, device=lens.device)).to(torch.float32)
    chord = ctx.chord_lengths(sub, float_average=True)
    xv = x0.clone()
    xu = torch.zeros(ctx.V, dtype=torch.float64, device=lens.device)
    cnt = torch.zeros(ctx.V, dtype=torch.int32, device=lens.device)
    rule = pct_amd.PCT_LISTS_ERROR_RED

    def timed_lists(flags, k):
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(ctx.stream)
        for _ in range(k):
            ctx.drop_lists_iteration(offsets, voxels, errors, chord, sub.wepl, xv, xu, cnt, B, p.lambda_, rule | flags)
        ev1.record(ctx.stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / k

    timed_lists(0, 1)                                   # warm-up
    ms_sort = timed_lists(0, 1)                         # a call on new lists: their sort included
    ms = timed_lists(pct_amd.PCT_LISTS_REUSE, steps)    # later iterations over the same lists
    res["research_lists_drop_blocks_error_red"] = {
        "ms_per_iteration": round(ms, 3), "ms_first_iteration_with_sort": round(ms_sort, 3), "lists": m,
        "intersections": E, "lists_per_s": round(m / (ms / 1e3), 1),
        "note": "one wavefront per list forms the terms and one lane adds them in list order, and the entries "
                "are sorted by (block, voxel) once per list set, so the result equals the reference loop bit for "
                "bit; not the production path"}
    log("research forms: " + ", ".join(f"{k} {v['ms_per_iteration']:.2f} ms" for k, v in res.items()))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--histories", type=int, default=None, help="override the scan size")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU time of the oracle sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched 12-scan form")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-path form")
    ap.add_argument("--no-large", action="store_true", help="skip the 256x256x48 / 50 M learned-form case")
    ap.add_argument("--no-research", action="store_true", help="skip the research update rules and the SM hull")
    ap.add_argument("--quick", action="store_true",
                    help="only the timed iterations (for counter runs, where every dispatch is slow)")
    ap.add_argument("--form", default="default",
                    choices=["default", "cached", "sequential", "precision", "learned", "walks"],
                    help="with --quick: the form of the timed iterations (default: MLP beside the block loop; "
                         "cached: the block loop alone; sequential: MLP then block loop; precision; learned: "
                         "network trained first on 10 M exact-MLP paths; walks: no iteration, the exact, "
                         "precision and surrogate MLP walks once over all histories, one launch each, the "
                         "surrogate with an untrained network over the whole input range)")
    ap.add_argument("--block-form", type=int, default=0, choices=[0, 1, 2, 3],
                    help="block loop: 1 one launch per block recomputing the previous block's updates, 2 two "
                         "launches per block, 3 one launch per block waiting for the previous block's stored "
                         "updates (0: the library's default)")
    ap.add_argument("--mlp-waves", type=int, default=0,
                    help="resident MLP wavefronts per SIMD beside the block loop (0: the library's default)")
    ap.add_argument("--no-cfg1", action="store_true", help="skip the 100x100x16 / 2 M CPU-and-GPU pipeline")
    ap.add_argument("--oracle-images", default=None,
                    help="directory of the CPU oracle's images of this scan after iterations 1, 6, 12 "
                         "(python -m oracle.full_image): every form's RSP error against them")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="5 device copies of 256 MiB and nothing else: the reference kernel for the HBM counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the multi-GPU sharding of the DROP blocks is designed but "
                                   "not built", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import pct_amd
    from pct_amd import configs

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}))
        sys.exit(1)
    if args.copy_calibration:
        n = 64 << 20
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 5}), flush=True)
        return
    dev = torch.device("cuda:0")
    pre = configs.PRESETS[args.config]
    p = configs.params(args.config)
    scfg = configs.synth_cfg(args.config, n_histories=args.histories)
    ctx = pct_amd.Context(p, dev)

    # ---- setup (untimed): the reference's main() up to the DROP loop ---------
    # intersections, binning, 3-sigma cuts, FBP, MSC hull, x0 = FBP * hull,
    # MLP endpoints -- every stage on the GPU (pct_amd.pipeline)
    from pct_amd import pipeline
    t0 = time.time()
    raw = ctx.generate(scfg)
    torch.cuda.synchronize()
    t_pre = time.time()
    pre = pipeline.preprocess(ctx, raw)
    torch.cuda.synchronize()
    t_pre = time.time() - t_pre
    # again, each stage alone between two events: time, algorithmic bytes and
    # HBM GB/s per stage (the same results; this run's are the ones used)
    del pre
    torch.cuda.empty_cache()
    timings = []
    pre = pipeline.preprocess(ctx, raw, timings=timings)
    def stage_entry(ms, nbytes, traffic):
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"ms": round(ms, 4), "bytes": int(nbytes), "gb_s": round(gbs, 1),
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": int(traffic),
                "traffic_gb_s": round(traffic / (ms * 1e-3) / 1e9, 1)}

    stages = {name: stage_entry(ms, nbytes, traffic) for name, ms, nbytes, traffic in timings}
    log("preprocessing stages: " + ", ".join(f"{k} {v['ms']:.3f} ms {v['gb_s']:.0f} GB/s" for k, v in stages.items()))
    sm_stages = None
    if not args.no_research:
        # the space-modelling hull in place of MSC (same scan, stages alone)
        t_sm = []
        pre_sm = pipeline.preprocess(ctx, raw, hull_kind="sm", timings=t_sm)
        sm_stages = {}
        for name, ms, nbytes, traffic in t_sm:
            if name in ("sm_counts", "sm_edge", "hull_select"):
                sm_stages[name] = stage_entry(ms, nbytes, traffic)
        sm_stages["hull_voxels"] = int(np.unpackbits(pre_sm.hull.cpu().numpy().view(np.uint8)).sum())
        del pre_sm
        log("SM hull stages: " + ", ".join(f"{k} {v['ms']:.3f} ms" for k, v in sm_stages.items() if isinstance(v, dict)))
    del raw
    hist = pre.hist
    n_r = hist.n
    x = pre.x0.clone()
    x0_np = pre.x0.cpu().numpy()
    B = p.drop_block_size
    if args.block_form:
        pct_amd.set_tuning(ctx, block_form=args.block_form)
    if args.mlp_waves:
        pct_amd.set_tuning(ctx, mlp_waves=args.mlp_waves)
    if args.quick and args.form == "walks":
        from pct_amd import learned
        learned.install(ctx, learned.init_params(1, dev), [-float("inf")] * 12, [float("inf")] * 12)
        res = {}
        for name, mode in (("exact", pct_amd.PCT_MLP_FP64), ("precision", pct_amd.PCT_MLP_TABLE),
                           ("surrogate", pct_amd.PCT_MLP_SURROGATE)):
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(ctx.stream)
            offs, vox, lens = ctx.mlp_paths(hist, mode=mode)
            e1.record(ctx.stream)
            torch.cuda.synchronize()
            res[name] = {"ms": round(e0.elapsed_time(e1), 3), "voxels": int(lens.to(torch.int64).sum())}
            del offs, vox, lens
        print(json.dumps({"quick": True, "form": "walks", "histories": n_r, "walks": res}), flush=True)
        return
    t1 = time.time()
    ctx.drop_prepare(hist, B)
    ctx.sync()
    t_plan = time.time() - t1
    sum_n, sum_vb = pct_amd.drop_stats(ctx, hist, B)
    _, total_slots = ctx.mlp_capacity(hist)
    sum_steps = total_slots - n_r
    log(f"setup {time.time() - t0:.1f}s (preprocess {t_pre:.2f}s, plan {t_plan:.2f}s): N={scfg.n_histories} N_r={n_r} "
        f"mean n={sum_n / n_r:.1f} sum|V_B|={sum_vb} steps={sum_steps}")

    def timed(mode, steps, img=None):
        img = x if img is None else img
        phases = []
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(ctx.stream)
        for _ in range(steps):
            ctx.drop_iteration(hist, img, B, p.lambda_, mode)
            phases.append(pct_amd.last_timings(ctx))
        ev1.record(ctx.stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1), np.array(phases)

    # ---- timed DROP iterations (default: MLP recomputed every iteration) ----
    pct_amd.set_profiling(ctx, True)
    form_mode = 0
    if args.quick and args.form == "cached":
        form_mode = pct_amd.PCT_DROP_CACHED
    elif args.quick and args.form == "sequential":
        pct_amd.set_tuning(ctx, overlap=2)
    elif args.quick and args.form == "precision":
        form_mode = pct_amd.PCT_PRECISION
        ctx.drop_prepare(hist, B, pct_amd.PCT_MLP_TABLE)
    elif args.quick and args.form == "learned":
        from pct_amd import learned
        scans = learned.training_scans(configs.LEARNED_TRAIN_SCANS, configs.LEARNED_TRAIN_HISTORIES_PER_SCAN)
        F, Y = learned.training_set(configs.LEARNED_TRAIN_PATHS, scans)
        net = learned.train_network(ctx, F, Y, epochs=20, batch=8192, lr=3e-3, seed=3)
        del F, Y
        net.install(ctx)
        learned.calibrate(ctx, hist, stride=64, tol=0.03 * 2 * p.recon_cyl_radius / p.columns)
        form_mode = pct_amd.PCT_LEARNED
        ctx.drop_prepare(hist, B, pct_amd.PCT_MLP_SURROGATE)
    for _ in range(args.warmup):
        ctx.drop_iteration(hist, x, B, p.lambda_, form_mode)
    ctx.sync()
    t_start = time.perf_counter()
    gpu_ms, ph = timed(form_mode, args.steps)
    wall = time.perf_counter() - t_start
    ms_per_step = gpu_ms / args.steps
    value = n_r * args.steps / (gpu_ms / 1e3)
    _, mlp_ms, blk_ms = ph.mean(axis=0)
    log(f"timed {args.steps} iterations: {ms_per_step:.2f} ms/iter (wall {wall / args.steps * 1e3:.2f} ms); "
        f"MLP {mlp_ms:.2f} ms, blocks {blk_ms:.2f} ms")
    if args.quick:
        print(json.dumps({"metric": "histories/s per DROP iteration (MLP+update), RSP RMSE vs FP64 oracle, 1 GPU",
                          "value": round(value, 1), "unit": "histories/s", "ms_per_step": round(ms_per_step, 3),
                          "steps": args.steps, "quick": True, "form": args.form,
                          "phases_ms": [round(float(v), 3) for v in ph.mean(axis=0)]}), flush=True)
        return
    # each phase alone, for its roofline: the block loop without the MLP pass
    # (cached lists: k_block_delta + k_block_apply only), and the MLP pass
    # followed by the block loop on one stream (no overlap)
    ms_cached, _ = timed(pct_amd.PCT_DROP_CACHED, 1)
    pct_amd.set_tuning(ctx, overlap=2)
    ms_seq, ph_seq = timed(0, 1)
    pct_amd.set_tuning(ctx, overlap=0)
    mlp_alone_ms = float(ph_seq[0][1])
    blk_alone_ms = ms_cached
    # the reference's structure (atomic scatter + volume scan per block), for comparison
    ms_atomic, _ = timed(pct_amd.PCT_DROP_ATOMIC, 1)
    log(f"block loop alone {ms_cached:.2f} ms, MLP alone {mlp_alone_ms:.2f} ms, sequential iteration {ms_seq:.2f} ms, "
        f"reference-structure (atomic) iteration {ms_atomic:.2f} ms")

    oracle_imgs = None
    if args.oracle_images:
        oracle_imgs = {it: np.load(os.path.join(args.oracle_images, f"x{it}.npy")) for it in (1, 6, 12)}
        assert np.array_equal(np.load(os.path.join(args.oracle_images, "x0.npy")), x0_np), "oracle x0 differs"
    vs_oracle = {}

    # ---- precision form: tabulated MLP + FP32 stepping + FP32/Kahan update ----
    # RSP error vs the exact FP64 form after iterations 1, 6 and 12 from x0 (the
    # exact form equals the CPU oracle bit for bit, see rmse_vs_oracle below)
    snaps = (1, 6, 12)
    ref = {}
    xe = pre.x0.clone()
    for it in range(1, snaps[-1] + 1):
        ctx.drop_iteration(hist, xe, B, p.lambda_)
        if it in snaps:
            ref[it] = xe.cpu().numpy()
    t1 = time.time()
    ctx.drop_prepare(hist, B, pct_amd.PCT_MLP_TABLE)
    ctx.sync()
    t_plan_p = time.time() - t1
    xs = pre.x0.clone()
    for _ in range(args.warmup):
        ctx.drop_iteration(hist, xs, B, p.lambda_, pct_amd.PCT_PRECISION)
    ms_prec, ph_p = timed(pct_amd.PCT_PRECISION, args.steps, xs)
    ms_prec /= args.steps
    _, mlp_ms_p, blk_ms_p = ph_p.mean(axis=0)
    in_hull = x0_np != 0
    if oracle_imgs:
        vs_oracle["fp64"] = _rel_errors(ref, oracle_imgs, x0_np)
    prec_err = {}
    prec_snaps = {}
    xp = pre.x0.clone()
    for it in range(1, snaps[-1] + 1):
        ctx.drop_iteration(hist, xp, B, p.lambda_, pct_amd.PCT_PRECISION)
        if it in snaps:
            prec_snaps[it] = xp.cpu().numpy()
            d = xp.cpu().numpy().astype(np.float64) - ref[it].astype(np.float64)
            r = float(np.sqrt(np.mean(d[in_hull] ** 2)))
            prec_err[f"iteration_{it}"] = {"rmse": r, "rmse_rel": r / float(np.mean(np.abs(ref[it][in_hull]))),
                                           "max_abs": float(np.max(np.abs(d)))}
    if oracle_imgs:
        vs_oracle["precision"] = _rel_errors(prec_snaps, oracle_imgs, x0_np)
    log(f"precision form: {ms_prec:.2f} ms/iter (MLP {mlp_ms_p:.2f} ms, blocks {blk_ms_p:.2f} ms), error {prec_err}")
    precision = {"histories_per_s": round(n_r / (ms_prec / 1e3), 1), "ms_per_iteration": round(ms_prec, 3),
                 "phases_ms": {"mlp": round(float(mlp_ms_p), 3), "blocks": round(float(blk_ms_p), 3),
                               "plan_setup_once": round(t_plan_p * 1e3, 1)},
                 "speedup_vs_fp64_form": round(ms_per_step / ms_prec, 3),
                 "error_vs_fp64": prec_err,
                 "error_region": "voxels of the hull (x0 != 0); reference image: exact FP64 form"}

    # ---- research update rules (robust variants, list loop) ----
    research = None
    if not args.no_research:
        try:
            ctx.drop_prepare(hist, B)
            research = research_forms(ctx, hist, pre.x0, B, p, args.steps, timed, log)
            research["sm_hull_stages"] = sm_stages
        except Exception as e:
            research = {"error": f"{type(e).__name__}: {e}"}
            log(f"research forms failed: {e}")

    # ---- learned-path form: network trained here on 10 M exact-MLP paths ----
    learned_res, net = None, None
    if not args.no_learned:
        try:
            learned_res, net = learned_form(ctx, hist, pre.x0, ref, B, p, ms_per_step, args, timed, log, oracle_imgs)
            if oracle_imgs and "error_vs_cpu_oracle" in learned_res:
                vs_oracle["learned"] = learned_res["error_vs_cpu_oracle"]
        except Exception as e:
            learned_res = {"error": f"{type(e).__name__}: {e}"}
            log(f"learned form failed: {e}")

    # algorithmic work per iteration: 48 B per history record, per path entry
    # the 4 B list index streamed, a 4 B x gather and a 4 B scatter, 24 B per
    # voxel touched by a block
    bytes_update = 48.0 * n_r + 12.0 * sum_n + 24.0 * sum_vb
    nb = (n_r + B - 1) // B
    upd_gbs = bytes_update / (blk_alone_ms / 1e3) / 1e9
    mlp_tflops = MLP_FLOPS_PER_STEP * float(sum_steps) / (mlp_alone_ms / 1e3) / 1e12
    roof_update = {"bound": "hbm", "achieved": round(upd_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(upd_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                   "kernel": "k_block_delta+k_block_apply, one pair per DROP block (hipGraph replay)",
                   "us_per_block": round(blk_alone_ms * 1e3 / nb, 3), "bytes_alg_per_block": round(bytes_update / nb),
                   "ms_per_iteration": round(float(blk_alone_ms), 3), "bytes_alg": bytes_update}
    roof_mlp = {"bound": "fp64-valu", "achieved": round(mlp_tflops, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(mlp_tflops / FP64_PEAK_TFLOPS, 4), "traffic": None,
                "kernel": "k_mlp_paths<exact_walk> (FP64 VALU; peak = MI355X FP64 78.6 TFLOP/s)",
                "ms_per_iteration": round(mlp_alone_ms, 3), "flops_alg": MLP_FLOPS_PER_STEP * float(sum_steps)}
    # measured HBM traffic of the block pair (FETCH_SIZE + WRITE_SIZE, scaled by
    # a plain copy) from the committed counter profile of this geometry
    tpath = os.path.join(ROOT, "profiles", f"hbm_traffic_{args.config}.json")
    if os.path.exists(tpath):
        tr = json.load(open(tpath))["kernels"]
        per_block = sum(tr[k]["fetch_bytes_per_dispatch"] + tr[k]["write_bytes_per_dispatch"]
                        for k in ("k_block_delta", "k_block_apply"))
        roof_update["traffic"] = per_block
        roof_update["traffic_source"] = os.path.relpath(tpath, ROOT) + " (bytes per block: delta + apply)"
    dominant = roof_mlp if mlp_alone_ms >= blk_alone_ms else roof_update

    # ---- CPU oracle on a bounded sample, and RMSE of the GPU vs the oracle ----
    cpu = None
    rmse = None
    if not args.no_cpu_baseline:
        import oracle
        hh = hist.numpy()
        # calibrate on one block, then size the sample to ~cpu_seconds
        m0 = min(B, n_r)
        hs = oracle.HostHistories(arrays={k: v[:m0] for k, v in hh.items()})
        t = time.perf_counter()
        oracle.drop(p, hs, x0_np, iterations=1)
        per_hist = (time.perf_counter() - t) / m0
        m = int(min(n_r, max(m0, args.cpu_seconds / per_hist)))
        m = max(B, (m // B) * B) if m >= B else m
        hs = oracle.HostHistories(arrays={k: v[:m] for k, v in hh.items()})
        t = time.perf_counter()
        x_cpu = oracle.drop(p, hs, x0_np, iterations=1)
        cpu_t = time.perf_counter() - t
        cpu = {"value": round(m / cpu_t, 1), "unit": "histories/s per DROP iteration", "cores": 1, "kind": "port",
               "sample": f"first {m} reconstruction histories ({m // B} DROP blocks of {B}), one iteration from the "
                         f"same x0, FP64 oracle (MLP + DROP_blocks/DROP_update), single thread, {cpu_t:.1f} s",
               "cpu_model": _cpu_model()}
        # the same sample with the per-block MLP work spread over the host
        # cores this GPU's share of the host allows (16 per GPU on the MI355X
        # boxes, whatever nproc shows)
        threads = min(16, os.cpu_count() or 1)
        t = time.perf_counter()
        x_mt = oracle.drop(p, hs, x0_np, iterations=1, threads=threads)
        mt_t = time.perf_counter() - t
        cpu["multithreaded"] = {"value": round(m / mt_t, 1), "cores": threads, "host_nproc": os.cpu_count(),
                                "seconds": round(mt_t, 2),
                                "identical_to_single_thread": bool(np.array_equal(x_mt, x_cpu))}
        # the GPU on the same sample from the same x0
        x_g = torch.from_numpy(x0_np).to(dev)
        ctx.drop_iteration(hist.head(m), x_g, B, p.lambda_)
        diff = x_g.cpu().numpy().astype(np.float64) - x_cpu.astype(np.float64)
        touched = x_cpu != x0_np
        rmse = {"rmse": float(np.sqrt(np.mean(diff[touched] ** 2))) if touched.any() else 0.0,
                "max_abs": float(np.max(np.abs(diff))), "voxels": int(touched.sum()),
                "sample": f"one iteration over the first {m} histories, GPU FP32 image vs FP64 oracle"}

    # ---- batched form (12 scans), after releasing this scan's plan ----
    batched = None
    post_cut = pre.counts["post_cut"]
    if not args.no_batched:
        ctx.close()
        del hist, pre, x
        torch.cuda.empty_cache()
        try:
            batched = batched_form(log)
        except Exception as e:        # e.g. device memory: reported, the headline stands
            batched = {"error": f"{type(e).__name__}: {e}"}
            log(f"batched form failed: {e}")

    if net is not None and not args.no_large and isinstance(learned_res, dict):
        ctx.close()                     # the 200x200x32 plan is released first
        torch.cuda.empty_cache()
        try:
            learned_res["large"] = learned_large(net, args.steps, log)
        except Exception as e:
            learned_res["large"] = {"error": f"{type(e).__name__}: {e}"}
            log(f"learned large case failed: {e}")

    cfg1 = None
    if not args.no_cfg1 and args.config == "cfg2":
        try:
            cfg1 = cfg1_pipeline(log, min(16, os.cpu_count() or 1))
        except Exception as e:
            cfg1 = {"error": f"{type(e).__name__}: {e}"}
            log(f"cfg1 pipeline failed: {e}")

    out = {
        "metric": "histories/s per DROP iteration (MLP+update), RSP RMSE vs FP64 oracle, 1 GPU",
        "value": round(value, 1),
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {p.columns}x{p.rows}x{p.slices} voxels, "
                               f"{scfg.n_histories} histories / 90 angles, DROP block {B}, lambda {p.lambda_:g}, "
                               f"exact FP64 MLP, FP32 image, x0 = FBP * MSC hull",
                   "histories": scfg.n_histories, "post_cut_histories": post_cut,
                   "reconstruction_histories": n_r,
                   "mean_path_voxels": round(sum_n / n_r, 2), "sum_block_voxels": sum_vb,
                   "drop_blocks": (n_r + B - 1) // B},
        "phases_ms": {"mlp_span": round(float(mlp_ms), 3), "block_loop_span": round(float(blk_ms), 3),
                      "note": "the MLP check runs beside the block loop and joins it at the end; alone they take "
                              f"{mlp_alone_ms:.2f} and {blk_alone_ms:.2f} ms",
                      "plan_setup_once": round(t_plan * 1e3, 1), "preprocess_once": round(t_pre * 1e3, 1)},
        "preprocess_stages": {"note": "each stage alone between two HIP events; bytes = the arrays the "
                                      "reference's kernels for the stage read and write, once at their stored "
                                      "width (pct_amd/pipeline.py STAGE_BYTES), gb_s and frac of "
                                      f"{HBM_PEAK_GBS:g} GB/s from them; traffic = the bytes this implementation's "
                                      "own arrays move (sort passes, records, cut windows; STAGE_TRAFFIC)",
                              **stages},
        "precision_form": precision,
        "batched_form": batched,
        "learned_form": learned_res,
        "research_forms": research,
        "other_forms": {"sequential_mlp_then_blocks_ms": round(ms_seq, 3),
                        "cached_lists_histories_per_s": round(n_r / (ms_cached / 1e3), 1),
                        "cached_lists_ms": round(ms_cached, 3),
                        "atomic_reference_structure_histories_per_s": round(n_r / (ms_atomic / 1e3), 1),
                        "atomic_reference_structure_ms": round(ms_atomic, 3)},
        "roofline": dominant,
        "roofline_update": roof_update,
        "roofline_mlp": roof_mlp,
        "rmse_vs_oracle": rmse,
        "rmse_vs_full_cpu_oracle": ({"source": args.oracle_images, **vs_oracle} if oracle_imgs else None),
        "cfg1_cpu_pipeline": cfg1,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
