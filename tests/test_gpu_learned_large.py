"""GPU: the learned-path form on its second evaluation case, 256x256x48 voxels
with 50 M histories (a random-ellipsoid phantom; ~42 M reconstruction
histories), where the oracle cannot walk every history in test time:

  - on samples spread over the scan, every surrogate-mode voxel list is the
    oracle's FP32 surrogate walk of the same network outputs or, for the
    histories the guard sends back, the oracle's exact FP64 walk; the
    fallback counter lies between the histories only the exact list explains
    and all but those only the surrogate list explains;
  - three DROP iterations of the learned form (surrogate MLP recomputed every
    iteration, FP32/Kahan update) stay within 0.2 % relative RMSE of the exact
    FP64 form's image, while moving the image far more than that.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_learned_form_at_256x256x48_with_50m_histories():
    import oracle
    import pct_amd
    from pct_amd import configs, learned, pipeline
    sc = configs.learned_large()
    ctx = pct_amd.Context(sc.params)
    torch = ctx.torch
    pre = pipeline.preprocess(ctx, ctx.generate(sc.synth))
    hist, p = pre.hist, sc.params
    print(f"{sc.name}: {hist.n} reconstruction histories", flush=True)
    assert hist.n > 35_000_000
    F, Y = learned.training_set(1_000_000, learned.training_scans(8, 200_000))
    net = learned.train_network(ctx, F, Y, epochs=20, batch=4096, lr=3e-3, seed=3)
    del F, Y
    net.install(ctx)
    cal = learned.calibrate(ctx, hist, stride=64, tol=0.03 * 2 * p.recon_cyl_radius / p.columns)
    print("calibration:", {k: cal[k] for k in ("samples", "fallback_rate", "dev_median_cm", "dev_p99_cm",
                                               "list_mismatch_rate", "cells_fallback")}, flush=True)
    # voxel lists on samples spread over the scan
    n_each, bad, only_exact, only_sur, total_fb = 400, 0, 0, 0, 0
    for s in np.linspace(0, hist.n - n_each, 5).astype(np.int64):
        part = hist.slice(int(s), n_each)
        offs, vox, lens = ctx.mlp_paths(part, 0, n_each, mode=pct_amd.PCT_MLP_SURROGATE)
        fb, tot = learned.fallback_stats(ctx)
        assert tot == n_each
        total_fb += fb
        Fp, _ = learned.features(ctx, part, 0, n_each, targets=False)
        out = learned.forward(ctx, Fp).cpu().numpy()
        offs, vox, lens = offs.cpu().numpy(), vox.cpu().numpy().view(np.uint32), lens.cpu().numpy()
        hh = oracle.HostHistories(arrays=part.numpy())
        exact = oracle.mlp_paths(p, hh, 0, n_each)
        for i in range(n_each):
            got = vox[offs[i]:offs[i] + lens[i]]
            sur, _, _ = oracle.surrogate_path(p, hh, i, out[i])
            is_sur, is_exact = np.array_equal(got, sur), np.array_equal(got, exact[i])
            bad += not (is_sur or is_exact)
            only_exact += is_exact and not is_sur
            only_sur += is_sur and not is_exact
        assert only_exact <= total_fb <= 5 * n_each - only_sur
    print(f"samples: {5 * n_each}, fallbacks {total_fb}, lists only the exact walk gives {only_exact}, "
          f"only the surrogate {only_sur}", flush=True)
    assert bad == 0, f"{bad} lists are neither the surrogate's nor the exact walk's"
    assert only_sur > 0
    # three iterations, learned form against the exact FP64 form
    B = p.drop_block_size
    ref, got = pre.x0.clone(), pre.x0.clone()
    for _ in range(3):
        ctx.drop_iteration(hist, ref, B, p.lambda_, pct_amd.PCT_MLP_FP64)
    for _ in range(3):
        ctx.drop_iteration(hist, got, B, p.lambda_, pct_amd.PCT_LEARNED)
    fb, tot = learned.fallback_stats(ctx)
    r, g, x0 = ref.cpu().numpy(), got.cpu().numpy(), pre.x0.cpu().numpy()
    inside = x0 != 0
    rel = np.sqrt(np.mean((g[inside] - r[inside]) ** 2)) / np.mean(np.abs(r[inside]))
    moved = np.sqrt(np.mean((r[inside] - x0[inside]) ** 2)) / np.mean(np.abs(r[inside]))
    print(f"learned vs FP64 after 3 iterations: relative RMSE {rel:.3e} (the iterations moved the image by "
          f"{moved:.3e}); fallbacks {fb} of {tot}", flush=True)
    assert tot >= 3 * hist.n and fb < 0.05 * tot
    assert rel < 2e-3 and moved > 10 * rel
    ctx.close()
