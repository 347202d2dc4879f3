"""GPU: the learned-path form through the C ABI (include/pct_net.h).

  - training data (inputs + exact-MLP knot targets) bit-identical to the oracle;
  - the training kernels' gradient (BF16 MFMA, FP32 accumulation) equals the
    oracle network's gradient with the same BF16 rounding, to FP32
    accumulation error; one Adam step equals the oracle's;
  - the inference kernel equals the oracle forward;
  - training on exact-MLP paths lowers the loss by orders of magnitude;
  - the surrogate walk's voxel lists are bit-identical to the oracle's FP32
    surrogate walk on the same network outputs; histories outside the trained
    range take the exact walk and give the exact lists; fallback counters;
  - ablation: the trained network against the method without it (the Hermite
    path alone) on path deviation, list agreement and calibration fallback;
  - a DROP iteration with the surrogate stays close to the FP64 image, and its
    recomputed lists match the plan in every iteration.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_TRAIN = 1_000_000


def _scan(preset, n, seed=None, device="cuda"):
    import pct_amd
    from pct_amd import configs, learned
    ctx = pct_amd.Context(configs.params(preset))
    hist = learned.scan_histories(ctx, configs.synth_cfg(preset, n_histories=n, seed=seed))
    return ctx, hist


def _host(hist):
    import oracle
    return oracle.HostHistories(arrays=hist.numpy())


@pytest.fixture(scope="module")
def cfg2():
    return _scan("cfg2", 30_000)


@pytest.fixture(scope="module")
def trained(cfg2):
    """a network trained on exact-MLP paths of random-ellipsoid phantoms (8
    scans of other grids; the evaluation scan is not among them)"""
    from pct_amd import learned
    ctx, _ = cfg2
    F, Y = learned.training_set(N_TRAIN, learned.training_scans(8, 200_000))
    net = learned.train_network(ctx, F, Y, epochs=30, batch=4096, lr=3e-3, seed=3)
    print("training loss per epoch:", net.epoch_loss)
    return net, F, Y


def test_features_bit_exact(cfg2):
    import oracle
    from pct_amd import configs, learned
    ctx, hist = cfg2
    F, Y = learned.features(ctx, hist)
    hh = _host(hist)
    Fo, Yo, _, _ = oracle.net_features(configs.params("cfg2"), hh)
    assert np.array_equal(F.cpu().numpy(), Fo)
    assert np.array_equal(Y.cpu().numpy(), Yo)


@pytest.mark.parametrize("n", [64, 1000])
def test_gradient_matches_oracle(cfg2, n):
    import torch
    from oracle import net as onet
    from pct_amd import learned
    ctx, hist = cfg2
    F, Y = learned.features(ctx, hist, 0, n)
    params = learned.init_params(5, ctx.device)
    params[768:832] = 0.05 * torch.arange(64, device=ctx.device, dtype=torch.float32).sin()
    g, loss = learned.gradient(ctx, params, F, Y)
    go, lo = onet.gradient(params.cpu().numpy(), F.cpu().numpy(), Y.cpu().numpy(), bf16=True)
    g = g.cpu().numpy().astype(np.float64)
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    assert rel < 2e-3, rel
    # every block of the parameter vector, not only the large ones
    for a, b in ((0, 768), (768, 832), (832, 4928), (4928, 4992), (4992, 5504), (5504, 5512)):
        r = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-30)
        assert r < 1e-2, (a, b, r)


def test_adam_step_matches_oracle(cfg2):
    from oracle import net as onet
    from pct_amd import learned
    ctx, hist = cfg2
    F, Y = learned.features(ctx, hist, 0, 4096)
    p0 = learned.init_params(9, ctx.device)
    g, _ = learned.gradient(ctx, p0, F, Y)
    p = p0.clone()
    learned.train(ctx, p, F, Y, epochs=1, batch=4096, lr=1e-3, lr_min=1e-3, seed=7)
    exp, _, _ = onet.adam_step(p0.cpu().numpy(), np.zeros(onet.P), np.zeros(onet.P), g.cpu().numpy(), 1, 1e-3)
    got = p.cpu().numpy()
    # the step is lr * g / (|g| + eps): equal wherever |g| is not near eps
    big = np.abs(g.cpu().numpy()) > 1e-6
    np.testing.assert_allclose(got[big], exp[big], rtol=0, atol=2e-6)


def test_forward_matches_oracle(cfg2, trained):
    from oracle import net as onet
    from pct_amd import learned
    ctx, hist = cfg2
    net, _, _ = trained
    net.install(ctx)
    F, _ = learned.features(ctx, hist, 0, 777)
    out = learned.forward(ctx, F).cpu().numpy()
    ref = onet.forward(net.params.cpu().numpy(), F.cpu().numpy(), bf16=True)["y"]
    assert np.abs(out - ref).max() <= 2e-3 * max(1.0, np.abs(ref).max())


def test_training_lowers_the_loss(trained):
    net, F, Y = trained
    assert net.epoch_loss[-1] < 0.01 * net.epoch_loss[0]
    # mean squared error of the knot offsets below (1.5 um)^2, in 0.01-cm units
    assert net.epoch_loss[-1] < 0.0225


def test_surrogate_lists_match_oracle_walk(cfg2, trained):
    import oracle
    import pct_amd
    from pct_amd import configs, learned
    ctx, hist = cfg2
    net, _, _ = trained
    learned.install(ctx, net.params, [-np.inf] * 12, [np.inf] * 12)
    n = 3000
    offs, vox, lens = ctx.mlp_paths(hist, 0, n, mode=pct_amd.PCT_MLP_SURROGATE)
    fb, tot = learned.fallback_stats(ctx)
    assert fb == 0 and tot == n
    F, _ = learned.features(ctx, hist, 0, n)
    out = learned.forward(ctx, F).cpu().numpy()
    offs, vox, lens = offs.cpu().numpy(), vox.cpu().numpy().view(np.uint32), lens.cpu().numpy()
    p, hh = configs.params("cfg2"), _host(hist)
    bad = 0
    for i in range(n):
        path, _, _ = oracle.surrogate_path(p, hh, i, out[i])
        bad += not np.array_equal(vox[offs[i]:offs[i] + lens[i]], path)
    assert bad == 0, f"{bad} of {n} surrogate lists differ from the oracle's walk"


def test_out_of_range_histories_take_the_exact_walk(cfg2, trained):
    import oracle
    import pct_amd
    from pct_amd import configs, learned
    ctx, hist = cfg2
    net, _, _ = trained
    n = 2000
    F, _ = learned.features(ctx, hist, 0, n)
    # a range that excludes histories with t_in < 0
    lo, hi = [-np.inf] * 12, [np.inf] * 12
    lo[6] = 0.0
    learned.install(ctx, net.params, lo, hi)
    offs, vox, lens = ctx.mlp_paths(hist, 0, n, mode=pct_amd.PCT_MLP_SURROGATE)
    fb, tot = learned.fallback_stats(ctx)
    outside = (F[:, 6] < 0).cpu().numpy()
    assert fb == int(outside.sum()) and tot == n and 0 < fb < n
    offs, vox, lens = offs.cpu().numpy(), vox.cpu().numpy().view(np.uint32), lens.cpu().numpy()
    exact = oracle.mlp_paths(configs.params("cfg2"), _host(hist), 0, n)
    for i in np.nonzero(outside)[0][:400]:
        assert np.array_equal(vox[offs[i]:offs[i] + lens[i]], exact[i])


def test_network_gains_over_hermite_path(cfg2, trained):
    """ablation: the same surrogate walk with and without the network"""
    from pct_amd import learned
    ctx, hist = cfg2
    net, F, _ = trained
    lo, hi = learned.input_range(F, 0.02)
    learned.install(ctx, net.params, lo, hi)
    with_net = learned.calibrate(ctx, hist, stride=4)
    learned.install_hermite(ctx, lo, hi)
    without = learned.calibrate(ctx, hist, stride=4)
    summary = {k: (with_net[k], without[k]) for k in ("dev_median_cm", "dev_p99_cm", "dev_max_cm",
                                                       "list_mismatch_rate")}
    print("network vs Hermite path:", summary)
    assert with_net["samples"] == without["samples"] > 5000
    # largest deviation of a path from the exact MLP: a few um instead of ~100 um
    assert with_net["dev_median_cm"] < 0.1 * without["dev_median_cm"]
    assert with_net["dev_p99_cm"] < 0.1 * without["dev_p99_cm"]
    assert with_net["dev_p99_cm"] < 3e-3
    # voxel lists equal to the exact walk's for most histories (the Hermite
    # path alone gets about one in ten right)
    assert with_net["list_mismatch_rate"] < 0.6 * without["list_mismatch_rate"]
    assert with_net["list_mismatch_rate"] < 0.45
    # calibration at 30 um: the network keeps its cells, the Hermite path loses them
    learned.install(ctx, net.params, lo, hi)
    cal_net = learned.calibrate(ctx, hist, stride=4, tol=3e-3, min_samples=8)
    learned.install_hermite(ctx, lo, hi)
    cal_h = learned.calibrate(ctx, hist, stride=4, tol=3e-3, min_samples=8)
    pop = np.array(cal_net["cell_samples"])
    kept_net = pop[np.array(cal_net["cell_fallback"]) == 0].sum() / pop.sum()
    kept_h = pop[np.array(cal_h["cell_fallback"]) == 0].sum() / pop.sum()
    print("share of histories in cells kept at 30 um: network", kept_net, "Hermite", kept_h)
    assert kept_net > 0.9 and kept_h < 0.2, (kept_net, kept_h)


def test_learned_drop_iterations(cfg2, trained):
    import pct_amd
    from pct_amd import learned
    ctx, hist = cfg2
    torch = ctx.torch
    net, F, _ = trained
    lo, hi = learned.input_range(F, 0.02)
    learned.install(ctx, net.params, lo, hi)
    learned.calibrate(ctx, hist, stride=8, tol=3e-3)
    x0 = ctx.initial_iterate(torch.ones(ctx.V, dtype=torch.float32, device=ctx.device), ctx.hull_init())
    ref, got = x0.clone(), x0.clone()
    for it in range(4):
        ctx.drop_iteration(hist, ref, 800, 1e-2)
    learned.fallback_stats(ctx)
    for it in range(4):
        # plan (one MLP pass) on the first call, then lists recomputed and
        # checked against the plan in every iteration
        ctx.drop_iteration(hist, got, 800, 1e-2, pct_amd.PCT_LEARNED)
    fb, tot = learned.fallback_stats(ctx)
    r, g, x = ref.cpu().numpy(), got.cpu().numpy(), x0.cpu().numpy()
    inside = x != 0
    rel = np.sqrt(np.mean((g[inside] - r[inside]) ** 2)) / np.mean(np.abs(r[inside]))
    print("fallbacks", fb, "of", tot, "relative RMSE vs FP64 image", rel)
    assert tot == 4 * hist.n and fb < 0.1 * tot
    assert 0 < rel < 2e-3, rel
    assert np.sqrt(np.mean((r - x) ** 2)) > 10 * np.sqrt(np.mean((g - r) ** 2))


def test_list_rule_cells_fall_back_to_exact_lists(cfg2, trained):
    """the north star's fallback rule: with PCT_NET_RULE_LISTS every cell in
    which a sampled history's surrogate voxel list disagrees with the exact
    walk's (threshold 0) goes to the exact walk; the histories of those cells
    then get the exact walk's lists bit for bit (oracle.mlp_paths), the
    others the surrogate's; the per-cell disagreement shares are reported"""
    import oracle
    import pct_amd
    from pct_amd import configs, learned
    ctx, hist = cfg2
    net, F, _ = trained
    lo, hi = learned.input_range(F, 0.02)
    learned.install(ctx, net.params, lo, hi)
    n = 3000
    cal = learned.calibrate(ctx, hist, stride=1, start=0, count=n, rule="lists", threshold=0.0, min_samples=1)
    mis = np.array(cal["cell_list_mismatch"])
    fbc = np.array(cal["cell_fallback"]).astype(bool)
    pop = np.array(cal["cell_samples"])
    print("cells with samples", int((pop > 0).sum()), "falling back", int(fbc.sum()),
          "list disagreement per cell", np.round(mis[pop > 0], 3))
    assert np.all(fbc[(pop > 0) & (mis > 0)])            # any disagreement sends the cell back
    assert not np.any(fbc[(pop > 0) & (mis == 0)])       # ... and only that (min_samples 1)
    assert cal["surrogate_mismatch"] == 0                 # every kept sample's list is the exact one
    offs, vox, lens = ctx.mlp_paths(hist, 0, n, mode=pct_amd.PCT_MLP_SURROGATE)
    fb, tot = learned.fallback_stats(ctx)
    assert tot == n and fb == cal["samples"] - cal["surrogate"]
    offs, vox, lens = offs.cpu().numpy(), vox.cpu().numpy().view(np.uint32), lens.cpu().numpy()
    exact = oracle.mlp_paths(configs.params("cfg2"), _host(hist), 0, n)
    bad = sum(not np.array_equal(vox[offs[i]:offs[i] + lens[i]], exact[i]) for i in range(n))
    assert bad == 0, f"{bad} of {n} lists differ from the exact walk's"
    # the deviation rule keeps these cells (the contrast the bench reports)
    cal_dev = learned.calibrate(ctx, hist, stride=1, start=0, count=n, tol=3e-3, min_samples=1)
    assert cal_dev["cells_fallback"] < cal["cells_fallback"]
