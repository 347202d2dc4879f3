"""GPU: pct_drop_block_update_noalloc (the allocation-free form of
pct_drop_block_update, its MLP lists in the library's fixed device array)
gives what pct_drop_block_update gives, through the C ABI: the per-voxel
counts bit for bit, the float x_update sums to float-atomic tolerance (both
scatter with hardware float atomics, whose order varies), on two calls in a
row, the second on a smaller range of histories; sizes beyond its limits are
refused; in the surrogate MLP mode, with part of the histories sent to the
exact walk through the fallback list the fixed array also holds, the same
holds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mini_scan():
    import pct_amd
    from pct_amd import configs
    p = configs.params("mini")
    ctx = pct_amd.Context(p)
    hist, missed = ctx.recon_volume_intersections(ctx.generate(configs.synth_cfg("mini")))
    hist = ctx.compact(hist, (missed == 0).to(ctx.torch.uint8))
    hull = ctx.hull_init()
    hist = ctx.compact(hist, ctx.mlp_endpoints(hist, hull))
    x0 = ctx.initial_iterate(ctx.torch.ones(ctx.V, dtype=ctx.torch.float32, device=ctx.device), hull)
    return p, ctx, hist, x0


def test_noalloc_block_update_matches_block_update():
    import pct_amd
    p, ctx, hist, x0 = _mini_scan()
    torch = ctx.torch
    assert hist.n > 9000
    lam = 1e-2
    for start, count in ((0, 8192), (1500, 700)):
        print(f"noalloc block update: histories [{start}, {start + count})", flush=True)
        out = []
        for fn in (ctx.drop_block_update, ctx.drop_block_update_noalloc):
            xu = torch.zeros(ctx.V, dtype=torch.float32, device=ctx.device)
            cnt = torch.zeros(ctx.V, dtype=torch.int32, device=ctx.device)
            fn(hist, start, count, x0, xu, cnt, lam)
            ctx.sync()
            out.append((xu.cpu().numpy(), cnt.cpu().numpy()))
        (xu_a, c_a), (xu_b, c_b) = out
        assert c_a.sum() > 20 * count
        assert np.array_equal(c_a, c_b)
        scale = float(np.max(np.abs(xu_a)))
        assert scale > 0
        np.testing.assert_allclose(xu_b, xu_a, rtol=1e-5, atol=1e-6 * scale)
        # the image update that follows gives the same image
        xa, xb = x0.clone(), x0.clone()
        for x, (xu, c) in ((xa, out[0]), (xb, out[1])):
            ctx.drop_image_update(x, torch.from_numpy(xu).to(ctx.device), torch.from_numpy(c).to(ctx.device))
        np.testing.assert_allclose(xb.cpu().numpy(), xa.cpu().numpy(), rtol=1e-6, atol=1e-6)
    xu = torch.zeros(ctx.V, dtype=torch.float32, device=ctx.device)
    cnt = torch.zeros(ctx.V, dtype=torch.int32, device=ctx.device)
    with pytest.raises(pct_amd.PctError):
        ctx.drop_block_update_noalloc(hist, 0, pct_amd._lib.PCT_NOALLOC_MAX_HISTORIES + 1, x0, xu, cnt, lam)
    ctx.close()
    p2 = configs_params_with_long_paths()
    ctx2 = pct_amd.Context(p2)
    with pytest.raises(pct_amd.PctError):
        ctx2.drop_block_update_noalloc(hist, 0, 10, x0, xu, cnt, lam)
    ctx2.close()


def configs_params_with_long_paths():
    from pct_amd import configs
    p = configs.params("mini")
    p.max_intersections = 1024
    return p


def test_noalloc_surrogate_with_fallbacks():
    import pct_amd
    from pct_amd import learned
    p, ctx, hist, x0 = _mini_scan()
    torch = ctx.torch
    lam = 1e-2
    # an untrained network over a range that excludes histories with t_in < 0:
    # those take the exact walk from the fallback list
    lo, hi = [-float("inf")] * 12, [float("inf")] * 12
    lo[6] = 0.0
    learned.install(ctx, learned.init_params(5, ctx.device), lo, hi)
    for start, count in ((0, 8192), (2000, 900)):
        print(f"noalloc surrogate block update: histories [{start}, {start + count})", flush=True)
        out, fallbacks = [], []
        for fn in (ctx.drop_block_update, ctx.drop_block_update_noalloc):
            xu = torch.zeros(ctx.V, dtype=torch.float32, device=ctx.device)
            cnt = torch.zeros(ctx.V, dtype=torch.int32, device=ctx.device)
            learned.fallback_stats(ctx)
            fn(hist, start, count, x0, xu, cnt, lam, pct_amd.PCT_MLP_SURROGATE)
            ctx.sync()
            fallbacks.append(learned.fallback_stats(ctx))
            out.append((xu.cpu().numpy(), cnt.cpu().numpy()))
        (xu_a, c_a), (xu_b, c_b) = out
        fb, tot = fallbacks[0]
        assert tot == count and 0 < fb < count and fallbacks[1] == fallbacks[0]
        assert np.array_equal(c_a, c_b)
        scale = float(np.max(np.abs(xu_a)))
        assert scale > 0
        np.testing.assert_allclose(xu_b, xu_a, rtol=1e-5, atol=1e-6 * scale)
    ctx.close()
