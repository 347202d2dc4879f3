"""GPU: the batched form (12 independent scans in one pair of launches per
block step, .cu:6538-6866 per scan) against the CPU oracle and, at the full
size of the batched configuration, through size-independent properties.

  - oracle: all 12 grids of configs.batch_scans() (64x64x8 ... 256x256x48,
    random-ellipsoid phantoms) with a few thousand histories each (what the
    oracle finishes in seconds), block sizes 800 / 1600 / 3200 / 6400 across
    the scans: each scan's batched image after two iterations equals
    oracle.drop (DROP_blocks + DROP_update, .cu:8023-8055) to float ulps,
    nearly all voxels bit-identical;
  - full size (0.5-30 M histories per scan, 66 M reconstruction histories):
    the batch equals each scan's own pct_drop_iteration bit for bit, lambda = 0
    leaves every image unchanged, and recomputed lists give the same images as
    cached lists.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [800, 1600, 3200, 6400]


def _prep_gpu(sc):
    import pct_amd
    ctx = pct_amd.Context(sc.params)
    hist, missed = ctx.recon_volume_intersections(ctx.generate(sc.synth))
    hist = ctx.compact(hist, (missed == 0).to(ctx.torch.uint8))
    hull = ctx.hull_init()
    hist = ctx.compact(hist, ctx.mlp_endpoints(hist, hull))
    x0 = ctx.initial_iterate(ctx.torch.ones(ctx.V, dtype=ctx.torch.float32, device=ctx.device), hull)
    return ctx, hist, x0


def _prep_oracle(sc):
    import oracle
    import pct_amd
    p = sc.params
    raw = pct_amd.generate_host(sc.synth).numpy()
    h, missed = oracle.recon_volume_intersections(p, raw)
    h = h.select(missed == 0)
    hull = oracle.hull_init(p)
    inter = oracle.mlp_endpoints(p, h, hull)
    return h.select(inter == 1), hull.astype(np.float32)


def test_batch_of_all_twelve_grids_matches_oracle():
    import oracle
    import pct_amd
    from pct_amd import configs
    scans = configs.batch_scans()
    assert len(scans) == 12
    ctxs, hists, xs, refs, sizes, lams = [], [], [], [], [], []
    for k, sc in enumerate(scans):
        sc.synth.n_histories = 3000 + 700 * k
        print(f"scan {k}: {sc.params.columns}x{sc.params.rows}x{sc.params.slices}, "
              f"{sc.synth.n_histories} histories", flush=True)
        ctx, hist, x0 = _prep_gpu(sc)
        h, x0h = _prep_oracle(sc)
        assert h.n == hist.n
        assert np.array_equal(x0.cpu().numpy(), x0h)
        B, lam = SIZES[k % 4], 1e-2
        refs.append((x0h, oracle.drop(sc.params, h, x0h, iterations=2, block_size=B, lam=lam, threads=8)))
        ctxs.append(ctx), hists.append(hist), xs.append(x0.clone()), sizes.append(B), lams.append(lam)
    batch = pct_amd.Batch(ctxs)
    batch.prepare(hists, sizes, 0)
    for _ in range(2):
        batch.iteration(hists, xs, sizes, lams, 0)
    tol = 4 * np.spacing(np.float32(2.0))
    for k, (x, (x0h, exp)) in enumerate(zip(xs, refs)):
        got = x.cpu().numpy()
        assert np.sqrt(np.mean((exp - x0h) ** 2)) > 1e-4, k
        assert np.max(np.abs(got - exp)) <= 4 * tol, (k, float(np.max(np.abs(got - exp))))
        assert np.mean(got == exp) > 0.99, k
    batch.close()
    for c in ctxs:
        c.close()


def test_full_size_batch_properties():
    import pct_amd
    from pct_amd import configs, pipeline
    scans = configs.batch_scans()
    ctxs, hists, x0s = [], [], []
    for k, sc in enumerate(scans):
        ctx = pct_amd.Context(sc.params)
        pre = pipeline.preprocess(ctx, ctx.generate(sc.synth))
        ctxs.append(ctx), hists.append(pre.hist), x0s.append(pre.x0)
        print(f"scan {k}: {sc.name}, {pre.hist.n} reconstruction histories", flush=True)
    n = sum(h.n for h in hists)
    assert n > 50_000_000
    sizes = [SIZES[k % 4] for k in range(12)]
    lam = [1e-4] * 12
    batch = pct_amd.Batch(ctxs)
    batch.prepare(hists, sizes, 0)
    xs = [x.clone() for x in x0s]
    batch.iteration(hists, xs, sizes, lam, 0)                       # recomputed lists
    xc = [x.clone() for x in x0s]
    batch.iteration(hists, xc, sizes, lam, pct_amd.PCT_DROP_CACHED)  # cached lists
    print("batch iterations done", flush=True)
    for k, (ctx, h, x0, B) in enumerate(zip(ctxs, hists, x0s, sizes)):
        r = x0.clone()
        ctx.drop_iteration(h, r, B, 1e-4, pct_amd.PCT_DROP_CACHED)
        assert ctx.torch.equal(xs[k], r), f"scan {k}: batch != own iteration"
        assert ctx.torch.equal(xc[k], r), f"scan {k}: cached != recomputed"
        assert not ctx.torch.equal(r, x0), k
    xz = [x.clone() for x in x0s]
    batch.iteration(hists, xz, sizes, [0.0] * 12, pct_amd.PCT_DROP_CACHED)
    for k in range(12):
        assert ctxs[k].torch.equal(xz[k], x0s[k]), f"scan {k}: lambda 0 changed the image"
    batch.close()
    for c in ctxs:
        c.close()
