"""GPU: the batched form (pct_batch_*) reproduces each scan's own
pct_drop_iteration bit for bit -- different grids, block sizes and history
counts in one batch, the FP64 and the precision forms, graph and eager
launches, recomputed and cached lists."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_scans(n_scans=4, n_hist=24_000):
    """the first scans of the batched configuration with few histories:
    intersections -> cylinder hull -> endpoints -> compaction, x0 = hull"""
    import pct_amd
    from pct_amd import configs
    out = []
    for k, sc in enumerate(configs.batch_scans()[:n_scans]):
        sc.synth.n_histories = n_hist + 7000 * k
        ctx = pct_amd.Context(sc.params)
        hist, missed = ctx.recon_volume_intersections(ctx.generate(sc.synth))
        hist = ctx.compact(hist, (missed == 0).to(ctx.torch.uint8))
        hull = ctx.hull_init()
        hist = ctx.compact(hist, ctx.mlp_endpoints(hist, hull))
        x0 = ctx.initial_iterate(ctx.torch.ones(ctx.V, dtype=ctx.torch.float32, device=ctx.device), hull)
        out.append((ctx, hist, x0))
    return out


@pytest.fixture(scope="module")
def scans():
    return _small_scans()


# knobs of the first context that lay the batch's block steps out in chunks
# (each with its own MLP launches; with overlap 1 each chunk waits for its
# check) instead of the default single chunk
CHUNK_KNOBS = {"fp64_chunks": dict(chunk_first=1, chunk_max=4),
               "precision_chunks_overlap1": dict(overlap=1, chunk_first=1, chunk_max=2)}


@pytest.mark.parametrize("mode_name", ["fp64", "precision", "fp64_eager", "fp64_cached", "fp64_persistent",
                                       "precision_persistent", "fp64_persistent_eager", "fp64_persistent_cached",
                                       "fp64_chunks", "precision_chunks_overlap1"])
def test_batch_equals_single_scans(scans, mode_name):
    import pct_amd
    P = pct_amd.PCT_DROP_PERSISTENT
    mode = {"fp64": 0, "precision": pct_amd.PCT_PRECISION, "fp64_eager": pct_amd.PCT_DROP_NO_GRAPH,
            "fp64_cached": pct_amd.PCT_DROP_CACHED, "fp64_persistent": P,
            "precision_persistent": pct_amd.PCT_PRECISION | P, "fp64_persistent_eager": P | pct_amd.PCT_DROP_NO_GRAPH,
            "fp64_persistent_cached": P | pct_amd.PCT_DROP_CACHED, "fp64_chunks": 0,
            "precision_chunks_overlap1": pct_amd.PCT_PRECISION}[mode_name]
    single = mode & ~P
    knobs = CHUNK_KNOBS.get(mode_name)
    if knobs:
        pct_amd.set_tuning(scans[0][0], **knobs)
    try:
        _batch_vs_single(scans, mode, single)
    finally:
        if knobs:
            pct_amd.set_tuning(scans[0][0], **{k: 0 for k in knobs})


def _batch_vs_single(scans, mode, single):
    import pct_amd
    sizes = [800, 1600, 3200, 6400][:len(scans)]
    lams = [1e-2, 2e-2, 5e-3, 1e-2][:len(scans)]
    ref = []
    for (ctx, hist, x0), B, lam in zip(scans, sizes, lams):
        x = x0.clone()
        for _ in range(3):
            ctx.drop_iteration(hist, x, B, lam, single)
        ref.append(x.cpu().numpy())
    batch = pct_amd.Batch([c for c, _, _ in scans])
    hists = [h for _, h, _ in scans]
    xs = [x0.clone() for _, _, x0 in scans]
    batch.prepare(hists, sizes, mode & 0xff)
    for _ in range(3):
        batch.iteration(hists, xs, sizes, lams, mode)
    for k, (x, r) in enumerate(zip(xs, ref)):
        got = x.cpu().numpy()
        assert np.any(got != scans[k][2].cpu().numpy())
        assert np.array_equal(got, r), f"scan {k}: {np.count_nonzero(got != r)} voxels differ"
    batch.close()


def test_batch_replays_and_changed_block_sizes(scans):
    """graph replays with eager work between them, then new block sizes
    (plans rebuilt for the changed scans only) -- still equal to single scans"""
    import pct_amd
    batch = pct_amd.Batch([c for c, _, _ in scans])
    hists = [h for _, h, _ in scans]
    for sizes in ([3200] * len(scans), [6400, 800, 3200, 1600][:len(scans)]):
        xs = [x0.clone() for _, _, x0 in scans]
        for _ in range(2):
            batch.iteration(hists, xs, sizes, None, 0)
            for x in xs:
                x.mul_(1.0)
        for (ctx, hist, x0), B, x in zip(scans, sizes, xs):
            r = x0.clone()
            for _ in range(2):
                ctx.drop_iteration(hist, r, B, None, 0)
            assert np.array_equal(x.cpu().numpy(), r.cpu().numpy())
    batch.close()


def test_batch_follows_plans_rebuilt_by_their_own_context(scans):
    """a scan whose histories change in place and whose plan its own context
    rebuilds (pct_drop_iteration) is picked up by the next batch call: the
    batch's captured graphs are dropped and every scan still equals its own
    iteration; a context listed twice is refused"""
    import pct_amd
    with pytest.raises(pct_amd.PctError):
        pct_amd.Batch([scans[0][0], scans[0][0]])
    ctxs = [c for c, _, _ in scans[:3]]
    hists = [h for _, h, _ in scans[:3]]
    sizes = [3200, 1600, 800]
    batch = pct_amd.Batch(ctxs)
    xs = [x0.clone() for _, _, x0 in scans[:3]]
    batch.iteration(hists, xs, sizes, None, pct_amd.PCT_DROP_CACHED)
    # scan 1: histories changed in place, plan rebuilt through its own context
    h1 = hists[1]
    saved = h1.x_exit[:200].clone()
    h1.x_exit[:200] += 0.3
    ctxs[1].drop_invalidate()
    x_own = scans[1][2].clone()
    ctxs[1].drop_iteration(h1, x_own, sizes[1], None, pct_amd.PCT_DROP_CACHED)
    xs = [x0.clone() for _, _, x0 in scans[:3]]
    batch.iteration(hists, xs, sizes, None, pct_amd.PCT_DROP_CACHED)
    for (ctx, hist, x0), B, x in zip(scans[:3], sizes, xs):
        r = x0.clone()
        ctx.drop_iteration(hist, r, B, None, pct_amd.PCT_DROP_CACHED)
        assert np.array_equal(x.cpu().numpy(), r.cpu().numpy())
    batch.close()
    h1.x_exit[:200].copy_(saved)
    ctxs[1].drop_invalidate()


def test_persistent_barrier_that_gives_up_leaves_images_unchanged(scans):
    """the persistent batch's grid barriers, forced from the host to give up
    (a target no arrival count reaches, two polls): the iteration returns the
    error and every image keeps its value; without the knob the next
    iteration equals the two-launch form's"""
    import pct_amd
    sizes = [3200] * len(scans)
    hists = [h for _, h, _ in scans]
    c0 = scans[0][0]
    batch = pct_amd.Batch([c for c, _, _ in scans])
    xs = [x0.clone() for _, _, x0 in scans]
    before = [x.cpu().numpy() for x in xs]
    P = pct_amd.PCT_DROP_PERSISTENT | pct_amd.PCT_DROP_CACHED
    pct_amd.set_tuning(c0, wait_give_up=1)
    try:
        with pytest.raises(pct_amd.PctError, match="gave up"):
            batch.iteration(hists, xs, sizes, None, P)
        for x, b in zip(xs, before):
            assert np.array_equal(x.cpu().numpy().view(np.uint32), b.view(np.uint32))
        pct_amd.set_tuning(c0, wait_give_up=0)
        batch.iteration(hists, xs, sizes, None, P)
    finally:
        pct_amd.set_tuning(c0, wait_give_up=0)
    ref = [x0.clone() for _, _, x0 in scans]
    batch.iteration(hists, ref, sizes, None, pct_amd.PCT_DROP_CACHED)
    for x, r, b in zip(xs, ref, before):
        assert np.array_equal(x.cpu().numpy().view(np.uint32), r.cpu().numpy().view(np.uint32))
        assert not np.array_equal(x.cpu().numpy(), b)
    batch.close()
