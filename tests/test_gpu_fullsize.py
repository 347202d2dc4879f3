"""GPU checks at the full size of the headline workload (200x200x32 voxels,
20 M histories -> ~17 M reconstruction histories, 5325 DROP blocks): the
oracle cannot run there in test time, so size-independent properties:

  - the schedules of an iteration -- MLP lists recomputed beside the block
    loop in chunks (default), cached lists, MLP first then the block loop,
    other chunkings, one or two launches per block -- give bit-identical
    images over two iterations, also for the research variants and the
    precision form (one and two launches per block);
  - lambda = 0 leaves the image unchanged;
  - the plan's recomputed lists match the cached ones (no plan flag);
  - FBP x0 is zero outside the hull and the reconstruction histories all
    start on a hull voxel (first_voxel inside the image).
The bench compares one iteration over the first 276 blocks of this scan
with the oracle (rmse_vs_oracle).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    import pct_amd
    from pct_amd import configs, pipeline
    p = configs.params("cfg2")
    ctx = pct_amd.Context(p)
    raw = ctx.generate(configs.synth_cfg("cfg2"))
    pre = pipeline.preprocess(ctx, raw)
    del raw
    yield p, ctx, pre
    ctx.close()


def test_full_size_schedules_agree(full):
    import pct_amd
    p, ctx, pre = full
    torch = ctx.torch
    hist, B, V = pre.hist, p.drop_block_size, ctx.V
    assert hist.n > 15_000_000
    fv = hist.first_voxel.to(torch.int64)
    assert int(fv.min()) >= 0 and int(fv.max()) < V
    bits = pre.hull.cpu().numpy().view(np.uint32)
    idx = np.arange(V)
    in_hull = ((bits[idx >> 5] >> (idx & 31)) & 1).astype(bool)
    x0 = pre.x0.cpu().numpy()
    assert in_hull.sum() > 100_000 and np.all(x0[~in_hull] == 0) and np.any(x0[in_hull] != 0)
    assert np.all(in_hull[fv.cpu().numpy()])
    print(f"cfg2 full size: {hist.n} reconstruction histories", flush=True)
    images = {}
    for name, mode, knobs in (("default", pct_amd.PCT_MLP_FP64, {}),
                              ("cached", pct_amd.PCT_MLP_FP64 | pct_amd.PCT_DROP_CACHED, {}),
                              ("sequential", pct_amd.PCT_MLP_FP64, {"overlap": 2}),
                              ("per-chunk waits", pct_amd.PCT_MLP_FP64, {"overlap": 1}),
                              ("chunks", pct_amd.PCT_MLP_FP64, {"chunk_first": 5, "chunk_max": 40}),
                              ("one launch per block", pct_amd.PCT_MLP_FP64 | pct_amd.PCT_DROP_CACHED,
                               {"block_form": 1}),
                              ("waiting one launch per block", pct_amd.PCT_MLP_FP64, {"block_form": 3}),
                              ("waiting, cached", pct_amd.PCT_MLP_FP64 | pct_amd.PCT_DROP_CACHED,
                               {"block_form": 3})):
        pct_amd.set_tuning(ctx, overlap=0, chunk_first=0, chunk_max=0, mlp_waves=0, block_form=0)
        pct_amd.set_tuning(ctx, **knobs)
        x = pre.x0.clone()
        for _ in range(2):
            ctx.drop_iteration(hist, x, B, p.lambda_, mode)
        images[name] = x.cpu().numpy()
        print(f"  {name} done", flush=True)
    pct_amd.set_tuning(ctx, overlap=0, chunk_first=0, chunk_max=0, mlp_waves=0, block_form=0)
    ref = images["default"]
    assert np.sum(ref != pre.x0.cpu().numpy()) > 100_000
    for name, img in images.items():
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), name
    x = pre.x0.clone()
    ctx.drop_iteration(hist, x, B, 0.0, pct_amd.PCT_MLP_FP64)
    assert torch.equal(x, pre.x0)


@pytest.mark.parametrize("form", ["robust", "robust2", "precision"])
def test_full_size_forms_schedules_agree(full, form):
    """the research variants and the precision form: lists recomputed beside
    the block loop and cached lists give the same image"""
    import pct_amd
    p, ctx, pre = full
    mode = {"robust": pct_amd.PCT_MLP_FP64 | pct_amd.PCT_DROP_ROBUST,
            "robust2": pct_amd.PCT_MLP_FP64 | pct_amd.PCT_DROP_ROBUST2,
            "precision": pct_amd.PCT_PRECISION}[form]
    out = []
    for extra, form in ((0, 0), (pct_amd.PCT_DROP_CACHED, 0), (pct_amd.PCT_DROP_CACHED, 1), (0, 3)):
        pct_amd.set_tuning(ctx, block_form=form)
        x = pre.x0.clone()
        for _ in range(2):
            ctx.drop_iteration(pre.hist, x, p.drop_block_size, p.lambda_, mode | extra)
        out.append(x.cpu().numpy())
    pct_amd.set_tuning(ctx, block_form=0)
    assert np.sum(out[0] != pre.x0.cpu().numpy()) > 100_000
    for o in out[1:]:
        assert np.array_equal(out[0].view(np.uint32), o.view(np.uint32))
