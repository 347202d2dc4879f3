"""GPU: scans whose DROP plan does not fit run in chunks of blocks.

  - a plan limit (PCT_TUNE_PLAN_LIMIT_MB) below what the mini scan's plan needs
    makes pct_drop_iteration run the reference's structure (atomic scatter and
    volume scan per block) over chunks of whole blocks -- many chunks at that
    limit -- and gives the plan form's image to float-atomic tolerance;
    pct_drop_prepare reports the capacity; without the limit the plan is back;
  - a scan of the reference's own size, NUM_RECON_HISTORIES = 105 642 524
    histories (include/pCT_data.h:130), on the 200x200x32 grid: the plan would
    exceed the device's memory, one iteration runs in chunks and updates the
    image (finite, moved, its RMS change of the order of the 20 M-history scan's).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _to_dev(h_host):
    import torch
    import pct_amd
    d = pct_amd.Histories(h_host.n, "cuda")
    for f, arr in h_host.arrays().items():
        t = torch.from_numpy(np.ascontiguousarray(arr).view(np.int32) if arr.dtype == np.uint32 else arr)
        getattr(d, f).copy_(t.to("cuda"))
    return d


def test_plan_limit_falls_back_to_chunks(mini_host):
    import pct_amd
    ctx = pct_amd.Context(mini_host["p"])
    torch = ctx.torch
    d = _to_dev(mini_host["hr"])
    x0 = mini_host["hull"].astype(np.float32)
    lam = 1e-2
    out = {}
    for name, limit in (("plan", 0), ("chunks", 1), ("plan again", 0)):
        pct_amd.set_tuning(ctx, plan_limit_mb=limit)
        x = torch.from_numpy(x0.copy()).cuda()
        for _ in range(2):
            ctx.drop_iteration(d, x, lam=lam)
        out[name] = x.cpu().numpy()
        if limit:
            with pytest.raises(pct_amd.PctError, match="chunks of blocks"):
                ctx.drop_prepare(d)
    x = torch.from_numpy(x0.copy()).cuda()
    for _ in range(2):
        ctx.drop_iteration(d, x, lam=lam, mode=pct_amd.PCT_DROP_ATOMIC)      # one chunk
    out["atomic"] = x.cpu().numpy()
    assert np.array_equal(out["plan"], out["plan again"])
    assert np.any(out["chunks"] != x0)
    for k in ("chunks", "atomic"):
        assert np.max(np.abs(out[k] - out["plan"])) < 1e-5, k
    ctx.close()


def test_reference_scale_scan_reconstructs():
    import pct_amd
    from pct_amd import configs, pipeline
    p = configs.params("cfg2")
    n = 105_642_524
    ctx = pct_amd.Context(p)
    torch = ctx.torch
    pre = pipeline.preprocess(ctx, ctx.generate(configs.synth_cfg("cfg2", n_histories=n)))
    hist = pre.hist
    print(f"{n} histories -> {hist.n} reconstruction histories", flush=True)
    assert hist.n > 80_000_000
    with pytest.raises(pct_amd.PctError, match="chunks of blocks"):
        ctx.drop_prepare(hist)
    x = pre.x0.clone()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(ctx.stream)
    ctx.drop_iteration(hist, x, p.drop_block_size, p.lambda_)
    e1.record(ctx.stream)
    torch.cuda.synchronize()
    got, x0 = x.cpu().numpy(), pre.x0.cpu().numpy()
    inside = x0 != 0
    moved = float(np.sqrt(np.mean((got[inside] - x0[inside]) ** 2)))
    print(f"one iteration in chunks: {e0.elapsed_time(e1):.0f} ms, RMS change {moved:.3e}", flush=True)
    assert np.all(np.isfinite(got))
    assert 1e-4 < moved < 1.0
    ctx.close()
