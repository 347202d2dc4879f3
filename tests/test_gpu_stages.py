"""GPU parity of the stages that feed the DROP loop, through the C ABI:
binning, per-bin statistics, 3-sigma cuts and sinogram (bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg1_sample():
    import oracle
    import pct_amd
    from pct_amd import configs
    p = configs.params("cfg1")
    cfg = configs.synth_cfg("cfg1", n_histories=400_000)
    raw = pct_amd.generate_host(cfg).numpy()
    h, missed = oracle.recon_volume_intersections(p, dict(raw))
    r = oracle.bin_stats_and_cuts(p, h, missed)
    return dict(p=p, cfg=cfg, h=h, missed=missed, r=r)


def test_binning_stats_cuts_bit_exact(cfg1_sample):
    import pct_amd
    p, r = cfg1_sample["p"], cfg1_sample["r"]
    ctx = pct_amd.Context(p)
    raw = ctx.generate(cfg1_sample["cfg"])
    hist, missed = ctx.recon_volume_intersections(raw)
    in_range = ctx.binning(hist, missed)
    st = ctx.bin_statistics(hist, missed, in_range)
    got_stats = {k: v.cpu().numpy().copy() for k, v in st.items()}
    failed, sino = ctx.statistical_cuts(hist, missed, st)
    assert np.array_equal(missed.cpu().numpy(), cfg1_sample["missed"])
    assert np.array_equal(hist.bin_num.cpu().numpy(), cfg1_sample["h"].bin_num)
    assert np.array_equal(in_range.cpu().numpy(), r["in_range"])
    for k in ("counts", "mean_wepl", "mean_rel_ut", "mean_rel_uv", "stddev_wepl", "stddev_rel_ut", "stddev_rel_uv"):
        a, b = got_stats[k], r[k]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{k}: {np.count_nonzero(a != b)} bins differ"
    assert np.array_equal(failed.cpu().numpy(), r["failed"])
    assert np.array_equal(sino.cpu().numpy().view(np.uint32), r["sinogram"].view(np.uint32))
    assert np.array_equal(st["counts"].cpu().numpy(), r["cut_counts"])
    ctx.close()
