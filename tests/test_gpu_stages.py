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


@pytest.fixture(scope="module")
def hull_sample():
    """cfg1 geometry, 600 k histories, a 16 cm water cylinder (so that
    protons pass through air inside the reconstruction cylinder)"""
    import oracle
    import pct_amd
    from pct_amd import configs
    p = configs.params("cfg1")
    cfg = configs.synth_cfg("cfg1", n_histories=600_000)
    cfg.cyl_r[0] = 8.0
    raw = pct_amd.generate_host(cfg).numpy()
    h, missed = oracle.recon_volume_intersections(p, dict(raw))
    r = oracle.bin_stats_and_cuts(p, h, missed)
    counts = oracle.msc_counts(p, h, missed)
    edge = oracle.msc_edge(p, counts)
    hull = oracle.hull_select(p, edge)
    sc = oracle.space_carve(p, h, missed, oracle.hull_init(p))
    filt = oracle.fbp_filter(p, r["sinogram"])
    fbp = oracle.backprojection(p, filt)
    fh = oracle.fbp_hull(p, fbp)
    x0 = oracle.initial_iterate(p, fbp, hull)
    sm = oracle.sm_counts(p, h, missed)
    sm_edge, sm_thr = oracle.sm_edge(p, sm)
    return dict(p=p, cfg=cfg, missed=missed, counts=counts, edge=edge, hull=hull, sc=sc, sino=r["sinogram"],
                filt=filt, fbp=fbp, fh=fh, x0=x0, sm=sm, sm_edge=sm_edge, sm_thr=sm_thr,
                sm_hull=oracle.hull_select(p, sm_edge))


def _unpack(bits, V):
    b = bits.cpu().numpy().view(np.uint32)
    idx = np.arange(V)
    return ((b[idx >> 5] >> (idx & 31)) & 1).astype(np.uint8)


def test_hull_and_fbp_bit_exact(hull_sample):
    import pct_amd
    s = hull_sample
    p = s["p"]
    V = p.columns * p.rows * p.slices
    ctx = pct_amd.Context(p)
    torch = ctx.torch
    raw = ctx.generate(s["cfg"])
    hist, missed = ctx.recon_volume_intersections(raw)
    counts = ctx.msc_counts(hist, missed)
    assert (s["counts"] > 1).sum() > 1000               # carving happened
    assert np.array_equal(counts.cpu().numpy(), s["counts"])
    edge = ctx.msc_edge(counts)
    assert np.array_equal(edge.cpu().numpy(), s["edge"])
    hull = ctx.hull_select(edge)
    assert np.array_equal(_unpack(hull, V), s["hull"])
    sc = ctx.space_carve(hist, missed, ctx.hull_init())
    assert np.array_equal(_unpack(sc, V), s["sc"])
    sino = torch.from_numpy(s["sino"]).cuda()
    filt = ctx.fbp_filter(sino)
    assert np.array_equal(filt.cpu().numpy().view(np.uint32), s["filt"].view(np.uint32))
    fbp = ctx.backproject(filt)
    assert np.array_equal(fbp.cpu().numpy().view(np.uint32), s["fbp"].view(np.uint32))
    assert np.array_equal(_unpack(ctx.fbp_threshold(fbp), V), s["fh"])
    x0 = ctx.initial_iterate(fbp, hull)
    assert np.array_equal(x0.cpu().numpy().view(np.uint32), s["x0"].view(np.uint32))
    ctx.close()


@pytest.mark.parametrize("cols, rows, slices", [(40, 40, 8), (100, 100, 4), (200, 200, 3), (37, 53, 5)])
def test_backprojection_tiles_bit_exact(cols, rows, slices):
    """the tiled backprojection (footprints staged in LDS, or read from
    global memory where a tile's footprint exceeds the staging area, as on
    the 40-column grid's 5 mm voxels) equals the oracle bit for bit on a random
    filtered sinogram, including grids whose sides are not multiples of the
    tile"""
    import oracle
    import pct_amd
    from pct_amd import configs
    p = configs.params("mini")
    p.columns, p.rows, p.slices = cols, rows, slices
    p.recon_cyl_height = slices * p.slice_thickness
    n = p.t_bins * p.v_bins * p.angular_bins
    sf = np.random.default_rng(cols * 7 + slices).normal(0.0, 1.0, n).astype(np.float32)
    exp = oracle.backprojection(p, sf)
    ctx = pct_amd.Context(p)
    got = ctx.backproject(ctx.torch.from_numpy(sf).cuda()).cpu().numpy()
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), int(np.sum(got != exp))
    ctx.close()


def test_sm_hull_bit_exact(hull_sample):
    """space modelling: counts, per-slice thresholds, edge and selected hull
    equal the oracle's; then counts with ties of the largest difference in
    several workgroups' tiles of a slice (the first pixel in order wins) and a
    flat slice (threshold 0)"""
    import oracle
    import pct_amd
    s = hull_sample
    p = s["p"]
    C, R, S = p.columns, p.rows, p.slices
    V = C * R * S
    ctx = pct_amd.Context(p)
    torch = ctx.torch
    hist, missed = ctx.recon_volume_intersections(ctx.generate(s["cfg"]))
    sm = ctx.sm_counts(hist, missed)
    assert (s["sm"] > 1).sum() > 1000 and np.all(s["sm_thr"] > 0)
    assert np.array_equal(sm.cpu().numpy(), s["sm"])
    thr = torch.full((S,), -1, dtype=torch.int32, device="cuda")
    edge = ctx.sm_edge(sm, thr)
    assert np.array_equal(thr.cpu().numpy(), s["sm_thr"])
    assert np.array_equal(edge.cpu().numpy(), s["sm_edge"])
    assert np.array_equal(_unpack(ctx.hull_select(edge), V), s["sm_hull"])
    rng = np.random.default_rng(11)
    counts = rng.integers(0, 6, V).astype(np.int32)
    c3 = counts.reshape(S, R, C)
    c3[0] = 3
    c3[1] = 10
    for (r, c, top) in ((R - 3, C - 3, 25), (R // 2, C // 2, 21), (2, 2, 18)):   # last tile first: same drop 8
        c3[1, r - 1:r + 2, c - 1:c + 2] = top - 8
        c3[1, r, c] = top
    for scale in (1.0, 0.5):
        q = p.copy()
        q.sm_scale_threshold = scale
        exp, exp_thr = oracle.sm_edge(q, counts)
        assert exp_thr[0] == 0 and exp_thr[1] == 18
        cq = pct_amd.Context(q)
        thr.fill_(-1)
        got = cq.sm_edge(torch.from_numpy(counts).cuda(), thr)
        assert np.array_equal(thr.cpu().numpy(), exp_thr)
        assert np.array_equal(got.cpu().numpy(), exp)
        cq.close()
    # in place is refused (other workgroups read the counts being thresholded)
    assert ctx.L.pct_sm_edge(ctx._ctx, pct_amd._ptr(sm), pct_amd._ptr(sm), None) < 0
    ctx.close()


@pytest.mark.parametrize("config", ["mini", "cfg2", "mini300", "mini1"])
def test_sm_counts_random_segments(config):
    """SM counts of random entry/exit segments inside the volume, a quarter of
    them steep (the slice index of a walk going up stops at 0 and a voxel
    recurs), a few missed or below the WEPL threshold, equal the oracle's;
    on the cfg2 grid the walks use the two-slice LDS window of 156 KB, on a
    300x300 grid (one slice of 16-bit counters is 180 KB) no window, on a
    single-slice grid a one-slice window"""
    import oracle
    import pct_amd
    import torch
    from pct_amd import configs
    p = configs.params("mini" if config in ("mini300", "mini1") else config)
    if config == "mini300":
        p.columns = p.rows = 300
    if config == "mini1":
        p.slices = 1
        p.recon_cyl_height = p.slice_thickness
    rng = np.random.default_rng(31)
    n = 40_000
    r, hz = p.recon_cyl_radius, p.recon_cyl_height / 2
    h = oracle.HostHistories(n)
    ang = rng.uniform(0, 2 * np.pi, (2, n))
    rad = r * np.sqrt(rng.uniform(0, 0.98, (2, n)))
    h.x_entry[:], h.y_entry[:] = rad[0] * np.cos(ang[0]), rad[0] * np.sin(ang[0])
    h.x_exit[:], h.y_exit[:] = rad[1] * np.cos(ang[1]), rad[1] * np.sin(ang[1])
    h.z_entry[:] = rng.uniform(-hz, hz, n) * 0.9
    steep = rng.uniform(size=n) < 0.25
    h.z_exit[:] = np.where(steep, rng.choice([-1.3, 1.3], n) * hz, h.z_entry + rng.normal(0, 0.05, n))
    short = steep & (rng.uniform(size=n) < 0.5)             # steep and short in x/y
    h.x_exit[short] = h.x_entry[short] + rng.normal(0, 0.05, short.sum())
    h.y_exit[short] = h.y_entry[short] + rng.normal(0, 0.05, short.sum())
    h.wepl[:] = rng.uniform(0, 30, n).astype(np.float32)
    missed = (rng.uniform(size=n) < 0.02).astype(np.uint8)
    exp = oracle.sm_counts(p, h, missed)
    assert (exp - oracle.hull_init(p)).max() > 1
    ctx = pct_amd.Context(p)
    d = pct_amd.Histories(n, "cuda")
    for f, a in h.arrays().items():
        a = np.ascontiguousarray(a)
        getattr(d, f).copy_(torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).cuda())
    got = ctx.sm_counts(d, torch.from_numpy(missed).cuda())
    assert np.array_equal(got.cpu().numpy(), exp)
    ctx.close()


def test_sm_pipeline_matches_oracle():
    """raw scan -> SM hull -> x0 -> reconstruction histories on the GPU
    equal the oracle's preprocessing with the SM hull"""
    import oracle
    import pct_amd
    from pct_amd import configs, pipeline
    p = configs.params("cfg1")
    cfg = configs.synth_cfg("cfg1", n_histories=300_000)
    cfg.cyl_r[0] = 8.0
    raw = pct_amd.generate_host(cfg).numpy()
    recon, x0, hull, _ = oracle.preprocess(p, raw, hull_kind="sm")
    ctx = pct_amd.Context(p)
    pre = pipeline.preprocess(ctx, ctx.generate(cfg), hull_kind="sm")
    V = p.columns * p.rows * p.slices
    assert hull.sum() > 1000
    assert np.array_equal(_unpack(pre.hull, V), hull)
    assert np.array_equal(pre.x0.cpu().numpy().view(np.uint32), x0.view(np.uint32))
    assert pre.hist.n == recon.n
    assert np.array_equal(pre.hist.numpy()["first_voxel"], recon.arrays()["first_voxel"])
    ctx.close()


def test_full_pipeline_matches_oracle():
    """raw scan -> cuts -> FBP -> MSC hull -> x0 -> endpoints -> 3 DROP
    iterations, GPU (pct_amd.pipeline) vs the oracle's main()-equivalent"""
    import oracle
    import pct_amd
    from pct_amd import configs, pipeline
    p = configs.params("cfg1")
    p.lambda_ = 1e-3
    cfg = configs.synth_cfg("cfg1", n_histories=300_000)
    cfg.cyl_r[0] = 8.0
    raw = pct_amd.generate_host(cfg).numpy()
    recon, x0, hull, fbp = oracle.preprocess(p, raw)
    ctx = pct_amd.Context(p)
    pre = pipeline.preprocess(ctx, ctx.generate(cfg))
    V = p.columns * p.rows * p.slices
    assert pre.hist.n == recon.n
    assert np.array_equal(_unpack(pre.hull, V), hull)
    assert np.array_equal(pre.x0.cpu().numpy().view(np.uint32), x0.view(np.uint32))
    got = pre.hist.numpy()
    for f in ("x_entry", "y_entry", "z_exit", "first_voxel", "wepl", "bin_num"):
        assert np.array_equal(np.asarray(got[f]).view(np.uint32), np.asarray(recon.arrays()[f]).view(np.uint32)), f
    x = pipeline.image_reconstruction(ctx, pre, iterations=3)
    exp = oracle.drop(p, recon, x0, iterations=3, threads=8)
    d = np.abs(x.cpu().numpy().astype(np.float64) - exp)
    assert d.max() <= 1e-5
    assert np.mean(d == 0) > 0.99
    ctx.close()


def test_timed_pipeline_gives_the_same_results():
    """the per-stage timing path of pipeline.preprocess runs the same stages:
    results bit-identical to the untimed run, one positive time and byte
    count per stage"""
    import pct_amd
    from pct_amd import configs, pipeline
    p = configs.params("cfg1")
    ctx = pct_amd.Context(p)
    raw = ctx.generate(configs.synth_cfg("cfg1", n_histories=200_000))
    a = pipeline.preprocess(ctx, raw)
    timings = []
    b = pipeline.preprocess(ctx, raw, timings=timings)
    names = [t[0] for t in timings]
    assert names == ["recon_volume_intersections", "binning", "msc_counts", "bin_statistics", "statistical_cuts",
                     "compact_post_cut", "fbp_filter", "backproject", "msc_edge", "hull_select", "initial_iterate",
                     "mlp_endpoints", "compact_reconstruction"]
    assert all(ms > 0 and nbytes > 0 and traffic > 0 for _, ms, nbytes, traffic in timings)
    assert a.hist.n == b.hist.n and a.counts == b.counts
    assert np.array_equal(a.x0.cpu().numpy().view(np.uint32), b.x0.cpu().numpy().view(np.uint32))
    assert np.array_equal(a.hull.cpu().numpy(), b.hull.cpu().numpy())
    ga, gb = a.hist.numpy(), b.hist.numpy()
    for f in ga:
        assert np.array_equal(np.asarray(ga[f]).view(np.uint32), np.asarray(gb[f]).view(np.uint32)), f
    ctx.close()


def test_pipeline_from_pctd_files(tmp_path):
    """a scan stored as PCTD projection files (mm, one file per gantry angle)
    and read back the reference's way, then preprocessed on the GPU: equal to
    the oracle's preprocessing of the same arrays; the x0 text dump equals the
    dump of the oracle's x0"""
    import oracle
    import pct_amd
    from pct_amd import configs, pctd, pipeline
    p = configs.params("cfg1")
    cfg = configs.synth_cfg("cfg1", n_histories=200_000)
    pctd.write_scan(pct_amd.generate_host(cfg).numpy(), str(tmp_path), "/in", phantom_name="cfg1")
    host = pctd.read_scan(str(tmp_path), "/in")
    recon, x0, hull, fbp = oracle.preprocess(p, host)
    ctx = pct_amd.Context(p)
    raw = pctd.read_scan(str(tmp_path), "/in", device=ctx.device)
    pre = pipeline.preprocess(ctx, raw)
    assert pre.hist.n == recon.n > 0
    assert np.array_equal(pre.x0.cpu().numpy().view(np.uint32), x0.view(np.uint32))
    got = pre.hist.numpy()
    for f in ("x_entry", "z_exit", "wepl", "first_voxel", "bin_num"):
        assert np.array_equal(np.asarray(got[f]).view(np.uint32), np.asarray(recon.arrays()[f]).view(np.uint32)), f
    d = str(tmp_path) + "/"
    pctd.array_to_disk(pre.x0, "x0_gpu", d, "out", p.columns, p.rows, p.slices)
    pctd.array_to_disk(x0, "x0_oracle", d, "out", p.columns, p.rows, p.slices)
    assert open(d + "out/x0_gpu.txt").read() == open(d + "out/x0_oracle.txt").read()
    ctx.close()
