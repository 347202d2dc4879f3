"""GPU edge cases through the C ABI, against the CPU oracle:

  - an empty scan through every stage (no history anywhere): the per-bin and
    per-voxel outputs are what the oracle gives for no data, and every
    per-history array may be absent;
  - ragged DROP blocks: block sizes that do not divide the history count, a
    block larger than the scan, one history per block, and a history count
    that is not a multiple of the 64 lanes of a wavefront;
  - a single history;
  - a scan in which every history misses the reconstruction volume.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _to_dev(h_host, device="cuda"):
    import torch
    import pct_amd
    d = pct_amd.Histories(h_host.n, device)
    for f, arr in h_host.arrays().items():
        t = torch.from_numpy(np.ascontiguousarray(arr).view(np.int32) if arr.dtype == np.uint32 else arr)
        getattr(d, f).copy_(t.to(device))
    return d


def test_empty_scan_through_every_stage():
    import oracle
    import pct_amd
    from pct_amd import configs, pipeline
    p = configs.params("mini")
    cfg = configs.synth_cfg("mini", n_histories=0)
    V = p.columns * p.rows * p.slices
    recon_o, x0_o, hull_o, fbp_o = oracle.preprocess(p, pct_amd.generate_host(cfg).numpy())
    assert recon_o.n == 0
    ctx = pct_amd.Context(p)
    torch = ctx.torch
    raw = ctx.generate(cfg)
    hist, missed = ctx.recon_volume_intersections(raw)
    assert hist.n == 0 and missed.numel() == 0
    in_range = ctx.binning(hist, missed)
    st = ctx.bin_statistics(hist, missed, in_range)
    for k, v in st.items():
        assert int(torch.count_nonzero(v)) == 0, k
    failed, sino = ctx.statistical_cuts(hist, missed, st)
    assert failed.numel() == 0 and int(torch.count_nonzero(sino)) == 0
    counts = ctx.msc_counts(hist, missed)
    init = torch.empty_like(counts)
    ctx._check(ctx.L.pct_msc_init(ctx._ctx, pct_amd._ptr(init)), "pct_msc_init")
    assert torch.equal(counts, init)
    assert torch.equal(ctx.space_carve(hist, missed, ctx.hull_init()), ctx.hull_init())
    post = ctx.compact(hist, (failed == 0).to(torch.uint8))
    assert post.n == 0
    assert ctx.mlp_endpoints(post, ctx.hull_init()).numel() == 0
    offs, vox, lens = ctx.mlp_paths(post)
    assert lens.numel() == 0 and offs.cpu().tolist() == [0]
    x = torch.rand(V, device=ctx.device)
    x_before = x.clone()
    ctx.drop_iteration(post, x, p.drop_block_size, 1e-2)
    assert torch.equal(x, x_before)
    # the whole pipeline: x0 = FBP of an empty sinogram on the MSC hull, as the oracle
    pre = pipeline.preprocess(ctx, raw)
    assert pre.hist.n == 0
    assert np.array_equal(pre.fbp.cpu().numpy().view(np.uint32), fbp_o.view(np.uint32))
    assert np.array_equal(pre.x0.cpu().numpy().view(np.uint32), x0_o.view(np.uint32))
    ctx.close()


@pytest.mark.parametrize("n, block", [(1000 + 37, 300), (1000 + 37, 5000), (200 + 3, 1), (64 * 7 + 1, 64)])
def test_ragged_drop_blocks_match_oracle(mini_host, n, block):
    import oracle
    import pct_amd
    p = mini_host["p"]
    hr = mini_host["hr"].select(np.arange(mini_host["hr"].n) < n)
    assert hr.n == n
    x0 = mini_host["hull"].astype(np.float32)
    lam = 1e-2
    exp = oracle.drop(p, hr, x0, iterations=2, block_size=block, lam=lam)
    ctx = pct_amd.Context(p)
    x = ctx.torch.from_numpy(x0.copy()).cuda()
    d = _to_dev(hr)
    ctx.drop_iteration(d, x, block, lam)
    ctx.drop_iteration(d, x, block, lam)
    got = x.cpu().numpy()
    assert np.abs(got - x0).max() > 1e-3                  # the update is far above float rounding
    assert np.max(np.abs(got - exp)) <= 8 * np.spacing(np.float32(2.0))
    assert np.array_equal(got != x0, exp != x0)
    ctx.close()


def test_single_history(mini_host):
    import oracle
    import pct_amd
    p = mini_host["p"]
    hr = mini_host["hr"].select(np.arange(mini_host["hr"].n) == 5)
    ctx = pct_amd.Context(p)
    d = _to_dev(hr)
    offs, vox, lens = ctx.mlp_paths(d)
    exp = oracle.mlp_paths(p, hr)
    assert np.array_equal(vox.cpu().numpy().view(np.uint32)[:int(lens[0])], exp[0])
    x0 = mini_host["hull"].astype(np.float32)
    x = ctx.torch.from_numpy(x0.copy()).cuda()
    ctx.drop_iteration(d, x, p.drop_block_size, 1e-2)
    ref = oracle.drop(p, hr, x0, iterations=1, lam=1e-2)
    assert np.max(np.abs(x.cpu().numpy() - ref)) <= 2 * np.spacing(np.float32(2.0))
    assert np.count_nonzero(x.cpu().numpy() != x0) == len(np.unique(exp[0]))
    ctx.close()


def test_every_history_misses():
    """tracker hits far outside the reconstruction cylinder: every history is
    missed, the cuts keep none, and the pipeline ends with no reconstruction
    history -- as in the oracle"""
    import oracle
    import pct_amd
    from pct_amd import configs, pipeline
    p = configs.params("mini")
    cfg = configs.synth_cfg("mini", n_histories=5000)
    raw_h = pct_amd.generate_host(cfg).numpy()
    for f in ("t_in_1", "t_in_2", "t_out_1", "t_out_2"):
        raw_h[f] = (raw_h[f] + np.float32(500.0)).astype(np.float32)
    recon_o, x0_o, hull_o, fbp_o = oracle.preprocess(p, raw_h)
    _, missed_o = oracle.recon_volume_intersections(p, dict(raw_h))
    assert missed_o.all() and recon_o.n == 0
    ctx = pct_amd.Context(p)
    torch = ctx.torch
    raw = pct_amd.RawScan(len(raw_h["wepl"]), ctx.device)
    for f, a in raw_h.items():
        getattr(raw, f).copy_(torch.from_numpy(np.ascontiguousarray(a)).to(ctx.device))
    hist, missed = ctx.recon_volume_intersections(raw)
    assert np.array_equal(missed.cpu().numpy(), missed_o)
    pre = pipeline.preprocess(ctx, raw)
    assert pre.hist.n == 0 and pre.counts["post_cut"] == 0
    assert np.array_equal(pre.x0.cpu().numpy().view(np.uint32), x0_o.view(np.uint32))
    ctx.close()


def test_sm_research_chord_on_empty_and_degenerate_input(mini_host):
    """the SM hull, the chord lengths and the research list loop on no
    histories, on lists without intersections, and on single lists with
    blocks of one -- as the oracle"""
    import oracle
    import pct_amd
    from pct_amd import configs, pipeline
    p = configs.params("mini")
    V = p.columns * p.rows * p.slices
    ctx = pct_amd.Context(p)
    torch = ctx.torch
    cfg = configs.synth_cfg("mini", n_histories=0)
    raw = ctx.generate(cfg)
    hist, missed = ctx.recon_volume_intersections(raw)
    counts = ctx.sm_counts(hist, missed)
    init = oracle.hull_init(p).astype(np.int32)
    assert np.array_equal(counts.cpu().numpy(), init)
    thr = torch.full((p.slices,), -1, dtype=torch.int32, device="cuda")
    edge = ctx.sm_edge(counts, thr)
    exp_edge, exp_thr = oracle.sm_edge(p, init)
    assert np.array_equal(edge.cpu().numpy(), exp_edge) and np.array_equal(thr.cpu().numpy(), exp_thr)
    assert ctx.chord_lengths(hist).numel() == 0
    pre = pipeline.preprocess(ctx, raw, hull_kind="sm")
    _, x0_o, hull_o, _ = oracle.preprocess(p, pct_amd.generate_host(cfg).numpy(), hull_kind="sm")
    assert np.array_equal(pre.x0.cpu().numpy().view(np.uint32), x0_o.view(np.uint32))
    # lists: none; five without intersections; one list per block
    x0 = (1.0 + 0.1 * np.sin(np.arange(V))).astype(np.float32)
    x = torch.from_numpy(x0.copy()).cuda()
    xu = torch.zeros(V, dtype=torch.float64, device="cuda")
    cnt = torch.zeros(V, dtype=torch.int32, device="cuda")
    one = torch.zeros(1, dtype=torch.float32, device="cuda")
    empty_offs = torch.zeros(1, dtype=torch.int64, device="cuda")
    none = torch.zeros(0, dtype=torch.int32, device="cuda")
    ctx.drop_lists_iteration(empty_offs, none, one, one.double(), one, x, xu, cnt, 2, 0.01, pct_amd.PCT_LISTS_ERROR_RED)
    five = torch.zeros(6, dtype=torch.int64, device="cuda")
    ctx.drop_lists_iteration(five, none, one, torch.full((5,), 0.3, dtype=torch.float64, device="cuda"),
                             torch.ones(5, device="cuda"), x, xu, cnt, 2, 0.01, pct_amd.PCT_LISTS_ERROR_RED)
    assert torch.equal(x.cpu(), torch.from_numpy(x0)) and int(cnt.sum()) == 0
    hr = mini_host["hr"].select(np.arange(mini_host["hr"].n) < 3)
    lists = oracle.mlp_paths(p, hr)
    offsets, vox = oracle.lists_csr(lists)
    rng = np.random.default_rng(2)
    err = rng.normal(0, 1, vox.size).astype(np.float32)
    chord = np.array([0.31, 0.29, 0.33])
    for variant in range(5):
        xo, xuo, cno = x0.copy(), np.zeros(V), np.zeros(V, np.uint32)
        oracle.drop_lists_research(V, offsets, vox, err, chord, hr.wepl, xo, xuo, cno, 1, 0.05, p.eta, p.psi_sign,
                                   variant)
        xd = torch.from_numpy(x0.copy()).cuda()
        xu.zero_()
        cnt.zero_()
        ctx.drop_lists_iteration(torch.from_numpy(offsets).cuda(), torch.from_numpy(vox.view(np.int32)).cuda(),
                                 torch.from_numpy(err).cuda(), torch.from_numpy(chord).cuda(),
                                 torch.from_numpy(np.ascontiguousarray(hr.wepl)).cuda(), xd, xu, cnt, 1, 0.05, variant)
        assert np.array_equal(xd.cpu().numpy().view(np.uint32), xo.view(np.uint32)), variant
        assert int(cnt.sum()) == 0 and np.sum(xo != x0) > 0
    ctx.close()


@pytest.mark.parametrize("block_form", [2, 1, 3])
def test_drop_on_a_grid_of_partial_tiles(block_form):
    """a 38x38x5 grid: neither dimension a multiple of the plan's 4 x 8 voxel
    tiles, so the tiled image is padded; the image against oracle.drop for
    every block-loop form"""
    import oracle
    import pct_amd
    from pct_amd import configs
    p = configs.params("mini", columns=38, rows=38, slices=5)
    p.recon_cyl_height = p.slices * p.slice_thickness
    cfg = configs.synth_cfg("mini", n_histories=12_000, seed=configs.SEED_BASE + 77)
    raw = pct_amd.generate_host(cfg).numpy()
    h_all, missed = oracle.recon_volume_intersections(p, dict(raw))
    h = h_all.select(missed == 0)
    hull = oracle.hull_init(p)
    inter = oracle.mlp_endpoints(p, h, hull)
    hr = h.select(inter == 1)
    assert hr.n > 2000
    x0 = hull.astype(np.float32)
    lam, block = 1e-2, 700
    exp = oracle.drop(p, hr, x0, iterations=2, block_size=block, lam=lam)
    ctx = pct_amd.Context(p)
    pct_amd.set_tuning(ctx, block_form=block_form)
    d = _to_dev(hr)
    x = ctx.torch.from_numpy(x0.copy()).cuda()
    ctx.drop_iteration(d, x, block, lam)
    ctx.drop_iteration(d, x, block, lam)
    got = x.cpu().numpy()
    assert np.abs(got - x0).max() > 1e-3
    assert np.max(np.abs(got - exp)) <= 8 * np.spacing(np.float32(2.0))
    assert np.array_equal(got != x0, exp != x0)
    ctx.close()
