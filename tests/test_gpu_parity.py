"""GPU parity of the DROP hot path against the CPU oracle (through the C ABI).

Bit-exact: the shared math library (host vs device), the synthetic generator,
recon-volume intersections, the cylinder hull, MLP endpoints, compaction and
MLP voxel lists. Floating point: the DROP image (float atomics on the GPU vs
double accumulation in storage order in the oracle).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx_mini(mini_host):
    import pct_amd
    return pct_amd.Context(mini_host["p"])


def _to_dev(h_host, device="cuda"):
    import torch
    import pct_amd
    d = pct_amd.Histories(h_host.n, device)
    for f, arr in h_host.arrays().items():
        t = torch.from_numpy(np.ascontiguousarray(arr).view(np.int32) if arr.dtype == np.uint32 else arr)
        getattr(d, f).copy_(t.to(device))
    return d


@pytest.mark.parametrize("func", ["log", "sin", "cos", "tan", "atan", "atan2", "sinf", "cosf", "atan2f", "sqrt", "div"])
def test_math_host_device_bit_equal(ctx_mini, func):
    import oracle
    import pct_amd
    rng = np.random.default_rng(7)
    n = 200_000
    if func == "log":
        a = np.concatenate([rng.uniform(1e-4, 40, n // 2), 10 ** rng.uniform(-300, 300, n // 2)])
    elif func in ("sqrt",):
        a = 10 ** rng.uniform(-300, 300, n)
    elif func in ("atan",):
        a = np.concatenate([rng.uniform(-3, 3, n // 2), 10 ** rng.uniform(-8, 8, n // 2)])
    else:
        a = rng.uniform(-13, 13, n)
    b = rng.uniform(-13, 13, n) if func in ("atan2", "atan2f", "div") else None
    host = oracle.math_eval(func, a, b)
    dev = pct_amd.selftest_math(ctx_mini, oracle.MATH_FUNCS[func], a, b)
    assert np.array_equal(host.view(np.uint64), dev.view(np.uint64)), \
        f"{func}: {np.count_nonzero(host.view(np.uint64) != dev.view(np.uint64))} of {n} differ"


def test_shared_reciprocal_quotients_are_ieee(ctx_mini):
    """the exact MLP step's quotients with one reciprocal per denominator (the
    compiler's division without its scale and fixup steps) equal the IEEE
    quotient for operands of magnitude in [2^-300, 2^300], log-uniform with
    both signs, including the operands near powers of two and with
    significands of all ones; a zero numerator over a positive denominator
    gives +0; the logarithm with that division equals pct_log"""
    import oracle
    import pct_amd
    rng = np.random.default_rng(11)
    n = 400_000
    mag = lambda k: np.ldexp(rng.uniform(1.0, 2.0, k), rng.integers(-300, 300, k))  # noqa: E731
    sign = lambda k: np.where(rng.random(k) < 0.5, -1.0, 1.0)  # noqa: E731
    a, b = sign(n) * mag(n), sign(n) * mag(n)
    edge = np.ldexp(1.0, rng.integers(-300, 300, 4000))
    ones = np.nextafter(np.ldexp(2.0, rng.integers(-300, 299, 4000)), 0.0)
    a = np.concatenate([a, edge, ones, np.zeros(100)])
    b = np.concatenate([b, ones, edge, mag(100)])
    host = a / b
    assert np.array_equal(host.view(np.uint64), oracle.math_eval("div_shared_rcp", a, b).view(np.uint64))
    dev = pct_amd.selftest_math(ctx_mini, oracle.MATH_FUNCS["div_shared_rcp"], a, b)
    bad = np.count_nonzero(host.view(np.uint64) != dev.view(np.uint64))
    assert bad == 0, f"{bad} of {a.size} quotients differ"
    x = np.concatenate([rng.uniform(1e-4, 40, n // 2), 10 ** rng.uniform(-300, 300, n // 2), [1.0, 2.0, 0.5]])
    h = oracle.math_eval("log", x)
    d = pct_amd.selftest_math(ctx_mini, oracle.MATH_FUNCS["log_rcp_div"], x)
    assert np.array_equal(h.view(np.uint64), d.view(np.uint64))


def test_synth_device_equals_host(ctx_mini, mini_host):
    raw_dev = ctx_mini.generate(mini_host["cfg"]).numpy()
    for k, v in mini_host["raw"].items():
        assert np.array_equal(np.asarray(v).view(np.uint32), raw_dev[k].view(np.uint32)), k


def test_recon_volume_intersections_bit_exact(ctx_mini, mini_host):
    raw = ctx_mini.generate(mini_host["cfg"])
    hist, missed = ctx_mini.recon_volume_intersections(raw)
    got = hist.numpy()
    exp = mini_host["h_all"].arrays()
    assert np.array_equal(missed.cpu().numpy(), mini_host["missed"])
    keep = mini_host["missed"] == 0
    for f in ("x_entry", "y_entry", "z_entry", "x_exit", "y_exit", "z_exit", "xy_entry_angle", "xz_entry_angle",
              "xy_exit_angle", "xz_exit_angle", "wepl"):
        a, b = exp[f][keep].view(np.uint32), got[f][keep].view(np.uint32)
        assert np.array_equal(a, b), f"{f}: {np.count_nonzero(a != b)} differ"


def test_hull_init_bit_exact(ctx_mini, mini_host):
    bits = ctx_mini.hull_init().cpu().numpy().view(np.uint32)
    V = mini_host["hull"].size
    unpacked = ((bits[np.arange(V) >> 5] >> (np.arange(V) & 31)) & 1).astype(np.uint8)
    assert np.array_equal(unpacked, mini_host["hull"])


def test_endpoints_and_compaction_bit_exact(ctx_mini, mini_host):
    d = _to_dev(mini_host["h_pre"])
    bits = ctx_mini.hull_init()
    inter = ctx_mini.mlp_endpoints(d, bits).cpu().numpy()
    assert np.array_equal(inter, mini_host["inter"])
    comp = ctx_mini.compact(d, ctx_mini.torch.from_numpy(inter).cuda())
    got = comp.numpy()
    exp = mini_host["hr"].arrays()
    assert comp.n == mini_host["hr"].n
    for f in ("x_entry", "y_entry", "z_entry", "x_exit", "y_exit", "z_exit", "first_voxel", "wepl"):
        assert np.array_equal(np.asarray(exp[f]).view(np.uint32), np.asarray(got[f]).view(np.uint32)), f


def test_mlp_paths_bit_exact(ctx_mini, mini_host):
    import oracle
    hr = mini_host["hr"]
    d = _to_dev(hr)
    offs, vox, lens = ctx_mini.mlp_paths(d)
    offs, vox, lens = offs.cpu().numpy(), vox.cpu().numpy().view(np.uint32), lens.cpu().numpy()
    exp = oracle.mlp_paths(mini_host["p"], hr)
    bad = 0
    for k, e in enumerate(exp):
        g = vox[offs[k]:offs[k] + lens[k]]
        if len(g) != len(e) or not np.array_equal(g, e):
            bad += 1
    assert bad == 0, f"{bad} of {len(exp)} voxel lists differ"


def test_mlp_paths_subrange(ctx_mini, mini_host):
    import oracle
    hr = mini_host["hr"]
    d = _to_dev(hr)
    start, count = 1234, 777
    offs, vox, lens = ctx_mini.mlp_paths(d, start, count)
    offs, vox, lens = offs.cpu().numpy(), vox.cpu().numpy().view(np.uint32), lens.cpu().numpy()
    exp = oracle.mlp_paths(mini_host["p"], hr, start, count)
    for k, e in enumerate(exp):
        assert np.array_equal(vox[offs[k]:offs[k] + lens[k]], e), k


def _rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


def test_drop_iteration_matches_oracle(ctx_mini, mini_host):
    """Plan path: FP64 deltas, FP64 per-voxel sums in history order (the
    oracle's order); only the a_i.x summation order differs (wave tree vs
    sequential), so the images agree to float rounding."""
    import oracle
    torch = ctx_mini.torch
    p = mini_host["p"]
    hr = mini_host["hr"]
    x0 = mini_host["hull"].astype(np.float32)
    lam = 1e-2   # large relaxation so that the update is far above float rounding
    exp1 = oracle.drop(p, hr, x0, iterations=1, lam=lam)
    exp3 = oracle.drop(p, hr, x0, iterations=3, lam=lam)
    d = _to_dev(hr)
    x = torch.from_numpy(x0.copy()).cuda()
    ctx_mini.drop_iteration(d, x, lam=lam)
    got1 = x.cpu().numpy()
    ctx_mini.drop_iteration(d, x, lam=lam)
    ctx_mini.drop_iteration(d, x, lam=lam)
    got3 = x.cpu().numpy()
    step1 = _rmse(exp1, x0)
    assert step1 > 1e-3
    tol = 4 * np.spacing(np.float32(2.0))          # a few float ulps at |x| <= 2
    assert np.max(np.abs(got1 - exp1)) <= tol
    assert np.max(np.abs(got3 - exp3)) <= 4 * tol
    assert np.mean(got1 == exp1) > 0.99             # nearly all voxels bit-identical
    assert np.array_equal(got1 != x0, exp1 != x0)   # same voxels touched


def test_drop_modes_agree(ctx_mini, mini_host):
    """recompute (default), cached lists, eager launches, the two one-launch
    block forms and the reference-style atomic scatter give the same image (the
    last to float-atomic tolerance)."""
    import pct_amd
    torch = ctx_mini.torch
    d = _to_dev(mini_host["hr"])
    x0 = mini_host["hull"].astype(np.float32)
    lam = 1e-2
    out = {}
    for name, mode, form in [("default", 0, 0), ("cached", pct_amd.PCT_DROP_CACHED, 0),
                             ("eager", pct_amd.PCT_DROP_NO_GRAPH, 0), ("one-launch", 0, 1),
                             ("one-launch eager", pct_amd.PCT_DROP_NO_GRAPH, 1),
                             ("waiting", 0, 3), ("waiting cached", pct_amd.PCT_DROP_CACHED, 3),
                             ("waiting eager", pct_amd.PCT_DROP_NO_GRAPH, 3),
                             ("f32", pct_amd.PCT_DROP_F32, 0), ("f32 one-launch", pct_amd.PCT_DROP_F32, 1),
                             ("f32 waiting", pct_amd.PCT_DROP_F32, 3),
                             ("atomic", pct_amd.PCT_DROP_ATOMIC, 0)]:
        pct_amd.set_tuning(ctx_mini, block_form=form)
        x = torch.from_numpy(x0.copy()).cuda()
        ctx_mini.drop_iteration(d, x, lam=lam, mode=mode)
        ctx_mini.drop_iteration(d, x, lam=lam, mode=mode)
        out[name] = x.cpu().numpy()
    pct_amd.set_tuning(ctx_mini, block_form=0)
    for name in ("cached", "eager", "one-launch", "one-launch eager", "waiting", "waiting cached", "waiting eager"):
        assert np.array_equal(out["default"].view(np.uint32), out[name].view(np.uint32)), name
    for name in ("f32 one-launch", "f32 waiting"):
        assert np.array_equal(out["f32"].view(np.uint32), out[name].view(np.uint32)), name
    assert np.max(np.abs(out["default"] - out["atomic"])) < 1e-5


def test_waiting_form_wait_that_gives_up_leaves_x_unchanged(ctx_mini, mini_host):
    """block form 3's bounded waits, forced from the host to give up (they
    target a stamp no update writes, two polls each): the iteration returns
    the error and x keeps its value; without the knob the next iteration runs
    and equals the two-launch form's"""
    import pct_amd
    torch = ctx_mini.torch
    d = _to_dev(mini_host["hr"])
    x0 = mini_host["hull"].astype(np.float32)
    lam = 1e-2
    pct_amd.set_tuning(ctx_mini, block_form=3, wait_give_up=1)
    try:
        x = torch.from_numpy(x0.copy()).cuda()
        with pytest.raises(pct_amd.PctError, match="gave up"):
            ctx_mini.drop_iteration(d, x, lam=lam, mode=pct_amd.PCT_DROP_CACHED)
        assert np.array_equal(x.cpu().numpy().view(np.uint32), x0.view(np.uint32))
        pct_amd.set_tuning(ctx_mini, wait_give_up=0)
        ctx_mini.drop_iteration(d, x, lam=lam, mode=pct_amd.PCT_DROP_CACHED)
        got = x.cpu().numpy()
    finally:
        pct_amd.set_tuning(ctx_mini, block_form=0, wait_give_up=0)
    x2 = torch.from_numpy(x0.copy()).cuda()
    ctx_mini.drop_iteration(d, x2, lam=lam, mode=pct_amd.PCT_DROP_CACHED)
    assert np.array_equal(got.view(np.uint32), x2.cpu().numpy().view(np.uint32))
    assert not np.array_equal(got, x0)


def test_graph_replays_match_eager(ctx_mini, mini_host):
    """The captured block loop replayed several times, with eager work and
    synchronisations in between, matches the eager launches every time."""
    import pct_amd
    torch = ctx_mini.torch
    d = _to_dev(mini_host["hr"])
    x0 = mini_host["hull"].astype(np.float32)
    xg = torch.from_numpy(x0.copy()).cuda()
    xe = torch.from_numpy(x0.copy()).cuda()
    for it in range(4):
        ctx_mini.drop_iteration(d, xg, lam=1e-2, mode=pct_amd.PCT_DROP_CACHED)
        ctx_mini.drop_iteration(d, xe, lam=1e-2, mode=pct_amd.PCT_DROP_CACHED | pct_amd.PCT_DROP_NO_GRAPH)
        torch.cuda.synchronize()
        _ = (xg * 2).sum().item()      # unrelated eager work between replays
        assert torch.equal(xg, xe), f"replay {it} differs"


def test_plan_detects_changed_inputs(mini_host):
    """a recomputed MLP pass whose lists differ from the plan is an error"""
    import pct_amd
    ctx = pct_amd.Context(mini_host["p"])
    torch = ctx.torch
    d = _to_dev(mini_host["hr"])
    x = torch.from_numpy(mini_host["hull"].astype(np.float32)).cuda()
    ctx.drop_iteration(d, x, lam=1e-2)
    before = x.clone()
    d.x_exit[:100] += 0.5           # same arrays, different content
    with pytest.raises(pct_amd.PctError):
        ctx.drop_iteration(d, x, lam=1e-2)
    assert torch.equal(x, before)   # x restored to its value before the call
    ctx.drop_iteration(d, x, lam=1e-2)          # the plan is rebuilt for the new content
    assert not torch.equal(x, before)
    # cached lists after an in-place change: the caller invalidates the plan
    x2 = before.clone()
    ctx.drop_invalidate()
    ctx.drop_iteration(d, x2, lam=1e-2, mode=pct_amd.PCT_DROP_CACHED)
    x3 = before.clone()
    ctx2 = pct_amd.Context(mini_host["p"])
    ctx2.drop_iteration(d, x3, lam=1e-2, mode=pct_amd.PCT_DROP_CACHED)
    assert torch.equal(x2, x3)
    ctx2.close()
    ctx.close()


def test_drop_block_api_matches_iteration(ctx_mini, mini_host):
    """pct_drop_block_update + pct_drop_image_update per block == pct_drop_iteration."""
    torch = ctx_mini.torch
    hr = mini_host["hr"]
    d = _to_dev(hr)
    x0 = mini_host["hull"].astype(np.float32)
    lam = 1e-2
    x_it = torch.from_numpy(x0.copy()).cuda()
    ctx_mini.drop_iteration(d, x_it, lam=lam)
    x = torch.from_numpy(x0.copy()).cuda()
    xu = torch.zeros_like(x)
    cnt = torch.zeros(x.numel(), dtype=torch.int32, device="cuda")
    B = mini_host["p"].drop_block_size
    for s in range(0, hr.n, B):
        ctx_mini.drop_block_update(d, s, min(B, hr.n - s), x, xu, cnt, lam)
        ctx_mini.drop_image_update(x, xu, cnt)
    a, b = x.cpu().numpy(), x_it.cpu().numpy()
    assert np.max(np.abs(a - b)) < 1e-5
    assert int(cnt.abs().sum()) == 0 and float(xu.abs().sum()) == 0.0


def test_tuning_knobs_do_not_change_results(ctx_mini, mini_host):
    """overlapped / sequential MLP and any chunking of the block loop give the
    same image bit for bit"""
    import pct_amd
    torch = ctx_mini.torch
    d = _to_dev(mini_host["hr"])
    x0 = mini_host["hull"].astype(np.float32)
    out = []
    for knobs in ({}, {"overlap": 1}, {"overlap": 2}, {"chunk_first": 1, "chunk_max": 2},
                  {"overlap": 1, "chunk_first": 1, "chunk_max": 2}, {"chunk_first": 3, "mlp_waves": 1}):
        pct_amd.set_tuning(ctx_mini, overlap=0, chunk_first=0, chunk_max=0, mlp_waves=0)
        pct_amd.set_tuning(ctx_mini, **knobs)
        x = torch.from_numpy(x0.copy()).cuda()
        for _ in range(3):
            ctx_mini.drop_iteration(d, x, lam=1e-2)
        out.append(x.cpu().numpy())
    pct_amd.set_tuning(ctx_mini, overlap=0, chunk_first=0, chunk_max=0, mlp_waves=0)
    for o in out[1:]:
        assert np.array_equal(out[0], o)


@pytest.mark.parametrize("columns, launch", [(256, "sequential"), (256, "beside"), (64, "sequential"),
                                             (512, "sequential")])
def test_mlp_lists_bit_exact_on_other_grids(columns, launch):
    """MLP voxel lists against the oracle on grids whose per-depth tables take
    the different launch shapes: 256 columns (a 40 KB table: three
    workgroups per CU with the table in LDS, at the full occupancy and beside
    the block loop), 64 columns (a 10 KB table, four workgroups) and 512
    columns (an 81 KB table, read from global memory)"""
    import oracle
    import pct_amd
    from pct_amd import configs
    slices = 8 if columns > 256 else 16
    p = configs.params("cfg2", columns=columns, rows=columns, slices=slices)
    p.recon_cyl_height = slices * p.slice_thickness
    p.max_intersections = max(p.max_intersections, 4 * columns)   # paths of ~2 columns of voxels
    cfg = configs.synth_cfg("cfg2", n_histories=60_000, seed=configs.SEED_BASE + 40 + columns)
    raw = pct_amd.generate_host(cfg).numpy()
    h_all, missed = oracle.recon_volume_intersections(p, dict(raw))
    h = h_all.select(missed == 0)
    inter = oracle.mlp_endpoints(p, h, oracle.hull_init(p))
    hr = h.select(inter == 1).select(np.arange(int(inter.sum())) < 6000)
    ctx = pct_amd.Context(p)
    d = _to_dev(hr)
    if launch == "beside":
        # the DROP plan's lists come from the single-pass MLP; a recomputing
        # iteration walks them again beside the block loop and checks them
        # against the plan's hash (an error on any difference)
        x = ctx.torch.zeros(ctx.V, dtype=ctx.torch.float32, device=ctx.device)
        ctx.drop_iteration(d, x, 1000, 1e-3)
        ctx.drop_iteration(d, x, 1000, 1e-3)
    for table in (False, True):            # exact walk and precision form
        mode = pct_amd.PCT_MLP_TABLE if table else pct_amd.PCT_MLP_FP64
        offs, vox, lens = ctx.mlp_paths(d, mode=mode)
        offs, vox, lens = offs.cpu().numpy(), vox.cpu().numpy().view(np.uint32), lens.cpu().numpy()
        exp = oracle.mlp_paths(p, hr, table=table)
        bad = sum(not np.array_equal(vox[offs[k]:offs[k] + lens[k]], e) for k, e in enumerate(exp))
        assert bad == 0, f"{bad} of {len(exp)} voxel lists differ (table={table})"
    ctx.close()


@pytest.mark.parametrize("variant", [1, 2])
def test_robust_drop_variants_match_oracle(ctx_mini, mini_host, variant):
    """the research variants of the update (PCT_DROP_ROBUST: DROP_blocks_red,
    PCT_DROP_ROBUST2: DROP_blocks_robust2 + DROP_update_robust2) against the
    oracle's restatement, on an image that is not 0/1 so that psi_v varies;
    invalid combinations are refused"""
    import oracle
    import pct_amd
    torch = ctx_mini.torch
    p = mini_host["p"]
    assert (p.eta, p.psi_sign) == (2.5, 1)          # the reference's defaults (data.h:369-371)
    hr = mini_host["hr"]
    k = np.arange(mini_host["hull"].size)
    x0 = (mini_host["hull"] * (1.0 + 0.3 * np.sin(0.37 * k))).astype(np.float32)
    flag = pct_amd.PCT_DROP_ROBUST if variant == 1 else pct_amd.PCT_DROP_ROBUST2
    lam = 1e-2
    exp = oracle.drop(p, hr, x0, iterations=2, lam=lam, variant=variant)
    plain = oracle.drop(p, hr, x0, iterations=2, lam=lam)
    d = _to_dev(hr)
    x = torch.from_numpy(x0.copy()).cuda()
    ctx_mini.drop_iteration(d, x, lam=lam, mode=flag)
    ctx_mini.drop_iteration(d, x, lam=lam, mode=flag)
    got = x.cpu().numpy()
    assert np.abs(exp - plain).max() > 100 * np.spacing(np.float32(1.0))   # the variant changes the image
    assert np.max(np.abs(got - exp)) <= 8 * np.spacing(np.float32(2.0))
    assert np.mean(got == exp) > 0.98
    assert np.array_equal(got != x0, exp != x0)
    for bad in (flag | pct_amd.PCT_DROP_F32, flag | pct_amd.PCT_DROP_ATOMIC,
                pct_amd.PCT_DROP_ROBUST | pct_amd.PCT_DROP_ROBUST2):
        with pytest.raises(pct_amd.PctError):
            ctx_mini.drop_iteration(d, x, lam=lam, mode=bad)
