"""GPU: the research loop over given lists (pct_drop_lists_iteration) and the
chord lengths (pct_chord_lengths), through the C ABI, against the oracle:

  - chord lengths bit-exact in both averaging forms;
  - every update rule bit-exact over two iterations whose last block is
    partial (the carried x_update / counts included), on the MLP lists of
    the mini scan with random path errors, for PSI_SIGN = +1 and -1;
  - the same through MLP_paths / MLP_path_error files and the ResearchLoop
    driver, lists in history_sequence order;
  - lists with a voxel outside the image, and offsets not starting at 0,
    are errors.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lists_case(mini_host):
    import oracle
    from pct_amd import configs
    p = configs.params("mini")
    hr = mini_host["hr"].select(np.arange(mini_host["hr"].n) < 1500)
    lists = oracle.mlp_paths(p, hr)
    offsets, vox = oracle.lists_csr(lists)
    rng = np.random.default_rng(21)
    err = rng.normal(0, 0.5, vox.size).astype(np.float32)
    V = p.columns * p.rows * p.slices
    k = np.arange(V)
    x0 = (mini_host["hull"] * (1.0 + 0.3 * np.sin(0.37 * k))).astype(np.float32)
    return dict(p=p, hr=hr, lists=lists, offsets=offsets, vox=vox, err=err, V=V, x0=x0)


def _chords(oracle, p, hr, fn):
    return np.array([fn(p, float(hr.xy_entry_angle[i]), float(hr.xy_exit_angle[i]), float(hr.xz_entry_angle[i]),
                        float(hr.xz_exit_angle[i])) for i in range(hr.n)])


def _dev_hist(hr):
    import pct_amd
    import torch
    d = pct_amd.Histories(hr.n, "cuda")
    for f, a in hr.arrays().items():
        a = np.ascontiguousarray(a)
        getattr(d, f).copy_(torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).cuda())
    return d


def test_chord_lengths_bit_exact(lists_case):
    import oracle
    import pct_amd
    c = lists_case
    ctx = pct_amd.Context(c["p"])
    d = _dev_hist(c["hr"])
    for flt, fn in ((False, oracle.chord_length), (True, oracle.chord_length_float_avg)):
        got = ctx.chord_lengths(d, float_average=flt).cpu().numpy()
        exp = _chords(oracle, c["p"], c["hr"], fn)
        assert np.array_equal(got.view(np.uint64), exp.view(np.uint64)), flt
    ctx.close()


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("psi_sign", [1, -1])
def test_lists_iteration_bit_exact(lists_case, variant, psi_sign):
    import oracle
    import pct_amd
    import torch
    c = lists_case
    p = c["p"].copy()
    p.eta, p.psi_sign = 2.5, psi_sign
    ctx = pct_amd.Context(p)
    V, hr = c["V"], c["hr"]
    chord = _chords(oracle, p, hr, oracle.chord_length_float_avg)
    block, lam = 128, 0.02                              # 1500 = 11 x 128 + 92: a partial last block
    x, xu, cnt = c["x0"].copy(), np.zeros(V), np.zeros(V, np.uint32)
    dev = dict(offsets=torch.from_numpy(c["offsets"]).cuda(), voxels=torch.from_numpy(c["vox"].view(np.int32)).cuda(),
               errors=torch.from_numpy(c["err"]).cuda(), chord=torch.from_numpy(chord).cuda(),
               wepl=torch.from_numpy(np.ascontiguousarray(hr.wepl)).cuda())
    xd = torch.from_numpy(c["x0"].copy()).cuda()
    xud = torch.zeros(V, dtype=torch.float64, device="cuda")
    cntd = torch.zeros(V, dtype=torch.int32, device="cuda")
    for it in range(2):
        oracle.drop_lists_research(V, c["offsets"], c["vox"], c["err"], chord, hr.wepl, x, xu, cnt, block, lam,
                                   2.5, psi_sign, variant)
        rule = variant | (pct_amd.PCT_LISTS_REUSE if it > 0 else 0)   # the sorted order of call 1 reused
        ctx.drop_lists_iteration(dev["offsets"], dev["voxels"], dev["errors"], dev["chord"], dev["wepl"], xd, xud, cntd,
                                 block, lam, rule)
        assert np.array_equal(xd.cpu().numpy().view(np.uint32), x.view(np.uint32)), f"iteration {it}"
        assert np.array_equal(xud.cpu().numpy().view(np.uint64), xu.view(np.uint64))
        assert np.array_equal(cntd.cpu().numpy().view(np.uint32), cnt)
    assert np.sum(x != c["x0"]) > 1000 and cnt.sum() > 0
    ctx.close()


def test_research_loop_from_files(lists_case, tmp_path):
    import oracle
    import pct_amd
    import torch
    from pct_amd import research
    c = lists_case
    p, hr, V = c["p"], c["hr"], c["V"]
    seq = research.history_sequence(hr.n)
    assert np.array_equal(np.sort(seq), np.arange(hr.n)) and seq[1] == research.PRIME_OFFSET % hr.n
    lists = [c["lists"][i] for i in seq]
    err = np.concatenate([c["err"][c["offsets"][i]:c["offsets"][i + 1]] for i in seq])
    d = str(tmp_path) + "/"
    pf, ef = research.paths_file_name(d, "out", 1), research.paths_file_name(d, "out", 1, "MLP_path_error")
    import os
    os.makedirs(d + "out")
    research.write_mlp_paths(pf, lists)
    research.write_path_errors(ef, err)
    offsets, vox = research.read_mlp_paths(pf)
    assert len(offsets) == hr.n + 1 and np.array_equal(vox, np.concatenate(lists))
    errors = research.read_path_errors(ef, offsets[-1])
    assert np.array_equal(errors, err)
    with open(pf, "rb") as f:
        assert f.readline() == b"%d\n" % hr.n
    ctx = pct_amd.Context(p)
    chord = ctx.chord_lengths(_dev_hist(hr), float_average=True).cpu().numpy()[seq]
    wepl = np.ascontiguousarray(hr.wepl[seq])
    loop = research.ResearchLoop(ctx, offsets, vox, wepl, chord, errors)
    xd = torch.from_numpy(c["x0"].copy()).cuda()
    x, xu, cnt = c["x0"].copy(), np.zeros(V), np.zeros(V, np.uint32)
    for _ in range(2):
        loop.iterate(xd, block_size=200, lam=0.01)
        oracle.drop_lists_research(V, offsets, vox, errors, chord, wepl, x, xu, cnt, 200, 0.01, p.eta, p.psi_sign, 2)
    assert np.array_equal(xd.cpu().numpy().view(np.uint32), x.view(np.uint32))
    # the reader stops at max_paths
    o2, v2 = research.read_mlp_paths(pf, max_paths=10)
    assert len(o2) == 11 and np.array_equal(v2, np.concatenate(lists[:10]))
    ctx.close()


def test_lists_iteration_errors(lists_case):
    import pct_amd
    import torch
    c = lists_case
    ctx = pct_amd.Context(c["p"])
    V = c["V"]
    offsets = torch.tensor([0, 3, 5], dtype=torch.int64, device="cuda")
    voxels = torch.tensor([1, 2, V + 7, 4, 5], dtype=torch.int32, device="cuda")
    chord = torch.full((2,), 0.3, dtype=torch.float64, device="cuda")
    wepl = torch.ones(2, dtype=torch.float32, device="cuda")
    x = torch.ones(V, dtype=torch.float32, device="cuda")
    xu = torch.zeros(V, dtype=torch.float64, device="cuda")
    cnt = torch.zeros(V, dtype=torch.int32, device="cuda")
    with pytest.raises(pct_amd.PctError):
        ctx.drop_lists_iteration(offsets, voxels, None, chord, wepl, x, xu, cnt, 2, 0.01, pct_amd.PCT_LISTS_DROP)
    with pytest.raises(pct_amd.PctError):               # error variants need the path errors
        ctx.drop_lists_iteration(offsets, voxels, None, chord, wepl, x, xu, cnt, 2, 0.01, pct_amd.PCT_LISTS_ERROR)
    with pytest.raises(pct_amd.PctError):
        ctx.drop_lists_iteration(offsets[1:].contiguous(), voxels, None, chord, wepl, x, xu, cnt, 2, 0.01, 0)
    # the context stays usable
    voxels[2] = 3
    x.fill_(1.0); xu.zero_(); cnt.zero_()
    ctx.drop_lists_iteration(offsets, voxels, None, chord, wepl, x, xu, cnt, 2, 0.01, pct_amd.PCT_LISTS_DROP)
    assert int(cnt.sum()) == 0 and float((x != 1).sum()) == 5
    ctx.close()
