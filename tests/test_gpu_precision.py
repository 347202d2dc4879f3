"""GPU parity of the precision form (PCT_MLP_TABLE, PCT_DROP_F32) through the
C ABI:

  - the tabulated-MLP voxel lists are bit-identical to the oracle's precision
    form (same FP64 tables, same FP32 stepping, no contraction on either side);
  - the FP32/Kahan DROP update on exact lists stays within FP32 rounding of the
    oracle's FP64 DROP;
  - the whole precision form (lists + FP32 update) stays within the stated RMSE
    of the FP64 oracle image over several iterations;
  - switching forms on one context rebuilds the plan (the FP64 form afterwards
    is again bit-identical to its own earlier result).
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


def _rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


def test_precision_lists_bit_exact(ctx_mini, mini_host):
    import oracle
    import pct_amd
    hr = mini_host["hr"]
    d = _to_dev(hr)
    offs, vox, lens = ctx_mini.mlp_paths(d, mode=pct_amd.PCT_MLP_TABLE)
    offs, vox, lens = offs.cpu().numpy(), vox.cpu().numpy().view(np.uint32), lens.cpu().numpy()
    exp = oracle.mlp_paths(mini_host["p"], hr, table=True)
    bad = sum(1 for k, e in enumerate(exp) if not np.array_equal(vox[offs[k]:offs[k] + lens[k]], e))
    assert bad == 0, f"{bad} of {len(exp)} voxel lists differ"


def test_f32_update_matches_fp64_oracle(ctx_mini, mini_host):
    import oracle
    import pct_amd
    torch = ctx_mini.torch
    p, hr = mini_host["p"], mini_host["hr"]
    x0 = mini_host["hull"].astype(np.float32)
    lam = 1e-2
    exp = oracle.drop(p, hr, x0, iterations=2, lam=lam)
    x = torch.from_numpy(x0.copy()).cuda()
    d = _to_dev(hr)
    for _ in range(2):
        ctx_mini.drop_iteration(d, x, lam=lam, mode=pct_amd.PCT_MLP_FP64 | pct_amd.PCT_DROP_F32)
    got = x.cpu().numpy()
    assert _rmse(exp, x0) > 1e-3
    # FP32 residual and deltas: a few ulps of |x| <= 2 per iteration
    assert np.max(np.abs(got - exp)) <= 16 * np.spacing(np.float32(2.0))
    assert np.array_equal(got != x0, exp != x0)


def test_precision_form_rmse_vs_fp64_oracle(ctx_mini, mini_host):
    import oracle
    import pct_amd
    torch = ctx_mini.torch
    p, hr = mini_host["p"], mini_host["hr"]
    x0 = mini_host["hull"].astype(np.float32)
    lam = 1e-2
    d = _to_dev(hr)
    x = torch.from_numpy(x0.copy()).cuda()
    for it in (1, 2, 3):
        ctx_mini.drop_iteration(d, x, lam=lam, mode=pct_amd.PCT_PRECISION)
        exp = oracle.drop(p, hr, x0, iterations=it, lam=lam)
        got = x.cpu().numpy()
        touched = exp != x0
        rel_rmse = _rmse(got[touched], exp[touched]) / float(np.mean(np.abs(exp[touched])))
        assert rel_rmse <= 2e-3, (it, rel_rmse)      # the stated bound: 0.2 % of the image
        assert np.max(np.abs(got - exp)) < 0.05


def test_switching_forms_rebuilds_the_plan(ctx_mini, mini_host):
    import pct_amd
    torch = ctx_mini.torch
    d = _to_dev(mini_host["hr"])
    x0 = mini_host["hull"].astype(np.float32)
    res = []
    for mode in (pct_amd.PCT_MLP_FP64, pct_amd.PCT_PRECISION, pct_amd.PCT_MLP_FP64):
        x = torch.from_numpy(x0.copy()).cuda()
        ctx_mini.drop_iteration(d, x, lam=1e-2, mode=mode)
        ctx_mini.drop_iteration(d, x, lam=1e-2, mode=mode)
        res.append(x.cpu().numpy())
    assert np.array_equal(res[0], res[2])
    assert not np.array_equal(res[0], res[1])
