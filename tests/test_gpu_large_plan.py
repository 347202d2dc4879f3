"""GPU: a DROP plan whose list store holds more than 2^32 entries (one
dispatch cannot cover it with one work-item per entry) is built correctly,
also when it is rebuilt on the buffers of an earlier plan: the image after a
rebuild (block size changed and changed back) equals the first one bit for
bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_rebuilt_plan_beyond_2_32_entries():
    import pct_amd
    from pct_amd import configs, learned
    torch = pytest.importorskip("torch")
    ctx = pct_amd.Context(configs.params("cfg2"))
    hist = learned.scan_histories(ctx, configs.synth_cfg("cfg2", n_histories=32_000_000))
    x0 = ctx.initial_iterate(torch.ones(ctx.V, dtype=torch.float32, device=ctx.device), ctx.hull_init())
    E, _ = pct_amd.drop_stats(ctx, hist, 3200)
    assert E > 2 ** 32
    first = x0.clone()
    ctx.drop_iteration(hist, first, 3200, 1e-3)
    other = x0.clone()
    ctx.drop_iteration(hist, other, 6400, 1e-3)           # plan rebuilt on the same buffers
    again = x0.clone()
    ctx.drop_iteration(hist, again, 3200, 1e-3)           # and once more
    a, b = first.cpu().numpy(), again.cpu().numpy()
    assert np.any(a != x0.cpu().numpy())
    assert np.array_equal(a, b), f"{np.count_nonzero(a != b)} voxels differ after the rebuild"
    ctx.close()
