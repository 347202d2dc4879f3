import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def mini_host():
    """The mini scan through the CPU oracle: raw -> intersections -> cylinder
    hull -> endpoints -> compaction (numpy)."""
    import oracle
    import pct_amd
    from pct_amd import configs
    p = configs.params("mini")
    cfg = configs.synth_cfg("mini")
    raw = pct_amd.generate_host(cfg).numpy()
    h_all, missed = oracle.recon_volume_intersections(p, dict(raw))
    h = h_all.select(missed == 0)
    h_pre = h.select(np.ones(h.n, bool))          # copy: endpoints work in place
    hull = oracle.hull_init(p)
    inter = oracle.mlp_endpoints(p, h, hull)
    hr = h.select(inter == 1)
    return dict(p=p, cfg=cfg, raw=raw, h_all=h_all, missed=missed, h_pre=h_pre, h_post=h, hull=hull, inter=inter,
                hr=hr)
