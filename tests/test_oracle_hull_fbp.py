"""Pins the oracle's FBP, hull and initial-iterate stages:

  - Shepp-Logan filter (.cu:3036-3069): a delta sinogram reproduces one
    weighted kernel row exactly; the kernel's closed form at s=0 and s=1;
    linearity;
  - backprojection (.cu:3155-3230): a vectorised numpy evaluation of the same
    formula on a small grid (float64, no per-angle float rounding);
  - SC/MSC voxel walk (.cu:3338-3558): the walk's voxels equal the voxels a
    dense sampling of the entry->exit segment passes through;
  - MSC edge detection / hull selection / averaging filter / FBP threshold /
    x0 = FBP * hull against numpy;
  - SM (.cu:3598-3789): only non-missed histories with WEPL >= 6 cm are
    counted, each by the walk above; the edge detection against a vectorised
    numpy restatement (3x3 differences, np.argmax's first maximum per slice,
    scaled threshold, rim), on counts with many ties, a flat slice and a
    non-unit scale; taking the last maximum instead of the first would change
    the thresholds of that data.
"""
import math

import numpy as np
import pytest

import oracle
from pct_amd import configs


def _p_small(c=16, s=2):
    p = configs.params("mini")
    p.columns = p.rows = c
    p.slices = s
    p.recon_cyl_height = s * p.slice_thickness
    return p


def test_filter_delta_and_kernel():
    p = configs.params("mini")
    T, V, A = p.t_bins, p.v_bins, p.angular_bins
    sino = np.zeros(T * V * A, np.float32)
    v_bin, a_bin, t0 = 7, 33, 120
    strip = v_bin * A * T + a_bin * T
    sino[strip + t0] = 2.5
    out = oracle.fbp_filter(p, sino)
    ts, sr = p.t_bin_size, p.source_radius
    v = (v_bin - V // 2) * p.v_bin_size + p.v_bin_size / 2.0
    t = (t0 - T // 2) * ts + ts / 2.0
    scale = sr / math.sqrt(sr * sr + t * t + v * v)
    sep = np.arange(T) - t0
    kern = 1 / ((ts * math.pi * ts * math.pi) * (1.0 - 4.0 * sep * sep))
    exp = (ts * 2.5 * kern * scale).astype(np.float32)
    np.testing.assert_array_equal(out[strip:strip + T], exp)
    # all other rows stay zero
    assert np.count_nonzero(out) == np.count_nonzero(exp)
    k0 = 1 / ((ts * math.pi) ** 2)
    assert kern[t0] == pytest.approx(k0) and kern[t0 + 1] == pytest.approx(-k0 / 3)


def test_filter_linearity():
    p = configs.params("mini")
    n = p.t_bins * p.v_bins * p.angular_bins
    rng = np.random.default_rng(2)
    s1 = rng.uniform(0, 20, n).astype(np.float32)
    s2 = rng.uniform(0, 20, n).astype(np.float32)
    f = oracle.fbp_filter(p, (2 * s1 + s2).astype(np.float32))
    g = 2 * oracle.fbp_filter(p, s1).astype(np.float64) + oracle.fbp_filter(p, s2)
    np.testing.assert_allclose(f, g, rtol=1e-4, atol=1e-2)


def test_backprojection_formula():
    p = _p_small(16, 2)
    T, V, A = p.t_bins, p.v_bins, p.angular_bins
    rng = np.random.default_rng(3)
    sf = rng.normal(0, 1, T * V * A).astype(np.float32)
    img = oracle.backprojection(p, sf).reshape(p.slices, p.rows, p.columns)
    R, sr = p.recon_cyl_radius, p.source_radius
    w = 2 * R / p.columns
    delta = math.radians(p.gantry_angle_interval)
    for s in range(p.slices):
        for r in range(p.rows):
            for c in range(p.columns):
                x = -R + (c + 0.5) * w
                y = R - (r + 0.5) * w
                z = -p.recon_cyl_height / 2 + (s + 0.5) * p.slice_thickness
                if x * x + y * y > R * R:
                    assert img[s, r, c] == np.float32(p.rsp_air)
                    continue
                acc = 0.0
                for a in range(A):
                    ca, sa = math.cos(a * delta), math.sin(a * delta)
                    u, t = x * ca + y * sa, -x * sa + y * ca
                    dnt = (t - u * (t / (sr + u))) / p.t_bin_size + T / 2.0
                    dnv = (z - u * (z / (sr + u))) / p.v_bin_size + V / 2.0
                    tb, vb = math.floor(dnt), math.floor(dnv)
                    eta, eps = dnt - tb, dnv - vb
                    b = tb + a * T + vb * A * T
                    f = lambda k: float(sf[k]) if 0 <= k < sf.size else 0.0  # noqa: E731
                    if vb == V - 1:
                        val = (1 - eta) * f(b) + eta * f(b + 1)
                    else:
                        val = ((1 - eta) * (1 - eps) * f(b) + eta * (1 - eps) * f(b + 1)
                               + (1 - eta) * eps * f(b + A * T) + eta * eps * f(b + A * T + 1))
                    acc += (sr / (sr + u)) ** 2 * val
                assert img[s, r, c] == pytest.approx(acc * delta, rel=1e-4, abs=1e-4)


def _walk_hist(points):
    h = oracle.HostHistories(len(points))
    for k, (a, b) in enumerate(points):
        h.x_entry[k], h.y_entry[k], h.z_entry[k] = a
        h.x_exit[k], h.y_exit[k], h.z_exit[k] = b
    h.wepl[:] = 0.0
    return h


def _dense_voxels(p, a, b, n=200_000):
    a = np.array(a, np.float32).astype(np.float64)
    b = np.array(b, np.float32).astype(np.float64)
    s = np.linspace(0, 1, n)[:, None]
    pts = a + s * (b - a)
    w = 2 * p.recon_cyl_radius / p.columns
    vx = (np.abs(pts[:, 0] + p.recon_cyl_radius) / w).astype(int)
    vy = (np.abs(pts[:, 1] - p.recon_cyl_radius) / w).astype(int)
    vz = (np.abs(pts[:, 2] - p.recon_cyl_height / 2) / p.slice_thickness).astype(int)
    v = vx + vy * p.columns + vz * p.columns * p.rows
    out_v = v[-1]
    first_out = np.argmax(v == out_v)
    return set(np.unique(v[:first_out]).tolist())


def test_carve_walk_matches_dense_sampling():
    p = configs.params("mini")
    rng = np.random.default_rng(8)
    pts = []
    for _ in range(60):
        th = rng.uniform(0, 2 * np.pi)
        d = rng.uniform(-7, 7)
        zz = rng.uniform(-0.9, 0.9, 2)
        ca, sa = np.cos(th), np.sin(th)
        L = np.sqrt(100 - d * d) * 0.95
        a = (d * -sa - L * ca, d * ca - L * sa, zz[0])
        b = (d * -sa + L * ca, d * ca + L * sa, zz[1])
        pts.append((a, b))
    h = _walk_hist(pts)
    base = oracle.hull_init(p).astype(np.int32)
    exact = 0
    for k in range(len(pts)):
        one = h.select(np.arange(h.n) == k)
        counts = oracle.msc_counts(p, one, np.zeros(1, np.uint8))
        walked = set(np.nonzero(counts - base)[0].tolist())
        dense = _dense_voxels(p, pts[k][0], pts[k][1])
        diff = walked ^ dense
        assert len(diff) <= 2, (k, sorted(diff))
        exact += len(diff) == 0
        # SC clears exactly the walked voxels
        carved = oracle.space_carve(p, one, np.zeros(1, np.uint8), oracle.hull_init(p))
        assert set(np.nonzero(oracle.hull_init(p) - carved)[0].tolist()) == walked & set(np.nonzero(base)[0].tolist())
    assert exact >= 0.9 * len(pts)


def test_msc_only_counts_air_histories():
    p = configs.params("mini")
    h = _walk_hist([((-9.0, 0.3, 0.1), (9.0, 0.4, 0.1))] * 3)
    h.wepl[:] = [0.0, 5.0, 0.0]
    missed = np.array([0, 0, 1], np.uint8)
    c = oracle.msc_counts(p, h, missed) - oracle.hull_init(p)
    one = oracle.msc_counts(p, h.select(np.array([True, False, False])), np.zeros(1, np.uint8)) - oracle.hull_init(p)
    np.testing.assert_array_equal(c, one)


def test_edge_select_threshold_x0_numpy():
    p = _p_small(24, 3)
    C, R, S = p.columns, p.rows, p.slices
    rng = np.random.default_rng(4)
    counts = rng.integers(0, 200, C * R * S).astype(np.int32)
    edge = oracle.msc_edge(p, counts).reshape(S, R, C)
    cnt = counts.reshape(S, R, C)
    w = 2 * p.recon_cyl_radius / C
    exp = np.ones((S, R, C), np.int32)
    for s in range(S):
        for r in range(R):
            for c in range(C):
                md = 0
                if 0 < r < R - 1 and 0 < c < C - 1:
                    md = max(0, int((cnt[s, r, c] - cnt[s, r - 1:r + 2, c - 1:c + 2]).max()))
                v = 0 if md > p.msc_diff_thresh else 1
                x = np.float32((c - C // 2 + 0.5) * w)
                y = np.float32((R // 2 - r - 0.5) * w)
                if float(np.float32(x * x + y * y)) >= (p.recon_cyl_radius - w / 2) ** 2:
                    v = 0
                exp[s, r, c] = v
    np.testing.assert_array_equal(edge, exp)
    # hull selection: counts != 0, then the clipped 3x3 box average > 0.1
    hull = oracle.hull_select(p, (counts % 3).astype(np.int32)).reshape(S, R, C)
    nz = (counts % 3 != 0).reshape(S, R, C).astype(np.float64)
    for s in range(S):
        for r in range(R):
            for c in range(C):
                win = nz[s, max(r - 1, 0):min(r + 1, R - 1) + 1, max(c - 1, 0):min(c + 1, C - 1) + 1]
                assert hull[s, r, c] == (win.sum() > win.size * p.hull_filter_threshold)
    img = rng.uniform(0, 1.5, C * R * S).astype(np.float32)
    fh = oracle.fbp_hull(p, img).reshape(S, R, C)
    cyl = oracle.hull_init(p).reshape(S, R, C).astype(bool)
    np.testing.assert_array_equal(fh, ((img.reshape(S, R, C) > p.fbp_threshold) & cyl).astype(np.uint8))
    x0 = oracle.initial_iterate(p, img, hull.reshape(-1))
    np.testing.assert_array_equal(x0, img * hull.reshape(-1).astype(np.float32))


def test_sm_counts_only_crossing_histories():
    """SM_GPU (.cu:3615): non-missed histories with WEPL >= SM_LOWER_THRESHOLD;
    NaN and 5.9 cm are not counted, 6.0 is"""
    p = configs.params("mini")
    assert p.sm_lower_threshold == 6.0 and p.sm_scale_threshold == 1.0
    pts = [((-9.0, 0.3, 0.1), (9.0, 0.4, 0.1)), ((0.2, -9.0, 0.5), (0.3, 9.0, -0.4)),
           ((-8.0, -6.0, 1.0), (7.0, 6.5, -1.0)), ((-9.0, 2.0, 0.0), (9.0, -2.0, 0.0)),
           ((-6.0, 7.0, 0.2), (6.0, -7.0, 0.3))]
    h = _walk_hist(pts)
    h.wepl[:] = [6.0, 5.9, 20.0, float("nan"), 8.0]
    missed = np.array([0, 0, 0, 0, 1], np.uint8)
    got = oracle.sm_counts(p, h, missed) - oracle.hull_init(p)
    # the same walks counted by the MSC code with a threshold that admits all
    q = configs.params("mini")
    q.msc_threshold = 1e30
    sel = np.array([True, False, True, False, False])
    exp = oracle.msc_counts(q, h.select(sel), np.zeros(2, np.uint8)) - oracle.hull_init(q)
    np.testing.assert_array_equal(got, exp)
    assert got.sum() > 30


def _sm_edge_numpy(p, counts, last=False):
    C, R, S = p.columns, p.rows, p.slices
    cnt = counts.reshape(S, R, C).astype(np.int64)
    diff = np.zeros_like(cnt)
    core = cnt[:, 1:-1, 1:-1]
    for dr in (-1, 0, 1):
        for dc in (-1, 0, 1):
            nb = cnt[:, 1 + dr:R - 1 + dr, 1 + dc:C - 1 + dc]
            diff[:, 1:-1, 1:-1] = np.maximum(diff[:, 1:-1, 1:-1], core - nb)
    flat = diff.reshape(S, -1)
    if last:
        arg = flat.shape[1] - 1 - np.argmax(flat[:, ::-1], axis=1)
    else:
        arg = np.argmax(flat, axis=1)
    thr = np.where(flat.max(axis=1) > 0, cnt.reshape(S, -1)[np.arange(S), arg], 0)
    out = (cnt > p.sm_scale_threshold * thr[:, None, None]).astype(np.int32)
    w = 2 * p.recon_cyl_radius / C
    x = ((np.arange(C) - C // 2 + 0.5) * w).astype(np.float32)
    y = ((R // 2 - np.arange(R) - 0.5) * w).astype(np.float32)
    s2 = (x[None, :] * x[None, :] + y[:, None] * y[:, None]).astype(np.float32)
    rim = s2.astype(np.float64) >= (p.recon_cyl_radius - w / 2) ** 2
    out[:, rim] = 0
    return out.reshape(-1), thr.astype(np.int32)


@pytest.mark.parametrize("scale", [1.0, 0.5])
def test_sm_edge_numpy(scale):
    p = _p_small(24, 4)
    p.sm_scale_threshold = scale
    C, R, S = p.columns, p.rows, p.slices
    rng = np.random.default_rng(7)
    counts = rng.integers(0, 6, C * R * S).astype(np.int32)
    counts.reshape(S, R, C)[2] = 3                        # a flat slice: no positive difference
    tie = counts.reshape(S, R, C)[1]                      # two pixels with the same largest drop (8)
    tie[:] = 10                                           # but different counts: 18 first, 25 later
    tie[3, 3] = 18
    tie[14:17, 14:17] = 17
    tie[15, 15] = 25
    counts.reshape(S, R, C)[3, 5:9, 5:9] += 40           # one clear step
    out, thr = oracle.sm_edge(p, counts)
    exp, exp_thr = _sm_edge_numpy(p, counts)
    np.testing.assert_array_equal(thr, exp_thr)
    np.testing.assert_array_equal(out, exp)
    assert thr[2] == 0 and out.reshape(S, R, C)[2].sum() > 0 and thr[1] == 18
    # the ties make the first maximum matter
    _, last_thr = _sm_edge_numpy(p, counts, last=True)
    assert not np.array_equal(last_thr, exp_thr)


def test_sm_hull_of_a_scan():
    """on the synthetic scan the SM hull covers the object: the pipeline's
    SM hull overlaps the MSC hull in most voxels"""
    from pct_amd import configs as cf
    import pct_amd
    p = cf.params("mini")
    raw = pct_amd.generate_host(cf.synth_cfg("mini", n_histories=60_000)).numpy()
    _, _, sm, _ = oracle.preprocess(p, raw, hull_kind="sm")
    _, _, msc, _ = oracle.preprocess(p, raw, hull_kind="msc")
    assert sm.sum() > 0.3 * msc.sum()
    assert (sm & msc).sum() > 0.8 * min(sm.sum(), msc.sum())
