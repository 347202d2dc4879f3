"""Pins the oracle's binning, statistics, 3-sigma cuts and sinogram
(.cu:2507-2918) against numpy on the same histories:

  - bin numbers from the chord midpoint and path angle (numpy float64 with the
    reference's float midpoints);
  - counts = number of in-range kept histories per bin; sum of counts check;
  - means and sample standard deviations (N-1) vs numpy float64 to float
    precision; a singleton bin keeps deviation 0 and its member is always cut
    (strict angle test);
  - the sinogram is the mean WEPL of the survivors of each bin;
  - the out-of-range quirk: such histories carry bin 0 and are not counted.
"""
import math

import numpy as np
import pytest

import oracle
import pct_amd
from pct_amd import configs


@pytest.fixture(scope="module")
def binned():
    p = configs.params("cfg1")
    cfg = configs.synth_cfg("cfg1", n_histories=400_000)
    raw = pct_amd.generate_host(cfg).numpy()
    h, missed = oracle.recon_volume_intersections(p, raw)
    r = oracle.bin_stats_and_cuts(p, h, missed)
    return p, h, missed, r


def _numpy_bins(p, h):
    f = np.float32
    xm = ((h.x_entry + h.x_exit) / f(2)).astype(np.float64)
    ym = ((h.y_entry + h.y_exit) / f(2)).astype(np.float64)
    zm = ((h.z_entry + h.z_exit) / f(2)).astype(np.float64)
    ang = np.arctan2((h.y_exit - h.y_entry).astype(np.float32), (h.x_exit - h.x_entry).astype(np.float32))
    ang = ang.astype(np.float32).astype(np.float64)
    ang = np.where(ang < 0, ang + 2 * math.pi, ang)
    ab = ((ang * (180.0 / math.pi) / p.angular_bin_size) + 0.5).astype(np.int64) % p.angular_bins
    a = ab * p.angular_bin_size * (math.pi / 180.0)
    t = ym * np.cos(a) - xm * np.sin(a)
    tb = np.trunc(t / p.t_bin_size + p.t_bins // 2).astype(np.int64)
    vb = np.trunc(zm / p.v_bin_size + p.v_bins // 2).astype(np.int64)
    ok = (tb >= 0) & (vb >= 0) & (tb < p.t_bins) & (vb < p.v_bins)
    return np.where(ok, tb + ab * p.t_bins + vb * p.t_bins * p.angular_bins, 0), ok


def test_bin_numbers(binned):
    p, h, missed, r = binned
    b, ok = _numpy_bins(p, h)
    kept = missed == 0
    agree = (b == h.bin_num) & (ok == r["in_range"].astype(bool))
    # numpy's atan2 may differ from the shared one by an ulp at a bin edge
    assert agree[kept].mean() > 0.9999
    assert np.all(h.bin_num[~r["in_range"].astype(bool)] == 0)


def _rel(h):
    def wrap(a):
        a = np.where(a > np.pi, a - 2 * np.pi, a)
        return np.where(a < -np.pi, a + 2 * np.pi, a)
    return (wrap((h.xy_exit_angle - h.xy_entry_angle).astype(np.float64)),
            wrap((h.xz_exit_angle - h.xz_entry_angle).astype(np.float64)))


def test_counts_means_stddevs(binned):
    p, h, missed, r = binned
    kept = missed == 0
    inr = r["in_range"].astype(bool)
    stat = kept & inr
    nb = p.t_bins * p.v_bins * p.angular_bins
    cnt = np.bincount(h.bin_num[stat], minlength=nb)
    np.testing.assert_array_equal(cnt, r["counts"])
    assert cnt.sum() == stat.sum()
    ut, uv = _rel(h)
    s_w = np.bincount(h.bin_num[stat], weights=h.wepl[stat].astype(np.float64), minlength=nb)
    s_u = np.bincount(h.bin_num[stat], weights=ut[stat], minlength=nb)
    nz = cnt > 0
    np.testing.assert_allclose(r["mean_wepl"][nz], s_w[nz] / cnt[nz], rtol=2e-5, atol=1e-6)
    np.testing.assert_allclose(r["mean_rel_ut"][nz], s_u[nz] / cnt[nz], rtol=1e-3, atol=1e-6)
    # sample standard deviation over bins with >= 2 members and no bin-0 quirk
    multi = np.where(cnt >= 2)[0]
    multi = multi[multi != 0][:3000]
    order = np.argsort(h.bin_num[stat], kind="stable")
    bins_sorted = h.bin_num[stat][order]
    w_sorted = h.wepl[stat][order].astype(np.float64)
    starts = np.searchsorted(bins_sorted, multi)
    ends = np.searchsorted(bins_sorted, multi, side="right")
    sd = np.array([np.std(w_sorted[a:e], ddof=1) for a, e in zip(starts, ends)])
    np.testing.assert_allclose(r["stddev_wepl"][multi], sd, rtol=2e-3, atol=2e-4)


def test_singletons_are_cut_and_sinogram_is_survivor_mean(binned):
    p, h, missed, r = binned
    kept = missed == 0
    inr = r["in_range"].astype(bool)
    cnt = r["counts"]
    single = kept & inr & (cnt[h.bin_num] == 1) & (h.bin_num != 0)
    assert single.sum() > 10
    assert np.all(r["failed"][single] == 1)
    assert np.all(r["failed"][~kept] == 1)
    surv = r["failed"] == 0
    nb = cnt.size
    c2 = np.bincount(h.bin_num[surv], minlength=nb)
    np.testing.assert_array_equal(c2, r["cut_counts"])
    s = np.bincount(h.bin_num[surv], weights=h.wepl[surv].astype(np.float64), minlength=nb)
    nz = c2 > 0
    np.testing.assert_allclose(r["sinogram"][nz], s[nz] / c2[nz], rtol=2e-5)
    assert np.all(r["sinogram"][~nz] == 0)
    # cuts keep most histories of well-populated bins
    busy = kept & inr & (cnt[h.bin_num] >= 8)
    assert r["failed"][busy].mean() < 0.2


def _one_bin(n, wepl, ut, uv):
    h = oracle.HostHistories(n)
    for k in range(n):
        h.x_entry[k], h.y_entry[k], h.z_entry[k] = -9.0, 0.05, 0.1
        h.x_exit[k], h.y_exit[k], h.z_exit[k] = 9.0, 0.05, 0.1
    h.xy_entry_angle[:] = 0.0
    h.xz_entry_angle[:] = 0.0
    h.xy_exit_angle[:] = ut
    h.xz_exit_angle[:] = uv
    h.wepl[:] = wepl
    return h


def test_cut_rule_on_hand_built_bins():
    p = configs.params("mini")
    rng = np.random.default_rng(5)
    # identical vertical angles: sigma 0 and the strict angle test cuts all
    h = _one_bin(4, [20.0, 20.0, 20.0, 26.0], [0.001, 0.002, 0.001, 0.002], 0.0)
    r = oracle.bin_stats_and_cuts(p, h, np.zeros(4, np.uint8))
    assert len(set(h.bin_num.tolist())) == 1 and r["counts"][h.bin_num[0]] == 4
    assert r["mean_wepl"][h.bin_num[0]] == pytest.approx(21.5)
    assert r["stddev_wepl"][h.bin_num[0]] == pytest.approx(3.0)        # sample sd, N-1
    assert r["failed"].tolist() == [1, 1, 1, 1]
    # 30 members near 20 cm and one at 30 cm: only the WEPL outlier is cut
    n = 31
    w = np.concatenate([20.0 + 0.1 * rng.standard_normal(30), [30.0]])
    ut = 0.002 * rng.standard_normal(n)
    uv = 0.002 * rng.standard_normal(n)
    h = _one_bin(n, w, ut, uv)
    r = oracle.bin_stats_and_cuts(p, h, np.zeros(n, np.uint8))
    assert r["failed"][-1] == 1
    assert r["failed"][:-1].sum() <= 1                                # angles: ~0.3 % beyond 3 sigma
    bn = h.bin_num[0]
    surv = r["failed"] == 0
    assert r["sinogram"][bn] == pytest.approx(np.mean(h.wepl[surv].astype(np.float64)), rel=1e-6)
