"""Pins the oracle of the learned-path form (oracle/oracle_net.cpp, oracle/net.py)
and the fixed tables it shares with the kernels (pct_amd/csrc/pct_net_basis.h):

  - the per-segment Hermite and clamped-spline cubics equal scipy's clamped
    CubicSpline and the Hermite polynomials, and the committed header is what
    its generator writes;
  - the exact MLP at an arbitrary depth reproduces the MLP loop's own steps
    bit for bit at the step depths (and is the same smooth curve between them);
  - inputs and knot targets against an independent numpy restatement of the
    geometry (rotation into the entry frame with the u_in > u_out flip);
  - the FP32 surrogate walk follows the FP64 Hermite + spline path to FP32
    rounding, with zero offsets the Hermite path, with the exact knot offsets
    the MLP to the spline's own error;
  - the numpy network: backward = finite differences of the forward (FP64),
    BF16 forward close to FP64, Adam's first step = lr * sign(g);
  - the calibration cell of a history.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import net as onet
from pct_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _recon_sample(preset, n):
    import pct_amd
    p = configs.params(preset)
    raw = pct_amd.generate_host(configs.synth_cfg(preset, n_histories=n)).numpy()
    h_all, missed = oracle.recon_volume_intersections(p, dict(raw))
    h = h_all.select(missed == 0)
    inter = oracle.mlp_endpoints(p, h, oracle.hull_init(p))
    return p, h.select(inter == 1)


@pytest.fixture(scope="module")
def cfg2_sample():
    return _recon_sample("cfg2", 6000)


def _entry_frame(h, i):
    """independent restatement: (u_in, t_in, t_out, v_in, v_out, dT, dV, u2)"""
    a_in, a_out = float(h.xy_entry_angle[i]), float(h.xy_exit_angle[i])
    xi, yi, xo, yo = (float(h.x_entry[i]), float(h.y_entry[i]), float(h.x_exit[i]), float(h.y_exit[i]))
    ca, sa = np.cos(a_in), np.sin(a_in)
    u_in, u_out = ca * xi + sa * yi, ca * xo + sa * yo
    if u_in > u_out:
        a_in += np.pi
        a_out += np.pi
        ca, sa = np.cos(a_in), np.sin(a_in)
        u_in, u_out = ca * xi + sa * yi, ca * xo + sa * yo
    t_in, t_out = ca * yi - sa * xi, ca * yo - sa * xo
    return dict(ca=ca, sa=sa, u_in=u_in, t_in=t_in, t_out=t_out, v_in=float(h.z_entry[i]), v_out=float(h.z_exit[i]),
                dT=a_out - a_in, dV=float(h.xz_exit_angle[i]) - float(h.xz_entry_angle[i]), u2=abs(u_out - u_in))


def test_basis_tables_match_scipy_and_header():
    import importlib.util
    from scipy.interpolate import CubicSpline
    spec = importlib.util.spec_from_file_location("gen", os.path.join(ROOT, "pct_amd", "csrc", "gen_net_basis.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    with open(os.path.join(ROOT, "pct_amd", "csrc", "pct_net_basis.h")) as f:
        assert f.read() == gen.render(), "pct_net_basis.h is stale: run gen_net_basis.py"
    h01, h11, phi = gen.tables()
    kn = np.linspace(0, 1, 6)
    sig = np.linspace(0, 0.999999, 997)
    seg = np.minimum((sig * 5).astype(int), 4)
    tau = sig * 5 - seg

    def ev(tab):
        c = np.array([[float(x) for x in row] for row in tab])[seg]
        return c[:, 0] + tau * (c[:, 1] + tau * (c[:, 2] + tau * c[:, 3]))
    np.testing.assert_allclose(ev(h01), 3 * sig ** 2 - 2 * sig ** 3, atol=1e-14)
    np.testing.assert_allclose(ev(h11), sig ** 3 - sig ** 2, atol=1e-14)
    for j in range(4):
        y = np.zeros(6)
        y[j + 1] = 1.0
        ref = CubicSpline(kn, y, bc_type="clamped")(sig)
        np.testing.assert_allclose(ev(phi[j]), ref, atol=1e-12)
        # exactly: 1 at its knot, 0 at the other knots, zero slope at both ends
        assert [phi[j][k][0] for k in range(5)] == [1 if k == j + 1 else 0 for k in range(5)]
        assert sum(phi[j][4]) == 0
        assert phi[j][0][1] == 0 and phi[j][4][1] + 2 * phi[j][4][2] + 3 * phi[j][4][3] == 0


def test_exact_points_reproduce_the_mlp_steps(cfg2_sample):
    from scipy.interpolate import CubicSpline
    p, h = cfg2_sample
    for i in range(0, 300, 17):
        tr = oracle.mlp_trace(p, h, i)
        t, v = oracle.mlp_points(p, h, i, tr["u1"])
        np.testing.assert_array_equal(t, tr["t1"])
        np.testing.assert_array_equal(v, tr["v1"])
        # between the steps: the same smooth curve
        mid = 0.5 * (tr["u1"][3:-3:7] + tr["u1"][4:-2:7])
        tm, vm = oracle.mlp_points(p, h, i, mid)
        np.testing.assert_allclose(tm, CubicSpline(tr["u1"], tr["t1"])(mid), atol=2e-8)
        np.testing.assert_allclose(vm, CubicSpline(tr["u1"], tr["v1"])(mid), atol=2e-8)


def test_features_and_targets(cfg2_sample):
    p, h = cfg2_sample
    F, Y, cells, u2 = oracle.net_features(p, h)
    for i in range(0, 400, 13):
        e = _entry_frame(h, i)
        ref = np.array([(e["t_out"] - e["t_in"]) * 2, (e["v_out"] - e["v_in"]) * 2, e["dT"] * 20, e["dV"] * 20,
                        e["u2"] * 0.125 - 1.5, float(h.wepl[i]) * 0.125 - 1.5, e["t_in"] * 0.1, e["t_out"] * 0.1,
                        e["v_in"] * 0.25, e["v_out"] * 0.25, float(h.xz_entry_angle[i]) * 10,
                        float(h.xz_exit_angle[i]) * 10])
        np.testing.assert_allclose(F[i], ref, rtol=2e-6, atol=2e-6)
        s = np.arange(1, 5) / 5.0
        t, v = oracle.mlp_points(p, h, i, e["u2"] * s)
        h01, h11 = 3 * s * s - 2 * s ** 3, s ** 3 - s * s
        tH = e["t_in"] + (e["t_out"] - e["t_in"]) * h01 + e["u2"] * e["dT"] * h11
        vH = e["v_in"] + (e["v_out"] - e["v_in"]) * h01 + e["u2"] * e["dV"] * h11
        np.testing.assert_allclose(Y[i, :4], (t - tH) / 0.01, rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(Y[i, 4:], (v - vH) / 0.01, rtol=1e-5, atol=1e-5)
    # the offsets are real bending, not noise: up to a few 100 um
    assert 0.05 < np.abs(Y).mean() < 5 and np.abs(Y).max() < 20


def test_surrogate_walk_follows_the_spline_path(cfg2_sample):
    p, h = cfg2_sample
    F, Y, _, _ = oracle.net_features(p, h)
    worst_exact_knots = 0.0
    for i in range(0, 300, 11):
        e = _entry_frame(h, i)
        tr = oracle.mlp_trace(p, h, i)
        sigma = tr["u1"] / e["u2"]
        for out8 in (np.zeros(8, np.float32), Y[i]):
            path, ts, vs = oracle.surrogate_path(p, h, i, out8)
            ref_t = onet.spline_path(out8, e["t_in"], e["t_out"], e["u2"], e["dT"], sigma)
            ref_v = onet.spline_path(out8[4:], e["v_in"], e["v_out"], e["u2"], e["dV"], sigma)
            assert len(ts) == len(tr["u1"])
            np.testing.assert_allclose(ts, ref_t, atol=2e-5)
            np.testing.assert_allclose(vs, ref_v, atol=2e-5)
            assert path[0] == h.first_voxel[i]
        worst_exact_knots = max(worst_exact_knots, np.abs(ts - tr["t1"]).max(), np.abs(vs - tr["v1"]).max())
    # with the exact offsets the surrogate is the MLP up to the spline's error
    assert worst_exact_knots < 1.5e-3


def test_network_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = onet.init_params(7).astype(np.float64)
    params[768:832] = rng.normal(0, 0.1, 64)
    params[4928:4992] = rng.normal(0, 0.1, 64)
    params[5504:] = rng.normal(0, 0.1, 8)
    X = rng.normal(0, 1, (37, 12))
    Y = rng.normal(0, 1, (37, 8))
    g, _ = onet.gradient(params, X, Y, bf16=False)
    idx = rng.choice(onet.P, 120, replace=False)
    idx = np.concatenate([idx, [0, 767, 768, 831, 832, 4927, 4928, 4991, 4992, 5503, 5504, 5511]])
    for k in idx:
        e = np.zeros(onet.P)
        hstep = 1e-6
        e[k] = hstep
        fd = (onet.loss(params + e, X, Y, bf16=False) - onet.loss(params - e, X, Y, bf16=False)) / (2 * hstep)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, g[k])


def test_network_bf16_forward_and_adam_first_step():
    rng = np.random.default_rng(4)
    params = onet.init_params(1)
    X = rng.normal(0, 1, (64, 12)).astype(np.float32)
    y32 = onet.forward(params, X, bf16=False)["y"]
    y16 = onet.forward(params, X, bf16=True)["y"]
    assert np.abs(y16 - y32).max() < 0.03 * np.abs(y32).max()
    g = rng.normal(0, 1, onet.P).astype(np.float32)
    p1, m1, v1 = onet.adam_step(params, np.zeros(onet.P), np.zeros(onet.P), g, 1, 1e-3)
    np.testing.assert_allclose(params - p1, 1e-3 * np.sign(g) * np.abs(g) / (np.abs(g) + 1e-8), rtol=1e-4, atol=1e-9)
    assert onet.bf16(np.float32(1.0 + 2 ** -8)) == 1.0          # tie to even
    assert onet.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6 * 1  # ties round up to even mantissa


def test_calibration_cells():
    p = configs.params("mini")
    h = oracle.HostHistories(3)
    # straight track (no bending, u2 = 16): cell 8*5 + 0; a bent one; a short one
    vals = [dict(x_entry=-8.0, y_entry=0.1, z_entry=0.0, x_exit=8.0, y_exit=0.1, z_exit=0.0),
            dict(x_entry=-8.0, y_entry=0.1, z_entry=0.0, x_exit=8.0, y_exit=0.5, z_exit=0.0),
            dict(x_entry=-1.0, y_entry=0.0, z_entry=0.0, x_exit=1.0, y_exit=0.0, z_exit=0.03)]
    for i, d in enumerate(vals):
        for k, v in d.items():
            getattr(h, k)[i] = v
    _, _, cells, u2 = oracle.net_features(p, h, targets=False)
    np.testing.assert_allclose(u2, [16.0, 16.0, 2.0], rtol=1e-6)
    # bending 0.4 cm: thresholds 0.02 * 2^k -> 5 of 7 at or below (0.02 .. 0.32)
    assert list(cells) == [8 * 5 + 0, 8 * 5 + 5, 8 * 0 + 1]
