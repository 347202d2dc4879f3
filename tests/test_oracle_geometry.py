"""Pins the oracle's geometry stages against closed forms and brute force:

  - recon-volume intersections (.cu:2169-2387): the entry/exit points are the
    line-circle intersections nearest the tracker planes (closed form in
    float64), rotated by the gantry angle; z from the uv slope; misses when
    |t| at closest approach exceeds the radius; the 5 cm t/v cut;
  - the cylinder hull (.cu:3316-3326): voxel-centre test;
  - MLP endpoints (.cu:4997-5277): the first hull voxel of a dense sampling of
    the walk line, and the hull entry point on that voxel's boundary.
"""
import math

import numpy as np
import pytest

import oracle
import pct_amd
from pct_amd import configs


def _raw_one(**kw):
    d = dict(gantry_angle=0, t_in_1=0.0, t_in_2=0.0, t_out_1=0.0, t_out_2=0.0, u_in_1=-25.0, u_in_2=-15.0,
             u_out_1=15.0, u_out_2=25.0, v_in_1=0.0, v_in_2=0.0, v_out_1=0.0, v_out_2=0.0, wepl=10.0)
    d.update(kw)
    return {k: np.array([v], np.int32 if k == "gantry_angle" else np.float32) for k, v in d.items()}


def _closed_form_entry(u0, t0, angle, R):
    """nearest intersection of the line through (u0,t0) with direction angle
    and the circle of radius R, in (u,t)"""
    du, dt = math.cos(angle), math.sin(angle)
    b = 2 * (u0 * du + t0 * dt)
    c = u0 * u0 + t0 * t0 - R * R
    disc = b * b - 4 * c
    if disc <= 0:
        return None
    s1, s2 = (-b - math.sqrt(disc)) / 2, (-b + math.sqrt(disc)) / 2
    s = s1 if abs(s1) <= abs(s2) else s2
    return u0 + s * du, t0 + s * dt


@pytest.mark.parametrize("gantry", [0, 36, 90, 132, 180, 268, 356])
def test_intersections_closed_form(gantry):
    p = configs.params("cfg1")
    rng = np.random.default_rng(gantry)
    for _ in range(20):
        t2, v2 = rng.uniform(-9, 9), rng.uniform(-1.5, 1.5)
        mt, mv = rng.uniform(-0.05, 0.05), rng.uniform(-0.01, 0.01)
        to1 = t2 + mt * 30 + rng.uniform(-0.3, 0.3)
        raw = _raw_one(gantry_angle=gantry, t_in_2=t2, t_in_1=t2 - 10 * mt, v_in_2=v2, v_in_1=v2 - 10 * mv,
                       t_out_1=to1, t_out_2=to1 + 10 * mt, v_out_1=v2 + 30 * mv, v_out_2=v2 + 40 * mv)
        h, missed = oracle.recon_volume_intersections(p, raw)
        rot = math.radians(gantry)
        a_in = math.atan2(float(np.float32(raw["t_in_2"][0] - raw["t_in_1"][0])), 10.0)
        e = _closed_form_entry(float(raw["u_in_2"][0]), float(raw["t_in_2"][0]), a_in, p.recon_cyl_radius)
        assert e is not None and missed[0] == 0
        x = math.cos(rot) * e[0] - math.sin(rot) * e[1]
        y = math.sin(rot) * e[0] + math.cos(rot) * e[1]
        assert h.x_entry[0] == pytest.approx(x, abs=2e-5)
        assert h.y_entry[0] == pytest.approx(y, abs=2e-5)
        assert h.xy_entry_angle[0] == pytest.approx(a_in + rot, abs=1e-6)
        assert math.hypot(h.x_entry[0], h.y_entry[0]) == pytest.approx(p.recon_cyl_radius, abs=2e-5)
        slope = float(np.float32(raw["v_in_2"][0] - raw["v_in_1"][0])) / 10.0
        z = float(raw["v_in_2"][0]) + slope * (e[0] - float(raw["u_in_2"][0]))
        assert h.z_entry[0] == pytest.approx(z, abs=2e-5)


def test_intersection_rotation_equivariance():
    """Rotating the gantry by 90 degrees rotates the entry point by 90 degrees."""
    p = configs.params("cfg1")
    raw0 = _raw_one(gantry_angle=0, t_in_2=3.0, t_in_1=2.9, t_out_1=3.5, t_out_2=3.6)
    raw1 = _raw_one(gantry_angle=90, t_in_2=3.0, t_in_1=2.9, t_out_1=3.5, t_out_2=3.6)
    h0, _ = oracle.recon_volume_intersections(p, raw0)
    h1, _ = oracle.recon_volume_intersections(p, raw1)
    assert h1.x_entry[0] == pytest.approx(-h0.y_entry[0], abs=2e-5)
    assert h1.y_entry[0] == pytest.approx(h0.x_entry[0], abs=2e-5)


def test_intersection_misses_and_cuts():
    p = configs.params("cfg1")
    # passes outside the cylinder (|t| = 10.5 > R)
    h, m = oracle.recon_volume_intersections(p, _raw_one(t_in_1=10.5, t_in_2=10.5, t_out_1=10.5, t_out_2=10.5))
    assert m[0] == 1
    # more than 5 cm lateral difference between entry and exit trackers
    h, m = oracle.recon_volume_intersections(p, _raw_one(t_in_1=0.0, t_in_2=0.0, t_out_1=5.5, t_out_2=5.5))
    assert m[0] == 1
    # z outside the cylinder at entry
    h, m = oracle.recon_volume_intersections(p, _raw_one(v_in_1=2.5, v_in_2=2.5, v_out_1=2.5, v_out_2=2.5))
    assert m[0] == 1
    # exits through the top face: exit point clipped to z = H/2
    h, m = oracle.recon_volume_intersections(p, _raw_one(v_in_1=1.2, v_in_2=1.5, v_out_1=4.5, v_out_2=5.5))
    assert m[0] == 0
    assert h.z_exit[0] == pytest.approx(p.recon_cyl_height / 2)
    assert abs(h.x_exit[0]) < p.recon_cyl_radius


def test_hull_init_voxel_centres():
    p = configs.params("mini")
    hull = oracle.hull_init(p).reshape(p.slices, p.rows, p.columns)
    vw = 2 * p.recon_cyl_radius / p.columns
    c = (np.arange(p.columns) - p.columns // 2 + 0.5) * vw
    r = (p.rows // 2 - np.arange(p.rows) - 0.5) * vw
    X, Y = np.meshgrid(c, r)
    exp = (X ** 2 + Y ** 2 < p.recon_cyl_radius ** 2).astype(np.uint8)
    for s in range(p.slices):
        np.testing.assert_array_equal(hull[s], exp)


def _dense_first_hull_voxel(p, hull, x0, y0, z0, axy, axz, entering, ds=2e-5):
    """brute force: sample the walk line densely (param x, the reference's
    z = z0 + tan(xz)(x - x0)) and return the first hull voxel"""
    if not entering:
        axy += math.pi
    vw = 2 * p.recon_cyl_radius / p.columns
    tx, tz = math.tan(axy), math.tan(axz)
    # parametrise by arc length in xy
    k = np.arange(int(40 / ds))
    x = x0 + k * (math.cos(axy) * ds)
    y = y0 + tx * (x - x0)
    z = z0 + tz * (x - x0)
    vx = (np.abs(x + p.recon_cyl_radius) / vw).astype(np.int64)
    vy = (np.abs(y - p.recon_cyl_radius) / vw).astype(np.int64)
    vz = (np.abs(z - p.recon_cyl_height / 2) / p.slice_thickness).astype(np.int64)
    out = (vx >= p.columns) | (vy >= p.rows) | (vz >= p.slices)
    v = np.where(out, 0, vx + vy * p.columns + vz * p.columns * p.rows)
    hit = (~out) & (hull[v] == 1)
    first_out = np.argmax(out) if out.any() else len(k)
    first_hit = np.argmax(hit) if hit.any() else len(k)
    if first_hit >= first_out:
        return None
    j = first_hit
    return int(v[j]), (x[j], y[j], z[j])


def test_endpoints_dense_sampling(mini_host):
    p = mini_host["p"]
    hull = mini_host["hull"].copy()
    # carve a smaller hull so the walks take many steps
    C, R, S = p.columns, p.rows, p.slices
    vw = 2 * p.recon_cyl_radius / C
    c = (np.arange(C) - C // 2 + 0.5) * vw
    X, Y = np.meshgrid(c, (R // 2 - np.arange(R) - 0.5) * vw)
    small = ((X ** 2 + Y ** 2) < 6.0 ** 2).astype(np.uint8)
    hull = np.broadcast_to(small, (S, R, C)).reshape(-1).copy()
    h = mini_host["h_pre"].head(60)
    h0 = h.select(np.ones(h.n, bool))
    inter = oracle.mlp_endpoints(p, h, hull)
    checked = 0
    for i in range(h.n):
        ent = _dense_first_hull_voxel(p, hull, float(h0.x_entry[i]), float(h0.y_entry[i]), float(h0.z_entry[i]),
                                      float(h0.xy_entry_angle[i]), float(h0.xz_entry_angle[i]), True)
        ext = _dense_first_hull_voxel(p, hull, float(h0.x_exit[i]), float(h0.y_exit[i]), float(h0.z_exit[i]),
                                      float(h0.xy_exit_angle[i]), float(h0.xz_exit_angle[i]), False)
        assert bool(inter[i]) == (ent is not None and ext is not None), i
        if inter[i]:
            assert int(h.first_voxel[i]) == ent[0], i
            assert h.x_entry[i] == pytest.approx(ent[1][0], abs=1e-3)
            assert h.y_entry[i] == pytest.approx(ent[1][1], abs=1e-3)
            checked += 1
    assert checked > 20
