"""CPU checks of the C ABI library and the host-side logic (no GPU compute):

  - libpct.so loads and exports every function declared in include/*.h;
  - the ctypes structs have the C layout (sizes/offsets from a compiled probe);
  - the product fails loudly when libpct.so is missing;
  - the shared math library is within a few ulp of numpy/glibc;
  - the synthetic generator: angle grouping, exact line integrals, misses,
    determinism and independence of the generated range.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import pct_amd
from pct_amd import _lib, configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(pct_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared_functions()
    assert len(declared) >= 15
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    # and the binding declares a signature for each of them
    assert declared <= set(_lib.SIGNATURES), declared - set(_lib.SIGNATURES)


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(f"""
#include <stddef.h>
#include "{ROOT}/include/pct.h"
#include "{ROOT}/include/pct_synth_api.h"
#include "{ROOT}/include/pct_net.h"
#include "{ROOT}/include/pct_io.h"
long v[] = {{ sizeof(pct_params), offsetof(pct_params, a), offsetof(pct_params, max_intersections),
             offsetof(pct_params, lambda), sizeof(pct_hist_soa), sizeof(pct_raw_soa), sizeof(pct_synth_cfg),
             offsetof(pct_synth_cfg, cyl_rsp), sizeof(pct_synth_soa), sizeof(pct_net_train_cfg),
             offsetof(pct_net_train_cfg, seed), sizeof(pct_net_calib), offsetof(pct_net_calib, cell_samples),
             offsetof(pct_net_calib, cell_p99), offsetof(pct_net_calib, cell_fallback),
             PCT_NET_PARAMS, PCT_NET_CELLS, PCT_MLP_SURROGATE, sizeof(pct_pctd_header),
             offsetof(pct_pctd_header, prepared_by), offsetof(pct_pctd_header, data_offset), PCT_TXT_BOOL, PCT_E_IO }};
""")
    so = tmp_path / "probe.so"
    subprocess.run(["gcc", "-shared", "-fPIC", "-o", str(so), str(src)], check=True)
    v = (ctypes.c_long * 23).in_dll(ctypes.CDLL(str(so)), "v")
    exp = [ctypes.sizeof(_lib.PctParams), _lib.PctParams.a.offset, _lib.PctParams.max_intersections.offset,
           _lib.PctParams.lambda_.offset, ctypes.sizeof(_lib.PctHistSoa), ctypes.sizeof(_lib.PctRawSoa),
           ctypes.sizeof(_lib.PctSynthCfg), _lib.PctSynthCfg.cyl_rsp.offset, ctypes.sizeof(_lib.PctSynthSoa),
           ctypes.sizeof(_lib.PctNetTrainCfg), _lib.PctNetTrainCfg.seed.offset, ctypes.sizeof(_lib.PctNetCalib),
           _lib.PctNetCalib.cell_samples.offset, _lib.PctNetCalib.cell_p99.offset,
           _lib.PctNetCalib.cell_fallback.offset, _lib.NET_PARAMS, _lib.NET_CELLS, _lib.PCT_MLP_SURROGATE,
           ctypes.sizeof(_lib.PctPctdHeader), _lib.PctPctdHeader.prepared_by.offset,
           _lib.PctPctdHeader.data_offset.offset, _lib.PCT_TXT_BOOL, -5]
    assert list(v) == exp


def test_default_params_are_the_reference_defaults():
    p = _lib.default_params()
    assert (p.columns, p.rows, p.slices) == (200, 200, 32)           # data.h:264-266
    assert (p.t_bins, p.v_bins, p.angular_bins) == (350, 36, 90)     # data.h:238-243
    assert p.max_intersections == 512 and p.drop_block_size == 3200
    assert p.lambda_ == pytest.approx(0.001)
    assert p.a[0] == 7.457 * float(np.float32(1e-6))                 # data.h:306 (float pow)
    assert p.e_0 == 13.6 and p.x_0 == 36.08


def test_product_fails_loudly_without_library(monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libpct.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError):
        _lib.lib()


def _ulp(a, b):
    return np.abs(a - b) / np.spacing(np.abs(b))


@pytest.mark.parametrize("func,ref,lo,hi,tol", [
    ("log", np.log, 1e-6, 60.0, 1), ("sin", np.sin, -13, 13, 1), ("cos", np.cos, -13, 13, 1),
    ("tan", np.tan, -1.5, 1.5, 4), ("atan", np.arctan, -50, 50, 1)])
def test_math_accuracy(func, ref, lo, hi, tol):
    rng = np.random.default_rng(1)
    x = rng.uniform(lo, hi, 100_000)
    got = oracle.math_eval(func, x)
    exp = ref(x)
    keep = np.abs(exp) > 1e-6
    assert _ulp(got[keep], exp[keep]).max() <= tol


def test_math_special_values():
    assert oracle.math_eval("log", [1.0])[0] == 0.0
    assert oracle.math_eval("log", [np.e])[0] == pytest.approx(1.0, rel=2e-16)
    assert oracle.math_eval("atan2", [0.0], [-1.0])[0] == np.pi
    assert oracle.math_eval("atan2", [1.0], [0.0])[0] == np.pi / 2
    assert oracle.math_eval("sin", [0.0])[0] == 0.0 and oracle.math_eval("cos", [0.0])[0] == 1.0
    assert np.isnan(oracle.math_eval("log", [-1.0])[0])
    x = np.array([1e-310, 5e-324])
    np.testing.assert_allclose(oracle.math_eval("log", x), np.log(x), rtol=1e-15)


def test_synth_angles_and_wepl():
    cfg = configs.synth_cfg("mini", n_histories=9020)
    raw = pct_amd.generate_host(cfg).numpy()
    ga = raw["gantry_angle"]
    # 9020 = 90 * 100 + 20: the first 20 angles get 101 histories
    counts = np.bincount(ga // 4, minlength=90)
    assert counts[:20].tolist() == [101] * 20 and counts[20:].tolist() == [100] * 70
    assert np.all(np.diff(ga) >= 0)
    w = raw["wepl"]
    assert np.all(w >= 0)
    assert 0.5 < np.mean(w > 0) < 1.0
    # straight-line WEPL through the water cylinder never exceeds its diameter
    # by more than the inserts and the noise allow
    assert w.max() < 20.0 * 1.6 * 1.05


def test_synth_line_integral_closed_form():
    """a central track at gantry 0 crosses 20 cm of water; the straggling is
    1 %: the mean WEPL of tracks with |t| < 0.2 cm is 20 cm within noise"""
    cfg = configs.synth_cfg("mini", n_histories=90 * 4000)
    cfg.n_cyl = 1           # water only
    raw = pct_amd.generate_host(cfg, 0, 4000).numpy()
    t = raw["t_in_2"]
    sel = np.abs(t) < 0.2
    assert sel.sum() > 20
    # chord of a line at lateral t through a circle of radius 10: 2 sqrt(100 - t^2),
    # scaled by the small cone-beam obliquity
    chord = 2 * np.sqrt(100 - t[sel].astype(np.float64) ** 2)
    rel = raw["wepl"][sel] / chord - 1
    assert abs(rel.mean()) < 0.01 and rel.std() < 0.02


def test_synth_range_independence():
    cfg = configs.synth_cfg("mini", n_histories=5000)
    full = pct_amd.generate_host(cfg).numpy()
    part = pct_amd.generate_host(cfg, 1234, 777).numpy()
    for k in full:
        np.testing.assert_array_equal(full[k][1234:1234 + 777], part[k])
    again = pct_amd.generate_host(cfg).numpy()
    for k in full:
        np.testing.assert_array_equal(full[k], again[k])


def test_synth_ellipsoid_line_integrals():
    """rotated ellipsoid phantoms (the batched configuration's): without
    straggling, the WEPL of a history is the RSP-weighted length of its
    entry line inside the bodies, checked against dense sampling of that line
    in numpy (gantry frame -> object frame -> ellipsoid frames)"""
    cfg = configs.synth_cfg("mini", n_histories=90 * 300)
    cfg.straggle = 0.0
    cfg.n_cyl = 0
    cfg.n_ell = 2
    body = dict(x=0.3, y=-0.2, z=0.1, a=9.0, b=6.5, c=3.0, phi=0.4, rsp=1.05)
    ins = dict(x=2.0, y=1.0, z=0.0, a=2.5, b=1.5, c=1.8, phi=-1.1, rsp=1.7)
    for k, e in enumerate((body, ins)):
        for f, v in e.items():
            getattr(cfg, "ell_" + f)[k] = v
    raw = pct_amd.generate_host(cfg, 0, cfg.n_histories).numpy()
    rng = np.random.default_rng(3)
    idx = rng.choice(cfg.n_histories, 60, replace=False)

    def inside(e, x, y, z):
        c, s = np.cos(e["phi"]), np.sin(e["phi"])
        dx, dy, dz = x - e["x"], y - e["y"], z - e["z"]
        lx, ly = (c * dx + s * dy) / e["a"], (-s * dx + c * dy) / e["b"]
        return lx ** 2 + ly ** 2 + (dz / e["c"]) ** 2 <= 1.0

    n_hit = 0
    for i in idx:
        th = np.deg2rad(float(raw["gantry_angle"][i]))
        u1, u2 = float(raw["u_in_1"][i]), float(raw["u_in_2"][i])
        t1, t2 = float(raw["t_in_1"][i]), float(raw["t_in_2"][i])
        v1, v2 = float(raw["v_in_1"][i]), float(raw["v_in_2"][i])
        mt, mv = (t2 - t1) / (u2 - u1), (v2 - v1) / (u2 - u1)
        u = np.linspace(-15.0, 15.0, 300001)
        t, v = t2 + mt * (u - u2), v2 + mv * (u - u2)
        x, y = np.cos(th) * u - np.sin(th) * t, np.sin(th) * u + np.cos(th) * t
        du = (u[1] - u[0]) * np.sqrt(1 + mt * mt + mv * mv)
        rsp = np.where(inside(body, x, y, v), body["rsp"], 0.0)
        rsp = np.where(inside(ins, x, y, v), ins["rsp"], rsp)
        w = rsp.sum() * du
        n_hit += w > 0
        assert abs(float(raw["wepl"][i]) - w) < 2e-3 * max(w, 1.0) + 2e-3, (i, raw["wepl"][i], w)
    assert n_hit > 30


def test_tuner_candidates():
    """the batched form's block-size candidates: one uniform assignment per
    allowed size first, then step-balanced ones, no duplicates, only allowed
    sizes; the balanced assignment at the largest scan's smallest step count
    is among them"""
    import math
    from pct_amd.tuner import candidate_assignments
    sizes = [800, 1600, 3200, 6400]
    hist = [540_629, 6_193_135, 1_218_385, 11_938_713, 3_969_720]
    cands = candidate_assignments(hist, sizes)
    assert cands[:4] == [tuple([b] * len(hist)) for b in sizes]
    assert len(cands) == len(set(cands))
    for c in cands:
        assert all(b in sizes for b in c)
    # the balanced assignment for the largest scan's minimum step count gives
    # the smaller scans smaller blocks than the largest one
    t = math.ceil(max(hist) / sizes[-1])
    bal = tuple(next((b for b in sizes if math.ceil(h / b) <= t), sizes[-1]) for h in hist)
    assert bal in cands and min(bal) < max(bal)
