"""PCTD version-0 projection files and array_2_disk text dumps (include/pct_io.h,
pct_amd/pctd.py): host code, run on the CPU.

  - a file assembled byte by byte from the documented layout (.cu:1006-1020,
    1772-1799) is parsed field by field; the event arrays land in the raw
    tracker fields in file order; mm -> cm is the float times the double 0.1
    rounded to float (.cu:1863-1878); every event gets the gantry angle
    (int(projection angle) + 270) % 360 with C's truncating int() and % (.cu:1917);
  - writer -> reader round trips, an empty file, long strings;
  - malformed files are errors: wrong magic, another version, short data,
    too small a buffer, a missing file;
  - a synthetic scan written one file per gantry angle (in mm) and read back
    in the reference's file order;
  - text dumps against the stream's default formatting (%g, 6 digits), both
    file layouts, a partial last row, bool masks.
"""
import ctypes
import math
import os
import struct

import numpy as np
import pytest

from pct_amd import _lib, pctd

RAW = _lib.RAW_FLOAT_FIELDS
FILE_ORDER = ("t_in_1", "t_in_2", "t_out_1", "t_out_2", "v_in_1", "v_in_2", "v_out_1", "v_out_2",
              "u_in_1", "u_in_2", "u_out_1", "u_out_2", "wepl")


def _c_mod(a, b):
    return a - b * int(a / b)            # C's % (truncating division)


def _pctd_bytes(angle, arrays, energy=200.0, gen=1500000000, pre=1500000100,
                strings=(b"water phantom", b"Geant4", b"tester"), version=0):
    n = len(arrays["wepl"])
    out = b"PCTD" + struct.pack("<iiffii", version, n, angle, energy, gen, pre)
    for s in strings:
        out += struct.pack("<i", len(s)) + s
    for f in FILE_ORDER:
        out += np.asarray(arrays[f], np.float32).tobytes()
    return out


def _random_arrays(rng, n):
    return {f: rng.uniform(-1500, 1500, n).astype(np.float32) for f in RAW}


@pytest.mark.parametrize("angle", [0.0, 90.7, 359.9, -45.3, -300.5])
def test_layout_is_parsed(tmp_path, angle):
    rng = np.random.default_rng(1)
    arrays = _random_arrays(rng, 1000)
    data = _pctd_bytes(angle, arrays)
    path = tmp_path / "projection_000.bin"
    path.write_bytes(data)
    h = pctd.read_header(str(path))
    assert h["version"] == 0 and h["n_histories"] == 1000
    assert h["projection_angle"] == np.float32(angle) and h["beam_energy"] == 200.0
    assert (h["generation_date"], h["preprocess_date"]) == (1500000000, 1500000100)
    assert (h["phantom_name"], h["data_source"], h["prepared_by"]) == ("water phantom", "Geant4", "tester")
    assert h["data_offset"] == len(data) - 13 * 4 * 1000
    raw = pctd.read_file(str(path), data_in_mm=False)
    for f in RAW:
        assert np.array_equal(raw[f].view(np.uint32), arrays[f].view(np.uint32)), f
    cm = pctd.read_file(str(path), data_in_mm=True)
    for f in RAW:
        ref = (arrays[f].astype(np.float64) * 0.1).astype(np.float32)
        assert np.array_equal(cm[f].view(np.uint32), ref.view(np.uint32)), f
    g = _c_mod(math.trunc(float(np.float32(angle))) + 270, 360)
    assert np.all(cm["gantry_angle"] == g)


def test_writer_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    arrays = _random_arrays(rng, 777)
    p = str(tmp_path / "a.bin")
    long_name = "x" * 400
    pctd.write_file(p, arrays, 12.0, beam_energy=180.5, generation_date=7, preprocess_date=8, phantom_name=long_name,
                    data_source="synthetic", prepared_by="")
    h = pctd.read_header(p)
    assert h["n_histories"] == 777 and h["projection_angle"] == 12.0 and h["beam_energy"] == np.float32(180.5)
    assert h["phantom_name"] == "x" * 255 and h["data_source"] == "synthetic" and h["prepared_by"] == ""
    back = pctd.read_file(p, data_in_mm=False)
    for f in RAW:
        assert np.array_equal(back[f], arrays[f])
    assert np.all(back["gantry_angle"] == 282)
    # the bytes are the documented layout
    assert open(p, "rb").read() == _pctd_bytes(12.0, arrays, 180.5, 7, 8, (b"x" * 255, b"synthetic", b""))
    # an empty file
    e = str(tmp_path / "e.bin")
    pctd.write_file(e, {f: np.zeros(0, np.float32) for f in RAW}, 0.0)
    assert pctd.read_header(e)["n_histories"] == 0
    assert all(len(v) == 0 for v in pctd.read_file(e).values())


def test_malformed_files_are_errors(tmp_path):
    rng = np.random.default_rng(3)
    arrays = _random_arrays(rng, 50)
    good = _pctd_bytes(4.0, arrays)
    cases = {"magic": b"PCTX" + good[4:], "version": _pctd_bytes(4.0, arrays, version=1), "short": good[:-10],
             "header": good[:30]}
    for name, data in cases.items():
        p = tmp_path / f"{name}.bin"
        p.write_bytes(data)
        with pytest.raises(pctd.PctdError):
            pctd.read_file(str(p))
    with pytest.raises(pctd.PctdError):
        pctd.read_header(str(tmp_path / "missing.bin"))
    # a buffer smaller than the file's event count
    p = tmp_path / "ok.bin"
    p.write_bytes(good)
    buf = {"gantry_angle": np.zeros(10, np.int32), **{f: np.zeros(10, np.float32) for f in RAW}}
    got = ctypes.c_int64(0)
    rc = _lib.lib().pct_pctd_read(str(p).encode(), 1, 10, ctypes.byref(pctd._host_soa(buf)), ctypes.byref(got))
    assert rc == -1 and b"capacity" in _lib.lib().pct_io_last_error()


def test_synthetic_scan_round_trip(tmp_path):
    import pct_amd
    from pct_amd import configs
    raw = pct_amd.generate_host(configs.synth_cfg("mini", n_histories=9000)).numpy()
    paths = pctd.write_scan(raw, str(tmp_path), "/scan", base_name="projection", phantom_name="mini")
    assert len(paths) == len(np.unique(raw["gantry_angle"]))
    back = pctd.read_scan(str(tmp_path), "/scan", base_name="projection", angle_interval=4.0)
    # the reader's loop visits the files in the scan's own order (gantry 0, 4, ...)
    assert len(back["wepl"]) == 9000
    assert np.array_equal(back["gantry_angle"], raw["gantry_angle"])
    assert pctd.read_header(paths[0])["projection_angle"] == 90.0
    for f in RAW:
        mm = (raw[f] * np.float32(10.0)).astype(np.float32)
        ref = (mm.astype(np.float64) * 0.1).astype(np.float32)
        assert np.array_equal(back[f].view(np.uint32), ref.view(np.uint32)), f
        assert np.max(np.abs(back[f] - raw[f])) <= 2 * np.spacing(np.abs(raw[f]).max())


def _dump_text(values, fmt, x_max, y_max, z_range, elements):
    out, index = [], 0
    for z in z_range:
        for y in range(y_max):
            for x in range(x_max):
                index = x + y * x_max + z * x_max * y_max
                if index >= elements:
                    break
                out.append(fmt(values[index]) + " ")
            if index >= elements:
                break
            out.append("\n")
        if index >= elements:
            break
    return "".join(out)


def test_array_to_disk_format(tmp_path):
    rng = np.random.default_rng(4)
    vals = np.concatenate([rng.normal(0, 1, 20), [0.0, 1.0, -2.5, 1e-7, 123456789.0, 0.00113, 1.0 / 3.0, -0.0],
                           rng.uniform(0, 1e4, 2)]).astype(np.float32)
    d = str(tmp_path) + "/"
    g = lambda v: "%g" % float(v)                                    # noqa: E731
    pctd.array_to_disk(vals, "x", d, "out", 5, 3, 2)
    assert open(d + "out/x.txt").read() == _dump_text(vals, g, 5, 3, range(2), 30)
    pctd.array_to_disk(vals, "xs", d, "out", 5, 3, 2, single_file=False)
    for z in range(2):
        assert open(d + f"out/xs_{z}.txt").read() == _dump_text(vals, g, 5, 3, [z], 30)
    # a partial last row: no newline after it, later slices empty
    pctd.array_to_disk(vals, "p", d, "out", 5, 3, 2, elements=13, single_file=False)
    assert open(d + "out/p_0.txt").read() == _dump_text(vals, g, 5, 3, [0], 13)
    assert open(d + "out/p_1.txt").read() == ""
    # one slice in single-file mode is named as a per-slice file, as in the reference
    pctd.array_to_disk(vals[:15], "one", d, "out", 5, 3, 1)
    assert os.path.exists(d + "out/one_0.txt")
    # integers and bool masks
    ints = np.arange(-3, 9, dtype=np.int32)
    pctd.array_to_disk(ints, "i", d, "out", 4, 3, 1)
    assert open(d + "out/i_0.txt").read() == _dump_text(ints, str, 4, 3, [0], 12)
    mask = np.array([0, 3, 1, 0, 255, 1], np.uint8)
    pctd.array_to_disk(mask, "m", d, "out", 3, 2, 1)
    assert open(d + "out/m_0.txt").read() == "0 1 1 \n0 1 1 \n"
    pctd.array_to_disk(mask.astype(bool), "mb", d, "out", 3, 2, 1)
    assert open(d + "out/mb_0.txt").read() == "0 1 1 \n0 1 1 \n"
