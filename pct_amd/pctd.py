"""Scan files and text dumps around the path (include/pct_io.h).

PCTD version-0 projection files are what the reference reads (one file per
gantry angle, count_histories_v0 .cu:997-1097, read_data_chunk_v02
.cu:1769-1933); `read_scan` loads a directory of them the way its main loop
does (angles 0, interval, ... < 360, file "<directory><folder>/<base>_<angle
%03d><extension>") into a RawScan, and `write_scan` stores a scan (e.g. a
synthetic one) in that format. `array_to_disk` writes images, masks and
sinograms as the reference's array_2_disk text dumps (.cu:8092-8134), so that
they can be compared with the reference's outputs.

The parsing and formatting run in libpct.so (C++); this module only sizes the
buffers and names the files.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import PctPctdHeader, PctRawSoa, PctSynthSoa

RAW_FIELDS = _lib.RAW_FLOAT_FIELDS


class PctdError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        msg = _lib.lib().pct_io_last_error()
        raise PctdError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _b(s):
    return s.encode() if isinstance(s, str) else s


def scan_file_name(directory, folder, base_name, angle, extension=".bin"):
    """the reference's projection file name (.cu:1011, 1782)"""
    return "%s%s/%s_%03d%s" % (directory, folder, base_name, int(angle), extension)


def read_header(path):
    h = PctPctdHeader()
    _check(_lib.lib().pct_pctd_read_header(_b(path), ctypes.byref(h)), f"pct_pctd_read_header({path})")
    return {"version": h.version, "n_histories": h.n_histories, "projection_angle": h.projection_angle,
            "beam_energy": h.beam_energy, "generation_date": h.generation_date,
            "preprocess_date": h.preprocess_date, "phantom_name": h.phantom_name.decode(errors="replace"),
            "data_source": h.data_source.decode(errors="replace"),
            "prepared_by": h.prepared_by.decode(errors="replace"), "data_offset": h.data_offset}


def _host_soa(arrays, offset=0):
    s = PctSynthSoa()
    s.gantry_angle = arrays["gantry_angle"][offset:].ctypes.data
    for f in RAW_FIELDS:
        setattr(s, f, arrays[f][offset:].ctypes.data)
    return s


def read_file(path, data_in_mm=True):
    """one projection file -> dict of numpy arrays (cm when data_in_mm)"""
    n = read_header(path)["n_histories"]
    arrays = {"gantry_angle": np.zeros(max(n, 1), np.int32)}
    for f in RAW_FIELDS:
        arrays[f] = np.zeros(max(n, 1), np.float32)
    got = ctypes.c_int64(0)
    _check(_lib.lib().pct_pctd_read(_b(path), int(bool(data_in_mm)), n, ctypes.byref(_host_soa(arrays)),
                                    ctypes.byref(got)), f"pct_pctd_read({path})")
    return {k: v[:got.value] for k, v in arrays.items()}


def read_scan(directory, folder="", base_name="projection", angle_interval=4.0, extension=".bin", data_in_mm=True,
              device=None):
    """every projection file of a scan, in the reference's order (gantry
    positions 0, interval, ... < 360), concatenated: a dict of numpy arrays,
    or a RawScan on `device`"""
    paths, counts = [], []
    pos = 0
    while True:
        angle = int(pos * angle_interval)
        if angle >= 360:
            break
        p = scan_file_name(directory, folder, base_name, angle, extension)
        paths.append(p)
        counts.append(read_header(p)["n_histories"])
        pos += 1
    total = sum(counts)
    arrays = {"gantry_angle": np.zeros(max(total, 1), np.int32)}
    for f in RAW_FIELDS:
        arrays[f] = np.zeros(max(total, 1), np.float32)
    off = 0
    for p, n in zip(paths, counts):
        got = ctypes.c_int64(0)
        _check(_lib.lib().pct_pctd_read(_b(p), int(bool(data_in_mm)), n, ctypes.byref(_host_soa(arrays, off)),
                                        ctypes.byref(got)), f"pct_pctd_read({p})")
        off += got.value
    arrays = {k: v[:total] for k, v in arrays.items()}
    if device is None:
        return arrays
    import torch
    from . import RawScan
    raw = RawScan(total, device)
    for k, v in arrays.items():
        getattr(raw, k).copy_(torch.from_numpy(v))
    return raw


def write_file(path, arrays, projection_angle, beam_energy=200.0, generation_date=0, preprocess_date=0,
               phantom_name="", data_source="", prepared_by=""):
    """one version-0 projection file from a dict of numpy float32 arrays (the
    13 raw fields, written as they are)"""
    n = len(arrays["wepl"])
    h = PctPctdHeader()
    h.version = 0
    h.n_histories = n
    h.projection_angle = projection_angle
    h.beam_energy = beam_energy
    h.generation_date = generation_date
    h.preprocess_date = preprocess_date
    h.phantom_name = _b(phantom_name)[:_lib.PCTD_STR - 1]
    h.data_source = _b(data_source)[:_lib.PCTD_STR - 1]
    h.prepared_by = _b(prepared_by)[:_lib.PCTD_STR - 1]
    keep = [np.ascontiguousarray(arrays[f], np.float32) for f in RAW_FIELDS]
    s = PctRawSoa()
    s.gantry_angle = None
    for f, a in zip(RAW_FIELDS, keep):
        setattr(s, f, a.ctypes.data if n else None)
    s.n = n
    _check(_lib.lib().pct_pctd_write(_b(path), ctypes.byref(h), ctypes.byref(s)), f"pct_pctd_write({path})")


def write_scan(raw, directory, folder="", base_name="projection", extension=".bin", to_mm=True, **header):
    """a scan (RawScan or dict of numpy arrays, lengths in cm) as one file per
    gantry angle: the file of internal gantry angle g is named by g, the
    angle the reader's loop visits it at, and carries the projection angle
    (g + 90) % 360, which the reader maps back to g; lengths are stored in mm
    (x 10 in float) when to_mm. Returns the paths written, in the reader's
    order (for gantry angles on its grid, the scan's own order)."""
    if not isinstance(raw, dict):
        raw = raw.numpy()
    os.makedirs(os.path.join(directory + folder) or ".", exist_ok=True)
    g = np.asarray(raw["gantry_angle"])
    written = {}
    for angle in np.unique(g):
        sel = g == angle
        arrays = {f: np.asarray(raw[f])[sel].astype(np.float32) for f in RAW_FIELDS}
        if to_mm:
            arrays = {f: (a * np.float32(10.0)).astype(np.float32) for f, a in arrays.items()}
        path = scan_file_name(directory, folder, base_name, int(angle), extension)
        write_file(path, arrays, float((int(angle) + 90) % 360), **header)
        written[int(angle)] = path
    return [written[k] for k in sorted(written)]


_TXT_TYPES = {np.dtype(np.float32): _lib.PCT_TXT_F32, np.dtype(np.float64): _lib.PCT_TXT_F64,
              np.dtype(np.int32): _lib.PCT_TXT_I32, np.dtype(np.uint32): _lib.PCT_TXT_U32,
              np.dtype(np.bool_): _lib.PCT_TXT_BOOL, np.dtype(np.uint8): _lib.PCT_TXT_BOOL}


def array_to_disk(data, filename_base, directory, folder, x_max, y_max, z_max, elements=None, single_file=True):
    """array_2_disk (.cu:8092-8134) of a numpy array or tensor (uint8 / bool
    arrays are written as the reference's bool masks)"""
    if hasattr(data, "detach"):
        data = data.detach().cpu().numpy()
    a = np.ascontiguousarray(data).reshape(-1)
    if a.dtype not in _TXT_TYPES:
        raise PctdError(f"array_to_disk: unsupported dtype {a.dtype}")
    kind = _TXT_TYPES[a.dtype]
    if a.dtype == np.bool_:
        a = a.view(np.uint8)
    elements = a.size if elements is None else int(elements)
    if elements > a.size:
        raise PctdError("array_to_disk: more elements than the array holds")
    os.makedirs(directory + folder or ".", exist_ok=True)
    _check(_lib.lib().pct_array_to_disk(_b(filename_base), _b(directory), _b(folder), a.ctypes.data, kind,
                                        int(x_max), int(y_max), int(z_max), elements, int(bool(single_file))),
           "pct_array_to_disk")
