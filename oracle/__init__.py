"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of oracle/liboracle.so, the plain C++ restatement of the
reference (lynch829/pCT) built by oracle/build.py. Used only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker. The
product package pct_amd never imports this module.

All arrays are host numpy arrays; the structs are the ABI structs of
include/pct.h (shared with the product's binding).
"""
import ctypes
import os

import numpy as np

from pct_amd import _lib as abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            from . import build as _b
            _b.build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        L.oracle_recon_volume_intersections.argtypes = [P(abi.PctParams), P(abi.PctRawSoa), P(abi.PctHistSoa), vp]
        L.oracle_hull_init.argtypes = [P(abi.PctParams), vp]
        L.oracle_mlp_endpoints.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), vp, vp]
        for name in ("oracle_chord_length", "oracle_chord_length_float_avg"):
            getattr(L, name).argtypes = [P(abi.PctParams), ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float]
            getattr(L, name).restype = ctypes.c_double
        L.oracle_drop_lists_research.argtypes = [ctypes.c_int64, ctypes.c_int64, vp, vp, vp, vp, vp, vp, vp, vp,
                                                 ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_int,
                                                 ctypes.c_int]
        L.oracle_mlp_path.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), ctypes.c_int64, vp, P(ctypes.c_int)]
        L.oracle_mlp_paths.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), ctypes.c_int64, ctypes.c_int64, vp,
                                       ctypes.c_int64, vp]
        L.oracle_mlp_paths_table.argtypes = L.oracle_mlp_paths.argtypes
        L.oracle_drop.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), vp, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                  ctypes.c_int64]
        L.oracle_drop_mt.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), vp, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_int, ctypes.c_int64, ctypes.c_int]
        L.oracle_drop_variant_mt.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), vp, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.oracle_mlp_trace.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), ctypes.c_int64, vp, vp, vp, vp, vp,
                                       ctypes.c_int, P(ctypes.c_int), ctypes.c_int]
        L.oracle_drop_lists.argtypes = [ctypes.c_int64, ctypes.c_int64, vp, vp, vp, vp, vp, ctypes.c_int,
                                        ctypes.c_float, ctypes.c_int]
        L.oracle_binning.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), vp, vp, vp, vp, vp, vp]
        L.oracle_bin_statistics.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), vp, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_statistical_cuts.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), vp, vp, vp, vp, vp, vp, vp, vp,
                                              vp, vp]
        for name, n in (("oracle_space_carve", 2), ("oracle_msc_count", 2), ("oracle_sm_count", 2)):
            getattr(L, name).argtypes = [P(abi.PctParams), P(abi.PctHistSoa)] + [vp] * n
        for name, n in (("oracle_msc_edge", 2), ("oracle_sm_edge", 3), ("oracle_hull_select", 2), ("oracle_fbp_filter", 2),
                        ("oracle_backprojection", 2), ("oracle_fbp_hull", 2), ("oracle_initial_iterate", 3)):
            getattr(L, name).argtypes = [P(abi.PctParams)] + [vp] * n
        L.oracle_math.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int64]
        L.oracle_math.restype = None
        L.oracle_mlp_points.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), ctypes.c_int64, ctypes.c_int, vp, vp, vp]
        L.oracle_net_features.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), ctypes.c_int64, vp, vp, vp, vp]
        L.oracle_surrogate_path.argtypes = [P(abi.PctParams), P(abi.PctHistSoa), ctypes.c_int64, vp, vp, vp, vp,
                                            ctypes.c_int, P(ctypes.c_int)]
        _L = L
    return _L


MATH_FUNCS = {"log": 0, "sin": 1, "cos": 2, "tan": 3, "atan": 4, "atan2": 5, "sinf": 6, "cosf": 7, "atan2f": 8,
              "sqrt": 9, "div": 10, "div_shared_rcp": 11, "log_rcp_div": 12}


def math_eval(func, a, b=None):
    """Host evaluation of a pct_math.h function (see MATH_FUNCS)."""
    f = MATH_FUNCS[func] if isinstance(func, str) else func
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty_like(a)
    lib().oracle_math(f, _p(a), _p(b), _p(out), a.size)
    return out


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


class HostHistories:
    """numpy SoA mirror of pct_hist_soa."""

    def __init__(self, n=0, arrays=None):
        if arrays is not None:
            self.n = len(arrays["wepl"])
            for f in abi.HIST_FLOAT_FIELDS:
                setattr(self, f, np.ascontiguousarray(arrays[f], dtype=np.float32))
            self.bin_num = np.ascontiguousarray(arrays.get("bin_num", np.zeros(self.n)), dtype=np.int32)
            self.first_voxel = np.ascontiguousarray(arrays.get("first_voxel", np.zeros(self.n)), dtype=np.uint32)
        else:
            self.n = n
            for f in abi.HIST_FLOAT_FIELDS:
                setattr(self, f, np.zeros(n, np.float32))
            self.bin_num = np.zeros(n, np.int32)
            self.first_voxel = np.zeros(n, np.uint32)

    def soa(self):
        s = abi.PctHistSoa()
        for f in abi.HIST_FLOAT_FIELDS:
            setattr(s, f, getattr(self, f).ctypes.data)
        s.bin_num = self.bin_num.ctypes.data
        s.first_voxel = self.first_voxel.ctypes.data
        s.n = self.n
        return s

    def select(self, mask):
        h = HostHistories(int(mask.sum()))
        for f in abi.HIST_FLOAT_FIELDS + ("bin_num", "first_voxel"):
            setattr(h, f, np.ascontiguousarray(getattr(self, f)[mask]))
        return h

    def head(self, n):
        h = HostHistories(n)
        for f in abi.HIST_FLOAT_FIELDS + ("bin_num", "first_voxel"):
            setattr(h, f, np.ascontiguousarray(getattr(self, f)[:n]))
        return h

    def arrays(self):
        d = {f: getattr(self, f) for f in abi.HIST_FLOAT_FIELDS}
        d["bin_num"] = self.bin_num
        d["first_voxel"] = self.first_voxel
        return d


def raw_soa(raw: dict):
    """pct_raw_soa over a dict of numpy arrays (keeps them alive in the dict)."""
    s = abi.PctRawSoa()
    raw["gantry_angle"] = np.ascontiguousarray(raw["gantry_angle"], dtype=np.int32)
    s.gantry_angle = raw["gantry_angle"].ctypes.data
    for f in abi.RAW_FLOAT_FIELDS:
        raw[f] = np.ascontiguousarray(raw[f], dtype=np.float32)
        setattr(s, f, raw[f].ctypes.data)
    s.n = len(raw["wepl"])
    return s


def recon_volume_intersections(p, raw: dict):
    s = raw_soa(raw)
    h = HostHistories(s.n)
    missed = np.zeros(s.n, np.uint8)
    hs = h.soa()
    lib().oracle_recon_volume_intersections(ctypes.byref(p), ctypes.byref(s), ctypes.byref(hs), _p(missed))
    return h, missed


def bin_stats_and_cuts(p, h: HostHistories, missed: np.ndarray):
    """A2-A7 on all histories of h (bin_num written into h). Returns dict with
    in_range, counts (of the statistics), mean_*, stddev_*, failed, sinogram,
    cut_counts."""
    nb = p.t_bins * p.v_bins * p.angular_bins
    missed = np.ascontiguousarray(missed, np.uint8)
    in_range = np.zeros(h.n, np.uint8)
    counts = np.zeros(nb, np.int32)
    mw, mu, mv = (np.zeros(nb, np.float32) for _ in range(3))
    sw, su, sv = (np.zeros(nb, np.float32) for _ in range(3))
    hs = h.soa()
    L = lib()
    L.oracle_binning(ctypes.byref(p), ctypes.byref(hs), _p(missed), _p(in_range), _p(counts), _p(mw), _p(mu), _p(mv))
    L.oracle_bin_statistics(ctypes.byref(p), ctypes.byref(hs), _p(missed), _p(counts), _p(mw), _p(mu), _p(mv),
                            _p(sw), _p(su), _p(sv))
    stat_counts = counts.copy()
    failed = np.zeros(h.n, np.uint8)
    sino = np.zeros(nb, np.float32)
    L.oracle_statistical_cuts(ctypes.byref(p), ctypes.byref(hs), _p(missed), _p(counts), _p(mw), _p(mu), _p(mv),
                              _p(sw), _p(su), _p(sv), _p(failed), _p(sino))
    return dict(in_range=in_range, counts=stat_counts, mean_wepl=mw, mean_rel_ut=mu, mean_rel_uv=mv,
                stddev_wepl=sw, stddev_rel_ut=su, stddev_rel_uv=sv, failed=failed, sinogram=sino,
                cut_counts=counts)


def space_carve(p, h: HostHistories, missed, hull):
    hull = np.ascontiguousarray(hull, np.uint8).copy()
    hs = h.soa()
    lib().oracle_space_carve(ctypes.byref(p), ctypes.byref(hs), _p(np.ascontiguousarray(missed, np.uint8)), _p(hull))
    return hull


def msc_counts(p, h: HostHistories, missed):
    counts = hull_init(p).astype(np.int32)
    hs = h.soa()
    lib().oracle_msc_count(ctypes.byref(p), ctypes.byref(hs), _p(np.ascontiguousarray(missed, np.uint8)), _p(counts))
    return counts


def msc_edge(p, counts):
    out = np.zeros_like(counts, dtype=np.int32)
    lib().oracle_msc_edge(ctypes.byref(p), _p(np.ascontiguousarray(counts, np.int32)), _p(out))
    return out


def sm_counts(p, h: HostHistories, missed):
    counts = hull_init(p).astype(np.int32)
    hs = h.soa()
    lib().oracle_sm_count(ctypes.byref(p), ctypes.byref(hs), _p(np.ascontiguousarray(missed, np.uint8)), _p(counts))
    return counts


def sm_edge(p, counts):
    """(SM hull, per-slice thresholds)"""
    out = np.zeros_like(counts, dtype=np.int32)
    thresholds = np.zeros(p.slices, np.int32)
    lib().oracle_sm_edge(ctypes.byref(p), _p(np.ascontiguousarray(counts, np.int32)), _p(out), _p(thresholds))
    return out, thresholds


def hull_select(p, counts):
    hull = np.zeros(counts.size, np.uint8)
    lib().oracle_hull_select(ctypes.byref(p), _p(np.ascontiguousarray(counts, np.int32)), _p(hull))
    return hull


def fbp_filter(p, sino):
    out = np.zeros_like(sino, dtype=np.float32)
    lib().oracle_fbp_filter(ctypes.byref(p), _p(np.ascontiguousarray(sino, np.float32)), _p(out))
    return out


def backprojection(p, filtered):
    img = np.zeros(p.columns * p.rows * p.slices, np.float32)
    lib().oracle_backprojection(ctypes.byref(p), _p(np.ascontiguousarray(filtered, np.float32)), _p(img))
    return img


def fbp_hull(p, image):
    hull = np.zeros(image.size, np.uint8)
    lib().oracle_fbp_hull(ctypes.byref(p), _p(np.ascontiguousarray(image, np.float32)), _p(hull))
    return hull


def initial_iterate(p, fbp, hull):
    x0 = np.zeros(fbp.size, np.float32)
    lib().oracle_initial_iterate(ctypes.byref(p), _p(np.ascontiguousarray(fbp, np.float32)),
                                 _p(np.ascontiguousarray(hull, np.uint8)), _p(x0))
    return x0


def preprocess(p, raw: dict, hull_kind="msc"):
    """The reference's main() up to the DROP loop (.cu:24-226) on the CPU:
    returns (reconstruction histories, x0, hull byte mask, fbp image)."""
    h, missed = recon_volume_intersections(p, dict(raw))
    r = bin_stats_and_cuts(p, h, missed)
    if hull_kind == "msc":
        hull = hull_select(p, msc_edge(p, msc_counts(p, h, missed)))
    elif hull_kind == "sm":
        hull = hull_select(p, sm_edge(p, sm_counts(p, h, missed))[0])
    elif hull_kind == "sc":
        hull = space_carve(p, h, missed, hull_init(p))
    elif hull_kind == "cylinder":
        hull = hull_init(p)
    post = h.select(r["failed"] == 0)
    fbp = backprojection(p, fbp_filter(p, r["sinogram"]))
    if hull_kind == "fbp":
        hull = fbp_hull(p, fbp)
    x0 = initial_iterate(p, fbp, hull)
    inter = mlp_endpoints(p, post, hull)
    recon = post.select(inter == 1)
    return recon, x0, hull, fbp


def hull_init(p):
    hull = np.zeros(p.columns * p.rows * p.slices, np.uint8)
    lib().oracle_hull_init(ctypes.byref(p), _p(hull))
    return hull


def mlp_endpoints(p, h: HostHistories, hull: np.ndarray):
    inter = np.zeros(h.n, np.uint8)
    hs = h.soa()
    lib().oracle_mlp_endpoints(ctypes.byref(p), ctypes.byref(hs), _p(np.ascontiguousarray(hull, np.uint8)), _p(inter))
    return inter


def chord_length(p, xy_in, xy_out, xz_in, xz_out):
    return lib().oracle_chord_length(ctypes.byref(p), xy_in, xy_out, xz_in, xz_out)


def mlp_path(p, h: HostHistories, i: int):
    buf = np.zeros(p.max_intersections + 1, np.uint32)
    skipped = ctypes.c_int(0)
    hs = h.soa()
    n = lib().oracle_mlp_path(ctypes.byref(p), ctypes.byref(hs), i, _p(buf), ctypes.byref(skipped))
    if n < 0:
        raise OverflowError("MLP path longer than max_intersections")
    return buf[:n].copy(), skipped.value


def mlp_paths(p, h: HostHistories, start=0, count=None, stride=None, table=False):
    """Voxel lists of histories [start, start+count) as a list of arrays
    (table=True: the precision form, D4)."""
    count = h.n - start if count is None else count
    if start < 0 or count < 0 or start + count > h.n:
        raise ValueError(f"histories [{start}, {start + count}) outside [0, {h.n})")
    stride = p.max_intersections if stride is None else stride
    vox = np.zeros(count * stride, np.uint32)
    lens = np.zeros(count, np.uint32)
    hs = h.soa()
    fn = lib().oracle_mlp_paths_table if table else lib().oracle_mlp_paths
    rc = fn(ctypes.byref(p), ctypes.byref(hs), start, count, _p(vox), stride, _p(lens))
    if rc != 0:
        raise OverflowError("MLP path overflow in the oracle")
    return [vox[k * stride:k * stride + lens[k]].copy() for k in range(count)]


def mlp_trace(p, h: HostHistories, i: int, cap=4096, table=False):
    """Per-step (u_1, t_1, v_1, x_1, y_1) of the MLP loop of history i, and the
    path length (table=True: the precision form, whose t_1.. are FP32)."""
    arrs = [np.zeros(cap) for _ in range(5)]
    npath = ctypes.c_int(0)
    hs = h.soa()
    k = lib().oracle_mlp_trace(ctypes.byref(p), ctypes.byref(hs), i, *[_p(a) for a in arrs], cap, ctypes.byref(npath),
                               1 if table else 0)
    u1, t1, v1, x1, y1 = (a[:k] for a in arrs)
    return dict(u1=u1, t1=t1, v1=v1, x1=x1, y1=y1, n_path=npath.value)


def drop_lists(V, lists, chord, b, x0, block_size, lam, iterations=1):
    """DROP on explicit voxel lists (see oracle_drop_lists)."""
    offsets = np.zeros(len(lists) + 1, np.int64)
    offsets[1:] = np.cumsum([len(a) for a in lists])
    vox = np.concatenate([np.asarray(a, np.uint32) for a in lists]) if lists else np.zeros(1, np.uint32)
    x = np.ascontiguousarray(x0, np.float32).copy()
    lib().oracle_drop_lists(V, len(lists), _p(offsets), _p(vox), _p(np.ascontiguousarray(chord, np.float64)),
                            _p(np.ascontiguousarray(b, np.float32)), _p(x), int(block_size), float(lam), int(iterations))
    return x


def chord_length_float_avg(p, xy_in, xy_out, xz_in, xz_out):
    """EffectiveChordLength of the float averages (the CPU research call)"""
    return lib().oracle_chord_length_float_avg(ctypes.byref(p), xy_in, xy_out, xz_in, xz_out)


def lists_csr(lists):
    offsets = np.zeros(len(lists) + 1, np.int64)
    offsets[1:] = np.cumsum([len(a) for a in lists])
    vox = np.concatenate([np.asarray(a, np.uint32) for a in lists]) if lists else np.zeros(0, np.uint32)
    return offsets, vox


def drop_lists_research(V, offsets, voxels, errors, chord, b, x, x_update, counts, block_size, lam, eta=2.5,
                        psi_sign=1, variant=2):
    """One iteration of the research loop (oracle_drop_lists_research); x,
    x_update (float64) and counts (uint32) are updated in place."""
    assert x.dtype == np.float32 and x_update.dtype == np.float64 and counts.dtype == np.uint32
    assert x.flags.c_contiguous and x_update.flags.c_contiguous and counts.flags.c_contiguous
    keep = [np.ascontiguousarray(offsets, np.int64), np.ascontiguousarray(voxels, np.uint32),
            None if errors is None else np.ascontiguousarray(errors, np.float32),
            np.ascontiguousarray(chord, np.float64), np.ascontiguousarray(b, np.float32)]
    lib().oracle_drop_lists_research(int(V), len(offsets) - 1, *(_p(a) for a in keep), _p(x), _p(x_update),
                                     _p(counts), int(block_size), float(lam), float(eta), int(psi_sign), int(variant))


def drop(p, h: HostHistories, x0: np.ndarray, iterations=1, block_size=None, lam=None, max_histories=-1,
         threads=1, variant=0):
    """DROP (E6 semantics) from x0; returns the new float32 image. variant 1
    is DROP_blocks_red + DROP_update, 2 DROP_blocks_robust2 +
    DROP_update_robust2 (ETA = p.eta, PSI_SIGN = p.psi_sign). The result does
    not depend on `threads` (paths in parallel, accumulation in storage
    order)."""
    x = np.ascontiguousarray(x0, dtype=np.float32).copy()
    block_size = p.drop_block_size if block_size is None else block_size
    lam = p.lambda_ if lam is None else lam
    hs = h.soa()
    rc = lib().oracle_drop_variant_mt(ctypes.byref(p), ctypes.byref(hs), _p(x), int(block_size), float(lam),
                                      int(iterations), int(max_histories), int(threads), int(variant))
    if rc != 0:
        raise OverflowError("MLP path overflow in the oracle DROP")
    return x


# ---------------------------------------------------------------------------
# learned-path form (oracle_net.cpp; the network itself is oracle/net.py)
# ---------------------------------------------------------------------------
def mlp_points(p, h: HostHistories, i: int, u1):
    """exact FP64 MLP (t, v) of history i at the depths u1"""
    u1 = np.ascontiguousarray(u1, np.float64)
    t = np.zeros(u1.size)
    v = np.zeros(u1.size)
    hs = h.soa()
    lib().oracle_mlp_points(ctypes.byref(p), ctypes.byref(hs), i, u1.size, _p(u1), _p(t), _p(v))
    return t, v


def net_features(p, h: HostHistories, targets=True):
    """(features[n][12], targets[n][8] or None, cells[n], u_2[n])"""
    f = np.zeros((h.n, 12), np.float32)
    t = np.zeros((h.n, 8), np.float32) if targets else None
    cells = np.zeros(h.n, np.int32)
    u2 = np.zeros(h.n)
    hs = h.soa()
    lib().oracle_net_features(ctypes.byref(p), ctypes.byref(hs), h.n, _p(f), _p(t), _p(cells), _p(u2))
    return f, t, cells, u2


def surrogate_path(p, h: HostHistories, i: int, out8, cap=4096):
    """FP32 surrogate walk of history i for network outputs out8: (path,
    per-step t, per-step v)"""
    out8 = np.ascontiguousarray(out8, np.float32)
    buf = np.zeros(p.max_intersections + 1, np.uint32)
    tr = np.zeros(cap, np.float32)
    vr = np.zeros(cap, np.float32)
    nst = ctypes.c_int(0)
    hs = h.soa()
    n = lib().oracle_surrogate_path(ctypes.byref(p), ctypes.byref(hs), i, _p(out8), _p(buf), _p(tr), _p(vr), cap,
                                    ctypes.byref(nst))
    if n < 0:
        raise OverflowError("surrogate path longer than max_intersections")
    k = min(nst.value, cap)
    return buf[:n].copy(), tr[:k].copy(), vr[:k].copy()
