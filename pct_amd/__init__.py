"""pct_amd -- MI355X-native proton-CT reconstruction engine (DROP hot path).

Host-side mirror of the reference's stage functions (lynch829/pCT,
src/pCT_Reconstruction_Data_Segments.cu) over the C ABI of libpct.so
(include/pct.h). Device memory is held in torch tensors (plumbing only); all
compute runs in the hand-written HIP kernels of libpct.so. There is no CPU
fallback: every entry point raises if libpct.so is missing or a call fails.
"""
import ctypes

from . import _lib
from . import configs
from ._lib import (PCT_LISTS_DROP, PCT_LISTS_DROP_RED, PCT_LISTS_ERROR, PCT_LISTS_ERROR_RED,
                   PCT_LISTS_ERROR_RED_SP, PCT_LISTS_REUSE, PCT_DROP_ATOMIC, PCT_DROP_CACHED, PCT_DROP_F32, PCT_DROP_NO_GRAPH, PCT_DROP_ROBUST,
                   PCT_DROP_ROBUST2, PCT_DROP_PERSISTENT, PCT_LEARNED, PCT_MLP_FP64, PCT_MLP_SURROGATE, PCT_MLP_TABLE, PCT_PRECISION,
                   PctHistSoa, PctParams, PctRawSoa, PctSynthCfg, PctSynthSoa)

__all__ = ["Context", "Batch", "Histories", "RawScan", "configs", "PCT_MLP_FP64", "PCT_MLP_TABLE", "PCT_MLP_SURROGATE",
           "PCT_PRECISION", "PCT_LEARNED", "PCT_DROP_ROBUST", "PCT_DROP_ROBUST2", "PCT_DROP_PERSISTENT",
           "generate_host"]


class PctError(RuntimeError):
    pass


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class RawScan:
    """Raw tracker arrays (pct_raw_soa) in torch tensors."""

    def __init__(self, n, device):
        import torch
        self.n = n
        self.gantry_angle = torch.empty(n, dtype=torch.int32, device=device)
        for f in _lib.RAW_FLOAT_FIELDS:
            setattr(self, f, torch.empty(n, dtype=torch.float32, device=device))

    def soa(self):
        s = PctRawSoa()
        s.gantry_angle = self.gantry_angle.data_ptr()
        for f in _lib.RAW_FLOAT_FIELDS:
            setattr(s, f, getattr(self, f).data_ptr())
        s.n = self.n
        return s

    def synth_soa(self):
        s = PctSynthSoa()
        s.gantry_angle = self.gantry_angle.data_ptr()
        for f in _lib.RAW_FLOAT_FIELDS:
            setattr(s, f, getattr(self, f).data_ptr())
        return s

    def to(self, device):
        r = RawScan.__new__(RawScan)
        r.n = self.n
        r.gantry_angle = self.gantry_angle.to(device)
        for f in _lib.RAW_FLOAT_FIELDS:
            setattr(r, f, getattr(self, f).to(device))
        return r

    def numpy(self):
        d = {"gantry_angle": self.gantry_angle.cpu().numpy()}
        for f in _lib.RAW_FLOAT_FIELDS:
            d[f] = getattr(self, f).cpu().numpy()
        return d


class Histories:
    """Per-history SoA (pct_hist_soa, data.h:577-590) in torch tensors."""

    def __init__(self, n, device, zero=True):
        """zero=False leaves the 11 float fields uninitialised, for producers
        that write every one of them; bin_num and first_voxel start at 0
        unless `zero` is None (a compaction writes every field)"""
        import torch
        self.n = n
        alloc = torch.zeros if zero else torch.empty
        for f in _lib.HIST_FLOAT_FIELDS:
            setattr(self, f, alloc(n, dtype=torch.float32, device=device))
        alloc = torch.empty if zero is None else torch.zeros
        self.bin_num = alloc(n, dtype=torch.int32, device=device)
        self.first_voxel = alloc(n, dtype=torch.int32, device=device)   # u32 bit pattern

    def soa(self):
        s = PctHistSoa()
        for f in _lib.HIST_FLOAT_FIELDS:
            setattr(s, f, getattr(self, f).data_ptr())
        s.bin_num = self.bin_num.data_ptr()
        s.first_voxel = self.first_voxel.data_ptr()
        s.n = self.n
        return s

    def to(self, device):
        h = Histories.__new__(Histories)
        h.n = self.n
        for f in _lib.HIST_FLOAT_FIELDS + ("bin_num", "first_voxel"):
            setattr(h, f, getattr(self, f).to(device))
        return h

    def head(self, n):
        h = Histories.__new__(Histories)
        h.n = n
        for f in _lib.HIST_FLOAT_FIELDS + ("bin_num", "first_voxel"):
            setattr(h, f, getattr(self, f)[:n])
        return h

    def slice(self, start, n):
        """histories [start, start + n) as views of these arrays"""
        h = Histories.__new__(Histories)
        h.n = n
        for f in _lib.HIST_FLOAT_FIELDS + ("bin_num", "first_voxel"):
            setattr(h, f, getattr(self, f)[start:start + n])
        return h

    def numpy(self):
        d = {f: getattr(self, f).cpu().numpy() for f in _lib.HIST_FLOAT_FIELDS}
        d["bin_num"] = self.bin_num.cpu().numpy()
        d["first_voxel"] = self.first_voxel.cpu().numpy().view("uint32")
        return d


class _StreamOrdered:
    """libpct's functions, each call ordered after the work already queued on
    torch's current stream and before whatever torch queues there next.

    The context runs on a stream of its own (torch's default stream is the
    null stream, which cannot be captured into a hipGraph). Tensors are
    allocated and read back on torch's current stream, so every library call
    waits for that stream and that stream waits for the call: an event record
    and a device-side wait on each side, no host synchronisation."""

    def __init__(self, lib, stream, torch):
        self._lib, self._stream, self._torch = lib, stream, torch

    def __getattr__(self, name):
        fn = getattr(self._lib, name)
        stream, torch = self._stream, self._torch

        def call(*args):
            cur = torch.cuda.current_stream(stream.device)
            stream.wait_stream(cur)
            try:
                return fn(*args)
            finally:
                cur.wait_stream(stream)
        return call


class Context:
    """A pct_ctx on the current HIP device (pct_create, include/pct.h)."""

    def __init__(self, params: PctParams, device="cuda"):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise PctError("pct_amd.Context needs a GPU device")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        torch.cuda.set_device(self.device)
        self.params = params.copy()
        lib = _lib.lib()
        self.stream = torch.cuda.Stream(self.device)
        self._ctx = lib.pct_create(ctypes.byref(self.params), ctypes.c_void_p(self.stream.cuda_stream))
        if not self._ctx:
            raise PctError("pct_create failed")
        self.L = _StreamOrdered(lib, self.stream, torch)
        self.V = params.columns * params.rows * params.slices

    def close(self):
        if getattr(self, "_ctx", None):
            self.L._lib.pct_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.pct_last_error(self._ctx)
            raise PctError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def sync(self):
        self._check(self.L.pct_sync(self._ctx), "pct_sync")

    # -- synthetic input ----------------------------------------------------
    def generate(self, cfg: PctSynthCfg, start=0, count=None) -> RawScan:
        count = cfg.n_histories - start if count is None else count
        raw = RawScan(count, self.device)
        soa = raw.synth_soa()
        self._check(self.L.pct_synth_generate(self._ctx, ctypes.byref(cfg), start, count, ctypes.byref(soa)),
                    "pct_synth_generate")
        return raw

    # -- (a) recon-volume intersection (.cu:2066-2168 / 2169-2387) ------------
    def recon_volume_intersections(self, raw: RawScan):
        torch = self.torch
        hist = Histories(raw.n, self.device, zero=False)      # the kernel writes the 11 float fields
        missed = torch.empty(raw.n, dtype=torch.uint8, device=self.device)
        rs, hs = raw.soa(), hist.soa()
        self._check(self.L.pct_recon_volume_intersections(self._ctx, ctypes.byref(rs), ctypes.byref(hs), _ptr(missed)),
                    "pct_recon_volume_intersections")
        return hist, missed

    # -- (a) binning, statistics, cuts (.cu:2388-2918) -----------------------------
    @property
    def n_bins(self):
        p = self.params
        return p.t_bins * p.v_bins * p.angular_bins

    def new_bin_stats(self):
        torch = self.torch
        st = {"counts": torch.zeros(self.n_bins, dtype=torch.int32, device=self.device)}
        for f in _lib.BIN_STAT_FIELDS[1:]:
            st[f] = torch.zeros(self.n_bins, dtype=torch.float32, device=self.device)
        return st

    @staticmethod
    def _bin_soa(st):
        s = _lib.PctBinStats()
        for f in _lib.BIN_STAT_FIELDS:
            setattr(s, f, st[f].data_ptr())
        return s

    def binning(self, hist, missed):
        """binning_GPU: writes hist.bin_num, returns in_range (uint8)."""
        in_range = self.torch.empty(hist.n, dtype=self.torch.uint8, device=self.device)
        hs = hist.soa()
        self._check(self.L.pct_binning(self._ctx, ctypes.byref(hs), _ptr(missed), _ptr(in_range)), "pct_binning")
        return in_range

    def bin_statistics(self, hist, missed, in_range):
        st = self.new_bin_stats()
        hs, ss = hist.soa(), self._bin_soa(st)
        self._check(self.L.pct_bin_statistics(self._ctx, ctypes.byref(hs), _ptr(missed), _ptr(in_range),
                                              ctypes.byref(ss)), "pct_bin_statistics")
        return st

    def statistical_cuts(self, hist, missed, st):
        """3-sigma cuts + sinogram; returns (failed uint8, sinogram f32). st['counts'] becomes the survivors' count."""
        failed = self.torch.empty(hist.n, dtype=self.torch.uint8, device=self.device)
        sino = self.torch.zeros(self.n_bins, dtype=self.torch.float32, device=self.device)
        hs, ss = hist.soa(), self._bin_soa(st)
        self._check(self.L.pct_statistical_cuts(self._ctx, ctypes.byref(hs), _ptr(missed), ctypes.byref(ss),
                                                _ptr(failed), _ptr(sino)), "pct_statistical_cuts")
        return failed, sino

    # -- (c) hull ---------------------------------------------------------------
    def hull_init(self):
        words = (self.V + 31) // 32
        bits = self.torch.empty(words, dtype=self.torch.int32, device=self.device)
        self._check(self.L.pct_hull_init(self._ctx, _ptr(bits)), "pct_hull_init")
        return bits

    def space_carve(self, hist, missed, hull_bits):
        hs = hist.soa()
        self._check(self.L.pct_space_carve(self._ctx, ctypes.byref(hs), _ptr(missed), _ptr(hull_bits)),
                    "pct_space_carve")
        return hull_bits

    def msc_counts(self, hist, missed, counts=None):
        """MSC counts (initialized to the cylinder when counts is None)."""
        if counts is None:
            counts = self.torch.empty(self.V, dtype=self.torch.int32, device=self.device)
            self._check(self.L.pct_msc_init(self._ctx, _ptr(counts)), "pct_msc_init")
        hs = hist.soa()
        self._check(self.L.pct_msc_count(self._ctx, ctypes.byref(hs), _ptr(missed), _ptr(counts)), "pct_msc_count")
        return counts

    def msc_edge(self, counts):
        out = self.torch.empty_like(counts)
        self._check(self.L.pct_msc_edge(self._ctx, _ptr(counts), _ptr(out)), "pct_msc_edge")
        return out

    def sm_counts(self, hist, missed, counts=None):
        """SM counts (initialized to the cylinder when counts is None)."""
        if counts is None:
            counts = self.torch.empty(self.V, dtype=self.torch.int32, device=self.device)
            self._check(self.L.pct_msc_init(self._ctx, _ptr(counts)), "pct_msc_init")
        hs = hist.soa()
        self._check(self.L.pct_sm_count(self._ctx, ctypes.byref(hs), _ptr(missed), _ptr(counts)), "pct_sm_count")
        return counts

    def sm_edge(self, counts, thresholds=None):
        """SM edge detection; with a tensor as `thresholds` (slices int32)
        the per-slice thresholds are written into it."""
        out = self.torch.empty_like(counts)
        self._check(self.L.pct_sm_edge(self._ctx, _ptr(counts), _ptr(out), _ptr(thresholds)), "pct_sm_edge")
        return out

    def hull_select(self, counts):
        bits = self.torch.empty((self.V + 31) // 32, dtype=self.torch.int32, device=self.device)
        self._check(self.L.pct_hull_select(self._ctx, _ptr(counts), _ptr(bits)), "pct_hull_select")
        return bits

    # -- (b) FBP (.cu:2919-3261) and the initial iterate (.cu:7191) -------------
    def fbp_filter(self, sinogram):
        out = self.torch.empty_like(sinogram)
        self._check(self.L.pct_fbp_filter(self._ctx, _ptr(sinogram), _ptr(out)), "pct_fbp_filter")
        return out

    def backproject(self, filtered):
        img = self.torch.empty(self.V, dtype=self.torch.float32, device=self.device)
        self._check(self.L.pct_backproject(self._ctx, _ptr(filtered), _ptr(img)), "pct_backproject")
        return img

    def fbp_threshold(self, image):
        bits = self.torch.empty((self.V + 31) // 32, dtype=self.torch.int32, device=self.device)
        self._check(self.L.pct_fbp_threshold(self._ctx, _ptr(image), _ptr(bits)), "pct_fbp_threshold")
        return bits

    def initial_iterate(self, fbp_image, hull_bits):
        x0 = self.torch.empty(self.V, dtype=self.torch.float32, device=self.device)
        self._check(self.L.pct_initial_iterate(self._ctx, _ptr(fbp_image), _ptr(hull_bits), _ptr(x0)),
                    "pct_initial_iterate")
        return x0

    # -- (d) endpoints + compaction (reconstruction_cuts, .cu:6142-6300) --------
    def mlp_endpoints(self, hist: Histories, hull_bits):
        inter = self.torch.empty(hist.n, dtype=self.torch.uint8, device=self.device)
        hs = hist.soa()
        self._check(self.L.pct_mlp_endpoints(self._ctx, ctypes.byref(hs), _ptr(hull_bits), _ptr(inter)),
                    "pct_mlp_endpoints")
        return inter

    def compact(self, hist: Histories, keep) -> Histories:
        out = Histories(hist.n, self.device, zero=None)       # every field of every kept row is written
        n_kept = ctypes.c_int64(0)
        s, d = hist.soa(), out.soa()
        self._check(self.L.pct_compact(self._ctx, ctypes.byref(s), _ptr(keep), ctypes.byref(d), ctypes.byref(n_kept)),
                    "pct_compact")
        return out.head(n_kept.value)

    # -- (d) MLP paths ----------------------------------------------------------
    def mlp_capacity(self, hist: Histories, start=0, count=None):
        count = hist.n - start if count is None else count
        offs = self.torch.empty(count + 1, dtype=self.torch.int64, device=self.device)
        total = ctypes.c_int64(0)
        hs = hist.soa()
        self._check(self.L.pct_mlp_capacity(self._ctx, ctypes.byref(hs), start, count, _ptr(offs), ctypes.byref(total)),
                    "pct_mlp_capacity")
        return offs, total.value

    def mlp_paths(self, hist: Histories, start=0, count=None, mode=PCT_MLP_FP64):
        count = hist.n - start if count is None else count
        offs, total = self.mlp_capacity(hist, start, count)
        vox = self.torch.zeros(max(total, 1), dtype=self.torch.int32, device=self.device)
        lens = self.torch.zeros(max(count, 1), dtype=self.torch.int32, device=self.device)
        hs = hist.soa()
        self._check(self.L.pct_mlp_paths(self._ctx, ctypes.byref(hs), start, count, _ptr(offs), _ptr(vox), _ptr(lens),
                                         mode), "pct_mlp_paths")
        return offs, vox, lens[:count]

    # -- (e) DROP -----------------------------------------------------------------
    def drop_block_update(self, hist, start, count, x, x_update, counts, lam, mode=PCT_MLP_FP64):
        hs = hist.soa()
        self._check(self.L.pct_drop_block_update(self._ctx, ctypes.byref(hs), start, count, _ptr(x), _ptr(x_update),
                                                 _ptr(counts), float(lam), mode), "pct_drop_block_update")

    def drop_block_update_noalloc(self, hist, start, count, x, x_update, counts, lam, mode=PCT_MLP_FP64):
        """pct_drop_block_update_noalloc: the same update with the lists in the
        library's fixed device array (no allocation; at most 8192 histories)."""
        hs = hist.soa()
        self._check(self.L.pct_drop_block_update_noalloc(self._ctx, ctypes.byref(hs), start, count, _ptr(x),
                                                         _ptr(x_update), _ptr(counts), float(lam), mode),
                    "pct_drop_block_update_noalloc")

    def drop_image_update(self, x, x_update, counts):
        self._check(self.L.pct_drop_image_update(self._ctx, _ptr(x), _ptr(x_update), _ptr(counts)),
                    "pct_drop_image_update")

    def drop_prepare(self, hist, block_size=None, mode=PCT_MLP_FP64):
        block_size = self.params.drop_block_size if block_size is None else block_size
        hs = hist.soa()
        self._check(self.L.pct_drop_prepare(self._ctx, ctypes.byref(hs), int(block_size), mode), "pct_drop_prepare")

    def drop_iteration(self, hist, x, block_size=None, lam=None, mode=PCT_MLP_FP64):
        """One DROP iteration (mode: MLP mode | PCT_DROP_* flags)."""
        block_size = self.params.drop_block_size if block_size is None else block_size
        lam = self.params.lambda_ if lam is None else lam
        hs = hist.soa()
        self._check(self.L.pct_drop_iteration(self._ctx, ctypes.byref(hs), _ptr(x), int(block_size), float(lam), mode),
                    "pct_drop_iteration")

    def drop_invalidate(self):
        """Marks the DROP plan stale: the next drop_iteration rebuilds it (after
        changing histories in place when iterating with PCT_DROP_CACHED)."""
        self._check(self.L.pct_drop_invalidate(self._ctx), "pct_drop_invalidate")


    def chord_lengths(self, hist, float_average=False):
        """effective chord length per history (float64; float_average: the
        CPU research loop's float average of the angles)"""
        out = self.torch.empty(max(hist.n, 1), dtype=self.torch.float64, device=self.device)
        hs = hist.soa()
        self._check(self.L.pct_chord_lengths(self._ctx, ctypes.byref(hs), int(bool(float_average)), _ptr(out)),
                    "pct_chord_lengths")
        return out[:hist.n]

    def drop_lists_iteration(self, offsets, voxels, errors, chord, wepl, x, x_update, counts, block_size=None,
                             lam=None, variant=PCT_LISTS_ERROR_RED):
        """One iteration of the research loop over given lists (device
        tensors: offsets int64 n+1, voxels int32/uint32, errors float32 or
        None, chord float64, wepl float32; x float32, x_update float64 and
        counts int32 of V values, kept between calls)."""
        block_size = self.params.drop_block_size if block_size is None else block_size
        lam = self.params.lambda_ if lam is None else lam
        n = offsets.numel() - 1
        assert x.numel() == self.V and x_update.numel() == self.V and counts.numel() == self.V
        assert x_update.dtype == self.torch.float64 and counts.dtype in (self.torch.int32, self.torch.uint32)
        assert chord.numel() >= n and wepl.numel() >= n and (errors is None or errors.numel() >= voxels.numel())
        self._check(self.L.pct_drop_lists_iteration(self._ctx, n, _ptr(offsets), _ptr(voxels), _ptr(errors),
                                                    _ptr(chord), _ptr(wepl), _ptr(x), _ptr(x_update), _ptr(counts),
                                                    int(block_size), float(lam), int(variant)),
                    "pct_drop_lists_iteration")


def set_profiling(ctx: "Context", on: bool):
    ctx._check(ctx.L.pct_set_profiling(ctx._ctx, 1 if on else 0), "pct_set_profiling")


def last_timings(ctx: "Context"):
    """[capacity ms, MLP ms, block loop ms] of the last profiled iteration."""
    out = (ctypes.c_double * 3)()
    ctx._check(ctx.L.pct_last_timings(ctx._ctx, out), "pct_last_timings")
    return list(out)


class Batch:
    """Several scans reconstructed together (pct_batch_*, include/pct.h): one
    Context per scan (its own grid and parameters); the batch runs on the
    first context's stream."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        n = len(self.contexts)
        self.L = _lib.lib()
        arr = (ctypes.c_void_p * n)(*[c._ctx for c in self.contexts])
        self._b = self.L.pct_batch_create(arr, n)
        if not self._b:
            raise PctError("pct_batch_create failed (1..16 contexts on one device)")

    def close(self):
        if getattr(self, "_b", None):
            self.L.pct_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, fn, *args, what=""):
        c0 = self.contexts[0]
        torch = c0.torch
        cur = torch.cuda.current_stream(c0.device)
        for c in self.contexts:
            c.stream.wait_stream(cur)
            if c is not c0:
                c0.stream.wait_stream(c.stream)
        try:
            rc = fn(self._b, *args)
        finally:
            cur.wait_stream(c0.stream)
            for c in self.contexts[1:]:
                cur.wait_stream(c.stream)
        if rc != 0:
            msg = self.L.pct_last_error(c0._ctx)
            raise PctError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def _soas(self, hists):
        n = len(self.contexts)
        if len(hists) != n:
            raise PctError(f"{len(hists)} history sets for {n} scans")
        return (PctHistSoa * n)(*[h.soa() for h in hists])

    def prepare(self, hists, block_sizes, mode=PCT_MLP_FP64):
        n = len(self.contexts)
        self._call(self.L.pct_batch_prepare, self._soas(hists), (ctypes.c_int * n)(*block_sizes), int(mode),
                   what="pct_batch_prepare")

    def iteration(self, hists, xs, block_sizes, lambdas=None, mode=PCT_MLP_FP64):
        """One DROP iteration of every scan (xs: one float32 image per scan)."""
        n = len(self.contexts)
        lambdas = [c.params.lambda_ for c in self.contexts] if lambdas is None else lambdas
        self._call(self.L.pct_batch_iteration, self._soas(hists), (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs]),
                   (ctypes.c_int * n)(*block_sizes), (ctypes.c_float * n)(*lambdas), int(mode),
                   what="pct_batch_iteration")


TUNING_KEYS = {"mlp_waves": _lib.PCT_TUNE_MLP_WAVES, "chunk_first": _lib.PCT_TUNE_CHUNK_FIRST,
               "chunk_max": _lib.PCT_TUNE_CHUNK_MAX, "overlap": _lib.PCT_TUNE_OVERLAP,
               "block_form": _lib.PCT_TUNE_BLOCK_FORM, "plan_limit_mb": _lib.PCT_TUNE_PLAN_LIMIT_MB,
               "wait_give_up": _lib.PCT_TUNE_WAIT_GIVE_UP}


def set_tuning(ctx: "Context", **knobs):
    """Execution knobs of the DROP iteration (pct_set_tuning): mlp_waves,
    chunk_first, chunk_max, overlap (0: the MLP check beside the block loop,
    joined at the end; 1: beside, each chunk waiting for its check; 2:
    sequential), block_form (1 one
    launch per block recomputing the previous block's updates it needs, 2 two
    launches per block, 3 one launch per block waiting for the previous block's
    stored updates); plan_limit_mb; wait_give_up (diagnostics of block form 3 and
    of a persistent batch whose first context this is: their waits give up,
    to exercise the failure path); 0 = default."""
    for k, v in knobs.items():
        if k not in TUNING_KEYS:
            raise PctError(f"unknown tuning knob {k}")
        ctx._check(ctx.L.pct_set_tuning(ctx._ctx, TUNING_KEYS[k], int(v)), "pct_set_tuning")


def drop_stats(ctx: "Context", hist: "Histories", block_size: int):
    """(sum of path lengths, sum over blocks of touched voxels) of one iteration."""
    a, b = ctypes.c_int64(0), ctypes.c_int64(0)
    hs = hist.soa()
    ctx._check(ctx.L.pct_drop_stats(ctx._ctx, ctypes.byref(hs), int(block_size), ctypes.byref(a), ctypes.byref(b)),
               "pct_drop_stats")
    return a.value, b.value


def selftest_math(ctx: "Context", func: int, a, b=None):
    """Device evaluation of one pct_math.h function (diagnostics)."""
    import torch
    a = torch.as_tensor(a, dtype=torch.float64).to(ctx.device).contiguous()
    b = None if b is None else torch.as_tensor(b, dtype=torch.float64).to(ctx.device).contiguous()
    out = torch.empty_like(a)
    ctx._check(ctx.L.pct_selftest_math(ctx._ctx, func, _ptr(a), _ptr(b), _ptr(out), a.numel()), "pct_selftest_math")
    ctx.sync()
    return out.cpu().numpy()


def generate_host(cfg: PctSynthCfg, start=0, count=None) -> RawScan:
    """The synthetic scan generated on the host (bit-identical to the device)."""
    count = cfg.n_histories - start if count is None else count
    raw = RawScan(count, "cpu")
    soa = raw.synth_soa()
    rc = _lib.lib().pct_synth_generate_host(ctypes.byref(cfg), start, count, ctypes.byref(soa))
    if rc != 0:
        raise PctError(f"pct_synth_generate_host failed ({rc})")
    return raw
