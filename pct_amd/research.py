"""The reference's CPU research reconstruction (image_reconstruction,
.cu:7413-7570) on the GPU.

That loop visits the reconstruction histories in a cyclic order
(generate_history_sequence, .cu:7225-7230, PRIME_OFFSET data.h:131), takes
their MLP voxel lists from an MLP_paths file and per-intersection path errors
from an MLP_path_error file, and applies one of the research update rules per
list (DROP_blocks, DROP_blocks_red, DROP_blocks_error_red,
DROP_blocks_error, DROP_blocks_error_red_sp) with DROP_update after every
full block. Here the files are read and written by libpct (include/pct_io.h),
the iteration runs in pct_drop_lists_iteration (include/pct.h), and the
update state (x_update, counts) is kept between iterations as the
reference's globals are.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import PCT_LISTS_ERROR_RED, PCT_LISTS_REUSE

PRIME_OFFSET = 26410633


class ResearchIOError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        msg = _lib.lib().pct_io_last_error()
        raise ResearchIOError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _b(s):
    return s.encode() if isinstance(s, str) else s


def history_sequence(n, offset_prime=PRIME_OFFSET):
    """generate_history_sequence (.cu:7225): s[0] = 0, s[i] = (s[i-1] +
    offset_prime) % n, i.e. s[i] = i * offset_prime mod n"""
    if n <= 0:
        return np.zeros(0, np.int64)
    i = np.arange(n, dtype=np.uint64)
    return ((i * np.uint64(offset_prime % n)) % np.uint64(n)).astype(np.int64)


def paths_file_name(directory, folder, radius, name="MLP_paths"):
    """"<directory><folder>/<name>_r=<HULL_FILTER_RADIUS>.bin" (.cu:7479, 7509;
    the error file's name is "MLP_path_error")"""
    return "%s%s/%s_r=%d.bin" % (directory, folder, name, int(radius))


def write_mlp_paths(path, lists):
    """write_MLP_path for every list, after the path-count line"""
    offsets = np.zeros(len(lists) + 1, np.int64)
    offsets[1:] = np.cumsum([len(a) for a in lists])
    vox = np.ascontiguousarray(np.concatenate([np.asarray(a, np.uint32) for a in lists]) if lists
                               else np.zeros(1, np.uint32))
    _check(_lib.lib().pct_mlp_paths_file_write(_b(path), len(lists), offsets.ctypes.data, vox.ctypes.data),
           f"pct_mlp_paths_file_write({path})")


def read_mlp_paths(path, max_paths=None):
    """(offsets int64 n+1, voxels uint32) of the first min(header count,
    max_paths) paths"""
    nh, npaths, total = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    L = _lib.lib()
    _check(L.pct_mlp_paths_file_info(_b(path), ctypes.byref(nh), ctypes.byref(npaths), ctypes.byref(total)),
           f"pct_mlp_paths_file_info({path})")
    want = nh.value if max_paths is None else min(nh.value, int(max_paths))
    offsets = np.zeros(want + 1, np.int64)
    vox = np.zeros(max(total.value, 1), np.uint32)
    got = ctypes.c_int64()
    _check(L.pct_mlp_paths_file_read(_b(path), want, total.value, offsets.ctypes.data, vox.ctypes.data,
                                     ctypes.byref(got)), f"pct_mlp_paths_file_read({path})")
    return offsets[:got.value + 1], vox[:offsets[got.value]]


def write_path_errors(path, errors):
    e = np.ascontiguousarray(errors, np.float32)
    _check(_lib.lib().pct_path_errors_write(_b(path), e.size, e.ctypes.data if e.size else None),
           f"pct_path_errors_write({path})")


def read_path_errors(path, count):
    e = np.zeros(max(int(count), 1), np.float32)
    _check(_lib.lib().pct_path_errors_read(_b(path), int(count), e.ctypes.data), f"pct_path_errors_read({path})")
    return e[:int(count)]


class ResearchLoop:
    """image_reconstruction's iterations on one Context: lists and errors in
    device memory, the reconstruction histories' WEPL and chord lengths in
    list order (history_sequence order when the lists were written in it),
    x_update / counts kept between iterations."""

    def __init__(self, ctx, offsets, voxels, wepl, chord, errors=None, variant=PCT_LISTS_ERROR_RED):
        torch = ctx.torch
        dev = ctx.device
        self.ctx = ctx
        self.variant = variant
        self.offsets = torch.as_tensor(np.asarray(offsets, np.int64)).to(dev)
        self.voxels = torch.as_tensor(np.asarray(voxels, np.uint32).view(np.int32)).to(dev)
        self.errors = None if errors is None else torch.as_tensor(np.asarray(errors, np.float32)).to(dev)
        self.wepl = torch.as_tensor(wepl).to(device=dev, dtype=torch.float32)
        self.chord = torch.as_tensor(chord).to(device=dev, dtype=torch.float64)
        self.x_update = torch.zeros(ctx.V, dtype=torch.float64, device=dev)
        self.counts = torch.zeros(ctx.V, dtype=torch.int32, device=dev)
        self._sorted_for = None

    def iterate(self, x, block_size=None, lam=None):
        """one iteration; the lists never change, so from the second call on
        with the same block size their sorted order is reused"""
        block_size = self.ctx.params.drop_block_size if block_size is None else block_size
        rule = self.variant | (PCT_LISTS_REUSE if self._sorted_for == block_size else 0)
        self.ctx.drop_lists_iteration(self.offsets, self.voxels, self.errors, self.chord, self.wepl, x,
                                      self.x_update, self.counts, block_size, lam, rule)
        self._sorted_for = block_size
        return x
