"""CPU: the capacity rule of pct_drop_iteration (pct_drop_plan_bytes /
pct_drop_plan_choice): when a scan's DROP plan cannot be built on the device
-- its segment records index at most 2^22 voxels and 8191 histories per
block, and its peak bytes must fit the free memory -- the iteration runs the
reference's structure over chunks of blocks instead. Pure functions of libpct,
called without a device."""
import ctypes

import pytest

from pct_amd import _lib

GiB = 1 << 30


def _lib_or_skip():
    try:
        return _lib.lib()
    except Exception as e:       # no libpct built
        pytest.skip(f"libpct.so not loadable: {e}")


def test_plan_bytes_terms():
    L = _lib_or_skip()
    base = L.pct_drop_plan_bytes(0, 0, 0, 0, 0)
    assert base == 4 * GiB
    assert L.pct_drop_plan_bytes(0, 1000, 0, 0, 0) - base == 4 * 1000          # build store
    assert L.pct_drop_plan_bytes(0, 0, 1000, 0, 0) - base == 4 * 1000          # list store
    assert L.pct_drop_plan_bytes(0, 0, 1000, 0, 1) - base == 12 * 1000         # + previous-block records
    assert L.pct_drop_plan_bytes(0, 0, 0, 1000, 0) - base == 10 * 1000         # segment record + csc index
    assert L.pct_drop_plan_bytes(1000, 0, 0, 0, 0) - base == 128 * 1000         # per history


def _choice(L, n, slots, list_slots, entries, V, B, budget, prev=0):
    need = ctypes.c_int64(0)
    c = L.pct_drop_plan_choice(n, slots, list_slots, entries, V, B, budget, prev, ctypes.byref(need))
    return c, need.value


def test_plan_choice_limits():
    L = _lib_or_skip()
    big = 1 << 62
    assert _choice(L, 1000, 1000, 1000, 1000, 1 << 22, 8191, big)[0] == 1
    assert _choice(L, 1000, 1000, 1000, 1000, (1 << 22) + 1, 3200, big)[0] == 0     # voxel field
    assert _choice(L, 1000, 1000, 1000, 1000, 1000, 8192, big)[0] == 0              # history index field
    assert _choice(L, 1000, 1000, 1000, 1000, 1000, 0, big)[0] == 0
    c, need = _choice(L, 1000, 1000, 1000, 1000, 1000, 3200, big)
    assert _choice(L, 1000, 1000, 1000, 1000, 1000, 3200, need)[0] == 1            # exactly fits
    assert _choice(L, 1000, 1000, 1000, 1000, 1000, 3200, need - 1)[0] == 0


def test_plan_choice_at_the_configurations():
    """the 200x200x32 / 20 M scan and the 256x256x48 / 50 M scan build plans on
    a 288 GB device; a scan of the reference's own size (NUM_RECON_HISTORIES =
    105642524, include/pCT_data.h:130) on the 200x200x32 grid passes the first
    check (before its MLP pass) but not the second, and runs in chunks"""
    L = _lib_or_skip()
    hbm = 285 * 10**9
    # measured at 200x200x32: 17.04 M histories, 5.34e9 capacity slots, 4.5e9 list slots, 2.99e9 entries
    assert _choice(L, 17_036_957, 5_340_000_000, 0, 0, 1_280_000, 3200, hbm)[0] == 1
    assert _choice(L, 17_036_957, 5_340_000_000, 4_500_000_000, 2_988_000_000, 1_280_000, 3200, hbm)[0] == 1
    # 256x256x48 / 42.07 M reconstruction histories: ~403 slots, ~230 entries and ~300 list slots per history
    n = 42_074_591
    assert _choice(L, n, 403 * n, 300 * n, 230 * n, 256 * 256 * 48, 3200, hbm)[0] == 1
    # the reference's history count with the 200x200x32 proportions
    n = 105_642_524
    f = n / 17_036_957
    c1, need1 = _choice(L, n, int(5.34e9 * f), 0, 0, 1_280_000, 3200, hbm)
    c2, need2 = _choice(L, n, int(5.34e9 * f), int(4.5e9 * f), int(2.988e9 * f), 1_280_000, 3200, hbm)
    assert c1 == 1 and c2 == 0 and need2 > hbm > need1
