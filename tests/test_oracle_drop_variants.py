"""Pins the oracle's research DROP variants (oracle_drop_variant_mt, restating
DROP_blocks_red + DROP_update, .cu:8005-8055, and DROP_blocks_robust2 +
DROP_update_robust2, .cu:7871-7901):

  - against an independent dense formulation in numpy (per block an
    occurrence matrix of the histories' voxel lists; per-voxel psi from the
    block's starting image, per-history residuals, N_v = sum of L^2), to
    double rounding;
  - with ETA = 0 both variants are the same update, bit for bit, and within
    float rounding of the plain DROP;
  - the sign convention: with PSI_SIGN = +1 the robust2 update shrinks where
    x < 1; with -1 it changes sign wherever |psi| exceeds N_v.
"""
import numpy as np
import pytest

import oracle
from pct_amd import configs


@pytest.fixture(scope="module")
def block_case(mini_host):
    p = configs.params("mini")
    hr = mini_host["hr"].select(np.arange(mini_host["hr"].n) < 240)
    V = p.columns * p.rows * p.slices
    k = np.arange(V)
    x0 = (mini_host["hull"] * (1.0 + 0.3 * np.sin(0.37 * k))).astype(np.float32)
    return p, hr, x0


def _dense(p, hr, x0, block, lam, variant, eta, psi_sign):
    """one iteration, block by block, as matrices"""
    lists = oracle.mlp_paths(p, hr)
    L = np.array([oracle.chord_length(p, float(hr.xy_entry_angle[i]), float(hr.xy_exit_angle[i]),
                                      float(hr.xz_entry_angle[i]), float(hr.xz_exit_angle[i]))
                  for i in range(hr.n)])
    b = hr.wepl.astype(np.float64)
    x = x0.astype(np.float32).copy()
    V = x.size
    for s in range(0, hr.n, block):
        ids = range(s, min(s + block, hr.n))
        occ = np.zeros((len(ids), V))
        for r, i in enumerate(ids):
            np.add.at(occ[r], lists[i], 1.0)
        xd = x.astype(np.float64)
        n = occ.sum(axis=1)
        Lb = L[list(ids)]
        res = b[list(ids)] - Lb * (occ @ xd)
        aa = Lb * Lb * n
        counts = occ.sum(axis=0)
        hit = counts > 0
        if variant == 1:
            psi = (1.0 - xd) * eta * psi_sign
            upd = ((Lb * lam * res)[:, None] * occ / (aa[:, None] + psi[None, :])).sum(axis=0)
            new = xd + np.where(hit, upd / np.maximum(counts, 1), 0.0)
        else:
            upd = ((Lb * lam * res / aa)[:, None] * occ).sum(axis=0)
            N = ((Lb * Lb)[:, None] * occ).sum(axis=0)
            psi = psi_sign * (1.0 - x.astype(np.float32)).astype(np.float64) * eta
            scale = np.where(hit, N / np.where(hit, N + psi, 1.0), 0.0)
            new = xd + np.where(hit, scale * upd / np.maximum(counts, 1), 0.0)
        x = new.astype(np.float32)
    return x


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("psi_sign", [1, -1])
def test_variants_match_dense_formulation(block_case, variant, psi_sign):
    p, hr, x0 = block_case
    p.eta, p.psi_sign = 2.5, psi_sign
    lam = 0.02
    got = oracle.drop(p, hr, x0, iterations=1, block_size=80, lam=lam, variant=variant)
    exp = _dense(p, hr, x0, 80, np.float64(np.float32(lam)), variant, 2.5, psi_sign)
    moved = got != x0
    assert moved.sum() > 1000
    np.testing.assert_allclose(got, exp, rtol=0, atol=2 * np.spacing(np.float32(2.0)))
    assert np.array_equal(moved, exp != x0)


def test_eta_zero_reduces_to_plain(block_case):
    p, hr, x0 = block_case
    p.eta, p.psi_sign = 0.0, 1
    a = oracle.drop(p, hr, x0, iterations=2, block_size=80, lam=0.02, variant=1)
    b = oracle.drop(p, hr, x0, iterations=2, block_size=80, lam=0.02, variant=2)
    plain = oracle.drop(p, hr, x0, iterations=2, block_size=80, lam=0.02)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.max(np.abs(a - plain)) <= 4 * np.spacing(np.float32(2.0))
    assert np.mean(a == plain) > 0.99


def test_psi_sign_convention(block_case):
    """one block, so that every voxel is updated once from x0: with PSI_SIGN
    = +1, psi = (1 - x) ETA > 0 where x < 1 damps the robust2 update there
    (N / (N + psi) < 1); with -1 the denominator N - |psi| turns negative
    wherever |psi| > N, and the update changes sign"""
    p, hr, x0 = block_case
    plain_p = configs.params("mini")
    plain_p.eta = 0.0
    base = oracle.drop(plain_p, hr, x0, block_size=hr.n, lam=0.02, variant=2)
    p.eta, p.psi_sign = 2.5, 1
    damp = oracle.drop(p, hr, x0, block_size=hr.n, lam=0.02, variant=2)
    p.psi_sign = -1
    flip = oracle.drop(p, hr, x0, block_size=hr.n, lam=0.02, variant=2)
    low = (x0 > 0) & (x0 < 0.9) & (base != x0)
    assert low.sum() > 100
    u0, u1, u2 = (v[low].astype(np.float64) - x0[low] for v in (base, damp, flip))
    assert np.all(np.abs(u1) <= np.abs(u0) + 1e-7) and np.mean(np.abs(u1) < np.abs(u0)) > 0.9
    assert np.mean(np.sign(u2) == -np.sign(u0)) > 0.75
