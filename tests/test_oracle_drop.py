"""Pins the oracle's DROP (DROP_blocks .cu:8023-8043, DROP_update
.cu:8044-8055, per-history expression of block_update_GPU .cu:5822-5834):

  - one block on a 3-ray / 4-voxel system equals the dense-matrix formula
    x += D^-1 A^T diag(lambda (b - A x) / ||a_i||^2) with D = diag(counts);
  - block size 1 is Kaczmarz/ART with relaxation lambda and converges to the
    solution of a consistent system;
  - one step is linear in lambda;
  - the DROP of whole histories equals the DROP of their MLP voxel lists.
"""
import numpy as np
import pytest

import oracle


def _dense_drop(V, lists, L, b, x, block, lam, iters=1):
    x = x.astype(np.float64).copy()
    A = np.zeros((len(lists), V))
    for i, a in enumerate(lists):
        A[i, a] = L[i]
    for _ in range(iters):
        for s in range(0, len(lists), block):
            rows = range(s, min(s + block, len(lists)))
            upd = np.zeros(V)
            cnt = np.zeros(V)
            for i in rows:
                a = A[i]
                delta = lam * (b[i] - a @ x) / (a @ a) * L[i]
                upd[a != 0] += delta
                cnt[a != 0] += 1
            nz = cnt > 0
            x[nz] += upd[nz] / cnt[nz]
    return x


def test_toy_block_equals_dense_formula():
    lists = [[0, 1], [1, 2, 3], [0, 3]]
    L = np.array([0.2, 0.25, 0.3])
    b = np.array([1.0, 2.0, 0.5], np.float32)
    x0 = np.array([0.5, 1.0, 0.0, 2.0], np.float32)
    for block, lam in [(3, 0.5), (2, 1.0), (1, 0.3)]:
        got = oracle.drop_lists(4, lists, L, b, x0, block, lam, iterations=2)
        exp = _dense_drop(4, lists, L, b.astype(np.float64), x0, block, lam, iters=2)
        np.testing.assert_allclose(got, exp, rtol=1e-6, atol=1e-6)


def test_block_size_one_is_kaczmarz_and_converges():
    rng = np.random.default_rng(0)
    V = 6
    lists = [sorted(rng.choice(V, size=rng.integers(2, 5), replace=False)) for _ in range(12)]
    L = np.full(len(lists), 0.1)
    xs = rng.uniform(0.5, 1.5, V)
    A = np.zeros((len(lists), V))
    for i, a in enumerate(lists):
        A[i, a] = L[i]
    assert np.linalg.matrix_rank(A) == V
    b = (A @ xs).astype(np.float32)
    x = oracle.drop_lists(V, lists, L, b, np.ones(V, np.float32), 1, 1.0, iterations=400)
    np.testing.assert_allclose(x, xs, atol=2e-4)


def test_one_step_linear_in_lambda():
    lists = [[0, 1, 2], [2, 3], [1, 3, 4]]
    L = np.array([0.1, 0.1, 0.1])
    b = np.array([0.7, 0.2, 0.9], np.float32)
    x0 = np.array([1.0, 1.0, 1.0, 1.0, 1.0], np.float32)
    d1 = oracle.drop_lists(5, lists, L, b, x0, 3, 1e-2).astype(np.float64) - x0
    d2 = oracle.drop_lists(5, lists, L, b, x0, 3, 2e-2).astype(np.float64) - x0
    np.testing.assert_allclose(d2, 2 * d1, rtol=1e-4, atol=1e-7)


def test_history_drop_equals_list_drop(mini_host):
    p = mini_host["p"]
    hr = mini_host["hr"].head(5000)
    lists = oracle.mlp_paths(p, hr)
    L = np.array([oracle.chord_length(p, hr.xy_entry_angle[i], hr.xy_exit_angle[i], hr.xz_entry_angle[i],
                                      hr.xz_exit_angle[i]) for i in range(hr.n)])
    x0 = mini_host["hull"].astype(np.float32)
    a = oracle.drop(p, hr, x0, iterations=1, lam=1e-2)
    b = oracle.drop_lists(x0.size, lists, L, hr.wepl, x0, p.drop_block_size, 1e-2)
    np.testing.assert_array_equal(a, b)
    assert np.any(a != x0)


def test_partial_last_block_is_applied(mini_host):
    """every block, including the last partial one, updates the image"""
    p = mini_host["p"]
    hr = mini_host["hr"].head(p.drop_block_size + 17)
    x0 = mini_host["hull"].astype(np.float32)
    full = oracle.drop(p, hr, x0, iterations=1, lam=1e-2)
    first = oracle.drop(p, hr.head(p.drop_block_size), x0, iterations=1, lam=1e-2)
    assert np.any(full != first)


def test_threaded_drop_equals_single_thread(mini_host):
    p = mini_host["p"]
    hr = mini_host["hr"]
    x0 = mini_host["hull"].astype(np.float32)
    a = oracle.drop(p, hr, x0, iterations=2, lam=1e-2, threads=1)
    b = oracle.drop(p, hr, x0, iterations=2, lam=1e-2, threads=8)
    np.testing.assert_array_equal(a, b)
