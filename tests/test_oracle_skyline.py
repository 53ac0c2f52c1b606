"""Pins the oracle's skyline LDL^T (advlas_skyline_ldl_decomposite / _solve,
lib/hddm/advlas_skyline.c:79-167) against the textbook pivot-free complex-
symmetric factorisation: A = L D L^T reconstructed densely, and solves
against numpy's dense solver, on random variable-band profiles."""
import numpy as np
import pytest

from oracle import fw_oracle as fo


def random_profile(rng, n, maxw):
    idx1 = np.zeros(n, np.int32)
    for i in range(1, n):
        idx1[i] = max(0, i - rng.integers(0, maxw + 1))
    idx2 = np.zeros(n, np.int32)
    for i in range(1, n):
        idx2[i] = idx2[i - 1] + i - idx1[i]
    nlen = int(idx2[-1] + n)
    return idx1, idx2, nlen


def dense_from_sky(idx1, idx2, mat):
    n = len(idx1)
    A = np.zeros((n, n), complex)
    for i in range(n):
        for j in range(idx1[i], i + 1):
            A[i, j] = A[j, i] = mat[idx2[i] + j]
    return A


@pytest.mark.parametrize("n,maxw,seed", [(1, 0, 0), (2, 1, 1), (17, 5, 2), (60, 20, 3), (200, 40, 4)])
def test_factor_reconstructs(n, maxw, seed):
    rng = np.random.default_rng(seed)
    idx1, idx2, nlen = random_profile(rng, n, maxw)
    mat = rng.standard_normal(nlen) + 1j * rng.standard_normal(nlen)
    A = dense_from_sky(idx1, idx2, mat)
    A += np.diag(2.0 * np.abs(A).sum(axis=1) * np.exp(1j * rng.uniform(0, np.pi, n)))  # complex symmetric, dominant
    mat = np.array([A[i, j] for i in range(n) for j in range(idx1[i], i + 1)])
    fac = fo.sky_factor(idx1, idx2, mat)
    L = np.eye(n, dtype=complex)
    Dinv = np.zeros(n, complex)
    for i in range(n):
        for j in range(idx1[i], i):
            L[i, j] = fac[idx2[i] + j]
        Dinv[i] = fac[idx2[i] + i]
    R = L @ np.diag(1.0 / Dinv) @ L.T
    assert np.abs(R - A).max() < 1e-13 * np.abs(A).max()
    # fill stays inside the profile: L is zero left of idx1
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = fo.sky_solve(idx1, idx2, fac, b)
    xr = np.linalg.solve(A, b)
    assert np.abs(x - xr).max() < 1e-12 * np.abs(xr).max()


def test_indefinite_pivot_free():
    """Complex-symmetric indefinite matrix with a negative real pivot."""
    A = np.array([[1.0, 2.0, 0], [2.0, 1.0, 1j], [0, 1j, 3.0]], complex)
    idx1 = np.array([0, 0, 1], np.int32)
    idx2 = np.array([0, 1, 2], np.int32)
    mat = np.array([A[0, 0], A[1, 0], A[1, 1], A[2, 1], A[2, 2]])
    fac = fo.sky_factor(idx1, idx2, mat)
    assert np.isclose(1.0 / fac[2], -3.0)   # d_1 = 1 - 4
    b = np.array([1.0, 2.0, 3.0], complex)
    assert np.allclose(fo.sky_solve(idx1, idx2, fac, b), np.linalg.solve(A, b), rtol=1e-14)


def test_mutation_conjugated_factor_is_caught():
    """Using a Hermitian (conjugating) factor would not reproduce A."""
    rng = np.random.default_rng(5)
    idx1, idx2, nlen = random_profile(rng, 30, 8)
    mat = rng.standard_normal(nlen) + 1j * rng.standard_normal(nlen)
    A = dense_from_sky(idx1, idx2, mat) + np.diag(np.full(30, 40.0 + 10j))
    mat = np.array([A[i, j] for i in range(30) for j in range(idx1[i], i + 1)])
    fac = fo.sky_factor(idx1, idx2, mat)
    L = np.eye(30, dtype=complex)
    Dinv = np.array([fac[idx2[i] + i] for i in range(30)])
    for i in range(30):
        for j in range(idx1[i], i):
            L[i, j] = fac[idx2[i] + j]
    assert np.abs(L @ np.diag(1 / Dinv) @ L.conj().T - A).max() > 1e-3
