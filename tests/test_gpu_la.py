"""Dense complex linear algebra on the GPU vs numpy (LAPACK)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# (3001, 2999, 77) and (4096, 4096, 64) take the 128x128-tile kernel (at least
# 256 tiles), the n <= 8 shapes with k >= 256 the skinny kernel
@pytest.mark.parametrize("m,n,k", [(64, 64, 16), (100, 70, 37), (257, 130, 64), (5, 300, 3), (3001, 2999, 77),
                                   (4096, 4096, 64), (3001, 1, 500), (1000, 6, 2048), (777, 3, 300), (17, 8, 999),
                                   (4000, 2, 256)])
def test_zgemm_matches_numpy(m, n, k):
    import bem3_emf_b2_amd as bem
    rng = np.random.default_rng(m * 1000 + n + k)
    A, B, C0 = _rand(rng, m, k), _rand(rng, k, n), _rand(rng, m, n)
    got = bem.zgemm(A, B, C0, alpha=-1.0 + 0.5j, beta=0.3 - 1.0j)
    want = (-1.0 + 0.5j) * (A @ B) + (0.3 - 1.0j) * C0
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("n,nrhs", [(1, 1), (63, 2), (300, 6), (1000, 3), (3100, 2)])
def test_zgesv_matches_numpy(n, nrhs):
    import bem3_emf_b2_amd as bem
    rng = np.random.default_rng(n + 7 * nrhs)
    A = _rand(rng, n, n) + n * 0.1 * np.eye(n)
    B = _rand(rng, n, nrhs)
    X, LU = bem.zgesv(A, B, return_lu=True)
    want = np.linalg.solve(A, B)
    res = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert res < 1e-14
    assert np.abs(X - want).max() < 1e-10 * np.abs(want).max()
    # same pivot sequence as LAPACK's zgetrf: the packed factors agree
    import scipy.linalg as sl
    lu_ref, _ = sl.lu_factor(A)
    assert np.abs(LU - lu_ref).max() < 1e-9 * np.abs(lu_ref).max()


# nrhs <= 4 takes one one-launch triangular solve per factor (64-row blocks;
# 65, 200 and 4100 end on a partial block), 5 <= nrhs <= 8 two (4 columns,
# then the rest), nrhs = 9 the recursive solves
@pytest.mark.parametrize("n", [64, 65, 200, 1000, 4100])
@pytest.mark.parametrize("nrhs", [1, 3, 4, 5, 8, 9])
def test_few_rhs_solve(n, nrhs):
    import bem3_emf_b2_amd as bem
    rng = np.random.default_rng(3 * n + nrhs)
    A = _rand(rng, n, n) + n * 0.05 * np.eye(n)
    B = _rand(rng, n, nrhs)
    X = bem.zgesv(A, B)
    want = np.linalg.solve(A, B)
    assert np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X)) < 1e-14
    assert np.abs(X - want).max() < 1e-10 * np.abs(want).max()


@pytest.mark.parametrize("n", [65, 1000, 3000])
@pytest.mark.parametrize("nrhs", [1, 2, 4])
def test_few_rhs_solve_fp32(n, nrhs):
    import bem3_emf_b2_amd as bem
    rng = np.random.default_rng(5 * n + nrhs)
    A = _rand(rng, n, n) + n * 0.05 * np.eye(n)
    B = _rand(rng, n, nrhs)
    X = bem.cgesv(A, B).astype(np.complex128)
    assert np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X)) < 1e-5


@pytest.mark.parametrize("n", [130, 700])
def test_lu_pivot_sequence_random(n):
    """A random matrix needs row interchanges in almost every column; the
    packed factors then match LAPACK's only if every pivot matches."""
    import scipy.linalg as sl

    import bem3_emf_b2_amd as bem
    rng = np.random.default_rng(100 + n)
    A = _rand(rng, n, n)
    B = _rand(rng, n, 1)
    X, LU = bem.zgesv(A, B, return_lu=True)
    lu_ref, piv_ref = sl.lu_factor(A)
    assert (piv_ref != np.arange(n)).sum() > n // 2
    assert np.abs(LU - lu_ref).max() < 1e-10 * np.abs(lu_ref).max()
    assert np.abs(A @ X - B).max() < 1e-10 * np.abs(B).max() * np.linalg.cond(A)


def test_zgesv_needs_pivoting():
    import bem3_emf_b2_amd as bem
    A = np.array([[0.0, 1.0, 2.0], [1.0, 0.0, 3.0], [4.0, -3.0, 8.0]], dtype=complex)
    B = np.array([[1.0], [2.0], [3.0]], dtype=complex)
    X = bem.zgesv(A, B)
    assert np.allclose(A @ X, B, rtol=0, atol=1e-14)


def test_zgesv_reports_singular():
    import bem3_emf_b2_amd as bem
    A = np.zeros((4, 4), complex)
    with pytest.raises(bem.BemError):
        bem.zgesv(A, np.ones((4, 1), complex))
