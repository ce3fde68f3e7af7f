"""Dense complex linear algebra on the GPU vs numpy (LAPACK)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("m,n,k", [(64, 64, 16), (100, 70, 37), (257, 130, 64), (5, 300, 3)])
def test_zgemm_matches_numpy(m, n, k):
    import bem3_emf_b2_amd as bem
    rng = np.random.default_rng(m * 1000 + n + k)
    A, B, C0 = _rand(rng, m, k), _rand(rng, k, n), _rand(rng, m, n)
    got = bem.zgemm(A, B, C0, alpha=-1.0 + 0.5j, beta=0.3 - 1.0j)
    want = (-1.0 + 0.5j) * (A @ B) + (0.3 - 1.0j) * C0
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("n,nrhs", [(1, 1), (63, 2), (300, 6), (1000, 3)])
def test_zgesv_matches_numpy(n, nrhs):
    import bem3_emf_b2_amd as bem
    rng = np.random.default_rng(n + 7 * nrhs)
    A = _rand(rng, n, n) + n * 0.1 * np.eye(n)
    B = _rand(rng, n, nrhs)
    X = bem.zgesv(A, B)
    want = np.linalg.solve(A, B)
    res = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert res < 1e-14
    assert np.abs(X - want).max() < 1e-10 * np.abs(want).max()


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
