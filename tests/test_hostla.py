"""Host Hermitian Jacobi eigensolver of the library (the Davidson subspace solver) against
numpy.linalg.eigh: eigenvalues to 1e-12 relative to the norm, eigenvectors orthonormal and
A V = V diag(w) to the same level, on random, degenerate and rank-deficient matrices."""
import numpy as np
import pytest


@pytest.mark.parametrize("kind", ["random", "degenerate", "rank_deficient", "tiny"])
def test_host_heev(kind):
    from dirac_caspt2_amd.api import host_heev
    rng = np.random.default_rng(len(kind))
    n = {"tiny": 2}.get(kind, 37)
    Z = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    Q, _ = np.linalg.qr(Z)
    if kind == "degenerate":
        lam = np.repeat(rng.normal(size=(n + 1) // 2), 2)[:n]
    elif kind == "rank_deficient":
        lam = np.concatenate([rng.uniform(0.1, 1.0, n // 3), np.zeros(n - n // 3)])
    else:
        lam = rng.normal(size=n) * 10
    A = (Q * lam) @ Q.conj().T
    A = 0.5 * (A + A.conj().T)
    w, V = host_heev(A)
    scale = max(1.0, np.abs(lam).max())
    np.testing.assert_allclose(w, np.linalg.eigvalsh(A), atol=1e-12 * scale, rtol=0)
    np.testing.assert_allclose(V.conj().T @ V, np.eye(n), atol=1e-12)
    np.testing.assert_allclose(A @ V, V * w, atol=1e-11 * scale)
