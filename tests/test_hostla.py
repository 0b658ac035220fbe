"""Host Hermitian eigensolvers of the library (Householder + QL, cyclic Jacobi, and the
Davidson's selected-vector form) against numpy.linalg.eigh: eigenvalues to 1e-12 relative to the norm,
eigenvectors orthonormal and A V = V diag(w) to the same level, on random, degenerate,
rank-deficient, already-tridiagonal, diagonal and tiny matrices."""
import numpy as np
import pytest


@pytest.mark.parametrize("method", ["auto", "jacobi", "ql", "low"])
@pytest.mark.parametrize("kind", ["random", "degenerate", "rank_deficient", "tiny", "one", "tridiagonal", "diagonal",
                                  "davidson48", "gram96"])
def test_host_heev(kind, method):
    from dirac_caspt2_amd.api import host_heev
    rng = np.random.default_rng(len(kind))
    n = {"tiny": 2, "one": 1, "davidson48": 48, "gram96": 96}.get(kind, 37)
    Z = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    Q, _ = np.linalg.qr(Z)
    if kind == "degenerate":
        lam = np.repeat(rng.normal(size=(n + 1) // 2), 2)[:n]
    elif kind in ("rank_deficient", "gram96"):   # gram96: a metric with two thirds null space
        lam = np.concatenate([rng.uniform(0.1, 1.0, n // 3), np.zeros(n - n // 3)])
    elif kind == "davidson48":   # Rayleigh quotient of a CI matrix: clustered low end near -40
        lam = np.concatenate([-40.0 + np.repeat(rng.uniform(0, 0.01, 3), 2), -39.0 + rng.uniform(0, 5, n - 6)])
    else:
        lam = rng.normal(size=n) * 10
    A = (Q * lam) @ Q.conj().T
    if kind == "tridiagonal":    # complex Hermitian tridiagonal: no reflector is needed
        A = np.diag(rng.normal(size=n)).astype(complex)
        off = rng.normal(size=n - 1) + 1j * rng.normal(size=n - 1)
        A += np.diag(off, -1) + np.diag(off.conj(), 1)
        lam = np.linalg.eigvalsh(A)
    elif kind == "diagonal":
        A = np.diag(rng.normal(size=n)).astype(complex)
        lam = np.diag(A).real
    A = 0.5 * (A + A.conj().T)
    w, V = host_heev(A, method)
    scale = max(1.0, np.abs(lam).max())
    np.testing.assert_allclose(w, np.linalg.eigvalsh(A), atol=1e-12 * scale, rtol=0)
    if method == "low":   # the lowest min(n, 4) eigenvectors, zero columns after them
        k = min(n, 4)
        assert not V[:, k:].any()
        V, w = V[:, :k], w[:k]
    k = V.shape[1]
    np.testing.assert_allclose(V.conj().T @ V, np.eye(k), atol=1e-12)
    np.testing.assert_allclose(A @ V, V * w, atol=1e-11 * scale)
