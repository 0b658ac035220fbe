"""Device Hermitian eigensolvers through dc2_cdiag (diag.F90:66-128) against numpy.linalg.eigh:
this library's tridiagonalisation + tridiagonal divide and conquer ("native") and rocSOLVER's
divide and conquer,
on random, degenerate (Kramers-like pairs), rank-deficient, diagonal and tridiagonal
matrices, for sizes on both sides of every register-tile boundary (64, 128, 256, 512, 768).
Eigenvalues to 1e-12 of the norm, |A V - V W| to 1e-11 of the norm, V^H V = I to 1e-12."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [3, 4, 17, 64, 65, 128, 129, 144, 164, 192, 193, 256, 300, 513, 792, 1024]


def _matrix(kind, n, rng):
    Z = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    if kind == "diagonal":
        return np.diag(rng.normal(size=n)).astype(complex)
    if kind == "tridiagonal":
        A = np.diag(rng.normal(size=n)).astype(complex)
        off = rng.normal(size=n - 1) + 1j * rng.normal(size=n - 1)
        return A + np.diag(off, -1) + np.diag(off.conj(), 1)
    if kind == "random":
        return 0.5 * (Z + Z.conj().T)
    Q, _ = np.linalg.qr(Z)
    if kind == "pairs":
        lam = np.repeat(rng.normal(size=(n + 1) // 2), 2)[:n]
    else:   # rank deficient: a Gram matrix Phi Phi^H of rank n/3
        lam = np.concatenate([rng.uniform(1e-6, 3.0, n // 3), np.zeros(n - n // 3)])
    A = (Q * lam) @ Q.conj().T
    return 0.5 * (A + A.conj().T)


@pytest.fixture(scope="module")
def ctx():
    from dirac_caspt2_amd.api import Caspt2Context
    with Caspt2Context(0) as c:
        yield c


@pytest.mark.parametrize("solver", ["native", "dc"])
@pytest.mark.parametrize("kind", ["random", "pairs", "rank_deficient", "diagonal", "tridiagonal"])
@pytest.mark.parametrize("n", SIZES)
def test_cdiag(ctx, n, kind, solver):
    from dirac_caspt2_amd.api import cdiag
    from gpu_helpers import say
    say("cdiag", solver, kind, n)
    rng = np.random.default_rng(1000 * n + len(kind))
    A = _matrix(kind, n, rng)
    ctx.set_option("eigensolver", solver)
    before = ctx.eig_fallbacks()
    w, V, dimm = cdiag(ctx, A)
    if solver == "native":   # (rocSOLVER's D&C falls back to QR on some diagonal matrices)
        assert ctx.eig_fallbacks() == before, "the native solver fell back to rocSOLVER"
    scale = max(1.0, np.abs(np.linalg.eigvalsh(A)).max())
    np.testing.assert_allclose(w, np.linalg.eigvalsh(A), atol=1e-12 * scale, rtol=0)
    assert np.abs(V.conj().T @ V - np.eye(n)).max() < 1e-12
    assert np.abs(A @ V - V * w).max() < 1e-11 * scale


def test_upper_triangle_only(ctx):
    """Only the upper triangle is read (as LAPACK's UPLO = 'U'): garbage below it is ignored."""
    from dirac_caspt2_amd.api import cdiag
    rng = np.random.default_rng(7)
    n = 200
    A = _matrix("random", n, rng)
    B = A.copy()
    B[np.tril_indices(n, -1)] = 1e3 * (rng.normal(size=n * (n - 1) // 2) + 1j)
    ctx.set_option("eigensolver", "native")
    w, V, _ = cdiag(ctx, B)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(A), atol=1e-12 * np.abs(w).max(), rtol=0)
    assert np.abs(A @ V - V * w).max() < 1e-11 * np.abs(w).max()
    ctx.set_option("eigensolver", "auto")


BIG = [(1025, "random"), (1025, "pairs"), (1274, "rank_deficient"), (1281, "pairs"), (1536, "random"),
       (1792, "rank_deficient"), (2002, "pairs"), (2002, "rank_deficient"), (2048, "random")]


@pytest.mark.parametrize("n,kind", BIG)
def test_cdiag_one_column_per_wave(ctx, n, kind):
    """Sizes above 1024, where every wave of the tridiagonalisation holds one column (up to
    2048 / 8 = 256 workgroups); 1274 and 2002 are the A and C metrics of the 324-spinor
    workload."""
    from dirac_caspt2_amd.api import cdiag
    from gpu_helpers import say
    say("cdiag native", kind, n)
    rng = np.random.default_rng(7 * n + len(kind))
    A = _matrix(kind, n, rng)
    ctx.set_option("eigensolver", "native")
    before = ctx.eig_fallbacks()
    w, V, dimm = cdiag(ctx, A)
    assert ctx.eig_fallbacks() == before, "the native solver fell back to rocSOLVER"
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, np.abs(ref).max())
    np.testing.assert_allclose(w, ref, atol=1e-12 * scale, rtol=0)
    assert np.abs(V.conj().T @ V - np.eye(n)).max() < 1e-12
    assert np.abs(A @ V - V * w).max() < 1e-11 * scale
    ctx.set_option("eigensolver", "auto")


def _tridiag(kind, n, rng):
    """Real symmetric tridiagonals that stress the divide and conquer: close pairs (Wilkinson),
    clusters (glued Wilkinson), zero couplings (split matrices), equal diagonals, a graded
    spectrum and the 1-2-1 Toeplitz matrix (deflation at every level)."""
    if kind == "wilkinson":
        d = np.abs(np.arange(n) - (n - 1) / 2.0)
        e = np.ones(n - 1)
    elif kind == "glued":
        m = 21
        d = np.tile(np.abs(np.arange(m) - 10.0), n // m + 1)[:n]
        e = np.ones(n - 1)
        e[m - 1::m] = 1e-10
    elif kind == "split":
        d = rng.normal(size=n)
        e = rng.normal(size=n - 1)
        e[rng.random(n - 1) < 0.2] = 0.0
    elif kind == "equal":
        d = np.full(n, 0.5)
        e = 1e-9 * rng.normal(size=n - 1)
    elif kind == "graded":
        d = 10.0 ** (-np.linspace(0, 12, n))
        e = 0.3 * np.sqrt(d[:-1] * d[1:])
    elif kind == "toeplitz":
        d = np.full(n, 2.0)
        e = -np.ones(n - 1)
    else:
        d = rng.normal(size=n)
        e = rng.normal(size=n - 1)
    ph = np.exp(1j * rng.uniform(0, 2 * np.pi, n - 1))   # a complex Hermitian tridiagonal
    return np.diag(d).astype(complex) + np.diag(e * ph, -1) + np.diag(e * ph.conj(), 1)


TRI = [(k, n) for k in ("wilkinson", "glued", "split", "equal", "graded", "toeplitz")
       for n in (3, 31, 33, 97, 500, 1001)] + [("random", n) for n in (31, 32, 33, 63, 64, 65, 127, 2048)]


@pytest.mark.parametrize("tri", ["own", "rocsolver"])
@pytest.mark.parametrize("kind,n", TRI)
def test_tridiagonal_divide_and_conquer(ctx, kind, n, tri):
    """The tridiagonal stage on its own (the tridiagonalisation of a tridiagonal matrix is the
    identity up to phases): this library's divide and conquer and rocSOLVER dstedc, both to the
    bounds above; this library's must not fall back."""
    from dirac_caspt2_amd.api import cdiag
    from gpu_helpers import say
    say("tridiagonal", tri, kind, n)
    rng = np.random.default_rng(31 * n + len(kind))
    A = _tridiag(kind, n, rng)
    ctx.set_option("eigensolver", "native")
    ctx.set_option("tridiag_solver", tri)
    before = ctx.eig_fallbacks()
    try:
        w, V, _ = cdiag(ctx, A)
    finally:
        ctx.set_option("tridiag_solver", "own")
        ctx.set_option("eigensolver", "auto")
    assert ctx.eig_fallbacks() == before, "the native solver fell back to rocSOLVER"
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, np.abs(ref).max())
    np.testing.assert_allclose(w, ref, atol=1e-12 * scale, rtol=0)
    assert np.abs(V.conj().T @ V - np.eye(n)).max() < 1e-12
    assert np.abs(A @ V - V * w).max() < 1e-11 * scale
