"""MFMA FP64 complex GEMM of the library against a plain numpy product (all op combinations)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from dirac_caspt2_amd.api import Caspt2Context
    c = Caspt2Context(0)
    yield c
    c.close()


def _op(x, op):
    if op == "N":
        return x
    if op == "T":
        return x.T
    if op == "C":
        return np.conj(x).T
    if op == "R":
        return np.conj(x)


@pytest.mark.parametrize("opA", ["N", "T", "C"])
@pytest.mark.parametrize("opB", ["N", "R", "T", "C"])
@pytest.mark.parametrize("shape", [(1, 1, 1), (17, 33, 5), (64, 64, 16), (130, 70, 97), (12, 200, 280),
                                   (40, 24, 3001), (8, 3, 2500), (48, 48, 924),   # the last three split K
                                   # the block-Davidson shapes: small outputs over a long K, a short K,
                                   # at most 8 columns over a long K, and their edges
                                   (48, 6, 924), (6, 6, 924), (924, 1, 924), (45, 7, 63), (2048, 1, 3),
                                   (924, 6, 48), (924, 6, 64), (255, 8, 65), (924, 6, 924), (3000, 5, 4097),
                                   (33, 9, 700),
                                   # matrix-vector products over a long K (one wave per row)
                                   (3003, 1, 3003), (2600, 4, 256),
                                   # few rows over a long K (a workgroup per row)
                                   (48, 4, 3003), (1, 1, 1024), (7, 3, 1500),
                                   # K <= 16 with one output dimension <= 16 (index rotations
                                   # of the transforms: the MFMA kernel)
                                   (4100, 16, 16), (4100, 5, 12), (12, 4100, 12), (7, 3000, 16), (16, 300, 9)])
def test_zgemm_ops(ctx, opA, opB, shape):
    from dirac_caspt2_amd.api import DeviceArray, zgemm
    from gpu_helpers import say
    M, N, K = shape
    say(f"zgemm {opA}{opB} {shape}")
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    a_shape = (M, K) if opA == "N" else (K, M)
    b_shape = (K, N) if opB in "NR" else (N, K)
    A = rng.normal(size=a_shape) + 1j * rng.normal(size=a_shape)
    B = rng.normal(size=b_shape) + 1j * rng.normal(size=b_shape)
    C = rng.normal(size=(M, N)) + 1j * rng.normal(size=(M, N))
    alpha, beta = 0.7 - 0.3j, -0.2 + 0.5j
    ref = alpha * _op(A, opA) @ _op(B, opB) + beta * C
    dev = lambda x: DeviceArray(ctx, np.asfortranarray(x).ravel(order="F"))
    dA, dB, dC = dev(A), dev(B), dev(C)
    zgemm(ctx, opA, opB, M, N, K, alpha, dA, a_shape[0], 0, dB, b_shape[0], 0, beta, dC, M, 0, 1)
    ctx.synchronize()
    out = dC.to_host().reshape((M, N), order="F")
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12 * max(1, K))


def test_zgemm_batched(ctx):
    from dirac_caspt2_amd.api import DeviceArray, zgemm
    rng = np.random.default_rng(3)
    M, N, K, nb = 40, 24, 30, 5
    A = rng.normal(size=(nb, K, M)) + 1j * rng.normal(size=(nb, K, M))
    B = rng.normal(size=(nb, N, K)) + 1j * rng.normal(size=(nb, N, K))
    dA = DeviceArray(ctx, A)        # batch-major; each (M,K) column-major == (K,M) C-order
    dB = DeviceArray(ctx, B)
    dC = DeviceArray(ctx, np.zeros(nb * M * N, dtype=np.complex128))
    zgemm(ctx, "N", "N", M, N, K, 1.0 + 0j, dA, M, M * K, dB, K, K * N, 0j, dC, M, M * N, nb)
    ctx.synchronize()
    out = dC.to_host().reshape(nb, N, M)
    for b in range(nb):
        np.testing.assert_allclose(out[b].T, A[b].T @ B[b].T, atol=1e-12)


@pytest.mark.parametrize("opA,opB,shape,batch", [("C", "N", (48, 48, 924), 1), ("T", "N", (130, 70, 97), 1),
                                                  ("T", "R", (300, 164, 164), 1), ("N", "C", (200, 200, 600), 1),
                                                  ("N", "N", (40, 24, 30), 5)])
def test_autotuned_choice_still_exact(ctx, opA, opB, shape, batch):
    """The tuner installs the fastest variant / split-K for the shape (or keeps the built-in
    choice), and products of that shape still match numpy to the same bound."""
    from dirac_caspt2_amd.api import DeviceArray, gemm_clear_choices, zgemm
    M, N, K = shape
    r = ctx.gemm_tune(opA, opB, M, N, K, batch, 2)
    assert r["seconds"] > 0 and r["seconds"] <= r["builtin_seconds"] * 1.0000001
    assert (r["variant"], r["split"]) == (0, 0) or (1 <= r["variant"] <= 4 and r["split"] in (1, 2, 4, 8, 16, 32))
    rng = np.random.default_rng(M + N + K)
    a_shape = (M, K) if opA == "N" else (K, M)
    b_shape = (K, N) if opB in "NR" else (N, K)
    A = rng.normal(size=(batch,) + a_shape) + 1j * rng.normal(size=(batch,) + a_shape)
    B = rng.normal(size=(batch,) + b_shape) + 1j * rng.normal(size=(batch,) + b_shape)
    dev = lambda x: DeviceArray(ctx, np.concatenate([np.asfortranarray(x[b]).ravel(order="F") for b in range(batch)]))
    dA, dB = dev(A), dev(B)
    dC = DeviceArray(ctx, np.zeros(batch * M * N, dtype=np.complex128))
    zgemm(ctx, opA, opB, M, N, K, 1.0 + 0j, dA, a_shape[0], a_shape[0] * a_shape[1], dB, b_shape[0],
          b_shape[0] * b_shape[1], 0j, dC, M, M * N, batch)
    ctx.synchronize()
    out = dC.to_host().reshape(batch, N, M)
    for b in range(batch):
        np.testing.assert_allclose(out[b].T, _op(A[b], opA) @ _op(B[b], opB), rtol=0, atol=1e-12 * max(1, K))
    assert gemm_clear_choices() >= 0



@pytest.mark.parametrize("m,k", [(1, 1), (5, 3), (64, 64), (65, 300), (130, 97), (792, 1728), (200, 4096),
                                 (300, 0)])
@pytest.mark.parametrize("split", [0, 1, 3, 16])
def test_zherk_whole_hermitian(ctx, m, k, split):
    """A A^H from the upper tiles (split over k or not) and mirrored: the whole matrix
    matches numpy, the mirrored triangle is the exact conjugate transpose."""
    from dirac_caspt2_amd.api import DeviceArray, zherk
    from gpu_helpers import say
    say(f"zherk m={m} k={k} split={split}")
    rng = np.random.default_rng(m * 11 + k + split)
    A = rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))
    ldc = m + 3
    dA = DeviceArray(ctx, np.asfortranarray(A).ravel(order="F") if k else np.zeros(1, dtype=np.complex128))
    dC = DeviceArray(ctx, np.full(ldc * m, 7.0 + 7.0j))
    zherk(ctx, m, k, dA, max(m, 1), dC, ldc, split)
    ctx.synchronize()
    out = dC.to_host().reshape((m, ldc)).T
    np.testing.assert_allclose(out[:m], A @ A.conj().T, rtol=0, atol=1e-12 * max(1, k))
    iu = np.triu_indices(m, 1)
    assert np.array_equal(out[:m][(iu[1], iu[0])], np.conj(out[:m][iu]))
    assert np.all(out[m:] == 7.0 + 7.0j)   # rows past m untouched


@pytest.mark.parametrize("opA,opB,M,N,K,nb", [("C", "N", 64, 64, 792, 13), ("N", "N", 40, 24, 700, 5),
                                               ("T", "R", 17, 9, 1000, 3), ("N", "C", 130, 70, 520, 2)])
def test_zgemm_batched_split_k(ctx, opA, opB, M, N, K, nb):
    """Batched products with few output tiles over a long K are split over K (partials of
    every product summed in a fixed order into its own output): padded strides, alpha and
    beta, each product against numpy."""
    from dirac_caspt2_amd.api import DeviceArray, zgemm
    from gpu_helpers import say
    say(f"batched split-K {opA}{opB} {M}x{N}x{K} batch {nb}")
    rng = np.random.default_rng(M + N + K + nb)
    a_shape = (M, K) if opA == "N" else (K, M)
    b_shape = (K, N) if opB in "NR" else (N, K)
    sA, sB, sC = a_shape[0] * a_shape[1] + 5, b_shape[0] * b_shape[1] + 3, M * N + 7
    A = [rng.normal(size=a_shape) + 1j * rng.normal(size=a_shape) for _ in range(nb)]
    B = [rng.normal(size=b_shape) + 1j * rng.normal(size=b_shape) for _ in range(nb)]
    C = [rng.normal(size=(M, N)) + 1j * rng.normal(size=(M, N)) for _ in range(nb)]

    def pack(xs, stride):
        buf = np.full(stride * len(xs), 3.0 - 1.0j)
        for b, x in enumerate(xs):
            buf[b * stride:b * stride + x.size] = np.asfortranarray(x).ravel(order="F")
        return buf
    dA, dB, dC = DeviceArray(ctx, pack(A, sA)), DeviceArray(ctx, pack(B, sB)), DeviceArray(ctx, pack(C, sC))
    alpha, beta = 0.4 + 0.9j, -0.6 + 0.1j
    zgemm(ctx, opA, opB, M, N, K, alpha, dA, a_shape[0], sA, dB, b_shape[0], sB, beta, dC, M, sC, nb)
    ctx.synchronize()
    out = dC.to_host()
    for b in range(nb):
        ref = alpha * _op(A[b], opA) @ _op(B[b], opB) + beta * C[b]
        got = out[b * sC:b * sC + M * N].reshape((M, N), order="F")
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * K)
        assert np.all(out[b * sC + M * N:(b + 1) * sC] == 3.0 - 1.0j)   # padding untouched


@pytest.mark.parametrize("opA,opB", [("T", "N"), ("C", "R"), ("C", "N")])
@pytest.mark.parametrize("M,N,K", [(5000, 16, 16), (12, 5000, 12), (300, 11, 7)])
def test_zgemm_short_k_padded(ctx, opA, opB, M, N, K):
    """Short-K index rotations with leading dimensions larger than the operands: rows of C
    past M stay untouched."""
    from dirac_caspt2_amd.api import DeviceArray, zgemm
    rng = np.random.default_rng(M + 3 * N + K)
    lda, ldb, ldc = K + 3, K + 1, M + 2
    A = rng.normal(size=(lda, M)) + 1j * rng.normal(size=(lda, M))
    B = rng.normal(size=(ldb, N)) + 1j * rng.normal(size=(ldb, N))
    C = rng.normal(size=(ldc, N)) + 1j * rng.normal(size=(ldc, N))
    alpha, beta = 1.1 - 0.2j, 0.3 + 0.3j
    dev = lambda x: DeviceArray(ctx, np.asfortranarray(x).ravel(order="F"))
    dA, dB, dC = dev(A), dev(B), dev(C)
    zgemm(ctx, opA, opB, M, N, K, alpha, dA, lda, 0, dB, ldb, 0, beta, dC, ldc, 0, 1)
    ctx.synchronize()
    out = dC.to_host().reshape((ldc, N), order="F")
    ref = alpha * _op(A[:K], opA) @ _op(B[:K], opB) + beta * C[:M]
    np.testing.assert_allclose(out[:M], ref, rtol=0, atol=1e-12 * K)
    assert np.array_equal(out[M:], C[M:])
