"""dc2_zgemm_split_noalloc (the split-precision product with its scratch in a fixed device
array, allocating nothing) against dc2_zgemm_split (the same kernels with the context's
scratch): bit-identical outputs on two calls in a row, the second smaller, and a product
larger than the fixed array refused with status 1."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _operands(rng, M, N, K):
    At = (rng.normal(size=(M, K)) + 1j * rng.normal(size=(M, K))) * 10.0 ** rng.uniform(-2, 2, (M, 1))
    Bm = (rng.normal(size=(K, N)) + 1j * rng.normal(size=(K, N)))
    return np.ascontiguousarray(At), np.asfortranarray(Bm).ravel(order="F")


def test_noalloc_equals_allocating_variant_twice():
    from dirac_caspt2_amd.api import Caspt2Context, DeviceArray, zgemm_split
    from dirac_caspt2_amd._lib import check
    from gpu_helpers import say
    rng = np.random.default_rng(7)
    with Caspt2Context(0) as ctx:
        for (M, N, K, slices, opB) in ((3000, 164, 164, 6, "R"), (257, 33, 40, 7, "N")):
            say("noalloc", M, N, K, slices, opB)
            At, Bf = _operands(rng, M, N, K)
            dA, dB = DeviceArray(ctx, At), DeviceArray(ctx, Bf)
            dC0 = DeviceArray(ctx, np.full(M * N, np.nan + 0j))
            dC1 = DeviceArray(ctx, np.full(M * N, np.nan + 0j))
            zgemm_split(ctx, slices, opB, M, N, K, dA, K, dB, K, dC0, M)
            check("dc2_zgemm_split_noalloc",
                  ctx.lib.dc2_zgemm_split_noalloc(ctx.h, slices, opB.encode(), M, N, K, dA.ptr, K, dB.ptr, K,
                                                  dC1.ptr, M))
            check("dc2_synchronize", ctx.lib.dc2_synchronize(ctx.h))
            ref, got = dC0.to_host(), dC1.to_host()
            assert np.all(np.isfinite(got))
            assert np.array_equal(ref, got)
            exact = (At @ (np.conj(Bf.reshape((K, N), order="F")) if opB == "R" else Bf.reshape((K, N), order="F")))
            assert np.max(np.abs(got.reshape((M, N), order="F") - exact)) < 1e-9 * np.max(np.abs(exact))


def test_noalloc_refuses_products_beyond_the_fixed_scratch():
    from dirac_caspt2_amd.api import Caspt2Context, DeviceArray
    with Caspt2Context(0) as ctx:
        d = DeviceArray(ctx, np.zeros(16, dtype=np.complex128))
        st = ctx.lib.dc2_zgemm_split_noalloc(ctx.h, 6, b"N", 100000, 100, 1024, d.ptr, 1024, d.ptr, 1024, d.ptr,
                                             100000)
        assert st == 1
        assert b"scratch" in ctx.lib.dc2_last_error()
