"""Split-precision (Ozaki INT8) integral transforms.

* dc2_zgemm_split against a numpy complex product: every element within the scheme's
  truncation bound 8*K*(S+3)*2^(-7S)*max|row of A|*max|column of B| plus FP64 rounding;
* energies with split-precision transforms within 1e-8 Eh of the FP64 GPU path on every
  subspace (52- and 192-spinor workloads), and on the 324-spinor workload within 1e-8 Eh
  of the oracle's energies (tests/golden/synthetic_split.json), the FP64 GPU path within
  1e-9 Eh of them.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "synthetic_split.json")


@pytest.fixture(scope="module")
def ctx():
    from dirac_caspt2_amd.api import Caspt2Context
    c = Caspt2Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("slices", [4, 6, 7])
@pytest.mark.parametrize("opB", ["N", "R"])
@pytest.mark.parametrize("shape", [(1000, 40, 40), (3000, 164, 164), (257, 33, 280), (64, 16, 12)])
def test_zgemm_split_accuracy(ctx, slices, opB, shape):
    from dirac_caspt2_amd.api import DeviceArray, zgemm_split
    from gpu_helpers import say
    M, N, K = shape
    say("zgemm_split", slices, opB, shape)
    rng = np.random.default_rng(M + 3 * N + 7 * K + slices)
    At = (rng.normal(size=(M, K)) + 1j * rng.normal(size=(M, K))) * 10.0 ** rng.uniform(-3, 3, (M, 1))
    Bm = (rng.normal(size=(K, N)) + 1j * rng.normal(size=(K, N))) * 10.0 ** rng.uniform(-3, 3, (1, N))
    At[5 % M, :] = 0.0                                   # a zero row
    ref = At @ (np.conj(Bm) if opB == "R" else Bm)
    dA = DeviceArray(ctx, np.ascontiguousarray(At))       # row m contiguous: A[k + K*m]
    dB = DeviceArray(ctx, np.asfortranarray(Bm).ravel(order="F"))
    dC = DeviceArray(ctx, np.zeros(M * N, dtype=np.complex128))
    zgemm_split(ctx, slices, opB, M, N, K, dA, K, dB, K, dC, M)
    out = dC.to_host().reshape((M, N), order="F")
    rowmax = np.max(np.maximum(np.abs(At.real), np.abs(At.imag)), axis=1)
    colmax = np.max(np.maximum(np.abs(Bm.real), np.abs(Bm.imag)), axis=0)
    bound = 8 * K * (slices + 3) * 2.0 ** (-7 * slices) * np.outer(rowmax, colmax) \
        + 1e-14 * (np.abs(At) @ np.abs(Bm))
    err = np.abs(out - ref)
    assert np.all(err <= bound), float(np.max(err / np.maximum(bound, 1e-300)))


def _energies(sp, split, min_k=32):
    from gpu_helpers import context_from_synth
    with context_from_synth(sp) as c:
        c.set_option("split_slices", split)
        c.set_option("split_min_k", min_k)
        c.casci(sp.spec.selectroot)
        out = c.caspt2(sp.spec.selectroot, sp.spec.eshift)
        return out, c.transform_flops()


@pytest.mark.parametrize("workload", ["small", "fp64"])
def test_split_energies_vs_fp64(workload):
    from dirac_caspt2_amd import synth
    from gpu_helpers import say
    sp = synth.generate({"small": synth.CONFIG_SMALL, "fp64": synth.CONFIG_FP64}[workload])
    ref, _ = _energies(sp, 0)
    for slices, min_k in ((7, 16), (6, 16)):
        out, (fs, fd) = _energies(sp, slices, min_k)
        d = {X: abs(out.e2[X] - ref.e2[X]) for X in "ABCDEFGH"}
        say(workload, "slices", slices, "split share %.3f" % (fs / (fs + fd)), "max|dE2| %.2e" % max(d.values()), d)
        assert fs > fd
        assert max(d.values()) < 1e-8, d
        assert abs(out.total - ref.total) < 1e-8


@pytest.mark.parametrize("slices", [6, 7])
def test_split_324_spinors_vs_oracle(slices):
    """6 slices is what bench.py and the served runs use."""
    from dirac_caspt2_amd import synth
    from gpu_helpers import say
    if not os.path.exists(GOLDEN):
        pytest.skip("tests/golden/synthetic_split.json has not been generated")
    g = json.load(open(GOLDEN))
    sp = synth.generate(synth.CONFIG_SPLIT)
    fp64, _ = _energies(sp, 0)
    split, (fs, fd) = _energies(sp, slices)
    d64 = {X: abs(fp64.e2[X] - g["e2"][X]) for X in "ABCDEFGH"}
    dsp = {X: abs(split.e2[X] - g["e2"][X]) for X in "ABCDEFGH"}
    say("324 spinors, %d slices: fp64 vs oracle" % slices, d64, "split vs oracle", dsp, "split share %.3f" % (fs / (fs + fd)))
    assert max(d64.values()) < 1e-9 and abs(fp64.total - g["total"]) < 1e-9
    assert max(dsp.values()) < 1e-8 and abs(split.total - g["total"]) < 1e-8
