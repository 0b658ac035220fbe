"""Shared helpers of the GPU tests: load oracle problems into the HIP library."""
import sys

import numpy as np

from dirac_caspt2_amd.api import Caspt2Context
from dirac_caspt2_amd.problem import load_dense, load_factors
from oracle import caspt2_ref as R


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def context_from_problem(p: R.Problem, device: int = 0) -> Caspt2Context:
    ctx = Caspt2Context(device)
    dims = (p.ninact, p.nact, p.nsec, p.nelec)
    ras = None
    if p.ras is not None:
        r = p.ras
        ras = (r.ras1_size, r.ras2_size, r.ras3_size, r.ras1_max_hole, r.ras3_max_elec, r.min_hole_ras1,
               r.min_elec_ras3)
    load_dense(ctx, dims, p.h, lambda *s: p.eri.block(*s), p.ecore, nsymrpa=p.sym.n, multb=p.sym.multb,
               irpamo=p.sym.irpamo, totsym=p.totsym, ras=ras)
    return ctx


def context_from_synth(sp, device: int = 0, on_gpu: bool = True) -> Caspt2Context:
    s = sp.spec
    ctx = Caspt2Context(device)
    load_factors(ctx, (s.ninact, s.nact, s.nsec, s.nelec), sp.h, sp.B, sp.ecore, on_device=on_gpu)
    return ctx


def oracle_problem(sp) -> R.Problem:
    s = sp.spec
    return R.Problem(s.ninact, s.nact, s.nsec, s.nelec, sp.h, R.FactorIntegrals(sp.B), sp.ecore, s.selectroot,
                     eshift=s.eshift)


def assert_energies(gpu, ref: R.Caspt2Result, tol: float = 1e-9, label: str = ""):
    for X in "ABCDEFGH":
        d = abs(gpu.e2[X] - ref.e2[X])
        assert d < tol, f"{label} e2{X}: gpu {gpu.e2[X]!r} oracle {ref.e2[X]!r} |diff| {d:.3e}"
        d = abs(gpu.sumc2[X] - ref.sumc2[X])
        assert d < tol, f"{label} sumc2{X}: |diff| {d:.3e}"
    assert abs(gpu.e0 - ref.e0) < tol, f"{label} e0"
    assert abs(gpu.total - ref.total) < tol, f"{label} total {gpu.total} vs {ref.total}"
