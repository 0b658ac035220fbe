"""Load problems into a Caspt2Context: seeded synthetic factors or dense integrals.

Integral blocks are built in the column-major layout the C ABI expects.  For
the synthetic workloads the blocks (pq|rs) = sum_P B^P_pq B^P_rs are formed on
the GPU by the library itself (input preparation, outside any timed region).
"""
from __future__ import annotations

from typing import Callable

import numpy as np

from .api import BLOCK_SHAPES, Caspt2Context, needed_blocks, space_slices


def block_from_factors_numpy(B: np.ndarray, sl) -> np.ndarray:
    L = B[:, sl[0], sl[1]]
    R = B[:, sl[2], sl[3]]
    out = L.reshape(L.shape[0], -1).T @ R.reshape(R.shape[0], -1)
    return out.reshape(L.shape[1], L.shape[2], R.shape[1], R.shape[2])


def load_factors(ctx: Caspt2Context, dims, h: np.ndarray, B: np.ndarray, ecore: float,
                 on_device: bool = True):
    """dims = (ninact, nact, nsec, nelec).  on_device: form the blocks with the
    library's GEMM (dc2_set_integral_factors); otherwise build them with numpy."""
    ctx.set_spaces(*dims)
    ctx.set_one_electron(h, ecore)
    sls = space_slices(dims)
    if on_device:
        ctx.set_integral_factors(B)
    else:
        for name in needed_blocks(dims):
            sl = [sls[ch] for ch in BLOCK_SHAPES[name]]
            ctx.set_integral_block(name, block_from_factors_numpy(B, sl))


def load_dense(ctx: Caspt2Context, dims, h: np.ndarray, block: Callable, ecore: float,
               nsymrpa: int = 1, multb=None, irpamo=None, totsym: int = 1, ras=None):
    """block(s1, s2, s3, s4) -> numpy [i,j,k,l] for python slices of the CAS-order spinors."""
    ctx.set_spaces(*dims)
    if nsymrpa != 1:
        ctx.set_symmetry(nsymrpa, multb, irpamo, totsym)
    if ras is not None:
        ctx.set_ras(*ras)
    ctx.set_one_electron(h, ecore)
    sls = space_slices(dims)
    for name in needed_blocks(dims):
        ctx.set_integral_block(name, block(*[sls[ch] for ch in BLOCK_SHAPES[name]]))
