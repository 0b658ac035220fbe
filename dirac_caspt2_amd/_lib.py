"""ctypes binding of libdcaspt2_amd.so (the C ABI declared in include/dcaspt2_amd.h).

The product path has no fallback: when the shared library is missing the
import of the binding raises, and every call through it fails loudly.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdcaspt2_amd.so")

c_int = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
P = ctypes.c_void_p
PD = ctypes.POINTER(ctypes.c_double)
PI64 = ctypes.POINTER(ctypes.c_int64)
PI32 = ctypes.POINTER(ctypes.c_int32)

_SIGS = {
    "dc2_context_create": (c_int, [c_int, ctypes.POINTER(P)]),
    "dc2_context_destroy": (c_int, [P]),
    "dc2_last_error": (ctypes.c_char_p, []),
    "dc2_version": (ctypes.c_char_p, []),
    "dc2_set_spaces": (c_int, [P, c_int, c_int, c_int, c_int]),
    "dc2_set_symmetry": (c_int, [P, c_int, P, P, c_int]),
    "dc2_set_ras": (c_int, [P, c_int, c_int, c_int, c_int, c_int, c_int, c_int]),
    "dc2_set_one_electron": (c_int, [P, P, c_dbl]),
    "dc2_set_integral_block": (c_int, [P, ctypes.c_char_p, P, c_i64, c_int]),
    "dc2_casci": (c_int, [P, c_int, P, P]),
    "dc2_get_dets": (c_int, [P, P]),
    "dc2_get_ci_vector": (c_int, [P, c_int, P]),
    "dc2_caspt2": (c_int, [P, c_int, c_dbl, P, P, P]),
    "dc2_get_subspace_info": (c_int, [P, P, P]),
    "dc2_get_fock_rotation": (c_int, [P, P, P]),
    "dc2_get_rotated_ci": (c_int, [P, P]),
    "dc2_get_density": (c_int, [P, c_int, P, c_i64]),
    "dc2_get_transformed_block": (c_int, [P, ctypes.c_char_p, P, c_i64]),
    "dc2_get_timings": (c_int, [P, P, c_int]),
    "dc2_casmat": (c_int, [P, P]),
    "dc2_cdiag": (c_int, [P, P, c_int, P, P, c_dbl]),
    "dc2_zgemm": (c_int, [P, ctypes.c_char, ctypes.c_char, c_i64, c_i64, c_i64, P, P, c_i64, c_i64, P, c_i64, c_i64,
                          P, P, c_i64, c_i64, c_i64]),
    "dc2_zherk": (c_int, [P, c_i64, c_i64, P, c_i64, P, c_i64, c_int]),
    "dc2_synchronize": (c_int, [P]),
    "dc2_malloc": (c_int, [P, c_i64, ctypes.POINTER(P)]),
    "dc2_free": (c_int, [P, P]),
    "dc2_memcpy": (c_int, [P, P, P, c_i64, c_int]),
    "dc2_set_integral_factors": (c_int, [P, P, c_i64, c_int]),
    "dc2_set_profiling": (c_int, [P, c_int]),
    "dc2_set_option": (c_int, [P, ctypes.c_char_p, c_i64]),
    "dc2_get_gemm_op_stats": (c_int, [P, ctypes.c_char, ctypes.c_char, P]),
    "dc2_get_gemm_shapes": (c_int, [P, P, c_i64, P]),
    "dc2_get_graph_stats": (c_int, [P, P]),
    "dc2_gemm_tune": (c_int, [P, ctypes.c_char, ctypes.c_char, c_i64, c_i64, c_i64, c_i64, c_int, P]),
    "dc2_gemm_set_choice": (c_int, [ctypes.c_char, ctypes.c_char, c_i64, c_i64, c_i64, c_i64, c_int, c_int]),
    "dc2_gemm_clear_choices": (c_int, [P]),
    "dc2_get_integral_block_device": (c_int, [P, ctypes.c_char_p, P, P]),
    "dc2_set_integral_records": (c_int, [P, c_i64, P, P, c_int]),
    "dc2_mdcint_records": (c_int, [P, c_i64, P, P, P, P, P, P, P, c_i64, P, c_i64, P, c_int, c_i64, P, P, P, c_int]),
    "dc2_zgemm_split": (c_int, [P, c_int, ctypes.c_char, c_i64, c_i64, c_i64, P, c_i64, P, c_i64, P, c_i64]),
    "dc2_zgemm_split_noalloc": (c_int, [P, c_int, ctypes.c_char, c_i64, c_i64, c_i64, P, c_i64, P, c_i64, P, c_i64]),
    "dc2_get_transform_flops": (c_int, [P, P]),
    "dc2_get_casci_info": (c_int, [P, P]),
    "dc2_host_heev": (c_int, [c_int, P, P, c_int]),
    "dc2_fockivo": (c_int, [P, c_int, P, c_int, P, P, P, P]),
    "dc2_get_hetrd_stats": (c_int, [c_int, P]),
    "dc2_set_casci_result": (c_int, [P, c_i64, c_int, P, P, P]),
    "dc2_caspt2_restart": (c_int, [P, c_int, ctypes.c_double, c_int, P, P, P, P, P]),
    "dc2_request_footprint": (c_int, [c_int, c_int, c_int, c_int, P]),
    "dc2_server_create": (c_int, [c_int, c_int, c_i64, c_int, P]),
    "dc2_server_create_batched": (c_int, [c_int, c_int, c_i64, c_int, c_int, ctypes.c_double, P]),
    "dc2_server_destroy": (c_int, [P]),
    "dc2_server_add_molecule": (c_int, [P, c_int, c_int, c_int, c_int, P, c_i64, P]),
    "dc2_server_submit": (c_int, [P, c_int, P, ctypes.c_double, c_int, ctypes.c_double, P, c_int, P]),
    "dc2_server_wait": (c_int, [P, ctypes.c_double, P, P, P]),
    "dc2_server_stats": (c_int, [P, P]),
    "dc2_server_set_checkpoint_dir": (c_int, [P, ctypes.c_char_p]),
    "dc2_server_set_option": (c_int, [P, ctypes.c_char_p, c_i64]),
    "dc2_server_submit_restart": (c_int, [P, c_int, P, ctypes.c_double, c_int, ctypes.c_double, P, c_int, c_int, P, P,
                                          P]),
    "dc2_set_subspace_callback": (c_int, [P, P, P]),
    "dc2_write_restart": (c_int, [ctypes.c_char_p, c_int, c_int, c_int, P, P]),
    "dc2_get_gemm_stats": (c_int, [P, c_int, P]),
    "dc2_get_eig_fallbacks": (c_int, [P, P]),
}

EXPORTED = tuple(_SIGS)


class Dc2Error(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m dirac_caspt2_amd.build`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def check(fn: str, status: int) -> None:
    if status != 0:
        raise Dc2Error(fn, status, lib().dc2_last_error().decode(errors="replace"))
