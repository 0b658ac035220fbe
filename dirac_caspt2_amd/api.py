"""Host-side mirror of the reference's CASCI/CASPT2 driver over the C ABI.

The reference keeps the problem in module globals (module_global_variables.F90)
filled by its readers and then calls casci (casci.F90) and r4dcaspt2_tra
(r4dcaspt2_tra.F90).  ``Caspt2Context`` holds the same state inside a GPU
context: ``set_spaces`` / ``set_symmetry`` / ``set_ras`` / ``set_one_electron``
/ ``set_integral_block`` correspond to what read_mrconee, readint2_casci and
divide_2_elec_integral_into_subspaces produce; ``casci()`` and ``caspt2()``
run the two subprograms.  Arrays are column-major (Fortran order) complex128,
given as numpy arrays (host) or DeviceArray buffers (device).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check

SUBSPACES = "ABCDEFGH"
BLOCK_SHAPES = {
    "OOOO": "oooo", "F1": "ssoo", "F2": "soos",
    "A1": "aiaa", "A2": "aiii", "B": "aiai",
    "C1": "saaa", "C2": "saii", "C3": "siia", "D1": "siaa", "D2": "saai", "D3": "siii",
    "E": "siai", "F": "sasa", "G": "sisa", "H": "sisi",
}
# blocks the C ABI needs (A1/A2/B are cut out of OOOO on the device when absent)
REQUIRED_BLOCKS = ("OOOO", "F1", "F2", "C1", "C2", "C3", "D1", "D2", "D3", "E", "F", "G", "H")


@dataclass
class Caspt2Energies:
    e2: Dict[str, float]
    sumc2: Dict[str, float]
    e0: float
    e2all: float
    sumc2_all: float
    weight0: float
    total_second_order: float
    total: float
    info: Dict[str, tuple] = field(default_factory=dict)
    margin: Dict[str, float] = field(default_factory=dict)
    timings: List[float] = field(default_factory=list)


def _ptr(a) -> int:
    """Data pointer of a numpy array or a DeviceArray."""
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if isinstance(a, DeviceArray):
        return a.ptr
    return a.data_ptr()


class DeviceArray:
    """Device allocation made through the library's own HIP runtime (dc2_malloc)."""

    def __init__(self, ctx: "Caspt2Context", host: Optional[np.ndarray] = None, nbytes: int = 0):
        self.ctx = ctx
        if host is not None:
            host = np.ascontiguousarray(host)
            nbytes = host.nbytes
        p = ctypes.c_void_p()
        check("dc2_malloc", ctx.lib.dc2_malloc(ctx.h, int(nbytes), ctypes.byref(p)))
        self.ptr = p.value or 0
        self.nbytes = int(nbytes)
        if host is not None and nbytes:
            check("dc2_memcpy", ctx.lib.dc2_memcpy(ctx.h, self.ptr, host.ctypes.data, self.nbytes, 1))

    def to_host(self, dtype=np.complex128) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        if self.nbytes:
            check("dc2_memcpy", self.ctx.lib.dc2_memcpy(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes, 2))
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            check("dc2_free", self.ctx.lib.dc2_free(self.ctx.h, self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _fortran_complex(a: np.ndarray) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


class Caspt2Context:
    """One GPU context of libdcaspt2_amd (one problem resident in HBM at a time)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.lib()
        h = ctypes.c_void_p()
        check("dc2_context_create", self.lib.dc2_context_create(device, ctypes.byref(h)))
        self.h = h
        self.dims = None
        self._keep: Dict[str, object] = {}

    def close(self):
        if self.h:
            check("dc2_context_destroy", self.lib.dc2_context_destroy(self.h))
            self.h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- problem definition --------------------------------------------------
    def set_spaces(self, ninact: int, nact: int, nsec: int, nelec: int):
        check("dc2_set_spaces", self.lib.dc2_set_spaces(self.h, ninact, nact, nsec, nelec))
        self.dims = (ninact, nact, nsec, nelec)
        self._keep.clear()

    def set_symmetry(self, nsymrpa: int, multb: np.ndarray, irpamo: Sequence[int], totsym: int):
        mb = np.asfortranarray(np.asarray(multb, dtype=np.int64))
        ir = np.ascontiguousarray(np.asarray(irpamo, dtype=np.int64))
        check("dc2_set_symmetry", self.lib.dc2_set_symmetry(self.h, nsymrpa, _ptr(mb), _ptr(ir), totsym))

    def set_ras(self, ras1: int, ras2: int, ras3: int, ras1_max_hole: int, ras3_max_elec: int,
                min_hole_ras1: int = 0, min_elec_ras3: int = 0):
        check("dc2_set_ras", self.lib.dc2_set_ras(self.h, ras1, ras2, ras3, ras1_max_hole, ras3_max_elec,
                                                   min_hole_ras1, min_elec_ras3))

    def set_one_electron(self, h: np.ndarray, ecore: float):
        hf = _fortran_complex(h)
        check("dc2_set_one_electron", self.lib.dc2_set_one_electron(self.h, _ptr(hf), float(ecore)))

    def set_integral_block(self, name: str, data, borrow: bool = False):
        """data: numpy array [i,j,k,l] (any order, copied from host) or a DeviceArray
        tensor whose memory is already column-major (i fastest)."""
        if isinstance(data, np.ndarray):
            arr = _fortran_complex(data)
            where = 0
        else:
            arr = data
            where = 2 if borrow else 1
            if borrow:
                self._keep[name] = data
        n = int(np.prod(arr.shape)) if isinstance(arr, np.ndarray) else arr.nbytes // 16
        check("dc2_set_integral_block", self.lib.dc2_set_integral_block(self.h, name.encode(), _ptr(arr), n, where))

    def set_integral_records(self, idx: np.ndarray, val: np.ndarray):
        """MDCINTNEW-equivalent records: idx (n, 4) 1-based CAS-order labels of (ij|kl), val (n,)
        complex; every block is filled on the device from their 16 symmetry images."""
        ijkl = np.ascontiguousarray(idx, dtype=np.int32)
        v = np.ascontiguousarray(np.asarray(val, dtype=np.complex128))
        check("dc2_set_integral_records", self.lib.dc2_set_integral_records(self.h, ijkl.shape[0], _ptr(ijkl),
                                                                            _ptr(v), 0))

    def mdcint_records(self, ikr, jkr, nz, indk, indl, rklr, rkli, nkr: int, kr, d2c, scheme: int,
                       ingest: bool = False, want_records: bool = True):
        """dc2_mdcint_records: DIRAC's MDCINT records (per record ikr, jkr, nz; entries indk, indl,
        rklr, rkli or None for real integrals; kr[nkr + m], d2c[k]) -> MDCINTNEW-equivalent
        records (idx (n, 4) 1-based CAS labels, val (n,) complex) converted on the GPU, in the
        reference's order; ingest: they also fill every integral block."""
        i64 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int64))
        f64 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64))
        ikr, jkr, nz, indk, indl, kr, d2c = map(i64, (ikr, jkr, nz, indk, indl, kr, d2c))
        rklr = f64(rklr)
        rkli = None if rkli is None else f64(rkli)
        nrec = len(ikr)
        nout = ctypes.c_int64()
        args = [self.h, nrec, _ptr(ikr), _ptr(jkr), _ptr(nz), _ptr(indk), _ptr(indl), _ptr(rklr),
                None if rkli is None else _ptr(rkli), int(nkr), _ptr(kr), len(d2c), _ptr(d2c), int(scheme)]
        cap = int(np.sum(nz)) * 4
        if want_records:
            ijkl = np.zeros((max(cap, 1), 4), dtype=np.int32)
            val = np.zeros(max(cap, 1), dtype=np.complex128)
            check("dc2_mdcint_records", self.lib.dc2_mdcint_records(*args, cap, ctypes.byref(nout), _ptr(ijkl),
                                                                    _ptr(val), int(ingest)))
            n = nout.value
            return ijkl[:n].copy(), val[:n].copy()
        check("dc2_mdcint_records", self.lib.dc2_mdcint_records(*args, 0, ctypes.byref(nout), None, None,
                                                                int(ingest)))
        return nout.value

    def integral_block(self, name: str) -> np.ndarray:
        """Copy of a resident original-basis integral block (column-major, its four extents)."""
        p = ctypes.c_void_p()
        n = ctypes.c_int64()
        check("dc2_get_integral_block_device", self.lib.dc2_get_integral_block_device(self.h, name.encode(),
                                                                                      ctypes.byref(p), ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.complex128)
        check("dc2_memcpy", self.lib.dc2_memcpy(self.h, out.ctypes.data, p.value, 16 * n.value, 2))
        return out.reshape(block_shape(self.dims, name), order="F")

    def set_integral_factors(self, B: np.ndarray):
        """(pq|rs) = sum_P B[P,p,q] B[P,r,s]: every block is formed on the device."""
        Bf = np.asfortranarray(np.asarray(B, dtype=np.complex128))
        check("dc2_set_integral_factors", self.lib.dc2_set_integral_factors(self.h, _ptr(Bf), Bf.shape[0], 0))

    # --- CASCI ----------------------------------------------------------------
    def casci(self, nroot: int = 1):
        eig = np.zeros(max(nroot, 1))
        nd = ctypes.c_int64()
        check("dc2_casci", self.lib.dc2_casci(self.h, nroot, _ptr(eig), ctypes.byref(nd)))
        self.ndet = nd.value
        return eig[:min(nroot, nd.value)]

    def dets(self) -> np.ndarray:
        out = np.zeros(self.ndet, dtype=np.int64)
        check("dc2_get_dets", self.lib.dc2_get_dets(self.h, _ptr(out)))
        return out

    def ci_vector(self, root: int = 1) -> np.ndarray:
        out = np.zeros(self.ndet, dtype=np.complex128)
        check("dc2_get_ci_vector", self.lib.dc2_get_ci_vector(self.h, root, _ptr(out)))
        return out

    def casmat(self) -> np.ndarray:
        out = np.zeros((self.ndet, self.ndet), dtype=np.complex128, order="F")
        check("dc2_casmat", self.lib.dc2_casmat(self.h, _ptr(out)))
        return out

    # --- CASPT2 ---------------------------------------------------------------
    def caspt2(self, selectroot: int = 1, eshift: float = 0.0) -> Caspt2Energies:
        e2 = np.zeros(8)
        s2 = np.zeros(8)
        sc = np.zeros(6)
        check("dc2_caspt2", self.lib.dc2_caspt2(self.h, selectroot, float(eshift), _ptr(e2), _ptr(s2), _ptr(sc)))
        return self._energies(e2, s2, sc)

    def caspt2_restart(self, selectroot: int, eshift: float, next_subspace: str, e2_file: Dict[str, float],
                       sumc2_file: Dict[str, float]) -> Caspt2Energies:
        """CASPT2 resumed from a caspt2_restart file: subspaces before next_subspace ('A'..'H', or
        'I' for none) are skipped; e2_file/sumc2_file hold the file's values (absent = 0)."""
        first = "ABCDEFGHI".index(next_subspace)
        fe = np.array([e2_file.get(X, 0.0) for X in SUBSPACES])
        fs = np.array([sumc2_file.get(X, 0.0) for X in SUBSPACES])
        e2 = np.zeros(8)
        s2 = np.zeros(8)
        sc = np.zeros(6)
        check("dc2_caspt2_restart", self.lib.dc2_caspt2_restart(self.h, selectroot, float(eshift), first, _ptr(fe),
                                                                _ptr(fs), _ptr(e2), _ptr(s2), _ptr(sc)))
        return self._energies(e2, s2, sc)

    def set_casci_result(self, dets, coeffs: np.ndarray, ecas):
        """Use a CASCI result computed elsewhere (determinants in order, complex CI vectors
        ndet x nroot, energies without ecore) instead of casci()."""
        d = np.ascontiguousarray(np.asarray(dets, dtype=np.int64))
        c = np.asfortranarray(np.asarray(coeffs, dtype=np.complex128).reshape(len(d), -1))
        e = np.ascontiguousarray(np.asarray(ecas, dtype=np.float64))
        check("dc2_set_casci_result", self.lib.dc2_set_casci_result(self.h, len(d), c.shape[1], _ptr(d), _ptr(c),
                                                                    _ptr(e)))
        self.ndet = len(d)

    def _energies(self, e2, s2, sc) -> Caspt2Energies:
        info = np.zeros(32, dtype=np.int64)
        margin = np.zeros(8)
        check("dc2_get_subspace_info", self.lib.dc2_get_subspace_info(self.h, _ptr(info), _ptr(margin)))
        tm = np.zeros(12)
        check("dc2_get_timings", self.lib.dc2_get_timings(self.h, _ptr(tm), 12))
        return Caspt2Energies(
            e2={X: float(e2[k]) for k, X in enumerate(SUBSPACES)},
            sumc2={X: float(s2[k]) for k, X in enumerate(SUBSPACES)},
            e0=float(sc[0]), e2all=float(sc[1]), sumc2_all=float(sc[2]), weight0=float(sc[3]),
            total_second_order=float(sc[4]), total=float(sc[5]),
            info={X: tuple(int(v) for v in info[4 * k:4 * k + 4]) for k, X in enumerate(SUBSPACES)},
            margin={X: float(margin[k]) for k, X in enumerate(SUBSPACES)},
            timings=[float(t) for t in tm])

    def fock_rotation(self):
        n = sum(self.dims[:3])
        eps = np.zeros(n)
        U = np.zeros((n, n), dtype=np.complex128, order="F")
        check("dc2_get_fock_rotation", self.lib.dc2_get_fock_rotation(self.h, _ptr(eps), _ptr(U)))
        return eps, U

    def rotated_ci(self) -> np.ndarray:
        out = np.zeros(self.ndet, dtype=np.complex128)
        check("dc2_get_rotated_ci", self.lib.dc2_get_rotated_ci(self.h, _ptr(out)))
        return out

    def density(self, which: int) -> np.ndarray:
        na = self.dims[1]
        n = na * na if which in (0, 1) else na ** 4
        out = np.zeros(n, dtype=np.complex128)
        check("dc2_get_density", self.lib.dc2_get_density(self.h, which, _ptr(out), n))
        return out

    def transformed_block(self, name: str) -> np.ndarray:
        shape = block_shape(self.dims, name)
        out = np.zeros(shape, dtype=np.complex128, order="F")
        check("dc2_get_transformed_block",
              self.lib.dc2_get_transformed_block(self.h, name.encode(), _ptr(out), int(np.prod(shape))))
        return out

    def synchronize(self):
        check("dc2_synchronize", self.lib.dc2_synchronize(self.h))

    def set_profiling(self, on: bool):
        check("dc2_set_profiling", self.lib.dc2_set_profiling(self.h, 1 if on else 0))

    GEMM_TAGS = {"all": -1, "other": 0, "transform": 1, "density": 2, "vbuild": 3, "projection": 4, "factors": 5}

    def gemm_stats(self, tag: str = "all"):
        out = np.zeros(4)
        check("dc2_get_gemm_stats", self.lib.dc2_get_gemm_stats(self.h, self.GEMM_TAGS[tag], _ptr(out)))
        return {"launches": int(out[0]), "seconds": float(out[1]), "flops": float(out[2]), "max_flops": float(out[3])}

    def gemm_op_stats(self, opA: str, opB: str):
        """GEMM accounting of the launches of kernel instantiation k_zgemm<opA, opB>."""
        out = np.zeros(4)
        check("dc2_get_gemm_op_stats", self.lib.dc2_get_gemm_op_stats(self.h, opA.encode(), opB.encode(), _ptr(out)))
        return {"launches": int(out[0]), "seconds": float(out[1]), "flops": float(out[2]), "max_flops": float(out[3])}

    def graph_stats(self):
        """Captured Davidson graphs: {captures, replays, held, enabled}."""
        out = np.zeros(4, dtype=np.int64)
        check("dc2_get_graph_stats", self.lib.dc2_get_graph_stats(self.h, _ptr(out)))
        return dict(zip(("captures", "replays", "held", "enabled"), (int(x) for x in out)))

    def gemm_shapes(self):
        """Distinct shapes (opA, opB, M, N, K, batch) of the MFMA GEMM launches recorded while
        profiling was on."""
        n = ctypes.c_int64()
        check("dc2_get_gemm_shapes", self.lib.dc2_get_gemm_shapes(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros(6 * max(n.value, 1), dtype=np.int64)
        check("dc2_get_gemm_shapes", self.lib.dc2_get_gemm_shapes(self.h, _ptr(out), n.value, ctypes.byref(n)))
        return [(chr(r[0]), chr(r[1]), int(r[2]), int(r[3]), int(r[4]), int(r[5]))
                for r in out[:6 * n.value].reshape(-1, 6)]

    def gemm_tune(self, opA: str, opB: str, M: int, N: int, K: int, batch: int = 1, reps: int = 3):
        """Time every kernel variant / split-K candidate for this shape and install the fastest
        (dc2_gemm_tune): {variant, split (0, 0: built-in kept), seconds, builtin_seconds}."""
        out = np.zeros(4)
        check("dc2_gemm_tune", self.lib.dc2_gemm_tune(self.h, opA.encode(), opB.encode(), M, N, K, batch, reps,
                                                      _ptr(out)))
        return {"variant": int(out[0]), "split": int(out[1]), "seconds": float(out[2]),
                "builtin_seconds": float(out[3])}

    def casci_info(self):
        """(ndet, roots held, Davidson iterations of the last casci; 0 = dense solver)."""
        out = np.zeros(3, dtype=np.int64)
        check("dc2_get_casci_info", self.lib.dc2_get_casci_info(self.h, _ptr(out)))
        return tuple(int(x) for x in out)

    def transform_flops(self):
        """(split-precision, FP64) transform flops of the last caspt2()."""
        out = np.zeros(2)
        check("dc2_get_transform_flops", self.lib.dc2_get_transform_flops(self.h, _ptr(out)))
        return float(out[0]), float(out[1])

    OPTIONS = {"eigensolver": {"auto": -1, "qr": 0, "dc": 1, "dcj": 2, "jacobi": 3, "native": 4},
               "ac_metric": {"auto": 0, "formula": 1, "determinant": 2, "antisymmetric": 3,
                             "antisymmetric_det": 4},
               "lanes": {}, "schur": {"off": 0, "on": 1}, "split_slices": {}, "split_min_k": {},
               "casci_solver": {"auto": 0, "dense": 1, "davidson": 2}, "graphs": {"off": 0, "on": 1},
               "tridiag_solver": {"own": 0, "rocsolver": 1}, "lane_plan": {"auto": -1}, "c_thin": {"off": 0, "on": 1},
               "davidson_block": {}, "davidson_maxsub": {}, "early_v": {"off": 0, "on": 1}}

    def set_option(self, key: str, value):
        """Options: eigensolver auto|qr|dc|dcj|jacobi|native, ac_metric auto|formula|determinant|
        antisymmetric|antisymmetric_det,
        lanes 1..4 (concurrent subspace streams), schur on|off, graphs on|off (captured
        Davidson iterations), tridiag_solver own|rocsolver (the tridiagonal stage of the native
        eigensolver)."""
        v = self.OPTIONS[key][value] if isinstance(value, str) else int(value)
        check("dc2_set_option", self.lib.dc2_set_option(self.h, key.encode(), v))

    def fockivo(self, numh: int, mo_energy: np.ndarray, irreps: Sequence[int]):
        """dc2_fockivo: (IVO Fock matrix of the secondary spinors, per-irrep eigenvalues at the
        spinors' positions, block-diagonal eigenvector matrix U)."""
        ns = self.dims[2]
        eps = np.ascontiguousarray(mo_energy, dtype=np.float64)
        irr = np.ascontiguousarray(irreps, dtype=np.int32)
        fock = np.zeros((ns, ns), dtype=np.complex128, order="F")
        w = np.zeros(ns)
        U = np.zeros((ns, ns), dtype=np.complex128, order="F")
        check("dc2_fockivo", self.lib.dc2_fockivo(self.h, int(numh), _ptr(eps), len(irr), _ptr(irr), _ptr(fock),
                                                  _ptr(w), _ptr(U)))
        return fock, w, U

    def eig_fallbacks(self) -> int:
        n = ctypes.c_int64()
        check("dc2_get_eig_fallbacks", self.lib.dc2_get_eig_fallbacks(self.h, ctypes.byref(n)))
        return n.value


def cdiag(ctx: Caspt2Context, c: np.ndarray, threshold: float = 0.0):
    """diag.F90:66-128 on the GPU: returns (w, eigenvectors, dimm)."""
    a = _fortran_complex(c).copy(order="F")
    n = a.shape[0]
    w = np.zeros(n)
    dimm = ctypes.c_int32()
    check("dc2_cdiag", ctx.lib.dc2_cdiag(ctx.h, _ptr(a), n, ctypes.byref(dimm), _ptr(w), float(threshold)))
    return w, a, dimm.value


def zgemm(ctx: Caspt2Context, opA: str, opB: str, M: int, N: int, K: int, alpha: complex, A, lda: int, sA: int,
          B, ldb: int, sB: int, beta: complex, C, ldc: int, sC: int, batch: int = 1):
    """C = alpha op(A) op(B) + beta C on device memory (DeviceArray buffers)."""
    al = np.array([alpha.real, alpha.imag])
    be = np.array([beta.real, beta.imag])
    check("dc2_zgemm", ctx.lib.dc2_zgemm(ctx.h, opA.encode(), opB.encode(), M, N, K, _ptr(al), _ptr(A), lda, sA,
                                         _ptr(B), ldb, sB, _ptr(be), _ptr(C), ldc, sC, batch))


def zherk(ctx: Caspt2Context, m: int, k: int, A, lda: int, C, ldc: int, split: int = 0):
    """C = A A^H (m x m, the whole Hermitian matrix) on device memory; split 0 lets the
    library choose the number of k chunks, 1 leaves k unsplit."""
    check("dc2_zherk", ctx.lib.dc2_zherk(ctx.h, m, k, _ptr(A), lda, _ptr(C), ldc, split))


def space_sizes(dims):
    ni, na, ns = dims[:3]
    return {"i": ni, "a": na, "s": ns, "o": ni + na}


def space_slices(dims):
    ni, na, ns = dims[:3]
    return {"i": slice(0, ni), "a": slice(ni, ni + na), "s": slice(ni + na, ni + na + ns), "o": slice(0, ni + na)}


def block_shape(dims, name: str):
    sz = space_sizes(dims)
    return tuple(sz[ch] for ch in BLOCK_SHAPES[name])


def needed_blocks(dims) -> List[str]:
    ni, na, ns = dims[:3]
    names = ["OOOO"]
    if ns:
        names += ["F1", "F2", "C1"]
        if ni:
            names += ["C2", "C3", "D1", "D2", "D3", "E", "F", "G", "H"]
        else:
            names += ["F"]
    return [n for n in names if all(v > 0 for v in block_shape(dims, n))]


@dataclass
class ServedResult:
    id: int
    status: int
    e2: Dict[str, float]
    sumc2: Dict[str, float]
    e0: float
    e2all: float
    sumc2_all: float
    weight0: float
    total2: float
    total: float
    t_submit: float
    t_start: float
    t_end: float
    slot: int
    error: str = ""

    @property
    def latency(self) -> float:
        return self.t_end - self.t_submit


class Server:
    """Continuous-batching request server of the library (dc2_server_*): nslots concurrent
    energies on one GPU, molecules resident in HBM, admission by device-memory footprint."""

    def __init__(self, device: int = 0, nslots: int = 4, mem_budget: int = 200 << 30, lanes_per_slot: int = 1,
                 eig_batch: int = 0, eig_window: float = 0.002):
        """eig_batch > 1 batches the eigensolves of concurrent requests (groups of equal size,
        flushed when all running requests wait, when full, or eig_window seconds late)."""
        self.lib = _lib.lib()
        h = ctypes.c_void_p()
        check("dc2_server_create", self.lib.dc2_server_create_batched(device, nslots, int(mem_budget), lanes_per_slot,
                                                                      eig_batch, float(eig_window), ctypes.byref(h)))
        self.h = h.value
        self.dims: Dict[int, tuple] = {}

    def add_molecule(self, dims, B: np.ndarray) -> int:
        """dims = (ninact, nact, nsec, nelec); (pq|rs) = sum_P B[P,p,q] B[P,r,s]."""
        Bf = np.asfortranarray(np.asarray(B, dtype=np.complex128))
        mid = ctypes.c_int32()
        check("dc2_server_add_molecule", self.lib.dc2_server_add_molecule(self.h, *[int(x) for x in dims], _ptr(Bf),
                                                                          Bf.shape[0], ctypes.byref(mid)))
        self.dims[mid.value] = tuple(dims)
        return mid.value

    def set_option(self, key: str, value: int):
        """dc2_set_option on every slot's context (dc2_server_set_option)."""
        check("dc2_server_set_option", self.lib.dc2_server_set_option(self.h, key.encode(), int(value)))

    def set_checkpoint_dir(self, path: Optional[str]):
        """Per-subspace checkpoints <path>/<request id>/caspt2_restart_1_<root> (None: off)."""
        check("dc2_server_set_checkpoint_dir",
              self.lib.dc2_server_set_checkpoint_dir(self.h, (path or "").encode()))

    def submit(self, mol: int, h: np.ndarray, ecore: float, selectroot: int = 1, eshift: float = 0.0,
               ras=None, split_slices: int = 0, restart=None) -> int:
        """restart: (first, e2_file, sumc2_file) to resume from a caspt2_restart file's values
        (subspaces before first, 0 = A .. 8 = none, taken from the file)."""
        hf = _fortran_complex(h)
        r = None if ras is None else np.ascontiguousarray(ras, dtype=np.int32)
        rid = ctypes.c_int64()
        first, e2f, s2f = 0, None, None
        if restart is not None:
            first = int(restart[0])
            e2f = np.ascontiguousarray([restart[1].get(X, 0.0) for X in SUBSPACES], dtype=np.float64)
            s2f = np.ascontiguousarray([restart[2].get(X, 0.0) for X in SUBSPACES], dtype=np.float64)
        check("dc2_server_submit", self.lib.dc2_server_submit_restart(
            self.h, mol, _ptr(hf), float(ecore), selectroot, float(eshift), None if r is None else _ptr(r),
            int(split_slices), first, None if e2f is None else _ptr(e2f), None if s2f is None else _ptr(s2f),
            ctypes.byref(rid)))
        return rid.value

    def wait(self, timeout: float = 1.0) -> Optional[ServedResult]:
        rid = ctypes.c_int64()
        st = ctypes.c_int32()
        out = np.zeros(26)
        check("dc2_server_wait", self.lib.dc2_server_wait(self.h, float(timeout), ctypes.byref(rid),
                                                          ctypes.byref(st), _ptr(out)))
        if rid.value == 0:
            return None
        err = self.lib.dc2_last_error().decode() if st.value else ""
        return ServedResult(rid.value, st.value, dict(zip(SUBSPACES, out[0:8])), dict(zip(SUBSPACES, out[8:16])),
                            *[float(x) for x in out[16:22]], float(out[22]), float(out[23]), float(out[24]),
                            int(out[25]), err)

    def stats(self) -> Dict[str, int]:
        out = np.zeros(7, dtype=np.int64)
        check("dc2_server_stats", self.lib.dc2_server_stats(self.h, _ptr(out)))
        return dict(zip(("pending", "running", "completed", "in_use", "peak", "eig_batches", "eig_jobs"),
                        (int(x) for x in out)))

    def close(self):
        if self.h:
            check("dc2_server_destroy", self.lib.dc2_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def zgemm_split(ctx: Caspt2Context, slices: int, opB: str, M, N, K, A: DeviceArray, lda, B: DeviceArray, ldb,
                C: DeviceArray, ldc):
    """Split-precision C = A^T-view * op(B) (dc2_zgemm_split) on device arrays."""
    check("dc2_zgemm_split", ctx.lib.dc2_zgemm_split(ctx.h, slices, opB.encode(), M, N, K, A.ptr, lda, B.ptr, ldb,
                                                     C.ptr, ldc))


def hetrd_stats(n: int):
    """(calls, summed device seconds) of the tridiagonalisation kernel for n x n matrices in
    this process; n = 0 resets the counts."""
    out = np.zeros(2)
    check("dc2_get_hetrd_stats", _lib.lib().dc2_get_hetrd_stats(n, _ptr(out)))
    return int(out[0]), float(out[1])


def write_restart(path: str, totsym: int, selectroot: int, next_subspace: int, e2: Dict[str, float],
                  sumc2: Dict[str, float]):
    """dc2_write_restart: the caspt2_restart file for the subspaces before next_subspace (0..8)."""
    a = np.ascontiguousarray([e2.get(X, 0.0) for X in SUBSPACES], dtype=np.float64)
    b = np.ascontiguousarray([sumc2.get(X, 0.0) for X in SUBSPACES], dtype=np.float64)
    check("dc2_write_restart", _lib.lib().dc2_write_restart(path.encode(), totsym, selectroot, next_subspace, _ptr(a),
                                                            _ptr(b)))


HOST_HEEV_METHODS = {"auto": 0, "jacobi": 1, "ql": 2, "zheev": 3, "low": 4}


def host_heev(a: np.ndarray, method: str = "auto"):
    """Host Hermitian eigensolver of the library (the Davidson subspace solver): (w, V).
    method: "auto" (QL, Jacobi if QL fails), "jacobi", "ql", "zheev" or "low" (every eigenvalue,
    the eigenvectors of the min(n, 4) lowest only, the other columns zero: the Davidson's)."""
    A = np.asfortranarray(np.asarray(a, dtype=np.complex128)).copy(order="F")
    n = A.shape[0]
    w = np.zeros(n)
    check("dc2_host_heev", _lib.lib().dc2_host_heev(n, _ptr(A), _ptr(w), HOST_HEEV_METHODS[method]))
    return w, A


def gemm_set_choice(opA: str, opB: str, M: int, N: int, K: int, batch: int, variant: int, split: int):
    """Install a GEMM kernel choice for one exact shape (dc2_gemm_set_choice)."""
    check("dc2_gemm_set_choice", _lib.lib().dc2_gemm_set_choice(opA.encode(), opB.encode(), M, N, K, batch, variant,
                                                                split))


def gemm_clear_choices() -> int:
    """Empty the GEMM choice table; returns how many choices it held."""
    n = ctypes.c_int64()
    check("dc2_gemm_clear_choices", _lib.lib().dc2_gemm_clear_choices(ctypes.byref(n)))
    return n.value


def request_footprint(ninact: int, nact: int, nsec: int, nelec: int) -> int:
    """Device bytes one served request holds while it runs (the admission estimate)."""
    b = ctypes.c_int64()
    check("dc2_request_footprint", _lib.lib().dc2_request_footprint(ninact, nact, nsec, nelec, ctypes.byref(b)))
    return b.value
