"""CPU oracle for the DIRAC-CASPT2 hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``dirac_caspt2_amd``)
imports, links or executes anything from this package.  It is used by
``tests/`` (as the checker), by ``__graft_entry__.smoke()`` (as the checker)
and by the ``cpu_baseline`` leg of ``bench.py`` (timed beside the GPU path).

Modules
-------
caspt2_ref
    numpy restatement of the reference algorithm: determinant space
    (search_cas_configuration.F90, ras_det_check.F90, one_e_exct.F90),
    CASCI matrix (casmat.F90), Fock build and diagonalisation
    (fockcasci.F90, fockdiag.F90, diag.F90), CI rotation (trac.F90, sdet.F90),
    e0 (calce0.F90), four-index transform (module_intra.F90), one-electron
    transform (tramo.F90), reduced densities (density_matrix.F90) and the
    A-H subspace solvers (solve_[A-H]_subspace.F90, cutoff.F90,
    ulambda_s_half.F90).
dirac_files
    readers for the reference's input files (active.inp, MRCONEE, MDCINT*)
    producing dense integrals in the reference's energy ("CAS") order.

Parity pinning status of each function is recorded in the module docstrings
and in the tests that pin them (tests/test_oracle_*.py).
"""
