! iso_c_binding interface of libdcaspt2_amd (include/dcaspt2_amd.h) for a Fortran host.
!
! The bind(C) interfaces follow the C prototypes one for one.  The wrappers below
! them carry the names and argument meaning of the reference subroutines they
! replace, with the module_global_variables state passed explicitly (the
! reference keeps it in module globals, module_global_variables.F90:21-183):
!
!   dcaspt2_amd_casci          casci            (casci.F90:3-72)
!   dcaspt2_amd_casmat         casmat_complex   (casmat.F90:3-222)
!   dcaspt2_amd_cdiag          cdiag            (diag.F90:66-128)
!   dcaspt2_amd_r4dcaspt2_tra  r4dcaspt2_tra    (r4dcaspt2_tra.F90:4-482)
!
! A non-zero status ends the program with that status as its exit code, as
! stop_with_errorcode does (module_error.F90:15-39).
module dcaspt2_amd
    use, intrinsic :: iso_c_binding
    implicit none
    private
    public :: c_ptr, c_null_ptr
    public :: dc2_context_create, dc2_context_destroy, dc2_version, dc2_set_spaces, dc2_set_symmetry, dc2_set_ras, &
              dc2_set_one_electron, dc2_set_integral_block, dc2_set_integral_records, dc2_mdcint_records, &
              dc2_set_integral_factors, &
              dc2_casci, dc2_caspt2, dc2_set_casci_result, dc2_caspt2_restart, &
              dc2_get_subspace_info, dc2_cdiag, dc2_casmat, dc2_synchronize
    public :: dc2_get_dets, dc2_get_ci_vector, dc2_request_footprint, dc2_server_create_batched, &
              dc2_server_add_molecule, dc2_server_submit, dc2_server_wait, dc2_server_stats, dc2_server_destroy
    public :: dcaspt2_amd_check, dcaspt2_amd_last_error, dcaspt2_amd_open, dcaspt2_amd_close, dcaspt2_amd_casci, &
              dcaspt2_amd_casmat, dcaspt2_amd_cdiag, dcaspt2_amd_r4dcaspt2_tra

    interface
        integer(c_int) function dc2_context_create(device, ctx) bind(C, name="dc2_context_create")
            import :: c_int, c_int32_t, c_ptr
            integer(c_int32_t), value :: device
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function dc2_context_destroy(ctx) bind(C, name="dc2_context_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function dc2_last_error_c() bind(C, name="dc2_last_error")
            import :: c_ptr
        end function
        type(c_ptr) function dc2_version() bind(C, name="dc2_version")
            import :: c_ptr
        end function
        integer(c_int) function dc2_set_spaces(ctx, ninact, nact, nsec, nelec) bind(C, name="dc2_set_spaces")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: ninact, nact, nsec, nelec
        end function
        integer(c_int) function dc2_set_symmetry(ctx, nsymrpa, multb, irpamo, totsym) bind(C, name="dc2_set_symmetry")
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nsymrpa, totsym
            integer(c_int64_t), intent(in) :: multb(*), irpamo(*)
        end function
        integer(c_int) function dc2_set_ras(ctx, ras1_size, ras2_size, ras3_size, ras1_max_hole, ras3_max_elec, &
                                            min_hole_ras1, min_elec_ras3) bind(C, name="dc2_set_ras")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: ras1_size, ras2_size, ras3_size, ras1_max_hole, ras3_max_elec, &
                                         min_hole_ras1, min_elec_ras3
        end function
        integer(c_int) function dc2_set_one_electron(ctx, h, ecore) bind(C, name="dc2_set_one_electron")
            import :: c_int, c_double, c_double_complex, c_ptr
            type(c_ptr), value :: ctx
            complex(c_double_complex), intent(in) :: h(*)
            real(c_double), value :: ecore
        end function
        integer(c_int) function dc2_set_integral_block(ctx, name, data, count, where) &
            bind(C, name="dc2_set_integral_block")
            import :: c_int, c_int32_t, c_int64_t, c_char, c_double_complex, c_ptr
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: name(*)
            complex(c_double_complex), intent(in) :: data(*)
            integer(c_int64_t), value :: count
            integer(c_int32_t), value :: where
        end function
        integer(c_int) function dc2_set_integral_records(ctx, n, ijkl, values, where) &
            bind(C, name="dc2_set_integral_records")
            import :: c_int, c_int32_t, c_int64_t, c_double_complex, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n
            integer(c_int32_t), intent(in) :: ijkl(*)
            complex(c_double_complex), intent(in) :: values(*)
            integer(c_int32_t), value :: where
        end function
        ! create_newmdcint.F90:86-177 on the GPU: the MDCINT records as read_one_line_of_mdcint
        ! reads them (entries of all records concatenated), kr(-nkr:nkr), indmo_dirac_to_cas
        ! (0:nmo); rkli, ijkl and values may be c_null_ptr (real integrals / no copy back)
        integer(c_int) function dc2_mdcint_records(ctx, nrec, ikr, jkr, nz, indk, indl, rklr, rkli, nkr, kr, &
                                                   nd2c, d2c, scheme, cap, nout, ijkl, values, ingest) &
            bind(C, name="dc2_mdcint_records")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: nrec, nkr, nd2c, cap
            integer(c_int64_t), intent(in) :: ikr(*), jkr(*), nz(*), indk(*), indl(*), kr(*), d2c(*)
            real(c_double), intent(in) :: rklr(*)
            type(c_ptr), value :: rkli, ijkl, values
            integer(c_int32_t), value :: scheme, ingest
            integer(c_int64_t), intent(out) :: nout
        end function
        integer(c_int) function dc2_set_casci_result(ctx, ndet, nroot, dets, coeffs, ecas) &
            bind(C, name="dc2_set_casci_result")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_double_complex, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: ndet
            integer(c_int32_t), value :: nroot
            integer(c_int64_t), intent(in) :: dets(*)
            complex(c_double_complex), intent(in) :: coeffs(*)
            real(c_double), intent(in) :: ecas(*)
        end function
        integer(c_int) function dc2_caspt2_restart(ctx, selectroot, eshift, first, e2_file, sumc2_file, e2, sumc2, &
                                                   scalars) bind(C, name="dc2_caspt2_restart")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: selectroot, first
            real(c_double), value :: eshift
            real(c_double), intent(in) :: e2_file(8), sumc2_file(8)
            real(c_double), intent(out) :: e2(8), sumc2(8), scalars(6)
        end function
        integer(c_int) function dc2_set_integral_factors(ctx, b, naux, where) bind(C, name="dc2_set_integral_factors")
            import :: c_int, c_int32_t, c_int64_t, c_double_complex, c_ptr
            type(c_ptr), value :: ctx
            complex(c_double_complex), intent(in) :: b(*)
            integer(c_int64_t), value :: naux
            integer(c_int32_t), value :: where
        end function
        integer(c_int) function dc2_casci(ctx, nroot, eigen, ndet) bind(C, name="dc2_casci")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nroot
            real(c_double), intent(out) :: eigen(*)
            integer(c_int64_t), intent(out) :: ndet
        end function
        integer(c_int) function dc2_caspt2(ctx, selectroot, eshift, e2, sumc2, scalars) bind(C, name="dc2_caspt2")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: selectroot
            real(c_double), value :: eshift
            real(c_double), intent(out) :: e2(8), sumc2(8), scalars(6)
        end function
        integer(c_int) function dc2_get_subspace_info(ctx, info, margin) bind(C, name="dc2_get_subspace_info")
            import :: c_int, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(out) :: info(32)
            real(c_double), intent(out) :: margin(8)
        end function
        integer(c_int) function dc2_casmat(ctx, mat) bind(C, name="dc2_casmat")
            import :: c_int, c_double_complex, c_ptr
            type(c_ptr), value :: ctx
            complex(c_double_complex), intent(out) :: mat(*)
        end function
        integer(c_int) function dc2_cdiag(ctx, c, n, dimm, w, threshold) bind(C, name="dc2_cdiag")
            import :: c_int, c_int32_t, c_double, c_double_complex, c_ptr
            type(c_ptr), value :: ctx
            complex(c_double_complex), intent(inout) :: c(*)
            integer(c_int32_t), value :: n
            integer(c_int32_t), intent(out) :: dimm
            real(c_double), intent(out) :: w(*)
            real(c_double), value :: threshold
        end function
        integer(c_int) function dc2_synchronize(ctx) bind(C, name="dc2_synchronize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        ! CASCI result readers (CIDATA contents: determinants and CI vectors)
        integer(c_int) function dc2_get_dets(ctx, dets) bind(C, name="dc2_get_dets")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(out) :: dets(*)
        end function
        integer(c_int) function dc2_get_ci_vector(ctx, root, c) bind(C, name="dc2_get_ci_vector")
            import :: c_int, c_int32_t, c_double_complex, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: root
            complex(c_double_complex), intent(out) :: c(*)
        end function
        ! request server: many energies on one GPU (no counterpart in the reference)
        integer(c_int) function dc2_request_footprint(ninact, nact, nsec, nelec, bytes) &
                bind(C, name="dc2_request_footprint")
            import :: c_int, c_int32_t, c_int64_t
            integer(c_int32_t), value :: ninact, nact, nsec, nelec
            integer(c_int64_t), intent(out) :: bytes
        end function
        integer(c_int) function dc2_server_create_batched(device, nslots, mem_budget, lanes_per_slot, eig_batch, &
                                                          eig_window_s, srv) bind(C, name="dc2_server_create_batched")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            integer(c_int32_t), value :: device, nslots, lanes_per_slot, eig_batch
            integer(c_int64_t), value :: mem_budget
            real(c_double), value :: eig_window_s
            type(c_ptr), intent(out) :: srv
        end function
        integer(c_int) function dc2_server_add_molecule(srv, ninact, nact, nsec, nelec, b, naux, mol_id) &
                bind(C, name="dc2_server_add_molecule")
            import :: c_int, c_int32_t, c_int64_t, c_double_complex, c_ptr
            type(c_ptr), value :: srv
            integer(c_int32_t), value :: ninact, nact, nsec, nelec
            complex(c_double_complex), intent(in) :: b(*)
            integer(c_int64_t), value :: naux
            integer(c_int32_t), intent(out) :: mol_id
        end function
        integer(c_int) function dc2_server_submit(srv, mol_id, h, ecore, selectroot, eshift, ras, flags, req_id) &
                bind(C, name="dc2_server_submit")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_double_complex, c_ptr
            type(c_ptr), value :: srv
            integer(c_int32_t), value :: mol_id, selectroot, flags
            complex(c_double_complex), intent(in) :: h(*)
            real(c_double), value :: ecore, eshift
            type(c_ptr), value :: ras     ! c_loc of integer(c_int32_t) ras(7), or c_null_ptr
            integer(c_int64_t), intent(out) :: req_id
        end function
        integer(c_int) function dc2_server_wait(srv, timeout_s, req_id, req_status, out) &
                bind(C, name="dc2_server_wait")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: srv
            real(c_double), value :: timeout_s
            integer(c_int64_t), intent(out) :: req_id
            integer(c_int32_t), intent(out) :: req_status
            real(c_double), intent(out) :: out(26)
        end function
        integer(c_int) function dc2_server_stats(srv, out) bind(C, name="dc2_server_stats")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: srv
            integer(c_int64_t), intent(out) :: out(7)
        end function
        integer(c_int) function dc2_server_destroy(srv) bind(C, name="dc2_server_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: srv
        end function
    end interface

contains

    ! message of the last failed call on this thread
    function dcaspt2_amd_last_error() result(msg)
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: s(:)
        integer :: n
        p = dc2_last_error_c()
        call c_f_pointer(p, s, [4096])
        n = 0
        do while (n < 4096)
            if (s(n + 1) == c_null_char) exit
            n = n + 1
        end do
        allocate (character(len=n) :: msg)
        msg = transfer(s(1:n), msg)
    end function

    ! stop_with_errorcode(status) for a failed library call (module_error.F90:15-39)
    subroutine dcaspt2_amd_check(status, what)
        integer(c_int), intent(in) :: status
        character(*), intent(in) :: what
        if (status == 0) return
        write (*, '(a,a,a,i0,a,a)') 'dcaspt2_amd: ', what, ' failed with status ', status, ': ', dcaspt2_amd_last_error()
        error stop status
    end subroutine

    ! context on a device with the spinor spaces set
    subroutine dcaspt2_amd_open(ctx, device, ninact, nact, nsec, nelec)
        type(c_ptr), intent(out) :: ctx
        integer, intent(in) :: device, ninact, nact, nsec, nelec
        call dcaspt2_amd_check(dc2_context_create(int(device, c_int32_t), ctx), 'dc2_context_create')
        call dcaspt2_amd_check(dc2_set_spaces(ctx, int(ninact, c_int32_t), int(nact, c_int32_t), int(nsec, c_int32_t), &
                                              int(nelec, c_int32_t)), 'dc2_set_spaces')
    end subroutine

    subroutine dcaspt2_amd_close(ctx)
        type(c_ptr), intent(inout) :: ctx
        call dcaspt2_amd_check(dc2_context_destroy(ctx), 'dc2_context_destroy')
        ctx = c_null_ptr
    end subroutine

    ! casci (casci.F90:31-72): determinants, CI matrix, eigenpairs; eigen(k) = ecas(k) + ecore
    subroutine dcaspt2_amd_casci(ctx, nroot, eigen, ndet)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: nroot
        real(8), intent(out) :: eigen(nroot)
        integer(8), intent(out) :: ndet
        call dcaspt2_amd_check(dc2_casci(ctx, int(nroot, c_int32_t), eigen, ndet), 'casci')
    end subroutine

    ! casmat_complex(mat) (casmat.F90:3): the CI matrix of the last casci
    subroutine dcaspt2_amd_casmat(ctx, mat)
        type(c_ptr), intent(in) :: ctx
        complex(8), intent(out) :: mat(:, :)
        call dcaspt2_amd_check(dc2_casmat(ctx, mat), 'casmat_complex')
    end subroutine

    ! cdiag(c, dimn, dimm, w, thr) (diag.F90:66-128): eigenvectors overwrite c
    subroutine dcaspt2_amd_cdiag(ctx, c, dimn, dimm, w, thr)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: dimn
        complex(8), intent(inout) :: c(dimn, dimn)
        integer, intent(out) :: dimm
        real(8), intent(out) :: w(dimn)
        real(8), intent(in) :: thr
        integer(c_int32_t) :: m
        call dcaspt2_amd_check(dc2_cdiag(ctx, c, int(dimn, c_int32_t), m, w, thr), 'cdiag')
        dimm = int(m)
    end subroutine

    ! r4dcaspt2_tra (r4dcaspt2_tra.F90:164-442) for root selectroot: per subspace
    ! e2_subspace / sumc2_subspace (A..H), e2all, sumc2, weight0 and the total energy
    subroutine dcaspt2_amd_r4dcaspt2_tra(ctx, selectroot, eshift, e2_subspace, sumc2_subspace, e2all, sumc2, weight0, &
                                         total)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: selectroot
        real(8), intent(in) :: eshift
        real(8), intent(out) :: e2_subspace(8), sumc2_subspace(8), e2all, sumc2, weight0, total
        real(8) :: scal(6)
        call dcaspt2_amd_check(dc2_caspt2(ctx, int(selectroot, c_int32_t), eshift, e2_subspace, sumc2_subspace, scal), &
                               'r4dcaspt2_tra')
        e2all = scal(2)
        sumc2 = scal(3)
        weight0 = scal(4)
        total = scal(6)
    end subroutine
end module dcaspt2_amd
