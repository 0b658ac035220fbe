! Same-name replacements of the reference's two subprogram entry points, for linking into the
! reference program in place of its r4dcasci.F90 and r4dcaspt2_tra.F90:
!
!   subroutine r4dcasci        (r4dcasci.F90:3-134)
!   subroutine r4dcaspt2_tra   (r4dcaspt2_tra.F90:4-482)
!
! Both take no arguments and work on the module_global_variables state, as the originals do,
! so the driver (dcaspt2_run_subprograms.F90:57,62), the input parser and the MRCONEE/MDCINT
! readers stay untouched.  The work runs on the GPU through the C ABI (module dcaspt2_amd):
!
!   r4dcasci       the reference's own create_newmdcint writes MDCINTNEW (untouched); its
!                  records are read here and handed to dc2_set_integral_records, which fills
!                  every integral block on the device (module_2integrals.F90:395-434 and
!                  divide_2_elec_integral_into_subspaces.F90 in the original); dc2_casci then
!                  stands for casci (casci.F90:3-160): eigen, cir/cii, dict_cas_idx(_reverse)
!                  are filled and CIDATA_sym<totsym> is written in casci.F90's layout.
!   r4dcaspt2_tra  the reference's read_cidata fills cir/cii/eigen/dict_cas_idx (untouched);
!                  when the library context does not hold this symmetry's CASCI (another
!                  process, or another symmetry since), the integrals are loaded again and
!                  dc2_set_casci_result hands the file's CASCI result over.  dc2_caspt2
!                  (dc2_caspt2_restart after read_and_validate_restart_file) computes
!                  e2_subspace, sumc2_subspace, e2all and sumc2, printed as the original does.
!
! A failed library call ends the program through stop_with_errorcode(status)
! (module_error.F90:15-39).  One library context per process (device 0, or the value of the
! environment variable DCASPT2_AMD_DEVICE).
module dcaspt2_amd_dropin
    use, intrinsic :: iso_c_binding
    use, intrinsic :: iso_fortran_env, only: int32, int64
    use dcaspt2_amd
    implicit none
    private
    public :: dropin_context, dropin_load_problem, dropin_check, dropin_realonly, dropin_ctx_totsym

    type(c_ptr), save :: ctx = c_null_ptr
    integer, save :: dropin_ctx_totsym = -1        ! total symmetry whose CASCI the context holds
    logical, save :: dropin_realonly = .false.     ! MDCINTNEW holds real parts only

contains

    subroutine dropin_check(status, what)
        use module_error, only: stop_with_errorcode
        integer(c_int), intent(in) :: status
        character(*), intent(in) :: what
        if (status == 0) return
        print '(a,a,a,i0,a,a)', ' dcaspt2_amd: ', what, ' failed with status ', status, ': ', dcaspt2_amd_last_error()
        call stop_with_errorcode(int(status, int32))
    end subroutine

    function dropin_context() result(c)
        type(c_ptr) :: c
        character(len=16) :: env
        integer :: dev, st
        if (.not. c_associated(ctx)) then
            dev = 0
            call get_environment_variable("DCASPT2_AMD_DEVICE", env, status=st)
            if (st == 0) read (env, *) dev
            call dropin_check(dc2_context_create(int(dev, c_int32_t), ctx), 'dc2_context_create')
        end if
        c = ctx
    end function

    ! spaces, symmetry, RAS limits, one-electron integrals and every two-electron block of the
    ! current module_global_variables state, the latter from the MDCINTNEW records
    subroutine dropin_load_problem()
        use module_global_variables
        type(c_ptr) :: c
        integer(c_int64_t), allocatable :: mult(:), irp(:)
        complex(c_double_complex), allocatable :: h(:, :)
        integer :: n2, i, j
        c = dropin_context()
        call dropin_check(dc2_set_spaces(c, int(ninact, c_int32_t), int(nact, c_int32_t), int(nsec, c_int32_t), &
                                         int(nelec, c_int32_t)), 'dc2_set_spaces')
        if (nsymrpa > 1) then
            n2 = int(2*nsymrpa)
            allocate (mult(n2*n2), irp(nmo))
            do j = 1, n2
                do i = 1, n2
                    mult(i + n2*(j - 1)) = multb(i, j)
                end do
            end do
            irp(:) = irpamo(1:nmo)
            call dropin_check(dc2_set_symmetry(c, int(nsymrpa, c_int32_t), mult, irp, int(totsym, c_int32_t)), &
                              'dc2_set_symmetry')
        end if
        if (ras1_size /= 0 .or. ras3_size /= 0) then
            call dropin_check(dc2_set_ras(c, int(ras1_size, c_int32_t), int(ras2_size, c_int32_t), &
                                          int(ras3_size, c_int32_t), int(ras1_max_hole, c_int32_t), &
                                          int(ras3_max_elec, c_int32_t), int(min_hole_ras1, c_int32_t), &
                                          int(min_elec_ras3, c_int32_t)), 'dc2_set_ras')
        end if
        allocate (h(nmo, nmo))
        h(:, :) = cmplx(one_elec_int_r(1:nmo, 1:nmo), one_elec_int_i(1:nmo, 1:nmo), kind=c_double_complex)
        call dropin_check(dc2_set_one_electron(c, h, ecore), 'dc2_set_one_electron')
        call dropin_load_records(mdcintnew)
    end subroutine

    ! MDCINTNEW (create_newmdcint.F90:56,173-176,198): a header record, then one integral per
    ! record (i, j, nnz, k, l, re[, -im]) with int64 CAS-order labels, ended by (0, 0, 0).
    ! A record too short for the imaginary part marks a real-only file, as module_realonly
    ! decides from the record length.
    subroutine dropin_load_records(fname)
        character(*), intent(in) :: fname
        integer :: unit, ios
        integer(int64) :: i, j, nnz, k, l, n, cap
        real(8) :: re, im
        integer(c_int32_t), allocatable :: ijkl(:), t32(:)
        complex(c_double_complex), allocatable :: vals(:), tv(:)
        open (newunit=unit, file=fname, form='unformatted', status='old', action='read')
        read (unit)
        cap = 1024
        allocate (ijkl(4*cap), vals(cap))
        n = 0
        im = 0.0d0
        read (unit, iostat=ios) i, j, nnz, k, l, re, im
        if (ios /= 0) then      ! no imaginary part (or only the terminator)
            backspace (unit)
            read (unit, iostat=ios) i, j, nnz, k, l, re
            dropin_realonly = .true.
            im = 0.0d0
        else
            dropin_realonly = .false.
        end if
        ! the three-integer terminator is too short for a data record
        do while (ios == 0 .and. i /= 0)
            if (n == cap) then
                allocate (t32(8*cap), tv(2*cap))
                t32(1:4*cap) = ijkl
                tv(1:cap) = vals
                call move_alloc(t32, ijkl)
                call move_alloc(tv, vals)
                cap = 2*cap
            end if
            n = n + 1
            ijkl(4*n - 3) = int(i, c_int32_t)
            ijkl(4*n - 2) = int(j, c_int32_t)
            ijkl(4*n - 1) = int(k, c_int32_t)
            ijkl(4*n) = int(l, c_int32_t)
            vals(n) = cmplx(re, im, kind=c_double_complex)
            if (dropin_realonly) then
                read (unit, iostat=ios) i, j, nnz, k, l, re
            else
                read (unit, iostat=ios) i, j, nnz, k, l, re, im
            end if
        end do
        close (unit)
        call dropin_check(dc2_set_integral_records(dropin_context(), int(n, c_int64_t), ijkl, vals, 0_c_int32_t), &
                          'dc2_set_integral_records')
    end subroutine
end module dcaspt2_amd_dropin

! r4dcasci.F90:3-134: CASCI of the current symmetry on the GPU
subroutine r4dcasci
    use, intrinsic :: iso_c_binding
    use, intrinsic :: iso_fortran_env, only: int64
    use module_dict, only: add
    use module_global_variables
    use dcaspt2_amd
    use dcaspt2_amd_dropin
    implicit none
    type(c_ptr) :: c
    integer(c_int64_t) :: nd
    integer(c_int64_t), allocatable :: dets(:)
    complex(c_double_complex), allocatable :: v(:)
    real(8), allocatable :: ecas(:)
    integer :: k, unit_cidata
    integer(int64) :: i
    character(len=cidata_key_size) :: key
    character(len=len_convert_int_to_chr) :: chr_totsym

    if (rank == 0) then
        print *, ''
        print *, ' ENTER RELATIVISTIC CASCI PROGRAM (MI355X)'
        print *, ''
    end if
    call create_newmdcint
    call get_mdcint_filename(0)
    call dropin_load_problem()
    c = dropin_context()
    if (allocated(eigen)) deallocate (eigen)
    allocate (eigen(nroot))
    call dropin_check(dc2_casci(c, int(nroot, c_int32_t), eigen, nd), 'casci')
    ndet = int(nd)
    dropin_ctx_totsym = totsym
    iroot = selectroot
    allocate (dets(nd), v(nd), ecas(nroot))
    call dropin_check(dc2_get_dets(c, dets), 'dc2_get_dets')
    if (allocated(cir)) deallocate (cir)
    if (allocated(cii)) deallocate (cii)
    allocate (cir(nd, nroot), cii(nd, nroot))
    do k = 1, nroot
        call dropin_check(dc2_get_ci_vector(c, int(k, c_int32_t), v), 'dc2_get_ci_vector')
        cir(:, k) = real(v)
        cii(:, k) = aimag(v)
    end do
    do i = 1, nd
        call add(dict_cas_idx, i, dets(i))
        call add(dict_cas_idx_reverse, dets(i), i)
    end do
    ecas(:) = eigen(:) - ecore
    if (rank == 0) then
        do k = 1, nroot
            print '(" CASCI ENERGY FOR ",I0," STATE",I4,F30.15)', totsym, k, eigen(k)
        end do
        ! CIDATA_sym<totsym> in the layout of casci.F90:118-160
        write (chr_totsym, *) totsym
        open (newunit=unit_cidata, file='CIDATA_sym'//trim(adjustl(chr_totsym)), form='unformatted', &
              status='replace')
        key = 'ninact'; write (unit_cidata) key; write (unit_cidata) ninact
        key = 'nact'; write (unit_cidata) key; write (unit_cidata) nact
        key = 'nsec'; write (unit_cidata) key; write (unit_cidata) nsec
        key = 'nelec'; write (unit_cidata) key; write (unit_cidata) nelec
        key = 'ndet'; write (unit_cidata) key; write (unit_cidata) ndet
        key = 'nroot'; write (unit_cidata) key; write (unit_cidata) nroot
        key = 'totsym'; write (unit_cidata) key; write (unit_cidata) totsym
        key = 'ecas'; write (unit_cidata) key; write (unit_cidata) ecas(1:nroot)
        key = 'dict_cas_idx_values'; write (unit_cidata) key; write (unit_cidata) dets(1:nd)
        key = 'ci_coefficients'; write (unit_cidata) key
        write (unit_cidata) (cir(:, k), k=1, nroot)
        if (.not. dropin_realonly) write (unit_cidata) (cii(:, k), k=1, nroot)
        key = 'end'; write (unit_cidata) key
        close (unit_cidata)
        print *, 'End r4dcasci part'
    end if
    ! the original leaves these to be read back from CIDATA by r4dcaspt2_tra
    deallocate (cir, cii, eigen)
end subroutine r4dcasci

! r4dcaspt2_tra.F90:4-482: CASPT2 of root selectroot on the GPU
subroutine r4dcaspt2_tra
    use, intrinsic :: iso_c_binding
    use module_global_variables
    use dcaspt2_amd
    use dcaspt2_amd_dropin
    use module_dict, only: get_val
    use, intrinsic :: iso_fortran_env, only: int64
    implicit none
    type(c_ptr) :: c
    real(8) :: scal(6), e2f(8), s2f(8), weight0
    integer(c_int64_t), allocatable :: dets(:)
    complex(c_double_complex), allocatable :: coeffs(:, :)
    real(8), allocatable :: ecas(:)
    integer(int64) :: i
    integer :: k, nr, first
    character(len=1), parameter :: names(8) = ['a', 'b', 'c', 'd', 'e', 'f', 'g', 'h']

    if (rank == 0) then
        print *, ''
        print *, ' START RELATIVISIC CASPT2 PROGRAM (MI355X)'
        print *, ''
    end if
    if (ninact == 0 .and. nsec == 0) then
        if (rank == 0) print *, "The CASPT2 energy cannot be defined when ninact = 0 and nsec = 0."
        stop
    end if
    call get_mdcint_filename(0)
    call read_cidata
    iroot = selectroot
    c = dropin_context()
    if (dropin_ctx_totsym /= totsym) then
        call dropin_load_problem()
        nr = size(eigen)
        allocate (dets(ndet), coeffs(ndet, nr), ecas(nr))
        do i = 1, ndet
            dets(i) = get_val(dict_cas_idx, i)
        end do
        coeffs(:, :) = cmplx(cir(1:ndet, 1:nr), cii(1:ndet, 1:nr), kind=c_double_complex)
        ecas(:) = eigen(1:nr) - ecore
        call dropin_check(dc2_set_casci_result(c, int(ndet, c_int64_t), int(nr, c_int32_t), dets, coeffs, ecas), &
                          'dc2_set_casci_result')
        dropin_ctx_totsym = totsym
    end if
    if (enable_restart) then
        first = index('ABCDEFGH', next_subspace) - 1
        if (first < 0) first = 8
        e2f(:) = e2_subspace(:)
        s2f(:) = sumc2_subspace(:)
        call dropin_check(dc2_caspt2_restart(c, int(selectroot, c_int32_t), eshift, int(first, c_int32_t), e2f, s2f, &
                                             e2_subspace, sumc2_subspace, scal), 'r4dcaspt2_tra')
    else
        call dropin_check(dc2_caspt2(c, int(selectroot, c_int32_t), eshift, e2_subspace, sumc2_subspace, scal), &
                          'r4dcaspt2_tra')
    end if
    e2all = scal(2)
    sumc2 = scal(3)
    weight0 = scal(4)
    if (rank == 0) then
        do k = 1, 8
            print '(" e2",a1,"      = ",E25.15," a.u.")', names(k), e2_subspace(k)
            print '(" sumc2,",a1,"  = ",E25.15)', names(k), sumc2_subspace(k)
        end do
        print '(" CASPT2 ENERGY FOR ",I0," STATE, selectroot = ",I0)', totsym, selectroot
        print '(" c^2 is                         ",F30.20)', sumc2
        print '(" weight of 0th wave function is ",F30.20)', weight0
        print '(" Total second order energy is   ",F30.20," a.u.")', e2all - eshift*sumc2
        print '(" Total energy is                ",F30.20," a.u.")', e2all + eigen(iroot) - eshift*sumc2
        print *, ' '
        print *, 'END OF RELATIVISTIC CASPT2 PROGRAM'
    end if
    if (allocated(cir)) deallocate (cir)
    if (allocated(cii)) deallocate (cii)
    if (allocated(eigen)) deallocate (eigen)
end subroutine r4dcaspt2_tra
