! A Fortran host driving the HIP path through the iso_c_binding module.
!
!   drive_caspt2 version           print the library version (no GPU needed)
!   drive_caspt2 badarg            a call on a null context: prints the status
!   drive_caspt2 run <input>       CASCI + CASPT2 of a factorised-integral problem
!
! <input> is a stream file: int32 ninact, nact, nsec, nelec, selectroot, nroot;
! int64 naux; real(8) ecore, eshift; complex(8) h(nmo,nmo); complex(8) B(naux,nmo,nmo)
! with (pq|rs) = sum_P B(P,p,q) B(P,r,s).  Output lines: "casci <e>", "e2 <X> <e2> <sumc2>",
! "total <e>", all with 17 significant digits.
program drive_caspt2
    use dcaspt2_amd
    use, intrinsic :: iso_c_binding
    implicit none
    character(len=256) :: mode, path
    type(c_ptr) :: ctx
    character(kind=c_char), pointer :: s(:)
    integer(c_int32_t) :: ni, na, ns, ne, root, nroot
    integer(c_int64_t) :: naux, ndet
    real(8) :: ecore, eshift, e2(8), sumc2(8), e2all, s2, w0, total
    real(8), allocatable :: eig(:)
    complex(8), allocatable :: h(:, :), b(:, :, :)
    integer :: nmo, u, k, st
    character(len=1), parameter :: names(8) = ['A', 'B', 'C', 'D', 'E', 'F', 'G', 'H']

    call get_command_argument(1, mode)
    if (trim(mode) == 'version') then
        call c_f_pointer(dc2_version(), s, [64])
        k = 0
        do while (s(k + 1) /= c_null_char)
            k = k + 1
        end do
        write (*, '(a,64a1)') 'version ', s(1:k)
        stop
    end if
    if (trim(mode) == 'badarg') then
        st = dc2_set_spaces(c_null_ptr, 1, 2, 3, 1)
        write (*, '(a,i0,a,a)') 'status ', st, ' ', dcaspt2_amd_last_error()
        stop
    end if
    call get_command_argument(2, path)
    open (newunit=u, file=trim(path), access='stream', form='unformatted', status='old')
    read (u) ni, na, ns, ne, root, nroot
    read (u) naux
    read (u) ecore, eshift
    nmo = ni + na + ns
    allocate (h(nmo, nmo), b(naux, nmo, nmo), eig(nroot))
    read (u) h
    read (u) b
    close (u)
    call dcaspt2_amd_open(ctx, 0, int(ni), int(na), int(ns), int(ne))
    call dcaspt2_amd_check(dc2_set_one_electron(ctx, h, ecore), 'dc2_set_one_electron')
    call dcaspt2_amd_check(dc2_set_integral_factors(ctx, b, naux, 0_c_int32_t), 'dc2_set_integral_factors')
    call dcaspt2_amd_casci(ctx, int(nroot), eig, ndet)
    do k = 1, nroot
        write (*, '(a,es25.17)') 'casci ', eig(k)
    end do
    call dcaspt2_amd_r4dcaspt2_tra(ctx, int(root), eshift, e2, sumc2, e2all, s2, w0, total)
    do k = 1, 8
        write (*, '(a,a1,2es25.17)') 'e2 ', names(k), e2(k), sumc2(k)
    end do
    write (*, '(a,es25.17)') 'total ', total
    call dcaspt2_amd_close(ctx)
end program drive_caspt2
