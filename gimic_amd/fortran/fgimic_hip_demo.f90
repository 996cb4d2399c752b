!
! Fortran caller of libgimic_hip.so through the gimic_hip module: the
! reference's per-point C interface and the batched field loop.
!
!   fgimic_hip_demo MOL XDENS points.txt out.txt
!
! points.txt: npts, then npts lines x y z (bohr).  out.txt: per point the
! nine tensor elements from the batched call, then the nine from
! gimic_calc_jtensor (identical up to rounding).
!
program fgimic_hip_demo
    use iso_c_binding
    use gimic_hip
    implicit none

    character(512) :: molfile, densfile, ptsfile, outfile
    type(c_ptr) :: ctx
    integer :: npts, i, ierr
    real(c_double), allocatable :: pts(:, :), tens(:, :)
    real(c_double) :: jt(9)

    call get_command_argument(1, molfile)
    call get_command_argument(2, densfile)
    call get_command_argument(3, ptsfile)
    call get_command_argument(4, outfile)

    open(11, file=trim(ptsfile), status='old')
    read(11, *) npts
    allocate(pts(3, npts), tens(9, npts))
    do i = 1, npts
        read(11, *) pts(:, i)
    end do
    close(11)

    ! batched handle API (screening 1e-6, the C interface's threshold)
    ierr = gimic_ctx_create_files(ctx, c_string(molfile), c_string(densfile), 0_c_int, &
                                  1.0d-6)
    if (ierr /= 0) stop 'gimic_ctx_create_files failed'
    call hip_calc_jtensors(ctx, pts, tens, 'total', GIMIC_PREC_FP64, ierr)
    if (ierr /= 0) stop 'gimic_jtensor_batch failed'
    ierr = gimic_ctx_destroy(ctx)

    ! the reference's C interface, one point at a time
    call gimic_init(c_string(molfile), c_string(densfile))
    call gimic_set_spin(c_string('total'))
    open(12, file=trim(outfile), status='replace')
    do i = 1, npts
        call gimic_calc_jtensor(pts(:, i), jt)
        write(12, '(18es25.16)') tens(:, i), jt
    end do
    close(12)
    call gimic_finalize()
end program
