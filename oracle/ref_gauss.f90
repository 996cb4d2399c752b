!
! Oracle driver for the reference quadrature (test infrastructure only).
! This project's own program; it calls the reference's setup_gauss_data
! (src/libgimic/gaussint.f90:267-319) and prints the composite nodes and
! weights at full precision.
!
! stdin: a b order npts quadr      (quadr = gauss | lobatto)
! stdout: npts lines 'pt wgt' in es25.17
!
program ref_gauss
    use kinds_module
    use gaussint_module
    implicit none

    real(DP) :: a, b
    integer :: order, npts, i
    character(16) :: quadr
    type(gdata_t) :: g

    read(*,*) a, b, order, npts, quadr
    allocate(g%pts(npts), g%wgt(npts))
    call setup_gauss_data(a, b, order, g, trim(quadr))
    do i = 1, npts
        write(*,'(2es25.17)') g%pts(i), g%wgt(i)
    end do
end program
