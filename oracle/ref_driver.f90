!
! Oracle driver for the reference GIMIC library (test infrastructure only).
!
! This program is this project's own code. It is compiled with flang against
! the reference's libgimic modules (built from the reference sources in place
! by oracle/Makefile) and calls the reference's per-point current-density
! routine exactly the way the reference field loop does:
!   new_basis/new_dens/read_dens  (src/fgimic/gimic.F90:141-194)
!   new_jtensor + ctensor per point, one jtensor_t per OpenMP thread
!                                 (src/fgimic/jfield.f90:114-129)
!
! Control input (stdin, one item per line):
!   molfile            MOL/INTGRL basis file
!   densmode           'xdens' (text XDENS via read_dens, dens.f90:56-135)
!                      or 'bin' (raw column-major doubles, set via set_dens /
!                      set_pdens; no UHF scaling and no TURBOMOLE reorder)
!   densfile
!   uhf                0/1
!   screening          threshold; <= 0 switches screening off
!   giao diamag paramag  0/1 each
!   spin               total | alpha | beta | spindens
!   pointsfile         stream: int32 npts, then 3*npts doubles (bohr)
!   outfile            stream: 9*npts doubles, ct(m,b) column-major per point
! Prints 'loop_seconds <t> threads <n> points <npts>' on stdout.
!
program ref_driver
    use omp_lib
    use globals_module
    use settings_module
    use teletype_module
    use basis_class
    use dens_class
    use jtensor_class
    implicit none

    character(256) :: molfile, densmode, densfile, pointsfile, outfile
    character(16) :: spin
    character(80) :: xdfile
    integer :: uhf, giao, dia, para, npts, i, n, nspin, s, b, nthr
    real(DP) :: thrs, t0, t1
    real(DP), allocatable :: pts(:,:), out(:,:), mat(:,:)
    type(molecule_t), save :: mol
    type(dens_t), save :: xdens
    type(jtensor_t) :: jt

    read(*,'(a)') molfile
    read(*,'(a)') densmode
    read(*,'(a)') densfile
    read(*,*) uhf
    read(*,*) thrs
    read(*,*) giao, dia, para
    read(*,'(a)') spin
    read(*,'(a)') pointsfile
    read(*,'(a)') outfile

    call set_debug_level(0)
    settings%use_spherical = .false.
    settings%is_uhf = (uhf /= 0)
    settings%use_giao = (giao /= 0)
    settings%use_diamag = (dia /= 0)
    settings%use_paramag = (para /= 0)

    call new_basis(mol, trim(molfile), thrs)
    call new_dens(xdens, mol)
    n = get_ncgto(mol)
    if (trim(densmode) == 'xdens') then
        xdfile = trim(densfile)
        call read_dens(xdens, xdfile)
    else
        nspin = 1
        if (settings%is_uhf) nspin = 2
        allocate(mat(n,n))
        open(31, file=trim(densfile), access='stream', form='unformatted', &
             status='old')
        do s = 1, nspin
            read(31) mat
            call set_dens(xdens, mat, s)
            do b = 1, 3
                read(31) mat
                call set_pdens(xdens, b, mat, s)
            end do
        end do
        close(31)
        deallocate(mat)
    end if

    open(32, file=trim(pointsfile), access='stream', form='unformatted', &
         status='old')
    read(32) npts
    allocate(pts(3,npts), out(9,npts))
    read(32) pts
    close(32)

    t0 = omp_get_wtime()
    nthr = 1
!$omp parallel default(shared) private(jt, i)
!$omp single
    nthr = omp_get_num_threads()
!$omp end single
    call new_jtensor(jt, mol, xdens)
!$omp do schedule(dynamic, 16)
    do i = 1, npts
        call ctensor(jt, pts(:,i), out(:,i), trim(spin))
    end do
!$omp end do
    call del_jtensor(jt)
!$omp end parallel
    t1 = omp_get_wtime()

    open(33, file=trim(outfile), access='stream', form='unformatted', &
         status='replace')
    write(33) out
    close(33)
    write(*,'(a,es16.8,a,i6,a,i10,a,i8)') 'loop_seconds ', t1-t0, &
        ' threads ', nthr, ' points ', npts, ' nbf ', n
end program
