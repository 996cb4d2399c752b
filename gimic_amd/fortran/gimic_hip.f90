!
! iso_c_binding interface to libgimic_hip.so for Fortran callers such as the
! fgimic driver.
!
! Part 1 declares the reference's C entry points (gimic_interface.h:9-17,
! gausspoints.h:10-11) exactly as the reference's own bind(c) routines take
! them: every argument by reference.  Part 2 declares the batched handle API
! (include/gimic_hip.h).  hip_calc_jtensors is the drop-in for the per-point
! loop of fgimic's calc_jtensors (src/fgimic/jfield.f90:114-129): one call
! for all grid points, tensors stored as the reference's tens(9, npts).
!
module gimic_hip
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: GIMIC_SPIN_TOTAL = 0, GIMIC_SPIN_ALPHA = 1
    integer(c_int), parameter :: GIMIC_SPIN_BETA = 2, GIMIC_SPIN_SPINDENS = 3
    integer(c_int), parameter :: GIMIC_PREC_FP64 = 0, GIMIC_PREC_MIXED = 1

    interface
        ! ---- Part 1: the reference's C interface ----
        subroutine gimic_init(molfile, densfile) bind(c, name='gimic_init')
            import :: c_char
            character(kind=c_char), dimension(*), intent(in) :: molfile, densfile
        end subroutine
        subroutine gimic_finalize() bind(c, name='gimic_finalize')
        end subroutine
        subroutine gimic_set_uhf(uhf) bind(c, name='gimic_set_uhf')
            import :: c_int
            integer(c_int), intent(in) :: uhf
        end subroutine
        subroutine gimic_set_magnet(b) bind(c, name='gimic_set_magnet')
            import :: c_double
            real(c_double), dimension(3), intent(in) :: b
        end subroutine
        subroutine gimic_set_spin(s) bind(c, name='gimic_set_spin')
            import :: c_char
            character(kind=c_char), dimension(*), intent(in) :: s
        end subroutine
        subroutine gimic_set_screening(thrs) bind(c, name='gimic_set_screening')
            import :: c_double
            real(c_double), intent(in) :: thrs
        end subroutine
        subroutine gimic_calc_jtensor(r, jt) bind(c, name='gimic_calc_jtensor')
            import :: c_double
            real(c_double), dimension(3), intent(in) :: r
            real(c_double), dimension(9), intent(out) :: jt
        end subroutine
        subroutine gimic_calc_jvector(r, jv) bind(c, name='gimic_calc_jvector')
            import :: c_double
            real(c_double), dimension(3), intent(in) :: r
            real(c_double), dimension(3), intent(out) :: jv
        end subroutine
        subroutine gimic_calc_modj(r, mj) bind(c, name='gimic_calc_modj')
            import :: c_double
            real(c_double), dimension(3), intent(in) :: r
            real(c_double), intent(out) :: mj
        end subroutine
        subroutine gimic_get_gauss_points(a, b, npts, order, pts, wgts) &
                bind(c, name='gimic_get_gauss_points')
            import :: c_double, c_int
            real(c_double), intent(in) :: a, b
            integer(c_int), intent(in) :: npts, order
            real(c_double), dimension(*), intent(out) :: pts, wgts
        end subroutine

        ! ---- Part 2: batched handle API ----
        function gimic_last_error() bind(c, name='gimic_last_error') result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function
        function gimic_ctx_create_files(ctx, molfile, densfile, uhf, thrs) &
                bind(c, name='gimic_ctx_create_files') result(rc)
            import :: c_ptr, c_char, c_int, c_double
            type(c_ptr), intent(out) :: ctx
            character(kind=c_char), dimension(*), intent(in) :: molfile, densfile
            integer(c_int), value :: uhf
            real(c_double), value :: thrs
            integer(c_int) :: rc
        end function
        function gimic_ctx_set_flags(ctx, giao, diamag, paramag) &
                bind(c, name='gimic_ctx_set_flags') result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: giao, diamag, paramag
            integer(c_int) :: rc
        end function
        function gimic_jtensor_batch(ctx, npts, r, jt, spin, precision) &
                bind(c, name='gimic_jtensor_batch') result(rc)
            import :: c_ptr, c_int, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: npts
            real(c_double), dimension(3, *), intent(in) :: r
            real(c_double), dimension(9, *), intent(out) :: jt
            integer(c_int), value :: spin, precision
            integer(c_int) :: rc
        end function
        function gimic_flux_planes(ctx, nplanes, offs, r, w, normals, bfields, centers, &
                radii, spin, precision, out, out_mod, out_acid, jt) &
                bind(c, name='gimic_flux_planes') result(rc)
            import :: c_ptr, c_int, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nplanes
            integer(c_int64_t), dimension(*), intent(in) :: offs
            real(c_double), dimension(3, *), intent(in) :: r
            real(c_double), dimension(*), intent(in) :: w
            real(c_double), dimension(3, *), intent(in) :: normals, bfields, centers
            real(c_double), dimension(*), intent(in) :: radii
            integer(c_int), value :: spin, precision
            real(c_double), dimension(3, *), intent(out) :: out
            type(c_ptr), value :: out_mod, out_acid, jt
            integer(c_int) :: rc
        end function
        ! the same sums over product grids generated on the device: origin,
        ! axis vectors (basv columns, grid.f90) and 1-D rules per plane
        function gimic_flux_plane_grids(ctx, nplanes, n1, n2, origins, axes, p3w3, &
                nodes1, wts1, nodes2, wts2, normals, bfields, centers, radii, spin, &
                precision, out, out_mod, out_acid) &
                bind(c, name='gimic_flux_plane_grids') result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nplanes
            integer(c_int), dimension(*), intent(in) :: n1, n2
            real(c_double), dimension(3, *), intent(in) :: origins
            real(c_double), dimension(9, *), intent(in) :: axes
            real(c_double), dimension(2, *), intent(in) :: p3w3
            real(c_double), dimension(*), intent(in) :: nodes1, wts1, nodes2, wts2
            real(c_double), dimension(3, *), intent(in) :: normals, bfields, centers
            real(c_double), dimension(*), intent(in) :: radii
            integer(c_int), value :: spin, precision
            real(c_double), dimension(3, *), intent(out) :: out
            type(c_ptr), value :: out_mod, out_acid
            integer(c_int) :: rc
        end function
        ! get_property (jfield.f90:584-929): shielding sums per nucleus and atom
        ! block, magnetizability sums per atom block, 5 values each
        function gimic_ctx_properties(ctx, npts, r, w, npart, part_offs, nnuc, nuclei, &
                spin, precision, shield, magnet, jt) &
                bind(c, name='gimic_ctx_properties') result(rc)
            import :: c_ptr, c_int, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: npts
            real(c_double), dimension(3, *), intent(in) :: r
            real(c_double), dimension(*), intent(in) :: w
            integer(c_int), value :: npart
            integer(c_int64_t), dimension(*), intent(in) :: part_offs
            integer(c_int), value :: nnuc
            real(c_double), dimension(3, *), intent(in) :: nuclei
            integer(c_int), value :: spin, precision
            real(c_double), dimension(5, *), intent(out) :: shield
            real(c_double), dimension(5, *), intent(out) :: magnet
            type(c_ptr), value :: jt
            integer(c_int) :: rc
        end function
        function gimic_ctx_destroy(ctx) bind(c, name='gimic_ctx_destroy') result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
    end interface

contains

    ! NUL-terminated copy of a Fortran string
    function c_string(s) result(cs)
        character(*), intent(in) :: s
        character(kind=c_char), dimension(len_trim(s) + 1) :: cs
        integer :: i
        do i = 1, len_trim(s)
            cs(i) = s(i:i)
        end do
        cs(len_trim(s) + 1) = c_null_char
    end function

    function spin_code(spin) result(k)
        character(*), intent(in) :: spin
        integer(c_int) :: k
        select case (trim(spin))
            case ('alpha')
                k = GIMIC_SPIN_ALPHA
            case ('beta')
                k = GIMIC_SPIN_BETA
            case ('spindens')
                k = GIMIC_SPIN_SPINDENS
            case default
                k = GIMIC_SPIN_TOTAL
        end select
    end function

    ! tens(:, n) = J tensor at pts(:, n), as calc_jtensors stores it
    ! (jfield.f90:120-125), for all points in one batched call
    subroutine hip_calc_jtensors(ctx, pts, tens, spin, precision, ierr)
        type(c_ptr), intent(in) :: ctx
        real(c_double), dimension(:, :), intent(in) :: pts
        real(c_double), dimension(:, :), intent(out) :: tens
        character(*), intent(in) :: spin
        integer(c_int), intent(in) :: precision
        integer, intent(out) :: ierr
        ierr = gimic_jtensor_batch(ctx, int(size(pts, 2), c_int64_t), pts, tens, &
                                   spin_code(spin), precision)
    end subroutine

end module
