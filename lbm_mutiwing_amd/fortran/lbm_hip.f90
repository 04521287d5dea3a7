! lbm_hip.f90 - iso_c_binding interface of liblbm_hip.so (include/lbm_hip.h).
!
! The reference's step loop calls Fortran subroutines that share module
! globals (main.f90:4-159, solution.f90:272).  A host program that uses this
! module instead holds one opaque solver handle and calls the C ABI, in the
! same bind(C) style as the reference's libzmq binding (zmq_comm.f90:23-81).
! Each interface below names the C entry point it binds; include/lbm_hip.h
! documents which reference routine each one replaces.
module lbm_hip
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: LBM_MAX_WINGS = 8
    integer(c_int), parameter :: LBM_MOTION_COLS_PER_WING = 9
    integer(c_int), parameter :: LBM_MAX_MOTION_COLS = 1 + 9 * LBM_MAX_WINGS
    integer(c_int), parameter :: LBM_MAX_FORCE_ROW = 4 + 12 * LBM_MAX_WINGS
    integer(c_int), parameter :: LBM_MOTION_FILE = 1, LBM_MOTION_STREAM = 2
    integer(c_int), parameter :: LBM_PRECISION_FP64 = 0, LBM_PRECISION_FP32_SHIFTED = 1
    integer(c_int), parameter :: LBM_OK = 0, LBM_E_INVALID = 1, LBM_E_STATE = 2, &
                                 LBM_E_DEVICE = 3, LBM_E_RANGE = 4, LBM_E_NUMERIC = 5

    ! struct lbm_config; same member order and types as the C declaration
    type, bind(C) :: lbm_config
        integer(c_int32_t) :: nx, ny, wing_count, points_per_wing
        real(c_double) :: scale, ox, oy, uu, re
        real(c_double) :: s(9)
        real(c_double) :: wing_x0, wing_dx, wing_y0, xc, yc, xp, yp
        integer(c_int32_t) :: motion_source, precision
        type(c_ptr) :: xb0, yb0
    end type

    ! rigid-body coupling of wing 1 (include/lbm_hip.h, lbm_dynamics)
    type, bind(C) :: lbm_dynamics
        integer(c_int32_t) :: enable, closed_boundary, rk4, reserved
        real(c_double) :: solid_den, area, par1, par2, alphad
    end type

    interface
        integer(c_int) function lbm_create(cfg, h) bind(C, name='lbm_create')
            import :: c_int, c_ptr, lbm_config
            type(lbm_config), intent(in) :: cfg
            type(c_ptr), intent(out) :: h
        end function

        integer(c_int) function lbm_destroy(h) bind(C, name='lbm_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function lbm_reset(h) bind(C, name='lbm_reset')
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function lbm_set_motion_table(h, rows, nrows, ncols, end_time) &
                bind(C, name='lbm_set_motion_table')
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: rows(*)
            integer(c_int32_t), value :: nrows, ncols
            real(c_double), intent(out) :: end_time
        end function

        integer(c_int) function lbm_set_stream_segment(h, start_row, target_row, ncols) &
                bind(C, name='lbm_set_stream_segment')
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: start_row(*), target_row(*)
            integer(c_int32_t), value :: ncols
        end function

        integer(c_int) function lbm_build_stream_target_row(wing_count, start_row, target_time, &
                delta_row, target_row) bind(C, name='lbm_build_stream_target_row')
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: wing_count
            real(c_double), intent(in) :: start_row(*), delta_row(*)
            real(c_double), value :: target_time
            real(c_double), intent(out) :: target_row(*)
        end function

        ! force_rows / motion_rows: c_null_ptr, or c_loc of a (row_len, steps)
        ! / (1+3W, steps) array, one column per step
        integer(c_int) function lbm_advance(h, n_last, force_rows, motion_rows, row_sum, last_row, err) &
                bind(C, name='lbm_advance')
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: n_last
            type(c_ptr), value :: force_rows, motion_rows
            real(c_double), intent(out) :: row_sum(*), last_row(*), err
        end function

        integer(c_int) function lbm_get_state(h, n_next, movenum, flag2, row_len, adjoint_n) &
                bind(C, name='lbm_get_state')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: n_next, movenum, flag2, row_len, adjoint_n
        end function

        integer(c_int) function lbm_get_fields(h, rho, u, v, pre) bind(C, name='lbm_get_fields')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: rho(*), u(*), v(*), pre(*)
        end function

        integer(c_int) function lbm_get_mass(h, mass) bind(C, name='lbm_get_mass')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: mass
        end function

        ! allocation-free variant: f and mass are device addresses, stream a hipStream_t
        integer(c_int) function lbm_get_mass_noalloc(f, nx, ny, precision, mass, stream) &
                bind(C, name='lbm_get_mass_noalloc')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: f, mass, stream
            integer(c_int32_t), value :: nx, ny, precision
        end function

        integer(c_int) function lbm_get_populations(h, f, f0) bind(C, name='lbm_get_populations')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: f(*), f0(*)
        end function

        integer(c_int) function lbm_set_populations(h, f, f0) bind(C, name='lbm_set_populations')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: f(*), f0(*)
        end function

        integer(c_int) function lbm_get_markers(h, xb1, yb1, xb2, yb2) bind(C, name='lbm_get_markers')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: xb1(*), yb1(*), xb2(*), yb2(*)
        end function

        integer(c_int) function lbm_get_ibm(h, capacity, adjoint_ij, wing_id, wing_force, wing_moment) &
                bind(C, name='lbm_get_ibm')
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: capacity
            integer(c_int32_t), intent(out) :: adjoint_ij(*), wing_id(*)
            real(c_double), intent(out) :: wing_force(*), wing_moment(*)
        end function

        integer(c_int) function lbm_set_step(h, n_next, movenum) bind(C, name='lbm_set_step')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: n_next, movenum
        end function

        integer(c_int) function lbm_set_body_rel_pos(h, wing, pos) bind(C, name='lbm_set_body_rel_pos')
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: wing
            real(c_double), value :: pos
        end function

        integer(c_int) function lbm_get_body_rel(h, rel) bind(C, name='lbm_get_body_rel')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: rel(*)
        end function

        integer(c_int) function lbm_set_body_rel(h, rel) bind(C, name='lbm_set_body_rel')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: rel(*)
        end function

        integer(c_int) function lbm_set_dynamics(h, d) bind(C, name='lbm_set_dynamics')
            import :: c_int, c_ptr, lbm_dynamics
            type(c_ptr), value :: h
            type(lbm_dynamics), intent(in) :: d
        end function

        ! mode: 0 explicit inverse, 1 band substitution in coupled steps (default), 2 in every step
        integer(c_int) function lbm_set_ibm_solve(h, mode) bind(C, name='lbm_set_ibm_solve')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: mode
        end function

        integer(c_int) function lbm_last_error(h, buf, n) bind(C, name='lbm_last_error')
            import :: c_int, c_ptr, c_char, c_size_t
            type(c_ptr), value :: h
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_size_t), value :: n
        end function
    end interface

contains

    ! Message of the last failed call on h (c_null_ptr: the last failed
    ! lbm_create), as a Fortran string.
    function lbm_error_text(h) result(text)
        type(c_ptr), intent(in) :: h
        character(len=512) :: text
        character(kind=c_char) :: buf(512)
        integer(c_int) :: rc
        integer :: i

        buf = c_null_char
        rc = lbm_last_error(h, buf, int(size(buf), c_size_t))
        text = ''
        do i = 1, size(buf)
            if (buf(i) == c_null_char) exit
            text(i:i) = buf(i)
        enddo
    end function

end module lbm_hip
