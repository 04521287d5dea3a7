! lbm_host_io.f90 - the host's files: marker files, the motion table, the
! per-step force / motion rows and the Tecplot flow snapshot.
!
!   marker file of wing k   boundary.dat (k = 1), boundary<k-1>.dat (k > 1),
!                           one '(2f14.8)' record (x, y) per marker, chord
!                           units (airfoil.f90:190-242)
!   motion_params.dat       list-directed rows of 1 + 9W columns, blank lines
!                           skipped, time column non-decreasing
!                           (mode.f90:104-178)
!   force / motion rows     one '(1000f15.8)' record per step
!                           (solution.f90:386-402)
!   F<cycle>.dat            Tecplot: the flow zone sampled every second node,
!                           then the airfoil FE-triangle zone
!                           (solution.f90:132-241)
module lbm_host_io
    use iso_c_binding
    implicit none
    private
    public :: marker_file_name, read_markers, read_motion_table, write_rows, write_tecplot, make_dir

    integer, parameter :: STRIDE_X = 2, STRIDE_Y = 2   ! main.f90:38
    integer, parameter :: LINE_BYTES = 8192

contains

    function marker_file_name(dir, wing) result(path)
        character(len=*), intent(in) :: dir
        integer, intent(in) :: wing
        character(len=512) :: path
        character(len=16) :: idx

        if (wing == 1) then
            path = trim(dir) // '/boundary.dat'
        else
            write(idx, '(i0)') wing - 1
            path = trim(dir) // '/boundary' // trim(idx) // '.dat'
        endif
    end function

    subroutine read_markers(path, npts, x, y, msg, ok)
        character(len=*), intent(in) :: path
        integer, intent(in) :: npts
        real(c_double), intent(out) :: x(npts), y(npts)
        character(len=*), intent(out) :: msg
        logical, intent(out) :: ok
        integer :: u, ios, i

        ok = .false.
        msg = ''
        open(newunit=u, file=trim(path), status='old', action='read', iostat=ios)
        if (ios /= 0) then
            msg = '无法打开边界文件: ' // trim(path)
            return
        endif
        read(u, '(2f14.8)', iostat=ios) (x(i), y(i), i = 1, npts)
        close(u)
        if (ios /= 0) then
            msg = '边界文件格式错误: ' // trim(path)
            return
        endif
        ok = .true.
    end subroutine

    ! table(1:ncols, 1:nrows): one column per row of the file, so that the
    ! memory order is the row-major layout lbm_set_motion_table takes.
    subroutine read_motion_table(path, ncols, table, nrows, msg, ok)
        character(len=*), intent(in) :: path
        integer, intent(in) :: ncols
        real(c_double), allocatable, intent(out) :: table(:, :)
        integer, intent(out) :: nrows
        character(len=*), intent(out) :: msg
        logical, intent(out) :: ok
        character(len=LINE_BYTES) :: line
        character(len=16) :: ctext
        integer :: u, ios, pass, r

        ok = .false.
        msg = ''
        nrows = 0
        write(ctext, '(i0)') ncols
        open(newunit=u, file=trim(path), status='old', action='read', iostat=ios)
        if (ios /= 0) then
            msg = '无法打开运动参数文件: ' // trim(path)
            return
        endif
        ! pass 1 counts and validates the rows, pass 2 stores them
        do pass = 1, 2
            r = 0
            do
                read(u, '(a)', iostat=ios) line
                if (ios /= 0) exit
                if (len_trim(line) == 0) cycle
                r = r + 1
                if (pass == 1) then
                    block
                        real(c_double) :: probe(ncols)
                        read(line, *, iostat=ios) probe
                    end block
                else
                    read(line, *, iostat=ios) table(:, r)
                endif
                if (ios /= 0) then
                    msg = '运动参数文件格式错误，列数应为: ' // trim(ctext) // ' ' // trim(path)
                    close(u)
                    return
                endif
            enddo
            if (pass == 1) then
                if (r < 1) then
                    msg = '运动参数文件为空: ' // trim(path)
                    close(u)
                    return
                endif
                nrows = r
                allocate(table(ncols, nrows))
                rewind(u)
            endif
        enddo
        close(u)
        do r = 2, nrows
            if (table(1, r) < table(1, r - 1)) then
                msg = '运动参数文件时间列必须非递减: ' // trim(path)
                return
            endif
        enddo
        ok = .true.
    end subroutine

    ! rows(1:ncols, 1:nrows), one '(1000f15.8)' record per column
    subroutine write_rows(u, rows, ncols, nrows)
        integer, intent(in) :: u, ncols, nrows
        real(c_double), intent(in) :: rows(ncols, nrows)
        integer :: r
        do r = 1, nrows
            write(u, '(1000f15.8)') rows(:, r)
        enddo
    end subroutine

    ! Zone headers are one record each: integers in 12 columns, reals with 17
    ! significant digits (F25.17 for 1e-3 <= |x| < 1e8 and zero, ES25.16
    ! otherwise), the layout of list-directed output without its line breaks.
    function int_text(v) result(t)
        integer, intent(in) :: v
        character(len=12) :: t
        write(t, '(i12)') v
    end function

    function real_text(x) result(t)
        real(c_double), intent(in) :: x
        character(len=25) :: t
        if (x == 0d0 .or. (abs(x) >= 1d-3 .and. abs(x) < 1d8)) then
            write(t, '(f25.17)') x
        else
            write(t, '(es25.16)') x
        endif
    end function

    subroutine make_dir(path)
        character(len=*), intent(in) :: path
        integer :: stat
        call execute_command_line('mkdir -p ' // trim(path), exitstat=stat)
    end subroutine

    ! Tecplot snapshot of one lattice.  u, v, pre and body are (nx, ny) with
    ! the reference's indices; velocities are written in units of uu, the
    ! vorticity dv/dx - du/dy in units of uu per chord with central
    ! differences inside and second-order one-sided ones on the edges.  The
    ! x coordinates include the frame shift movenum*scale of move_mesh.
    subroutine write_tecplot(path, cycle, nx, ny, ox, oy, scale, uu, movenum, u, v, pre, body, &
                             wing_count, points_per_wing, xb1, yb1)
        character(len=*), intent(in) :: path
        integer, intent(in) :: cycle, nx, ny, movenum, wing_count, points_per_wing
        real(c_double), intent(in) :: ox, oy, scale, uu
        real(c_double), intent(in) :: u(nx, ny), v(nx, ny), pre(nx, ny)
        integer(c_int32_t), intent(in) :: body(nx, ny)
        real(c_double), intent(in) :: xb1(*), yb1(*)
        integer :: fu, i, j, w, k, first, mid, npts
        real(c_double) :: dvx, duy, shift, tsol

        shift = dble(movenum) * scale
        tsol = cycle * uu * scale
        npts = wing_count * points_per_wing
        open(newunit=fu, file=trim(path), status='replace', action='write')
        write(fu, '(a)') 'VARIABLES = "X" "Y" "U" "V" "Presure" "Vorticity" "body" "dvordt" '
        write(fu, '(a)') ' ZONE T="Flow feild", I=' // int_text((nx - 1) / STRIDE_X + 1) // ',J=' // &
            int_text((ny - 1) / STRIDE_Y + 1) // ',SOLUTIONTIME=' // real_text(tsol)
        do j = 1, ny, STRIDE_Y
            do i = 1, nx, STRIDE_X
                if (i == nx) then
                    dvx = (1.5d0 * v(i, j) - 2.0d0 * v(i - 1, j) + 0.5d0 * v(i - 2, j)) / (uu * scale)
                else if (i == 1) then
                    dvx = (-1.5d0 * v(i, j) + 2.0d0 * v(i + 1, j) - 0.5d0 * v(i + 2, j)) / (uu * scale)
                else
                    dvx = (v(i + 1, j) - v(i - 1, j)) / (uu * 2.0d0 * scale)
                endif
                if (j == ny) then
                    duy = (1.5d0 * u(i, j) - 2.0d0 * u(i, j - 1) + 0.5d0 * u(i, j - 2)) / (uu * scale)
                else if (j == 1) then
                    duy = (-1.5d0 * u(i, j) + 2.0d0 * u(i, j + 1) - 0.5d0 * u(i, j + 2)) / (uu * scale)
                else
                    duy = (u(i, j + 1) - u(i, j - 1)) / (uu * 2.0d0 * scale)
                endif
                ! the last column (dvordt) is never computed by the reference
                write(fu, '(25e16.8)') ox + dble(i - 1) * scale + shift, oy + dble(j - 1) * scale, &
                    u(i, j) / uu, v(i, j) / uu, pre(i, j), dvx - duy, dble(body(i, j)), 0.0d0
            enddo
        enddo
        write(fu, '(a)') ' ZONE T="airfoil", NODES=' // int_text(npts + wing_count) // ' ELEMENTS,=' // &
            int_text(wing_count * (points_per_wing - 1)) // ', DATAPACKING=POINT, ZONETYPE=FETRIANGLE' // &
            ',SOLUTIONTIME=' // real_text(tsol)
        do k = 1, npts
            write(fu, '(8f14.8)') xb1(k) + shift, yb1(k), 255d0, 0d0, 0d0, 0d0, 0d0, 0d0
        enddo
        do w = 1, wing_count
            first = 1 + (w - 1) * points_per_wing
            mid = first + (points_per_wing - 1) / 2
            write(fu, '(8f14.8)') 0.5d0 * ((xb1(first) + shift) + (xb1(mid) + shift)), &
                0.5d0 * (yb1(first) + yb1(mid)), 255d0, 0d0, 0d0, 0d0, 0d0, 0d0
        enddo
        do w = 1, wing_count
            first = 1 + (w - 1) * points_per_wing
            do k = first, first + points_per_wing - 2
                write(fu, '(3i5.3)') k, k + 1, npts + w
            enddo
        enddo
        close(fu)
    end subroutine

end module lbm_host_io
