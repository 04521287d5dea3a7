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
!   saved_flow_feild.dat    restart checkpoint, unformatted stream (the
!   (+ _copy)               reference's form='binary'): see write_checkpoint
!                           (solution.f90:1616-1851)
module lbm_host_io
    use iso_c_binding
    implicit none
    private
    public :: marker_file_name, read_markers, read_motion_table, write_rows, write_tecplot, make_dir
    public :: checkpoint_ring, ring_from_populations, write_checkpoint, read_checkpoint, checkpoint_bytes

    ! The boundary records of the checkpoint: the previous-step boundary
    ! populations leftf2/rightf2 (rows 2..ny-1), upf2/downf2 (all columns),
    ! which the live step never changes after init (solution.f90:60-63), and
    ! the f0 ring (main.f90:174, solution.f90:1704-1715).  (ny, 9) and (nx, 9).
    type :: checkpoint_ring
        real(c_double), allocatable :: left2(:, :), f0_left(:, :), right2(:, :), f0_right(:, :)
        real(c_double), allocatable :: up2(:, :), f0_top(:, :), down2(:, :), f0_bottom(:, :)
    end type

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

    ! The ring records as init leaves them: f (nx, ny, 9) just after init, where
    ! f = f0 = the equilibrium everywhere (solution.f90:55-63).
    subroutine ring_from_populations(nx, ny, f, ring)
        integer, intent(in) :: nx, ny
        real(c_double), intent(in) :: f(nx, ny, 9)
        type(checkpoint_ring), intent(out) :: ring
        allocate(ring%left2(ny, 9), ring%f0_left(ny, 9), ring%right2(ny, 9), ring%f0_right(ny, 9))
        allocate(ring%up2(nx, 9), ring%f0_top(nx, 9), ring%down2(nx, 9), ring%f0_bottom(nx, 9))
        ring%left2 = f(1, :, :)
        ring%f0_left = f(1, :, :)
        ring%right2 = f(nx, :, :)
        ring%f0_right = f(nx, :, :)
        ring%up2 = f(:, ny, :)
        ring%f0_top = f(:, ny, :)
        ring%down2 = f(:, 1, :)
        ring%f0_bottom = f(:, 1, :)
    end subroutine

    integer(c_int64_t) function checkpoint_bytes(wing_count, nx, ny)
        integer, intent(in) :: wing_count, nx, ny
        checkpoint_bytes = 12_c_int64_t + 88_c_int64_t * wing_count + 88_c_int64_t * nx * ny + &
                           288_c_int64_t * (ny - 2) + 288_c_int64_t * nx
    end function

    ! Checkpoint as saveflowfield writes it (solution.f90:1660-1686): the
    ! step, the wing count, the eleven rigid-body arrays of wing_count
    ! doubles (body_rel(:, 1..11) = body_rel_a, _a_prev, _v, _v_prev, _s,
    ! _s_prev, _pos, _theta_prev, _theta, _temp_theta_prev, _temp_theta, the
    ! layout of lbm_get_body_rel), movenum, then u, v, f(1:9) node by node
    ! (i fastest), then per row j = 2..ny-1 leftf2, f0(1,j), rightf2,
    ! f0(nx,j), then per column i = 1..nx upf2, f0(i,ny), downf2, f0(i,1).
    ! Integers are 4 bytes, reals 8, no record markers.
    subroutine write_checkpoint(path, cycle, wing_count, body_rel, movenum, nx, ny, u, v, f, ring, ok)
        character(len=*), intent(in) :: path
        integer, intent(in) :: cycle, wing_count, movenum, nx, ny
        real(c_double), intent(in) :: body_rel(wing_count, 11), u(nx, ny), v(nx, ny), f(nx, ny, 9)
        type(checkpoint_ring), intent(in) :: ring
        logical, intent(out) :: ok
        real(c_double), allocatable :: buf(:, :)
        integer :: fu, ios, i, j

        ok = .false.
        open(newunit=fu, file=trim(path), access='stream', form='unformatted', status='replace', &
             action='write', iostat=ios)
        if (ios /= 0) return
        allocate(buf(11, nx))
        write(fu, iostat=ios) int(cycle, c_int32_t), int(wing_count, c_int32_t)
        if (wing_count > 0 .and. ios == 0) write(fu, iostat=ios) body_rel
        if (ios == 0) write(fu, iostat=ios) int(movenum, c_int32_t)
        do j = 1, ny
            if (ios /= 0) exit
            do i = 1, nx
                buf(1, i) = u(i, j)
                buf(2, i) = v(i, j)
                buf(3:11, i) = f(i, j, :)
            enddo
            write(fu, iostat=ios) buf
        enddo
        do j = 2, ny - 1
            if (ios /= 0) exit
            write(fu, iostat=ios) ring%left2(j, :), ring%f0_left(j, :), ring%right2(j, :), ring%f0_right(j, :)
        enddo
        do i = 1, nx
            if (ios /= 0) exit
            write(fu, iostat=ios) ring%up2(i, :), ring%f0_top(i, :), ring%down2(i, :), ring%f0_bottom(i, :)
        enddo
        close(fu)
        ok = ios == 0
    end subroutine

    ! readflowfield (solution.f90:1721-1787): the checkpoint of a case of
    ! wing_count wings on an nx x ny lattice.  A missing file, a file of the
    ! wrong size or another wing count fails with a message.
    subroutine read_checkpoint(path, wing_count, nx, ny, cycle, body_rel, movenum, u, v, f, ring, msg, ok)
        character(len=*), intent(in) :: path
        integer, intent(in) :: wing_count, nx, ny
        integer, intent(out) :: cycle, movenum
        real(c_double), intent(out) :: body_rel(wing_count, 11), u(nx, ny), v(nx, ny), f(nx, ny, 9)
        type(checkpoint_ring), intent(out) :: ring
        character(len=*), intent(out) :: msg
        logical, intent(out) :: ok
        real(c_double), allocatable :: buf(:, :)
        integer(c_int32_t) :: i4, w4, m4
        integer(c_int64_t) :: fsize
        logical :: exists
        integer :: fu, ios, i, j

        ok = .false.
        msg = ''
        inquire(file=trim(path), exist=exists, size=fsize)
        if (.not. exists) then
            msg = '无法打开续算存档文件: ' // trim(path)
            return
        endif
        if (fsize /= checkpoint_bytes(wing_count, nx, ny)) then
            msg = '续算存档文件大小与网格不一致: ' // trim(path)
            return
        endif
        open(newunit=fu, file=trim(path), access='stream', form='unformatted', status='old', &
             action='read', iostat=ios)
        if (ios /= 0) then
            msg = '无法打开续算存档文件: ' // trim(path)
            return
        endif
        read(fu, iostat=ios) i4, w4
        if (ios == 0 .and. w4 /= wing_count) then
            close(fu)
            msg = '续算文件中的翼数量与当前编译设置不一致'
            return
        endif
        allocate(buf(11, nx))
        body_rel = 0d0
        if (ios == 0 .and. wing_count > 0) read(fu, iostat=ios) body_rel
        if (ios == 0) read(fu, iostat=ios) m4
        do j = 1, ny
            if (ios /= 0) exit
            read(fu, iostat=ios) buf
            do i = 1, nx
                u(i, j) = buf(1, i)
                v(i, j) = buf(2, i)
                f(i, j, :) = buf(3:11, i)
            enddo
        enddo
        allocate(ring%left2(ny, 9), ring%f0_left(ny, 9), ring%right2(ny, 9), ring%f0_right(ny, 9))
        allocate(ring%up2(nx, 9), ring%f0_top(nx, 9), ring%down2(nx, 9), ring%f0_bottom(nx, 9))
        ring%left2 = 0d0
        ring%f0_left = 0d0
        ring%right2 = 0d0
        ring%f0_right = 0d0
        do j = 2, ny - 1
            if (ios /= 0) exit
            read(fu, iostat=ios) ring%left2(j, :), ring%f0_left(j, :), ring%right2(j, :), ring%f0_right(j, :)
        enddo
        do i = 1, nx
            if (ios /= 0) exit
            read(fu, iostat=ios) ring%up2(i, :), ring%f0_top(i, :), ring%down2(i, :), ring%f0_bottom(i, :)
        enddo
        close(fu)
        if (ios /= 0) then
            msg = '续算存档文件读取失败: ' // trim(path)
            return
        endif
        ! the corners of the f0 ring are the end columns' rows 1 and ny
        ring%f0_left(1, :) = ring%f0_bottom(1, :)
        ring%f0_left(ny, :) = ring%f0_top(1, :)
        ring%f0_right(1, :) = ring%f0_bottom(nx, :)
        ring%f0_right(ny, :) = ring%f0_top(nx, :)
        cycle = i4
        movenum = m4
        ok = .true.
    end subroutine

end module lbm_host_io
