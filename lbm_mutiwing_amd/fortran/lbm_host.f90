! lbm_host.f90 - host program of the multi-airfoil LBM/IBM solver on one
! MI355X, over the C ABI of liblbm_hip.so.
!
! It keeps the reference program's two run modes (main.f90:235-441,
! driver.f90:1-408):
!   file mode (default)    motion from motion_params.dat; steps n = 1, 2, ...
!                          while (n-1)*dt <= the table's last time; one
!                          force.dat and one airfoil_motion.dat record per step
!   online mode            LBM_MOTION_SOURCE=ZMQ: a ZeroMQ REP socket on
!                          LBM_ZMQ_ENDPOINT (default tcp://*:5555) takes
!                          RESET / STEP / SAVE / STOP and answers each one
!                          (FORCE / STATUS / ERROR); per-episode rows go to
!                          force_episode_<k>.dat and airfoil_motion_episode_<k>.dat,
!                          SAVE writes episode_<k>/F<cycle>.dat
! File mode also writes the restart checkpoint saved_flow_feild.dat and its
! copy every checkpoint_every steps (main.f90:436-439) and, with
! restart = .true., continues from one (main.f90:343-347, the goon branch).
! The reference fixes the case at compile time (main.f90:4-159,
! ini_mesh.f90:15-67); here it is read from the namelist /lbm_case/ in the
! file named by the first argument (default lbm_case.nml).  All solver work
! is on the GPU; the host only reads inputs, moves rows and formats text.
! Exit status 1 on any failed library call in file mode (the reference
! stops), ERROR replies in online mode.
program lbm_host
    use iso_c_binding
    use lbm_hip
    use lbm_front
    use lbm_host_io
    implicit none

    ! ---- case (namelist /lbm_case/)
    integer :: nx = 0, ny = 0, wing_count = 2, points_per_wing = 0, precision = 0
    integer :: rows_per_call = 100, print_every = 100, checkpoint_every = 100000
    logical :: restart = .false.
    real(c_double) :: scale = 0.01d0, ox = 0d0, oy = 0d0, uu = 0.05d0, re = 1000d0
    real(c_double) :: s(9) = [0d0, 1.4d0, 1.4d0, 0d0, 1.2d0, 0d0, 1.2d0, 0d0, 0d0]
    real(c_double) :: wing_x0 = 0d0, wing_dx = 1.5d0, wing_y0 = 0.5d0
    real(c_double) :: xc = 0d0, yc = 0d0, xp = 0d0, yp = 0d0
    character(len=256) :: boundary_dir = '.', motion_file = 'motion_params.dat', output_dir = 'output'
    character(len=256) :: restart_dir = ''
    ! rigid-body coupling of wing 1: the reference's compile-time switches and
    ! constants (main.f90:98, 122-151; if_rk4 boundary_condition.f90:1527-1529)
    logical :: solve_dynamic_equations = .false., closed_boundary = .true., if_rk4 = .true.
    real(c_double) :: solid_den = 10d0, area_airfoil = 6.807946476488794d-2, par1 = 0d0, par2 = 0d0
    real(c_double) :: alphad = 10.00d0 * 3.1415926535897932384626433832795d0 / 180.0d0
    namelist /lbm_case/ nx, ny, scale, ox, oy, uu, re, s, wing_count, points_per_wing, &
        wing_x0, wing_dx, wing_y0, xc, yc, xp, yp, precision, boundary_dir, motion_file, &
        output_dir, rows_per_call, print_every, checkpoint_every, restart, restart_dir, &
        solve_dynamic_equations, closed_boundary, if_rk4, solid_den, area_airfoil, par1, par2, alphad

    ! per-step row files: force.dat and airfoil_motion.dat (file mode) or the
    ! current episode's force_episode_<k>.dat / airfoil_motion_episode_<k>.dat
    integer, parameter :: UNIT_FORCE = 101, UNIT_MOTION = 102

    type(lbm_config) :: cfg
    type(lbm_dynamics) :: dyn
    type(c_ptr) :: h = c_null_ptr
    real(c_double), allocatable, target :: xb0(:), yb0(:)
    character(len=512) :: case_file, text, path
    character(len=16) :: env_mode
    integer :: u, ios, w, npts, stat, env_status
    logical :: ok, online

    case_file = 'lbm_case.nml'
    if (command_argument_count() >= 1) call get_command_argument(1, case_file)
    open(newunit=u, file=trim(case_file), status='old', action='read', iostat=ios)
    if (ios /= 0) call die('cannot open case file ' // trim(case_file))
    read(u, nml=lbm_case, iostat=ios)
    close(u)
    if (ios /= 0) call die('bad namelist /lbm_case/ in ' // trim(case_file))
    if (wing_count < 0 .or. wing_count > LBM_MAX_WINGS) call die('wing_count out of range')

    env_mode = ''
    call get_environment_variable('LBM_MOTION_SOURCE', env_mode, status=env_status)
    online = env_status == 0 .and. (trim(env_mode) == 'ZMQ' .or. trim(env_mode) == 'zmq')

    ! ---- markers (read_boundary): wing k from its own file
    npts = wing_count * points_per_wing
    allocate(xb0(max(1, npts)), yb0(max(1, npts)))
    do w = 1, wing_count
        path = marker_file_name(boundary_dir, w)
        call read_markers(path, points_per_wing, xb0(1 + (w - 1) * points_per_wing:), &
                          yb0(1 + (w - 1) * points_per_wing:), text, ok)
        if (.not. ok) call die(text)
    enddo

    cfg%nx = nx
    cfg%ny = ny
    cfg%wing_count = wing_count
    cfg%points_per_wing = points_per_wing
    cfg%scale = scale
    cfg%ox = ox
    cfg%oy = oy
    cfg%uu = uu
    cfg%re = re
    cfg%s = s
    cfg%wing_x0 = wing_x0
    cfg%wing_dx = wing_dx
    cfg%wing_y0 = wing_y0
    cfg%xc = xc
    cfg%yc = yc
    cfg%xp = xp
    cfg%yp = yp
    cfg%precision = precision
    cfg%xb0 = c_loc(xb0)
    cfg%yb0 = c_loc(yb0)
    if (online) then
        cfg%motion_source = LBM_MOTION_STREAM
    else
        cfg%motion_source = LBM_MOTION_FILE
    endif
    if (lbm_create(cfg, h) /= LBM_OK) call die('lbm_create: ' // trim(lbm_error_text(c_null_ptr)))
    if (solve_dynamic_equations) then
        dyn%enable = 1
        dyn%closed_boundary = merge(1, 0, closed_boundary)
        dyn%rk4 = merge(1, 0, if_rk4)
        dyn%reserved = 0
        dyn%solid_den = solid_den
        dyn%area = area_airfoil
        dyn%par1 = par1
        dyn%par2 = par2
        dyn%alphad = alphad
        call check(lbm_set_dynamics(h, dyn), 'lbm_set_dynamics')
    endif

    call make_dir(output_dir)
    if (online) then
        call run_online()
    else
        call run_file_mode()
    endif
    stat = lbm_destroy(h)

contains

    subroutine die(msg)
        character(len=*), intent(in) :: msg
        write(*, '(a)') trim(msg)
        if (c_associated(h)) stat = lbm_destroy(h)
        stop 1
    end subroutine

    subroutine check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc /= LBM_OK) call die(what // ': ' // trim(lbm_error_text(h)))
    end subroutine

    function row_len() result(n)
        integer :: n
        n = 4 + 12 * wing_count
    end function

    ! Advance through step n_last, writing the per-step force and motion rows
    ! to units UNIT_FORCE and UNIT_MOTION.  Returns the row sum, the last row
    ! and the number of steps run.
    subroutine advance_to(n_last, rsum, lrow, steps, err, rc)
        integer, intent(in) :: n_last
        real(c_double), intent(out) :: rsum(LBM_MAX_FORCE_ROW), lrow(LBM_MAX_FORCE_ROW), err
        integer, intent(out) :: steps
        integer(c_int), intent(out) :: rc
        real(c_double), allocatable, target :: frows(:, :), mrows(:, :)
        integer(c_int32_t) :: n_next, movenum, flag2, rl, nadj

        rsum = 0d0
        lrow = 0d0
        err = 0d0
        rc = lbm_get_state(h, n_next, movenum, flag2, rl, nadj)
        if (rc /= LBM_OK) return
        steps = max(0, n_last - n_next + 1)
        if (steps == 0) return
        allocate(frows(row_len(), steps), mrows(1 + 3 * wing_count, steps))
        rc = lbm_advance(h, int(n_last, c_int32_t), c_loc(frows), c_loc(mrows), rsum, lrow, err)
        if (rc /= LBM_OK) return
        call write_rows(UNIT_FORCE, frows, row_len(), steps)
        call write_rows(UNIT_MOTION, mrows, 1 + 3 * wing_count, steps)
    end subroutine

    ! ------------------------------------------------------------ file mode
    subroutine run_file_mode()
        real(c_double), allocatable :: table(:, :)
        real(c_double) :: end_time, dt, rsum(LBM_MAX_FORCE_ROW), lrow(LBM_MAX_FORCE_ROW), err
        type(checkpoint_ring) :: ring
        integer :: nrows, n_end, n_next, n_last, steps
        integer(c_int) :: rc

        call read_motion_table(motion_file, 1 + 9 * wing_count, table, nrows, text, ok)
        if (.not. ok) call die(text)
        call check(lbm_set_motion_table(h, table, int(nrows, c_int32_t), int(1 + 9 * wing_count, c_int32_t), &
                                        end_time), 'lbm_set_motion_table')
        ! last step: the loop runs n while (n-1)*dt <= end_time + 1e-12, with
        ! the time computed as the reference does (main.f90:374-376)
        dt = scale * uu
        n_end = 0
        do while (dble(n_end) * dt <= end_time + 1.0d-12)
            n_end = n_end + 1
        enddo

        if (restart) then
            call restore_checkpoint(ring, n_next)
        else
            call init_ring(ring)
            n_next = 1
        endif
        open(UNIT_FORCE, file=trim(output_dir) // '/force.dat', status='replace', action='write')
        open(UNIT_MOTION, file=trim(output_dir) // '/airfoil_motion.dat', status='replace', action='write')
        do while (n_next <= n_end)
            n_last = min(n_end, n_next + rows_per_call - 1)
            if (checkpoint_every > 0) n_last = min(n_last, ((n_next - 1) / checkpoint_every + 1) * checkpoint_every)
            call advance_to(n_last, rsum, lrow, steps, err, rc)
            call check(rc, 'lbm_advance')
            if (print_every > 0) then
                if (n_last / print_every > (n_next - 1) / print_every) &
                    write(*, '(a,i0,a,es12.5)') 'step ', n_last, '  err ', err
            endif
            if (checkpoint_every > 0) then
                if (mod(n_last, checkpoint_every) == 0) call save_checkpoint(n_last, ring)
            endif
            n_next = n_last + 1
        enddo
        close(UNIT_FORCE)
        close(UNIT_MOTION)
        write(*, '(a,i0,a)') 'file mode done: ', n_end, ' steps'
    end subroutine

    ! The boundary records of a fresh start: the populations init leaves.
    subroutine init_ring(ring)
        type(checkpoint_ring), intent(out) :: ring
        real(c_double), allocatable :: f(:, :, :), f0(:, :, :)
        allocate(f(nx, ny, 9), f0(nx, ny, 9))
        call check(lbm_get_populations(h, f, f0), 'lbm_get_populations')
        call ring_from_populations(nx, ny, f, ring)
    end subroutine

    ! saveflowfield (solution.f90:1616-1719): the same checkpoint twice.
    subroutine save_checkpoint(n, ring)
        integer, intent(in) :: n
        type(checkpoint_ring), intent(inout) :: ring
        real(c_double), allocatable :: f(:, :, :), f0(:, :, :), uf(:, :), vf(:, :), rho(:, :), pre(:, :)
        real(c_double) :: rel(wing_count, 11)
        integer(c_int32_t) :: n_next, movenum, flag2, rl, nadj
        character(len=32), parameter :: names(2) = [character(len=32) :: 'saved_flow_feild', 'saved_flow_feild_copy']
        integer :: q

        allocate(f(nx, ny, 9), f0(nx, ny, 9), uf(nx, ny), vf(nx, ny), rho(nx, ny), pre(nx, ny))
        call check(lbm_get_state(h, n_next, movenum, flag2, rl, nadj), 'lbm_get_state')
        call check(lbm_get_fields(h, rho, uf, vf, pre), 'lbm_get_fields')
        call check(lbm_get_populations(h, f, f0), 'lbm_get_populations')
        if (wing_count > 0) call check(lbm_get_body_rel(h, rel), 'lbm_get_body_rel')
        ! the f0 ring as the library holds it (move_mesh shifts it, rows included)
        ring%f0_left = f0(1, :, :)
        ring%f0_right = f0(nx, :, :)
        ring%f0_bottom = f0(:, 1, :)
        ring%f0_top = f0(:, ny, :)
        do q = 1, 2
            call write_checkpoint(trim(output_dir) // '/' // trim(names(q)) // '.dat', n, wing_count, rel, &
                                  int(movenum), nx, ny, uf, vf, f, ring, ok)
            if (.not. ok) call die('cannot write ' // trim(output_dir) // '/' // trim(names(q)) // '.dat')
        enddo
    end subroutine

    ! readflowfield (solution.f90:1721-1787): the checkpoint in restart_dir
    ! (default output_dir), its copy if that fails; the next step is the
    ! saved step + 1 (main.f90:343-347).
    subroutine restore_checkpoint(ring, n_next)
        type(checkpoint_ring), intent(out) :: ring
        integer, intent(out) :: n_next
        real(c_double), allocatable :: f(:, :, :), f0(:, :, :), uf(:, :), vf(:, :)
        real(c_double) :: rel(wing_count, 11)
        character(len=512) :: dir
        integer :: cycle, movenum

        dir = restart_dir
        if (len_trim(dir) == 0) dir = output_dir
        allocate(f(nx, ny, 9), f0(nx, ny, 9), uf(nx, ny), vf(nx, ny))
        call read_checkpoint(trim(dir) // '/saved_flow_feild.dat', wing_count, nx, ny, cycle, rel, movenum, &
                             uf, vf, f, ring, text, ok)
        if (ok) then
            write(*, '(a)') 'sucess!'
        else
            write(*, '(a)') trim(text)
            write(*, '(a)') '启动备用存档！'
            call read_checkpoint(trim(dir) // '/saved_flow_feild_copy.dat', wing_count, nx, ny, cycle, rel, &
                                 movenum, uf, vf, f, ring, text, ok)
            if (.not. ok) call die(text)
            write(*, '(a)') '读取保存的流场数据成功!'
        endif
        f0 = 0d0
        f0(1, :, :) = ring%f0_left
        f0(nx, :, :) = ring%f0_right
        f0(:, 1, :) = ring%f0_bottom
        f0(:, ny, :) = ring%f0_top
        call check(lbm_set_step(h, int(cycle + 1, c_int32_t), int(movenum, c_int32_t)), 'lbm_set_step')
        if (wing_count > 0) call check(lbm_set_body_rel(h, rel), 'lbm_set_body_rel')
        call check(lbm_set_populations(h, f, f0), 'lbm_set_populations')
        n_next = cycle + 1
    end subroutine

    ! ---------------------------------------------------------- online mode
    subroutine run_online()
        character(len=256) :: endpoint
        real(c_double) :: row(LBM_MAX_MOTION_COLS), pending(LBM_MAX_MOTION_COLS)
        real(c_double) :: start_row(LBM_MAX_MOTION_COLS), target_row(LBM_MAX_MOTION_COLS)
        real(c_double) :: next_row(LBM_MAX_MOTION_COLS)
        integer :: cmd, st, episode, ncols
        logical :: stopping, have_pending, files_open

        endpoint = 'tcp://*:5555'
        call get_environment_variable('LBM_ZMQ_ENDPOINT', text, status=st)
        if (st == 0 .and. len_trim(text) > 0) endpoint = text
        call front_open(endpoint, st)
        if (st /= 0) then
            write(text, '(a,i0)') 'ZeroMQ set-up failed, stage ', st
            call die(text)
        endif
        write(*, '(a)') 'online mode on ' // trim(endpoint) // ': waiting for RESET / STEP / SAVE / STOP'
        flush(6)

        ncols = 1 + 9 * wing_count
        episode = 0
        files_open = .false.
        stopping = .false.
        have_pending = .false.
        do while (.not. stopping)
            ! between episodes: only RESET starts one
            if (have_pending) then
                cmd = CMD_RESET
                row = pending
                have_pending = .false.
            else
                call front_next_command(wing_count, cmd, row, ok)
                if (.not. ok) cycle
            endif
            select case (cmd)
            case (CMD_STOP)
                call front_reply_status('STOPPING')
                exit
            case (CMD_OUTPUT)
                call front_reply_error('online mode requires RESET before SAVE')
                cycle
            case (CMD_RESET)
            case default
                call front_reply_error('online mode requires RESET before STEP')
                cycle
            end select
            if (abs(row(1)) > 1.0d-12) then
                call front_reply_error('RESET command must provide initial time 0.0')
                cycle
            endif

            ! new episode (driver.f90:114-147)
            episode = episode + 1
            call close_episode_files(files_open)
            call open_episode_files(episode, files_open)
            start_row = 0d0
            start_row(1) = row(1)
            target_row = start_row
            if (.not. reply_on_error(lbm_reset(h))) cycle
            if (.not. reply_on_error(lbm_set_stream_segment(h, start_row, target_row, int(ncols, c_int32_t)))) cycle
            if (.not. run_segment(row(1))) then
                call close_episode_files(files_open)
                cycle
            endif

            ! inside an episode (driver.f90:149-196)
            do
                call front_next_command(wing_count, cmd, row, ok)
                if (.not. ok) cycle
                if (cmd == CMD_STOP) then
                    call front_reply_status('STOPPING')
                    stopping = .true.
                    exit
                endif
                if (cmd == CMD_RESET) then
                    pending = row
                    have_pending = .true.
                    exit
                endif
                if (cmd == CMD_OUTPUT) then
                    call save_flow(episode)
                    cycle
                endif
                if (cmd /= CMD_STEP) then
                    call front_reply_error('expected STEP, SAVE, RESET or STOP')
                    cycle
                endif
                start_row = target_row
                if (lbm_build_stream_target_row(int(wing_count, c_int32_t), start_row, row(1), row, &
                                                next_row) /= LBM_OK) then
                    call front_reply_error('STEP 时间必须大于等于当前段起始时刻')
                    cycle
                endif
                target_row = 0d0
                target_row(1:ncols) = next_row(1:ncols)
                if (.not. reply_on_error(lbm_set_stream_segment(h, start_row, target_row, int(ncols, c_int32_t)))) &
                    cycle
                if (.not. run_segment(row(1))) then
                    ! the failed STEP's target is dropped (driver.f90:185-188)
                    target_row = start_row
                    st = lbm_set_stream_segment(h, start_row, target_row, int(ncols, c_int32_t))
                endif
            enddo
            call close_episode_files(files_open)
        enddo
        call close_episode_files(files_open)
        call front_close()
    end subroutine

    ! ERROR reply with the library's message for a failed call
    logical function reply_on_error(rc)
        integer(c_int), intent(in) :: rc
        reply_on_error = rc == LBM_OK
        if (.not. reply_on_error) call front_reply_error(trim(lbm_error_text(h)))
    end function

    ! Advance to target_time and send the FORCE reply, or the ERROR reply
    ! (advance_zmq_segment + finalize_force_reply_average,
    ! driver.f90:270-289, 322-367): the average over the segment's steps of
    ! columns 2 .. 1+3W+3, column 1 the segment end time, the remaining
    ! (motion) columns from the last step.
    logical function run_segment(target_time) result(done)
        real(c_double), intent(in) :: target_time
        real(c_double) :: dt, rsum(LBM_MAX_FORCE_ROW), lrow(LBM_MAX_FORCE_ROW), err, reply(LBM_MAX_FORCE_ROW)
        integer :: target_cycle, steps, last_col
        integer(c_int32_t) :: n_next, movenum, flag2, rl, nadj
        integer(c_int) :: rc

        done = .false.
        dt = scale * uu
        if (dt <= 0d0) then
            call front_reply_error('求解器时间步无效')
            return
        endif
        target_cycle = nint(target_time / dt) + 1
        if (abs(dble(target_cycle - 1) * dt - target_time) > 1.0d-10) then
            call front_reply_error('STEP 时间必须与 LBM 时间步严格对齐')
            return
        endif
        rc = lbm_get_state(h, n_next, movenum, flag2, rl, nadj)
        if (.not. reply_on_error(rc)) return
        if (target_cycle < n_next) then
            call front_reply_error('STEP 时间必须大于等于当前待推进时刻')
            return
        endif
        call advance_to(target_cycle, rsum, lrow, steps, err, rc)
        if (.not. reply_on_error(rc)) return
        if (steps <= 0) then
            call front_reply_force(reply, 0)
        else
            reply = lrow
            last_col = min(row_len(), 1 + 3 * wing_count + 3)
            reply(2:last_col) = rsum(2:last_col) / dble(steps)
            reply(1) = target_time
            call front_reply_force(reply, row_len())
        endif
        done = .true.
    end function

    ! SAVE (driver.f90:291-320): not inside the three steps after a
    ! move_mesh; the snapshot of the last completed step.
    subroutine save_flow(episode)
        integer, intent(in) :: episode
        real(c_double), allocatable :: rho(:, :), uf(:, :), vf(:, :), pre(:, :)
        real(c_double), allocatable :: xb1(:), yb1(:), xb2(:), yb2(:)
        integer(c_int32_t), allocatable :: body(:, :), adj(:, :), wid(:)
        real(c_double) :: wf(2 * LBM_MAX_WINGS), wm(LBM_MAX_WINGS)
        integer(c_int32_t) :: n_next, movenum, flag2, rl, nadj
        integer :: cycle, k
        character(len=32) :: ctext
        character(len=512) :: dir

        if (.not. reply_on_error(lbm_get_state(h, n_next, movenum, flag2, rl, nadj))) return
        if (flag2 /= 0) then
            call front_reply_error('当前时刻处于移网格后的过渡阶段，请在下一次 STEP 后再保存流场')
            return
        endif
        cycle = max(1, n_next - 1)
        allocate(rho(nx, ny), uf(nx, ny), vf(nx, ny), pre(nx, ny), body(nx, ny))
        allocate(xb1(max(1, npts)), yb1(max(1, npts)), xb2(max(1, npts)), yb2(max(1, npts)))
        allocate(adj(2, max(1, nadj)), wid(max(1, nadj)))
        if (.not. reply_on_error(lbm_get_fields(h, rho, uf, vf, pre))) return
        if (.not. reply_on_error(lbm_get_markers(h, xb1, yb1, xb2, yb2))) return
        ! body = 1 at the adjoint nodes of the last step (airfoil.f90:322-329)
        body = 0
        if (nadj > 0) then
            if (.not. reply_on_error(lbm_get_ibm(h, nadj, adj, wid, wf, wm))) return
            do k = 1, nadj
                body(adj(1, k), adj(2, k)) = 1
            enddo
        endif
        write(ctext, '(i0)') episode
        dir = trim(output_dir) // '/episode_' // trim(ctext)
        call make_dir(dir)
        write(ctext, '(i0)') cycle
        call write_tecplot(trim(dir) // '/F' // trim(ctext) // '.dat', cycle, nx, ny, ox, oy, scale, uu, &
                           int(movenum), uf, vf, pre, body, wing_count, points_per_wing, xb1, yb1)
        call front_reply_status('FLOW_SAVED F' // trim(ctext) // '.dat')
    end subroutine

    subroutine open_episode_files(episode, files_open)
        integer, intent(in) :: episode
        logical, intent(out) :: files_open
        character(len=32) :: ctext
        write(ctext, '(i0)') episode
        open(UNIT_FORCE, file=trim(output_dir) // '/force_episode_' // trim(ctext) // '.dat', &
             status='replace', action='write')
        open(UNIT_MOTION, file=trim(output_dir) // '/airfoil_motion_episode_' // trim(ctext) // '.dat', &
             status='replace', action='write')
        files_open = .true.
    end subroutine

    subroutine close_episode_files(files_open)
        logical, intent(inout) :: files_open
        if (files_open) then
            close(UNIT_FORCE)
            close(UNIT_MOTION)
        endif
        files_open = .false.
    end subroutine

end program lbm_host
