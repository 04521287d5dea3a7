! Test driver of the host's checkpoint I/O (module lbm_host_io), no GPU work.
!   checkpoint_check write <path> <nx> <ny> <wings> <cycle> <movenum>
!       writes a checkpoint of known values (all exact in binary):
!       u = i + j/1024, v = -i - j/2048, f(k) = k + i/1024 + j/2^20,
!       ring record r (1..8 in file order) = 16 r + k + (i or j)/1024,
!       body_rel(w, q) = q + w/8
!   checkpoint_check copy <in> <out> <nx> <ny> <wings>
!       read_checkpoint <in>, then write_checkpoint <out>; prints the
!       error text and exits 1 if the read fails
program checkpoint_check
    use iso_c_binding
    use lbm_host_io
    implicit none
    character(len=512) :: mode, path, path2, arg, msg
    integer :: nx, ny, w, cycle, movenum, i, j, k
    real(c_double), allocatable :: u(:, :), v(:, :), f(:, :, :)
    real(c_double), allocatable :: rel(:, :)
    type(checkpoint_ring) :: ring
    logical :: ok

    call get_command_argument(1, mode)
    call get_command_argument(2, path)
    if (trim(mode) == 'write') then
        call get_command_argument(3, arg); read(arg, *) nx
        call get_command_argument(4, arg); read(arg, *) ny
        call get_command_argument(5, arg); read(arg, *) w
        call get_command_argument(6, arg); read(arg, *) cycle
        call get_command_argument(7, arg); read(arg, *) movenum
        allocate(u(nx, ny), v(nx, ny), f(nx, ny, 9))
        allocate(ring%left2(ny, 9), ring%f0_left(ny, 9), ring%right2(ny, 9), ring%f0_right(ny, 9))
        allocate(ring%up2(nx, 9), ring%f0_top(nx, 9), ring%down2(nx, 9), ring%f0_bottom(nx, 9))
        do j = 1, ny
            do i = 1, nx
                u(i, j) = dble(i) + dble(j) / 1024d0
                v(i, j) = -dble(i) - dble(j) / 2048d0
                do k = 1, 9
                    f(i, j, k) = dble(k) + dble(i) / 1024d0 + dble(j) / 1048576d0
                enddo
            enddo
        enddo
        do k = 1, 9
            do j = 1, ny
                ring%left2(j, k) = 16d0 + k + j / 1024d0
                ring%f0_left(j, k) = 32d0 + k + j / 1024d0
                ring%right2(j, k) = 48d0 + k + j / 1024d0
                ring%f0_right(j, k) = 64d0 + k + j / 1024d0
            enddo
            do i = 1, nx
                ring%up2(i, k) = 80d0 + k + i / 1024d0
                ring%f0_top(i, k) = 96d0 + k + i / 1024d0
                ring%down2(i, k) = 112d0 + k + i / 1024d0
                ring%f0_bottom(i, k) = 128d0 + k + i / 1024d0
            enddo
        enddo
        allocate(rel(w, 11))
        do k = 1, 11
            do i = 1, w
                rel(i, k) = dble(k) + dble(i) / 8d0
            enddo
        enddo
        call write_checkpoint(path, cycle, w, rel, movenum, nx, ny, u, v, f, ring, ok)
        if (.not. ok) stop 1
    else
        call get_command_argument(3, path2)
        call get_command_argument(4, arg); read(arg, *) nx
        call get_command_argument(5, arg); read(arg, *) ny
        call get_command_argument(6, arg); read(arg, *) w
        allocate(u(nx, ny), v(nx, ny), f(nx, ny, 9), rel(w, 11))
        call read_checkpoint(path, w, nx, ny, cycle, rel, movenum, u, v, f, ring, msg, ok)
        if (.not. ok) then
            write(*, '(a)') trim(msg)
            stop 1
        endif
        call write_checkpoint(path2, cycle, w, rel, movenum, nx, ny, u, v, f, ring, ok)
        if (.not. ok) stop 1
    endif
end program checkpoint_check
