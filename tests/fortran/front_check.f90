! Test driver of the host's command front end (module lbm_front), no GPU
! work: reads the wing count, then one message per line from stdin.  A line
! "#FORCE n v1 .. vn" prints the FORCE reply of those values; any other line
! is parsed as a command and prints "<cmd> <ok> <error text>" and the parsed
! row (1 + 3W values, es24.16).
program front_check
    use iso_c_binding
    use lbm_front
    implicit none
    character(len=MESSAGE_BYTES) :: line, msg
    character(len=256) :: err
    character(len=32) :: tok
    real(c_double) :: row(73), vals(200)
    integer :: ios, cmd, w, i, n, length
    logical :: ok

    read(*, *) w
    do
        read(*, '(a)', iostat=ios) line
        if (ios /= 0) exit
        if (line(1:6) == '#FORCE') then
            read(line(7:), *) n
            read(line(7:), *) n, vals(1:n)
            call format_force_reply(vals, n, msg, length)
            write(*, '(a)') msg(1:length)
        else
            call parse_command(line, w, cmd, row, ok, err)
            write(*, '(i0,1x,l1,1x,a)') cmd, ok, trim(err)
            msg = ''
            do i = 1, 1 + 3 * w
                write(tok, '(es24.16)') row(i)
                msg = trim(msg) // ' ' // trim(adjustl(tok))
            enddo
            write(*, '(a)') trim(msg)
        endif
    enddo
end program front_check
