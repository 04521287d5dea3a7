! lbm_front.f90 - the online command front end of the host: one ZeroMQ REP
! socket, text commands in, one text reply per command.
!
! Wire protocol (zmq_comm.f90:151-281, unchanged so that existing clients
! work):
!   RESET|START [t]          new episode from t (only t = 0 is accepted)
!   STEP t d(1..3W)          advance to time t; per-wing (domega, dxv, dyv)
!   SAVE|OUTPUT              write the flow snapshot
!   STOP|EXIT                end the process
! Replies: "FORCE <n> v1 .. vn" with each value in es24.16 form,
! left-adjusted; "STATUS <text>"; "ERROR <text>".  Commands are read with
! list-directed input, so blanks, commas and d-exponents are accepted, and
! keywords are case-insensitive.  Messages are at most 8192 bytes.
!
! libzmq is bound through bind(C) interfaces to its C API (zmq.h of libzmq
! 4.x: ZMQ_REP = 4, ZMQ_LINGER = 17).
module lbm_front
    use iso_c_binding
    implicit none
    private

    public :: front_open, front_close, front_next_command, parse_command
    public :: front_reply_force, front_reply_status, front_reply_error, format_force_reply
    public :: CMD_INVALID, CMD_RESET, CMD_STEP, CMD_STOP, CMD_OUTPUT, MESSAGE_BYTES

    integer, parameter :: CMD_INVALID = 0, CMD_RESET = 1, CMD_STEP = 2, CMD_STOP = 3, CMD_OUTPUT = 4
    integer, parameter :: MESSAGE_BYTES = 8192
    integer(c_int), parameter :: ZMQ_REP = 4_c_int, ZMQ_LINGER = 17_c_int

    type(c_ptr), save :: ctx = c_null_ptr, sock = c_null_ptr

    interface
        type(c_ptr) function zmq_ctx_new() bind(C, name='zmq_ctx_new')
            import :: c_ptr
        end function
        integer(c_int) function zmq_ctx_term(context) bind(C, name='zmq_ctx_term')
            import :: c_ptr, c_int
            type(c_ptr), value :: context
        end function
        type(c_ptr) function zmq_socket(context, kind) bind(C, name='zmq_socket')
            import :: c_ptr, c_int
            type(c_ptr), value :: context
            integer(c_int), value :: kind
        end function
        integer(c_int) function zmq_setsockopt(s, option, value, length) bind(C, name='zmq_setsockopt')
            import :: c_ptr, c_int, c_size_t
            type(c_ptr), value :: s, value
            integer(c_int), value :: option
            integer(c_size_t), value :: length
        end function
        integer(c_int) function zmq_bind(s, endpoint) bind(C, name='zmq_bind')
            import :: c_ptr, c_int, c_char
            type(c_ptr), value :: s
            character(kind=c_char), intent(in) :: endpoint(*)
        end function
        integer(c_int) function zmq_send(s, buf, length, flags) bind(C, name='zmq_send')
            import :: c_ptr, c_int, c_size_t, c_char
            type(c_ptr), value :: s
            character(kind=c_char), intent(in) :: buf(*)
            integer(c_size_t), value :: length
            integer(c_int), value :: flags
        end function
        integer(c_int) function zmq_recv(s, buf, length, flags) bind(C, name='zmq_recv')
            import :: c_ptr, c_int, c_size_t, c_char
            type(c_ptr), value :: s
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_size_t), value :: length
            integer(c_int), value :: flags
        end function
        integer(c_int) function zmq_close(s) bind(C, name='zmq_close')
            import :: c_ptr, c_int
            type(c_ptr), value :: s
        end function
    end interface

contains

    ! Create the context and the REP socket, set LINGER 0 and bind to
    ! endpoint.  status: 0 ok, 1 context, 2 socket, 3 option, 4 bind
    ! (zmq_comm.f90:84-125 reports the same stages).
    subroutine front_open(endpoint, status)
        character(len=*), intent(in) :: endpoint
        integer, intent(out) :: status
        integer(c_int), target :: linger
        character(kind=c_char) :: ep(len_trim(endpoint) + 1)
        integer :: i

        status = 0
        ctx = zmq_ctx_new()
        if (.not. c_associated(ctx)) then
            status = 1
            return
        endif
        sock = zmq_socket(ctx, ZMQ_REP)
        if (.not. c_associated(sock)) then
            status = 2
            return
        endif
        linger = 0_c_int
        if (zmq_setsockopt(sock, ZMQ_LINGER, c_loc(linger), c_sizeof(linger)) /= 0) then
            status = 3
            return
        endif
        do i = 1, len_trim(endpoint)
            ep(i) = endpoint(i:i)
        enddo
        ep(len_trim(endpoint) + 1) = c_null_char
        if (zmq_bind(sock, ep) /= 0) status = 4
    end subroutine

    subroutine front_close()
        integer(c_int) :: rc
        if (c_associated(sock)) rc = zmq_close(sock)
        if (c_associated(ctx)) rc = zmq_ctx_term(ctx)
        sock = c_null_ptr
        ctx = c_null_ptr
    end subroutine

    ! Block for the next message and parse it.  A message that does not
    ! parse is answered here with its ERROR reply and ok = .false.; the caller
    ! then waits for the next one (zmq_comm.f90:151-181).
    subroutine front_next_command(wing_count, cmd, row, ok)
        integer, intent(in) :: wing_count
        integer, intent(out) :: cmd
        real(c_double), intent(out) :: row(:)
        logical, intent(out) :: ok
        character(len=MESSAGE_BYTES) :: message
        character(len=256) :: error_text
        character(kind=c_char) :: buf(MESSAGE_BYTES)
        integer(c_int) :: got
        integer :: i

        cmd = CMD_INVALID
        row = 0.0d0
        ok = .false.
        got = zmq_recv(sock, buf, int(MESSAGE_BYTES, c_size_t), 0_c_int)
        if (got < 0) return
        message = ' '
        do i = 1, min(int(got), MESSAGE_BYTES)
            message(i:i) = buf(i)
        enddo
        call parse_command(message, wing_count, cmd, row, ok, error_text)
        if (.not. ok) call front_reply_error(trim(error_text))
    end subroutine

    ! parse_command_message semantics (zmq_comm.f90:225-281): the keyword is
    ! the first list-directed item; RESET takes an optional time, STEP takes
    ! exactly 1 + 3*wing_count values into row(1:1+3W).
    subroutine parse_command(message, wing_count, cmd, row, ok, error_text)
        character(len=*), intent(in) :: message
        integer, intent(in) :: wing_count
        integer, intent(out) :: cmd
        real(c_double), intent(out) :: row(:)
        logical, intent(out) :: ok
        character(len=*), intent(out) :: error_text
        character(len=32) :: key
        integer :: ios, i, code, nvals

        cmd = CMD_INVALID
        row = 0.0d0
        ok = .false.
        error_text = ''
        key = ''
        read(message, *, iostat=ios) key
        if (ios /= 0) then
            error_text = '无法解析 ZeroMQ 指令'
            return
        endif
        do i = 1, len_trim(key)
            code = iachar(key(i:i))
            if (code >= iachar('a') .and. code <= iachar('z')) key(i:i) = achar(code - 32)
        enddo

        select case (trim(key))
        case ('RESET', 'START')
            if (len_trim(message) /= len_trim(key)) then
                read(message, *, iostat=ios) key, row(1)
                if (ios /= 0) then
                    error_text = 'RESET 指令缺少初始时间'
                    return
                endif
            endif
            cmd = CMD_RESET
        case ('STEP')
            nvals = 1 + 3 * wing_count
            read(message, *, iostat=ios) key, row(1:nvals)
            if (ios /= 0) then
                error_text = 'STEP 指令缺少目标时刻或速度增量'
                return
            endif
            cmd = CMD_STEP
        case ('SAVE', 'OUTPUT')
            cmd = CMD_OUTPUT
        case ('STOP', 'EXIT')
            cmd = CMD_STOP
        case default
            error_text = '未知 ZeroMQ 指令: ' // trim(key)
            return
        end select
        ok = .true.
    end subroutine

    ! "FORCE <n> v1 .. vn", values in es24.16, left-adjusted
    ! (zmq_comm.f90:183-203).
    subroutine format_force_reply(values, n, message, length)
        real(c_double), intent(in) :: values(*)
        integer, intent(in) :: n
        character(len=*), intent(out) :: message
        integer, intent(out) :: length
        character(len=32) :: tok
        integer :: i

        write(tok, '(i0)') n
        message = 'FORCE ' // trim(tok)
        length = len_trim(message)
        do i = 1, n
            write(tok, '(es24.16)') values(i)
            tok = adjustl(tok)
            message(length + 1:) = ' ' // trim(tok)
            length = length + 1 + len_trim(tok)
        enddo
    end subroutine

    subroutine front_reply_force(values, n)
        real(c_double), intent(in) :: values(*)
        integer, intent(in) :: n
        character(len=MESSAGE_BYTES) :: message
        integer :: length
        call format_force_reply(values, n, message, length)
        call send_text(message(1:length))
    end subroutine

    subroutine front_reply_status(text)
        character(len=*), intent(in) :: text
        call send_text('STATUS ' // trim(text))
    end subroutine

    subroutine front_reply_error(text)
        character(len=*), intent(in) :: text
        call send_text('ERROR ' // trim(text))
    end subroutine

    subroutine send_text(text)
        character(len=*), intent(in) :: text
        character(kind=c_char) :: buf(max(1, len_trim(text)))
        integer(c_int) :: rc
        integer :: i, n

        n = len_trim(text)
        if (n <= 0) return
        do i = 1, n
            buf(i) = text(i:i)
        enddo
        rc = zmq_send(sock, buf, int(n, c_size_t), 0_c_int)
    end subroutine

end module lbm_front
