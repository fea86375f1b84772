! Test harness (not reference code): prints the reference hilbert module's keys so the
! C oracle can be checked against it.  Usage: hilbert_drv M1 M2 M3 mode
!   mode 0: compact_p_to_h for every p in [0,2^M1)x[0,2^M2)x[0,2^M3) (x fastest),
!           then compact_h_to_p for every h in [0, 2^(M1+M2+M3)), as z-hex integers
!   mode 1: p_to_h / h_to_p with m = M1 (uncompacted cube of side 2^M1)
program hilbert_drv
  use hilbert
  implicit none
  integer :: M(3), p(3), q(3), mode, h, i, j, k, n
  character(len=32) :: arg
  do i = 1, 3
     call get_command_argument(i, arg)
     read(arg, *) M(i)
  end do
  call get_command_argument(4, arg)
  read(arg, *) mode
  if (mode == 0) then
     do k = 0, 2**M(3) - 1
        do j = 0, 2**M(2) - 1
           do i = 0, 2**M(1) - 1
              p = [i, j, k]
              h = compact_p_to_h(p, M)
              write(*, '(i0)') h
           end do
        end do
     end do
     n = 2**(M(1)+M(2)+M(3))
     do h = 0, n - 1
        q = compact_h_to_p(h, M)
        write(*, '(i0,1x,i0,1x,i0)') q
     end do
  else
     do k = 0, 2**M(1) - 1
        do j = 0, 2**M(1) - 1
           do i = 0, 2**M(1) - 1
              p = [i, j, k]
              h = p_to_h(p, M(1))
              write(*, '(i0)') h
           end do
        end do
     end do
     n = 2**(3*M(1))
     do h = 0, n - 1
        q = h_to_p(h, M(1))
        write(*, '(i0,1x,i0,1x,i0)') q
     end do
  end if
end program hilbert_drv
