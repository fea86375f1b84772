! Test harness (not reference code): evaluates the reference interaction module on
! inputs read from stdin and prints the results bit-exactly (IEEE-754 bits as hex).
! stdin: n, then n lines "d1 d2 d3 r_sq eps sigma" given as 16-digit hex bit patterns.
! stdout per line: f1 f2 f3 e_raw cut_energy (hex), where e_raw = lj_energy(r_sq,eps,sigma,0)
! and cut_energy comes from lj_params_init with cut = sqrt(r_sq) (1x1 parameter set).
program interaction_drv
  use interaction
  use iso_fortran_env, only: int64, real64
  implicit none
  integer :: n, i
  integer(int64) :: b(6)
  real(real64) :: v(6), f(3), e
  type(lj_params_t) :: lj
  read(*, *) n
  do i = 1, n
     read(*, '(6(z16,1x))') b
     v = transfer(b, v)
     f = lj_force(v(1:3), v(4), v(5), v(6))
     e = lj_energy(v(4), v(5), v(6), 0.d0)
     call lj%init(reshape([v(5)], [1, 1]), reshape([v(6)], [1, 1]), reshape([sqrt(v(4))], [1, 1]))
     write(*, '(5(z16.16,1x))') transfer(f(1), 1_int64), transfer(f(2), 1_int64), &
          transfer(f(3), 1_int64), transfer(e, 1_int64), transfer(lj%cut_energy(1,1), 1_int64)
  end do
end program interaction_drv
