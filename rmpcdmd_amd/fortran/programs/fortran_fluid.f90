! A pure SRD fluid through the Fortran module API (the collide / stream / shift / sort
! loop of test/setup_simple_fluid.f90): Maxwell-Boltzmann velocities without net
! momentum, then collisions.  Checks: total momentum stays zero to rounding, the SRD
! temperature stays at T within 5%.  Exit status 0 on success.
program fortran_fluid
  use rmpcdmd_module
  use threefry_module
  use rmpcd_capi, only: rmpcd_init, check
  implicit none
  type(cell_system_t) :: cells
  type(particle_system_t) :: p
  type(threefry_rng_t), allocatable :: state(:)
  integer :: L(3), N, i, k
  double precision :: v_com(3), te, tau
  call check(rmpcd_init(0))
  allocate(state(1))
  call threefry_rng_init(state, 7_8)
  L = [16, 12, 10]
  N = 10*L(1)*L(2)*L(3)
  tau = 0.1d0
  call p%init(N)
  do i = 1, N
     do k = 1, 3
        p%vel(k, i) = threefry_normal(state(1))
     end do
  end do
  v_com = sum(p%vel, dim=2)/N
  p%vel = p%vel - spread(v_com, 2, N)
  p%force = 0
  p%species = 1
  call cells%init(L, 1.d0)
  call p%random_placement(cells%edges, state=state(1))
  call p%sort(cells)
  do i = 1, 50
     call simple_mpcd_step(p, cells, state)
     call mpcd_stream_periodic(p, cells, tau)
     call cells%random_shift(state(1))
     call p%sort(cells)
  end do
  v_com = sum(p%vel, dim=2)/N
  te = compute_temperature(p, cells)
  write(*, '(a,3es12.3,a,f8.4)') 'fortran_fluid: v_com', v_com, '  T', te
  if (maxval(abs(v_com)) > 1d-12) error stop 'momentum not conserved'
  if (abs(te - 1) > 0.05d0) error stop 'temperature off'
  write(*, '(a)') 'fortran_fluid ok'
end program fortran_fluid
