! Forced flow between walls through the Fortran module API: the loop of
! programs/poiseuille_flow.f90 (stream with the x force and bounce-back z walls, md_vel,
! apply_pbc, grid shift, sort, wall collision with virtual wall particles), then the
! thermostat variants (MPC-AT in wall_mpcd_step and simple_mpcd_step, rescale_cells,
! bulk_reaction_endothermic).  Checks: the flow develops along x and is faster in the
! middle of the channel than next to the walls, the temperature stays near T, each
! thermostat keeps the total momentum where it should, and the endothermic reaction
! converts particles.  Exit status 0 on success.
program fortran_walls
  use rmpcdmd_module
  use threefry_module
  use rmpcd_capi, only: rmpcd_init, check
  implicit none
  type(cell_system_t) :: cells
  type(particle_system_t) :: p
  type(threefry_rng_t), allocatable :: state(:)
  integer :: L(3), N, i, k, n_mid, n_edge, n_conv
  double precision :: tau, g, te, wall_t(2), wall_v(3,2), v_mid, v_edge, p0(3), p1(3)
  call check(rmpcd_init(0))
  allocate(state(1))
  call threefry_rng_init(state, 11_8)
  L = [12, 6, 10]
  N = 10*L(1)*L(2)*L(3)
  tau = 0.1d0
  g = 0.02d0
  call p%init(N)
  do i = 1, N
     do k = 1, 3
        p%vel(k, i) = threefry_normal(state(1))
     end do
  end do
  p%vel = p%vel - spread(sum(p%vel, dim=2)/N, 2, N)
  p%force = 0
  p%species = 1
  call p%random_placement(L*1.d0, state=state(1))
  call cells%init(L, 1.d0, has_walls=.true.)
  cells%origin(3) = -0.5d0
  call p%sort(cells)
  p%force(1,:) = g
  p%force_old(1,:) = g
  wall_v = 0
  wall_t = 1
  cells%bc = [PERIODIC_BC, PERIODIC_BC, BOUNCE_BACK_BC]
  do i = 1, 300
     call mpcd_stream_xforce_yzwall(p, cells, tau, g)
     call md_vel(p, tau)
     call apply_pbc(p, cells%edges)
     call cells%random_shift(state(1))
     call p%sort(cells)
     call wall_mpcd_step(p, cells, state, wall_temperature=wall_t, wall_v=wall_v, wall_n=[10, 10])
  end do
  te = compute_temperature(p, cells)
  v_mid = 0
  v_edge = 0
  n_mid = 0
  n_edge = 0
  do i = 1, N
     if (abs(p%pos(3, i) - L(3)/2.d0) < 2) then
        v_mid = v_mid + p%vel(1, i)
        n_mid = n_mid + 1
     else if (p%pos(3, i) < 1.5d0 .or. p%pos(3, i) > L(3) - 1.5d0) then
        v_edge = v_edge + p%vel(1, i)
        n_edge = n_edge + 1
     end if
  end do
  v_mid = v_mid / n_mid
  v_edge = v_edge / n_edge
  write(*, '(a,f8.4,a,f8.4,a,f8.4)') 'fortran_walls: vx middle', v_mid, '  vx walls', v_edge, '  T', te
  if (.not. (v_mid > 0.05d0 .and. v_mid > 1.5d0*v_edge)) error stop 'no channel flow'
  if (abs(te - 1) > 0.1d0) error stop 'temperature off'

  ! the thermostat branch of the wall collision keeps the cell means (keep_cell_v)
  p0 = sum(p%vel, dim=2)
  call wall_mpcd_step(p, cells, state, thermostat=.true., bulk_temperature=1.d0, keep_cell_v=.true.)
  p1 = sum(p%vel, dim=2)
  if (maxval(abs(p1 - p0)) > 1d-8*N) error stop 'wall thermostat moved the total momentum'
  ! MPC-AT in simple_mpcd_step, with hydro: momentum kept
  call simple_mpcd_step(p, cells, state, thermostat=.true., T=1.d0, hydro=.true.)
  p0 = sum(p%vel, dim=2)
  if (maxval(abs(p0 - p1)) > 1d-8*N) error stop 'MPC-AT moved the total momentum'
  ! rescale_cells: the SRD temperature becomes T
  call rescale_cells(p, cells, state, 0.8d0)
  te = compute_temperature(p, cells)
  write(*, '(a,f10.6)') 'fortran_walls: T after rescale_cells', te
  if (abs(te - 0.8d0) > 1d-9) error stop 'rescale_cells'
  ! bulk_reaction_endothermic converts some A into B
  cells%is_reac = .true.
  call bulk_reaction_endothermic(p, cells, 1, 2, 1.d0, tau, state, 0.5d0)
  n_conv = count(p%species == 2)
  write(*, '(a,i8)') 'fortran_walls: converted', n_conv
  if (n_conv == 0) error stop 'no endothermic conversion'
  write(*, '(a)') 'fortran_walls ok'
end program fortran_walls
