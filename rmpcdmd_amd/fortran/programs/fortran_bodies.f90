! Linked colloids through the Fortran module API:
!  - the known answer of test/test_elastic_network_0.f90 (energy and forces of three
!    links from particle 1, one across the periodic box);
!  - a tetrahedron held by RATTLE as in test/setup_simple_rattle.f90: free flight with
!    md_pos, rattle_body_pos, then rattle_body_vel; the six bond lengths stay put, the
!    momentum is kept;
!  - a rigid body of six beads spinning freely (rigid_body_t with quaternions): its
!    angular momentum stays fixed and the beads' distances to the centre are kept.
! Exit status 0 on success.
program fortran_bodies
  use rmpcdmd_module
  use rmpcd_capi, only: rmpcd_init, check
  implicit none
  type(particle_system_t) :: p, t, b
  type(rigid_body_t) :: body
  integer :: links(2,3), tl(2,6), i, j
  double precision :: L(3), d(3), e, td(6), p0(3), p1(3), r(3), L0(3), r0(6), dt
  call check(rmpcd_init(0))

  ! test_elastic_network_0
  call p%init(4)
  L = 5
  p%force = 0
  p%pos = 1
  p%pos(1, 2) = 3
  p%pos(2, 3) = 2
  p%pos(1, 4) = 4.5d0
  p%species = 1
  links(:, 1) = [1, 2]
  links(:, 2) = [1, 3]
  links(:, 3) = [1, 4]
  d = [0.5d0, 1.d0, 1.d0]
  e = elastic_network(p, links, d, 7.d0, L)
  write(*, '(a,f12.8)') 'fortran_bodies: elastic energy', e
  if (abs(e - (1.5d0**2 + 0.5d0**2)*7/2) > 1d-12) error stop 'elastic energy'
  if (maxval(abs(p%force(:, 1) - [7.d0, 0.d0, 0.d0])) > 1d-12) error stop 'elastic force'
  if (maxval(abs(sum(p%force, dim=2))) > 1d-12) error stop 'elastic net force'

  ! a tetrahedron held by RATTLE
  call t%init(4, 1, [3.d0])
  L = 12
  t%species = 1
  t%pos(:, 1) = [11.5d0, 6.d0, 6.d0]
  t%pos(:, 2) = [12.5d0, 6.d0, 6.d0]  ! across the x face once folded
  t%pos(:, 3) = [12.d0, 6.9d0, 6.d0]
  t%pos(:, 4) = [12.d0, 6.3d0, 6.8d0]
  t%pos = modulo(t%pos, 12.d0)
  tl = reshape([1,2, 1,3, 1,4, 2,3, 2,4, 3,4], [2, 6])
  do i = 1, 6
     r = t%pos(:, tl(1, i)) - t%pos(:, tl(2, i))
     r = r - L*nint(r/L)
     td(i) = norm2(r)
  end do
  t%vel = reshape([0.3d0, -0.1d0, 0.2d0, -0.2d0, 0.4d0, 0.1d0, 0.1d0, 0.1d0, -0.3d0, 0.d0, -0.2d0, 0.4d0], [3, 4])
  t%force = 0
  p0 = sum(t%vel, dim=2)
  dt = 0.05d0
  do j = 1, 50
     t%pos_rattle = t%pos
     t%pos = t%pos + dt*t%vel
     call rattle_body_pos(t, tl, td, dt, L, 1d-12)
     call apply_pbc(t, L)
     call rattle_body_vel(t, tl, td, dt, L, 1d-12)
  end do
  do i = 1, 6
     r = t%pos(:, tl(1, i)) - t%pos(:, tl(2, i))
     r = r - L*nint(r/L)
     if (abs(norm2(r) - td(i)) > 1d-10) error stop 'rattle bond length'
  end do
  p1 = sum(t%vel, dim=2)
  write(*, '(a,3es12.3)') 'fortran_bodies: rattle momentum change', p1 - p0
  if (maxval(abs(p1 - p0)) > 1d-12) error stop 'rattle momentum'

  ! a free rigid body
  call b%init(6, 2, [1.d0, 2.d0])
  L = 20
  b%species = [1, 1, 2, 2, 1, 1]
  b%pos = reshape([1.d0, 0.d0, 0.d0, -1.d0, 0.d0, 0.d0, 0.d0, 1.5d0, 0.d0, 0.d0, -1.5d0, 0.d0, &
       0.d0, 0.d0, 2.d0, 0.d0, 0.d0, -2.d0], [3, 6])
  b%pos = b%pos + 10
  b%vel = 0
  b%force = 0
  b%image = 0
  call body%init(b, 1, 6, L)
  body%L = [0.4d0, 1.1d0, -0.7d0]
  L0 = body%L
  do i = 1, 6
     r0(i) = norm2(body%pos_body(:, i))
  end do
  do j = 1, 100
     call body%compute_force_torque(b)
     call body%vv1(b, 0.01d0, 1d-12, L)
     call body%vv2(b, 0.01d0)
  end do
  write(*, '(a,4f10.6)') 'fortran_bodies: rigid body q', body%q
  if (maxval(abs(body%L - L0)) > 1d-14) error stop 'rigid body L'
  if (abs(norm2(body%q) - 1) > 1d-14) error stop 'rigid body |q|'
  do i = 1, 6
     r = b%pos(:, i) - body%pos
     r = r - L*nint(r/L)
     if (abs(norm2(r) - r0(i)) > 1d-12) error stop 'rigid body shape'
  end do
  write(*, '(a)') 'fortran_bodies ok'
end program fortran_bodies
