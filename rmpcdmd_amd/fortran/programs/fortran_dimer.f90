! A few steps of the single-dimer loop (programs/single_dimer_pbc.f90:306-383) written
! against the Fortran module API: velocity Verlet, RATTLE, neighbour list, LJ forces.
! Checks: the bond length stays d, total momentum (solvent + dimer) is conserved.
program fortran_dimer
  use rmpcdmd_module
  use threefry_module
  use rmpcd_capi, only: rmpcd_init, check
  implicit none
  type(cell_system_t) :: cells
  type(particle_system_t) :: solvent, colloids
  type(neighbor_list_t) :: neigh
  type(lj_params_t) :: slj, clj
  type(threefry_rng_t) :: state(1)
  integer :: L(3), N, i, k, j
  double precision :: sigma, d, dt, e1, e2, skin, mass(2), eps(2,2), sig(2,2), p0(3), p1(3), r(3)
  double precision :: so_max, co_max
  call check(rmpcd_init(0))
  call threefry_rng_init(state, 11_8)
  L = [16, 16, 16]
  N = 10*L(1)*L(2)*L(3)
  sigma = 2
  d = 4.5d0
  dt = 0.01d0
  skin = 2
  eps = 1
  sig = sigma
  call slj%init(eps, sig, sig*2**(1.d0/6.d0))
  sig = 2*sigma
  call clj%init(eps, sig, sig*2**(1.d0/6.d0))
  mass = 10*sigma**3*4*3.14159265/3
  call solvent%init(N, 2)
  call colloids%init(2, 2, mass)
  colloids%species = [1, 2]
  colloids%vel = 0
  colloids%force = 0
  call cells%init(L, 1.d0)
  colloids%pos(:, 1) = cells%edges/2
  colloids%pos(:, 2) = cells%edges/2
  colloids%pos(1, 2) = colloids%pos(1, 2) + d
  do i = 1, N
     do k = 1, 3
        solvent%vel(k, i) = threefry_normal(state(1))
     end do
  end do
  solvent%vel = solvent%vel - spread(sum(solvent%vel, dim=2)/N, 2, N)
  solvent%force = 0
  solvent%species = 1
  call solvent%random_placement(cells%edges, colloids, slj, state(1))
  call solvent%sort(cells)
  call neigh%init(2, 4000)
  call neigh%make_stencil(cells, sigma*2**(1.d0/6.d0) + skin)
  call neigh%update_list(colloids, solvent, sigma*2**(1.d0/6.d0) + skin, cells, slj)
  e1 = compute_force(colloids, solvent, neigh, cells%edges, slj)
  e2 = compute_force_n2(colloids, cells%edges, clj)
  solvent%force_old = solvent%force
  colloids%force_old = colloids%force
  solvent%pos_old = solvent%pos
  colloids%pos_old = colloids%pos
  p0 = sum(solvent%vel, dim=2) + mass(1)*colloids%vel(:, 1) + mass(2)*colloids%vel(:, 2)
  do j = 1, 20
     call md_pos(solvent, dt)
     colloids%pos_rattle = colloids%pos
     do k = 1, 2
        colloids%pos(:, k) = colloids%pos(:, k) + dt*colloids%vel(:, k) + dt**2*colloids%force(:, k)/(2*mass(k))
     end do
     call rattle_dimer_pos(colloids, d, dt, cells%edges)
     so_max = solvent%maximum_displacement()
     co_max = colloids%maximum_displacement()
     if (co_max >= skin*0.1d0 .or. so_max >= skin*0.89d0) then
        call apply_pbc(solvent, cells%edges)
        call apply_pbc(colloids, cells%edges)
        call solvent%sort(cells)
        call neigh%update_list(colloids, solvent, sigma*2**(1.d0/6.d0) + skin, cells, slj)
        solvent%pos_old = solvent%pos
        colloids%pos_old = colloids%pos
     end if
     call switch(solvent%force, solvent%force_old)
     call switch(colloids%force, colloids%force_old)
     solvent%force = 0
     colloids%force = 0
     e1 = compute_force(colloids, solvent, neigh, cells%edges, slj)
     e2 = compute_force_n2(colloids, cells%edges, clj)
     call md_vel(solvent, dt)
     do k = 1, 2
        colloids%vel(:, k) = colloids%vel(:, k) + dt*(colloids%force(:, k) + colloids%force_old(:, k))/(2*mass(k))
     end do
     call rattle_dimer_vel(colloids, d, dt, cells%edges)
  end do
  p1 = sum(solvent%vel, dim=2) + mass(1)*colloids%vel(:, 1) + mass(2)*colloids%vel(:, 2)
  r = rel_pos(colloids%pos(:, 1), colloids%pos(:, 2), cells%edges)
  write(*, '(a,3es12.3,a,f12.8)') 'fortran_dimer: momentum drift', p1 - p0, '  |r12|', norm2(r)
  if (maxval(abs(p1 - p0)) > 1d-9) error stop 'momentum not conserved'
  if (abs(norm2(r) - d) > 1d-9) error stop 'bond length not kept'
  write(*, '(a)') 'fortran_dimer ok'
end program fortran_dimer
