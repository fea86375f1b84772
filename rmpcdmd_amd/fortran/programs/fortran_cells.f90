! Cell sort through the Fortran module API on the GPU, checked the way
! test/test_cell_system_0.f90 checks it: cell_start(1) = 1, the last padded cell starts
! at N+1, the counts add up to N, and every sorted particle lies in the cell whose
! compact Hilbert index it is filed under.  Exit status 0 on success.
program fortran_cells
  use rmpcdmd_module
  use threefry_module
  use rmpcd_capi, only: rmpcd_init, check
  implicit none
  type(cell_system_t) :: cells
  type(particle_system_t) :: p
  type(threefry_rng_t) :: state(1)
  integer, parameter :: N = 200000
  integer :: L(3), i, h, c1(3), c2(3), nbad
  call check(rmpcd_init(0))
  call threefry_rng_init(state, 20261202_8)
  L = [60, 30, 24]
  call cells%init(L, 1.d0)
  call p%init(N)
  p%species = 1
  call p%random_placement(cells%edges, state=state(1))
  call cells%random_shift(state(1))
  call p%sort(cells)
  if (cells%cell_start(1) /= 1) error stop 'cell_start(1) /= 1'
  if (cells%cell_start(cells%N) /= N + 1) error stop 'last cell_start /= N+1'
  if (sum(cells%cell_count) /= N) error stop 'counts do not add up'
  nbad = 0
  h = 1
  do i = 1, N
     do while (i >= cells%cell_start(h) + cells%cell_count(h))
        h = h + 1
     end do
     c1 = cells%cartesian_indices(p%pos(:, i))
     c2 = compact_h_to_p(h - 1, cells%M)
     if (any(c1 /= c2)) nbad = nbad + 1
     if (p%id_to_idx(p%id(i)) /= i) nbad = nbad + 1
  end do
  if (nbad /= 0) error stop 'particles filed in the wrong cell'
  write(*, '(a,i0,a)') 'fortran_cells ok: ', N, ' particles sorted on the GPU'
end program fortran_cells
