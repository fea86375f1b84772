! fortran_shims -- the host-side shims on their own (no GPU): writes an H5MD text file
! through h5md_module/hdf5 the way the reference's set-up programs do
! (test/setup_simple_fluid.f90:65-79, 119-123), reads it back with
! particle_system_t%init_from_file (src/particle_system.f90:186-248), bins a profile_t
! (src/common.f90:180-225), and checks the tester's own bookkeeping.  Run in a scratch
! directory: it writes shim_check.h5.txt there.
program fortran_shims
  use hdf5
  use h5md_module
  use tester
  use common
  use particle_system
  implicit none
  type(tester_t) :: test, t2
  type(h5md_file_t) :: f
  type(h5md_element_t) :: e_pos, e_n, elem
  integer(HID_T) :: g, g2, box_group
  type(particle_system_t) :: p, q
  type(profile_t) :: prof
  double precision :: x(3, 5)
  integer :: sp(5), i, error

  call test% init()
  call h5open_f(error)
  x = reshape([(dble(i)*0.5d0 + 1.d0/3.d0, i = 1, 15)], [3, 5])
  sp = [1, 2, 1, 2, 2]

  call f% create('shim_check.h5', 'RMPCDMD:fortran_shims', '1.0', 'nobody')
  call h5gcreate_f(f% particles, 'fluid', g, error)
  call test% assert_equal(error, 0)
  call h5gcreate_f(g, 'box', box_group, error)
  call h5md_write_attribute(box_group, 'dimension', 3)
  call h5md_write_attribute(box_group, 'boundary', ['periodic', 'periodic', 'periodic'])
  call elem% create_fixed(box_group, 'edges', [8.d0, 5.d0, 5.d0])
  call h5gclose_f(box_group, error)
  call e_pos% create_time(g, 'position', x, ior(H5MD_TIME, H5MD_STORE_TIME))
  do i = 1, 3
     call e_pos% append(x + i, i, i*0.1d0)
  end do
  call e_pos% close()
  call h5gclose_f(g, error)
  call h5gcreate_f(f% particles, 'still', g2, error)
  call elem% create_fixed(g2, 'position', x)
  call elem% create_fixed(g2, 'species', sp)
  call h5gclose_f(g2, error)
  call e_n% create_time(f% observables, 'n', 5, H5MD_LINEAR, step=10, time=0.5d0)
  do i = 1, 4
     call e_n% append(5 + i)
  end do
  call e_n% close()
  call f% close()

  ! frame 1 (0-based) of the time-dependent positions, then the fixed set with species
  call p% init_from_file('shim_check.h5', 'fluid', H5MD_TIME, 1)
  call test% assert_equal(p% Nmax, 5)
  call test% assert_close(reshape(p% pos, [15]), reshape(x + 2, [15]))
  call q% init_from_file('shim_check.h5', 'still', H5MD_FIXED)
  call test% assert_equal(q% Nmax, 5)
  call test% assert_equal(q% species, sp)
  call test% assert_close(reshape(q% pos, [15]), reshape(x, [15]))

  ! profile_t: bin floor((x - xmin)/dx) + 1, out-of-range values dropped
  call prof% init(0.d0, 1.d0, 4)
  call prof% bin(0.3d0, 2.d0)
  call prof% bin(0.26d0, 1.d0)
  call prof% bin(1.5d0, 7.d0)
  call prof% bin(-0.01d0, 7.d0)
  call prof% bin(0.99d0, 3.d0)
  call test% assert_equal(prof% count, [0, 2, 0, 1])
  call test% assert_close(prof% data, [0.d0, 3.d0, 0.d0, 3.d0])
  call prof% norm()
  call test% assert_close(prof% data, [0.d0, 1.5d0, 0.d0, 3.d0])
  call test% assert_close(cross([1.d0, 0.d0, 0.d0], [0.d0, 1.d0, 0.d0]), [0.d0, 0.d0, 1.d0])

  ! the tester counts failures: one of each kind, and the passes around them
  call t2% init(tolerance64=1d-12)
  call t2% assert_equal(1, 2)
  call t2% assert_equal(.true., .false.)
  call t2% assert_close(1.d0, 1.d0 + 1d-9)
  call t2% assert_close([1.d0, 2.d0], [1.d0, 2.d0 + 1d-13])
  call t2% assert_positive(-1.d0)
  call t2% assert_positive([0, 1])
  call test% assert_equal(t2% n_tests, 6)
  call test% assert_equal(t2% n_errors, 4)
  call t2% print(errorstop=.false.)

  call test% print()
end program fortran_shims
