! rmpcdmd_module -- the module the reference's programs use (src/rmpcdmd_module.f90.in:1-19):
! every module of the time step, and the particle output.  The correlator, polar and
! planar field and quaternion modules of the reference are not part of this library.
module rmpcdmd_module
  use common
  use hilbert
  use interaction
  use cell_system
  use particle_system
  use particle_system_io
  use neighbor_list
  use mpcd
  use md
  implicit none
  character(len=*), parameter :: RMPCDMD_REVISION = 'rmpcdmd-amd 0.1 (gfx950)'
end module rmpcdmd_module
