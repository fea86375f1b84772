! The single_dimer_pbc driver with the reference's command line and parameter files:
!
!     single_dimer_pbc input.parameters output seed
!
! Parameters are read with ParseText under the reference's names
! (programs/single_dimer_pbc.f90:109-190); the loop runs equilibration_loops + N_loop + 1
! collision cycles of N_MD steps on the device (the fused program of the C ABI,
! rmpcd_dimer_*), and after each sampled cycle it records what the reference's thermo
! group holds: temperature, potential, kinetic and total energy, centre-of-mass velocity,
! plus the dimer's unfolded centre.  HDF5 is not available, so the record is a text
! table (`output`) with the same columns instead of an H5MD file.  The collision is the
! reference's default (no thermostat, hydrodynamics kept); a parameter file that asks
! for the thermostat stops with a message.
!
! Environment: RMPCD_FORM=fp32 runs the solvent in the FP32-mixed form (default FP64).
program single_dimer_pbc
  use iso_c_binding
  use rmpcd_capi
  use ParseText
  implicit none
  type(PTo) :: config
  type(rmpcd_dimer_config_t) :: cfg
  type(rmpcd_dimer_stats_t) :: st
  type(rmpcd_particles_t) :: solvent, colloids
  type(rmpcd_box_t) :: box
  type(rmpcd_cells_t) :: cells
  type(c_ptr) :: sim
  character(len=1024) :: input_file, output_file, arg, form
  integer(c_int64_t) :: seed
  integer :: N_loop, equilibration_loops, colloid_sampling, i, u, iostat, count
  logical :: do_hydro, do_thermostat, bulk_rmpcd
  double precision :: temperature, kin_e, v_com(3), com(3), e_pot
  double precision, allocatable, target :: sv(:,:), cp(:,:), cv(:,:)
  real(c_float), allocatable, target :: sv32(:,:)
  integer(c_int32_t), allocatable, target :: cim(:,:)

  if (command_argument_count() /= 3) then
     write(*,*) 'Usage:'
     write(*,*) '    single_dimer_pbc input output seed'
     write(*,*) '    where input is the filename of the parameters file, output is the name'
     write(*,*) '    of the observables table and seed is a signed 64-bit integer or auto'
     error stop
  end if
  call get_command_argument(1, input_file)
  call get_command_argument(2, output_file)
  call get_command_argument(3, arg)
  if (trim(arg) == 'auto') then
     call system_clock(count)
     seed = int(count, c_int64_t)*2654435761_c_int64_t
  else
     read(arg, *, iostat=iostat) seed
     if (iostat /= 0) then
        write(*,*) 'Error when reading the seed value (third command-line argument)'
        error stop
     end if
  end if
  call check(rmpcd_init(0))

  call PTparse(config, input_file, 11)
  cfg%T = PTread_d(config, 'T')
  cfg%L = PTread_ivec(config, 'L', 3)
  cfg%rho = PTread_i(config, 'rho')
  cfg%tau = PTread_d(config, 'tau')
  cfg%prob = PTread_d(config, 'probability')
  bulk_rmpcd = PTread_l(config, 'bulk_rmpcd')
  cfg%bulk_rate = PTread_d(config, 'bulk_rate')
  cfg%n_md = PTread_i(config, 'N_MD')
  colloid_sampling = PTread_i(config, 'colloid_sampling')
  if (modulo(cfg%n_md, colloid_sampling) /= 0) then
     error stop 'colloid_sampling must divide N_MD with no remainder'
  end if
  N_loop = PTread_i(config, 'N_loop')
  equilibration_loops = PTread_i(config, 'equilibration_loops')
  cfg%sigma_C = PTread_d(config, 'sigma_C')
  cfg%sigma_N = PTread_d(config, 'sigma_N')
  cfg%epsilon_C = PTread_dvec(config, 'epsilon_C', 2)
  cfg%epsilon_N = PTread_dvec(config, 'epsilon_N', 2)
  cfg%epsilon_CC = PTread_d(config, 'epsilon_C_C')
  cfg%epsilon_NC = PTread_d(config, 'epsilon_N_C')
  cfg%epsilon_NN = PTread_d(config, 'epsilon_N_N')
  cfg%d = PTread_d(config, 'd')
  do_hydro = PTread_l(config, 'do_hydro')
  do_thermostat = PTread_l(config, 'do_thermostat')
  call PTkill(config)
  if (do_thermostat .or. .not. do_hydro) then
     error stop 'single_dimer_pbc: the device program collides without thermostat (do_thermostat = F, do_hydro = T)'
  end if
  cfg%bulk_rmpcd = merge(1, 0, bulk_rmpcd)
  cfg%seed = seed
  cfg%mode = RMPCD_SIM_FUSED
  call get_environment_variable('RMPCD_FORM', form)
  cfg%form = merge(RMPCD_FP32, RMPCD_FP64, trim(form) == 'fp32')
  call check(rmpcd_dimer_create(cfg, sim))
  call check(rmpcd_dimer_state(sim, 0_c_int32_t, solvent, colloids, box, cells, st))

  open(newunit=u, file=trim(output_file), status='replace', action='write')
  write(u, '(a)') '# RMPCDMD::single_dimer_pbc (MI355X), parameters from '//trim(input_file)
  write(u, '(a,i0,a,i0,a,i0,a,i0)') '# L ', cfg%L(1), ' ', cfg%L(2), ' ', cfg%L(3), '  rho ', cfg%rho
  write(u, '(a,i0,a,es24.16)') '# N_MD ', cfg%n_md, '  dt ', st%dt
  write(u, '(a,2es24.16)') '# mass ', st%mass
  write(u, '(a,i0)') '# seed ', seed
  write(u, '(a)') '# loop temperature e_potential e_kinetic e_total v_com_x v_com_y v_com_z com_x com_y com_z'

  write(*,*) 'Running for', equilibration_loops, '+', N_loop, 'loops'
  write(*,*) 'mass', st%mass
  allocate(cp(2, 3), cv(2, 3), cim(2, 3))
  if (cfg%form == RMPCD_FP32) then
     allocate(sv32(st%n_solvent, 3))
  else
     allocate(sv(st%n_solvent, 3))
  end if
  do i = 0, N_loop + equilibration_loops
     if (modulo(i, 20) == 0) write(*, '(i05)', advance='no') i
     call check(rmpcd_dimer_run(sim, 1_c_int32_t))
     if (i < equilibration_loops) cycle
     call check(rmpcd_dimer_state(sim, 0_c_int32_t, solvent, colloids, box, cells, st))
     call check(rmpcd_compute_temperature(box, cells, solvent, temperature))
     if (cfg%form == RMPCD_FP32) then
        call check(rmpcd_memcpy(c_loc(sv32), solvent%vel, int(size(sv32), c_size_t)*4, D2H))
        kin_e = sum(real(sv32, c_double)**2)/2
        v_com = sum(real(sv32, c_double), dim=1)
     else
        call check(rmpcd_memcpy(c_loc(sv), solvent%vel, int(size(sv), c_size_t)*8, D2H))
        kin_e = sum(sv**2)/2
        v_com = sum(sv, dim=1)
     end if
     call check(rmpcd_memcpy(c_loc(cp), colloids%pos, int(size(cp), c_size_t)*8, D2H))
     call check(rmpcd_memcpy(c_loc(cv), colloids%vel, int(size(cv), c_size_t)*8, D2H))
     call check(rmpcd_memcpy(c_loc(cim), colloids%image, int(size(cim), c_size_t)*4, D2H))
     kin_e = kin_e + (st%mass(1)*sum(cv(1, :)**2) + st%mass(2)*sum(cv(2, :)**2))/2
     v_com = (v_com + st%mass(1)*cv(1, :) + st%mass(2)*cv(2, :))/(st%n_solvent + st%mass(1) + st%mass(2))
     com = (cp(1, :) + cim(1, :)*box%edges + cp(2, :) + cim(2, :)*box%edges)/2
     e_pot = st%e1 + st%e2
     write(u, '(i8,10es24.16)') i, temperature, e_pot, kin_e, e_pot + kin_e, v_com, com
  end do
  write(*, *)
  write(*, '(a,i0)') 'n extra sorting ', st%n_extra_sorting
  close(u)
  call check(rmpcd_dimer_destroy(sim))
end program single_dimer_pbc
