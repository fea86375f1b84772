! Reads a parameter file of the reference's drivers with ParseText and prints the values
! single_dimer_pbc asks for, one `name value...` per line, in full precision.  No GPU:
! it checks the parameter reader on its own.
!
!     fortran_params file.parameters
program fortran_params
  use ParseText
  implicit none
  type(PTo) :: config
  character(len=1024) :: input_file
  integer :: L(3)
  double precision :: e(2)
  call get_command_argument(1, input_file)
  call PTparse(config, input_file, 11)
  L = PTread_ivec(config, 'L', 3)
  write(*, '(a,3(1x,i0))') 'L', L
  write(*, '(a,1x,i0)') 'rho', PTread_i(config, 'rho')
  write(*, '(a,1x,es24.16)') 'T', PTread_d(config, 'T')
  write(*, '(a,1x,es24.16)') 'tau', PTread_d(config, 'tau')
  write(*, '(a,1x,es24.16)') 'probability', PTread_d(config, 'probability')
  write(*, '(a,1x,l1)') 'bulk_rmpcd', PTread_l(config, 'bulk_rmpcd')
  write(*, '(a,1x,es24.16)') 'bulk_rate', PTread_d(config, 'bulk_rate')
  write(*, '(a,1x,i0)') 'N_MD', PTread_i(config, 'N_MD')
  write(*, '(a,1x,i0)') 'N_loop', PTread_i(config, 'N_loop')
  write(*, '(a,1x,i0)') 'colloid_sampling', PTread_i(config, 'colloid_sampling')
  write(*, '(a,1x,i0)') 'equilibration_loops', PTread_i(config, 'equilibration_loops')
  write(*, '(a,1x,es24.16)') 'sigma_N', PTread_d(config, 'sigma_N')
  write(*, '(a,1x,es24.16)') 'sigma_C', PTread_d(config, 'sigma_C')
  write(*, '(a,1x,es24.16)') 'd', PTread_d(config, 'd')
  e = PTread_dvec(config, 'epsilon_N', 2)
  write(*, '(a,2(1x,es24.16))') 'epsilon_N', e
  e = PTread_dvec(config, 'epsilon_C', 2)
  write(*, '(a,2(1x,es24.16))') 'epsilon_C', e
  write(*, '(a,1x,es24.16)') 'epsilon_N_N', PTread_d(config, 'epsilon_N_N')
  write(*, '(a,1x,es24.16)') 'epsilon_N_C', PTread_d(config, 'epsilon_N_C')
  write(*, '(a,1x,es24.16)') 'epsilon_C_C', PTread_d(config, 'epsilon_C_C')
  write(*, '(a,1x,l1)') 'do_hydro', PTread_l(config, 'do_hydro')
  write(*, '(a,1x,l1)') 'do_thermostat', PTread_l(config, 'do_thermostat')
  call PTkill(config)
end program fortran_params
