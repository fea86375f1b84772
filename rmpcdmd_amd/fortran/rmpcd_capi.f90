! rmpcd_capi.f90 -- iso_c_binding view of include/rmpcd_capi.h (the C ABI of the HIP
! library) for the Fortran modules that keep the RMPCDMD module API.
module rmpcd_capi
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: RMPCD_OK = 0, RMPCD_ERR_OVERFLOW = -3
  integer(c_int32_t), parameter :: RMPCD_FP64 = 0, RMPCD_FP32 = 1
  integer(c_int32_t), parameter :: H2D = 0, D2H = 1, D2D = 2

  type, bind(c) :: rmpcd_box_t
     integer(c_int32_t) :: L(3), M(3), ncells, has_walls
     real(c_double) :: a, edges(3), origin(3)
     integer(c_int32_t) :: bc(3)
     real(c_double) :: max_v
  end type rmpcd_box_t

  type, bind(c) :: rmpcd_particles_t
     integer(c_int64_t) :: n
     integer(c_int32_t) :: form, pad_
     type(c_ptr) :: pos = c_null_ptr, vel = c_null_ptr, force = c_null_ptr, force_old = c_null_ptr
     type(c_ptr) :: image = c_null_ptr, id = c_null_ptr, species = c_null_ptr, flags = c_null_ptr
  end type rmpcd_particles_t

  type, bind(c) :: rmpcd_cells_t
     type(c_ptr) :: cell_count = c_null_ptr, cell_start = c_null_ptr
     type(c_ptr) :: is_md = c_null_ptr, is_reac = c_null_ptr
  end type rmpcd_cells_t

  integer(c_int32_t), parameter :: RMPCD_SIM_REFERENCE = 0, RMPCD_SIM_FUSED = 1

  !> rmpcd_dimer_config_t: the device-resident single-dimer program
  type, bind(c) :: rmpcd_dimer_config_t
     integer(c_int32_t) :: L(3) = 0, rho = 0, n_md = 0, form = 0, mode = 1, bulk_rmpcd = 0
     real(c_double) :: T = 0, tau = 0, sigma_C = 0, sigma_N = 0, d = 0
     real(c_double) :: epsilon_C(2) = 0, epsilon_N(2) = 0
     real(c_double) :: epsilon_CC = 0, epsilon_NC = 0, epsilon_NN = 0
     real(c_double) :: prob = 0, bulk_rate = 0, skin = 0
     integer(c_int32_t) :: nmax_list = 0, n_replicas = 0
     integer(c_int64_t) :: seed = 0
     type(c_ptr) :: replica_L = c_null_ptr
     integer(c_int32_t) :: stream_blocks = 0, sort_blocks = 0, lazy = 0, extra_sort_blocks = 0
     integer(c_int32_t) :: sort_order = 0, groups = 0, collide_cells = 0
  end type rmpcd_dimer_config_t

  type, bind(c) :: rmpcd_dimer_stats_t
     integer(c_int64_t) :: n_solvent, md_steps, cycles, n_extra_sorting, n_replicas, n_solvent_total
     real(c_double) :: e1, e2, dt, mass(2)
     integer(c_int32_t) :: nmax, error
     type(c_ptr) :: list, nlist, pos_old, colloid_pos_old
  end type rmpcd_dimer_stats_t

  !> rmpcd_rigid_body_t, in device memory
  type, bind(c) :: rmpcd_rigid_body_t
     integer(c_int32_t) :: i_start, i_stop
     real(c_double) :: mass
     real(c_double) :: pos(3), vel(3), force(3), q(4), L(3), torque(3), I_body(3), omega_body(3)
     type(c_ptr) :: pos_body
  end type rmpcd_rigid_body_t

  type, bind(c) :: rmpcd_lj_t
     integer(c_int32_t) :: n1, n2
     real(c_double) :: eps(16), sigma(16), cut(16), cut_sq(16), cut_energy(16)
  end type rmpcd_lj_t

  interface
     function rmpcd_last_error() bind(c) result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function rmpcd_last_error
     integer(c_int) function rmpcd_init(device) bind(c)
       import :: c_int
       integer(c_int), value :: device
     end function rmpcd_init
     integer(c_int) function rmpcd_malloc(ptr, bytes) bind(c)
       import :: c_int, c_ptr, c_size_t
       type(c_ptr) :: ptr
       integer(c_size_t), value :: bytes
     end function rmpcd_malloc
     integer(c_int) function rmpcd_free(ptr) bind(c)
       import :: c_int, c_ptr
       type(c_ptr), value :: ptr
     end function rmpcd_free
     integer(c_int) function rmpcd_memcpy(dst, src, bytes, kind) bind(c)
       import :: c_int, c_ptr, c_size_t, c_int32_t
       type(c_ptr), value :: dst, src
       integer(c_size_t), value :: bytes
       integer(c_int32_t), value :: kind
     end function rmpcd_memcpy
     integer(c_int32_t) function rmpcd_compact_p_to_h(p, M) bind(c)
       import :: c_int32_t
       integer(c_int32_t), intent(in) :: p(3), M(3)
     end function rmpcd_compact_p_to_h
     subroutine rmpcd_compact_h_to_p(h, M, p) bind(c)
       import :: c_int32_t
       integer(c_int32_t), value :: h
       integer(c_int32_t), intent(in) :: M(3)
       integer(c_int32_t), intent(out) :: p(3)
     end subroutine rmpcd_compact_h_to_p
     integer(c_int) function rmpcd_box_init(box, L, a, has_walls) bind(c)
       import :: c_int, c_int32_t, c_double, rmpcd_box_t
       type(rmpcd_box_t) :: box
       integer(c_int32_t), intent(in) :: L(3)
       real(c_double), value :: a
       integer(c_int32_t), value :: has_walls
     end function rmpcd_box_init
     integer(c_int) function rmpcd_count_particles(box, p, cells) bind(c)
       import :: c_int, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_particles_t), intent(in) :: p
       type(rmpcd_cells_t) :: cells
     end function rmpcd_count_particles
     integer(c_int) function rmpcd_sort(box, src, dst, id_to_idx, cells, apply_pbc) bind(c)
       import :: c_int, c_ptr, c_int32_t, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_particles_t), intent(in) :: src
       type(rmpcd_particles_t) :: dst
       type(c_ptr), value :: id_to_idx
       type(rmpcd_cells_t) :: cells
       integer(c_int32_t), value :: apply_pbc
     end function rmpcd_sort
     integer(c_int) function rmpcd_maximum_displacement(p, pos_old, r) bind(c)
       import :: c_int, c_ptr, c_double, rmpcd_particles_t
       type(rmpcd_particles_t), intent(in) :: p
       type(c_ptr), value :: pos_old
       real(c_double) :: r
     end function rmpcd_maximum_displacement
     integer(c_int) function rmpcd_cell_maximum_displacement(box, cells, p, pos_old, delta_t, r) bind(c)
       import :: c_int, c_ptr, c_double, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t), intent(in) :: cells
       type(rmpcd_particles_t), intent(in) :: p
       type(c_ptr), value :: pos_old
       real(c_double), value :: delta_t
       real(c_double) :: r
     end function rmpcd_cell_maximum_displacement
     integer(c_int) function rmpcd_simple_mpcd_step(box, cells, p, axes, seed, step, alpha_given, &
          alpha, thermostat, T, hydro) bind(c)
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t), intent(in) :: cells
       type(rmpcd_particles_t) :: p
       type(c_ptr), value :: axes
       integer(c_int64_t), value :: seed, step
       integer(c_int32_t), value :: alpha_given, thermostat, hydro
       real(c_double), value :: alpha, T
     end function rmpcd_simple_mpcd_step
     integer(c_int) function rmpcd_compute_temperature(box, cells, p, te) bind(c)
       import :: c_int, c_double, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t), intent(in) :: cells
       type(rmpcd_particles_t), intent(in) :: p
       real(c_double) :: te
     end function rmpcd_compute_temperature
     integer(c_int) function rmpcd_compute_temperature_profile(box, cells, p, te, xmin, dx, n_bins, tz_data, &
          tz_count) bind(c)
       import :: c_int, c_int32_t, c_double, c_ptr, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t), intent(in) :: cells
       type(rmpcd_particles_t), intent(in) :: p
       real(c_double) :: te
       real(c_double), value :: xmin, dx
       integer(c_int32_t), value :: n_bins
       type(c_ptr), value :: tz_data, tz_count
     end function rmpcd_compute_temperature_profile
     integer(c_int) function rmpcd_mpcd_stream_periodic(box, p, dt) bind(c)
       import :: c_int, c_double, rmpcd_box_t, rmpcd_particles_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_particles_t) :: p
       real(c_double), value :: dt
     end function rmpcd_mpcd_stream_periodic
     integer(c_int) function rmpcd_bulk_reaction(box, cells, p, from, to, rate, tau, draws, seed, step) bind(c)
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t), intent(in) :: cells
       type(rmpcd_particles_t) :: p
       integer(c_int32_t), value :: from, to
       real(c_double), value :: rate, tau
       type(c_ptr), value :: draws
       integer(c_int64_t), value :: seed, step
     end function rmpcd_bulk_reaction
     integer(c_int) function rmpcd_bulk_reaction_endothermic(box, cells, p, from, to, rate, tau, delta_u, draws, &
          seed, step) bind(c)
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t), intent(in) :: cells
       type(rmpcd_particles_t) :: p
       integer(c_int32_t), value :: from, to
       real(c_double), value :: rate, tau, delta_u
       type(c_ptr), value :: draws
       integer(c_int64_t), value :: seed, step
     end function rmpcd_bulk_reaction_endothermic
     integer(c_int) function rmpcd_rescale_cells(box, cells, p, T) bind(c)
       import :: c_int, c_double, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t), intent(in) :: cells
       type(rmpcd_particles_t) :: p
       real(c_double), value :: T
     end function rmpcd_rescale_cells
     integer(c_int) function rmpcd_wall_mpcd_step(box, cells, p, walls_given, wall_temperature, wall_n, &
          thermostat, bulk_temperature, keep_cell_v, alpha_given, alpha, axes, vnorm, pnorm, seed, step) bind(c)
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t), intent(in) :: cells
       type(rmpcd_particles_t) :: p
       integer(c_int32_t), value :: walls_given, thermostat, keep_cell_v, alpha_given
       type(c_ptr), value :: wall_temperature, wall_n, axes, vnorm, pnorm
       real(c_double), value :: bulk_temperature, alpha
       integer(c_int64_t), value :: seed, step
     end function rmpcd_wall_mpcd_step
     integer(c_int) function rmpcd_mpcd_stream_xforce_yzwall(box, p, dt, g) bind(c)
       import :: c_int, c_double, rmpcd_box_t, rmpcd_particles_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_particles_t) :: p
       real(c_double), value :: dt, g
     end function rmpcd_mpcd_stream_xforce_yzwall
     integer(c_int) function rmpcd_mpcd_stream_nogravity_zwall(box, p, dt) bind(c)
       import :: c_int, c_double, rmpcd_box_t, rmpcd_particles_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_particles_t) :: p
       real(c_double), value :: dt
     end function rmpcd_mpcd_stream_nogravity_zwall
     integer(c_int) function rmpcd_md_pos(p, dt) bind(c)
       import :: c_int, c_double, rmpcd_particles_t
       type(rmpcd_particles_t) :: p
       real(c_double), value :: dt
     end function rmpcd_md_pos
     integer(c_int) function rmpcd_md_vel(p, dt) bind(c)
       import :: c_int, c_double, rmpcd_particles_t
       type(rmpcd_particles_t) :: p
       real(c_double), value :: dt
     end function rmpcd_md_vel
     integer(c_int) function rmpcd_apply_pbc(p, edges) bind(c)
       import :: c_int, c_double, rmpcd_particles_t
       type(rmpcd_particles_t) :: p
       real(c_double), intent(in) :: edges(3)
     end function rmpcd_apply_pbc
     integer(c_int) function rmpcd_cell_md_pos(box, cells, p, dt, md_flag) bind(c)
       import :: c_int, c_double, c_int32_t, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t), intent(in) :: cells
       type(rmpcd_particles_t) :: p
       real(c_double), value :: dt
       integer(c_int32_t), value :: md_flag
     end function rmpcd_cell_md_pos
     integer(c_int) function rmpcd_cell_md_vel(box, cells, p, dt, md_flag) bind(c)
       import :: c_int, c_double, c_int32_t, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t), intent(in) :: cells
       type(rmpcd_particles_t) :: p
       real(c_double), value :: dt
       integer(c_int32_t), value :: md_flag
     end function rmpcd_cell_md_vel
     integer(c_int) function rmpcd_cell_md_pos_wall(box, cells, p, dt, md_flag, axis) bind(c)
       import :: c_int, c_double, c_int32_t, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t), intent(in) :: cells
       type(rmpcd_particles_t) :: p
       real(c_double), value :: dt
       integer(c_int32_t), value :: md_flag, axis
     end function rmpcd_cell_md_pos_wall
     integer(c_int) function rmpcd_rattle_dimer_pos(c, pos_rattle, mass, n_mass, d, dt, edges) bind(c)
       import :: c_int, c_ptr, c_double, c_int32_t, rmpcd_particles_t
       type(rmpcd_particles_t) :: c
       type(c_ptr), value :: pos_rattle
       real(c_double), intent(in) :: mass(*)
       integer(c_int32_t), value :: n_mass
       real(c_double), value :: d, dt
       real(c_double), intent(in) :: edges(3)
     end function rmpcd_rattle_dimer_pos
     integer(c_int) function rmpcd_rattle_dimer_vel(c, mass, n_mass, d, dt, edges) bind(c)
       import :: c_int, c_double, c_int32_t, rmpcd_particles_t
       type(rmpcd_particles_t) :: c
       real(c_double), intent(in) :: mass(*)
       integer(c_int32_t), value :: n_mass
       real(c_double), value :: d, dt
       real(c_double), intent(in) :: edges(3)
     end function rmpcd_rattle_dimer_vel
     integer(c_int) function rmpcd_rattle_body_pos(p, pos_rattle, links, distances, n_links, mass, n_mass, &
          dt, edges, precision, iterations) bind(c)
       import :: c_int, c_ptr, c_double, c_int32_t, rmpcd_particles_t
       type(rmpcd_particles_t) :: p
       type(c_ptr), value :: pos_rattle
       integer(c_int32_t), intent(in) :: links(*)
       real(c_double), intent(in) :: distances(*), mass(*)
       integer(c_int32_t), value :: n_links, n_mass
       real(c_double), value :: dt, precision
       real(c_double), intent(in) :: edges(3)
       integer(c_int32_t) :: iterations
     end function rmpcd_rattle_body_pos
     integer(c_int) function rmpcd_rattle_body_vel(p, links, distances, n_links, mass, n_mass, dt, edges, &
          precision, iterations) bind(c)
       import :: c_int, c_double, c_int32_t, rmpcd_particles_t
       type(rmpcd_particles_t) :: p
       integer(c_int32_t), intent(in) :: links(*)
       real(c_double), intent(in) :: distances(*), mass(*)
       integer(c_int32_t), value :: n_links, n_mass
       real(c_double), value :: dt, precision
       real(c_double), intent(in) :: edges(3)
       integer(c_int32_t) :: iterations
     end function rmpcd_rattle_body_vel
     integer(c_int) function rmpcd_elastic_network(p, links, distances, n_links, k, edges, energy) bind(c)
       import :: c_int, c_double, c_int32_t, rmpcd_particles_t
       type(rmpcd_particles_t) :: p
       integer(c_int32_t), intent(in) :: links(*)
       real(c_double), intent(in) :: distances(*)
       integer(c_int32_t), value :: n_links
       real(c_double), value :: k
       real(c_double), intent(in) :: edges(3)
       real(c_double) :: energy
     end function rmpcd_elastic_network
     integer(c_int) function rmpcd_rigid_body_init(body, p, mass, n_mass, edges, i_start, i_stop, pos_body) bind(c)
       import :: c_int, c_ptr, c_double, c_int32_t, rmpcd_particles_t
       type(c_ptr), value :: body, pos_body
       type(rmpcd_particles_t), intent(in) :: p
       real(c_double), intent(in) :: mass(*), edges(3)
       integer(c_int32_t), value :: n_mass, i_start, i_stop
     end function rmpcd_rigid_body_init
     integer(c_int) function rmpcd_rigid_body_compute_force_torque(body, p) bind(c)
       import :: c_int, c_ptr, rmpcd_particles_t
       type(c_ptr), value :: body
       type(rmpcd_particles_t), intent(in) :: p
     end function rmpcd_rigid_body_compute_force_torque
     integer(c_int) function rmpcd_rigid_body_vv1(body, p, dt, threshold, edges, iterations) bind(c)
       import :: c_int, c_ptr, c_double, c_int32_t, rmpcd_particles_t
       type(c_ptr), value :: body
       type(rmpcd_particles_t) :: p
       real(c_double), value :: dt, threshold
       real(c_double), intent(in) :: edges(3)
       integer(c_int32_t) :: iterations
     end function rmpcd_rigid_body_vv1
     integer(c_int) function rmpcd_rigid_body_vv2(body, p, dt) bind(c)
       import :: c_int, c_ptr, c_double, rmpcd_particles_t
       type(c_ptr), value :: body
       type(rmpcd_particles_t) :: p
       real(c_double), value :: dt
     end function rmpcd_rigid_body_vv2
     integer(c_int) function rmpcd_dimer_create(cfg, sim) bind(c)
       import :: c_int, c_ptr, rmpcd_dimer_config_t
       type(rmpcd_dimer_config_t), intent(in) :: cfg
       type(c_ptr) :: sim
     end function rmpcd_dimer_create
     integer(c_int) function rmpcd_dimer_run(sim, n_cycles) bind(c)
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: sim
       integer(c_int32_t), value :: n_cycles
     end function rmpcd_dimer_run
     integer(c_int) function rmpcd_dimer_state(sim, replica, solvent, colloids, box, cells, stats) bind(c)
       import :: c_int, c_ptr, c_int32_t, rmpcd_particles_t, rmpcd_box_t, rmpcd_cells_t, rmpcd_dimer_stats_t
       type(c_ptr), value :: sim
       integer(c_int32_t), value :: replica
       type(rmpcd_particles_t) :: solvent, colloids
       type(rmpcd_box_t) :: box
       type(rmpcd_cells_t) :: cells
       type(rmpcd_dimer_stats_t) :: stats
     end function rmpcd_dimer_state
     integer(c_int) function rmpcd_dimer_destroy(sim) bind(c)
       import :: c_int, c_ptr
       type(c_ptr), value :: sim
     end function rmpcd_dimer_destroy
     integer(c_int) function rmpcd_lj_params_init(lj, n1, n2, eps, sigma, cut) bind(c)
       import :: c_int, c_double, c_int32_t, rmpcd_lj_t
       type(rmpcd_lj_t) :: lj
       integer(c_int32_t), value :: n1, n2
       real(c_double), intent(in) :: eps(*), sigma(*), cut(*)
     end function rmpcd_lj_params_init
     integer(c_int32_t) function rmpcd_make_stencil(a, cut, stencil, cap) bind(c)
       import :: c_int32_t, c_double
       real(c_double), value :: a, cut
       integer(c_int32_t) :: stencil(*)
       integer(c_int32_t), value :: cap
     end function rmpcd_make_stencil
     integer(c_int) function rmpcd_update_list(box, cells, points, solvent, radius, stencil, n_stencil, &
          lj, nmax, list, n) bind(c)
       import :: c_int, c_ptr, c_double, c_int32_t, rmpcd_box_t, rmpcd_particles_t, rmpcd_cells_t
       type(rmpcd_box_t), intent(in) :: box
       type(rmpcd_cells_t) :: cells
       type(rmpcd_particles_t), intent(in) :: points, solvent
       real(c_double), value :: radius
       type(c_ptr), value :: stencil, lj, list, n
       integer(c_int32_t), value :: n_stencil, nmax
     end function rmpcd_update_list
     integer(c_int) function rmpcd_compute_force(ps1, ps2, list, n, nmax, L, lj, e) bind(c)
       import :: c_int, c_ptr, c_double, c_int32_t, rmpcd_particles_t, rmpcd_lj_t
       type(rmpcd_particles_t) :: ps1, ps2
       type(c_ptr), value :: list, n
       integer(c_int32_t), value :: nmax
       real(c_double), intent(in) :: L(3)
       type(rmpcd_lj_t), intent(in) :: lj
       real(c_double) :: e
     end function rmpcd_compute_force
     integer(c_int) function rmpcd_compute_force_n2(ps, L, lj, e) bind(c)
       import :: c_int, c_double, rmpcd_particles_t, rmpcd_lj_t
       type(rmpcd_particles_t) :: ps
       real(c_double), intent(in) :: L(3)
       type(rmpcd_lj_t), intent(in) :: lj
       real(c_double) :: e
     end function rmpcd_compute_force_n2
     subroutine rmpcd_threefry2x64_20(ctr, key, out) bind(c)
       import :: c_int64_t
       integer(c_int64_t), intent(in) :: ctr(2), key(2)
       integer(c_int64_t), intent(out) :: out(2)
     end subroutine rmpcd_threefry2x64_20
  end interface

contains

  !> stop with the library's message when a call fails (the reference's error stop)
  subroutine check(rc)
    integer(c_int), intent(in) :: rc
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == RMPCD_OK) return
    call c_f_pointer(rmpcd_last_error(), msg, [512])
    text = ''
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    write(*, '(a)') trim(text)
    error stop 'rmpcdmd_amd: library call failed'
  end subroutine check

end module rmpcd_capi
