! particle_system_io -- the particle and thermodynamic output of the drivers
! (src/particle_system_io.f90: thermo_t, particle_system_io_info_t,
! particle_system_io_t), written through h5md_module (a plain-text H5MD stand-in).
! Same types, members and argument meaning as the reference; each stored field of a
! particle set becomes one element of its group, time-dependent (H5MD_TIME or
! H5MD_LINEAR, with optional step/time offsets) or fixed.
module particle_system_io
  use hdf5
  use h5md_module
  use particle_system
  implicit none
  private

  public :: particle_system_io_t, particle_system_io_info_t, thermo_t

  !> buffered observables, written every n_buffer appends (src/particle_system_io.f90:25-36)
  type thermo_t
     integer :: n_buffer
     integer :: idx
     double precision, allocatable, dimension(:) :: temperature
     double precision, allocatable, dimension(:) :: potential_energy
     double precision, allocatable, dimension(:) :: kinetic_energy
     double precision, allocatable, dimension(:) :: internal_energy
     double precision, allocatable, dimension(:,:) :: center_of_mass_velocity
   contains
     procedure :: init => thermo_init
     procedure :: append => thermo_append
  end type thermo_t

  !> how one field is stored (src/particle_system_io.f90:38-45)
  type particle_system_io_info_t
     integer :: mode
     integer :: step
     integer :: step_offset = -1
     double precision :: time
     double precision :: time_offset = -1
     logical :: store
  end type particle_system_io_info_t

  !> the elements of one particle group (src/particle_system_io.f90:47-66)
  type particle_system_io_t
     integer :: Nmax
     integer :: error
     integer(HID_T) :: group
     type(h5md_element_t) :: box
     type(h5md_element_t) :: position
     type(h5md_element_t) :: image
     type(h5md_element_t) :: velocity
     type(h5md_element_t) :: force
     type(h5md_element_t) :: id
     type(h5md_element_t) :: species
     type(particle_system_io_info_t) :: position_info
     type(particle_system_io_info_t) :: image_info
     type(particle_system_io_info_t) :: velocity_info
     type(particle_system_io_info_t) :: force_info
     type(particle_system_io_info_t) :: id_info
     type(particle_system_io_info_t) :: species_info
   contains
     procedure :: init => ps_init
     procedure :: close => ps_close
  end type particle_system_io_t

contains

  ! the storage of one field: time-dependent (with the increments, and both offsets or
  ! neither, of a linear mode) or fixed; d2 for (3, N) reals, i2 for (3, N) integers,
  ! i1 for (N) integers
  subroutine check_offsets(info, what)
    type(particle_system_io_info_t), intent(in) :: info
    character(len=*), intent(in) :: what
    if ( (info%step_offset < 0) .and. (info%time_offset > 0) .or. &
         (info%step_offset > 0) .and. (info%time_offset < 0) ) then
       write(*, '(a)') 'use both or neither of step_offset and time_offset in ps_init '//what
       stop
    end if
  end subroutine check_offsets

  subroutine make_d2(e, g, name, x, info)
    type(h5md_element_t), intent(inout) :: e
    integer(HID_T), intent(in) :: g
    character(len=*), intent(in) :: name
    double precision, intent(in) :: x(:,:)
    type(particle_system_io_info_t), intent(in) :: info
    call check_offsets(info, name)
    if (iand(info%mode, H5MD_TIME) == H5MD_TIME) then
       call e%create_time(g, name, x, info%mode)
    else if (iand(info%mode, H5MD_LINEAR) == H5MD_LINEAR) then
       if (info%step_offset > 0) then
          call e%create_time(g, name, x, info%mode, info%step, info%time, step_offset=info%step_offset, &
               time_offset=info%time_offset)
       else
          call e%create_time(g, name, x, info%mode, info%step, info%time)
       end if
    else if (iand(info%mode, H5MD_FIXED) == H5MD_FIXED) then
       call e%create_fixed(g, name, x)
    else
       write(*, '(a)') 'unknown storage for '//name//' in particle_system_io init'
       stop
    end if
  end subroutine make_d2

  subroutine make_i2(e, g, name, x, info)
    type(h5md_element_t), intent(inout) :: e
    integer(HID_T), intent(in) :: g
    character(len=*), intent(in) :: name
    integer, intent(in) :: x(:,:)
    type(particle_system_io_info_t), intent(in) :: info
    call check_offsets(info, name)
    if (iand(info%mode, H5MD_TIME) == H5MD_TIME) then
       call e%create_time(g, name, x, info%mode)
    else if (iand(info%mode, H5MD_LINEAR) == H5MD_LINEAR) then
       if (info%step_offset > 0) then
          call e%create_time(g, name, x, info%mode, info%step, info%time, step_offset=info%step_offset, &
               time_offset=info%time_offset)
       else
          call e%create_time(g, name, x, info%mode, info%step, info%time)
       end if
    else if (iand(info%mode, H5MD_FIXED) == H5MD_FIXED) then
       call e%create_fixed(g, name, x)
    else
       write(*, '(a)') 'unknown storage for '//name//' in particle_system_io init'
       stop
    end if
  end subroutine make_i2

  subroutine make_i1(e, g, name, x, info)
    type(h5md_element_t), intent(inout) :: e
    integer(HID_T), intent(in) :: g
    character(len=*), intent(in) :: name
    integer, intent(in) :: x(:)
    type(particle_system_io_info_t), intent(in) :: info
    call check_offsets(info, name)
    if (iand(info%mode, H5MD_TIME) == H5MD_TIME) then
       call e%create_time(g, name, x, info%mode)
    else if (iand(info%mode, H5MD_LINEAR) == H5MD_LINEAR) then
       if (info%step_offset > 0) then
          call e%create_time(g, name, x, info%mode, info%step, info%time, step_offset=info%step_offset, &
               time_offset=info%time_offset)
       else
          call e%create_time(g, name, x, info%mode, info%step, info%time)
       end if
    else if (iand(info%mode, H5MD_FIXED) == H5MD_FIXED) then
       call e%create_fixed(g, name, x)
    else
       write(*, '(a)') 'unknown storage for '//name//' in particle_system_io init'
       stop
    end if
  end subroutine make_i1

  !> the group particles/<name> and an element for each stored field
  !! (src/particle_system_io.f90:70-222)
  subroutine ps_init(this, h5md_file, name, ps)
    class(particle_system_io_t), intent(inout) :: this
    type(h5md_file_t), intent(inout) :: h5md_file
    character(len=*), intent(in) :: name
    type(particle_system_t), intent(in) :: ps
    this% Nmax = ps% Nmax
    call h5gcreate_f(h5md_file% particles, name, this% group, this% error)
    if (this%position_info%store) call make_d2(this%position, this%group, 'position', ps%pos, this%position_info)
    if (this%velocity_info%store) call make_d2(this%velocity, this%group, 'velocity', ps%vel, this%velocity_info)
    if (this%force_info%store) call make_d2(this%force, this%group, 'force', ps%force, this%force_info)
    if (this%image_info%store) call make_i2(this%image, this%group, 'image', ps%image, this%image_info)
    if (this%id_info%store) call make_i1(this%id, this%group, 'id', ps%id, this%id_info)
    if (this%species_info%store) call make_i1(this%species, this%group, 'species', ps%species, this%species_info)
  end subroutine ps_init

  subroutine ps_close(this)
    class(particle_system_io_t), intent(inout) :: this
    call this% position% close()
    call h5gclose_f(this% group, this% error)
  end subroutine ps_close

  !> five linear time series under observables/ (src/particle_system_io.f90:230-293)
  subroutine thermo_init(this, datafile, n_buffer, step, time, step_offset, time_offset)
    class(thermo_t), intent(out) :: this
    type(h5md_file_t), intent(inout) :: datafile
    integer, intent(in) :: n_buffer
    integer, intent(in) :: step
    double precision, intent(in) :: time
    integer, intent(in), optional :: step_offset
    double precision, intent(in), optional :: time_offset
    type(h5md_element_t) :: e
    double precision :: dummy, dummy_vec(3)
    integer :: mode, k
    character(len=24), parameter :: names(4) = [character(len=24) :: 'temperature', 'potential_energy', &
         'kinetic_energy', 'internal_energy']
    if (n_buffer <= 0) error stop 'n_buffer non-positive in thermo_init'
    if (present(step_offset) .neqv. present(time_offset)) then
       write(*, '(a)') 'in thermo_init, use both or neither of step_offset and time_offset'
       stop
    end if
    this% n_buffer = n_buffer
    this% idx = 0
    mode = ior(H5MD_LINEAR, H5MD_STORE_TIME)
    allocate(this% temperature(n_buffer))
    allocate(this% potential_energy(n_buffer))
    allocate(this% kinetic_energy(n_buffer))
    allocate(this% internal_energy(n_buffer))
    allocate(this% center_of_mass_velocity(3, n_buffer))
    dummy = 0
    dummy_vec = 0
    do k = 1, 4
       if (present(step_offset)) then
          call e%create_time(datafile%observables, trim(names(k)), dummy, mode, step, time, &
               step_offset=step_offset, time_offset=time_offset)
       else
          call e%create_time(datafile%observables, trim(names(k)), dummy, mode, step, time)
       end if
       call e%close()
    end do
    if (present(step_offset)) then
       call e%create_time(datafile%observables, 'center_of_mass_velocity', dummy_vec, mode, step, time, &
            step_offset=step_offset, time_offset=time_offset)
    else
       call e%create_time(datafile%observables, 'center_of_mass_velocity', dummy_vec, mode, step, time)
    end if
    call e%close()
  end subroutine thermo_init

  !> buffer one set of observables (unless add is false); write the buffer when it is
  !! full or when force is true (src/particle_system_io.f90:295-353)
  subroutine thermo_append(this, datafile, temperature, potential_energy, kinetic_energy, internal_energy, &
       center_of_mass_velocity, add, force)
    class(thermo_t), intent(inout) :: this
    type(h5md_file_t), intent(inout) :: datafile
    double precision, intent(in) :: temperature, potential_energy, kinetic_energy, internal_energy
    double precision, intent(in) :: center_of_mass_velocity(3)
    logical, intent(in), optional :: add, force
    type(h5md_element_t) :: e
    logical :: do_add, do_append
    integer :: i
    do_add = .true.
    if (present(add)) do_add = add
    if (do_add) then
       i = this%idx + 1
       this%idx = i
       this%temperature(i) = temperature
       this%potential_energy(i) = potential_energy
       this%kinetic_energy(i) = kinetic_energy
       this%internal_energy(i) = internal_energy
       this%center_of_mass_velocity(:,i) = center_of_mass_velocity
    end if
    do_append = (this%idx == this%n_buffer)
    if (present(force)) then
       if (force .and. this%idx > 0) do_append = .true.
    end if
    if (.not. do_append) return
    call e%open_time(datafile%observables, 'temperature')
    call e%append_buffer(this%temperature, force_size=this%idx)
    call e%close()
    call e%open_time(datafile%observables, 'potential_energy')
    call e%append_buffer(this%potential_energy, force_size=this%idx)
    call e%close()
    call e%open_time(datafile%observables, 'kinetic_energy')
    call e%append_buffer(this%kinetic_energy, force_size=this%idx)
    call e%close()
    call e%open_time(datafile%observables, 'internal_energy')
    call e%append_buffer(this%internal_energy, force_size=this%idx)
    call e%close()
    call e%open_time(datafile%observables, 'center_of_mass_velocity')
    call e%append_buffer(this%center_of_mass_velocity, force_size=this%idx)
    call e%close()
    this%idx = 0
  end subroutine thermo_append

end module particle_system_io
