! rmpcdmd.f90 -- the RMPCDMD module API (common, hilbert, threefry_module, interaction,
! cell_system, particle_system, neighbor_list, mpcd, md; rmpcdmd_module, which also
! brings particle_system_io, is in rmpcdmd_module.f90) over the HIP library.  Types keep the reference's public members as host arrays, so drivers read and
! write them as before; each library routine moves the arrays it needs to the device,
! runs the kernels through the C ABI and brings back what changed.  Indices stay 1-based
! on the host (ids, cell_start, list entries) and 0-based on the device.
!
! Random numbers: threefry_module gives the drivers a Threefry-2x64-20 stream per state
! element (key = (seed, element), counter incremented per draw); library routines that
! draw on the device (simple_mpcd_step, bulk_reaction) key the device generator with the
! seed of state(1) and a call counter.

module common
  use iso_c_binding
  implicit none
  private
  public :: rel_pos, switch, timer_t, pi, cross, profile_t, histogram_t
  public :: timer_list_t, args_t, get_input_filename, get_input_args, numbered_string, max_path_length
  public :: REAC_BIT, MD_BIT, WALL_BIT, PAST_MD_BIT, OUTBOUND_BIT, CATALYZED_BIT, ENZYME_REGION_BIT
  public :: REAC_MASK, MD_MASK, WALL_MASK, CATALYZED_MASK, ENZYME_REGION_MASK

  double precision, parameter :: pi = 4*atan(1.d0)
  ! flag bits (src/common.f90:49-61)
  integer, parameter :: REAC_BIT = 1, MD_BIT = 2, WALL_BIT = 3, PAST_MD_BIT = 4
  integer, parameter :: OUTBOUND_BIT = 5, CATALYZED_BIT = 6, ENZYME_REGION_BIT = 7
  integer, parameter :: REAC_MASK = 2, MD_MASK = 4, WALL_MASK = 8, CATALYZED_MASK = 64
  integer, parameter :: ENZYME_REGION_MASK = 128

  integer, parameter :: max_path_length = 255

  type timer_t
     double precision :: tic_time = 0
     double precision :: total = 0
     character(len=32) :: name = ''
   contains
     procedure :: init => timer_init
     procedure :: tic
     procedure :: tac
  end type timer_t

  !> timers collected for output (src/common.f90:112-126, 365-402)
  type timer_pointer_t
     type(timer_t), pointer :: p => null()
  end type timer_pointer_t
  type timer_list_t
     type(timer_pointer_t), allocatable :: timers(:)
     integer :: current_idx
   contains
     generic, public :: init => timer_list_init
     procedure, private :: timer_list_init
     generic, public :: append => timer_list_append
     procedure, private :: timer_list_append
     generic, public :: write => timer_list_write
     procedure, private :: timer_list_write
  end type timer_list_t

  !> the drivers' command line: input output seed (src/common.f90:133-137, 308-333)
  type args_t
     character(len=max_path_length) :: input_file
     character(len=max_path_length) :: output_file
     integer(c_int64_t) :: seed
  end type args_t

  !> binned averages along one coordinate (src/common.f90:66-80, 180-225)
  type profile_t
     double precision, allocatable :: data(:)
     integer, allocatable :: count(:)
     double precision :: xmin
     double precision :: dx
     integer :: n
   contains
     generic, public :: init => profile_init
     procedure, private :: profile_init
     generic, public :: bin => profile_bin
     procedure, private :: profile_bin
     generic, public :: reset => profile_reset
     procedure, private :: profile_reset
     generic, public :: norm => profile_norm
     procedure, private :: profile_norm
  end type profile_t

  !> histogram per species along one coordinate (src/common.f90:82-95, 227-262)
  type histogram_t
     double precision, allocatable :: data(:,:)
     double precision :: xmin
     double precision :: dx
     integer :: n, n_species
   contains
     generic, public :: init => histogram_init
     procedure, private :: histogram_init
     generic, public :: bin => histogram_bin
     procedure, private :: histogram_bin
  end type histogram_t

  interface switch
     module procedure :: switch_d2
     module procedure :: switch_i1
     module procedure :: switch_i2
  end interface switch

contains

  !> minimum image with a single shift (src/common.f90:162-178)
  pure function rel_pos(x, y, L) result(r)
    double precision, intent(in) :: x(3), y(3), L(3)
    double precision :: r(3)
    integer :: k
    r = x - y
    do k = 1, 3
       if (r(k) < -0.5d0*L(k)) then
          r(k) = r(k) + L(k)
       else if (r(k) > 0.5d0*L(k)) then
          r(k) = r(k) - L(k)
       end if
    end do
  end function rel_pos

  subroutine switch_d2(p1, p2)
    double precision, pointer, dimension(:,:), intent(inout) :: p1, p2
    double precision, pointer, dimension(:,:) :: p
    p => p1
    p1 => p2
    p2 => p
  end subroutine switch_d2

  subroutine switch_i1(p1, p2)
    integer, pointer, dimension(:), intent(inout) :: p1, p2
    integer, pointer, dimension(:) :: p
    p => p1
    p1 => p2
    p2 => p
  end subroutine switch_i1

  subroutine switch_i2(p1, p2)
    integer, pointer, dimension(:,:), intent(inout) :: p1, p2
    integer, pointer, dimension(:,:) :: p
    p => p1
    p1 => p2
    p2 => p
  end subroutine switch_i2

  !> the cross product x1 x x2 (src/common.f90:404-412)
  pure function cross(x1, x2) result(r)
    double precision, intent(in) :: x1(3), x2(3)
    double precision :: r(3)
    r(1) = x1(2)*x2(3) - x1(3)*x2(2)
    r(2) = x1(3)*x2(1) - x1(1)*x2(3)
    r(3) = x1(1)*x2(2) - x1(2)*x2(1)
  end function cross

  subroutine profile_init(this, xmin, xmax, n)
    class(profile_t), intent(out) :: this
    double precision, intent(in) :: xmin, xmax
    integer, intent(in) :: n
    this% n = n
    this% xmin = xmin
    this% dx = (xmax - xmin) / n
    if (this% dx <= 0) error stop 'negative step in profile_init'
    allocate(this% data(n))
    this% data = 0
    allocate(this% count(n))
    this% count = 0
  end subroutine profile_init

  !> bin floor((x - xmin)/dx) + 1 gets value and one count when it is in range
  subroutine profile_bin(this, x, value)
    class(profile_t), intent(inout) :: this
    double precision, intent(in) :: x, value
    integer :: idx
    idx = floor( (x - this% xmin) / this% dx ) + 1
    if ( ( idx > 0 ) .and. ( idx <= this% n ) ) then
       this% data(idx) = this% data(idx) + value
       this% count(idx) = this% count(idx) + 1
    end if
  end subroutine profile_bin

  subroutine profile_reset(this)
    class(profile_t), intent(inout) :: this
    this% data = 0
    this% count = 0
  end subroutine profile_reset

  subroutine profile_norm(this)
    class(profile_t), intent(inout) :: this
    where (this% count > 0)
       this% data = this% data / this% count
    end where
  end subroutine profile_norm

  subroutine histogram_init(this, xmin, xmax, n, n_species)
    class(histogram_t), intent(out) :: this
    double precision, intent(in) :: xmin, xmax
    integer, intent(in) :: n
    integer, optional, intent(in) :: n_species
    this% n_species = 1
    if (present(n_species)) this% n_species = n_species
    this% n = n
    this% xmin = xmin
    this% dx = (xmax - xmin) / n
    if (this% dx <= 0) error stop 'negative step in histogram_init'
    allocate(this% data(this% n_species, n))
    this% data = 0
  end subroutine histogram_init

  subroutine histogram_bin(this, x, s)
    class(histogram_t), intent(inout) :: this
    double precision, intent(in) :: x
    integer, intent(in), optional :: s
    integer :: idx, s_var
    s_var = 1
    if (present(s)) s_var = s
    idx = floor( (x - this% xmin) / this% dx ) + 1
    if ( ( idx > 0 ) .and. ( idx <= this% n ) ) then
       this% data(s_var, idx) = this% data(s_var, idx) + 1
    end if
  end subroutine histogram_bin

  character(len=max_path_length) function get_input_filename() result(r)
    if (command_argument_count() < 1) error stop 'missing argument for parameter file'
    call get_command_argument(1, r)
  end function get_input_filename

  type(args_t) function get_input_args() result(args)
    integer :: iostat
    character(max_path_length) :: r
    if (command_argument_count() /= 3) then
       write(*,*) 'Usage: program input output seed'
       write(*,*) '    input: the parameter file; output: the name of the H5MD output file;'
       write(*,*) '    seed: a signed 64-bit integer'
       error stop
    end if
    call get_command_argument(1, args%input_file)
    call get_command_argument(2, args%output_file)
    call get_command_argument(3, r)
    read(r, *, iostat=iostat) args%seed
    if (iostat /= 0) then
       write(*,*) 'Error when reading the seed value (third command-line argument)'
       error stop
    end if
  end function get_input_args

  subroutine timer_list_init(this, n)
    class(timer_list_t), intent(out) :: this
    integer, intent(in) :: n
    allocate(this%timers(n))
    this%current_idx = 0
  end subroutine timer_list_init

  subroutine timer_list_append(this, timer_target)
    class(timer_list_t), intent(inout) :: this
    type(timer_t), target, intent(in) :: timer_target
    this%current_idx = this%current_idx + 1
    if (this%current_idx > size(this%timers)) error stop 'exceeded timer_list_t size'
    this%timers(this%current_idx)%p => timer_target
  end subroutine timer_list_append

  !> one dataset per timer (its total) in group, and their sum in total_out
  subroutine timer_list_write(this, group, total_out)
    use hdf5, only: HID_T
    use h5md_module, only: h5md_write_dataset
    class(timer_list_t), intent(inout) :: this
    integer(HID_T), intent(inout) :: group
    double precision, optional, intent(out) :: total_out
    double precision :: total
    integer :: i
    total = 0
    do i = 1, size(this%timers)
       if (associated(this%timers(i)%p)) then
          call h5md_write_dataset(group, trim(this%timers(i)%p%name), this%timers(i)%p%total)
          total = total + this%timers(i)%p%total
       end if
    end do
    if (present(total_out)) total_out = total
  end subroutine timer_list_write

  !> base_string followed by index on length digits with leading zeros (src/common.f90:447-462)
  function numbered_string(base_string, index, length) result(s)
    character(len=*), intent(in) :: base_string
    integer, intent(in) :: index
    integer, intent(in) :: length
    character(len=:), allocatable :: s
    character(len=16) :: fmt
    allocate(character(len=len_trim(base_string) + length) :: s)
    write(fmt, '(a,i0,a,i0,a)') '(a,i', length, '.', length, ')'
    write(s, fmt) trim(base_string), index
  end function numbered_string

  subroutine timer_init(this, name, system_name)
    class(timer_t), intent(out) :: this
    character(len=*), intent(in) :: name
    character(len=*), optional, intent(in) :: system_name
    this%name = name
    if (present(system_name)) then
       if (len(system_name) > 0) this%name = system_name//' '//name
    end if
    this%total = 0
  end subroutine timer_init

  subroutine tic(this)
    class(timer_t), intent(inout) :: this
    integer(8) :: c, r
    call system_clock(c, r)
    this%tic_time = dble(c)/dble(r)
  end subroutine tic

  subroutine tac(this)
    class(timer_t), intent(inout) :: this
    integer(8) :: c, r
    call system_clock(c, r)
    this%total = this%total + dble(c)/dble(r) - this%tic_time
  end subroutine tac

end module common

module hilbert
  use iso_c_binding
  use rmpcd_capi, only: rmpcd_compact_p_to_h, rmpcd_compact_h_to_p
  implicit none
  private
  public :: p_to_h, h_to_p, compact_p_to_h, compact_h_to_p
contains
  ! the compact Hilbert key (src/hilbert.f90:241-307); with equal bit counts on the three
  ! axes the compact key is the plain one (src/hilbert.f90:139-186)
  function compact_p_to_h(p, m) result(h)
    integer, intent(in) :: p(3), m(3)
    integer :: h
    h = rmpcd_compact_p_to_h(int(p, c_int32_t), int(m, c_int32_t))
  end function compact_p_to_h

  function compact_h_to_p(h, m) result(p)
    integer, intent(in) :: h, m(3)
    integer :: p(3)
    integer(c_int32_t) :: q(3)
    call rmpcd_compact_h_to_p(int(h, c_int32_t), int(m, c_int32_t), q)
    p = q
  end function compact_h_to_p

  function p_to_h(p, m) result(h)
    integer, intent(in) :: p(3), m
    integer :: h
    h = compact_p_to_h(p, [m, m, m])
  end function p_to_h

  function h_to_p(h, m) result(p)
    integer, intent(in) :: h, m
    integer :: p(3)
    p = compact_h_to_p(h, [m, m, m])
  end function h_to_p
end module hilbert

module threefry_module
  use iso_c_binding
  use rmpcd_capi, only: rmpcd_threefry2x64_20
  implicit none
  private
  public :: threefry_rng_t, threefry_rng_init, threefry_double, threefry_normal

  type, bind(c) :: threefry_rng_t
     integer(c_int64_t) :: key(2) = 0
     integer(c_int64_t) :: ctr(2) = 0
     integer(c_int64_t) :: has_spare = 0
     real(c_double) :: spare = 0
  end type threefry_rng_t

contains

  subroutine threefry_rng_init(state, seed)
    type(threefry_rng_t), intent(out) :: state(:)
    integer(c_int64_t), intent(in) :: seed
    integer :: i
    do i = 1, size(state)
       state(i)%key = [seed, int(i, c_int64_t)]
       state(i)%ctr = 0
       state(i)%has_spare = 0
    end do
  end subroutine threefry_rng_init

  !> uniform double in [0,1) from the top 53 bits of the next block
  function threefry_double(state) result(u)
    type(threefry_rng_t), intent(inout) :: state
    double precision :: u
    integer(c_int64_t) :: out(2)
    call rmpcd_threefry2x64_20(state%ctr, state%key, out)
    state%ctr(1) = state%ctr(1) + 1
    u = dble(shiftr(out(1), 11)) * 2.d0**(-53)
  end function threefry_double

  !> standard normal by Box-Muller, the second value kept for the next call
  function threefry_normal(state) result(x)
    type(threefry_rng_t), intent(inout) :: state
    double precision :: x, u1, u2, r
    double precision, parameter :: two_pi = 8*atan(1.d0)
    if (state%has_spare /= 0) then
       state%has_spare = 0
       x = state%spare
       return
    end if
    u1 = 1.d0 - threefry_double(state)
    u2 = threefry_double(state)
    r = sqrt(-2.d0*log(u1))
    x = r*cos(two_pi*u2)
    state%spare = r*sin(two_pi*u2)
    state%has_spare = 1
  end function threefry_normal
end module threefry_module

module interaction
  use iso_c_binding
  use rmpcd_capi
  implicit none
  private
  public :: lj_params_t, lj_force, lj_energy

  type lj_params_t
     integer :: n1 = 0, n2 = 0
     double precision, allocatable :: epsilon(:,:), sigma(:,:), sigma_sq(:,:)
     double precision, allocatable :: cut(:,:), cut_sq(:,:), cut_energy(:,:), shift(:)
     type(rmpcd_lj_t) :: c
   contains
     procedure :: init => lj_params_init
  end type lj_params_t

contains

  !> lj_params_init (src/interaction.f90:38-80); the C library fills the table it uses
  subroutine lj_params_init(this, epsilon, sigma, cut, shift)
    class(lj_params_t), intent(out) :: this
    double precision, intent(in) :: epsilon(:,:), sigma(:,:), cut(:,:)
    double precision, intent(in), optional :: shift(:)
    integer :: n1, n2, i, j
    if (size(epsilon, 1) /= size(sigma, 1) .or. size(epsilon, 1) /= size(cut, 1) .or. &
         size(epsilon, 2) /= size(sigma, 2) .or. size(epsilon, 2) /= size(cut, 2)) then
       error stop 'unequal sizes for arguments to lj_params_init'
    end if
    n1 = size(epsilon, 1)
    n2 = size(epsilon, 2)
    this%n1 = n1
    this%n2 = n2
    call check(rmpcd_lj_params_init(this%c, int(n1, c_int32_t), int(n2, c_int32_t), &
         reshape(epsilon, [n1*n2]), reshape(sigma, [n1*n2]), reshape(cut, [n1*n2])))
    allocate(this%epsilon(n1, n2), this%sigma(n1, n2), this%sigma_sq(n1, n2), this%cut(n1, n2))
    allocate(this%cut_sq(n1, n2), this%cut_energy(n1, n2), this%shift(n1))
    this%epsilon = epsilon
    this%sigma = sigma
    this%sigma_sq = sigma**2
    this%cut = cut
    do j = 1, n2
       do i = 1, n1
          this%cut_sq(i, j) = this%c%cut_sq(i + n1*(j-1))
          this%cut_energy(i, j) = this%c%cut_energy(i + n1*(j-1))
       end do
    end do
    if (present(shift)) then
       this%shift = shift
    else
       this%shift = 0
    end if
  end subroutine lj_params_init

  !> LJ force (src/interaction.f90:83-96)
  pure function lj_force(d, r_sq, epsilon, sigma) result(f)
    double precision, intent(in) :: d(3), r_sq, epsilon, sigma
    double precision :: f(3), s6
    s6 = sigma**6/r_sq**3
    f = (24.d0*epsilon*s6/r_sq*(2.d0*s6 - 1.d0))*d
  end function lj_force

  !> LJ energy (src/interaction.f90:126-136)
  pure elemental function lj_energy(r_sq, epsilon, sigma, cut_energy) result(e)
    double precision, intent(in) :: r_sq, epsilon, sigma, cut_energy
    double precision :: e, s6
    s6 = sigma**6/r_sq**3
    e = 4.d0*epsilon*(s6**2 - s6) - cut_energy
  end function lj_energy
end module interaction

!> device buffers and host<->device moves shared by the modules below
module rmpcd_mirror
  use iso_c_binding
  use rmpcd_capi
  implicit none
  private
  public :: dbuf_t, ensure, put_vec, get_vec, put_ivec, get_ivec, put_i, get_i, put_l, get_l, put_d, &
       get_d, put_i_shift, get_i_shift

  !> a grow-only device buffer
  type dbuf_t
     type(c_ptr) :: p = c_null_ptr
     integer(c_size_t) :: bytes = 0
  end type dbuf_t

contains

  subroutine ensure(b, bytes)
    type(dbuf_t), intent(inout) :: b
    integer(c_size_t), intent(in) :: bytes
    if (b%bytes >= bytes .and. c_associated(b%p)) return
    if (c_associated(b%p)) call check(rmpcd_free(b%p))
    call check(rmpcd_malloc(b%p, max(bytes, 256_c_size_t)))
    b%bytes = max(bytes, 256_c_size_t)
  end subroutine ensure

  !> host (3,n) -> device planes x, y, z
  subroutine put_vec(b, a)
    type(dbuf_t), intent(inout) :: b
    double precision, intent(in) :: a(:,:)
    double precision, allocatable, target :: t(:,:)
    integer(c_size_t) :: nb
    nb = int(size(a), c_size_t)*8
    call ensure(b, nb)
    allocate(t(size(a, 2), 3))
    t = transpose(a)
    call check(rmpcd_memcpy(b%p, c_loc(t), nb, H2D))
  end subroutine put_vec

  subroutine get_vec(b, a)
    type(dbuf_t), intent(in) :: b
    double precision, intent(inout) :: a(:,:)
    double precision, allocatable, target :: t(:,:)
    allocate(t(size(a, 2), 3))
    call check(rmpcd_memcpy(c_loc(t), b%p, int(size(a), c_size_t)*8, D2H))
    a = transpose(t)
  end subroutine get_vec

  subroutine put_ivec(b, a)
    type(dbuf_t), intent(inout) :: b
    integer, intent(in) :: a(:,:)
    integer(c_int32_t), allocatable, target :: t(:,:)
    integer(c_size_t) :: nb
    nb = int(size(a), c_size_t)*4
    call ensure(b, nb)
    allocate(t(size(a, 2), 3))
    t = transpose(a)
    call check(rmpcd_memcpy(b%p, c_loc(t), nb, H2D))
  end subroutine put_ivec

  subroutine get_ivec(b, a)
    type(dbuf_t), intent(in) :: b
    integer, intent(inout) :: a(:,:)
    integer(c_int32_t), allocatable, target :: t(:,:)
    allocate(t(size(a, 2), 3))
    call check(rmpcd_memcpy(c_loc(t), b%p, int(size(a), c_size_t)*4, D2H))
    a = transpose(t)
  end subroutine get_ivec

  subroutine put_i(b, a)
    type(dbuf_t), intent(inout) :: b
    integer, intent(in) :: a(:)
    integer(c_int32_t), allocatable, target :: t(:)
    integer(c_size_t) :: nb
    nb = int(size(a), c_size_t)*4
    call ensure(b, nb)
    t = a
    call check(rmpcd_memcpy(b%p, c_loc(t), nb, H2D))
  end subroutine put_i

  subroutine get_i(b, a)
    type(dbuf_t), intent(in) :: b
    integer, intent(inout) :: a(:)
    integer(c_int32_t), allocatable, target :: t(:)
    allocate(t(size(a)))
    call check(rmpcd_memcpy(c_loc(t), b%p, int(size(a), c_size_t)*4, D2H))
    a = t
  end subroutine get_i

  !> 1-based host index -> 0-based device index, and back
  subroutine put_i_shift(b, a)
    type(dbuf_t), intent(inout) :: b
    integer, intent(in) :: a(:)
    call put_i(b, a - 1)
  end subroutine put_i_shift

  subroutine get_i_shift(b, a)
    type(dbuf_t), intent(in) :: b
    integer, intent(inout) :: a(:)
    call get_i(b, a)
    a = a + 1
  end subroutine get_i_shift

  subroutine put_l(b, a)
    type(dbuf_t), intent(inout) :: b
    logical, intent(in) :: a(:)
    integer(c_int8_t), allocatable, target :: t(:)
    integer(c_size_t) :: nb
    nb = int(size(a), c_size_t)
    call ensure(b, nb)
    allocate(t(size(a)))
    where (a)
       t = 1_c_int8_t
    elsewhere
       t = 0_c_int8_t
    end where
    call check(rmpcd_memcpy(b%p, c_loc(t), nb, H2D))
  end subroutine put_l

  subroutine get_l(b, a)
    type(dbuf_t), intent(in) :: b
    logical, intent(inout) :: a(:)
    integer(c_int8_t), allocatable, target :: t(:)
    allocate(t(size(a)))
    call check(rmpcd_memcpy(c_loc(t), b%p, int(size(a), c_size_t), D2H))
    a = t /= 0
  end subroutine get_l

  subroutine put_d(b, a)
    type(dbuf_t), intent(inout) :: b
    double precision, intent(in), target :: a(:)
    double precision, allocatable, target :: t(:)
    integer(c_size_t) :: nb
    nb = int(size(a), c_size_t)*8
    call ensure(b, nb)
    t = a
    call check(rmpcd_memcpy(b%p, c_loc(t), nb, H2D))
  end subroutine put_d

  subroutine get_d(b, a)
    type(dbuf_t), intent(in) :: b
    double precision, intent(inout) :: a(:)
    double precision, allocatable, target :: t(:)
    allocate(t(size(a)))
    call check(rmpcd_memcpy(c_loc(t), b%p, int(size(a), c_size_t)*8, D2H))
    a = t
  end subroutine get_d
end module rmpcd_mirror

module cell_system
  use iso_c_binding
  use rmpcd_capi
  use rmpcd_mirror
  use hilbert
  implicit none
  private
  public :: cell_system_t, PERIODIC_BC, BOUNCE_BACK_BC, SPECULAR_BC

  integer, parameter :: PERIODIC_BC = 1, BOUNCE_BACK_BC = 2, SPECULAR_BC = 3

  type cell_system_t
     integer :: L(3)
     double precision :: edges(3)
     integer :: N
     double precision :: a
     double precision :: origin(3)
     integer, allocatable :: cell_count(:), cell_start(:)
     logical, allocatable :: is_md(:), is_reac(:)
     double precision :: max_v = 0
     integer :: M(3)
     integer :: bc(3) = PERIODIC_BC
     logical :: has_walls
     type(dbuf_t) :: d_count, d_start, d_is_md, d_is_reac, d_pos
   contains
     procedure :: init
     procedure :: count_particles
     procedure :: sort_particles
     procedure :: random_shift
     procedure :: cartesian_indices
     procedure :: box => cells_box
     procedure :: to_device => cells_to_device
     procedure :: from_device => cells_from_device
     procedure :: view => cells_view
  end type cell_system_t

contains

  !> cell_system%init (src/cell_system.f90:49-84)
  subroutine init(this, L, a, has_walls)
    class(cell_system_t), intent(out) :: this
    integer, intent(in) :: L(3)
    double precision, intent(in) :: a
    logical, intent(in), optional :: has_walls
    type(rmpcd_box_t) :: b
    integer :: w
    w = 0
    this%has_walls = .false.
    if (present(has_walls)) this%has_walls = has_walls
    if (this%has_walls) w = 1
    call check(rmpcd_box_init(b, int(L, c_int32_t), a, int(w, c_int32_t)))
    this%L = b%L
    this%M = b%M
    this%N = b%ncells
    this%edges = b%edges
    this%a = a
    this%origin = 0
    this%bc = PERIODIC_BC
    allocate(this%cell_count(this%N), this%cell_start(this%N), this%is_md(this%N), this%is_reac(this%N))
    this%cell_count = 0
    this%cell_start = 1
    this%is_md = .false.
    this%is_reac = .true.
  end subroutine init

  function cells_box(this) result(b)
    class(cell_system_t), intent(in) :: this
    type(rmpcd_box_t) :: b
    b%L = this%L
    b%M = this%M
    b%ncells = this%N
    b%has_walls = merge(1, 0, this%has_walls)
    b%a = this%a
    b%edges = this%edges
    b%origin = this%origin
    b%bc = this%bc
    b%max_v = this%max_v
  end function cells_box

  subroutine cells_to_device(this)
    class(cell_system_t), intent(inout) :: this
    call put_i(this%d_count, this%cell_count)
    call put_i_shift(this%d_start, this%cell_start)
    call put_l(this%d_is_md, this%is_md)
    call put_l(this%d_is_reac, this%is_reac)
  end subroutine cells_to_device

  subroutine cells_from_device(this)
    class(cell_system_t), intent(inout) :: this
    call get_i(this%d_count, this%cell_count)
    call get_i_shift(this%d_start, this%cell_start)
    call get_l(this%d_is_md, this%is_md)
    call get_l(this%d_is_reac, this%is_reac)
  end subroutine cells_from_device

  function cells_view(this) result(c)
    class(cell_system_t), intent(in) :: this
    type(rmpcd_cells_t) :: c
    c%cell_count = this%d_count%p
    c%cell_start = this%d_start%p
    c%is_md = this%d_is_md%p
    c%is_reac = this%d_is_reac%p
  end function cells_view

  !> cell_system%count_particles (src/cell_system.f90:86-129)
  subroutine count_particles(this, position)
    class(cell_system_t), intent(inout) :: this
    double precision, intent(in) :: position(:, :)
    type(rmpcd_particles_t) :: p
    type(rmpcd_cells_t) :: c
    call ensure(this%d_count, int(this%N, c_size_t)*4)
    call ensure(this%d_start, int(this%N, c_size_t)*4)
    call ensure(this%d_is_md, int(this%N, c_size_t))
    call ensure(this%d_is_reac, int(this%N, c_size_t))
    call put_vec(this%d_pos, position)
    p%n = size(position, 2)
    p%form = RMPCD_FP64
    p%pos = this%d_pos%p
    p%vel = this%d_pos%p
    p%species = this%d_pos%p
    p%flags = this%d_pos%p
    c = this%view()
    call check(rmpcd_count_particles(this%box(), p, c))
    call get_i(this%d_count, this%cell_count)
    call get_i_shift(this%d_start, this%cell_start)
  end subroutine count_particles

  !> cell_system%sort_particles (src/cell_system.f90:131-151): the order within a cell is
  !! the stable (single-thread) one
  subroutine sort_particles(this, position_old, position_new)
    class(cell_system_t), intent(inout) :: this
    double precision, intent(in) :: position_old(:, :)
    double precision, intent(out) :: position_new(:, :)
    integer :: i, idx
    integer, allocatable :: start(:)
    start = this%cell_start
    do i = 1, size(position_old, 2)
       idx = compact_p_to_h(this%cartesian_indices(position_old(:, i)), this%M) + 1
       position_new(:, start(idx)) = position_old(:, i)
       start(idx) = start(idx) + 1
    end do
  end subroutine sort_particles

  !> cell_system%random_shift (src/cell_system.f90:153-163)
  subroutine random_shift(this, state)
    use threefry_module
    class(cell_system_t), intent(inout) :: this
    type(threefry_rng_t), intent(inout) :: state
    this%origin(1) = threefry_double(state) - 1
    this%origin(2) = threefry_double(state) - 1
    this%origin(3) = threefry_double(state) - 1
  end subroutine random_shift

  !> cell_system%cartesian_indices (src/cell_system.f90:165-175)
  function cartesian_indices(this, position) result(p)
    class(cell_system_t), intent(in) :: this
    double precision, intent(in) :: position(3)
    integer :: p(3)
    p = floor((position/this%a) - this%origin)
    if (p(1) == this%L(1)) p(1) = 0
    if (p(2) == this%L(2)) p(2) = 0
    if ((.not. this%has_walls) .and. (p(3) == this%L(3))) p(3) = 0
  end function cartesian_indices
end module cell_system

module particle_system
  use iso_c_binding
  use rmpcd_capi
  use rmpcd_mirror
  use common
  use interaction
  use cell_system
  implicit none
  private
  public :: particle_system_t, cell_maximum_displacement, compute_radial_histogram, correct_radial_histogram

  type particle_system_t
     integer :: Nmax
     integer :: n_species
     double precision, allocatable :: mass(:)
     double precision, contiguous, pointer :: pos1(:,:), pos2(:,:), pos(:,:), pos_old(:,:), pos_pointer(:,:)
     double precision, allocatable :: pos_rattle(:,:)
     double precision, contiguous, pointer :: vel1(:,:), vel2(:,:), vel(:,:), vel_old(:,:), vel_pointer(:,:)
     double precision, contiguous, pointer :: force1(:,:), force2(:,:), force3(:,:), force4(:,:)
     double precision, contiguous, pointer :: force(:,:), force_store(:,:), force_old(:,:)
     double precision, contiguous, pointer :: force_old_store(:,:), force_pointer(:,:)
     integer, contiguous, pointer :: id1(:), id2(:), id(:), id_old(:), id_pointer(:)
     integer, contiguous, pointer :: species1(:), species2(:), species(:), species_old(:), species_pointer(:)
     integer, contiguous, pointer :: image1(:,:), image2(:,:), image(:,:), image_old(:,:), image_pointer(:,:)
     integer, contiguous, pointer :: flags1(:), flags2(:), flags(:), flags_old(:)
     integer, allocatable :: id_to_idx(:)
     type(timer_t), pointer :: time_stream, time_step, time_sort, time_count, time_ct
     type(timer_t), pointer :: time_md_pos, time_md_vel, time_self_force, time_max_disp
     type(timer_t), pointer :: time_apply_pbc, time_rattle_pos, time_rattle_vel, time_elastic
     ! device mirror: set a is the working copy, set b the sort destination
     type(dbuf_t) :: a_pos, a_vel, a_force, a_force_old, a_image, a_id, a_species, a_flags
     type(dbuf_t) :: b_pos, b_vel, b_force, b_force_old, b_image, b_id, b_species, b_flags
     type(dbuf_t) :: d_aux, d_id_to_idx
   contains
     procedure :: init
     procedure :: init_from_file
     procedure :: random_placement
     procedure :: sort
     procedure :: maximum_displacement
     procedure :: upload
     procedure :: download
     procedure :: view
  end type particle_system_t

contains

  !> particle_system%init_from_file (src/particle_system.f90:186-248): positions (and
  !! species, when stored) of particles/<group_name> in a file written by h5md_module,
  !! either the fixed element (mode H5MD_FIXED) or frame idx (0-based) of the
  !! time-dependent one (H5MD_TIME or H5MD_LINEAR)
  subroutine init_from_file(this, filename, group_name, mode, idx)
    use h5md_module
    use hdf5
    class(particle_system_t), intent(out) :: this
    character(len=*), intent(in) :: filename, group_name
    integer, intent(in) :: mode
    integer, intent(in), optional :: idx
    type(h5md_file_t) :: f
    type(h5md_element_t) :: e
    integer(HID_T) :: g
    double precision, allocatable :: pos(:,:)
    integer, allocatable :: species(:)
    call f% open(filename, H5F_ACC_RDONLY_F)
    call h5md_check_valid(f% particles, 'particles group not found')
    call h5md_check_exists(f% particles, group_name, 'group '//group_name//' not found')
    call h5gopen_f(f% particles, group_name, g, f% error)
    if (mode == H5MD_FIXED) then
       call e% read_fixed(g, 'position', pos)
       call this% init(size(pos, 2), 2)
       this% pos = pos
       if (has_fixed(g, 'species')) then
          call e% read_fixed(g, 'species', species)
          this% species = species
       end if
    else if ( (mode == H5MD_TIME) .or. (mode == H5MD_LINEAR) ) then
       if (.not. present(idx) ) error stop 'missing idx in init_from_file'
       call e% open_time(g, 'position')
       call this% init(e% Nmax)
       call read_frame_d2(e, idx, this% pos)
       call e% close()
    end if
    call h5gclose_f(g, f% error)
    call f% close()
  contains
    logical function has_fixed(loc, name)
      integer(HID_T), intent(in) :: loc
      character(len=*), intent(in) :: name
      character(len=4096) :: line
      integer :: u, ios
      character(len=512) :: path
      has_fixed = .false.
      u = h5t_unit(h5t_file(loc))
      path = trim(h5t_path(loc))//'/'//trim(name)//' '
      rewind(u)
      do
         read(u, '(a)', iostat=ios) line
         if (ios /= 0) exit
         if (line(1:6) == 'fixed ' .and. index(line(7:), trim(path)//' ') == 1) then
            has_fixed = .true.
            exit
         end if
      end do
    end function has_fixed
  end subroutine init_from_file

  !> particle_system%init (src/particle_system.f90:81-184)
  subroutine init(this, Nmax, n_species, mass, system_name)
    class(particle_system_t), intent(out) :: this
    integer, intent(in) :: Nmax
    integer, intent(in), optional :: n_species
    double precision, intent(in), optional :: mass(:)
    character(len=*), intent(in), optional :: system_name
    character(len=24) :: nm
    integer :: i
    this%Nmax = Nmax
    this%n_species = 1
    if (present(n_species)) this%n_species = n_species
    if (present(mass)) this%mass = mass
    allocate(this%pos1(3, Nmax), this%pos2(3, Nmax), this%pos_rattle(3, Nmax))
    this%pos => this%pos1
    this%pos_old => this%pos2
    allocate(this%vel1(3, Nmax), this%vel2(3, Nmax))
    this%vel => this%vel1
    this%vel_old => this%vel2
    allocate(this%force1(3, Nmax), this%force2(3, Nmax), this%force3(3, Nmax), this%force4(3, Nmax))
    this%force => this%force1
    this%force_store => this%force2
    this%force_old => this%force3
    this%force_old_store => this%force4
    allocate(this%id1(Nmax), this%id2(Nmax))
    this%id => this%id1
    this%id_old => this%id2
    do i = 1, Nmax
       this%id(i) = i
    end do
    allocate(this%species1(Nmax), this%species2(Nmax))
    this%species => this%species1
    this%species_old => this%species2
    allocate(this%image1(3, Nmax), this%image2(3, Nmax))
    this%image => this%image1
    this%image_old => this%image2
    this%image = 0
    allocate(this%flags1(Nmax), this%flags2(Nmax))
    this%flags => this%flags1
    this%flags_old => this%flags2
    this%flags = 0
    allocate(this%id_to_idx(Nmax))
    nm = ''
    if (present(system_name)) nm = system_name
    allocate(this%time_stream, this%time_step, this%time_sort, this%time_count, this%time_ct)
    allocate(this%time_md_pos, this%time_md_vel, this%time_self_force, this%time_max_disp)
    allocate(this%time_apply_pbc, this%time_rattle_pos, this%time_rattle_vel, this%time_elastic)
    call this%time_stream%init('stream', trim(nm))
    call this%time_step%init('step', trim(nm))
    call this%time_sort%init('sort', trim(nm))
    call this%time_count%init('count', trim(nm))
    call this%time_ct%init('compute_temperature', trim(nm))
    call this%time_md_pos%init('md_pos', trim(nm))
    call this%time_md_vel%init('md_vel', trim(nm))
    call this%time_self_force%init('self force', trim(nm))
    call this%time_max_disp%init('maximum disp', trim(nm))
    call this%time_rattle_pos%init('rattle_pos', trim(nm))
    call this%time_rattle_vel%init('rattle_vel', trim(nm))
    call this%time_elastic%init('elastic_network', trim(nm))
    call this%time_apply_pbc%init('apply_pbc', trim(nm))
  end subroutine init

  !> host arrays -> device working set; 'what' lists the members: p v f o i d s g
  !! (pos, vel, force, force_old, image, id, species, flags)
  subroutine upload(this, what)
    class(particle_system_t), intent(inout) :: this
    character(len=*), intent(in) :: what
    if (index(what, 'p') > 0) call put_vec(this%a_pos, this%pos)
    if (index(what, 'v') > 0) call put_vec(this%a_vel, this%vel)
    if (index(what, 'f') > 0) call put_vec(this%a_force, this%force)
    if (index(what, 'o') > 0) call put_vec(this%a_force_old, this%force_old)
    if (index(what, 'i') > 0) call put_ivec(this%a_image, this%image)
    if (index(what, 'd') > 0) call put_i_shift(this%a_id, this%id)
    if (index(what, 's') > 0) call put_i(this%a_species, this%species)
    if (index(what, 'g') > 0) call put_i(this%a_flags, this%flags)
  end subroutine upload

  subroutine download(this, what)
    class(particle_system_t), intent(inout) :: this
    character(len=*), intent(in) :: what
    if (index(what, 'p') > 0) call get_vec(this%a_pos, this%pos)
    if (index(what, 'v') > 0) call get_vec(this%a_vel, this%vel)
    if (index(what, 'f') > 0) call get_vec(this%a_force, this%force)
    if (index(what, 'o') > 0) call get_vec(this%a_force_old, this%force_old)
    if (index(what, 'i') > 0) call get_ivec(this%a_image, this%image)
    if (index(what, 'd') > 0) call get_i_shift(this%a_id, this%id)
    if (index(what, 's') > 0) call get_i(this%a_species, this%species)
    if (index(what, 'g') > 0) call get_i(this%a_flags, this%flags)
  end subroutine download

  !> the C view of the working set (members not uploaded point to the position buffer)
  function view(this) result(p)
    class(particle_system_t), intent(inout) :: this
    type(rmpcd_particles_t) :: p
    integer(c_size_t) :: n3, n1
    n3 = int(3*this%Nmax, c_size_t)*8
    n1 = int(this%Nmax, c_size_t)*4
    call ensure(this%a_pos, n3)
    call ensure(this%a_vel, n3)
    call ensure(this%a_force, n3)
    call ensure(this%a_force_old, n3)
    call ensure(this%a_image, 3*n1)
    call ensure(this%a_id, n1)
    call ensure(this%a_species, n1)
    call ensure(this%a_flags, n1)
    p%n = this%Nmax
    p%form = RMPCD_FP64
    p%pos = this%a_pos%p
    p%vel = this%a_vel%p
    p%force = this%a_force%p
    p%force_old = this%a_force_old%p
    p%image = this%a_image%p
    p%id = this%a_id%p
    p%species = this%a_species%p
    p%flags = this%a_flags%p
  end function view

  !> random_placement (src/particle_system.f90:250-302)
  subroutine random_placement(this, L, other, lj_params, state)
    use threefry_module
    class(particle_system_t), intent(inout) :: this
    double precision, intent(in) :: L(3)
    type(particle_system_t), intent(inout), optional :: other
    type(lj_params_t), intent(in), optional :: lj_params
    type(threefry_rng_t), intent(inout) :: state
    integer :: i, j, n_obst
    double precision :: x(3), rsq
    logical :: tooclose
    if (present(other) .neqv. present(lj_params)) &
         error stop 'other and lj_params must be present/absent together'
    n_obst = 0
    if (present(other)) n_obst = other%Nmax
    do i = 1, this%Nmax
       tooclose = .true.
       do while (tooclose)
          x(1) = threefry_double(state)*L(1)
          x(2) = threefry_double(state)*L(2)
          x(3) = threefry_double(state)*L(3)
          tooclose = .false.
          do j = 1, n_obst
             rsq = sum(rel_pos(x, other%pos(:, j), L)**2)
             if (rsq < lj_params%cut_sq(this%species(i), other%species(j))) then
                tooclose = .true.
                exit
             end if
          end do
       end do
       this%pos(:, i) = x
    end do
  end subroutine random_placement

  !> particle_system%sort (src/particle_system.f90:304-355) on the device, then the
  !! reference's pointer switches
  subroutine sort(this, cells)
    class(particle_system_t), intent(inout) :: this
    type(cell_system_t), intent(inout) :: cells
    type(rmpcd_particles_t) :: src, dst
    type(rmpcd_cells_t) :: c
    integer(c_size_t) :: n3, n1
    call this%time_count%tic()
    call this%upload('pvfoidsg')
    src = this%view()
    n3 = int(3*this%Nmax, c_size_t)*8
    n1 = int(this%Nmax, c_size_t)*4
    call ensure(this%b_pos, n3)
    call ensure(this%b_vel, n3)
    call ensure(this%b_force, n3)
    call ensure(this%b_force_old, n3)
    call ensure(this%b_image, 3*n1)
    call ensure(this%b_id, n1)
    call ensure(this%b_species, n1)
    call ensure(this%b_flags, n1)
    call ensure(this%d_id_to_idx, n1)
    dst = src
    dst%pos = this%b_pos%p
    dst%vel = this%b_vel%p
    dst%force = this%b_force%p
    dst%force_old = this%b_force_old%p
    dst%image = this%b_image%p
    dst%id = this%b_id%p
    dst%species = this%b_species%p
    dst%flags = this%b_flags%p
    call ensure(cells%d_count, int(cells%N, c_size_t)*4)
    call ensure(cells%d_start, int(cells%N, c_size_t)*4)
    call ensure(cells%d_is_md, int(cells%N, c_size_t))
    call ensure(cells%d_is_reac, int(cells%N, c_size_t))
    c = cells%view()
    call this%time_count%tac()
    call this%time_sort%tic()
    call check(rmpcd_sort(cells%box(), src, dst, this%d_id_to_idx%p, c, 0))
    call get_vec(this%b_pos, this%pos_old)
    call get_vec(this%b_vel, this%vel_old)
    call get_vec(this%b_force, this%force_store)
    call get_vec(this%b_force_old, this%force_old_store)
    call get_ivec(this%b_image, this%image_old)
    call get_i_shift(this%b_id, this%id_old)
    call get_i(this%b_species, this%species_old)
    call get_i(this%b_flags, this%flags_old)
    call get_i_shift(this%d_id_to_idx, this%id_to_idx)
    call get_i(cells%d_count, cells%cell_count)
    call get_i_shift(cells%d_start, cells%cell_start)
    call this%time_sort%tac()
    call switch(this%pos, this%pos_old)
    call switch(this%image, this%image_old)
    call switch(this%vel, this%vel_old)
    call switch(this%force, this%force_store)
    call switch(this%force_old, this%force_old_store)
    call switch(this%id, this%id_old)
    call switch(this%species, this%species_old)
    call switch(this%flags, this%flags_old)
  end subroutine sort

  !> particle_system%maximum_displacement (src/particle_system.f90:357-378)
  function maximum_displacement(this) result(r)
    class(particle_system_t), intent(inout) :: this
    double precision :: r
    type(rmpcd_particles_t) :: p
    call this%time_max_disp%tic()
    call this%upload('ps')
    p = this%view()
    call put_vec(this%d_aux, this%pos_old)
    call check(rmpcd_maximum_displacement(p, this%d_aux%p, r))
    call this%time_max_disp%tac()
  end function maximum_displacement

  !> cell_maximum_displacement (src/particle_system.f90:380-410)
  function cell_maximum_displacement(cells, p, delta_t) result(r)
    class(cell_system_t), intent(inout) :: cells
    class(particle_system_t), intent(inout) :: p
    double precision, intent(in) :: delta_t
    double precision :: r
    type(rmpcd_particles_t) :: v
    call p%upload('ps')
    v = p%view()
    call put_vec(p%d_aux, p%pos_old)
    call cells%to_device()
    call check(rmpcd_cell_maximum_displacement(cells%box(), cells%view(), v, p%d_aux%p, delta_t, r))
  end function cell_maximum_displacement
  !> the radial distribution of solvent (species > 0) around x1 into hist, bins per
  !! species, from the cells of the box [-max_i, max_i]^3 around x1's cell
  !! (src/particle_system.f90:447-493); reads the host arrays of the last sort
  subroutine compute_radial_histogram(hist, x1, L, solvent, cells)
    use hilbert
    type(histogram_t), intent(inout) :: hist
    double precision, intent(in), dimension(3) :: x1, L
    type(particle_system_t), intent(in) :: solvent
    type(cell_system_t), intent(in) :: cells
    integer :: max_i, i, j, k, cell_idx, i_particle, s, start, count
    integer :: p(3)
    double precision :: delta_x(3), norm, r_max
    r_max = hist%xmin + hist%dx*hist%n
    p = cells%cartesian_indices(x1)
    max_i = floor(r_max/cells%a) + 1
    do i = -max_i, max_i
       do j = -max_i, max_i
          do k = -max_i, max_i
             cell_idx = compact_p_to_h(modulo(p + [i, j, k], cells%L), cells%M) + 1
             start = cells%cell_start(cell_idx)
             count = cells%cell_count(cell_idx)
             do i_particle = start, start + count - 1
                s = solvent%species(i_particle)
                if (s <= 0) cycle
                delta_x = rel_pos(solvent%pos(:,i_particle), x1, L)
                norm = norm2(delta_x)
                if (norm < r_max) call hist%bin(norm, s)
             end do
          end do
       end do
    end do
  end subroutine compute_radial_histogram

  !> divide bin i by the volume of its shell, 4 pi r^2 dx at its centre
  !! (src/particle_system.f90:495-506)
  subroutine correct_radial_histogram(hist)
    type(histogram_t), intent(inout) :: hist
    integer :: i
    double precision :: r
    do i = 1, hist%n
       r = hist%xmin + (i-0.5d0) * hist%dx
       hist%data(:,i) = hist%data(:,i) / (4*pi*r**2*hist%dx)
    end do
  end subroutine correct_radial_histogram

end module particle_system

module neighbor_list
  use iso_c_binding
  use rmpcd_capi
  use rmpcd_mirror
  use common
  use cell_system
  use particle_system
  use interaction
  implicit none
  private
  public :: neighbor_list_t, compute_force, compute_force_n2

  type neighbor_list_t
     integer :: Npoints, Nmax
     integer, allocatable :: list(:, :), n(:), stencil(:,:)
     type(timer_t) :: time_update, time_force
     type(dbuf_t) :: d_list, d_n, d_stencil
   contains
     procedure :: init
     procedure :: update_list
     procedure :: make_stencil
  end type neighbor_list_t

contains

  !> neighbor_list%init (src/neighbor_list.f90:38-53)
  subroutine init(this, Npoints, Nmax)
    class(neighbor_list_t), intent(out) :: this
    integer, intent(in) :: Npoints, Nmax
    this%Nmax = Nmax
    this%Npoints = Npoints
    allocate(this%list(Nmax, Npoints), this%n(Npoints))
    this%list = 0
    this%n = 0
    call this%time_update%init('neigh list update')
    call this%time_force%init('neigh list force')
  end subroutine init

  !> make_stencil (src/neighbor_list.f90:271-308)
  subroutine make_stencil(this, cells, cut)
    class(neighbor_list_t), intent(inout) :: this
    type(cell_system_t), intent(in) :: cells
    double precision :: cut
    integer(c_int32_t) :: n
    integer(c_int32_t), allocatable :: st(:)
    integer(c_int32_t) :: dummy(1)
    n = rmpcd_make_stencil(cells%a, cut, dummy, 0_c_int32_t)
    allocate(st(3*n))
    n = rmpcd_make_stencil(cells%a, cut, st, n)
    if (allocated(this%stencil)) deallocate(this%stencil)
    allocate(this%stencil(3, n))
    this%stencil = reshape(st, [3, int(n)])
    call put_i(this%d_stencil, st)
  end subroutine make_stencil

  !> update_list (src/neighbor_list.f90:76-157)
  subroutine update_list(this, system1, system2, radius, cells, lj)
    class(neighbor_list_t), intent(inout) :: this
    type(particle_system_t), intent(inout) :: system1, system2
    type(cell_system_t), intent(inout) :: cells
    double precision, intent(in) :: radius
    type(lj_params_t), intent(in), optional, target :: lj
    type(rmpcd_particles_t) :: p1, p2
    type(rmpcd_cells_t) :: c
    type(rmpcd_lj_t), target :: ljc
    type(c_ptr) :: ljp
    integer(c_int) :: rc
    integer, allocatable :: flat(:)
    call this%time_update%tic()
    call system1%upload('ps')
    call system2%upload('ps')
    p1 = system1%view()
    p2 = system2%view()
    call cells%to_device()
    c = cells%view()
    call ensure(this%d_list, int(this%Nmax, c_size_t)*this%Npoints*4)
    call ensure(this%d_n, int(this%Npoints, c_size_t)*4)
    ljp = c_null_ptr
    if (present(lj)) then
       ljc = lj%c
       ljp = c_loc(ljc)
    end if
    rc = rmpcd_update_list(cells%box(), c, p1, p2, radius, this%d_stencil%p, &
         int(size(this%stencil, 2), c_int32_t), ljp, int(this%Nmax, c_int32_t), this%d_list%p, this%d_n%p)
    if (rc == RMPCD_ERR_OVERFLOW) error stop 'maximum of neighbor list reached'
    call check(rc)
    allocate(flat(size(this%list)))
    call get_i_shift(this%d_list, flat)
    this%list = reshape(flat, shape(this%list))
    call get_i(this%d_n, this%n)
    call get_l(cells%d_is_md, cells%is_md)
    call get_l(cells%d_is_reac, cells%is_reac)
    call this%time_update%tac()
  end subroutine update_list

  !> compute_force (src/neighbor_list.f90:159-225)
  function compute_force(ps1, ps2, n_list, L, lj_params) result(e)
    type(particle_system_t), intent(inout) :: ps1, ps2
    type(neighbor_list_t), intent(inout) :: n_list
    double precision, intent(in) :: L(3)
    type(lj_params_t), intent(in) :: lj_params
    double precision :: e
    type(rmpcd_particles_t) :: p1, p2
    call n_list%time_force%tic()
    call ps1%upload('psf')
    call ps2%upload('psfg')
    p1 = ps1%view()
    p2 = ps2%view()
    call put_i_shift(n_list%d_list, reshape(n_list%list, [size(n_list%list)]))
    call put_i(n_list%d_n, n_list%n)
    call check(rmpcd_compute_force(p1, p2, n_list%d_list%p, n_list%d_n%p, int(n_list%Nmax, c_int32_t), L, &
         lj_params%c, e))
    call ps1%download('f')
    call ps2%download('fg')
    call n_list%time_force%tac()
  end function compute_force

  !> compute_force_n2 (src/neighbor_list.f90:227-264)
  function compute_force_n2(ps, L, lj_params) result(e)
    type(particle_system_t), intent(inout) :: ps
    double precision, intent(in) :: L(3)
    type(lj_params_t), intent(in) :: lj_params
    double precision :: e
    type(rmpcd_particles_t) :: p
    call ps%time_self_force%tic()
    call ps%upload('psf')
    p = ps%view()
    call check(rmpcd_compute_force_n2(p, L, lj_params%c, e))
    call ps%download('f')
    call ps%time_self_force%tac()
  end function compute_force_n2
end module neighbor_list

module mpcd
  use iso_c_binding
  use rmpcd_capi
  use rmpcd_mirror, only: dbuf_t, put_d, get_d, put_i, get_i
  use common
  use particle_system
  use cell_system
  use threefry_module
  implicit none
  private
  public :: simple_mpcd_step, compute_temperature, mpcd_stream_periodic, bulk_reaction
  public :: mpcd_stream_nogravity_zwall, rand_sphere, rescale_cells, bulk_reaction_endothermic
  public :: wall_mpcd_step, mpcd_stream_xforce_yzwall, compute_rho, compute_vx

contains

  !> rand_sphere (src/mpcd.f90:43-62), Marsaglia's method on the host stream
  function rand_sphere(state) result(n)
    type(threefry_rng_t), intent(inout) :: state
    double precision :: n(3), s, alpha
    do
       n(1) = 2*threefry_double(state) - 1
       n(2) = 2*threefry_double(state) - 1
       s = n(1)**2 + n(2)**2
       if (s < 1.d0) exit
    end do
    alpha = 2.d0*sqrt(1.d0 - s)
    n(1) = n(1)*alpha
    n(2) = n(2)*alpha
    n(3) = 1.d0 - 2.d0*s
  end function rand_sphere

  !> the device generator of a call: seed of state(1), one counter value per call
  subroutine device_draw_key(state, seed, step)
    type(threefry_rng_t), intent(inout) :: state(:)
    integer(c_int64_t), intent(out) :: seed, step
    seed = ieor(state(1)%key(1), shiftl(state(1)%key(2), 32))
    step = state(1)%ctr(2)
    state(1)%ctr(2) = state(1)%ctr(2) + 1
  end subroutine device_draw_key

  !> simple_mpcd_step (src/mpcd.f90:71-183): SRD rotation or, with thermostat, MPC-AT
  subroutine simple_mpcd_step(particles, cells, state, alpha, thermostat, T, hydro)
    class(particle_system_t), intent(inout) :: particles
    class(cell_system_t), intent(inout) :: cells
    type(threefry_rng_t), intent(inout) :: state(:)
    double precision, intent(in), optional :: alpha
    logical, intent(in), optional :: thermostat
    double precision, intent(in), optional :: T
    logical, intent(in), optional :: hydro
    type(rmpcd_particles_t) :: p
    integer(c_int64_t) :: seed, step
    integer(c_int32_t) :: ag, th, hy
    double precision :: a, tt
    ag = 0
    a = 0
    if (present(alpha)) then
       ag = 1
       a = alpha
    end if
    th = 0
    tt = 1
    if (present(thermostat)) then
       if (thermostat) th = 1
       if (.not. present(T)) stop 'no temperature given for thermostat in simple_mpcd_step'
       tt = T
    end if
    hy = 1
    if (present(hydro)) then
       if (.not. hydro) hy = 0
    end if
    call particles%time_step%tic()
    if (th == 1) then
       call particles%upload('vsd')  ! the thermostat's draws are keyed by particle id
    else
       call particles%upload('vs')
    end if
    p = particles%view()
    call cells%to_device()
    call device_draw_key(state, seed, step)
    call check(rmpcd_simple_mpcd_step(cells%box(), cells%view(), p, c_null_ptr, seed, step, ag, a, th, tt, hy))
    call particles%download('v')
    call particles%time_step%tac()
  end subroutine simple_mpcd_step

  !> compute_temperature (src/mpcd.f90:342-407); the z profile is not provided
  function compute_temperature(particles, cells, tz) result(te)
    type(particle_system_t), intent(inout) :: particles
    type(cell_system_t), intent(inout) :: cells
    type(profile_t), intent(inout), optional :: tz
    double precision :: te
    type(rmpcd_particles_t) :: p
    type(dbuf_t) :: d_data, d_count
    call particles%time_ct%tic()
    call particles%upload('vs')
    p = particles%view()
    call cells%to_device()
    if (present(tz)) then
       if (.not. (allocated(tz% data) .and. allocated(tz% count))) error stop 'profile_t: data not allocated'
       ! the profile's bins on the device, added to there, and back
       call put_d(d_data, tz% data)
       call put_i(d_count, tz% count)
       call check(rmpcd_compute_temperature_profile(cells%box(), cells%view(), p, te, tz% xmin, tz% dx, &
            int(tz% n, c_int32_t), d_data%p, d_count%p))
       call get_d(d_data, tz% data)
       call get_i(d_count, tz% count)
       call check(rmpcd_free(d_data%p))
       call check(rmpcd_free(d_count%p))
    else
       call check(rmpcd_compute_temperature(cells%box(), cells%view(), p, te))
    end if
    call particles%time_ct%tac()
  end function compute_temperature

  !> density profile along z: every particle's z binned (src/mpcd.f90:410-422)
  subroutine compute_rho(particles, rhoz)
    type(particle_system_t), intent(in) :: particles
    type(histogram_t), intent(inout) :: rhoz
    integer :: i
    if (.not. allocated(rhoz% data)) error stop 'histogram_t: data not allocated'
    do i = 1, particles% Nmax
       call rhoz% bin(particles% pos(3, i))
    end do
  end subroutine compute_rho

  !> x-velocity profile along z (src/mpcd.f90:424-437)
  subroutine compute_vx(particles, vx)
    type(particle_system_t), intent(in) :: particles
    type(profile_t), intent(inout) :: vx
    integer :: i
    if (.not. allocated(vx% data)) error stop 'histogram_t: data not allocated'
    do i = 1, particles% Nmax
       call vx% bin(particles% pos(3, i), particles% vel(1, i))
    end do
  end subroutine compute_vx

  !> mpcd_stream_periodic (src/mpcd.f90:440-457)
  subroutine mpcd_stream_periodic(particles, cells, dt)
    type(particle_system_t), intent(inout) :: particles
    type(cell_system_t), intent(in) :: cells
    double precision, intent(in) :: dt
    type(rmpcd_particles_t) :: p
    call particles%time_stream%tic()
    call particles%upload('pvsi')
    p = particles%view()
    call check(rmpcd_mpcd_stream_periodic(cells%box(), p, dt))
    call particles%download('pi')
    call particles%time_stream%tac()
  end subroutine mpcd_stream_periodic

  !> bulk_reaction (src/mpcd.f90:605-643)
  subroutine bulk_reaction(p, c, from, to, rate, tau, state)
    type(particle_system_t), intent(inout) :: p
    type(cell_system_t), intent(inout) :: c
    integer, intent(in) :: from, to
    double precision, intent(in) :: rate, tau
    type(threefry_rng_t), intent(inout) :: state(:)
    type(rmpcd_particles_t) :: v
    integer(c_int64_t) :: seed, step
    call p%upload('vs')
    v = p%view()
    call c%to_device()
    call device_draw_key(state, seed, step)
    call check(rmpcd_bulk_reaction(c%box(), c%view(), v, int(from, c_int32_t), int(to, c_int32_t), rate, tau, &
         c_null_ptr, seed, step))
    call p%download('s')
  end subroutine bulk_reaction

  !> bulk_reaction_endothermic (src/mpcd.f90:646-700)
  subroutine bulk_reaction_endothermic(p, c, from, to, rate, tau, state, delta_u)
    type(particle_system_t), intent(inout) :: p
    type(cell_system_t), intent(inout) :: c
    integer, intent(in) :: from, to
    double precision, intent(in) :: rate, tau, delta_u
    type(threefry_rng_t), intent(inout) :: state(:)
    type(rmpcd_particles_t) :: v
    integer(c_int64_t) :: seed, step
    call p%upload('vs')
    v = p%view()
    call c%to_device()
    call device_draw_key(state, seed, step)
    call check(rmpcd_bulk_reaction_endothermic(c%box(), c%view(), v, int(from, c_int32_t), int(to, c_int32_t), &
         rate, tau, delta_u, c_null_ptr, seed, step))
    call p%download('vs')
  end subroutine bulk_reaction_endothermic

  !> rescale_cells (src/mpcd.f90:748-801)
  subroutine rescale_cells(particles, cells, state, T)
    class(particle_system_t), intent(inout) :: particles
    class(cell_system_t), intent(inout) :: cells
    type(threefry_rng_t), intent(inout) :: state(:)
    double precision, intent(in) :: T
    type(rmpcd_particles_t) :: p
    call particles%time_step%tic()
    call particles%upload('vs')
    p = particles%view()
    call cells%to_device()
    call check(rmpcd_rescale_cells(cells%box(), cells%view(), p, T))
    call particles%download('v')
    call particles%time_step%tac()
  end subroutine rescale_cells

  !> mpcd_stream_nogravity_zwall (src/mpcd.f90:702-745)
  subroutine mpcd_stream_nogravity_zwall(particles, cells, dt)
    type(particle_system_t), intent(inout) :: particles
    type(cell_system_t), intent(in) :: cells
    double precision, intent(in) :: dt
    type(rmpcd_particles_t) :: p
    call particles%time_stream%tic()
    call particles%upload('pvfsg')
    p = particles%view()
    call check(rmpcd_mpcd_stream_nogravity_zwall(cells%box(), p, dt))
    call particles%download('pvg')
    call particles%time_stream%tac()
  end subroutine mpcd_stream_nogravity_zwall

  !> wall_mpcd_step (src/mpcd.f90:193-337); wall_v is accepted and, as in the reference,
  !! not used
  subroutine wall_mpcd_step(particles, cells, state, wall_temperature, wall_v, wall_n, &
       thermostat, bulk_temperature, alpha, keep_cell_v)
    class(particle_system_t), intent(inout) :: particles
    class(cell_system_t), intent(inout) :: cells
    type(threefry_rng_t), intent(inout) :: state(:)
    double precision, optional, intent(in) :: wall_temperature(2)
    double precision, optional, intent(in) :: wall_v(3,2)
    integer, optional, intent(in) :: wall_n(2)
    logical, intent(in), optional :: thermostat
    double precision, intent(in), optional :: bulk_temperature
    double precision, intent(in), optional :: alpha
    logical, intent(in), optional :: keep_cell_v
    double precision, target :: wt(2)
    integer(c_int32_t), target :: wn(2)
    type(rmpcd_particles_t) :: p
    integer(c_int64_t) :: seed, step
    integer(c_int32_t) :: wg, th, kc, ag
    double precision :: bt, a
    logical :: all_present, all_absent
    all_present = present(wall_temperature) .and. present(wall_v) .and. present(wall_n)
    all_absent = .not. present(wall_temperature) .and. .not. present(wall_v) .and. .not. present(wall_n)
    if ( .not. (all_present .or. all_absent) ) &
         error stop 'wall parameters must be all present or all absent in wall_mpcd_step'
    wg = 0
    wt = 0
    wn = 0
    if (all_present) then
       wg = 1
       wt = wall_temperature
       wn = int(wall_n, c_int32_t)
    end if
    th = 0
    bt = 1
    if (present(thermostat)) then
       if (thermostat) then
          th = 1
          if (.not. present(bulk_temperature)) &
               error stop 'thermostat requested but no temperature given in wall_mpcd_step'
          bt = bulk_temperature
       end if
    end if
    kc = 1
    if (present(keep_cell_v)) then
       if (.not. present(thermostat)) stop 'requiring keep_cell_v and not setting thermostat in wall_mpcd_step'
       if ((.not. keep_cell_v) .and. th == 0) &
            stop 'requiring .not. keep_cell_v and .not. thermostat in wall_mpcd_step'
       if (.not. keep_cell_v) kc = 0
    end if
    ag = 0
    a = 0
    if (present(alpha)) then
       ag = 1
       a = alpha
    end if
    call particles%time_step%tic()
    if (th == 1) then
       call particles%upload('vsd')  ! the thermostat's draws are keyed by particle id
    else
       call particles%upload('vs')
    end if
    p = particles%view()
    call cells%to_device()
    call device_draw_key(state, seed, step)
    call check(rmpcd_wall_mpcd_step(cells%box(), cells%view(), p, wg, c_loc(wt), c_loc(wn), th, bt, kc, ag, a, &
         c_null_ptr, c_null_ptr, c_null_ptr, seed, step))
    call particles%download('v')
    call particles%time_step%tac()
  end subroutine wall_mpcd_step

  !> mpcd_stream_xforce_yzwall (src/mpcd.f90:459-511)
  subroutine mpcd_stream_xforce_yzwall(particles, cells, dt, g)
    type(particle_system_t), intent(inout) :: particles
    type(cell_system_t), intent(in) :: cells
    double precision, intent(in) :: dt
    double precision, intent(in) :: g
    type(rmpcd_particles_t) :: p
    call particles%time_stream%tic()
    call particles%upload('pvsg')
    p = particles%view()
    call check(rmpcd_mpcd_stream_xforce_yzwall(cells%box(), p, dt, g))
    call particles%download('pvg')
    call particles%time_stream%tac()
  end subroutine mpcd_stream_xforce_yzwall
end module mpcd

module md
  use iso_c_binding
  use rmpcd_capi
  use rmpcd_mirror
  use common
  use particle_system
  use cell_system
  implicit none
  private
  public :: md_pos, md_vel, apply_pbc, cell_md_pos, cell_md_vel, cell_md_pos_ywall, cell_md_pos_zwall
  public :: rattle_dimer_pos, rattle_dimer_vel
  public :: rattle_body_pos, rattle_body_vel, elastic_network, rigid_body_t

  !> rigid_body_t (src/md.f90:39-57); the fields are the host copy of the device state,
  !! written back after every call and read before it
  type rigid_body_t
     integer :: i_start
     integer :: i_stop
     double precision :: mass
     double precision :: pos(3)
     double precision, allocatable :: pos_body(:,:)
     double precision :: vel(3)
     double precision :: force(3)
     double precision :: q(4)
     double precision :: L(3)
     double precision :: torque(3)
     double precision :: I_body(3)
     double precision :: omega_body(3)
     type(dbuf_t), private :: d_body, d_pos_body
   contains
     procedure :: init => rigid_body_init
     procedure :: compute_force_torque => rigid_body_compute_force_torque
     procedure :: vv1 => rigid_body_vv1
     procedure :: vv2 => rigid_body_vv2
  end type rigid_body_t

contains

  !> md_pos (src/md.f90:61-77)
  subroutine md_pos(particles, dt)
    type(particle_system_t), intent(inout) :: particles
    double precision, intent(in) :: dt
    type(rmpcd_particles_t) :: p
    call particles%time_md_pos%tic()
    call particles%upload('pvfs')
    p = particles%view()
    call check(rmpcd_md_pos(p, dt))
    call particles%download('p')
    call particles%time_md_pos%tac()
  end subroutine md_pos

  !> md_vel (src/md.f90:219-237)
  subroutine md_vel(particles, dt)
    type(particle_system_t), intent(inout) :: particles
    double precision, intent(in) :: dt
    type(rmpcd_particles_t) :: p
    call particles%time_md_vel%tic()
    call particles%upload('vfosg')
    p = particles%view()
    call check(rmpcd_md_vel(p, dt))
    call particles%download('vg')
    call particles%time_md_vel%tac()
  end subroutine md_vel

  !> apply_pbc (src/md.f90:202-217)
  subroutine apply_pbc(particles, edges)
    type(particle_system_t), intent(inout) :: particles
    double precision, intent(in) :: edges(3)
    type(rmpcd_particles_t) :: p
    call particles%time_apply_pbc%tic()
    call particles%upload('pvsgi')
    p = particles%view()
    call check(rmpcd_apply_pbc(p, edges))
    call particles%download('pi')
    call particles%time_apply_pbc%tac()
  end subroutine apply_pbc

  subroutine cell_md_common(cells, particles, dt, md_flag, axis)
    type(cell_system_t), intent(inout) :: cells
    type(particle_system_t), intent(inout) :: particles
    double precision, intent(in) :: dt
    logical, intent(in) :: md_flag
    integer, intent(in) :: axis
    type(rmpcd_particles_t) :: p
    integer(c_int32_t) :: f
    f = merge(1, 0, md_flag)
    call particles%upload('pvfsg')
    p = particles%view()
    call cells%to_device()
    if (axis == 0) then
       call check(rmpcd_cell_md_pos(cells%box(), cells%view(), p, dt, f))
    else
       call check(rmpcd_cell_md_pos_wall(cells%box(), cells%view(), p, dt, f, int(axis, c_int32_t)))
    end if
    call particles%download('pv')
  end subroutine cell_md_common

  !> cell_md_pos (src/md.f90:79-102)
  subroutine cell_md_pos(cells, particles, dt, md_flag)
    type(cell_system_t), intent(inout) :: cells
    type(particle_system_t), intent(inout) :: particles
    double precision, intent(in) :: dt
    logical, intent(in) :: md_flag
    call particles%time_md_pos%tic()
    call cell_md_common(cells, particles, dt, md_flag, 0)
    call particles%time_md_pos%tac()
  end subroutine cell_md_pos

  !> cell_md_pos_ywall (src/md.f90:104-151)
  subroutine cell_md_pos_ywall(cells, particles, dt, md_flag)
    type(cell_system_t), intent(inout) :: cells
    type(particle_system_t), intent(inout) :: particles
    double precision, intent(in) :: dt
    logical, intent(in) :: md_flag
    call particles%time_md_pos%tic()
    call cell_md_common(cells, particles, dt, md_flag, 1)
    call particles%time_md_pos%tac()
  end subroutine cell_md_pos_ywall

  !> cell_md_pos_zwall (src/md.f90:153-200)
  subroutine cell_md_pos_zwall(cells, particles, dt, md_flag)
    type(cell_system_t), intent(inout) :: cells
    type(particle_system_t), intent(inout) :: particles
    double precision, intent(in) :: dt
    logical, intent(in) :: md_flag
    call particles%time_md_pos%tic()
    call cell_md_common(cells, particles, dt, md_flag, 2)
    call particles%time_md_pos%tac()
  end subroutine cell_md_pos_zwall

  !> cell_md_vel (src/md.f90:239-263)
  subroutine cell_md_vel(cells, particles, dt, md_flag)
    type(cell_system_t), intent(inout) :: cells
    type(particle_system_t), intent(inout) :: particles
    double precision, intent(in) :: dt
    logical, intent(in) :: md_flag
    type(rmpcd_particles_t) :: p
    call particles%time_md_vel%tic()
    call particles%upload('vfosg')
    p = particles%view()
    call cells%to_device()
    call check(rmpcd_cell_md_vel(cells%box(), cells%view(), p, dt, merge(1, 0, md_flag)))
    call particles%download('vg')
    call particles%time_md_vel%tac()
  end subroutine cell_md_vel

  !> rattle_dimer_pos (src/md.f90:265-296)
  subroutine rattle_dimer_pos(p, d, dt, edges)
    type(particle_system_t), intent(inout) :: p
    double precision, intent(in) :: d, dt, edges(3)
    type(rmpcd_particles_t) :: v
    call p%time_rattle_pos%tic()
    call p%upload('pvs')
    v = p%view()
    call put_vec(p%d_aux, p%pos_rattle)
    call check(rmpcd_rattle_dimer_pos(v, p%d_aux%p, p%mass, int(size(p%mass), c_int32_t), d, dt, edges))
    call p%download('pv')
    call p%time_rattle_pos%tac()
  end subroutine rattle_dimer_pos

  !> rattle_dimer_vel (src/md.f90:298-319)
  subroutine rattle_dimer_vel(p, d, dt, edges)
    type(particle_system_t), intent(inout) :: p
    double precision, intent(in) :: d, dt, edges(3)
    type(rmpcd_particles_t) :: v
    call p%time_rattle_vel%tic()
    call p%upload('pvs')
    v = p%view()
    call check(rmpcd_rattle_dimer_vel(v, p%mass, int(size(p%mass), c_int32_t), d, dt, edges))
    call p%download('v')
    call p%time_rattle_vel%tac()
  end subroutine rattle_dimer_vel

  !> the link list with zero-based indices, as the C ABI takes it
  function links0(links) result(l0)
    integer, intent(in) :: links(:,:)
    integer(c_int32_t), allocatable :: l0(:)
    l0 = int(reshape(links, [size(links)]) - 1, c_int32_t)
  end function links0

  !> rattle_body_pos (src/md.f90:321-408); pos_rattle comes back unfolded as in the reference
  subroutine rattle_body_pos(p, links, distances, dt, edges, precision)
    type(particle_system_t), intent(inout) :: p
    integer, intent(in) :: links(:,:)
    double precision, intent(in) :: distances(:)
    double precision, intent(in) :: dt, edges(3), precision
    type(rmpcd_particles_t) :: v
    integer(c_int32_t) :: it
    call p%time_rattle_pos%tic()
    call p%upload('pvs')
    v = p%view()
    call put_vec(p%d_aux, p%pos_rattle)
    call check(rmpcd_rattle_body_pos(v, p%d_aux%p, links0(links), distances, int(size(distances), c_int32_t), &
         p%mass, int(size(p%mass), c_int32_t), dt, edges, precision, it))
    call p%download('pv')
    call get_vec(p%d_aux, p%pos_rattle)
    call p%time_rattle_pos%tac()
    if (it >= 1000) write(*,*) 'rattle_max reached in rattle_body_pos'
  end subroutine rattle_body_pos

  !> rattle_body_vel (src/md.f90:410-480)
  subroutine rattle_body_vel(p, links, distances, dt, edges, precision)
    type(particle_system_t), intent(inout) :: p
    integer, intent(in) :: links(:,:)
    double precision, intent(in) :: distances(:)
    double precision, intent(in) :: dt, edges(3), precision
    type(rmpcd_particles_t) :: v
    integer(c_int32_t) :: it
    call p%time_rattle_vel%tic()
    call p%upload('pvs')
    v = p%view()
    call check(rmpcd_rattle_body_vel(v, links0(links), distances, int(size(distances), c_int32_t), &
         p%mass, int(size(p%mass), c_int32_t), dt, edges, precision, it))
    call p%download('pv')
    call p%time_rattle_vel%tac()
    if (it >= 1000) write(*,*) 'rattle_max reached in rattle_body_vel'
  end subroutine rattle_body_vel

  !> elastic_network (src/md.f90:516-544)
  function elastic_network(p, links, distances, k, edges) result(e)
    type(particle_system_t), intent(inout) :: p
    integer, intent(in) :: links(:,:)
    double precision, intent(in) :: distances(:)
    double precision, intent(in) :: k
    double precision, intent(in) :: edges(3)
    double precision :: e
    type(rmpcd_particles_t) :: v
    call p%time_elastic%tic()
    call p%upload('pfo')
    v = p%view()
    call check(rmpcd_elastic_network(v, links0(links), distances, int(size(distances), c_int32_t), k, edges, e))
    call p%download('f')
    call p%time_elastic%tac()
  end function elastic_network

  subroutine body_push(this)
    class(rigid_body_t), intent(inout) :: this
    type(rmpcd_rigid_body_t), target :: h
    h%i_start = this%i_start - 1
    h%i_stop = this%i_stop - 1
    h%mass = this%mass
    h%pos = this%pos
    h%vel = this%vel
    h%force = this%force
    h%q = this%q
    h%L = this%L
    h%torque = this%torque
    h%I_body = this%I_body
    h%omega_body = this%omega_body
    h%pos_body = this%d_pos_body%p
    call check(rmpcd_memcpy(this%d_body%p, c_loc(h), int(c_sizeof(h), c_size_t), H2D))
  end subroutine body_push

  subroutine body_pull(this)
    class(rigid_body_t), intent(inout) :: this
    type(rmpcd_rigid_body_t), target :: h
    call check(rmpcd_memcpy(c_loc(h), this%d_body%p, int(c_sizeof(h), c_size_t), D2H))
    this%mass = h%mass
    this%pos = h%pos
    this%vel = h%vel
    this%force = h%force
    this%q = h%q
    this%L = h%L
    this%torque = h%torque
    this%I_body = h%I_body
    this%omega_body = h%omega_body
  end subroutine body_pull

  !> rigid_body_init (src/md.f90:547-588); pos_body(:, i - i_start + 1) is particle i
  subroutine rigid_body_init(this, ps, i_start, i_stop, edges)
    class(rigid_body_t), intent(inout) :: this
    type(particle_system_t), intent(inout) :: ps
    integer, intent(in) :: i_start
    integer, intent(in) :: i_stop
    double precision, intent(in) :: edges(3)
    type(rmpcd_particles_t) :: v
    type(rmpcd_rigid_body_t) :: h
    this%i_start = i_start
    this%i_stop = i_stop
    call ensure(this%d_body, int(c_sizeof(h), c_size_t))
    call ensure(this%d_pos_body, int(3*(i_stop - i_start + 1), c_size_t)*8)
    call ps%upload('pvsi')
    v = ps%view()
    call check(rmpcd_rigid_body_init(this%d_body%p, v, ps%mass, int(size(ps%mass), c_int32_t), edges, &
         int(i_start - 1, c_int32_t), int(i_stop - 1, c_int32_t), this%d_pos_body%p))
    call body_pull(this)
    if (allocated(this%pos_body)) deallocate(this%pos_body)
    allocate(this%pos_body(3, i_stop - i_start + 1))
    call get_rows(this%d_pos_body, this%pos_body)
  end subroutine rigid_body_init

  !> body-frame rows (3, n) stored contiguously on the device
  subroutine get_rows(b, a)
    type(dbuf_t), intent(in) :: b
    double precision, intent(inout), target :: a(:,:)
    call check(rmpcd_memcpy(c_loc(a), b%p, int(size(a), c_size_t)*8, D2H))
  end subroutine get_rows

  !> rigid_body_compute_force_torque (src/md.f90:591-607)
  subroutine rigid_body_compute_force_torque(this, ps)
    class(rigid_body_t), intent(inout) :: this
    type(particle_system_t), intent(inout) :: ps
    type(rmpcd_particles_t) :: v
    call ps%upload('fo')
    v = ps%view()
    call body_push(this)
    call check(rmpcd_rigid_body_compute_force_torque(this%d_body%p, v))
    call body_pull(this)
  end subroutine rigid_body_compute_force_torque

  !> rigid_body_vv1 (src/md.f90:613-673)
  subroutine rigid_body_vv1(this, ps, dt, treshold, edges)
    class(rigid_body_t), intent(inout) :: this
    type(particle_system_t), intent(inout) :: ps
    double precision, intent(in) :: dt
    double precision, intent(in) :: treshold
    double precision, intent(in) :: edges(3)
    type(rmpcd_particles_t) :: v
    integer(c_int32_t) :: it
    call ps%upload('p')
    v = ps%view()
    call body_push(this)
    call check(rmpcd_rigid_body_vv1(this%d_body%p, v, dt, treshold, edges, it))
    call body_pull(this)
    call ps%download('p')
  end subroutine rigid_body_vv1

  !> rigid_body_vv2 (src/md.f90:679-700)
  subroutine rigid_body_vv2(this, ps, dt)
    class(rigid_body_t), intent(inout) :: this
    type(particle_system_t), intent(inout) :: ps
    double precision, intent(in) :: dt
    type(rmpcd_particles_t) :: v
    v = ps%view()
    call body_push(this)
    call check(rmpcd_rigid_body_vv2(this%d_body%p, v, dt))
    call body_pull(this)
    call ps%download('v')
  end subroutine rigid_body_vv2
end module md

!> the umbrella module the drivers use (src/rmpcdmd_module.f90.in)
