! h5md_module -- the interface of the H5MD writer/reader that the reference's drivers and
! tests use (h5md_file_t, h5md_element_t, h5md_write_attribute and the storage-mode
! constants; e.g. test/setup_simple_fluid.f90:26-29, 65-79, 119-123), over a plain-text
! file instead of HDF5, which this library does not depend on.  The reference's module
! is an absent submodule (fortran_h5md); this one keeps its names and argument meaning.
!
! File format (written to <name>.txt, one record per line, values one per line in
! Fortran array order, doubles with 17 significant digits):
!   H5MD-TEXT 1
!   creator <creator> <version>        author <author>
!   group <path>
!   attribute <object path> <name> <int|double|string> <count>, then the values
!   fixed <path> <int|double> <rank> <dims...>, then the values
!   time <path> <int|double> <rank> <dims...> mode <mode> step <step> time <time>
!   frame <path> <step> <time>, then the values of one append
module h5md_module
  use hdf5
  implicit none
  private

  public :: h5md_file_t, h5md_element_t, h5md_write_attribute, h5md_check_valid, h5md_check_exists
  public :: read_frame_d2, h5md_write_dataset
  ! storage modes (bit flags, combined with ior as the drivers do)
  integer, parameter, public :: H5MD_TIME = 1, H5MD_LINEAR = 2, H5MD_FIXED = 4, H5MD_STORE_TIME = 8

  type h5md_file_t
     integer(HID_T) :: id = -1
     integer(HID_T) :: particles = -1, observables = -1, parameters = -1, connectivity = -1
     integer :: error = 0
     integer :: f = 0
   contains
     procedure :: create => file_create
     procedure :: open => file_open
     procedure :: close => file_close
  end type h5md_file_t

  type h5md_element_t
     integer(HID_T) :: id = -1
     integer(HID_T) :: v = -1
     integer :: error = 0
     integer :: Nmax = 0
     integer :: mode = 0
     integer :: step = 0, step_offset = 0, n_frames = 0
     double precision :: time = 0, time_offset = 0
     integer :: f = 0
     integer :: reg = 0  ! its record in the table of time-dependent elements being written
     character(len=512) :: path = ''
   contains
     generic :: create_fixed => fixed_d0, fixed_i0, fixed_d1, fixed_i1, fixed_d2, fixed_i2
     procedure, private :: fixed_d0, fixed_i0, fixed_d1, fixed_i1, fixed_d2, fixed_i2
     generic :: create_time => time_d0, time_i0, time_d1, time_i1, time_d2, time_i2
     procedure, private :: time_d0, time_i0, time_d1, time_i1, time_d2, time_i2
     generic :: append => app_d0, app_i0, app_d1, app_i1, app_d2, app_i2
     procedure, private :: app_d0, app_i0, app_d1, app_i1, app_d2, app_i2
     generic :: append_buffer => buf_d1, buf_d2
     procedure, private :: buf_d1, buf_d2
     generic :: read_fixed => read_fixed_d2, read_fixed_i1
     procedure, private :: read_fixed_d2, read_fixed_i1
     procedure :: open_time
     procedure :: close => element_close
  end type h5md_element_t

  ! the time-dependent elements of the files being written: an element closed and opened
  ! again (open_time) goes on with its frame count and storage mode
  type time_rec_t
     integer :: f = 0
     character(len=512) :: path = ''
     integer :: mode = 0, step = 0, step_offset = 0, n_frames = 0, Nmax = 0
     double precision :: time = 0, time_offset = 0
  end type time_rec_t
  type(time_rec_t), allocatable, save :: recs(:)
  integer, save :: n_recs = 0

  interface h5md_write_dataset
     module procedure :: dset_d0, dset_i0, dset_d1, dset_i1
  end interface h5md_write_dataset

  interface h5md_write_attribute
     module procedure :: attr_i0, attr_d0, attr_s0, attr_s1, attr_i1, attr_d1
  end interface h5md_write_attribute

contains

  ! ---- files ------------------------------------------------------------------------
  subroutine file_create(this, filename, creator, creator_version, author, author_email)
    class(h5md_file_t), intent(out) :: this
    character(len=*), intent(in) :: filename, creator, creator_version, author
    character(len=*), intent(in), optional :: author_email
    integer :: u, e
    call h5open_f(e)
    open(newunit=u, file=trim(filename)//'.txt', status='replace', action='write', iostat=this% error)
    if (this% error /= 0) error stop 'h5md_file_t: cannot create file'
    write(u, '(a)') 'H5MD-TEXT 1'
    write(u, '(a,1x,a,1x,a)') 'creator', trim(creator), trim(creator_version)
    if (present(author_email)) then
       write(u, '(a,1x,a,1x,a)') 'author', trim(author), trim(author_email)
    else
       write(u, '(a,1x,a)') 'author', trim(author)
    end if
    this% f = h5t_new_file(trim(filename)//'.txt', u, .true.)
    call roots(this)
  end subroutine file_create

  subroutine file_open(this, filename, flags)
    class(h5md_file_t), intent(out) :: this
    character(len=*), intent(in) :: filename
    integer, intent(in) :: flags
    integer :: u, e
    character(len=64) :: head
    character(len=256) :: msg
    call h5open_f(e)
    if (flags /= H5F_ACC_RDONLY_F) error stop 'h5md_file_t: only read-only opening is provided'
    open(newunit=u, file=trim(filename)//'.txt', status='old', action='read', iostat=this% error, iomsg=msg)
    if (this% error /= 0) then
       write(*, '(a)') trim(msg)
       error stop 'h5md_file_t: cannot open file'
    end if
    read(u, '(a)', iostat=e) head
    if (e /= 0 .or. head(1:11) /= 'H5MD-TEXT 1') error stop 'h5md_file_t: not an H5MD text file'
    this% f = h5t_new_file(trim(filename)//'.txt', u, .false.)
    call roots(this)
  end subroutine file_open

  subroutine roots(this)
    type(h5md_file_t), intent(inout) :: this
    this% id = h5t_new_object(this% f, '')
    this% particles = h5t_new_object(this% f, '/particles')
    this% observables = h5t_new_object(this% f, '/observables')
    this% parameters = h5t_new_object(this% f, '/parameters')
    this% connectivity = h5t_new_object(this% f, '/connectivity')
    this% error = 0
  end subroutine roots

  subroutine file_close(this)
    class(h5md_file_t), intent(inout) :: this
    integer :: i
    do i = 1, n_recs
       if (recs(i)% f == this% f) recs(i)% f = 0
    end do
    call h5t_close_file(this% f)
    this% f = 0
    this% id = -1
  end subroutine file_close

  subroutine h5md_check_valid(loc, msg)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: msg
    if (h5t_file(loc) <= 0) then
       write(*, *) msg
       error stop 'h5md: invalid object'
    end if
  end subroutine h5md_check_valid

  ! an object exists when a record of the file names its path (or a path below it)
  subroutine h5md_check_exists(loc, name, msg)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name, msg
    if (.not. path_in_file(h5t_file(loc), trim(h5t_path(loc))//'/'//trim(name))) then
       write(*, *) msg
       error stop 'h5md: object not found'
    end if
  end subroutine h5md_check_exists

  logical function path_in_file(f, path) result(found)
    integer, intent(in) :: f
    character(len=*), intent(in) :: path
    character(len=4096) :: line
    integer :: u, e, k
    found = .false.
    u = h5t_unit(f)
    if (u == 0) return
    rewind(u)
    do
       read(u, '(a)', iostat=e) line
       if (e /= 0) exit
       k = index(line, ' ')
       if (k == 0) cycle
       if (index(line(k + 1:), trim(path)) == 1) then
          found = .true.
          exit
       end if
    end do
  end function path_in_file

  ! ---- writing ----------------------------------------------------------------------
  integer function unit_of(loc, f) result(u)
    integer(HID_T), intent(in) :: loc
    integer, intent(out) :: f
    f = h5t_file(loc)
    if (f <= 0) error stop 'h5md: invalid location'
    u = h5t_unit(f)
  end function unit_of

  subroutine put_d(u, x)
    integer, intent(in) :: u
    double precision, intent(in) :: x(:)
    integer :: i
    do i = 1, size(x)
       write(u, '(es25.17e3)') x(i)
    end do
  end subroutine put_d

  subroutine put_i(u, x)
    integer, intent(in) :: u
    integer, intent(in) :: x(:)
    integer :: i
    do i = 1, size(x)
       write(u, '(i0)') x(i)
    end do
  end subroutine put_i

  subroutine header(this, loc, name, kind, tag, dims)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name, kind, tag
    integer, intent(in) :: dims(:)
    integer :: u, i
    u = unit_of(loc, this% f)
    this% path = trim(h5t_path(loc))//'/'//trim(name)
    this% id = h5t_new_object(this% f, trim(this% path))
    this% v = this% id
    write(u, '(a,1x,a,1x,a,1x,i0)', advance='no') tag, trim(this% path), kind, size(dims)
    do i = 1, size(dims)
       write(u, '(1x,i0)', advance='no') dims(i)
    end do
    if (size(dims) > 0) this% Nmax = dims(size(dims))
  end subroutine header

  subroutine fixed_d0(this, loc, name, data)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    double precision, intent(in) :: data
    integer :: empty(0)
    call header(this, loc, name, 'double', 'fixed', empty)
    write(h5t_unit(this% f), '(a)') ''
    call put_d(h5t_unit(this% f), [data])
  end subroutine fixed_d0
  subroutine fixed_i0(this, loc, name, data)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    integer, intent(in) :: data
    integer :: empty(0)
    call header(this, loc, name, 'int', 'fixed', empty)
    write(h5t_unit(this% f), '(a)') ''
    call put_i(h5t_unit(this% f), [data])
  end subroutine fixed_i0
  subroutine fixed_d1(this, loc, name, data)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    double precision, intent(in) :: data(:)
    call header(this, loc, name, 'double', 'fixed', shape(data))
    write(h5t_unit(this% f), '(a)') ''
    call put_d(h5t_unit(this% f), data)
  end subroutine fixed_d1
  subroutine fixed_i1(this, loc, name, data)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    integer, intent(in) :: data(:)
    call header(this, loc, name, 'int', 'fixed', shape(data))
    write(h5t_unit(this% f), '(a)') ''
    call put_i(h5t_unit(this% f), data)
  end subroutine fixed_i1
  subroutine fixed_d2(this, loc, name, data)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    double precision, intent(in) :: data(:, :)
    call header(this, loc, name, 'double', 'fixed', shape(data))
    write(h5t_unit(this% f), '(a)') ''
    call put_d(h5t_unit(this% f), reshape(data, [size(data)]))
  end subroutine fixed_d2
  subroutine fixed_i2(this, loc, name, data)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    integer, intent(in) :: data(:, :)
    call header(this, loc, name, 'int', 'fixed', shape(data))
    write(h5t_unit(this% f), '(a)') ''
    call put_i(h5t_unit(this% f), reshape(data, [size(data)]))
  end subroutine fixed_i2

  ! a time-dependent element: H5MD_TIME stores the step (and time) given to each append,
  ! H5MD_LINEAR the step and time increments given here, with the offsets of the first frame
  subroutine time_meta(this, mode, step, time, step_offset, time_offset)
    class(h5md_element_t), intent(inout) :: this
    integer, intent(in) :: mode
    integer, intent(in), optional :: step, step_offset
    double precision, intent(in), optional :: time, time_offset
    this% mode = mode
    this% step = 0
    this% time = 0
    this% step_offset = 0
    this% time_offset = 0
    this% n_frames = 0
    if (present(step)) this% step = step
    if (present(time)) this% time = time
    if (present(step_offset)) this% step_offset = step_offset
    if (present(time_offset)) this% time_offset = time_offset
    if (iand(mode, H5MD_LINEAR) /= 0 .and. .not. present(step)) error stop 'h5md: step missing for H5MD_LINEAR'
    write(h5t_unit(this% f), '(a,1x,i0,1x,a,1x,i0,1x,a,1x,es25.17e3)') ' mode', mode, 'step', this% step, &
         'time', this% time
    call register(this)
  end subroutine time_meta

  subroutine register(this)
    class(h5md_element_t), intent(inout) :: this
    type(time_rec_t), allocatable :: grown(:)
    if (.not. allocated(recs)) allocate(recs(16))
    if (n_recs == size(recs)) then
       allocate(grown(2 * size(recs)))
       grown(1:n_recs) = recs(1:n_recs)
       call move_alloc(grown, recs)
    end if
    n_recs = n_recs + 1
    recs(n_recs)% f = this% f
    recs(n_recs)% path = this% path
    recs(n_recs)% mode = this% mode
    recs(n_recs)% step = this% step
    recs(n_recs)% time = this% time
    recs(n_recs)% step_offset = this% step_offset
    recs(n_recs)% time_offset = this% time_offset
    recs(n_recs)% n_frames = 0
    recs(n_recs)% Nmax = this% Nmax
    this% reg = n_recs
  end subroutine register

  subroutine time_d0(this, loc, name, data, mode, step, time, step_offset, time_offset)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    double precision, intent(in) :: data
    integer, intent(in) :: mode
    integer, intent(in), optional :: step, step_offset
    double precision, intent(in), optional :: time, time_offset
    integer :: empty(0)
    call header(this, loc, name, 'double', 'time', empty)
    call time_meta(this, mode, step, time, step_offset, time_offset)
    if (.false.) print *, data
  end subroutine time_d0
  subroutine time_i0(this, loc, name, data, mode, step, time, step_offset, time_offset)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    integer, intent(in) :: data
    integer, intent(in) :: mode
    integer, intent(in), optional :: step, step_offset
    double precision, intent(in), optional :: time, time_offset
    integer :: empty(0)
    call header(this, loc, name, 'int', 'time', empty)
    call time_meta(this, mode, step, time, step_offset, time_offset)
    if (.false.) print *, data
  end subroutine time_i0
  subroutine time_d1(this, loc, name, data, mode, step, time, step_offset, time_offset)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    double precision, intent(in) :: data(:)
    integer, intent(in) :: mode
    integer, intent(in), optional :: step, step_offset
    double precision, intent(in), optional :: time, time_offset
    call header(this, loc, name, 'double', 'time', shape(data))
    call time_meta(this, mode, step, time, step_offset, time_offset)
  end subroutine time_d1
  subroutine time_i1(this, loc, name, data, mode, step, time, step_offset, time_offset)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    integer, intent(in) :: data(:)
    integer, intent(in) :: mode
    integer, intent(in), optional :: step, step_offset
    double precision, intent(in), optional :: time, time_offset
    call header(this, loc, name, 'int', 'time', shape(data))
    call time_meta(this, mode, step, time, step_offset, time_offset)
  end subroutine time_i1
  subroutine time_d2(this, loc, name, data, mode, step, time, step_offset, time_offset)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    double precision, intent(in) :: data(:, :)
    integer, intent(in) :: mode
    integer, intent(in), optional :: step, step_offset
    double precision, intent(in), optional :: time, time_offset
    call header(this, loc, name, 'double', 'time', shape(data))
    call time_meta(this, mode, step, time, step_offset, time_offset)
  end subroutine time_d2
  subroutine time_i2(this, loc, name, data, mode, step, time, step_offset, time_offset)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    integer, intent(in) :: data(:, :)
    integer, intent(in) :: mode
    integer, intent(in), optional :: step, step_offset
    double precision, intent(in), optional :: time, time_offset
    call header(this, loc, name, 'int', 'time', shape(data))
    call time_meta(this, mode, step, time, step_offset, time_offset)
  end subroutine time_i2

  ! the frame line of one append: its step and time (given, or from the linear mode)
  subroutine frame(this, step, time)
    class(h5md_element_t), intent(inout) :: this
    integer, intent(in), optional :: step
    double precision, intent(in), optional :: time
    integer :: s
    double precision :: t
    if (this% f <= 0 .or. this% reg <= 0) error stop 'h5md: append to an element that is not open'
    this% n_frames = recs(this% reg)% n_frames
    recs(this% reg)% n_frames = this% n_frames + 1
    if (iand(this% mode, H5MD_LINEAR) /= 0) then
       s = this% step_offset + this% n_frames * this% step
       t = this% time_offset + this% n_frames * this% time
    else
       if (.not. present(step)) error stop 'h5md: step missing for H5MD_TIME'
       s = step
       t = 0
       if (present(time)) t = time
    end if
    this% n_frames = this% n_frames + 1
    write(h5t_unit(this% f), '(a,1x,a,1x,i0,1x,es25.17e3)') 'frame', trim(this% path), s, t
  end subroutine frame

  ! the first force_size values of a buffer, one frame each (scalars, or columns)
  subroutine buf_d1(this, data, force_size)
    class(h5md_element_t), intent(inout) :: this
    double precision, intent(in) :: data(:)
    integer, intent(in), optional :: force_size
    integer :: i, n
    n = size(data)
    if (present(force_size)) n = min(n, force_size)
    do i = 1, n
       call app_d0(this, data(i))
    end do
  end subroutine buf_d1
  subroutine buf_d2(this, data, force_size)
    class(h5md_element_t), intent(inout) :: this
    double precision, intent(in) :: data(:, :)
    integer, intent(in), optional :: force_size
    integer :: i, n
    n = size(data, 2)
    if (present(force_size)) n = min(n, force_size)
    do i = 1, n
       call app_d1(this, data(:, i))
    end do
  end subroutine buf_d2

  subroutine app_d0(this, data, step, time)
    class(h5md_element_t), intent(inout) :: this
    double precision, intent(in) :: data
    integer, intent(in), optional :: step
    double precision, intent(in), optional :: time
    call frame(this, step, time)
    call put_d(h5t_unit(this% f), [data])
  end subroutine app_d0
  subroutine app_i0(this, data, step, time)
    class(h5md_element_t), intent(inout) :: this
    integer, intent(in) :: data
    integer, intent(in), optional :: step
    double precision, intent(in), optional :: time
    call frame(this, step, time)
    call put_i(h5t_unit(this% f), [data])
  end subroutine app_i0
  subroutine app_d1(this, data, step, time)
    class(h5md_element_t), intent(inout) :: this
    double precision, intent(in) :: data(:)
    integer, intent(in), optional :: step
    double precision, intent(in), optional :: time
    call frame(this, step, time)
    call put_d(h5t_unit(this% f), data)
  end subroutine app_d1
  subroutine app_i1(this, data, step, time)
    class(h5md_element_t), intent(inout) :: this
    integer, intent(in) :: data(:)
    integer, intent(in), optional :: step
    double precision, intent(in), optional :: time
    call frame(this, step, time)
    call put_i(h5t_unit(this% f), data)
  end subroutine app_i1
  subroutine app_d2(this, data, step, time)
    class(h5md_element_t), intent(inout) :: this
    double precision, intent(in) :: data(:, :)
    integer, intent(in), optional :: step
    double precision, intent(in), optional :: time
    call frame(this, step, time)
    call put_d(h5t_unit(this% f), reshape(data, [size(data)]))
  end subroutine app_d2
  subroutine app_i2(this, data, step, time)
    class(h5md_element_t), intent(inout) :: this
    integer, intent(in) :: data(:, :)
    integer, intent(in), optional :: step
    double precision, intent(in), optional :: time
    call frame(this, step, time)
    call put_i(h5t_unit(this% f), reshape(data, [size(data)]))
  end subroutine app_i2

  subroutine element_close(this)
    class(h5md_element_t), intent(inout) :: this
    integer :: e
    if (this% id > 0) call h5gclose_f(this% id, e)
    this% id = -1
    this% v = -1
    this% f = 0
  end subroutine element_close

  ! ---- plain datasets (a fixed element without the element) -----------------------------
  subroutine dset_d0(loc, name, value)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    double precision, intent(in) :: value
    type(h5md_element_t) :: e
    call e% create_fixed(loc, name, value)
    call e% close()
  end subroutine dset_d0
  subroutine dset_i0(loc, name, value)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    integer, intent(in) :: value
    type(h5md_element_t) :: e
    call e% create_fixed(loc, name, value)
    call e% close()
  end subroutine dset_i0
  subroutine dset_d1(loc, name, value)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    double precision, intent(in) :: value(:)
    type(h5md_element_t) :: e
    call e% create_fixed(loc, name, value)
    call e% close()
  end subroutine dset_d1
  subroutine dset_i1(loc, name, value)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    integer, intent(in) :: value(:)
    type(h5md_element_t) :: e
    call e% create_fixed(loc, name, value)
    call e% close()
  end subroutine dset_i1

  ! ---- attributes -------------------------------------------------------------------
  subroutine attr_head(loc, name, kind, n, u)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name, kind
    integer, intent(in) :: n
    integer, intent(out) :: u
    integer :: f
    u = unit_of(loc, f)
    write(u, '(a,1x,a,1x,a,1x,a,1x,i0)') 'attribute', trim(h5t_path(loc)), trim(name), kind, n
  end subroutine attr_head
  subroutine attr_i0(loc, name, value)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    integer, intent(in) :: value
    integer :: u
    call attr_head(loc, name, 'int', 1, u)
    call put_i(u, [value])
  end subroutine attr_i0
  subroutine attr_i1(loc, name, value)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    integer, intent(in) :: value(:)
    integer :: u
    call attr_head(loc, name, 'int', size(value), u)
    call put_i(u, value)
  end subroutine attr_i1
  subroutine attr_d0(loc, name, value)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    double precision, intent(in) :: value
    integer :: u
    call attr_head(loc, name, 'double', 1, u)
    call put_d(u, [value])
  end subroutine attr_d0
  subroutine attr_d1(loc, name, value)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    double precision, intent(in) :: value(:)
    integer :: u
    call attr_head(loc, name, 'double', size(value), u)
    call put_d(u, value)
  end subroutine attr_d1
  subroutine attr_s0(loc, name, value)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    character(len=*), intent(in) :: value
    integer :: u
    call attr_head(loc, name, 'string', 1, u)
    write(u, '(a)') trim(value)
  end subroutine attr_s0
  subroutine attr_s1(loc, name, value)
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    character(len=*), intent(in) :: value(:)
    integer :: u, i
    call attr_head(loc, name, 'string', size(value), u)
    do i = 1, size(value)
       write(u, '(a)') trim(value(i))
    end do
  end subroutine attr_s1

  ! ---- reading (particle_system_t%init_from_file) -----------------------------------
  ! the record line "<tag> <path> ..." of the n-th (0-based) such record; the unit is left
  ! after it.  For frames the tag is "frame" and the element's header gives the dims.
  logical function seek(u, tag, path, nth, line) result(found)
    integer, intent(in) :: u, nth
    character(len=*), intent(in) :: tag, path
    character(len=*), intent(out) :: line
    integer :: e, k, seen
    character(len=4096) :: buf
    found = .false.
    seen = 0
    rewind(u)
    do
       read(u, '(a)', iostat=e) buf
       if (e /= 0) exit
       k = len_trim(tag)
       if (buf(1:k + 1) /= tag(1:k)//' ') cycle
       if (index(buf(k + 2:), trim(path)//' ') /= 1) cycle
       if (seen == nth) then
          line = buf
          found = .true.
          return
       end if
       seen = seen + 1
    end do
  end function seek

  ! word k (1-based, blank separated) of line; list-directed input would stop at the '/'
  ! that starts a path
  function word(line, k) result(w)
    character(len=*), intent(in) :: line
    integer, intent(in) :: k
    character(len=512) :: w
    integer :: i, n, b
    w = ''
    n = 0
    i = 1
    do while (i <= len_trim(line))
       if (line(i:i) == ' ') then
          i = i + 1
          cycle
       end if
       b = i
       do while (i <= len_trim(line))
          if (line(i:i) == ' ') exit
          i = i + 1
       end do
       n = n + 1
       if (n == k) then
          w = line(b:i - 1)
          return
       end if
    end do
  end function word

  ! the dims of a "fixed" or "time" record: "<tag> <path> <kind> <rank> <dims...>"
  subroutine dims_of(line, skip, dims)
    character(len=*), intent(in) :: line
    integer, intent(in) :: skip  ! words before the rank
    integer, allocatable, intent(out) :: dims(:)
    integer :: rank, k
    read(word(line, skip + 1), *) rank
    allocate(dims(rank))
    do k = 1, rank
       read(word(line, skip + 1 + k), *) dims(k)
    end do
  end subroutine dims_of

  subroutine read_fixed_d2(this, loc, name, data)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    double precision, allocatable, intent(out) :: data(:, :)
    character(len=4096) :: line
    integer, allocatable :: dims(:)
    integer :: u, f, i, j
    u = unit_of(loc, f)
    this% f = f
    if (.not. seek(u, 'fixed', trim(h5t_path(loc))//'/'//trim(name), 0, line)) &
         error stop 'h5md: fixed element not found'
    call dims_of(line, 3, dims)
    if (size(dims) /= 2) error stop 'h5md: expected a rank-2 element'
    allocate(data(dims(1), dims(2)))
    do j = 1, dims(2)
       do i = 1, dims(1)
          read(u, *) data(i, j)
       end do
    end do
  end subroutine read_fixed_d2

  subroutine read_fixed_i1(this, loc, name, data)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    integer, allocatable, intent(out) :: data(:)
    character(len=4096) :: line
    integer, allocatable :: dims(:)
    integer :: u, f, i
    u = unit_of(loc, f)
    this% f = f
    if (.not. seek(u, 'fixed', trim(h5t_path(loc))//'/'//trim(name), 0, line)) &
         error stop 'h5md: fixed element not found'
    call dims_of(line, 3, dims)
    if (size(dims) /= 1) error stop 'h5md: expected a rank-1 element'
    allocate(data(dims(1)))
    do i = 1, dims(1)
       read(u, *) data(i)
    end do
  end subroutine read_fixed_i1

  ! a time-dependent element opened again: in a file being written, to append to it (its
  ! mode, increments and frame count as created); in a file read, Nmax from its header
  subroutine open_time(this, loc, name)
    class(h5md_element_t), intent(inout) :: this
    integer(HID_T), intent(in) :: loc
    character(len=*), intent(in) :: name
    character(len=4096) :: line
    integer, allocatable :: dims(:)
    integer :: u, f, i
    u = unit_of(loc, f)
    this% f = f
    this% path = trim(h5t_path(loc))//'/'//trim(name)
    do i = n_recs, 1, -1
       if (recs(i)% f == f .and. recs(i)% path == this% path) then
          this% reg = i
          this% mode = recs(i)% mode
          this% step = recs(i)% step
          this% time = recs(i)% time
          this% step_offset = recs(i)% step_offset
          this% time_offset = recs(i)% time_offset
          this% n_frames = recs(i)% n_frames
          this% Nmax = recs(i)% Nmax
          this% id = h5t_new_object(f, trim(this% path))
          this% v = this% id
          return
       end if
    end do
    if (.not. seek(u, 'time', trim(this% path), 0, line)) error stop 'h5md: time element not found'
    call dims_of(line, 3, dims)
    this% Nmax = 0
    if (size(dims) > 0) this% Nmax = dims(size(dims))
    this% id = h5t_new_object(f, trim(this% path))
    this% v = this% id
  end subroutine open_time

  ! frame idx (0-based) of an element opened with open_time, into data (its header's shape)
  subroutine read_frame_d2(this, idx, data)
    class(h5md_element_t), intent(inout) :: this
    integer, intent(in) :: idx
    double precision, intent(inout) :: data(:, :)
    character(len=4096) :: line
    integer :: u, i, j
    u = h5t_unit(this% f)
    if (.not. seek(u, 'frame', trim(this% path), idx, line)) error stop 'h5md: frame not found'
    do j = 1, size(data, 2)
       do i = 1, size(data, 1)
          read(u, *) data(i, j)
       end do
    end do
  end subroutine read_frame_d2

end module h5md_module

! hdf5_util_write_dataset: the drivers call it without an interface (programs/
! n_colloids_pbc.f90:99 writes the 64-bit seed); a scalar dataset of the given value
subroutine hdf5_util_write_dataset(loc, name, value)
  use iso_c_binding, only: c_int64_t
  use hdf5, only: HID_T, h5t_unit, h5t_file, h5t_path
  implicit none
  integer(HID_T), intent(in) :: loc
  character(len=*), intent(in) :: name
  integer(c_int64_t), intent(in) :: value
  integer :: u
  u = h5t_unit(h5t_file(loc))
  if (u == 0) error stop 'hdf5_util_write_dataset: invalid location'
  write(u, '(a,1x,a,1x,a)') 'fixed', trim(h5t_path(loc))//'/'//trim(name), 'int 0'
  write(u, '(i0)') value
end subroutine hdf5_util_write_dataset
