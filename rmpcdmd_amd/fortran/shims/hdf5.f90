! hdf5 -- the part of the HDF5 Fortran interface that the reference's drivers and tests
! call directly (h5open_f, h5close_f, h5gcreate_f, h5gopen_f, h5gclose_f and the HID_T
! kind; e.g. test/setup_simple_fluid.f90:43, 69-74, 141).  HDF5 is not a dependency of
! this library: groups here are paths inside the plain-text files written by
! h5md_module, and identifiers index a table of (file, path) pairs.
module hdf5
  use iso_c_binding, only: c_int64_t
  implicit none
  private

  integer, parameter, public :: HID_T = c_int64_t, HSIZE_T = c_int64_t
  integer, parameter, public :: H5F_ACC_RDONLY_F = 0, H5F_ACC_RDWR_F = 1, H5F_ACC_TRUNC_F = 2

  public :: h5open_f, h5close_f, h5gcreate_f, h5gopen_f, h5gclose_f
  ! for h5md_module: the table of files and objects
  public :: h5t_new_file, h5t_close_file, h5t_new_object, h5t_path, h5t_file, h5t_unit, h5t_file_name
  public :: h5t_file_valid

  integer, parameter :: PATH_LEN = 512
  type object_t
     character(len=PATH_LEN) :: path = ''
     integer :: file = 0
     logical :: used = .false.
  end type object_t
  type file_t
     character(len=PATH_LEN) :: name = ''
     integer :: unit = 0
     logical :: used = .false.
     logical :: writable = .false.
  end type file_t
  type(object_t), allocatable, save :: objects(:)
  type(file_t), allocatable, save :: files(:)

contains

  subroutine h5open_f(error)
    integer, intent(out) :: error
    if (.not. allocated(objects)) allocate(objects(64))
    if (.not. allocated(files)) allocate(files(8))
    error = 0
  end subroutine h5open_f

  subroutine h5close_f(error)
    integer, intent(out) :: error
    error = 0
  end subroutine h5close_f

  ! a group: the path loc/name in loc's file ("group" line in a writable file)
  subroutine h5gcreate_f(loc_id, name, grp_id, error)
    integer(HID_T), intent(in) :: loc_id
    character(len=*), intent(in) :: name
    integer(HID_T), intent(out) :: grp_id
    integer, intent(out) :: error
    integer :: f
    f = h5t_file(loc_id)
    if (f <= 0) then
       grp_id = -1
       error = -1
       return
    end if
    grp_id = h5t_new_object(f, trim(h5t_path(loc_id))//'/'//trim(name))
    if (files(f)% writable) write(files(f)% unit, '(a,1x,a)') 'group', trim(h5t_path(grp_id))
    error = 0
  end subroutine h5gcreate_f

  subroutine h5gopen_f(loc_id, name, grp_id, error)
    integer(HID_T), intent(in) :: loc_id
    character(len=*), intent(in) :: name
    integer(HID_T), intent(out) :: grp_id
    integer, intent(out) :: error
    integer :: f
    f = h5t_file(loc_id)
    if (f <= 0) then
       grp_id = -1
       error = -1
       return
    end if
    grp_id = h5t_new_object(f, trim(h5t_path(loc_id))//'/'//trim(name))
    error = 0
  end subroutine h5gopen_f

  subroutine h5gclose_f(grp_id, error)
    integer(HID_T), intent(in) :: grp_id
    integer, intent(out) :: error
    error = -1
    if (grp_id < 1 .or. .not. allocated(objects)) return
    if (grp_id > size(objects)) return
    objects(grp_id)% used = .false.
    error = 0
  end subroutine h5gclose_f

  ! ---- the table --------------------------------------------------------------------
  integer function h5t_new_file(name, unit, writable) result(f)
    character(len=*), intent(in) :: name
    integer, intent(in) :: unit
    logical, intent(in) :: writable
    type(file_t), allocatable :: grown(:)
    integer :: e
    if (.not. allocated(files)) allocate(files(8))
    do f = 1, size(files)
       if (.not. files(f)% used) exit
    end do
    if (f > size(files)) then
       allocate(grown(2 * size(files)))
       grown(1:size(files)) = files
       call move_alloc(grown, files)
    end if
    files(f)% name = name
    files(f)% unit = unit
    files(f)% used = .true.
    files(f)% writable = writable
    e = 0
  end function h5t_new_file

  subroutine h5t_close_file(f)
    integer, intent(in) :: f
    integer :: i
    if (f < 1 .or. f > size(files)) return
    if (files(f)% used) close(files(f)% unit)
    files(f)% used = .false.
    do i = 1, size(objects)
       if (objects(i)% file == f) objects(i)% used = .false.
    end do
  end subroutine h5t_close_file

  integer(HID_T) function h5t_new_object(f, path) result(id)
    integer, intent(in) :: f
    character(len=*), intent(in) :: path
    type(object_t), allocatable :: grown(:)
    integer :: i
    if (.not. allocated(objects)) allocate(objects(64))
    do i = 1, size(objects)
       if (.not. objects(i)% used) exit
    end do
    if (i > size(objects)) then
       allocate(grown(2 * size(objects)))
       grown(1:size(objects)) = objects
       call move_alloc(grown, objects)
    end if
    objects(i)% path = path
    objects(i)% file = f
    objects(i)% used = .true.
    id = i
  end function h5t_new_object

  function h5t_path(id) result(p)
    integer(HID_T), intent(in) :: id
    character(len=PATH_LEN) :: p
    p = ''
    if (.not. allocated(objects)) return
    if (id >= 1 .and. id <= size(objects)) p = objects(id)% path
  end function h5t_path

  integer function h5t_file(id) result(f)
    integer(HID_T), intent(in) :: id
    f = 0
    if (.not. allocated(objects)) return
    if (id < 1 .or. id > size(objects)) return
    if (objects(id)% used) f = objects(id)% file
  end function h5t_file

  ! the Fortran unit of file f (newunit numbers are negative); 0 when f is not open
  integer function h5t_unit(f) result(u)
    integer, intent(in) :: f
    u = 0
    if (.not. allocated(files)) return
    if (f >= 1 .and. f <= size(files)) then
       if (files(f)% used) u = files(f)% unit
    end if
  end function h5t_unit

  function h5t_file_name(f) result(n)
    integer, intent(in) :: f
    character(len=PATH_LEN) :: n
    n = ''
    if (f >= 1 .and. f <= size(files)) n = files(f)% name
  end function h5t_file_name

  logical function h5t_file_valid(f) result(ok)
    integer, intent(in) :: f
    ok = .false.
    if (.not. allocated(files)) return
    if (f >= 1 .and. f <= size(files)) ok = files(f)% used
  end function h5t_file_valid

end module hdf5
