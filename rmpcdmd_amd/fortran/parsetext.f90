!> Parameter files as the reference's drivers read them (the ParseText dependency of
!! programs/*.f90, an absent submodule): one `name = value` per line, values separated
!! by blanks, `#` starts a comment.  PTparse reads a file, PTread_d / _i / _l / _s /
!! _dvec / _ivec / _lvec return the value of a name and stop with a message when the name
!! is absent or its value does not read, PTkill forgets the file.  The optional `loc`
!! argument, where the drivers write each parameter to their HDF5 output, is accepted and
!! ignored: HDF5 is not part of this library.
module ParseText
  use iso_c_binding, only: c_int64_t
  implicit none
  private
  public :: PTo, PTparse, PTkill
  public :: PTread_d, PTread_i, PTread_l, PTread_s, PTread_dvec, PTread_ivec, PTread_lvec, PTread_c_int64

  integer, parameter :: LINE_LEN = 1024

  type PTo
     character(len=LINE_LEN), allocatable :: names(:), values(:)
     integer :: n = 0
  end type PTo

contains

  !> reads filename on unit u (the drivers pass 11)
  subroutine PTparse(this, filename, u)
    type(PTo), intent(out) :: this
    character(len=*), intent(in) :: filename
    integer, intent(in) :: u
    character(len=LINE_LEN) :: line
    integer :: iostat, n_lines, k, eq
    open(unit=u, file=trim(filename), status='old', action='read', iostat=iostat)
    if (iostat /= 0) then
       write(*, '(a,a)') 'PTparse: cannot open ', trim(filename)
       error stop
    end if
    n_lines = 0
    do
       read(u, '(a)', iostat=iostat) line
       if (iostat /= 0) exit
       n_lines = n_lines + 1
    end do
    allocate(this%names(n_lines), this%values(n_lines))
    rewind(u)
    this%n = 0
    do k = 1, n_lines
       read(u, '(a)', iostat=iostat) line
       if (iostat /= 0) exit
       eq = index(line, '#')
       if (eq > 0) line(eq:) = ' '
       eq = index(line, '=')
       if (eq == 0) cycle
       if (len_trim(line(:eq-1)) == 0) cycle
       this%n = this%n + 1
       this%names(this%n) = adjustl(line(:eq-1))
       this%values(this%n) = adjustl(line(eq+1:))
    end do
    close(u)
  end subroutine PTparse

  subroutine PTkill(this)
    type(PTo), intent(inout) :: this
    if (allocated(this%names)) deallocate(this%names, this%values)
    this%n = 0
  end subroutine PTkill

  !> the value text of name (the last occurrence wins)
  function value_of(this, name) result(v)
    type(PTo), intent(in) :: this
    character(len=*), intent(in) :: name
    character(len=LINE_LEN) :: v
    integer :: k
    do k = this%n, 1, -1
       if (trim(this%names(k)) == trim(name)) then
          v = this%values(k)
          return
       end if
    end do
    ! the message goes with the stop: value_of may run inside the caller's write
    error stop 'ParseText: parameter '//trim(name)//' not found'
  end function value_of

  subroutine bad_value(name)
    character(len=*), intent(in) :: name
    error stop 'ParseText: cannot read the value of '//trim(name)
  end subroutine bad_value

  double precision function PTread_d(this, name, loc) result(x)
    type(PTo), intent(in) :: this
    character(len=*), intent(in) :: name
    integer(c_int64_t), intent(in), optional :: loc
    integer :: iostat
    read(value_of(this, name), *, iostat=iostat) x
    if (iostat /= 0) call bad_value(name)
  end function PTread_d

  integer function PTread_i(this, name, loc) result(x)
    type(PTo), intent(in) :: this
    character(len=*), intent(in) :: name
    integer(c_int64_t), intent(in), optional :: loc
    integer :: iostat
    read(value_of(this, name), *, iostat=iostat) x
    if (iostat /= 0) call bad_value(name)
  end function PTread_i

  integer(c_int64_t) function PTread_c_int64(this, name, loc) result(x)
    type(PTo), intent(in) :: this
    character(len=*), intent(in) :: name
    integer(c_int64_t), intent(in), optional :: loc
    integer :: iostat
    read(value_of(this, name), *, iostat=iostat) x
    if (iostat /= 0) call bad_value(name)
  end function PTread_c_int64

  logical function PTread_l(this, name, loc) result(x)
    type(PTo), intent(in) :: this
    character(len=*), intent(in) :: name
    integer(c_int64_t), intent(in), optional :: loc
    integer :: iostat
    read(value_of(this, name), *, iostat=iostat) x
    if (iostat /= 0) call bad_value(name)
  end function PTread_l

  function PTread_s(this, name, loc) result(x)
    type(PTo), intent(in) :: this
    character(len=*), intent(in) :: name
    integer(c_int64_t), intent(in), optional :: loc
    character(len=:), allocatable :: x
    x = trim(value_of(this, name))
  end function PTread_s

  function PTread_dvec(this, name, n, loc) result(x)
    type(PTo), intent(in) :: this
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    integer(c_int64_t), intent(in), optional :: loc
    double precision :: x(n)
    integer :: iostat
    read(value_of(this, name), *, iostat=iostat) x
    if (iostat /= 0) call bad_value(name)
  end function PTread_dvec

  function PTread_ivec(this, name, n, loc) result(x)
    type(PTo), intent(in) :: this
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    integer(c_int64_t), intent(in), optional :: loc
    integer :: x(n)
    integer :: iostat
    read(value_of(this, name), *, iostat=iostat) x
    if (iostat /= 0) call bad_value(name)
  end function PTread_ivec

  function PTread_lvec(this, name, n, loc) result(x)
    type(PTo), intent(in) :: this
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    integer(c_int64_t), intent(in), optional :: loc
    logical :: x(n)
    integer :: iostat
    read(value_of(this, name), *, iostat=iostat) x
    if (iostat /= 0) call bad_value(name)
  end function PTread_lvec

end module ParseText
