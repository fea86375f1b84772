! tester -- the interface of the test helper the reference's Fortran tests use
! (test/test_*.f90: `use tester`, `type(tester_t)`, init/assert_*/print).  The helper is
! an absent third-party submodule of the reference (fortran_tester, CMakeLists.txt:13);
! this module restates its interface with the semantics written next to each routine,
! so that the reference's tests compile and link unmodified against this library.
module tester
  use iso_fortran_env, only: real32, real64, int32, int64
  implicit none
  private

  public :: tester_t

  type tester_t
     integer :: n_errors = 0
     integer :: n_tests = 0
     real(real64) :: tolerance64 = 2 * epsilon(1._real64)
     real(real32) :: tolerance32 = 2 * epsilon(1._real32)
   contains
     procedure :: init
     procedure :: print => print_summary
     generic, public :: assert_equal => eq_i32, eq_i64, eq_l, eq_r64, eq_i32_1d, eq_i64_1d, eq_l_1d, eq_r64_1d
     procedure, private :: eq_i32, eq_i64, eq_l, eq_r64, eq_i32_1d, eq_i64_1d, eq_l_1d, eq_r64_1d
     generic, public :: assert_close => close_r64, close_r32, close_r64_1d, close_r32_1d, close_r64_2d
     procedure, private :: close_r64, close_r32, close_r64_1d, close_r32_1d, close_r64_2d
     generic, public :: assert_positive => pos_i32, pos_r64, pos_i32_1d, pos_r64_1d
     procedure, private :: pos_i32, pos_r64, pos_i32_1d, pos_r64_1d
  end type tester_t

contains

  ! reset the counters; tolerance64/tolerance32 are the relative tolerances of assert_close
  subroutine init(this, tolerance64, tolerance32)
    class(tester_t), intent(out) :: this
    real(real64), intent(in), optional :: tolerance64
    real(real32), intent(in), optional :: tolerance32
    this% n_errors = 0
    this% n_tests = 0
    this% tolerance64 = 2 * epsilon(1._real64)
    this% tolerance32 = 2 * epsilon(1._real32)
    if (present(tolerance64)) this% tolerance64 = tolerance64
    if (present(tolerance32)) this% tolerance32 = tolerance32
  end subroutine init

  ! one line with the counts; an error stop when a test failed (unless errorstop is false)
  subroutine print_summary(this, errorstop)
    class(tester_t), intent(in) :: this
    logical, intent(in), optional :: errorstop
    logical :: stop_on_error
    stop_on_error = .true.
    if (present(errorstop)) stop_on_error = errorstop
    write(*, '(a,i0,a,i0,a)') 'tester: ', this% n_errors, ' error(s) for ', this% n_tests, ' test(s)'
    if (this% n_errors == 0) then
       write(*, '(a)') 'tester: all tests succeeded'
    else if (stop_on_error) then
       error stop 'tester: some tests failed'
    end if
  end subroutine print_summary

  subroutine record(this, ok)
    class(tester_t), intent(inout) :: this
    logical, intent(in) :: ok
    this% n_tests = this% n_tests + 1
    if (.not. ok) this% n_errors = this% n_errors + 1
  end subroutine record

  ! assert_equal: exact equality (arrays: same size and every element equal)
  subroutine eq_i32(this, a, b)
    class(tester_t), intent(inout) :: this
    integer(int32), intent(in) :: a, b
    call record(this, a == b)
  end subroutine eq_i32
  subroutine eq_i64(this, a, b)
    class(tester_t), intent(inout) :: this
    integer(int64), intent(in) :: a, b
    call record(this, a == b)
  end subroutine eq_i64
  subroutine eq_l(this, a, b)
    class(tester_t), intent(inout) :: this
    logical, intent(in) :: a, b
    call record(this, a .eqv. b)
  end subroutine eq_l
  subroutine eq_r64(this, a, b)
    class(tester_t), intent(inout) :: this
    real(real64), intent(in) :: a, b
    call record(this, a == b)
  end subroutine eq_r64
  subroutine eq_i32_1d(this, a, b)
    class(tester_t), intent(inout) :: this
    integer(int32), intent(in) :: a(:), b(:)
    if (size(a) /= size(b)) then
       call record(this, .false.)
    else
       call record(this, all(a == b))
    end if
  end subroutine eq_i32_1d
  subroutine eq_i64_1d(this, a, b)
    class(tester_t), intent(inout) :: this
    integer(int64), intent(in) :: a(:), b(:)
    if (size(a) /= size(b)) then
       call record(this, .false.)
    else
       call record(this, all(a == b))
    end if
  end subroutine eq_i64_1d
  subroutine eq_l_1d(this, a, b)
    class(tester_t), intent(inout) :: this
    logical, intent(in) :: a(:), b(:)
    if (size(a) /= size(b)) then
       call record(this, .false.)
    else
       call record(this, all(a .eqv. b))
    end if
  end subroutine eq_l_1d
  subroutine eq_r64_1d(this, a, b)
    class(tester_t), intent(inout) :: this
    real(real64), intent(in) :: a(:), b(:)
    if (size(a) /= size(b)) then
       call record(this, .false.)
    else
       call record(this, all(a == b))
    end if
  end subroutine eq_r64_1d

  ! assert_close: |a - b| <= tolerance * max(|a|, |b|), element by element; equal values
  ! (zeros included) always pass
  logical elemental function near64(a, b, tol)
    real(real64), intent(in) :: a, b, tol
    near64 = (a == b) .or. (abs(a - b) <= tol * max(abs(a), abs(b)))
  end function near64
  logical elemental function near32(a, b, tol)
    real(real32), intent(in) :: a, b, tol
    near32 = (a == b) .or. (abs(a - b) <= tol * max(abs(a), abs(b)))
  end function near32
  subroutine close_r64(this, a, b)
    class(tester_t), intent(inout) :: this
    real(real64), intent(in) :: a, b
    call record(this, near64(a, b, this% tolerance64))
  end subroutine close_r64
  subroutine close_r32(this, a, b)
    class(tester_t), intent(inout) :: this
    real(real32), intent(in) :: a, b
    call record(this, near32(a, b, this% tolerance32))
  end subroutine close_r32
  subroutine close_r64_1d(this, a, b)
    class(tester_t), intent(inout) :: this
    real(real64), intent(in) :: a(:), b(:)
    if (size(a) /= size(b)) then
       call record(this, .false.)
    else
       call record(this, all(near64(a, b, this% tolerance64)))
    end if
  end subroutine close_r64_1d
  subroutine close_r32_1d(this, a, b)
    class(tester_t), intent(inout) :: this
    real(real32), intent(in) :: a(:), b(:)
    if (size(a) /= size(b)) then
       call record(this, .false.)
    else
       call record(this, all(near32(a, b, this% tolerance32)))
    end if
  end subroutine close_r32_1d
  subroutine close_r64_2d(this, a, b)
    class(tester_t), intent(inout) :: this
    real(real64), intent(in) :: a(:, :), b(:, :)
    if (any(shape(a) /= shape(b))) then
       call record(this, .false.)
    else
       call record(this, all(near64(a, b, this% tolerance64)))
    end if
  end subroutine close_r64_2d

  ! assert_positive: every value >= 0
  subroutine pos_i32(this, a)
    class(tester_t), intent(inout) :: this
    integer(int32), intent(in) :: a
    call record(this, a >= 0)
  end subroutine pos_i32
  subroutine pos_r64(this, a)
    class(tester_t), intent(inout) :: this
    real(real64), intent(in) :: a
    call record(this, a >= 0)
  end subroutine pos_r64
  subroutine pos_i32_1d(this, a)
    class(tester_t), intent(inout) :: this
    integer(int32), intent(in) :: a(:)
    call record(this, all(a >= 0))
  end subroutine pos_i32_1d
  subroutine pos_r64_1d(this, a)
    class(tester_t), intent(inout) :: this
    real(real64), intent(in) :: a(:)
    call record(this, all(a >= 0))
  end subroutine pos_r64_1d

end module tester
