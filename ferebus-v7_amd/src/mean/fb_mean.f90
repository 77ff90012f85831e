!> Prior mean functions of the GP, evaluated once per training set.
!>
!> zero       src/mean/zero.f90:47-52
!> constant   src/mean/constant.f90:42-55 (sum(y)/n, src/utils/arith.f90:44-48)
!> linear     src/mean/least_squares.f90:54-93, src/mean/linear.f90:41-51
!> quadratic  src/mean/least_squares.f90:54-93, src/mean/quadratic.f90:41-52
!> The least-squares means fit y - ymin on x - xmin (xmin = the point of the
!> smallest y, first occurrence) without intercept, by Cholesky of the normal
!> equations; the quadratic mean uses the squared terms only.
module fb_mean
  use iso_c_binding, only: c_double
  implicit none
  private
  integer, parameter :: wp = c_double
  integer, parameter, public :: MEAN_ZERO = 0, MEAN_CONSTANT = 1, MEAN_LINEAR = 2, MEAN_QUADRATIC = 3

  type, public :: MeanFit
    integer :: kind = MEAN_CONSTANT
    real(wp) :: constant = 0.0_wp
    real(wp), allocatable :: beta(:), xmin(:)
    real(wp) :: ymin = 0.0_wp
    integer :: info = 0
  contains
    procedure :: fit => mean_fit
    procedure :: value => mean_value
    procedure :: write => mean_write
  end type MeanFit

  public :: mean_kind_from_name

contains

  integer function mean_kind_from_name(name) result(k)
    character(len=*), intent(in) :: name
    select case (trim(adjustl(name)))
    case ("zero"); k = MEAN_ZERO
    case ("constant"); k = MEAN_CONSTANT
    case ("linear"); k = MEAN_LINEAR
    case ("quadratic"); k = MEAN_QUADRATIC
    case default; k = -1
    end select
  end function mean_kind_from_name

  subroutine mean_fit(self, kind, x, y)
    class(MeanFit), intent(inout) :: self
    integer, intent(in) :: kind
    real(wp), intent(in) :: x(:, :), y(:)
    real(wp), allocatable :: a(:, :), aa(:, :), ab(:)
    integer :: n, d, imin, i, j
    self%kind = kind
    n = size(y)
    d = size(x, 2)
    select case (kind)
    case (MEAN_ZERO)
      self%constant = 0.0_wp
    case (MEAN_CONSTANT)
      self%constant = sum(y) / real(n, wp)
    case (MEAN_LINEAR, MEAN_QUADRATIC)
      imin = minloc(y, dim=1)
      self%xmin = x(imin, :)
      self%ymin = y(imin)
      allocate(a(n, d))
      do j = 1, d
        a(:, j) = x(:, j) - self%xmin(j)
      end do
      if (kind == MEAN_QUADRATIC) a = a * a
      aa = matmul(transpose(a), a)
      ab = matmul(transpose(a), y - self%ymin)
      call chol_solve(aa, ab, self%info)
      self%beta = ab
    end select
  end subroutine mean_fit

  !> m(x_i) for the training inputs.
  subroutine mean_value(self, x, m)
    class(MeanFit), intent(in) :: self
    real(wp), intent(in) :: x(:, :)
    real(wp), intent(out) :: m(:)
    real(wp), allocatable :: a(:, :)
    integer :: j
    select case (self%kind)
    case (MEAN_ZERO, MEAN_CONSTANT)
      m = self%constant
    case default
      allocate(a, mold=x)
      do j = 1, size(x, 2)
        a(:, j) = x(:, j) - self%xmin(j)
      end do
      if (self%kind == MEAN_QUADRATIC) a = a * a
      m = matmul(a, self%beta) + self%ymin
    end select
  end subroutine mean_value

  !> [mean] block of the .model file (src/mean/*.f90 write routines).
  subroutine mean_write(self, unit)
    class(MeanFit), intent(in) :: self
    integer, intent(in) :: unit
    character(len=64) :: fmt
    write(unit, '(A)') "[mean]"
    select case (self%kind)
    case (MEAN_ZERO)
      write(unit, '(A, 1x, A)') "type", "zero"
    case (MEAN_CONSTANT)
      write(unit, '(A, 1x, A)') "type", "constant"
      write(unit, '(A, 1x, G0)') "value", self%constant
    case default
      write(fmt, '(a,i0,a)') "(a,1x,(", size(self%beta), "(g0,1x)))"
      if (self%kind == MEAN_LINEAR) then
        write(unit, '(a, 1x, a)') "type", "linear"
      else
        write(unit, '(a, 1x, a)') "type", "quadratic"
      end if
      write(unit, trim(fmt)) "beta", self%beta
      write(unit, trim(fmt)) "xmin", self%xmin
      write(unit, '(a, 1x, g0)') "ymin", self%ymin
    end select
  end subroutine mean_write

  !> dposv('L') of a small SPD system: b <- a^{-1} b, info as LAPACK.
  subroutine chol_solve(a, b, info)
    real(wp), intent(inout) :: a(:, :), b(:)
    integer, intent(out) :: info
    integer :: n, j, i
    real(wp) :: ajj
    n = size(b)
    info = 0
    do j = 1, n
      ajj = a(j, j) - dot_product(a(j, 1:j-1), a(j, 1:j-1))
      if (.not. (ajj > 0.0_wp)) then
        info = j
        return
      end if
      ajj = sqrt(ajj)
      a(j, j) = ajj
      do i = j + 1, n
        a(i, j) = (a(i, j) - dot_product(a(i, 1:j-1), a(j, 1:j-1))) / ajj
      end do
    end do
    do j = 1, n
      b(j) = (b(j) - dot_product(a(j, 1:j-1), b(1:j-1))) / a(j, j)
    end do
    do j = n, 1, -1
      b(j) = (b(j) - dot_product(a(j+1:n, j), b(j+1:n))) / a(j, j)
    end do
  end subroutine chol_solve
end module fb_mean
