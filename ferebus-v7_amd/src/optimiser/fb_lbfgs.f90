!> L-BFGS polishing of one position (the swarm's global best) on the cost
!> -logL, driven by the adjoint gradient of the gfx950 library
!> (fb_swarm_loglik_grad).  The reference PSO has no gradient step (the
!> gradient flags of gpr.f90:272-282 are parsed but unused); this is an
!> optional addition, off by default: plain PSO stays the default and the
!> baseline its results are compared with.
!>
!> Limited-memory BFGS with the two-loop recursion (Nocedal 1980) and
!> H0 = (s'y / y'y) I from the newest pair (1 / |g| on the first step), and a
!> backtracking Armijo line search (sufficient decrease c1 = 1e-4, halving).
!> A trial point whose matrix is not positive definite (INFO > 0) counts as
!> a failed step and is halved like any other; a pair with s'y <= 0 is not
!> stored.  The full step t = 1 is tried alone; after a rejection the trial
!> steps 1/2, 1/4, ... are evaluated `batch` at a time as one swarm gradient
!> call (a few particles cost about as much as one on the GPU); the
!> accepted step is the first of the sequence that satisfies the Armijo
!> condition, as in a one-at-a-time search.
module fb_lbfgs
  use iso_c_binding, only: c_int, c_double
  use ferebus_hip, only: FB_OK, fb_error_message
  use fb_gpr, only: GaussianProcessRegressor
  implicit none
  private
  integer, parameter :: wp = c_double

  type, public :: LbfgsSettings
    integer :: steps = 5            ! accepted steps per polish
    integer :: memory = 5           ! stored (s, y) pairs
    real(wp) :: c1 = 1e-4_wp        ! Armijo constant
    integer :: max_backtracks = 30
    integer :: batch = 4            ! trial steps per gradient call
  end type LbfgsSettings

  public :: lbfgs_polish

contains

  !> Cost and gradient at one position; ok = positive definite.
  subroutine cost_grad(gp, x, f, g, ok, rc)
    type(GaussianProcessRegressor), intent(inout) :: gp
    real(wp), intent(in) :: x(:)
    real(wp), intent(out) :: f, g(:)
    logical, intent(out) :: ok
    integer, intent(out) :: rc
    real(wp) :: th(size(x), 1), c(1), gg(size(x), 1)
    integer(c_int) :: info(1)
    th(:, 1) = x
    call gp%swarm_grad(th, c, gg, info, rc)
    ok = rc == FB_OK .and. info(1) == 0
    f = c(1)
    g = gg(:, 1)
  end subroutine cost_grad

  !> Costs and gradients at the columns of xs; ok = positive definite.
  subroutine cost_grad_batch(gp, xs, f, g, ok, rc)
    type(GaussianProcessRegressor), intent(inout) :: gp
    real(wp), intent(in) :: xs(:, :)
    real(wp), intent(out) :: f(:), g(:, :)
    logical, intent(out) :: ok(:)
    integer, intent(out) :: rc
    integer(c_int) :: info(size(xs, 2))
    call gp%swarm_grad(xs, f, g, info, rc)
    ok = rc == FB_OK .and. info == 0
  end subroutine cost_grad_batch

  !> Polishes x (cost f on entry) in place.  evals: gradient evaluations
  !> used; accepted: steps taken.  rc /= 0 only on a library error.
  subroutine lbfgs_polish(gp, cfg, x, f, evals, accepted, rc)
    type(GaussianProcessRegressor), intent(inout) :: gp
    type(LbfgsSettings), intent(in) :: cfg
    real(wp), intent(inout) :: x(:), f
    integer, intent(out) :: evals, accepted, rc
    real(wp), allocatable :: s(:, :), y(:, :), rho(:), g(:), d(:), q(:), xt(:), gt(:), a(:)
    real(wp), allocatable :: xs(:, :), fs(:), gs(:, :), ts(:)
    logical, allocatable :: oks(:)
    real(wp) :: ft, t, gd, gamma, b, sy
    integer :: n, m, k, j, i, k0, nb, nt, newest
    logical :: ok
    n = size(x)
    m = max(cfg%memory, 1)
    allocate(s(n, m), y(n, m), rho(m), g(n), d(n), q(n), xt(n), gt(n), a(m))
    evals = 0
    accepted = 0
    k = 0                                        ! pairs stored
    newest = 0
    call cost_grad(gp, x, ft, g, ok, rc)
    evals = 1
    if (rc /= FB_OK .or. .not. ok) return
    f = ft
    do while (accepted < cfg%steps)
      ! two-loop recursion: d = -H g
      q = g
      do i = 0, k - 1
        j = modulo(newest - 1 - i, m) + 1
        a(j) = rho(j) * dot_product(s(:, j), q)
        q = q - a(j) * y(:, j)
      end do
      if (k > 0) then
        gamma = dot_product(s(:, newest), y(:, newest)) / dot_product(y(:, newest), y(:, newest))
      else
        gamma = 1.0_wp / max(norm2(g), tiny(1.0_wp))
      end if
      q = gamma * q
      do i = k - 1, 0, -1
        j = modulo(newest - 1 - i, m) + 1
        b = rho(j) * dot_product(y(:, j), q)
        q = q + (a(j) - b) * s(:, j)
      end do
      d = -q
      gd = dot_product(g, d)
      if (.not. (gd < 0.0_wp)) then              ! not a descent direction: steepest descent
        d = -gamma * g
        gd = dot_product(g, d)
        if (.not. (gd < 0.0_wp)) exit
      end if
      ! trial steps 0.5**k, k = k0 .. k0 + nt - 1, in one gradient call
      ok = .false.
      nb = max(1, min(cfg%batch, gp%max_swarm))
      k0 = 0
      do while (.not. ok .and. k0 <= cfg%max_backtracks)
        nt = 1                                   ! the full step alone, then nb at a time
        if (k0 > 0) nt = min(nb, cfg%max_backtracks + 1 - k0)
        allocate(xs(n, nt), fs(nt), gs(n, nt), oks(nt), ts(nt))
        do j = 1, nt
          ts(j) = 0.5_wp ** (k0 + j - 1)
          xs(:, j) = x + ts(j) * d
        end do
        call cost_grad_batch(gp, xs, fs, gs, oks, rc)
        evals = evals + nt
        if (rc /= FB_OK) return
        do j = 1, nt
          if (oks(j)) then
            if (fs(j) <= f + cfg%c1 * ts(j) * gd) then
              ok = .true.
              t = ts(j)
              xt = xs(:, j)
              ft = fs(j)
              gt = gs(:, j)
              exit
            end if
          end if
        end do
        deallocate(xs, fs, gs, oks, ts)
        k0 = k0 + nt
      end do
      if (.not. ok) exit                         ! no acceptable step: stop polishing
      sy = dot_product(xt - x, gt - g)
      if (sy > 0.0_wp) then
        newest = modulo(newest, m) + 1
        s(:, newest) = xt - x
        y(:, newest) = gt - g
        rho(newest) = 1.0_wp / sy
        k = min(k + 1, m)
      end if
      x = xt
      f = ft
      g = gt
      accepted = accepted + 1
    end do
  end subroutine lbfgs_polish
end module fb_lbfgs
