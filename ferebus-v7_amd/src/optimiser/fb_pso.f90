!> Particle swarm optimiser over whole-swarm log-likelihood evaluations.
!>
!> Semantics of the reference PSO (src/optimiser/pso/pso.f90,
!> src/optimiser/pso/particle.f90, src/optimiser/iter_state.f90):
!>  * initial swarm: positions U[min, max], velocities U[min-max, max-min]
!>    (pso.f90:291-292), evaluated once (pso.f90:294-308); a particle whose
!>    initial matrix is not positive definite keeps cost +Inf (the reference
!>    leaves it undefined, particle.f90:84-88 / likelihood.f90:181-196);
!>  * global update (pso.f90:382-472): r1, r2 ~ U[0,1)^(P x S) drawn before
!>    the swarm loop; v = w v + c1 r1 (pbest - x) + c2 r2 (gbest - x); x += v,
!>    no bounds (pso.f90:433);
!>  * get_cost (particle.f90:149-162): a particle whose matrix is not positive
!>    definite is re-randomised (position and velocity, particle.f90:117-128)
!>    and evaluated again until it succeeds -- here all failed particles of
!>    an iteration are retried together in one batched call;
!>  * pbest with <= (pso.f90:440-443), gbest in particle order with <=
!>    (pso.f90:454-459), best state with <= (iter_state.f90:62-82);
!>  * stopping: relative change of the best cost below `tolerance` for
!>    `stall_iterations` iterations (pso.f90:619-636), or the iteration cap
!>    (optimiser_type.f90:122-132).
!> Every evaluation of a whole swarm is one call of the gfx950 library.
!> Optional (off by default, lbfgs_every = 0): every lbfgs_every iterations
!> the global best is polished by a few L-BFGS steps on the adjoint gradient
!> (fb_lbfgs) and replaces gbest when the polished cost is lower.
module fb_pso
  use iso_c_binding, only: c_int, c_double
  use ieee_arithmetic, only: ieee_value, ieee_positive_inf
  use ferebus_hip, only: FB_OK, fb_error_message
  use fb_gpr, only: GaussianProcessRegressor
  use fb_lbfgs, only: LbfgsSettings, lbfgs_polish
  implicit none
  private
  integer, parameter :: wp = c_double

  type, public :: PsoSettings
    integer :: swarm_size = 50
    integer :: iterations = 1000
    real(wp) :: inertia_weight = 0.789_wp         ! config/optimiser/pso/pso.f90:48-53
    real(wp) :: cognitive_learning_rate = 1.494_wp
    real(wp) :: social_learning_rate = 1.494_wp
    logical :: relative_change = .false.
    real(wp) :: tolerance = 1e-6_wp               ! stopping_criteria.f90:60-62
    integer :: stall_iterations = 20
    integer :: seed = 1
    integer :: max_retries = 10000
    logical :: verbose = .true.
    logical :: trace = .false.                    ! one line per evaluated particle on trace_unit
    integer :: trace_unit = 0
    integer :: lbfgs_every = 0                    ! >0: polish gbest every this many iterations
    type(LbfgsSettings) :: lbfgs
  end type PsoSettings

  type, public :: PsoResult
    real(wp), allocatable :: best_position(:)
    real(wp) :: best_cost
    integer :: iterations = 0
    integer :: evaluations = 0
    integer :: rejections = 0
    integer :: best_iter = 0
    real(wp) :: opt_seconds = 0.0_wp
    integer :: polish_calls = 0
    integer :: polish_evals = 0
    integer :: polish_improvements = 0
    integer :: status = 0
    character(len=:), allocatable :: message
  end type PsoResult

  public :: run_pso

contains

  subroutine uniform(a, lo, hi)
    real(wp), intent(out) :: a(:, :)
    real(wp), intent(in) :: lo(:), hi(:)
    integer :: j
    call random_number(a)
    do j = 1, size(a, 2)
      a(:, j) = lo + a(:, j) * (hi - lo)
    end do
  end subroutine uniform

  subroutine seed_rng(seed)
    integer, intent(in) :: seed
    integer :: n, i
    integer, allocatable :: s(:)
    call random_seed(size=n)
    allocate(s(n))
    s = [(seed * 7919 + 17 * (i - 13), i = 1, n)]
    call random_seed(put=s)
  end subroutine seed_rng

  !> Evaluates the listed particles; failed ones are re-randomised and
  !> re-evaluated together until every particle succeeds (get_cost).
  subroutine evaluate(gp, cfg, x, v, lo, hi, idx, cost, res, it, retry)
    type(GaussianProcessRegressor), intent(inout) :: gp
    type(PsoSettings), intent(in) :: cfg
    real(wp), intent(inout) :: x(:, :), v(:, :)
    real(wp), intent(in) :: lo(:), hi(:)
    integer, intent(in) :: idx(:)
    real(wp), intent(inout) :: cost(:)
    type(PsoResult), intent(inout) :: res
    integer, intent(in) :: it
    logical, intent(in) :: retry
    integer, allocatable :: todo(:)
    real(wp), allocatable :: th(:, :), c(:), rp(:, :), rv(:, :)
    logical, allocatable :: ok(:)
    integer(c_int), allocatable :: info(:)
    integer :: rc, k, round
    todo = idx
    round = 0
    do while (size(todo) > 0)
      th = x(:, todo)
      allocate(c(size(todo)), ok(size(todo)), info(size(todo)))
      call gp%swarm_cost(th, c, ok, info, rc)
      if (rc /= FB_OK) then
        res%status = rc
        res%message = fb_error_message()
        return
      end if
      res%evaluations = res%evaluations + size(todo)
      do k = 1, size(todo)
        if (cfg%trace) write(cfg%trace_unit, '(I0,1X,I0,1X,I0,1X,I0,*(1X,ES24.16E3))') &
            it, round, todo(k), info(k), c(k), th(:, k)
        if (ok(k)) then
          cost(todo(k)) = c(k)
        else
          cost(todo(k)) = ieee_value(1.0_wp, ieee_positive_inf)
        end if
      end do
      res%rejections = res%rejections + count(.not. ok)
      if (.not. retry) exit
      todo = pack(todo, .not. ok)
      deallocate(c, ok, info)
      if (size(todo) == 0) exit
      round = round + 1
      if (round > cfg%max_retries) then
        res%status = -100
        res%message = "PSO: particles stayed non-positive-definite after max_retries re-randomisations"
        return
      end if
      ! particle.f90:117-128: new position in [min, max], velocity in +-(max - min)
      allocate(rp(size(x, 1), size(todo)), rv(size(x, 1), size(todo)))
      call uniform(rp, lo, hi)
      call uniform(rv, lo - hi, hi - lo)
      x(:, todo) = rp
      v(:, todo) = rv
      deallocate(rp, rv)
    end do
  end subroutine evaluate

  subroutine run_pso(gp, cfg, res)
    type(GaussianProcessRegressor), intent(inout) :: gp
    type(PsoSettings), intent(in) :: cfg
    type(PsoResult), intent(out) :: res
    real(wp), allocatable :: x(:, :), v(:, :), pbest(:, :), r1(:, :), r2(:, :)
    real(wp), allocatable :: cost(:), pcost(:), lo(:), hi(:), gbest(:)
    real(wp) :: gcost, stall_cost, fp
    real(wp), allocatable :: xp(:)
    integer :: p, s, i, it, stall_count, all_idx(cfg%swarm_size), ne, na, rc
    integer(8) :: t0, t1, rate
    p = gp%nparams
    s = cfg%swarm_size
    allocate(x(p, s), v(p, s), pbest(p, s), r1(p, s), r2(p, s), cost(s), pcost(s), lo(p), hi(p))
    call gp%search_bounds(lo, hi)
    call seed_rng(cfg%seed)
    call uniform(x, lo, hi)
    call uniform(v, lo - hi, hi - lo)
    all_idx = [(i, i = 1, s)]
    ! initial swarm: evaluated once, no re-randomisation (particle.f90:84-88)
    call evaluate(gp, cfg, x, v, lo, hi, all_idx, cost, res, 0, .false.)
    if (res%status /= 0) return
    pbest = x
    pcost = cost
    gcost = ieee_value(1.0_wp, ieee_positive_inf)
    allocate(gbest(p))
    gbest = x(:, 1)
    do i = 1, s                                     ! state%update in particle order
      if (cost(i) <= gcost) then
        gcost = cost(i)
        gbest = x(:, i)
      end if
    end do
    stall_cost = ieee_value(1.0_wp, ieee_positive_inf)
    stall_count = 0
    call system_clock(t0, rate)
    it = 1
    do
      if (cfg%verbose) print '(A,I0,A,ES23.15E3)', " Iter: ", it, " Best Cost: ", gcost
      call random_number(r1)
      call random_number(r2)
      v = cfg%inertia_weight * v + cfg%cognitive_learning_rate * r1 * (pbest - x) + &
          cfg%social_learning_rate * r2 * (spread(gbest, 2, s) - x)
      x = x + v
      call evaluate(gp, cfg, x, v, lo, hi, all_idx, cost, res, it, .true.)
      if (res%status /= 0) return
      do i = 1, s
        if (cost(i) <= pcost(i)) then
          pbest(:, i) = x(:, i)
          pcost(i) = cost(i)
        end if
      end do
      do i = 1, s
        if (cost(i) <= gcost) then
          gcost = cost(i)
          gbest = x(:, i)
          res%best_iter = it
        end if
      end do
      if (cfg%lbfgs_every > 0) then
        if (mod(it, cfg%lbfgs_every) == 0) then
          xp = gbest
          fp = gcost
          call lbfgs_polish(gp, cfg%lbfgs, xp, fp, ne, na, rc)
          if (rc /= FB_OK) then
            res%status = rc
            res%message = fb_error_message()
            return
          end if
          res%polish_calls = res%polish_calls + 1
          res%polish_evals = res%polish_evals + ne
          if (fp < gcost) then
            gcost = fp
            gbest = xp
            res%best_iter = it
            res%polish_improvements = res%polish_improvements + 1
          end if
          if (cfg%trace) write(cfg%trace_unit, '(A,1X,I0,1X,I0,1X,I0,*(1X,ES24.16E3))') &
              "polish", it, ne, na, fp, xp
        end if
      end if
      res%iterations = it
      if (cfg%relative_change) then
        if (abs((gcost - stall_cost) / stall_cost) <= cfg%tolerance) then
          stall_count = stall_count + 1
        else
          stall_count = 0
          stall_cost = gcost
        end if
        if (stall_count >= cfg%stall_iterations) exit
      end if
      if (it >= cfg%iterations) exit
      it = it + 1
    end do
    call system_clock(t1)
    res%opt_seconds = real(t1 - t0, wp) / real(rate, wp)
    res%best_position = gbest
    res%best_cost = gcost
  end subroutine run_pso
end module fb_pso
