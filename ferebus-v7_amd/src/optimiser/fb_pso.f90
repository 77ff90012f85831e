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
!>  * the swarm lives on the device (fb_pso_*): one iteration's selection of
!>    the previous bests, velocity / position update, build, factorisation
!>    and likelihood are one graph replay; r1, r2 and re-randomised
!>    particles come from this host's RNG;
!>  * pbest with <= (pso.f90:440-443), gbest in particle order with <=
!>    (pso.f90:454-459), best state with <= (iter_state.f90:62-82);
!>  * stopping: relative change of the best cost below `tolerance` for
!>    `stall_iterations` iterations (pso.f90:619-636), or the iteration cap
!>    (optimiser_type.f90:122-132).
!>  * swarm updaters: "global" (the above) or "forced" (pso.f90:485-558):
!>    particles updated one after another with their own r1, r2, each
!>    evaluated and selected before the next moves (gbest only when pbest
!>    improved); velocity components whose partial potential
!>    |v| + |gbest - x| is below delta are re-drawn in [-delta, delta].  The
!>    forced rule is sequential, so it runs on the host, one particle per
!>    evaluation;
!>  * parameter updaters (pso.f90:638-825): constant; unified / individual
!>    adaptive (learning automata over w, c1, c2,
!>    learning_automata.f90:48-116, rewarded when n_improved / n reaches
!>    success_threshold in integer division, pso.f90:760); random unified /
!>    individual (w ~ U[0,1], c1 ~ U[0, b], c2 ~ U[0, b - c1],
!>    b = 24 (1 - w^2) / (7 - 5 w)).  Unified variants update after every
!>    swarm update, individual ones after every particle of the forced
!>    updater (never under the global updater, pso.f90:445-447).  The device
!>    swarm takes the coefficients through fb_pso_set_coefficients.
!> Every evaluation of a whole swarm is one call of the gfx950 library.
!> Optional (off by default, lbfgs_every = 0): every lbfgs_every iterations
!> the global best is polished by a few L-BFGS steps on the adjoint gradient
!> (fb_lbfgs) and replaces gbest when the polished cost is lower.
module fb_pso
  use iso_c_binding, only: c_int, c_double, c_ptr, c_null_ptr
  use ieee_arithmetic, only: ieee_value, ieee_positive_inf
  use ferebus_hip, only: FB_OK, fb_error_message, fb_pso_create, fb_pso_destroy, fb_pso_init, fb_pso_step, &
                         fb_pso_retry, fb_pso_select, fb_pso_set_gbest, fb_pso_positions, &
                         fb_pso_set_coefficients
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
    character(len=32) :: swarm_updater = "global"          ! or "forced"
    real(wp) :: delta = 1e-7_wp                            ! forced updater (swarm_updater.f90:58)
    character(len=32) :: parameter_updater = "constant"    ! unified_adaptive, individual_adaptive,
                                                           ! random_unified, random_individual
    real(wp) :: amin(3) = [0.0_wp, 0.0_wp, 0.0_wp]         ! adaptive: w, c1, c2 ranges and
    real(wp) :: amax(3) = [1.0_wp, 4.0_wp, 4.0_wp]         ! action counts
    integer :: an(3) = [10, 10, 10]                        ! (parameter_updater.f90:62-75)
    real(wp) :: reward_step = 0.01_wp, penalty_step = 0.01_wp, success_threshold = 0.5_wp
  end type PsoSettings

  !> learning_automata.f90:48-116
  type :: LearningAutomata
    real(wp), allocatable :: values(:), prob(:)
    integer :: selected = 1
    real(wp) :: reward_step = 0.0_wp, penalty_step = 0.0_wp
  end type LearningAutomata

  !> The swarm's (w, c1, c2) and how they change (pso.f90:638-825).
  type :: ParamUpdater
    character(len=32) :: kind = "constant"
    logical :: per_swarm = .false., per_particle = .false.
    real(wp) :: coef(3) = 0.0_wp
    type(LearningAutomata) :: la(3)
    real(wp) :: threshold = 0.5_wp
    real(wp), allocatable :: previous(:)
  end type ParamUpdater

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

  !> Writes one trace line per listed particle.
  subroutine trace_lines(cfg, it, round, idx, info, c, th)
    type(PsoSettings), intent(in) :: cfg
    integer, intent(in) :: it, round, idx(:)
    integer(c_int), intent(in) :: info(:)
    real(wp), intent(in) :: c(:), th(:, :)
    integer :: k
    if (.not. cfg%trace) return
    do k = 1, size(idx)
      write(cfg%trace_unit, '(I0,1X,I0,1X,I0,1X,I0,*(1X,ES24.16E3))') it, round, idx(k), info(k), c(k), th(:, k)
    end do
  end subroutine trace_lines

  subroutine la_init(la, lo, hi, n, rstep, pstep)
    type(LearningAutomata), intent(out) :: la
    real(wp), intent(in) :: lo, hi, rstep, pstep
    integer, intent(in) :: n
    integer :: i
    allocate(la%values(n), la%prob(n))
    do i = 1, n
      la%values(i) = lo + real(i - 1, wp) * (hi - lo) / real(n - 1, wp)
      la%prob(i) = 1.0_wp / real(n, wp)
    end do
    la%reward_step = rstep
    la%penalty_step = pstep
  end subroutine la_init

  !> The last action whose closed interval holds u * sum(prob) (the loop of
  !> the reference's choose does not stop at the first match).
  function la_choose(la) result(v)
    type(LearningAutomata), intent(inout) :: la
    real(wp) :: v, u, rnd, prev, total
    integer :: i
    call random_number(u)
    total = 0.0_wp
    do i = 1, size(la%prob)
      total = total + la%prob(i)
    end do
    rnd = u * total
    prev = 0.0_wp
    do i = 1, size(la%prob)
      if (rnd >= prev .and. rnd <= prev + la%prob(i)) la%selected = i
      prev = prev + la%prob(i)
    end do
    v = la%values(la%selected)
  end function la_choose

  subroutine la_reward(la)
    type(LearningAutomata), intent(inout) :: la
    integer :: i
    do i = 1, size(la%prob)
      if (i == la%selected) then
        la%prob(i) = la%prob(i) + la%reward_step * (1.0_wp - la%prob(i))
      else
        la%prob(i) = la%prob(i) * (1.0_wp - la%reward_step)
      end if
    end do
  end subroutine la_reward

  subroutine la_penalise(la)
    type(LearningAutomata), intent(inout) :: la
    integer :: i
    do i = 1, size(la%prob)
      if (i == la%selected) then
        la%prob(i) = la%prob(i) * (1.0_wp - la%penalty_step)
      else
        la%prob(i) = la%penalty_step / real(size(la%prob) - 1, wp) + (1.0_wp - la%penalty_step) * la%prob(i)
      end if
    end do
  end subroutine la_penalise

  subroutine upd_random(u)
    type(ParamUpdater), intent(inout) :: u
    real(wp) :: r, b
    call random_number(r)
    u%coef(1) = r
    b = 24.0_wp * (1.0_wp - u%coef(1) * u%coef(1)) / (7.0_wp - 5.0_wp * u%coef(1))
    call random_number(r)
    u%coef(2) = 0.0_wp + r * (b - 0.0_wp)
    call random_number(r)
    u%coef(3) = 0.0_wp + r * ((b - u%coef(2)) - 0.0_wp)
  end subroutine upd_random

  subroutine upd_init(u, cfg, costs)
    type(ParamUpdater), intent(out) :: u
    type(PsoSettings), intent(in) :: cfg
    real(wp), intent(in) :: costs(:)
    integer :: j
    u%kind = cfg%parameter_updater
    u%coef = [cfg%inertia_weight, cfg%cognitive_learning_rate, cfg%social_learning_rate]
    u%per_swarm = u%kind == "unified_adaptive" .or. u%kind == "random_unified"
    u%per_particle = u%kind == "individual_adaptive" .or. u%kind == "random_individual"
    select case (trim(u%kind))
    case ("unified_adaptive", "individual_adaptive")
      do j = 1, 3
        call la_init(u%la(j), cfg%amin(j), cfg%amax(j), cfg%an(j), cfg%reward_step, cfg%penalty_step)
      end do
      u%threshold = cfg%success_threshold
      u%previous = costs
      do j = 1, 3
        u%coef(j) = la_choose(u%la(j))
      end do
    case ("random_unified", "random_individual")
      call upd_random(u)
    end select
  end subroutine upd_init

  !> update_parameters over particles first .. first + size(costs) - 1.
  subroutine upd_update(u, costs, first)
    type(ParamUpdater), intent(inout) :: u
    real(wp), intent(in) :: costs(:)
    integer, intent(in) :: first
    integer :: i, improved, j
    select case (trim(u%kind))
    case ("random_unified", "random_individual")
      call upd_random(u)
    case ("unified_adaptive", "individual_adaptive")
      improved = 0
      do i = 1, size(costs)
        if (costs(i) < u%previous(first + i - 1)) improved = improved + 1   ! pso.f90:752-757
        u%previous(first + i - 1) = costs(i)
      end do
      if (real(improved / size(costs), wp) >= u%threshold) then   ! integer division, pso.f90:760
        do j = 1, 3
          call la_reward(u%la(j))
        end do
      else
        do j = 1, 3
          call la_penalise(u%la(j))
        end do
      end if
      do j = 1, 3
        u%coef(j) = la_choose(u%la(j))
      end do
    end select
  end subroutine upd_update

  !> Failed particles of this iteration: re-randomised (particle.f90:117-128)
  !> and re-evaluated together on the device until every one succeeds
  !> (get_cost, particle.f90:149-162).
  subroutine retry_failed(ps, cfg, lo, hi, cost, info, res, it)
    type(c_ptr), intent(in) :: ps
    type(PsoSettings), intent(in) :: cfg
    real(wp), intent(in) :: lo(:), hi(:)
    real(wp), intent(inout) :: cost(:)
    integer(c_int), intent(inout) :: info(:)
    type(PsoResult), intent(inout) :: res
    integer, intent(in) :: it
    integer, allocatable :: todo(:)
    integer(c_int), allocatable :: idx0(:), inf(:)
    real(wp), allocatable :: rp(:, :), rv(:, :), c(:)
    integer :: round, k, i, p
    p = size(lo)
    todo = pack([(i, i = 1, size(cost))], info /= 0)
    round = 0
    do while (size(todo) > 0)
      res%rejections = res%rejections + size(todo)
      round = round + 1
      if (round > cfg%max_retries) then
        res%status = -100
        res%message = "PSO: particles stayed non-positive-definite after max_retries re-randomisations"
        return
      end if
      k = size(todo)
      allocate(rp(p, k), rv(p, k), c(k), inf(k))
      call uniform(rp, lo, hi)
      call uniform(rv, lo - hi, hi - lo)
      idx0 = int(todo - 1, c_int)
      res%status = fb_pso_retry(ps, int(k, c_int), idx0, rp, rv, c, inf)
      if (res%status /= FB_OK) then
        res%message = fb_error_message()
        return
      end if
      res%evaluations = res%evaluations + k
      call trace_lines(cfg, it, round, todo, inf, c, rp)
      cost(todo) = c
      info(todo) = inf
      todo = pack(todo, inf /= 0)
      deallocate(rp, rv, c, inf)
    end do
  end subroutine retry_failed

  !> The PSO over the device-resident swarm (fb_pso_*): each iteration's
  !> selection, update, build, factorisation and likelihood are one graph
  !> replay; r1, r2 and re-randomised particles come from this host's RNG.
  subroutine run_pso(gp, cfg, res)
    type(GaussianProcessRegressor), intent(inout) :: gp
    type(PsoSettings), intent(in) :: cfg
    type(PsoResult), intent(out) :: res
    real(wp), allocatable :: x(:, :), v(:, :), r1(:, :), r2(:, :)
    real(wp), allocatable :: cost(:), lo(:), hi(:), gbest(:)
    real(wp) :: gcost, stall_cost, fp, gdev
    real(wp), allocatable :: xp(:)
    integer(c_int), allocatable :: info(:)
    integer(c_int) :: gidx
    integer :: p, s, i, it, stall_count, ne, na, rc
    integer(8) :: t0, t1, rate
    type(c_ptr) :: ps
    type(ParamUpdater) :: upd
    if (trim(cfg%swarm_updater) == "forced") then
      call run_forced(gp, cfg, res)
      return
    end if
    p = gp%nparams
    s = cfg%swarm_size
    allocate(x(p, s), v(p, s), r1(p, s), r2(p, s), cost(s), info(s), lo(p), hi(p), gbest(p))
    call gp%search_bounds(lo, hi)
    call seed_rng(cfg%seed)
    call uniform(x, lo, hi)
    call uniform(v, lo - hi, hi - lo)
    res%status = fb_pso_create(ps, gp%h, int(s, c_int), cfg%inertia_weight, cfg%cognitive_learning_rate, &
                               cfg%social_learning_rate)
    if (res%status /= FB_OK) then
      res%message = fb_error_message()
      return
    end if
    ! initial swarm: evaluated once, no re-randomisation (particle.f90:84-88)
    res%status = fb_pso_init(ps, x, v, cost, info)
    if (res%status /= FB_OK) then
      res%message = fb_error_message()
      rc = fb_pso_destroy(ps)
      return
    end if
    res%evaluations = s
    res%rejections = count(info /= 0)
    call trace_lines(cfg, 0, 0, [(i, i = 1, s)], info, cost, x)
    gcost = ieee_value(1.0_wp, ieee_positive_inf)
    do i = 1, s                                     ! state%update in particle order
      if (cost(i) <= gcost) gcost = cost(i)
    end do
    call upd_init(upd, cfg, cost)
    stall_cost = ieee_value(1.0_wp, ieee_positive_inf)
    stall_count = 0
    call system_clock(t0, rate)
    it = 1
    do
      if (cfg%verbose) print '(A,I0,A,ES23.15E3)', " Iter: ", it, " Best Cost: ", gcost
      res%status = fb_pso_set_coefficients(ps, upd%coef(1), upd%coef(2), upd%coef(3))
      if (res%status /= FB_OK) exit
      call random_number(r1)
      call random_number(r2)
      res%status = fb_pso_step(ps, r1, r2, cost, info)
      if (res%status /= FB_OK) exit
      res%evaluations = res%evaluations + s
      if (cfg%trace) then
        res%status = fb_pso_positions(ps, x, c_null_ptr)
        if (res%status /= FB_OK) exit
        call trace_lines(cfg, it, 0, [(i, i = 1, s)], info, cost, x)
      end if
      call retry_failed(ps, cfg, lo, hi, cost, info, res, it)
      if (res%status /= 0) exit
      do i = 1, s                                   ! pbest / gbest are selected on the device
        if (cost(i) <= gcost) then
          gcost = cost(i)
          res%best_iter = it
        end if
      end do
      if (upd%per_swarm) call upd_update(upd, cost, 1)   ! next_iter_pso, pso.f90:563-565
      if (cfg%lbfgs_every > 0) then
        if (mod(it, cfg%lbfgs_every) == 0) then
          res%status = fb_pso_select(ps, gdev, gbest, gidx)
          if (res%status /= FB_OK) exit
          xp = gbest
          fp = gdev
          call lbfgs_polish(gp, cfg%lbfgs, xp, fp, ne, na, rc)
          if (rc /= FB_OK) then
            res%status = rc
            exit
          end if
          res%polish_calls = res%polish_calls + 1
          res%polish_evals = res%polish_evals + ne
          if (fp < gdev) then
            res%status = fb_pso_set_gbest(ps, xp, fp)
            if (res%status /= FB_OK) exit
            gcost = fp
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
    if (res%status == 0) res%status = fb_pso_select(ps, gdev, gbest, gidx)
    if (res%status /= 0) then
      if (.not. allocated(res%message)) res%message = fb_error_message()
      rc = fb_pso_destroy(ps)
      return
    end if
    call system_clock(t1)
    rc = fb_pso_destroy(ps)
    res%opt_seconds = real(t1 - t0, wp) / real(rate, wp)
    res%best_position = gbest
    res%best_cost = gdev
  end subroutine run_pso
  !> forced_swarm_update (pso.f90:485-558): the particles move one after
  !> another, each evaluated (get_cost with its re-randomisations,
  !> particle.f90:149-162) and selected before the next moves.  The state
  !> lives on the host; every particle is one evaluation of the library (a
  !> swarm of one), as the rule is sequential by definition.
  subroutine run_forced(gp, cfg, res)
    type(GaussianProcessRegressor), intent(inout) :: gp
    type(PsoSettings), intent(in) :: cfg
    type(PsoResult), intent(out) :: res
    real(wp), allocatable :: x(:, :), v(:, :), pbest(:, :), cost(:), pcost(:), lo(:), hi(:), gbest(:)
    real(wp), allocatable :: ra(:), rb(:), dv(:), pp(:), c1v(:)
    logical, allocatable :: ok(:)
    integer(c_int), allocatable :: info(:)
    real(wp) :: gcost, stall_cost, w, c1, c2
    integer :: p, s, i, k, it, stall_count, rnd, rc
    integer(8) :: t0, t1, rate
    type(ParamUpdater) :: upd
    p = gp%nparams
    s = cfg%swarm_size
    allocate(x(p, s), v(p, s), pbest(p, s), cost(s), pcost(s), info(s), ok(s), lo(p), hi(p), gbest(p))
    allocate(ra(p), rb(p), dv(p), pp(p), c1v(1))
    call gp%search_bounds(lo, hi)
    call seed_rng(cfg%seed)
    call uniform(x, lo, hi)
    call uniform(v, lo - hi, hi - lo)
    call system_clock(t0, rate)
    call gp%swarm_cost(x, cost, ok, info, rc)
    if (rc /= FB_OK) then
      res%status = rc
      res%message = fb_error_message()
      return
    end if
    where (info /= 0) cost = ieee_value(1.0_wp, ieee_positive_inf)
    res%evaluations = s
    res%rejections = count(info /= 0)
    call trace_lines(cfg, 0, 0, [(i, i = 1, s)], info, cost, x)
    pbest = x
    pcost = cost
    gcost = ieee_value(1.0_wp, ieee_positive_inf)
    gbest = x(:, 1)
    do i = 1, s
      if (cost(i) <= gcost) then
        gcost = cost(i)
        gbest = x(:, i)
      end if
    end do
    call upd_init(upd, cfg, cost)
    stall_cost = ieee_value(1.0_wp, ieee_positive_inf)
    stall_count = 0
    it = 1
    do
      if (cfg%verbose) print '(A,I0,A,ES23.15E3)', " Iter: ", it, " Best Cost: ", gcost
      do k = 1, s
        w = upd%coef(1)
        c1 = upd%coef(2)
        c2 = upd%coef(3)
        call random_number(ra)
        call random_number(rb)
        v(:, k) = w * v(:, k) + c1 * ra * (pbest(:, k) - x(:, k)) + c2 * rb * (gbest - x(:, k))
        pp = abs(v(:, k)) + abs(gbest - x(:, k))
        if (any(pp < cfg%delta)) then
          call random_number(dv)
          dv = -cfg%delta + dv * (cfg%delta - (-cfg%delta))
          v(:, k) = merge(dv, v(:, k), pp < cfg%delta)
        end if
        x(:, k) = x(:, k) + v(:, k)
        rnd = 0
        do
          call gp%swarm_cost(x(:, k:k), c1v, ok(k:k), info(k:k), rc)
          if (rc /= FB_OK) then
            res%status = rc
            res%message = fb_error_message()
            return
          end if
          res%evaluations = res%evaluations + 1
          if (info(k) /= 0) c1v(1) = ieee_value(1.0_wp, ieee_positive_inf)
          call trace_lines(cfg, it, rnd, [k], info(k:k), c1v, x(:, k:k))
          if (info(k) == 0) exit
          res%rejections = res%rejections + 1
          rnd = rnd + 1
          if (rnd > cfg%max_retries) then
            res%status = -100
            res%message = "PSO: a particle stayed non-positive-definite after max_retries re-randomisations"
            return
          end if
          call uniform(x(:, k:k), lo, hi)                ! particle.f90:117-128
          call uniform(v(:, k:k), lo - hi, hi - lo)
        end do
        cost(k) = c1v(1)
        if (cost(k) <= pcost(k)) then
          pbest(:, k) = x(:, k)
          pcost(k) = cost(k)
          if (cost(k) <= gcost) then
            gbest = x(:, k)
            gcost = cost(k)
            res%best_iter = it
          end if
        end if
        if (upd%per_particle) call upd_update(upd, cost(k:k), k)
      end do
      if (upd%per_swarm) call upd_update(upd, cost, 1)
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
  end subroutine run_forced
end module fb_pso
