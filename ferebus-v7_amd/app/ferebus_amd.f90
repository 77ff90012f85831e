!> ferebus_amd: trains GP models of atomic properties with PSO, evaluating
!> every swarm on the GPU through libferebus_hip.so.
!>
!> Same driver shape as the reference program (app/main.f90:23-132): for each
!> atom of the system, read SYSTEM_ATOM_TRAINING_SET.csv, for each property
!> build the GP, run the PSO (printing "Opt Time Taken"), and write
!> SYSTEM_property_ATOM.model.
!>
!> usage: ferebus_amd [-c ferebus.toml] [--trace FILE] [--seed N] [--plain]
program ferebus_amd
  use iso_c_binding, only: c_double
  use ferebus_hip, only: FB_OK, FB_LIKELIHOOD_MARGINAL, FB_LIKELIHOOD_CONCENTRATED, fb_error_message
  use fb_config, only: FerebusConfig, read_config, read_training_set
  use fb_kernel_expr, only: KernelProgram, compile_kernel
  use fb_mean, only: mean_kind_from_name
  use fb_gpr, only: GaussianProcessRegressor
  use fb_pso, only: PsoSettings, PsoResult, run_pso
  use fb_model_writer, only: ModelMeta, write_model, model_file_name
  implicit none
  integer, parameter :: wp = c_double
  type(FerebusConfig) :: cfg
  type(KernelProgram) :: prog
  type(GaussianProcessRegressor) :: gp
  type(PsoSettings) :: pso
  type(PsoResult) :: res
  type(ModelMeta) :: meta
  character(len=1024) :: arg, cfg_path, trace_path
  character(len=:), allocatable :: msg, csv
  real(wp), allocatable :: x(:, :), y(:, :)
  character(len=64), allocatable :: labels(:)
  logical :: ok, train
  integer :: i, a, p, rc, mean_kind, lik, u, seed
  integer(8) :: t0, t1, rate

  cfg_path = "ferebus.toml"
  trace_path = ""
  seed = 1
  i = 1
  do while (i <= command_argument_count())
    call get_command_argument(i, arg)
    select case (trim(arg))
    case ("-c", "--config")
      i = i + 1
      call get_command_argument(i, cfg_path)
    case ("--trace")
      i = i + 1
      call get_command_argument(i, trace_path)
    case ("--seed")
      i = i + 1
      call get_command_argument(i, arg)
      read(arg, *) seed
    case ("-h", "--help")
      print '(A)', "usage: ferebus_amd [-c ferebus.toml] [--trace FILE] [--seed N]"
      stop
    end select
    i = i + 1
  end do

  call read_config(trim(cfg_path), cfg, ok, msg)
  if (.not. ok) error stop "ferebus_amd: " // msg
  mean_kind = mean_kind_from_name(cfg%mean)
  if (mean_kind < 0) error stop "ferebus_amd: Unknown Mean Type: " // cfg%mean
  select case (cfg%likelihood)
  case ("marginal"); lik = FB_LIKELIHOOD_MARGINAL
  case ("concentrated"); lik = FB_LIKELIHOOD_CONCENTRATED
  case default; error stop "ferebus_amd: Unknown likelihood function: " // cfg%likelihood
  end select
  pso%swarm_size = cfg%swarm_size
  pso%iterations = cfg%iterations
  pso%inertia_weight = cfg%inertia_weight
  pso%cognitive_learning_rate = cfg%cognitive
  pso%social_learning_rate = cfg%social
  pso%relative_change = cfg%stopping == "relative_change"
  pso%swarm_updater = cfg%swarm_updater
  pso%delta = cfg%delta
  pso%parameter_updater = cfg%parameter_updater
  pso%amin = [cfg%min_w, cfg%min_c1, cfg%min_c2]
  pso%amax = [cfg%max_w, cfg%max_c1, cfg%max_c2]
  pso%an = [cfg%n_w, cfg%n_c1, cfg%n_c2]
  pso%reward_step = cfg%reward_step
  pso%penalty_step = cfg%penalty_step
  pso%success_threshold = cfg%success_threshold
  pso%tolerance = cfg%tolerance
  pso%stall_iterations = cfg%stall_iterations
  pso%lbfgs_every = cfg%lbfgs_every
  pso%lbfgs%steps = cfg%lbfgs_steps
  pso%seed = seed
  if (len_trim(trace_path) > 0) then
    open(newunit=u, file=trim(trace_path), status="replace", action="write")
    pso%trace_unit = u
    pso%trace = .true.
  end if

  do a = 1, size(cfg%atoms)
    csv = cfg%system_name // "_" // trim(cfg%atoms(a)%name) // "_TRAINING_SET.csv"
    call read_training_set(csv, cfg%nfeats, x, y, labels, ok, msg)
    if (.not. ok) error stop "ferebus_amd: " // msg
    call compile_kernel(cfg%kernel, cfg%kernels, size(x, 2), prog)
    if (allocated(prog%error)) error stop "ferebus_amd: " // prog%error
    do p = 1, size(labels)
      train = size(cfg%properties) == 0
      if (.not. train) train = any(cfg%properties == labels(p))
      if (.not. train) cycle
      call system_clock(t0, rate)
      call gp%init(x, y(:, p), prog, mean_kind, cfg%noise, cfg%optimise_noise, lik, cfg%swarm_size, rc)
      if (rc /= FB_OK) error stop "ferebus_amd: " // fb_error_message()
      gp%search_min = cfg%search_min
      gp%search_max = cfg%search_max
      gp%noise_max = cfg%noise_max
      print '(A)', "running pso"
      call run_pso(gp, pso, res)
      if (res%status /= 0) error stop "ferebus_amd: " // res%message
      print '(A,ES23.15E3)', " Best Cost: ", res%best_cost
      print '(A,*(1X,ES23.15E3))', " Best Params:", res%best_position
      print '(A,F12.4,A)', " Opt Time Taken = ", res%opt_seconds, " seconds"
      print '(A,I0,A,I0,A,I0)', " Iterations: ", res%iterations, "  evaluations: ", res%evaluations, &
            "  rejected: ", res%rejections
      if (res%polish_calls > 0) print '(A,I0,A,I0,A,I0)', " L-BFGS polishes: ", res%polish_calls, &
            "  gradient evaluations: ", res%polish_evals, "  improved gbest: ", res%polish_improvements
      meta%system_name = cfg%system_name
      meta%atom_name = trim(cfg%atoms(a)%name)
      meta%property = trim(labels(p))
      meta%natoms = cfg%natoms
      meta%alf = cfg%atoms(a)%alf
      meta%note_keys = cfg%note_keys
      meta%note_values = cfg%note_values
      open(newunit=u, file=model_file_name(meta), status="replace", action="write")
      call write_model(gp, res%best_position, meta, u, rc)
      close(u)
      if (rc /= 0) print '(A,I0)', " warning: best parameters factor with INFO ", rc
      call gp%cleanup()
      call system_clock(t1)
      print '(A,F12.4,A)', " Time Taken = ", real(t1 - t0, wp) / real(rate, wp), " seconds"
    end do
  end do
  if (pso%trace) close(pso%trace_unit)
end program ferebus_amd
