!> Host-side checks of the Fortran layer without a GPU: prints what the TOML
!> reader, the kernel-expression compiler, the CSV reader and the mean
!> functions produce, for tests/test_fortran_host.py to compare.
!> usage: test_host <ferebus.toml> <training.csv>
program test_host
  use iso_c_binding, only: c_double
  use fb_config, only: FerebusConfig, read_config, read_training_set
  use fb_kernel_expr, only: KernelProgram, compile_kernel, KernelConfigEntry
  use fb_mean, only: MeanFit, mean_kind_from_name
  implicit none
  type(FerebusConfig) :: cfg
  type(KernelProgram) :: prog
  type(MeanFit) :: mf
  character(len=1024) :: path, csv
  character(len=:), allocatable :: msg
  real(c_double), allocatable :: x(:, :), y(:, :), m(:)
  character(len=64), allocatable :: labels(:)
  logical :: ok
  integer :: k, kind
  type(KernelConfigEntry), allocatable :: extra(:)
  character(len=32), parameter :: exprs(6) = [character(len=32) :: "k1+k2", "k1*k2*k1", "(k1+k2)*k1", &
                                                "2*k1", "k1+k2+k1", "k1-k2"]

  call get_command_argument(1, path)
  call get_command_argument(2, csv)
  call read_config(trim(path), cfg, ok, msg)
  if (.not. ok) error stop msg
  print '(A,A)', "system.name=", cfg%system_name
  print '(A,I0)', "system.natoms=", cfg%natoms
  print '(A,I0)', "system.nfeats=", cfg%nfeats
  print '(A,I0)', "atoms=", size(cfg%atoms)
  print '(A,A)', "atom1.name=", trim(cfg%atoms(1)%name)
  print '(A,3(I0,1X))', "atom1.alf=", cfg%atoms(1)%alf
  print '(A,A)', "model.mean=", cfg%mean
  print '(A,A)', "model.kernel=", cfg%kernel
  print '(A,A)', "model.likelihood=", cfg%likelihood
  print '(A,ES24.16)', "model.noise=", cfg%noise
  print '(A,ES24.16)', "optimiser.search_min=", cfg%search_min
  print '(A,ES24.16)', "optimiser.search_max=", cfg%search_max
  print '(A,I0)', "pso.swarm_size=", cfg%swarm_size
  print '(A,I0)', "pso.iterations=", cfg%iterations
  print '(A,ES24.16)', "pso.inertia_weight=", cfg%inertia_weight
  print '(A,A)', "pso.stopping=", cfg%stopping
  print '(A,ES24.16)', "pso.tolerance=", cfg%tolerance
  print '(A,I0)', "pso.stall_iterations=", cfg%stall_iterations
  print '(A,A)', "pso.swarm_updater=", cfg%swarm_updater
  print '(A,ES24.16)', "pso.delta=", cfg%delta
  print '(A,A)', "pso.parameter_updater=", cfg%parameter_updater
  print '(A,3(ES24.16,1X))', "pso.adaptive_min=", cfg%min_w, cfg%min_c1, cfg%min_c2
  print '(A,3(ES24.16,1X))', "pso.adaptive_max=", cfg%max_w, cfg%max_c1, cfg%max_c2
  print '(A,3(I0,1X))', "pso.adaptive_n=", cfg%n_w, cfg%n_c1, cfg%n_c2
  print '(A,3(ES24.16,1X))', "pso.adaptive_steps=", cfg%reward_step, cfg%penalty_step, cfg%success_threshold
  print '(A,I0)', "kernels=", size(cfg%kernels)
  do k = 1, size(cfg%kernels)
    if (allocated(cfg%kernels(k)%active_dims)) then
      print '(A,A,A,A,A,*(I0,1X))', "kernel.", trim(cfg%kernels(k)%name), "=", trim(cfg%kernels(k)%ktype), &
            " dims ", cfg%kernels(k)%active_dims
    else
      print '(A,A,A,A)', "kernel.", trim(cfg%kernels(k)%name), "=", trim(cfg%kernels(k)%ktype)
    end if
  end do
  print '(A,I0)', "notes=", size(cfg%note_keys)
  do k = 1, size(cfg%note_keys)
    print '(A,A,A,A)', "note.", trim(cfg%note_keys(k)), "=", trim(cfg%note_values(k))
  end do

  ! kernel expressions over the configured kernels plus a periodic k2 on dim 3
  allocate(extra(2))
  extra(1)%name = "k1"
  extra(1)%ktype = "rbf"
  extra(1)%active_dims = [1, 2]
  extra(2)%name = "k2"
  extra(2)%ktype = "periodic"
  extra(2)%active_dims = [3]
  do k = 1, size(exprs)
    call compile_kernel(trim(exprs(k)), extra, 3, prog)
    if (allocated(prog%error)) then
      print '(A,A,A,A)', "expr ", trim(exprs(k)), " error ", prog%error
    else
      print '(A,A,A,*(I0,1X))', "expr ", trim(exprs(k)), " types ", prog%node_type(1:prog%n_nodes)
      print '(A,A,A,*(I0,1X))', "expr ", trim(exprs(k)), " ndims ", prog%node_ndim(1:prog%n_nodes)
      print '(A,A,A,A)', "expr ", trim(exprs(k)), " name ", prog%composition
      print '(A,A,A,I0)', "expr ", trim(exprs(k)), " nparams ", prog%n_params()
    end if
  end do

  call read_training_set(trim(csv), cfg%nfeats, x, y, labels, ok, msg)
  if (.not. ok) error stop msg
  print '(A,I0,1X,I0,1X,I0)', "csv.shape=", size(x, 1), size(x, 2), size(y, 2)
  print '(A,A,1X,A)', "csv.labels=", trim(labels(1)), trim(labels(size(labels)))
  print '(A,3ES24.16)', "csv.x1=", x(1, :)
  print '(A,ES24.16)', "csv.ylast1=", y(1, size(y, 2))
  allocate(m(size(x, 1)))
  do kind = 0, 3
    call mf%fit(kind, x, y(:, size(y, 2)))
    call mf%value(x, m)
    print '(A,I0,A,3ES24.16)', "mean", kind, "=", m(1), m(2), sum(m)
  end do
end program test_host
