!> Gaussian-process regressor over the gfx950 swarm log-likelihood library.
!>
!> Host side of GaussianProcessRegressor (src/model/gpr/gpr.f90): it owns the
!> training set, the cached mean (gpr.f90:168-171) and a library handle, and
!> evaluates cost = -logL for a whole swarm per call instead of one particle
!> per OpenMP thread (gpr.f90:258-289 called from pso.f90:419-452).
module fb_gpr
  use iso_c_binding, only: c_int, c_double, c_ptr, c_null_ptr, c_loc, c_associated
  use ferebus_hip
  use fb_mean, only: MeanFit, MEAN_ZERO, MEAN_CONSTANT
  use fb_kernel_expr, only: KernelProgram
  implicit none
  private
  integer, parameter :: wp = c_double

  type, public :: GaussianProcessRegressor
    type(c_ptr) :: h = c_null_ptr
    integer :: n = 0, d = 0, nparams = 0, nparams_kernel = 0, max_swarm = 0
    integer :: likelihood = FB_LIKELIHOOD_MARGINAL
    logical :: optimise_noise = .false.
    real(wp) :: noise = 1e-10_wp
    !> search box of the kernel parameters (optimiser search_min / search_max)
    real(wp) :: search_min = 0.0_wp, search_max = 1.0_wp
    !> noise search box, hard-coded to [1e-12, 1e-6] in gpr.f90:229,242
    real(wp) :: noise_min = 1e-12_wp, noise_max = 1e-6_wp
    type(KernelProgram) :: prog
    type(MeanFit) :: mean
    real(wp), allocatable :: x(:, :), y(:), m(:)
    integer(c_int), allocatable, private :: spec_type(:), spec_ndim(:), spec_dims(:)
  contains
    procedure :: init => gpr_init
    procedure :: swarm_cost => gpr_swarm_cost
    procedure :: particle_cost => gpr_particle_cost
    procedure :: swarm_grad => gpr_swarm_grad
    procedure :: search_bounds => gpr_search_bounds
    procedure :: weights => gpr_weights
    procedure :: cleanup => gpr_cleanup
  end type GaussianProcessRegressor

contains

  !> gpr.f90:124-183: kernel program, mean fit + cached value, device handle.
  subroutine gpr_init(self, x, y, prog, mean_kind, noise, optimise_noise, likelihood, max_swarm, rc)
    class(GaussianProcessRegressor), intent(inout), target :: self
    real(wp), intent(in) :: x(:, :), y(:)
    type(KernelProgram), intent(in) :: prog
    integer, intent(in) :: mean_kind, likelihood, max_swarm
    real(wp), intent(in) :: noise
    logical, intent(in) :: optimise_noise
    integer, intent(out) :: rc
    type(fb_kernel_spec) :: spec
    self%n = size(y)
    self%d = size(x, 2)
    self%max_swarm = max_swarm
    self%x = x
    self%y = y
    self%prog = prog
    self%noise = noise
    self%optimise_noise = optimise_noise
    self%likelihood = likelihood
    call self%mean%fit(mean_kind, x, y)
    allocate(self%m(self%n))
    call self%mean%value(x, self%m)
    self%spec_type = prog%node_type(1:prog%n_nodes)
    self%spec_ndim = prog%node_ndim(1:prog%n_nodes)
    if (size(prog%active_dims) > 0) then
      self%spec_dims = prog%active_dims
    else
      self%spec_dims = [0_c_int]
    end if
    spec%n_nodes = prog%n_nodes
    spec%node_type = c_loc(self%spec_type)
    spec%node_ndim = c_loc(self%spec_ndim)
    spec%active_dims = c_loc(self%spec_dims)
    rc = fb_create(self%h, int(self%n, c_int), int(self%d, c_int), self%x, self%y, c_loc(self%m), spec, &
                   noise, merge(1_c_int, 0_c_int, optimise_noise), int(likelihood, c_int), &
                   int(max_swarm, c_int))
    if (rc /= FB_OK) return
    self%nparams_kernel = prog%n_params()
    self%nparams = self%nparams_kernel + merge(1, 0, optimise_noise)
  end subroutine gpr_init

  !> cost(s) = -logL(theta(:, s)), success(s) = (INFO == 0)
  !> (gpr.f90:258-289, iter_result.f90:92-101), all particles in one call.
  subroutine gpr_swarm_cost(self, theta, cost, success, info, rc)
    class(GaussianProcessRegressor), intent(inout) :: self
    real(wp), intent(in) :: theta(:, :)
    real(wp), intent(out) :: cost(:)
    logical, intent(out) :: success(:)
    integer(c_int), intent(out) :: info(:)
    integer, intent(out) :: rc
    real(wp), allocatable :: ll(:)
    allocate(ll(size(theta, 2)))
    rc = fb_swarm_loglik(self%h, int(size(theta, 2), c_int), theta, ll, info)
    cost = -ll
    success = info == 0
  end subroutine gpr_swarm_cost

  !> The plain per-particle path, same results.
  subroutine gpr_particle_cost(self, theta, cost, success, info, rc)
    class(GaussianProcessRegressor), intent(inout) :: self
    real(wp), intent(in) :: theta(:, :)
    real(wp), intent(out) :: cost(:)
    logical, intent(out) :: success(:)
    integer(c_int), intent(out) :: info(:)
    integer, intent(out) :: rc
    real(wp), allocatable :: ll(:)
    allocate(ll(size(theta, 2)))
    rc = fb_particle_loglik(self%h, int(size(theta, 2), c_int), theta, ll, info)
    cost = -ll
    success = info == 0
  end subroutine gpr_particle_cost

  !> Gradient of the cost (= -d logL / d theta) of each particle.
  subroutine gpr_swarm_grad(self, theta, cost, grad, info, rc)
    class(GaussianProcessRegressor), intent(inout) :: self
    real(wp), intent(in) :: theta(:, :)
    real(wp), intent(out) :: cost(:), grad(:, :)
    integer(c_int), intent(out) :: info(:)
    integer, intent(out) :: rc
    real(wp), allocatable :: ll(:)
    allocate(ll(size(theta, 2)))
    rc = fb_swarm_loglik_grad(self%h, int(size(theta, 2), c_int), theta, ll, grad, info)
    cost = -ll
    grad = -grad
  end subroutine gpr_swarm_grad

  !> search_min_gpr / search_max_gpr (gpr.f90:204-246).
  subroutine gpr_search_bounds(self, lo, hi)
    class(GaussianProcessRegressor), intent(in) :: self
    real(wp), intent(out) :: lo(:), hi(:)
    lo(1:self%nparams_kernel) = self%search_min
    hi(1:self%nparams_kernel) = self%search_max
    if (self%optimise_noise) then
      lo(self%nparams) = self%noise_min
      hi(self%nparams) = self%noise_max
    end if
  end subroutine gpr_search_bounds

  !> alpha and [ll, logdet, mu_hat, sigma2_hat] of one theta (model output).
  subroutine gpr_weights(self, theta, alpha, out4, info, rc)
    class(GaussianProcessRegressor), intent(inout) :: self
    real(wp), intent(in) :: theta(:)
    real(wp), intent(out) :: alpha(:), out4(4)
    integer(c_int), intent(out) :: info
    integer, intent(out) :: rc
    rc = fb_get_alpha(self%h, theta, alpha, out4, info)
  end subroutine gpr_weights

  subroutine gpr_cleanup(self)
    class(GaussianProcessRegressor), intent(inout) :: self
    integer(c_int) :: rc
    if (c_associated(self%h)) rc = fb_destroy(self%h)
    self%h = c_null_ptr
  end subroutine gpr_cleanup
end module fb_gpr
