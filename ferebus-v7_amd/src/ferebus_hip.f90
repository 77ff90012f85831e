!> ISO_C_BINDING interface to libferebus_hip.so (include/ferebus_hip.h).
!>
!> The Fortran host side (model, optimiser) calls the gfx950 library through
!> these interfaces only: plain pointers, sizes and integer status codes.
module ferebus_hip
  use iso_c_binding, only: c_int, c_double, c_ptr, c_null_ptr, c_char, c_long_long, &
                           c_associated, c_loc, c_f_pointer
  implicit none
  private

  integer(c_int), parameter, public :: FB_OK = 0
  integer(c_int), parameter, public :: FB_E_ARG = -1, FB_E_SPEC = -2, FB_E_HIP = -3, &
                                       FB_E_NOMEM = -4, FB_E_NOGPU = -5, FB_E_CAPACITY = -6
  integer(c_int), parameter, public :: FB_LEAF_RBF = 1, FB_LEAF_PERIODIC = 2, &
                                       FB_LEAF_RBF_CYCLIC = 3, FB_LEAF_LINEAR = 4, &
                                       FB_LEAF_CONSTANT = 5, FB_OP_SUM = 16, FB_OP_PROD = 17
  integer(c_int), parameter, public :: FB_LIKELIHOOD_MARGINAL = 0, FB_LIKELIHOOD_CONCENTRATED = 1

  !> Postfix kernel program (fb_kernel_spec).
  type, bind(c), public :: fb_kernel_spec
    integer(c_int) :: n_nodes
    type(c_ptr) :: node_type
    type(c_ptr) :: node_ndim
    type(c_ptr) :: active_dims
  end type fb_kernel_spec

  public :: fb_create, fb_destroy, fb_nparams, fb_set_stream, fb_swarm_loglik, &
            fb_swarm_loglik_async, fb_swarm_loglik_static, fb_particle_loglik, fb_get_alpha, &
            fb_swarm_loglik_grad, fb_debug_matrix, &
            fb_multi_create, fb_multi_swarm_loglik, fb_multi_destroy, fb_multi_schedule, fb_multi_pso_step, &
            fb_pso_create, fb_pso_destroy, fb_pso_init, fb_pso_step, fb_pso_retry, fb_pso_select, &
            fb_pso_set_gbest, fb_pso_positions, fb_pso_set_coefficients, &
            fb_profile_enable, fb_profile_read, fb_factor_path, fb_factor_groups, fb_device_name, fb_last_error, &
            fb_error_message

  interface
    integer(c_int) function fb_create(h, n, d, x, y, mean, spec, noise, optimise_noise, &
                                      likelihood, max_swarm) bind(c, name="fb_create")
      import :: c_int, c_double, c_ptr, fb_kernel_spec
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: n, d
      real(c_double), intent(in) :: x(*), y(*)
      type(c_ptr), value :: mean
      type(fb_kernel_spec), intent(in) :: spec
      real(c_double), value :: noise
      integer(c_int), value :: optimise_noise, likelihood, max_swarm
    end function fb_create

    integer(c_int) function fb_destroy(h) bind(c, name="fb_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function fb_destroy

    integer(c_int) function fb_nparams(h, p) bind(c, name="fb_nparams")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: p
    end function fb_nparams

    integer(c_int) function fb_set_stream(h, stream) bind(c, name="fb_set_stream")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, stream
    end function fb_set_stream

    integer(c_int) function fb_swarm_loglik(h, s, theta, ll, info) bind(c, name="fb_swarm_loglik")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: s
      real(c_double), intent(in) :: theta(*)
      real(c_double), intent(out) :: ll(*)
      integer(c_int), intent(out) :: info(*)
    end function fb_swarm_loglik

    integer(c_int) function fb_swarm_loglik_async(h, s, d_theta, d_ll, d_info) &
        bind(c, name="fb_swarm_loglik_async")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: s
      type(c_ptr), value :: d_theta, d_ll, d_info
    end function fb_swarm_loglik_async

    ! no allocation: the library's fixed scratch array on the device (for
    ! callers that cannot allocate); stream = c_null_ptr for the handle's
    integer(c_int) function fb_swarm_loglik_static(h, s, d_theta, d_ll, d_info, stream) &
        bind(c, name="fb_swarm_loglik_static")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: s
      type(c_ptr), value :: d_theta, d_ll, d_info, stream
    end function fb_swarm_loglik_static

    ! one particle's covariance (stage 0) or Cholesky factor (stage 1), n x n
    integer(c_int) function fb_debug_matrix(h, theta, stage, out, info) bind(c, name="fb_debug_matrix")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(in) :: theta(*)
      integer(c_int), value :: stage
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: info
    end function fb_debug_matrix

    integer(c_int) function fb_particle_loglik(h, s, theta, ll, info) bind(c, name="fb_particle_loglik")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: s
      real(c_double), intent(in) :: theta(*)
      real(c_double), intent(out) :: ll(*)
      integer(c_int), intent(out) :: info(*)
    end function fb_particle_loglik

    integer(c_int) function fb_get_alpha(h, theta, alpha, out4, info) bind(c, name="fb_get_alpha")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(in) :: theta(*)
      real(c_double), intent(out) :: alpha(*), out4(4)
      integer(c_int), intent(out) :: info
    end function fb_get_alpha

    integer(c_int) function fb_swarm_loglik_grad(h, s, theta, ll, grad, info) &
        bind(c, name="fb_swarm_loglik_grad")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: s
      real(c_double), intent(in) :: theta(*)
      real(c_double), intent(out) :: ll(*), grad(*)
      integer(c_int), intent(out) :: info(*)
    end function fb_swarm_loglik_grad

    !> Several handles evaluated in one call (uneven N; schedules as in
    !> include/ferebus_hip.h).
    !> gps, theta, ll, info: arrays of c_ptr (c_loc of each handle's arrays).
    integer(c_int) function fb_multi_create(m, n_gps, gps) bind(c, name="fb_multi_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: m
      integer(c_int), value :: n_gps
      type(c_ptr), intent(in) :: gps(*)
    end function fb_multi_create

    integer(c_int) function fb_multi_swarm_loglik(m, s, theta, ll, info) bind(c, name="fb_multi_swarm_loglik")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), intent(in) :: s(*)
      type(c_ptr), intent(in) :: theta(*), ll(*), info(*)
    end function fb_multi_swarm_loglik

    integer(c_int) function fb_multi_destroy(m) bind(c, name="fb_multi_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function fb_multi_destroy

    !> Schedule chosen for the swarm sizes s (0 sequential, 1 persistent,
    !> 2 merged lock-step), -1 before a call with those sizes.
    integer(c_int) function fb_multi_schedule(m, s) bind(c, name="fb_multi_schedule")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), intent(in) :: s(*)
    end function fb_multi_schedule

    !> One PSO iteration of the swarms ps (c_null_ptr skips a handle), all
    !> evaluated in one captured graph; r1, r2, cost, info: arrays of c_ptr.
    integer(c_int) function fb_multi_pso_step(m, ps, r1, r2, cost, info) bind(c, name="fb_multi_pso_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      type(c_ptr), intent(in) :: ps(*), r1(*), r2(*), cost(*), info(*)
    end function fb_multi_pso_step

    ! device-resident particle swarm: one graph replay per PSO iteration
    integer(c_int) function fb_pso_create(p, h, s, w, c1, c2) bind(c, name="fb_pso_create")
      import :: c_int, c_double, c_ptr
      type(c_ptr), intent(out) :: p
      type(c_ptr), value :: h
      integer(c_int), value :: s
      real(c_double), value :: w, c1, c2
    end function fb_pso_create

    integer(c_int) function fb_pso_destroy(p) bind(c, name="fb_pso_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
    end function fb_pso_destroy

    integer(c_int) function fb_pso_init(p, x0, v0, cost, info) bind(c, name="fb_pso_init")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: p
      real(c_double), intent(in) :: x0(*), v0(*)
      real(c_double), intent(out) :: cost(*)
      integer(c_int), intent(out) :: info(*)
    end function fb_pso_init

    integer(c_int) function fb_pso_step(p, r1, r2, cost, info) bind(c, name="fb_pso_step")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: p
      real(c_double), intent(in) :: r1(*), r2(*)
      real(c_double), intent(out) :: cost(*)
      integer(c_int), intent(out) :: info(*)
    end function fb_pso_step

    integer(c_int) function fb_pso_retry(p, k, idx, x, v, cost, info) bind(c, name="fb_pso_retry")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: p
      integer(c_int), value :: k
      integer(c_int), intent(in) :: idx(*)
      real(c_double), intent(in) :: x(*), v(*)
      real(c_double), intent(out) :: cost(*)
      integer(c_int), intent(out) :: info(*)
    end function fb_pso_retry

    integer(c_int) function fb_pso_select(p, gcost, gbest, gidx) bind(c, name="fb_pso_select")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: p
      real(c_double), intent(out) :: gcost
      real(c_double), intent(out) :: gbest(*)
      integer(c_int), intent(out) :: gidx
    end function fb_pso_select

    integer(c_int) function fb_pso_set_gbest(p, gbest, gcost) bind(c, name="fb_pso_set_gbest")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: p
      real(c_double), intent(in) :: gbest(*)
      real(c_double), value :: gcost
    end function fb_pso_set_gbest

    integer(c_int) function fb_pso_set_coefficients(p, w, c1, c2) bind(c, name="fb_pso_set_coefficients")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: p
      real(c_double), value :: w, c1, c2
    end function fb_pso_set_coefficients

    integer(c_int) function fb_pso_positions(p, x, v) bind(c, name="fb_pso_positions")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: p
      real(c_double), intent(out) :: x(*)
      type(c_ptr), value :: v
    end function fb_pso_positions

    integer(c_int) function fb_profile_enable(h, enable) bind(c, name="fb_profile_enable")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: enable
    end function fb_profile_enable

    integer(c_int) function fb_profile_read(h, kind, total_ms, launches) bind(c, name="fb_profile_read")
      import :: c_int, c_double, c_ptr, c_long_long
      type(c_ptr), value :: h
      integer(c_int), value :: kind
      real(c_double), intent(out) :: total_ms
      integer(c_long_long), intent(out) :: launches
    end function fb_profile_read

    integer(c_int) function fb_factor_path(h, s, path) bind(c, name="fb_factor_path")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: s
      integer(c_int), intent(out) :: path
    end function fb_factor_path

    integer(c_int) function fb_factor_groups(h, s, groups) bind(c, name="fb_factor_groups")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: s
      integer(c_int), intent(out) :: groups
    end function fb_factor_groups

    integer(c_int) function fb_device_name(buf, len) bind(c, name="fb_device_name")
      import :: c_int, c_char
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int), value :: len
    end function fb_device_name

    type(c_ptr) function fb_last_error() bind(c, name="fb_last_error")
      import :: c_ptr
    end function fb_last_error
  end interface

contains

  !> Last error message of the library as a Fortran string.
  function fb_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), dimension(:), pointer :: chars
    type(c_ptr) :: p
    integer :: n
    p = fb_last_error()
    msg = ""
    if (.not. c_associated(p)) return
    call c_f_pointer(p, chars, [1024])
    n = 0
    do while (n < 1024)
      if (chars(n + 1) == achar(0)) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(chars(1:n), msg)
  end function fb_error_message
end module ferebus_hip
