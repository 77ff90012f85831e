/*
 * ferebus_hip.h -- C ABI of the MI355X (gfx950) swarm log-likelihood library.
 *
 * One call evaluates the Gaussian-process log-likelihood of every particle of a
 * PSO swarm: covariance build K(theta_s), Cholesky factorisation, log-det and
 * the quadratic form, for all s at once on one GPU, in FP64.
 *
 * Plain C: pointers, sizes and int return codes.  No HIP or torch types appear
 * in the signatures; streams are passed as `void*` (a hipStream_t, NULL = the
 * library's own stream).  Host arrays are caller owned and only read during the
 * call.  Every device buffer is owned by the handle and released by fb_destroy.
 *
 * Return codes: 0 = success, < 0 = error (see FB_E_*).  Per-particle Cholesky
 * status uses LAPACK INFO semantics: info[s] = 0 on success, k > 0 when the
 * leading minor of order k is not positive definite (pivot <= 0 or NaN).
 * ll[s] is unspecified when info[s] != 0.
 *
 * Reference seams replaced (popelier-group/FEREBUS-v7, Fortran):
 *   GaussianProcessRegressor%cost_function   src/model/gpr/gpr.f90:258-289
 *   GaussianProcessRegressor%R (compute_R)    src/model/gpr/gpr.f90:185-189
 *   Kernel%R and its overrides                src/kernels/kernel_module.f90:208-219,309-325,359-375
 *   Likelihood%ll (marginal / concentrated)   src/model/gpr/likelihood.f90:161-253
 *   potrf / potrs wrappers                    src/utils/linalg.f90:73-98,135-168
 * called per particle from Particle%get_cost  src/optimiser/pso/particle.f90:149-162
 * inside the OpenMP swarm loop                src/optimiser/pso/pso.f90:419-452
 */
#ifndef FEREBUS_HIP_H
#define FEREBUS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define FB_OK              0
#define FB_E_ARG          -1   /* invalid argument                          */
#define FB_E_SPEC         -2   /* malformed kernel specification            */
#define FB_E_HIP          -3   /* a HIP runtime call failed                 */
#define FB_E_NOMEM        -4   /* device allocation failed                  */
#define FB_E_NOGPU        -5   /* no usable gfx950 device                   */
#define FB_E_CAPACITY     -6   /* swarm larger than the handle was sized for */

/* ---- kernel composition ------------------------------------------------- *
 * The kernel expression of ferebus.toml (model.kernel, e.g. "k1+k2",
 * "k1*k2", "(k1+k2)*k3", "2*k1") is passed as a postfix program.  Leaves are
 * listed in the order the reference's parser meets them, which is also the
 * order of their parameters in theta (kernel_module.f90:227-245).
 *
 *   leaf type          parameters      hyperparameter transform (reference)
 *   FB_LEAF_RBF        ndim            l_d = exp(theta_d)      rbf_kernel.f90:113-119
 *   FB_LEAF_PERIODIC   ndim            l_d = exp(-2 theta_d)   periodic_kernel.f90:111-118
 *   FB_LEAF_RBF_CYCLIC ndim (= d)      l_d = exp(theta_d)      rbf_cyclic_kernel.f90:109-115
 *   FB_LEAF_LINEAR     ndim            c_d = exp(-theta_d)     linear_kernel.f90:94-100
 *   FB_LEAF_CONSTANT   1               value = theta           constant.f90:86-90
 * Operators: FB_OP_SUM (kernel_module.f90:309-325), FB_OP_PROD (:359-375).
 * With optimise_noise the nugget is the last theta entry (gpr.f90:191-202).
 */
#define FB_LEAF_RBF         1
#define FB_LEAF_PERIODIC    2
#define FB_LEAF_RBF_CYCLIC  3
#define FB_LEAF_LINEAR      4
#define FB_LEAF_CONSTANT    5
#define FB_OP_SUM          16
#define FB_OP_PROD         17

typedef struct fb_kernel_spec {
    int        n_nodes;      /* length of the postfix program                   */
    const int *node_type;    /* [n_nodes] FB_LEAF_* or FB_OP_*                  */
    const int *node_ndim;    /* [n_nodes] active dimensions per leaf (0 for ops) */
    const int *active_dims;  /* leaves' 1-based feature indices, concatenated in
                                program order (sum of node_ndim entries)        */
} fb_kernel_spec;

#define FB_LIKELIHOOD_MARGINAL      0   /* likelihood.f90:161-201 */
#define FB_LIKELIHOOD_CONCENTRATED  1   /* likelihood.f90:203-253 */

typedef struct fb_handle fb_handle;

/* Create a handle for one training set.
 *   x      n*d FP64, column major (x[i + n*f] = feature f of point i), as the
 *          reference stores training inputs (src/csv.f90:37-177).
 *   y      n FP64 targets.
 *   mean   n FP64 prior mean m(x_i) evaluated once on the host
 *          (gpr.f90:168-171); NULL means a zero mean.  Ignored by the
 *          concentrated likelihood, as in the reference (likelihood.f90:220-243).
 *   noise  nugget added to the diagonal (gpr.f90:185-189) unless
 *          optimise_noise != 0, in which case theta's last entry is used.
 *   max_swarm  largest swarm any later call will pass (sizes the K buffers).
 * Replaces GaussianProcessRegressor%init (gpr.f90:124-183) for the device side. */
int fb_create(fb_handle **out, int n, int d, const double *x, const double *y,
              const double *mean, const fb_kernel_spec *spec, double noise,
              int optimise_noise, int likelihood, int max_swarm);

/* Release every device buffer of the handle.  NULL is accepted. */
int fb_destroy(fb_handle *h);

/* Number of theta entries P per particle (GaussianProcessRegressor%nparams,
 * gpr.f90:248-256). */
int fb_nparams(const fb_handle *h, int *p);

/* Bind the handle to a HIP stream (hipStream_t passed as void*; NULL restores
 * the handle's own stream).  All later launches of the handle go there. */
int fb_set_stream(fb_handle *h, void *stream);

/* Swarm log-likelihood, batched path: every particle's build, factorisation and
 * likelihood in one pass over the whole swarm.
 *   theta  P*s FP64 host array, column s = particle s.
 *   ll     s log-likelihoods (not negated; cost = -ll as in gpr.f90:284-288).
 *   info   s LAPACK-style Cholesky INFO values.
 * Replaces s calls of cost_function (gpr.f90:258-289) from pso.f90:419-452. */
int fb_swarm_loglik(fb_handle *h, int s, const double *theta, double *ll, int *info);

/* Same as fb_swarm_loglik with device-resident theta / ll / info, enqueued on
 * the handle's stream without host synchronisation (for callers that keep the
 * swarm on the GPU or capture the iteration in a graph). */
int fb_swarm_loglik_async(fb_handle *h, int s, const double *d_theta,
                          double *d_ll, int *d_info);

/* The same evaluation as fb_swarm_loglik_async with no allocation at all:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Every work buffer (the swarm's K matrices,
 * hyperparameters, partial log-dets, 16x16 inverses, hand-over words,
 * residuals) lives in one 64 MiB array the library holds on each device
 * until the process ends; a swarm that needs more (about 64 MiB / (8 np^2),
 * np = n + right-hand sides rounded up to 128: 30 particles at n = 500, one
 * at n = 2000) is rejected with FB_E_CAPACITY.  The handle's inputs (x, y,
 * mean, kernel) are read, none of its buffers written.  One fixed lock-step
 * schedule (no tuning, no graph), bit-identical to fb_swarm_loglik.
 *   d_theta   P*s FP64 device array (layout of fb_swarm_loglik)
 *   d_ll, d_info  s entries each, device arrays (outputs)
 *   stream    hipStream_t as void*, NULL = the handle's stream; enqueued
 *             without host synchronisation.
 * Calls that overlap on one device (any handles, any streams) share the
 * array and race: serialise them. */
int fb_swarm_loglik_static(fb_handle *h, int s, const double *d_theta, double *d_ll, int *d_info,
                           void *stream);

/* Plain per-particle path: one launch sequence per particle, particles one after
 * the other, exactly what a line-by-line offload of get_cost would do.  Same
 * arguments and results as fb_swarm_loglik.  Kept as the comparison path. */
int fb_particle_loglik(fb_handle *h, int s, const double *theta, double *ll, int *info);

/* Model weights of one parameter vector, for the .model writer
 * (write_gpr gpr.f90:329-398, write_likelihood likelihood.f90:255-264).
 *   alpha  n FP64: K^{-1}(y - m)  (marginal) or K^{-1}(y - mu_hat) (concentrated)
 *   out4   [ll, logdet, mu_hat, sigma2_hat]; mu_hat and sigma2_hat are only
 *          meaningful for the concentrated likelihood.
 *   info   LAPACK INFO of the factorisation. */
int fb_get_alpha(fb_handle *h, const double *theta, double *alpha, double *out4, int *info);

/* The covariance matrix of one parameter vector as the device builds it
 * (stage 0: K(theta) with the nugget on the diagonal, compute_R
 * gpr.f90:185-189 with the kernel builders of kernel_module.f90:208-219,
 * stationary_kernel.f90:101-122, periodic_kernel.f90:181-206,
 * stationary_cyclic_kernel.f90:104-127), or its Cholesky factor as the
 * factorisation leaves it (stage 1: L of dpotrf('L'), linalg.f90:95).
 * A diagnostic for element-by-element checks of the build and the
 * factorisation; no counterpart in the reference's call path.
 *   out   n*n FP64, column major: the lower triangle and diagonal; the
 *         strictly upper part is set to 0.  With stage 1 and info != 0 the
 *         columns from the failing one on are unspecified.
 *   info  0 for stage 0; the factorisation's LAPACK INFO for stage 1. */
int fb_debug_matrix(fb_handle *h, const double *theta, int stage, double *out, int *info);

/* Gradient of the log-likelihood with respect to theta for s particles:
 * d logL / d theta_k = 1/2 tr((alpha alpha^T / s2 - K^{-1}) dK/dtheta_k)
 * (s2 = 1 marginal, sigma2_hat concentrated), evaluated without forming
 * dK/dtheta: K^{-1} = L^{-T} L^{-1} from the factor on the matrix cores, then
 * one fused pass over the lower triangle that recomputes each kernel leaf.
 * The reference parses a gradient flag but never computes one
 * (gpr.f90:272-282); this entry has no reference counterpart and serves
 * L-BFGS polishing of the swarm's best particle (src/optimiser/fb_lbfgs.f90).
 * Allocates a workspace of 2 s np^2 FP64 on the first call with this many
 * particles (np = N + right-hand sides, rounded up to 128).
 *   ll    s FP64, bit-identical to fb_swarm_loglik on the same theta
 *   grad  P*s FP64, column s = gradient of particle s (of ll, not of cost);
 *         NaN where info[s] != 0. */
int fb_swarm_loglik_grad(fb_handle *h, int s, const double *theta, double *ll,
                         double *grad, int *info);

/* ---- several training sets at once (uneven N) ------------------------------
 * The swarms of several handles -- e.g. one GP per atom or property, each
 * with its own N, kernel and likelihood -- are evaluated by one call, on
 * one of three schedules: each handle's own batched path one after another
 * (0), all factorisations in ONE persistent launch (1), or the chained
 * lock-step schedule with every handle's matrices in each of its launches
 * (panel p - (P - P_g) of handle g in launch p, so the handles' last panels
 * share the last launches), the whole evaluation replayed as one hipGraph
 * (2).  The first call for a set of swarm sizes times the schedules and
 * keeps the fastest (FB_MULTI_SCHEDULE=0/1/2, or FB_MULTI_PERSISTENT=1/0,
 * read at fb_multi_create, pins one).  Covariance builds and likelihood
 * reductions stay per handle.  Results are bit-identical to
 * fb_swarm_loglik on each handle.
 * No reference counterpart: the reference trains its GPs one after another
 * (main.f90 loops over atoms and properties; its kernel for one GP is the
 * path of fb_swarm_loglik).  The handles are borrowed (not owned, not to be
 * used concurrently while a multi evaluation runs) and must outlive m; all
 * work runs on the first handle's stream.
 *   s[g]       particles of handle g (0..its max_swarm; 0 skips it)
 *   theta[g]   P_g * s[g] FP64 (layout of fb_swarm_loglik)
 *   ll[g], info[g]  s[g] entries each */
typedef struct fb_multi fb_multi;
int fb_multi_create(fb_multi **m, int n_gps, fb_handle *const *gps);
int fb_multi_swarm_loglik(fb_multi *m, const int *s, const double *const *theta, double *const *ll,
                          int *const *info);
int fb_multi_destroy(fb_multi *m);
/* The schedule (0, 1, 2 as above) fb_multi_swarm_loglik uses for the swarm
 * sizes s[0..n_gps), or -1 before a call with those sizes has chosen one.
 * No reference counterpart (introspection of the schedule choice). */
int fb_multi_schedule(const fb_multi *m, const int *s);

/* Particle swarm on the device: positions, velocities, personal and global
 * bests stay in device memory, and one PSO iteration -- selection of the
 * previous iteration's bests, velocity / position update, covariance build,
 * factorisation, likelihood -- is one captured graph replay.  Replaces the
 * swarm loop of ParticleSwarm (src/optimiser/pso/pso.f90:382-472) around
 * Particle%get_cost (src/optimiser/pso/particle.f90:149-162):
 *   update  v = w v + c1 r1 (pbest - x) + c2 r2 (gbest - x); x = x + v, no
 *           bounds (pso.f90:422-434); r1, r2 come from the caller's RNG
 *           (drawn before the swarm loop, pso.f90:405-409);
 *   pbest   replaced when cost <= pbest cost (pso.f90:440-443);
 *   gbest   particles in index order, replaced when cost <= running gbest
 *           cost (pso.f90:454-459; iter_state.f90:62-82).
 * cost = -logL, +Inf where INFO != 0.  Layouts as fb_swarm_loglik: particle
 * k's P parameters at x[k * P .. k * P + P).  Selection is deferred until
 * the next fb_pso_step (or fb_pso_select), so that the caller can first
 * re-randomise and re-evaluate failed particles with fb_pso_retry, as
 * get_cost does.  The handle must outlive the swarm; calls on one handle
 * are not re-entrant.
 *   fb_pso_create   s particles (<= the handle's max_swarm), coefficients
 *   fb_pso_init     initial swarm x0, v0: one evaluation, no retries;
 *                   pbest = x0, gbest selected (pso.f90:291-327)
 *   fb_pso_step     one iteration with r1, r2 (P * s each, U[0,1)); costs
 *                   and INFO of every particle
 *   fb_pso_retry    new positions / velocities for the k particles idx
 *                   (0-based), re-evaluated together; their costs / INFO
 *   fb_pso_set_coefficients  (w, c1, c2) of the following steps: the swarm's
 *                   inertia weight and learning rates as a parameter updater
 *                   sets them (pso.f90:638-825; the update reads them from
 *                   device memory, so the captured iteration is kept)
 *   fb_pso_select   applies pending selections; gbest cost, position and
 *                   particle index (gidx may be NULL)
 *   fb_pso_set_gbest  replaces gbest (after polishing), pending costs
 *                   selected first
 *   fb_pso_positions  current positions and / or velocities (either NULL) */
typedef struct fb_pso fb_pso;
int fb_pso_create(fb_pso **p, fb_handle *h, int s, double w, double c1, double c2);
int fb_pso_destroy(fb_pso *p);
int fb_pso_init(fb_pso *p, const double *x0, const double *v0, double *cost, int *info);
int fb_pso_step(fb_pso *p, const double *r1, const double *r2, double *cost, int *info);
int fb_pso_retry(fb_pso *p, int k, const int *idx, const double *x, const double *v, double *cost, int *info);
int fb_pso_set_coefficients(fb_pso *p, double w, double c1, double c2);
int fb_pso_select(fb_pso *p, double *gcost, double *gbest, int *gidx);
int fb_pso_set_gbest(fb_pso *p, const double *gbest, double gcost);
int fb_pso_positions(fb_pso *p, double *x, double *v);

/* One PSO iteration of several swarms at once: ps[g] (or NULL to skip) is a
 * swarm created on handle g of m.  Each swarm's pending selection and its
 * velocity / position update (r1[g], r2[g]: P_g * s_g each; the swarm's
 * coefficients) as fb_pso_step, then one evaluation of all of them on the
 * schedule of fb_multi_swarm_loglik for these swarm sizes (chosen on the
 * first call, on the swarms' current positions; where that is the
 * persistent launch, the merged schedule is used), the whole iteration one
 * captured graph replay.  cost[g], info[g] (either array or entry may be
 * NULL) receive s_g entries each.  Each swarm ends in the state fb_pso_step
 * would leave, bit for bit; init, retries, selection and coefficients stay
 * per swarm (fb_pso_*).
 * No reference counterpart: the reference trains its GPs one after another
 * (app/main.f90:78-130), each with the swarm loop of pso.f90:382-472. */
int fb_multi_pso_step(fb_multi *m, fb_pso *const *ps, const double *const *r1, const double *const *r2,
                      double *const *cost, int *const *info);

/* Kernel timing: when enabled (1), each launch of the handle is bracketed by
 * HIP events and the per-kernel totals accumulate until read; enable = 2
 * brackets only the phases (kinds 0, 4, 5, 6, 8: no events between the
 * factorisation's launches).
 * kind: 0 covariance build; lock-step factorisation, per 128-column panel:
 *       1 trailing update + first diagonal block, 2 panel solve + second
 *       diagonal block, 3 second block's update + solve; 4 likelihood
 *       reduction, 5 whole swarm evaluation, 6 persistent (single-launch)
 *       factorisation; 7: *launches = the swarm evaluations the autotuner
 *       has timed on this handle so far (counted whether or not profiling
 *       is enabled, not reset; *total_ms = 0).  fb_pso_retry never tunes:
 *       its swarms of re-randomised particles take the nearest tuned size's
 *       path; 8 lock-step factorisation span, from the end of the build to
 *       the start of the likelihood (with stream groups, the launches of
 *       different groups overlap inside it, and kinds 1-3 add up durations
 *       on several streams); 9: *launches = the persistent evaluations that
 *       found an XCD queue no work-group served (the device does not
 *       dispatch the way the persistent schedule assumes) and were re-run on
 *       the lock-step path, which the handle then keeps (a counter, not
 *       reset; *total_ms = 0). */
int fb_profile_enable(fb_handle *h, int enable);
int fb_profile_read(fb_handle *h, int kind, double *total_ms, long long *launches);

/* The factorisation path the batched evaluation uses for a swarm of s on
 * this handle (no reference counterpart: a diagnostic of the autotuner).
 * The first fb_swarm_loglik / fb_swarm_loglik_grad / fb_pso call for a swarm
 * size times the candidates on its inputs and keeps the fastest; all of
 * them compute the same bits.  path: -1 not chosen yet, 0 persistent single
 * launch; lock-step launches with 1: 32-deep GEMM chunks and 64x128 update
 * tiles, 2: 16-deep chunks, 64x128 tiles, 3: 16-deep, 32x128, 4: 16-deep,
 * 32x64, 5: each panel's fastest of those three tile shapes and the
 * variants below (timed panel by panel).  With 16-deep chunks and swarms up to 160 the
 * lock-step schedule is chained: two launches per 128-column panel, the
 * panel's diagonal factorisation handed over between work-groups inside the
 * update launch (path 5 then also times each shape with two or three such
 * work-groups per particle).  Environment, read at fb_create:
 * FB_FACTOR_PATH=0..4 pins a path; FB_FACTOR_PATH=5 with
 * FB_PANEL_TILES=t0,t1,... (one code per 128-column panel: 0..2 the shapes
 * of paths 2..4, 4..6 the same with two work-groups per particle in the
 * chained update launch, 8 and 12 the 64x128 shape with three or two such
 * work-groups and the row solve fused into the update launch, 16 and 20 the
 * 64x128 shape with three or two such work-groups in 8-deep chunks, three
 * work-groups per CU instead of two; + 32 on a 64x128 code without 8: the
 * update launch also solves the previous panel's rows; 64 and 68 (+ 32) the
 * 64x128 shape in 8-deep chunks at four work-groups per CU) pins a
 * per-panel mix; FB_CHAIN=0 gives three launches per panel instead;
 * FB_TAIL=n sets how many trailing panels the chained schedule pre-updates
 * in the launch before them (default 3, 0 off); FB_PERSISTENT=0/1 pins path
 * 1/0;
 * FB_AUTOTUNE=0 keeps path 0 without timing; FB_TUNE_CACHE=<file> keeps the
 * choices across processes (JSON records keyed on device, padded order and
 * swarm size, with the per-panel shapes and the stream groups). */
int fb_factor_path(const fb_handle *h, int s, int *path);

/* Stream groups of the lock-step paths for a swarm of s: the swarm is built
 * once, its particles' factorisations run as `groups` groups on concurrent
 * streams (a launch of one group that leaves the chip partly idle in its
 * last round overlaps with the others'), then one likelihood launch.  The
 * autotuner times 1 .. 4 groups with the path it chose; FB_STREAMS=n (read
 * at fb_create) pins n; 1 before the first call for s.  The bits do not
 * depend on it.  No reference counterpart (introspection of the schedule). */
int fb_factor_groups(const fb_handle *h, int s, int *groups);

/* Library-wide information: device name into buf (len bytes), 0 on success. */
int fb_device_name(char *buf, int len);

/* Human-readable message for the last error on this thread. */
const char *fb_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* FEREBUS_HIP_H */
