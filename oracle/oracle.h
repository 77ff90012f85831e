/*
 * oracle.h -- CPU FP64 restatement of FEREBUS-v7's swarm log-likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is
 * compared against (tests/, __graft_entry__.smoke()) and the CPU baseline that
 * bench.py times beside it.  The product (libferebus_hip.so) never links,
 * loads or calls it.
 *
 * Every routine cites the reference file:line it restates.  Arithmetic is plain
 * FP64 in the reference's order of operations where the reference fixes one.
 */
#ifndef FEREBUS_ORACLE_H
#define FEREBUS_ORACLE_H

#include "../include/ferebus_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Number of theta entries (gpr.f90:248-256 + kernel nparams). -1 on a bad spec. */
int or_nparams(const fb_kernel_spec *spec, int optimise_noise);

/* Kernel matrix without nugget: K[i + ldk*j] for i >= j (lower triangle and
 * diagonal), evaluated leaf by leaf as the reference's R builders do and
 * combined by the postfix sum / product program.  Returns 0 or -1. */
int or_kernel_matrix(int n, int d, const double *x, const fb_kernel_spec *spec,
                     const double *theta, double *k, int ldk);

/* Unblocked lower Cholesky with LAPACK dpotf2 INFO semantics.  Returns INFO. */
int or_potrf(int n, double *a, int lda);

/* Solve L L^T x = b in place (dpotrs with UPLO='L', one right-hand side). */
void or_potrs(int n, const double *l, int lda, double *b);

/* 2 * sum(log L_ii), arith.f90:72-93. */
double or_logdet(int n, const double *l, int lda);

/* Full log-likelihood of one theta.  out4 (nullable) receives
 * [ll, logdet, mu_hat, sigma2_hat]; alpha (nullable) receives the weights.
 * Returns the Cholesky INFO. */
int or_loglik(int n, int d, const double *x, const double *y, const double *mean,
              const fb_kernel_spec *spec, double noise, int optimise_noise,
              int likelihood, const double *theta, double *alpha, double *out4);

/* The same for s particles (theta column s = particle s), OpenMP over
 * particles as pso.f90:419-452 does.  nthreads <= 0 keeps the runtime default. */
int or_swarm_loglik(int n, int d, const double *x, const double *y, const double *mean,
                    const fb_kernel_spec *spec, double noise, int optimise_noise,
                    int likelihood, int s, const double *theta, double *ll, int *info,
                    int nthreads);

/* Dense reference gradient d ll / d theta (length P) by
 * 1/2 sum_ij W_ij dK_ij/dtheta with W = alpha alpha^T - K^{-1} (marginal) or
 * alpha alpha^T / sigma2 - K^{-1} (concentrated), dK formed explicitly.
 * Returns the Cholesky INFO; ll_out (nullable) receives ll. */
int or_loglik_grad(int n, int d, const double *x, const double *y, const double *mean,
                   const fb_kernel_spec *spec, double noise, int optimise_noise,
                   int likelihood, const double *theta, double *ll_out, double *grad);

/* The kernel matrix of or_kernel_matrix evaluated in long double (expl, sinl,
 * no FP64 rounding of the exponents), returned as the nearest FP64 value
 * k_hi and the remainder k_lo (nullable): the truth each FP64 element is
 * measured against, in ulps of k_hi. */
int or_kernel_matrix_ext(int n, int d, const double *x, const fb_kernel_spec *spec,
                         const double *theta, double *k_hi, double *k_lo, int ldk);

/* Extended-precision truth for one given FP64 matrix: the same likelihood as
 * or_loglik (likelihood.f90:161-253 after dpotrf, linalg.f90:95) with the
 * Cholesky factorisation, log-det, solves and reductions in x87 long double
 * (64-bit significand, unit roundoff 5.4e-20), so that the likelihood of an
 * ill-conditioned K (a 1e-10 nugget) is known far below FP64 rounding.
 *   k      n x n FP64 (leading dimension ldk), lower triangle and diagonal
 *          read, nugget already on the diagonal
 *   out4   [ll, logdet, mu_hat, sigma2_hat] rounded to FP64
 *   lfac   (nullable) n x n: the long double factor rounded to FP64
 * Returns INFO (0, or the first non-positive pivot), -1 on allocation
 * failure. */
int or_loglik_ext(int n, const double *k, int ldk, const double *y, const double *mean,
                  int likelihood, double *out4, double *lfac);

/* Mean functions evaluated once per training set (src/mean/).
 * kind: 0 zero, 1 constant, 2 linear, 3 quadratic.  For the least-squares
 * means beta, xmin (d entries each) and ymin (1) are returned when non-NULL.
 * Returns 0, or the dposv INFO of the normal equations. */
int or_mean(int kind, int n, int d, const double *x, const double *y, double *m,
            double *beta, double *xmin, double *ymin);

/* ---- CPU baseline: the reference's own host algorithm -------------------- *
 * Precomputed per-feature pair tensors (distance_matrix.f90:64-142), then per
 * particle a dgemv over the active features + exp (stationary_kernel.f90:101-122,
 * periodic_kernel.f90:181-206), nugget (gpr.f90:185-189), LAPACK dpotrf /
 * dpotrs (linalg.f90:73-168) and the likelihood (likelihood.f90:161-253),
 * OpenMP over particles with one LAPACK thread each.
 * LAPACK / BLAS come from a host library opened with dlopen (symbols
 * <prefix>dpotrf_, <prefix>dpotrs_, <prefix>dgemv_); or_baseline_lapack()
 * reports 1 when they were found, 0 when the oracle's own routines are used. */
typedef struct or_baseline or_baseline;
or_baseline *or_baseline_create(int n, int d, const double *x, const double *y,
                                const double *mean, const fb_kernel_spec *spec,
                                double noise, int optimise_noise, int likelihood,
                                const char *lapack_path, const char *symbol_prefix);
int  or_baseline_lapack(const or_baseline *b);
int  or_baseline_swarm(or_baseline *b, int s, const double *theta, double *ll, int *info,
                       int nthreads);
void or_baseline_destroy(or_baseline *b);

#ifdef __cplusplus
}
#endif
#endif
