/*
 * oracle.c -- CPU FP64 restatement of FEREBUS-v7's swarm log-likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU path and the CPU baseline
 * of bench.py.  The product library never links or calls this code.
 *
 * Citations are to popelier-group/FEREBUS-v7 (src/...).
 */
#include "oracle.h"

#include <dlfcn.h>
#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants.f90:30-32: PI = 4*atan(1), TWOPI = 2*PI, LOG_TWOPI = log(TWOPI) */
#define OR_PI 3.14159265358979323846
#define OR_TWOPI (2.0 * OR_PI)

#define OR_MAX_NODES 64

typedef struct {
    int type;     /* FB_LEAF_* or FB_OP_* */
    int ndim;     /* active dimensions (leaves) */
    int dim_off;  /* offset into spec->active_dims */
    int par_off;  /* offset of the leaf's parameters in theta */
    int npar;     /* parameters of the leaf */
} or_node;

/* Walks the postfix program, assigning parameter offsets in leaf order, which
 * is the order kernel_module.f90:227-245 stacks parameters (k1 then k2). */
static int parse_spec(const fb_kernel_spec *sp, int d, or_node *nodes, int *npar_out)
{
    if (!sp || sp->n_nodes < 1 || sp->n_nodes > OR_MAX_NODES || !sp->node_type)
        return -1;
    int depth = 0, par = 0, doff = 0;
    for (int k = 0; k < sp->n_nodes; ++k) {
        int t = sp->node_type[k];
        int nd = sp->node_ndim ? sp->node_ndim[k] : 0;
        or_node *o = &nodes[k];
        o->type = t; o->ndim = 0; o->dim_off = doff; o->par_off = par; o->npar = 0;
        if (t == FB_OP_SUM || t == FB_OP_PROD) {
            if (depth < 2) return -1;
            depth -= 1;
            continue;
        }
        switch (t) {
        case FB_LEAF_RBF: case FB_LEAF_PERIODIC: case FB_LEAF_LINEAR:
            if (nd < 1) return -1;
            o->ndim = nd; o->npar = nd;
            break;
        case FB_LEAF_RBF_CYCLIC:
            /* stationary_cyclic_kernel.f90:104-127 sums over every feature, so
             * the leaf must cover all d features in order. */
            if (d > 0 && nd != d) return -1;
            o->ndim = nd; o->npar = nd;
            break;
        case FB_LEAF_CONSTANT:
            if (nd != 0) return -1;
            o->npar = 1;
            break;
        default:
            return -1;
        }
        if (o->ndim) {
            if (!sp->active_dims) return -1;
            for (int q = 0; q < o->ndim; ++q) {
                int f = sp->active_dims[doff + q];
                if (f < 1 || (d > 0 && f > d)) return -1;
                if (t == FB_LEAF_RBF_CYCLIC && f != q + 1) return -1;
            }
        }
        doff += o->ndim;
        par += o->npar;
        depth += 1;
    }
    if (depth != 1) return -1;
    *npar_out = par;
    return 0;
}

int or_nparams(const fb_kernel_spec *spec, int optimise_noise)
{
    or_node nodes[OR_MAX_NODES];
    int p;
    if (parse_spec(spec, 0, nodes, &p)) return -1;
    return p + (optimise_noise ? 1 : 0);
}

/* Cyclic wrap of distance_matrix.f90:119-120:
 *   diff = sign(1, diff) * (mod(|diff| + PI, TWOPI) - PI) */
static double cyclic_wrap(double diff)
{
    return copysign(1.0, diff) * (fmod(fabs(diff) + OR_PI, OR_TWOPI) - OR_PI);
}

/* One leaf's value at the pair (i, j). */
static double leaf_value(const or_node *o, const int *dims, int n, int d, const double *x,
                         const double *theta, int i, int j)
{
    const double *th = theta + o->par_off;
    const int *ad = dims + o->dim_off;
    double t = 0.0;
    switch (o->type) {
    case FB_LEAF_RBF:
        /* rbf_kernel.f90:113-119 (l = exp(theta)); distance_matrix.f90:86-87
         * (D = diff*diff); stationary_kernel.f90:110-121 (dgemv -0.5, exp). */
        for (int q = 0; q < o->ndim; ++q) {
            int f = ad[q] - 1;
            double diff = x[i + (size_t)n * f] - x[j + (size_t)n * f];
            t += exp(th[q]) * (diff * diff);
        }
        return exp(-0.5 * t);
    case FB_LEAF_PERIODIC:
        /* periodic_kernel.f90:111-118 (l = (1/exp(theta))**2);
         * distance_matrix.f90:140-141 (sin(sqrt(D)/2)**2);
         * periodic_kernel.f90:181-206 (dgemv -2, exp). */
        for (int q = 0; q < o->ndim; ++q) {
            int f = ad[q] - 1;
            double diff = x[i + (size_t)n * f] - x[j + (size_t)n * f];
            double dd = diff * diff;
            double s = sin(sqrt(dd) / 2.0);
            double l = 1.0 / exp(th[q]);
            l = l * l;
            t += l * (s * s);
        }
        return exp(-2.0 * t);
    case FB_LEAF_RBF_CYCLIC:
        /* distance_matrix.f90:113-125 (wrap on features 3, 6, 9, ...);
         * stationary_cyclic_kernel.f90:113-125 (sum over features, exp(-R/2)). */
        for (int f = 0; f < d; ++f) {
            double diff = x[i + (size_t)n * f] - x[j + (size_t)n * f];
            if ((f + 1) % 3 == 0) diff = cyclic_wrap(diff);
            t += exp(th[f]) * (diff * diff);
        }
        return exp(-0.5 * t);
    case FB_LEAF_LINEAR:
        /* linear_kernel.f90:70-76 with order = 0 (linear_kernel.f90:33):
         * sum((xi*xj)**0 / c) = sum(1/c), c = exp(-theta) (:94-100). */
        for (int q = 0; q < o->ndim; ++q) t += 1.0 / exp(-th[q]);
        return t;
    case FB_LEAF_CONSTANT:
        /* constant.f90:62-68, 86-90: the value is the raw parameter. */
        return th[0];
    default:
        return NAN;
    }
}

/* Postfix evaluation: R_sum (kernel_module.f90:309-325) adds, R_prod
 * (:359-375) multiplies the children's matrices elementwise. */
static double pair_value(const or_node *nodes, int nn, const int *dims, int n, int d,
                         const double *x, const double *theta, int i, int j)
{
    double st[OR_MAX_NODES];
    int sp = 0;
    for (int k = 0; k < nn; ++k) {
        const or_node *o = &nodes[k];
        if (o->type == FB_OP_SUM) { st[sp - 2] = st[sp - 2] + st[sp - 1]; --sp; }
        else if (o->type == FB_OP_PROD) { st[sp - 2] = st[sp - 2] * st[sp - 1]; --sp; }
        else st[sp++] = leaf_value(o, dims, n, d, x, theta, i, j);
    }
    return st[0];
}

int or_kernel_matrix(int n, int d, const double *x, const fb_kernel_spec *spec,
                     const double *theta, double *k, int ldk)
{
    or_node nodes[OR_MAX_NODES];
    int p;
    if (parse_spec(spec, d, nodes, &p)) return -1;
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i)
            k[i + (size_t)ldk * j] =
                pair_value(nodes, spec->n_nodes, spec->active_dims, n, d, x, theta, i, j);
    return 0;
}

/* LAPACK dpotf2, UPLO = 'L' (the routine behind linalg.f90:95 for one block):
 * column j: ajj = a(j,j) - dot(a(j,1:j-1), a(j,1:j-1)); stop with INFO = j
 * when ajj <= 0 or NaN; else a(j,j) = sqrt(ajj), the column below is updated
 * by a gemv and scaled by 1/ajj. */
int or_potrf(int n, double *a, int lda)
{
    for (int j = 0; j < n; ++j) {
        double ajj = a[j + (size_t)lda * j];
        for (int k = 0; k < j; ++k) {
            double v = a[j + (size_t)lda * k];
            ajj -= v * v;
        }
        if (ajj <= 0.0 || isnan(ajj)) {
            a[j + (size_t)lda * j] = ajj;
            return j + 1;
        }
        ajj = sqrt(ajj);
        a[j + (size_t)lda * j] = ajj;
        double *col = a + (size_t)lda * j;
        for (int k = 0; k < j; ++k) {
            double ajk = a[j + (size_t)lda * k];
            const double *ck = a + (size_t)lda * k;
            for (int i = j + 1; i < n; ++i) col[i] -= ck[i] * ajk;
        }
        double r = 1.0 / ajj;
        for (int i = j + 1; i < n; ++i) col[i] *= r;
    }
    return 0;
}

/* dpotrs('L'): L z = b, then L^T x = z (linalg.f90:135-168). */
void or_potrs(int n, const double *l, int lda, double *b)
{
    for (int j = 0; j < n; ++j) {
        b[j] /= l[j + (size_t)lda * j];
        double bj = b[j];
        const double *c = l + (size_t)lda * j;
        for (int i = j + 1; i < n; ++i) b[i] -= c[i] * bj;
    }
    for (int j = n - 1; j >= 0; --j) {
        const double *c = l + (size_t)lda * j;
        double t = b[j];
        for (int i = j + 1; i < n; ++i) t -= c[i] * b[i];
        b[j] = t / c[j];
    }
}

/* arith.f90:72-93 */
double or_logdet(int n, const double *l, int lda)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += log(l[i + (size_t)lda * i]);
    return 2.0 * s;
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* The likelihood of an already factored K (likelihood.f90:161-253). */
static void likelihood_from_factor(int n, const double *l, const double *y, const double *mean,
                                   int likelihood, double logdet, double *alpha_out,
                                   double *out4)
{
    double *r = (double *)malloc(sizeof(double) * n);
    double *a = (double *)malloc(sizeof(double) * n);
    double ll, mu = 0.0, s2 = 0.0;
    if (likelihood == FB_LIKELIHOOD_CONCENTRATED) {
        /* likelihood.f90:220-243: mu = 1'K^-1 y / 1'K^-1 1, r = y - mu,
         * sigma2 = r'K^-1 r / N, ll = -N/2 log sigma2 - logdet/2. */
        double *ry = (double *)malloc(sizeof(double) * n);
        double *r1 = (double *)malloc(sizeof(double) * n);
        memcpy(ry, y, sizeof(double) * n);
        or_potrs(n, l, n, ry);
        for (int i = 0; i < n; ++i) r1[i] = 1.0;
        or_potrs(n, l, n, r1);
        double sy = 0.0, s1 = 0.0;
        for (int i = 0; i < n; ++i) { sy += ry[i]; s1 += r1[i]; }
        mu = sy / s1;
        for (int i = 0; i < n; ++i) r[i] = y[i] - mu;
        memcpy(a, r, sizeof(double) * n);
        or_potrs(n, l, n, a);
        s2 = dot(n, r, a) / (double)n;
        ll = -0.5 * (double)n * log(s2) - 0.5 * logdet;
        free(ry); free(r1);
    } else {
        /* likelihood.f90:178-191: r = y - m, alpha = K^-1 r,
         * ll = -r'alpha/2 - logdet/2 - N log(2 pi)/2. */
        for (int i = 0; i < n; ++i) r[i] = y[i] - (mean ? mean[i] : 0.0);
        memcpy(a, r, sizeof(double) * n);
        or_potrs(n, l, n, a);
        ll = -0.5 * dot(n, r, a) - 0.5 * logdet - 0.5 * (double)n * log(OR_TWOPI);
    }
    if (alpha_out) memcpy(alpha_out, a, sizeof(double) * n);
    if (out4) { out4[0] = ll; out4[1] = logdet; out4[2] = mu; out4[3] = s2; }
    free(r); free(a);
}

static double nugget_of(const double *theta, int npar_kernel, double noise, int optimise_noise)
{
    /* gpr.f90:191-202: nugget = params(nkernel+1) when optimise_noise. */
    return optimise_noise ? theta[npar_kernel] : noise;
}

int or_loglik(int n, int d, const double *x, const double *y, const double *mean,
              const fb_kernel_spec *spec, double noise, int optimise_noise,
              int likelihood, const double *theta, double *alpha, double *out4)
{
    or_node nodes[OR_MAX_NODES];
    int pk;
    if (parse_spec(spec, d, nodes, &pk)) return -1;
    double *k = (double *)malloc(sizeof(double) * (size_t)n * n);
    if (!k) return -1;
    or_kernel_matrix(n, d, x, spec, theta, k, n);
    /* gpr.f90:185-189: K = K + nugget * I */
    double nug = nugget_of(theta, pk, noise, optimise_noise);
    for (int i = 0; i < n; ++i) k[i + (size_t)n * i] += nug;
    int info = or_potrf(n, k, n);
    if (info == 0)
        likelihood_from_factor(n, k, y, mean, likelihood, or_logdet(n, k, n), alpha, out4);
    free(k);
    return info;
}

int or_swarm_loglik(int n, int d, const double *x, const double *y, const double *mean,
                    const fb_kernel_spec *spec, double noise, int optimise_noise,
                    int likelihood, int s, const double *theta, double *ll, int *info,
                    int nthreads)
{
    int p = or_nparams(spec, optimise_noise);
    if (p < 0) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    (void)nthreads;
    #pragma omp parallel for schedule(dynamic)
    for (int q = 0; q < s; ++q) {
        double o4[4];
        info[q] = or_loglik(n, d, x, y, mean, spec, noise, optimise_noise, likelihood,
                            theta + (size_t)p * q, NULL, o4);
        ll[q] = info[q] == 0 ? o4[0] : NAN;
    }
    return 0;
}

/* ---- gradient (dense, for checking) -------------------------------------- */

/* Value and parameter derivatives of one leaf at (i, j).  g has P entries and
 * is only written at the leaf's own parameter positions. */
static double leaf_grad(const or_node *o, const int *dims, int n, int d, const double *x,
                        const double *theta, int i, int j, double *g)
{
    const double *th = theta + o->par_off;
    const int *ad = dims + o->dim_off;
    double v, t = 0.0;
    switch (o->type) {
    case FB_LEAF_RBF:
        for (int q = 0; q < o->ndim; ++q) {
            int f = ad[q] - 1;
            double diff = x[i + (size_t)n * f] - x[j + (size_t)n * f];
            t += exp(th[q]) * diff * diff;
        }
        v = exp(-0.5 * t);
        /* d/dtheta_q exp(-1/2 sum e^theta D) = -1/2 e^theta_q D_q k */
        for (int q = 0; q < o->ndim; ++q) {
            int f = ad[q] - 1;
            double diff = x[i + (size_t)n * f] - x[j + (size_t)n * f];
            g[o->par_off + q] = -0.5 * exp(th[q]) * diff * diff * v;
        }
        return v;
    case FB_LEAF_PERIODIC:
        for (int q = 0; q < o->ndim; ++q) {
            int f = ad[q] - 1;
            double diff = x[i + (size_t)n * f] - x[j + (size_t)n * f];
            double s = sin(fabs(diff) / 2.0);
            t += exp(-2.0 * th[q]) * s * s;
        }
        v = exp(-2.0 * t);
        /* l = e^{-2 theta}: d/dtheta_q exp(-2 sum l P) = 4 l_q P_q k */
        for (int q = 0; q < o->ndim; ++q) {
            int f = ad[q] - 1;
            double diff = x[i + (size_t)n * f] - x[j + (size_t)n * f];
            double s = sin(fabs(diff) / 2.0);
            g[o->par_off + q] = 4.0 * exp(-2.0 * th[q]) * s * s * v;
        }
        return v;
    case FB_LEAF_RBF_CYCLIC:
        for (int f = 0; f < d; ++f) {
            double diff = x[i + (size_t)n * f] - x[j + (size_t)n * f];
            if ((f + 1) % 3 == 0) diff = cyclic_wrap(diff);
            t += exp(th[f]) * diff * diff;
        }
        v = exp(-0.5 * t);
        for (int f = 0; f < d; ++f) {
            double diff = x[i + (size_t)n * f] - x[j + (size_t)n * f];
            if ((f + 1) % 3 == 0) diff = cyclic_wrap(diff);
            g[o->par_off + f] = -0.5 * exp(th[f]) * diff * diff * v;
        }
        return v;
    case FB_LEAF_LINEAR:
        for (int q = 0; q < o->ndim; ++q) { t += exp(th[q]); g[o->par_off + q] = exp(th[q]); }
        return t;
    case FB_LEAF_CONSTANT:
        g[o->par_off] = 1.0;
        return th[0];
    default:
        return NAN;
    }
}

int or_loglik_grad(int n, int d, const double *x, const double *y, const double *mean,
                   const fb_kernel_spec *spec, double noise, int optimise_noise,
                   int likelihood, const double *theta, double *ll_out, double *grad)
{
    or_node nodes[OR_MAX_NODES];
    int pk;
    if (parse_spec(spec, d, nodes, &pk)) return -1;
    int p = pk + (optimise_noise ? 1 : 0);
    size_t nn = (size_t)n * n;
    double *k = (double *)malloc(sizeof(double) * nn);
    double *kinv = (double *)calloc(nn, sizeof(double));
    double *alpha = (double *)malloc(sizeof(double) * n);
    or_kernel_matrix(n, d, x, spec, theta, k, n);
    double nug = nugget_of(theta, pk, noise, optimise_noise);
    for (int i = 0; i < n; ++i) k[i + (size_t)n * i] += nug;
    int info = or_potrf(n, k, n);
    if (info) { free(k); free(kinv); free(alpha); return info; }
    double o4[4];
    likelihood_from_factor(n, k, y, mean, likelihood, or_logdet(n, k, n), alpha, o4);
    if (ll_out) *ll_out = o4[0];
    for (int c = 0; c < n; ++c) {
        double *col = kinv + (size_t)n * c;
        col[c] = 1.0;
        or_potrs(n, k, n, col);
    }
    double scale = likelihood == FB_LIKELIHOOD_CONCENTRATED ? 1.0 / o4[3] : 1.0;
    for (int q = 0; q < p; ++q) grad[q] = 0.0;
    int nnode = spec->n_nodes;
    double *gst = (double *)malloc(sizeof(double) * (size_t)OR_MAX_NODES * (p + 1));
    double vst[OR_MAX_NODES];
    for (int j = 0; j < n; ++j) {
        for (int i = j; i < n; ++i) {
            int sp = 0;
            for (int t = 0; t < nnode; ++t) {
                const or_node *o = &nodes[t];
                if (o->type == FB_OP_SUM || o->type == FB_OP_PROD) {
                    double *g1 = gst + (size_t)(sp - 2) * p, *g2 = gst + (size_t)(sp - 1) * p;
                    double v1 = vst[sp - 2], v2 = vst[sp - 1];
                    if (o->type == FB_OP_SUM) {
                        for (int q = 0; q < p; ++q) g1[q] += g2[q];
                        vst[sp - 2] = v1 + v2;
                    } else {
                        for (int q = 0; q < p; ++q) g1[q] = v2 * g1[q] + v1 * g2[q];
                        vst[sp - 2] = v1 * v2;
                    }
                    --sp;
                } else {
                    double *g = gst + (size_t)sp * p;
                    for (int q = 0; q < p; ++q) g[q] = 0.0;
                    vst[sp] = leaf_grad(o, spec->active_dims, n, d, x, theta, i, j, g);
                    ++sp;
                }
            }
            double *g = gst;
            if (optimise_noise) g[pk] = (i == j) ? 1.0 : 0.0;
            double w = scale * alpha[i] * alpha[j] - kinv[i + (size_t)n * j];
            double mult = (i == j) ? 0.5 : 1.0;  /* 1/2 * (2 for the symmetric pair) */
            for (int q = 0; q < p; ++q) grad[q] += mult * w * g[q];
        }
    }
    free(gst); free(k); free(kinv); free(alpha);
    return 0;
}

/* ---- extended precision (x87 80-bit long double) ------------------------- */

/* One leaf's value at (i, j) in long double, the formulas of leaf_value
 * (rbf_kernel.f90:113-119, periodic_kernel.f90:111-118,
 * distance_matrix.f90:86-87, 113-125, 140-141, stationary_kernel.f90:101-122,
 * periodic_kernel.f90:181-206, stationary_cyclic_kernel.f90:104-127,
 * linear_kernel.f90:70-100, constant.f90:62-96) without FP64 rounding. */
static long double leaf_value_ext(const or_node *o, const int *dims, int n, int d, const double *x,
                                  const double *theta, int i, int j)
{
    const double *th = theta + o->par_off;
    const int *ad = dims + o->dim_off;
    const long double pi = 3.141592653589793238462643383279503L;
    long double t = 0.0L;
    switch (o->type) {
    case FB_LEAF_RBF:
        for (int q = 0; q < o->ndim; ++q) {
            int f = ad[q] - 1;
            long double diff = (long double)x[i + (size_t)n * f] - (long double)x[j + (size_t)n * f];
            t += expl((long double)th[q]) * diff * diff;
        }
        return expl(-0.5L * t);
    case FB_LEAF_PERIODIC:
        for (int q = 0; q < o->ndim; ++q) {
            int f = ad[q] - 1;
            long double diff = (long double)x[i + (size_t)n * f] - (long double)x[j + (size_t)n * f];
            long double s = sinl(fabsl(diff) / 2.0L);
            long double l = 1.0L / expl((long double)th[q]);
            t += l * l * s * s;
        }
        return expl(-2.0L * t);
    case FB_LEAF_RBF_CYCLIC:
        for (int f = 0; f < d; ++f) {
            long double diff = (long double)x[i + (size_t)n * f] - (long double)x[j + (size_t)n * f];
            if ((f + 1) % 3 == 0) diff = fmodl(fabsl(diff) + pi, 2.0L * pi) - pi;
            t += expl((long double)th[f]) * diff * diff;
        }
        return expl(-0.5L * t);
    case FB_LEAF_LINEAR:
        for (int q = 0; q < o->ndim; ++q) t += 1.0L / expl(-(long double)th[q]);
        return t;
    case FB_LEAF_CONSTANT:
        return (long double)th[0];
    default:
        return NAN;
    }
}

int or_kernel_matrix_ext(int n, int d, const double *x, const fb_kernel_spec *spec,
                         const double *theta, double *k_hi, double *k_lo, int ldk)
{
    or_node nodes[OR_MAX_NODES];
    int p;
    if (parse_spec(spec, d, nodes, &p)) return -1;
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) {
            long double st[OR_MAX_NODES];
            int sp = 0;
            for (int q = 0; q < spec->n_nodes; ++q) {
                const or_node *o = &nodes[q];
                if (o->type == FB_OP_SUM) { st[sp - 2] += st[sp - 1]; --sp; }
                else if (o->type == FB_OP_PROD) { st[sp - 2] *= st[sp - 1]; --sp; }
                else st[sp++] = leaf_value_ext(o, spec->active_dims, n, d, x, theta, i, j);
            }
            const double hi = (double)st[0];
            k_hi[i + (size_t)ldk * j] = hi;
            if (k_lo) k_lo[i + (size_t)ldk * j] = (double)(st[0] - (long double)hi);
        }
    return 0;
}

/* Cholesky of the given FP64 matrix in long double (64-bit significand),
 * left-looking: each entry starts from a(i,j) and the products of the
 * earlier columns are subtracted one k at a time, the order in which the
 * remainder shrinks toward the pivot (dpotf2's dot-product form,
 * linalg.f90:95).  Returns INFO (LAPACK semantics). */
static int potrf_ext(int n, long double *a)
{
    for (int j = 0; j < n; ++j) {
        long double *cj = a + (size_t)n * j;
        for (int k = 0; k < j; ++k) {
            const long double *ck = a + (size_t)n * k;
            const long double ajk = ck[j];
            for (int i = j; i < n; ++i) cj[i] -= ck[i] * ajk;
        }
        long double ajj = cj[j];
        if (!(ajj > 0.0L)) return j + 1;
        ajj = sqrtl(ajj);
        cj[j] = ajj;
        for (int i = j + 1; i < n; ++i) cj[i] /= ajj;
    }
    return 0;
}

/* L L^T x = b in place, long double (dpotrs('L'), linalg.f90:135-168). */
static void potrs_ext(int n, const long double *l, long double *b)
{
    for (int j = 0; j < n; ++j) {
        const long double *c = l + (size_t)n * j;
        b[j] /= c[j];
        for (int i = j + 1; i < n; ++i) b[i] -= c[i] * b[j];
    }
    for (int j = n - 1; j >= 0; --j) {
        const long double *c = l + (size_t)n * j;
        long double t = b[j];
        for (int i = j + 1; i < n; ++i) t -= c[i] * b[i];
        b[j] = t / c[j];
    }
}

int or_loglik_ext(int n, const double *k, int ldk, const double *y, const double *mean,
                  int likelihood, double *out4, double *lfac)
{
    size_t nn = (size_t)n * n;
    long double *a = (long double *)malloc(sizeof(long double) * nn);
    long double *r = (long double *)malloc(sizeof(long double) * n);
    long double *w = (long double *)malloc(sizeof(long double) * n);
    long double *v = (long double *)malloc(sizeof(long double) * n);
    if (!a || !r || !w || !v) { free(a); free(r); free(w); free(v); return -1; }
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) a[i + (size_t)n * j] = (long double)k[i + (size_t)ldk * j];
    int info = potrf_ext(n, a);
    if (lfac) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) lfac[i + (size_t)n * j] = i >= j ? (double)a[i + (size_t)n * j] : 0.0;
    }
    if (info == 0) {
        long double ld = 0.0L;
        for (int i = 0; i < n; ++i) ld += logl(a[i + (size_t)n * i]);   /* arith.f90:72-93 */
        ld *= 2.0L;
        long double ll, mu = 0.0L, s2 = 0.0L;
        if (likelihood == FB_LIKELIHOOD_CONCENTRATED) {
            /* likelihood.f90:220-243 */
            long double sy = 0.0L, s1 = 0.0L;
            for (int i = 0; i < n; ++i) { w[i] = y[i]; v[i] = 1.0L; }
            potrs_ext(n, a, w);
            potrs_ext(n, a, v);
            for (int i = 0; i < n; ++i) { sy += w[i]; s1 += v[i]; }
            mu = sy / s1;
            for (int i = 0; i < n; ++i) { r[i] = (long double)y[i] - mu; w[i] = r[i]; }
            potrs_ext(n, a, w);
            long double q = 0.0L;
            for (int i = 0; i < n; ++i) q += r[i] * w[i];
            s2 = q / (long double)n;
            ll = -0.5L * (long double)n * logl(s2) - 0.5L * ld;
        } else {
            /* likelihood.f90:178-191; r = y - m is formed in FP64 as both
             * the reference and the device do */
            for (int i = 0; i < n; ++i) { r[i] = (long double)(y[i] - (mean ? mean[i] : 0.0)); w[i] = r[i]; }
            potrs_ext(n, a, w);
            long double q = 0.0L;
            for (int i = 0; i < n; ++i) q += r[i] * w[i];
            ll = -0.5L * q - 0.5L * ld - 0.5L * (long double)n * logl(2.0L * 3.141592653589793238462643383279503L);
        }
        if (out4) { out4[0] = (double)ll; out4[1] = (double)ld; out4[2] = (double)mu; out4[3] = (double)s2; }
    }
    free(a); free(r); free(w); free(v);
    return info;
}

/* ---- mean functions (src/mean/) ------------------------------------------ */

int or_mean(int kind, int n, int d, const double *x, const double *y, double *m,
            double *beta_out, double *xmin_out, double *ymin_out)
{
    if (kind == 0) {                       /* zero.f90:47-52 */
        for (int i = 0; i < n; ++i) m[i] = 0.0;
        return 0;
    }
    if (kind == 1) {                       /* constant.f90:42-55, arith.f90:44-48 */
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += y[i];
        s /= (double)n;
        for (int i = 0; i < n; ++i) m[i] = s;
        return 0;
    }
    if (kind != 2 && kind != 3) return -1;
    /* least_squares.f90:54-93: imin = minloc(y); a = x - xmin; b = y - ymin;
     * a = a*a for the quadratic mean; beta = posv(a'a, a'b). */
    int imin = 0;
    for (int i = 1; i < n; ++i) if (y[i] < y[imin]) imin = i;
    double *a = (double *)malloc(sizeof(double) * (size_t)n * d);
    double *aa = (double *)calloc((size_t)d * d, sizeof(double));
    double *ab = (double *)calloc((size_t)d, sizeof(double));
    double *xmin = (double *)malloc(sizeof(double) * d);
    for (int f = 0; f < d; ++f) xmin[f] = x[imin + (size_t)n * f];
    double ymin = y[imin];
    for (int f = 0; f < d; ++f)
        for (int i = 0; i < n; ++i) {
            double v = x[i + (size_t)n * f] - xmin[f];
            if (kind == 3) v = v * v;
            a[i + (size_t)n * f] = v;
        }
    for (int f = 0; f < d; ++f) {
        for (int g = 0; g < d; ++g) {
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += a[i + (size_t)n * g] * a[i + (size_t)n * f];
            aa[g + (size_t)d * f] = s;
        }
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += a[i + (size_t)n * f] * (y[i] - ymin);
        ab[f] = s;
    }
    int info = or_potrf(d, aa, d);
    if (info == 0) or_potrs(d, aa, d, ab);
    /* linear.f90:41-51 / quadratic.f90:41-52: m = a*beta + ymin */
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int f = 0; f < d; ++f) s += a[i + (size_t)n * f] * ab[f];
        m[i] = s + ymin;
    }
    if (beta_out) memcpy(beta_out, ab, sizeof(double) * d);
    if (xmin_out) memcpy(xmin_out, xmin, sizeof(double) * d);
    if (ymin_out) *ymin_out = ymin;
    free(a); free(aa); free(ab); free(xmin);
    return info;
}

/* ---- CPU baseline --------------------------------------------------------- */

typedef void (*dpotrf_fn)(const char *, const int *, double *, const int *, int *, size_t);
typedef void (*dpotrs_fn)(const char *, const int *, const int *, const double *, const int *,
                          double *, const int *, int *, size_t);
typedef void (*dgemv_fn)(const char *, const int *, const int *, const double *, const double *,
                         const int *, const double *, const int *, const double *, double *,
                         const int *, size_t);
typedef void (*setthreads_fn)(int);

struct or_baseline {
    int n, d, likelihood, optimise_noise, npar_k, nnodes;
    double noise;
    or_node nodes[OR_MAX_NODES];
    int *dims;
    double *x, *y, *mean;
    double *dist;      /* (d, n, n): diff^2, lower triangle (distance_matrix.f90:64-93) */
    double *per;       /* (d, n, n): sin(sqrt(D)/2)^2 (distance_matrix.f90:130-142)     */
    double *cyc;       /* (d, n, n): cyclic diff^2 (distance_matrix.f90:95-128)          */
    void *lib;
    dpotrf_fn potrf;
    dpotrs_fn potrs;
    dgemv_fn gemv;
};

or_baseline *or_baseline_create(int n, int d, const double *x, const double *y,
                                const double *mean, const fb_kernel_spec *spec,
                                double noise, int optimise_noise, int likelihood,
                                const char *lapack_path, const char *prefix)
{
    or_baseline *b = (or_baseline *)calloc(1, sizeof(or_baseline));
    if (!b) return NULL;
    if (parse_spec(spec, d, b->nodes, &b->npar_k)) { free(b); return NULL; }
    b->n = n; b->d = d; b->likelihood = likelihood; b->optimise_noise = optimise_noise;
    b->noise = noise; b->nnodes = spec->n_nodes;
    int nd = 0;
    int need_d = 0, need_p = 0, need_c = 0;
    for (int k = 0; k < spec->n_nodes; ++k) {
        nd += b->nodes[k].ndim;
        if (b->nodes[k].type == FB_LEAF_RBF) need_d = 1;
        if (b->nodes[k].type == FB_LEAF_PERIODIC) need_p = 1;
        if (b->nodes[k].type == FB_LEAF_RBF_CYCLIC) need_c = 1;
    }
    b->dims = (int *)malloc(sizeof(int) * (nd > 0 ? nd : 1));
    if (nd) memcpy(b->dims, spec->active_dims, sizeof(int) * nd);
    b->x = (double *)malloc(sizeof(double) * (size_t)n * d);
    b->y = (double *)malloc(sizeof(double) * n);
    b->mean = (double *)malloc(sizeof(double) * n);
    memcpy(b->x, x, sizeof(double) * (size_t)n * d);
    memcpy(b->y, y, sizeof(double) * n);
    for (int i = 0; i < n; ++i) b->mean[i] = mean ? mean[i] : 0.0;
    size_t tn = (size_t)d * n * n;
    if (need_d || need_p) b->dist = (double *)calloc(tn, sizeof(double));
    if (need_p) b->per = (double *)calloc(tn, sizeof(double));
    if (need_c) b->cyc = (double *)calloc(tn, sizeof(double));
    #pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < n; ++j) {
        for (int i = j + 1; i < n; ++i) {
            size_t base = (size_t)d * (i + (size_t)n * j);
            for (int f = 0; f < d; ++f) {
                double diff = x[i + (size_t)n * f] - x[j + (size_t)n * f];
                if (b->dist) b->dist[base + f] = diff * diff;
                if (b->per) {
                    double s = sin(sqrt(diff * diff) / 2.0);
                    b->per[base + f] = s * s;
                }
                if (b->cyc) {
                    double c = (f + 1) % 3 == 0 ? cyclic_wrap(diff) : diff;
                    b->cyc[base + f] = c * c;
                }
            }
        }
    }
    if (lapack_path && lapack_path[0]) {
        b->lib = dlopen(lapack_path, RTLD_NOW | RTLD_LOCAL);
        if (b->lib) {
            char name[128];
            const char *pf = prefix ? prefix : "";
            snprintf(name, sizeof name, "%sdpotrf_", pf); b->potrf = (dpotrf_fn)dlsym(b->lib, name);
            snprintf(name, sizeof name, "%sdpotrs_", pf); b->potrs = (dpotrs_fn)dlsym(b->lib, name);
            snprintf(name, sizeof name, "%sdgemv_", pf);  b->gemv = (dgemv_fn)dlsym(b->lib, name);
            snprintf(name, sizeof name, "%sopenblas_set_num_threads", pf);
            setthreads_fn st = (setthreads_fn)dlsym(b->lib, name);
            if (st) st(1);
            if (!b->potrf || !b->potrs || !b->gemv) { b->potrf = NULL; b->potrs = NULL; b->gemv = NULL; }
        }
    }
    return b;
}

int or_baseline_lapack(const or_baseline *b) { return b && b->potrf ? 1 : 0; }

/* y(0:m-1) = alpha * A' v for A (nact x m), as dgemv('t') does. */
static void gemv_t(const or_baseline *b, int nact, int m, double alpha, const double *a,
                   const double *v, double *out)
{
    if (b->gemv) {
        const double zero = 0.0;
        const int one = 1;
        b->gemv("T", &nact, &m, &alpha, a, &nact, v, &one, &zero, out, &one, 1);
        return;
    }
    for (int c = 0; c < m; ++c) {
        double t = 0.0;
        for (int q = 0; q < nact; ++q) t += a[(size_t)nact * c + q] * v[q];
        out[c] = alpha * t;
    }
}

/* One particle through the reference's host algorithm. */
static int baseline_one(or_baseline *b, const double *theta, double *ll)
{
    int n = b->n, d = b->d;
    size_t nn = (size_t)n * n;
    double *stack[OR_MAX_NODES];
    int sp = 0;
    double *gather = NULL, *lvec = (double *)malloc(sizeof(double) * (d > 0 ? d : 1));
    int info = 0;
    for (int k = 0; k < b->nnodes; ++k) {
        const or_node *o = &b->nodes[k];
        if (o->type == FB_OP_SUM || o->type == FB_OP_PROD) {
            double *r1 = stack[sp - 2], *r2 = stack[sp - 1];
            if (o->type == FB_OP_SUM) for (size_t e = 0; e < nn; ++e) r1[e] += r2[e];
            else for (size_t e = 0; e < nn; ++e) r1[e] *= r2[e];
            free(r2);
            --sp;
            continue;
        }
        double *r = (double *)malloc(sizeof(double) * nn);
        const double *th = theta + o->par_off;
        const int *ad = b->dims + o->dim_off;
        if (o->type == FB_LEAF_RBF || o->type == FB_LEAF_PERIODIC) {
            const double *t3 = o->type == FB_LEAF_RBF ? b->dist : b->per;
            double alpha = o->type == FB_LEAF_RBF ? -0.5 : -2.0;
            for (int q = 0; q < o->ndim; ++q) {
                if (o->type == FB_LEAF_RBF) lvec[q] = exp(th[q]);
                else { double l = 1.0 / exp(th[q]); lvec[q] = l * l; }
            }
            int all = o->ndim == d;
            for (int q = 0; all && q < o->ndim; ++q) all = ad[q] == q + 1;
            if (all) {
                gemv_t(b, d, (int)nn, alpha, t3, lvec, r);
            } else {
                /* stationary_kernel.f90:114-119: gather a (nact, n, n) copy */
                if (!gather) gather = (double *)malloc(sizeof(double) * nn * o->ndim);
                else gather = (double *)realloc(gather, sizeof(double) * nn * o->ndim);
                for (size_t e = 0; e < nn; ++e)
                    for (int q = 0; q < o->ndim; ++q)
                        gather[e * o->ndim + q] = t3[e * d + ad[q] - 1];
                gemv_t(b, o->ndim, (int)nn, alpha, gather, lvec, r);
            }
            for (size_t e = 0; e < nn; ++e) r[e] = exp(r[e]);
        } else if (o->type == FB_LEAF_RBF_CYCLIC) {
            for (int f = 0; f < d; ++f) lvec[f] = exp(th[f]);
            for (int j = 0; j < n; ++j) {
                r[j + (size_t)n * j] = 0.0;
                for (int i = j + 1; i < n; ++i) {
                    const double *c = b->cyc + (size_t)d * (i + (size_t)n * j);
                    double t = 0.0;
                    for (int f = 0; f < d; ++f) t += lvec[f] * c[f];
                    r[i + (size_t)n * j] = t;
                    r[j + (size_t)n * i] = t;
                }
            }
            for (size_t e = 0; e < nn; ++e) r[e] = exp(-0.5 * r[e]);
        } else {
            for (int j = 0; j < n; ++j)
                for (int i = j; i < n; ++i)
                    r[i + (size_t)n * j] = leaf_value(o, b->dims, n, d, b->x, theta, i, j);
        }
        stack[sp++] = r;
    }
    double *k = stack[0];
    double nug = nugget_of(theta, b->npar_k, b->noise, b->optimise_noise);
    for (int i = 0; i < n; ++i) k[i + (size_t)n * i] += nug;
    if (b->potrf) {
        b->potrf("L", &n, k, &n, &info, 1);
    } else {
        info = or_potrf(n, k, n);
    }
    if (info == 0) {
        double logdet = or_logdet(n, k, n);
        double o4[4];
        if (b->potrs) {
            /* likelihood_from_factor with LAPACK solves */
            int one = 1, linfo = 0;
            double *r = (double *)malloc(sizeof(double) * n);
            double *a = (double *)malloc(sizeof(double) * n);
            if (b->likelihood == FB_LIKELIHOOD_CONCENTRATED) {
                double *ry = (double *)malloc(sizeof(double) * n);
                double *r1 = (double *)malloc(sizeof(double) * n);
                memcpy(ry, b->y, sizeof(double) * n);
                b->potrs("L", &n, &one, k, &n, ry, &n, &linfo, 1);
                for (int i = 0; i < n; ++i) r1[i] = 1.0;
                b->potrs("L", &n, &one, k, &n, r1, &n, &linfo, 1);
                double sy = 0.0, s1 = 0.0;
                for (int i = 0; i < n; ++i) { sy += ry[i]; s1 += r1[i]; }
                double mu = sy / s1;
                for (int i = 0; i < n; ++i) r[i] = b->y[i] - mu;
                memcpy(a, r, sizeof(double) * n);
                b->potrs("L", &n, &one, k, &n, a, &n, &linfo, 1);
                double s2 = dot(n, r, a) / (double)n;
                o4[0] = -0.5 * (double)n * log(s2) - 0.5 * logdet;
                free(ry); free(r1);
            } else {
                for (int i = 0; i < n; ++i) r[i] = b->y[i] - b->mean[i];
                memcpy(a, r, sizeof(double) * n);
                b->potrs("L", &n, &one, k, &n, a, &n, &linfo, 1);
                o4[0] = -0.5 * dot(n, r, a) - 0.5 * logdet - 0.5 * (double)n * log(OR_TWOPI);
            }
            free(r); free(a);
        } else {
            likelihood_from_factor(n, k, b->y, b->mean, b->likelihood, logdet, NULL, o4);
        }
        *ll = o4[0];
    } else {
        *ll = NAN;
    }
    free(k);
    free(gather);
    free(lvec);
    return info;
}

int or_baseline_swarm(or_baseline *b, int s, const double *theta, double *ll, int *info,
                      int nthreads)
{
    if (!b) return -1;
    int p = b->npar_k + (b->optimise_noise ? 1 : 0);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    (void)nthreads;
    /* pso.f90:419-452: one particle per OpenMP thread */
    #pragma omp parallel for schedule(dynamic)
    for (int q = 0; q < s; ++q)
        info[q] = baseline_one(b, theta + (size_t)p * q, &ll[q]);
    return 0;
}

void or_baseline_destroy(or_baseline *b)
{
    if (!b) return;
    free(b->dims); free(b->x); free(b->y); free(b->mean);
    free(b->dist); free(b->per); free(b->cyc);
    if (b->lib) dlclose(b->lib);
    free(b);
}
