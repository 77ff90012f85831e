// fb_pso.hip -- the particle swarm's state and its per-iteration update on
// the device, so that one PSO iteration (selection of the previous
// iteration's bests, velocity / position update, covariance build,
// factorisation, likelihood) is one captured graph.
//
// Rules of the reference PSO (src/optimiser/pso/pso.f90):
//   * update (pso.f90:422-434): v = w v + c1 r1 (pbest - x) + c2 r2 (gbest - x),
//     x = x + v, no bounds; r1, r2 ~ U[0,1)^(P x S) drawn on the host before
//     the swarm loop (pso.f90:405-409) and handed over per iteration, with
//     the swarm's current (w, c1, c2), which a parameter updater may change
//     between iterations (pso.f90:638-825);
//   * pbest (pso.f90:440-443): replaced when cost <= pbest cost;
//   * gbest (pso.f90:454-459): particles in index order, replaced when
//     cost <= running gbest cost, so the last particle holding the minimum
//     wins when the minimum is <= the previous gbest cost;
//   * a particle whose matrix is not positive definite costs +Inf here (the
//     host re-randomises and re-evaluates it before selection, as get_cost
//     does, particle.f90:149-162).
// Layout: particle k's parameters at x[k * P + j] (the theta layout of the
// swarm evaluation), so the evaluation reads the positions in place.
#include <algorithm>

#include "fb_common.h"

namespace fb {

// cost = -logL, +Inf for a failed factorisation; copies INFO and marks the
// costs pending.
__global__ void pso_cost_kernel(int S, const double *__restrict__ ll, const int *__restrict__ info_src,
                                double *__restrict__ cost, int *__restrict__ info, int *__restrict__ pending)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < S) {
        const int f = info_src[k];
        cost[k] = f == 0 ? -ll[k] : __builtin_inf();
        info[k] = f;
    }
    if (k == 0) *pending = 1;
}

// pbest / gbest selection of the pending costs (one work-group of 256).
// state[0] = gbest cost; gbest holds its position.  Runs only when the
// costs are pending, so a graph replay after an explicit selection leaves
// the bests alone.
__global__ void __launch_bounds__(256) pso_select_kernel(int S, int P, const double *__restrict__ x,
                                                         const double *__restrict__ cost,
                                                         double *__restrict__ pbest,
                                                         double *__restrict__ pcost,
                                                         double *__restrict__ gbest,
                                                         double *__restrict__ state,
                                                         int *__restrict__ pending, int *__restrict__ gidx)
{
    __shared__ double s_min[256];
    __shared__ int s_idx[256];
    __shared__ int s_go;
    const int t = threadIdx.x;
    if (t == 0) s_go = *pending;
    __syncthreads();
    if (!s_go) return;
    // minimum cost, ties to the highest index; NaN never selects
    double m = __builtin_inf();
    int mi = -1;
    for (int k = t; k < S; k += 256) {
        const double c = cost[k];
        if (c <= m) { m = c; mi = k; }   // k increases: a tie moves to the later particle
        if (c <= pcost[k]) {
            pcost[k] = c;
            for (int j = 0; j < P; ++j) pbest[(size_t)k * P + j] = x[(size_t)k * P + j];
        }
    }
    s_min[t] = m;
    s_idx[t] = mi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            const double a = s_min[t], b = s_min[t + w];
            const int ia = s_idx[t], ib = s_idx[t + w];
            if (ib >= 0 && (ia < 0 || b < a || (b == a && ib > ia))) { s_min[t] = b; s_idx[t] = ib; }
        }
        __syncthreads();
    }
    const int g = s_idx[0];
    const double gc = s_min[0];
    if (g >= 0 && gc <= state[0]) {
        for (int j = t; j < P; j += 256) gbest[j] = x[(size_t)g * P + j];
        if (t == 0) { state[0] = gc; *gidx = g; }
    }
    if (t == 0) *pending = 0;
}

// Velocity and position update of every particle with the coefficients
// (w, c1, c2) at coef -- device memory, so that a parameter updater can
// change them from one replay of the captured iteration to the next.
__global__ void pso_update_kernel(int S, int P, const double *__restrict__ coef,
                                  const double *__restrict__ r1, const double *__restrict__ r2,
                                  const double *__restrict__ pbest, const double *__restrict__ gbest,
                                  double *__restrict__ x, double *__restrict__ v)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= S * P) return;
    const int j = e % P;
    const double w = coef[0], c1 = coef[1], c2 = coef[2];
    const double xe = x[e];
    const double ve = w * v[e] + c1 * r1[e] * (pbest[e] - xe) + c2 * r2[e] * (gbest[j] - xe);
    v[e] = ve;
    x[e] = xe + ve;
}

// Replaces the positions / velocities of k listed particles (rows of xs, vs).
__global__ void pso_scatter_kernel(int k, int P, const int *__restrict__ idx, const double *__restrict__ xs,
                                   const double *__restrict__ vs, double *__restrict__ x, double *__restrict__ v)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= k * P) return;
    const int r = e / P, j = e % P;
    x[(size_t)idx[r] * P + j] = xs[e];
    v[(size_t)idx[r] * P + j] = vs[e];
}

// Copies the positions of k listed particles into consecutive rows.
__global__ void pso_gather_kernel(int k, int P, const int *__restrict__ idx, const double *__restrict__ x,
                                  double *__restrict__ out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= k * P) return;
    out[e] = x[(size_t)idx[e / P] * P + e % P];
}

// Writes the costs / INFO of k re-evaluated particles back to their slots.
__global__ void pso_put_cost_kernel(int k, const int *__restrict__ idx, const double *__restrict__ ll,
                                    const int *__restrict__ info_k, double *__restrict__ cost,
                                    int *__restrict__ info, int *__restrict__ pending)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < k) {
        cost[idx[r]] = info_k[r] == 0 ? -ll[r] : __builtin_inf();
        info[idx[r]] = info_k[r];
    }
    if (r == 0) *pending = 1;
}

static int blocks(long long n) { return (int)std::max(1LL, (n + 255) / 256); }

void launch_pso_cost(const PsoState &p, const double *ll, const int *info_src, hipStream_t st)
{
    hipLaunchKernelGGL(pso_cost_kernel, dim3(blocks(p.S)), dim3(256), 0, st, p.S, ll, info_src, p.cost, p.info,
                       p.pending);
}

void launch_pso_select(const PsoState &p, hipStream_t st)
{
    hipLaunchKernelGGL(pso_select_kernel, dim3(1), dim3(256), 0, st, p.S, p.P, p.x, p.cost, p.pbest, p.pcost,
                       p.gbest, p.state, p.pending, p.gidx);
}

void launch_pso_update(const PsoState &p, const double *r1, const double *r2, const double *coef, hipStream_t st)
{
    hipLaunchKernelGGL(pso_update_kernel, dim3(blocks((long long)p.S * p.P)), dim3(256), 0, st, p.S, p.P, coef,
                       r1, r2, p.pbest, p.gbest, p.x, p.v);
}

void launch_pso_scatter(const PsoState &p, int k, const int *idx, const double *xs, const double *vs,
                        hipStream_t st)
{
    hipLaunchKernelGGL(pso_scatter_kernel, dim3(blocks((long long)k * p.P)), dim3(256), 0, st, k, p.P, idx, xs,
                       vs, p.x, p.v);
}

void launch_pso_gather(const PsoState &p, int k, const int *idx, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(pso_gather_kernel, dim3(blocks((long long)k * p.P)), dim3(256), 0, st, k, p.P, idx, p.x,
                       out);
}

void launch_pso_put_cost(const PsoState &p, int k, const int *idx, const double *ll, const int *info_k,
                         hipStream_t st)
{
    hipLaunchKernelGGL(pso_put_cost_kernel, dim3(blocks(k)), dim3(256), 0, st, k, idx, ll, info_k, p.cost, p.info,
                       p.pending);
}

}  // namespace fb
