// fb_likelihood.hip -- log-likelihood of every factored particle.
//
// After the bordered factorisation the right-hand-side rows of each matrix
// hold z = L^{-1} rhs, so
//   marginal      (likelihood.f90:161-201)
//       r' K^{-1} r = z'z,  ll = -z'z/2 - logdet/2 - n log(2 pi)/2
//   concentrated  (likelihood.f90:203-253), rhs = (y, 1):
//       mu = z_1' z_y / z_1' z_1,  w = z_y - mu z_1,  sigma2 = w'w / n,
//       ll = -n log(sigma2)/2 - logdet/2
// with logdet = 2 sum log L_ii (arith.f90:72-93).  The residual vector
// (z or w) is kept for the weights alpha = L^{-T} w of the model output.
#include "fb_common.h"

namespace fb {

__device__ inline double block_sum(double v, double *red)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void __launch_bounds__(kThreads) likelihood_kernel(const double *__restrict__ K, int np, int n,
                                                              int nsteps, int likelihood,
                                                              const int *__restrict__ info,
                                                              const double *__restrict__ logdet_part,
                                                              double *__restrict__ ll,
                                                              double *__restrict__ extra,
                                                              double *__restrict__ zbuf)
{
    __shared__ double red[kThreads / 64];
    const int s = blockIdx.x;
    if (info[s] != 0) {
        if (threadIdx.x == 0) ll[s] = __builtin_nan("");
        return;
    }
    const double *A = K + (size_t)s * np * np;
    double ld = 0.0;
    for (int j = 0; j < nsteps; ++j) ld += logdet_part[(size_t)s * nsteps + j];
    const double logdet = 2.0 * ld;
    double *z = zbuf + (size_t)s * n;
    const double log_two_pi = log(kTwoPi);
    if (likelihood == 0) {
        double q = 0.0;
        for (int c = threadIdx.x; c < n; c += kThreads) {
            const double v = A[(size_t)c * np + n];
            z[c] = v;
            q += v * v;
        }
        q = block_sum(q, red);
        if (threadIdx.x == 0) {
            const double v = -0.5 * q - 0.5 * logdet - 0.5 * (double)n * log_two_pi;
            ll[s] = v;
            double *x4 = extra + (size_t)s * 4;
            x4[0] = v; x4[1] = logdet; x4[2] = 0.0; x4[3] = 0.0;
        }
    } else {
        double a = 0.0, b = 0.0;
        for (int c = threadIdx.x; c < n; c += kThreads) {
            const double zy = A[(size_t)c * np + n];
            const double z1 = A[(size_t)c * np + n + 1];
            a += z1 * zy;
            b += z1 * z1;
        }
        a = block_sum(a, red);
        b = block_sum(b, red);
        const double mu = a / b;
        double q = 0.0;
        for (int c = threadIdx.x; c < n; c += kThreads) {
            const double w = A[(size_t)c * np + n] - mu * A[(size_t)c * np + n + 1];
            z[c] = w;
            q += w * w;
        }
        q = block_sum(q, red);
        if (threadIdx.x == 0) {
            const double s2 = q / (double)n;
            const double v = -0.5 * (double)n * log(s2) - 0.5 * logdet;
            ll[s] = v;
            double *x4 = extra + (size_t)s * 4;
            x4[0] = v; x4[1] = logdet; x4[2] = mu; x4[3] = s2;
        }
    }
}

void launch_likelihood(const SwarmBuffers &b, hipStream_t st)
{
    hipLaunchKernelGGL(likelihood_kernel, dim3(b.S), dim3(kThreads), 0, st, b.K, b.np, b.n,
                       b.np / kNBI, b.likelihood, b.info, b.logdet_part, b.ll, b.extra, b.zbuf);
}

// alpha = L^{-T} w for one particle (weights of the model file), by blocks of
// 64 columns from the last: alpha_b = L_bb^{-T} (w_b - L(rest, b)^T alpha_rest).
__global__ void __launch_bounds__(kThreads) backsolve_kernel(const double *__restrict__ K, int np, int n,
                                                             const double *__restrict__ w,
                                                             double *__restrict__ alpha)
{
    extern __shared__ double al[];   // n entries
    __shared__ double part[kThreads / 64][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int c = tid; c < n; c += kThreads) al[c] = 0.0;
    __syncthreads();
    const int nb = ceil_div(n, 64);
    for (int b = nb - 1; b >= 0; --b) {
        const int c0 = b * 64;
        const int cn = min(64, n - c0);
        // t_c = w_c - sum_{i >= c0 + cn} L(i, c) alpha_i for the block's columns;
        // wave w sums rows i = c0 + cn + w, +4, ... for column c0 + lane
        double acc = 0.0;
        if (lane < cn) {
            const double *col = K + (size_t)(c0 + lane) * np;
            for (int i = c0 + cn + wave; i < n; i += kThreads / 64) acc += col[i] * al[i];
        }
        part[wave][lane] = acc;
        __syncthreads();
        if (wave == 0) {
            double t = (lane < cn) ? w[c0 + lane] - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane])) : 0.0;
            // triangular solve inside the block, last column first
            for (int c = cn - 1; c >= 0; --c) {
                const double lcc = K[(size_t)(c0 + c) * np + c0 + c];
                double v = __shfl(t, c, 64) / lcc;
                if (lane == c) t = v;
                // subtract L(c, k) v from columns k < c: L(c0+c, c0+k)
                if (lane < c) t -= K[(size_t)(c0 + lane) * np + c0 + c] * v;
                if (lane == c) al[c0 + c] = v;
            }
        }
        __syncthreads();
    }
    for (int c = tid; c < n; c += kThreads) alpha[c] = al[c];
}

void launch_backsolve(const SwarmBuffers &b, int particle, double *alpha, hipStream_t st)
{
    hipLaunchKernelGGL(backsolve_kernel, dim3(1), dim3(kThreads), (size_t)b.n * sizeof(double), st,
                       b.K + (size_t)particle * b.np * b.np, b.np, b.n, b.zbuf + (size_t)particle * b.n,
                       alpha);
}

}  // namespace fb
