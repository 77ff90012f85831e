// fb_grad.hip -- adjoint gradient of the log-likelihood for every particle.
//
//   d ll / d theta_p = 1/2 tr( W dK/dtheta_p ),   W = alpha alpha^T / s2 - K^{-1}
//
// marginal (likelihood.f90:161-201): alpha = K^{-1}(y - m), s2 = 1;
// concentrated (likelihood.f90:203-253): alpha = K^{-1}(y - mu 1), s2 =
// sigma2_hat (the mu and sigma2 terms vanish at their optimum).  The
// reference parses a gradient flag but computes none (gpr.f90:272-282); this
// is the derivative of the reference's own log-likelihood, pinned by central
// finite differences and by the oracle's dense-W gradient.
//
// After the batched factorisation (L in the lower triangle of each K, the
// inverses of the 16x16 diagonal blocks in dinv, w = L^{-1}(y - ...) in zbuf):
//   1. U = L^{-T} by recursive doubling over block sizes 64, 128, ... (the
//      diagonal blocks by the identity panel solve, then two tile-GEMM
//      launches per level, uinv_diag / uinv_t / uinv_ux kernels);
//   2. columns [n, np) of U cleared (zero_tail_columns_kernel);
//   3. alpha = L^{-T} w = U w, a coalesced product (ualpha_kernel);
//   4. -K^{-1} = -U U^T on the lower tiles (nkinv_kernel);
//   5. one fused pass over the lower triangle recomputes every leaf of K_ij
//      from the staged features, as the build does, and accumulates
//      sum_ij W_ij dK_ij/dtheta_p per parameter -- dK/dtheta is never formed
//      (grad_pair_kernel, a chunk of parameters per pass);
//   6. per-parameter sums over the tiles in a fixed order, times the
//      derivative of each hyperparameter map (grad_reduce_kernel).
//
// Leaf derivatives (K_ij = sum_t c_t exp(sum_{l in t} e_l) + nugget delta_ij,
// e_l = sum_q h_q f_q):
//   rbf / rbf-cyclic  h = -exp(theta)/2        dh/dtheta = h
//   periodic          h = -2 exp(-2 theta)     dh/dtheta = -2 h
//   linear            u = sum_q exp(theta_q)   du/dtheta_q = exp(theta_q)
//   constant          u = theta                du/dtheta = 1
//   nugget            dK_ij/dtheta = delta_ij
// so dK_ij/dtheta_q = dh_q f_q G_l with G_l = sum_{t containing l} c_t X_t
// (stationary leaf l), or du_q H_l with H_l = sum_{t containing l} c_t^{(l)} X_t
// (linear / constant leaf l, c_t^{(l)} the term's coefficient without u_l),
// X_t = exp(sum_{l in t} e_l).
#include "fb_tiles.h"

namespace fb {

// ---- 1. U = L^{-T} by recursive doubling ----------------------------------
//
// With L = [[A, 0], [B, C]] (A, C square blocks), L^{-1} = [[A^{-1}, 0],
// [-C^{-1} B A^{-1}, C^{-1}]].  Bottom up over block sizes m = 64, 128, ...:
// the diagonal 64x64 blocks first (U_kk = L_kk^{-T} by the identity panel
// solve, X_kk = L_kk^{-1} its transpose), then for each pair of m-blocks
//   T   =  U_A B^T                (m x mc)
//   U12 = -T X_C^T                (the pair's off-diagonal block of U)
//   X21 = -X_C T^T                (the same block of X = L^{-1})
// with both factors always read column major (U and X are kept so that no
// product needs a transposed operand).  U_A is upper and X_C lower
// triangular, so each tile's depth stops at the triangle.  Two launches per
// level, log2(Np / 64) levels: the chain of block-row steps it replaces was
// Np / 64 dependent launches.

// Level 0: diagonal blocks (grid: blocks x particles).
__global__ void __launch_bounds__(kThreads) uinv_diag_kernel(const double *__restrict__ K, double *__restrict__ U,
                                                            double *__restrict__ X, int np,
                                                            const double *__restrict__ dinv,
                                                            const int *__restrict__ info)
{
    __shared__ double ys[kNBI * kPLD];
    const int k = blockIdx.x, s = blockIdx.y;
    if (info[s] != 0) return;
    const size_t mat = (size_t)s * np * np;
    const double *Lkk = K + mat + (size_t)64 * k * np + 64 * k;
    double *Ukk = U + mat + (size_t)64 * k * np + 64 * k;
    trsm64_solve<true, true>(Ukk, np, Ukk, np, Lkk, np, dinv + ((size_t)s * (np / kNBI) + k) * kDinv, ys);
    __syncthreads();
    // X_kk(r, c) = U_kk(c, r): ys[c' * kPLD + r'] = U_kk(r', c')
    double *Xkk = X + mat + (size_t)64 * k * np + 64 * k;
    for (int e = threadIdx.x; e < kNBI * kNBI; e += kThreads) {
        const int r = e % kNBI, c = e / kNBI;
        Xkk[(size_t)c * np + r] = ys[r * kPLD + c];
    }
}

// T = U_A B^T for every pair of level m (grid: pairs * (m/64)^2 tiles x
// particles; T is kept in the scratch buffer at the position of U12).
__global__ void __launch_bounds__(kThreads) uinv_t_kernel(const double *__restrict__ K,
                                                         const double *__restrict__ U, double *__restrict__ Tb,
                                                         int np, int m, const int *__restrict__ info)
{
    __shared__ __attribute__((aligned(16))) double smem[TileShape<64, 64>::SMEM];
    const int s = blockIdx.y;
    if (info[s] != 0) return;
    const int tm = m / 64, per = tm * tm;
    const int q = blockIdx.x / per, u = blockIdx.x % per;
    const int a0 = 2 * m * q, mc = min(m, np - a0 - m);
    const int R = u / tm, Ct = u % tm;
    if (mc <= 0 || Ct * 64 >= mc) return;
    const size_t mat = (size_t)s * np * np;
    const int k0 = 64 * R;     // U_A(r, k) = 0 for k < r
    // T(r, c) = sum_k U_A(r, k) B(c, k);  U_A(r, k) at U[(a0 + k) np + a0 + r],
    // B(c, k) = L(a0 + m + c, a0 + k)
    tile_gemm_sub<64, 64, kKC, 2, false, true, true>(Tb + mat + (size_t)(a0 + m + 64 * Ct) * np + a0 + 64 * R, np,
                                                     U + mat + (size_t)(a0 + k0) * np + a0 + 64 * R, np,
                                                     K + mat + (size_t)(a0 + k0) * np + a0 + m + 64 * Ct, np,
                                                     m - k0, smem);
}

// U12 = -T X_C^T and X21 = -X_C T^T for every pair of level m (grid:
// 2 * pairs * (m/64)^2 tiles x particles).
__global__ void __launch_bounds__(kThreads) uinv_ux_kernel(double *__restrict__ U, double *__restrict__ X,
                                                          const double *__restrict__ Tb, int np, int m,
                                                          const int *__restrict__ info)
{
    __shared__ __attribute__((aligned(16))) double smem[TileShape<64, 64>::SMEM];
    const int s = blockIdx.y;
    if (info[s] != 0) return;
    const int tm = m / 64, per = tm * tm;
    const int which = blockIdx.x % 2, t = blockIdx.x / 2;
    const int q = t / per, u = t % per;
    const int a0 = 2 * m * q, mc = min(m, np - a0 - m);
    if (mc <= 0) return;
    const size_t mat = (size_t)s * np * np;
    // T(r, c) at Tb[(a0 + m + c) np + a0 + r] (r < m, c < mc);
    // X_C(r, c) at X[(a0 + m + c) np + a0 + m + r]
    const double *Tp = Tb + mat + (size_t)(a0 + m) * np + a0;
    const double *XC = X + mat + (size_t)(a0 + m) * np + a0 + m;
    if (which == 0) {
        // U12(r, c) = -sum_{k <= c} T(r, k) X_C(c, k): rows r < m, columns c < mc
        const int R = u / tm, Ct = u % tm;
        if (Ct * 64 >= mc) return;
        const int depth = min(64 * (Ct + 1), mc);
        tile_gemm_sub<64, 64, kKC, 2, false, true, false>(U + mat + (size_t)(a0 + m + 64 * Ct) * np + a0 + 64 * R, np,
                                                          Tp + 64 * R, np, XC + 64 * Ct, np, depth, smem);
    } else {
        // X21(r, c) = -sum_{k <= r} X_C(r, k) T(c, k): rows r < mc, columns c < m
        const int Rt = u / tm, C = u % tm;
        if (Rt * 64 >= mc) return;
        const int depth = min(64 * (Rt + 1), mc);
        tile_gemm_sub<64, 64, kKC, 2, false, true, false>(X + mat + (size_t)(a0 + 64 * C) * np + a0 + m + 64 * Rt, np,
                                                          XC + 64 * Rt, np, Tp + 64 * C, np, depth, smem);
    }
}

// Columns [n, np) of U are not part of L^{-T}: zero them before U U^T.
__global__ void zero_tail_columns_kernel(double *__restrict__ U, int np, int n, int S)
{
    const size_t per = (size_t)(np - n) * np;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < per * S; e += (size_t)gridDim.x * blockDim.x) {
        const size_t s = e / per, r = e % per;
        U[s * np * np + (size_t)n * np + r] = 0.0;
    }
}

// ---- 2. alpha = U w for every particle -------------------------------------

// Rows [64 R, 64 R + 64) of alpha = U w (U upper triangular, column major):
// lane = row, the four waves take every fourth column, so each wave reads
// 512 contiguous bytes per column; the wave sums are added in a fixed order.
__global__ void __launch_bounds__(kThreads) ualpha_kernel(const double *__restrict__ U, int np, int n,
                                                         const double *__restrict__ wbuf,
                                                         const int *__restrict__ info, double *__restrict__ alpha)
{
    __shared__ double part[kThreads / 64][64];
    const int R = blockIdx.x, s = blockIdx.y;
    if (info[s] != 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = 64 * R + lane;
    const double *Us = U + (size_t)s * np * np;
    const double *w = wbuf + (size_t)s * n;
    double acc = 0.0;
    for (int c = 64 * R + wave; c < n; c += kThreads / 64) acc += Us[(size_t)c * np + r] * w[c];
    part[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && r < n)
        alpha[(size_t)s * n + r] = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

// ---- 3. -K^{-1} = -U U^T on the lower tiles ------------------------------

__global__ void __launch_bounds__(kThreads) nkinv_kernel(const double *__restrict__ U, double *__restrict__ NK,
                                                        int np, int nbn, const int *__restrict__ info)
{
    __shared__ __attribute__((aligned(16))) double smem[TileShape<64, 64>::SMEM];
    const int b = blockIdx.x, s = blockIdx.y;
    if (info[s] != 0) return;
    int I = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= b) ++I;
    while (I * (I + 1) / 2 > b) --I;
    const int J = b - I * (I + 1) / 2;
    const double *Us = U + (size_t)s * np * np;
    // NK(I, J) = -sum_{k >= 64I} U(64I + r, k) U(64J + c, k)  (U is upper triangular)
    tile_gemm_sub<64, 64, kKC, 2, false, true>(NK + (size_t)s * np * np + (size_t)64 * J * np + 64 * I, np,
                               Us + (size_t)64 * I * np + 64 * I, np, Us + (size_t)64 * I * np + 64 * J, np,
                               64 * (nbn - I), smem);
}

// ---- 4. fused pass: sum_ij W_ij dK_ij/dtheta_p ----------------------------

constexpr int kGradChunk = 16;   // parameters per pass

struct GradArgs {
    int n, np, S, n_slots, P, ntiles, pass0;
    const double *slots;
    const double *hyp;      // per-particle record of the build (exponent weights, term coefficients, nugget)
    const double *theta;    // P x S
    const KernelProgram *prog;
    const double *alpha;    // S x n
    const double *NK;       // -K^{-1}, lower tiles
    const double *extra;    // S x 4: ll, logdet, mu, sigma2
    const int *info;
    int likelihood;
    double *partial;        // S x ntiles x P
    double expc[kExpConsts];   // constants of exp_nonpos (the build's exp)
};

template <int MAXL>
__device__ inline double sel(const double (&v)[MAXL], int l)
{
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < MAXL; ++k) r = (l == k) ? v[k] : r;
    return r;
}

template <int TB, int MAXL>
__global__ void __launch_bounds__(kThreads) grad_pair_kernel(GradArgs a)
{
    constexpr int CG = kThreads / TB, CPT = TB / CG;
    extern __shared__ double lds[];
    double *row_f = lds;                          // [n_slots][TB]
    double *col_f = lds + (size_t)a.n_slots * TB;
    __shared__ double cex[kMaxTerms][kMaxLeaves];  // term coefficient without leaf l
    __shared__ double red[kThreads / 64][kGradChunk];
    __shared__ int pleaf[kGradChunk], pq[kGradChunk];

    const int b = blockIdx.x, s = blockIdx.y;
    if (a.info[s] != 0) return;
    int ti = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
    while (ti * (ti + 1) / 2 > b) --ti;
    const int tj = b - ti * (ti + 1) / 2;
    const int p0 = a.pass0 + blockIdx.z * kGradChunk;
    const KernelProgram &p = *a.prog;
    const int hs = p.n_par_kernel + kMaxTerms + 1;
    const double *h = a.hyp + (size_t)s * hs;
    const double *th = a.theta + (size_t)s * p.n_par;

    for (int e = threadIdx.x; e < a.n_slots * TB; e += kThreads) {
        const int slot = e / TB, r = e % TB;
        const int gi = ti * TB + r, gj = tj * TB + r;
        row_f[e] = gi < a.n ? a.slots[(size_t)slot * a.n + gi] : 0.0;
        col_f[e] = gj < a.n ? a.slots[(size_t)slot * a.n + gj] : 0.0;
    }
    if (threadIdx.x < kGradChunk) {
        // parameter -> (leaf, index within the leaf); leaf -1: the nugget, -2: none
        const int q = p0 + threadIdx.x;
        int leaf = -2, qi = 0;
        if (q < p.n_par_kernel) {
            for (int l = 0; l < p.n_leaves; ++l) {
                const int off = p.leaf_par_off[l];
                const int np_l = (p.leaf_kind[l] == kConstant) ? 1 : p.leaf_ndim[l];
                if (q >= off && q < off + np_l) { leaf = l; qi = q - off; }
            }
        } else if (q < p.n_par && p.optimise_noise) {
            leaf = -1;
        }
        pleaf[threadIdx.x] = leaf;
        pq[threadIdx.x] = qi;
    }
    if (threadIdx.x == 0) {
        // values of the linear / constant leaves (hyper_kernel's leafval)
        double u[kMaxLeaves];
        for (int l = 0; l < kMaxLeaves; ++l) u[l] = 1.0;
        for (int l = 0; l < p.n_leaves; ++l) {
            const int off = p.leaf_par_off[l];
            if (p.leaf_kind[l] == kLinear) {
                double t = 0.0;
                for (int q = 0; q < p.leaf_ndim[l]; ++q) t += 1.0 / exp_cr(-th[off + q]);
                u[l] = t;
            } else if (p.leaf_kind[l] == kConstant) {
                u[l] = th[off];
            }
        }
        for (int t = 0; t < kMaxTerms; ++t)
            for (int l = 0; l < kMaxLeaves; ++l) {
                double c = 0.0;
                if (t < p.n_terms && l < p.n_leaves && ((p.term_mask[t] >> l) & 1u)) {
                    c = 1.0;
                    for (int l2 = 0; l2 < p.n_leaves; ++l2)
                        if (l2 != l && ((p.term_mask[t] >> l2) & 1u) &&
                            (p.leaf_kind[l2] == kLinear || p.leaf_kind[l2] == kConstant))
                            c *= u[l2];
                }
                cex[t][l] = c;
            }
    }
    __syncthreads();

    const double scale = a.likelihood == 1 ? 1.0 / a.extra[(size_t)s * 4 + 3] : 1.0;
    const double *al = a.alpha + (size_t)s * a.n;
    const double *nk = a.NK + (size_t)s * a.np * a.np;
    const int r = threadIdx.x % TB, cg = threadIdx.x / TB;
    const int i = ti * TB + r;
    double acc[kGradChunk];
#pragma unroll
    for (int k = 0; k < kGradChunk; ++k) acc[k] = 0.0;

    for (int m = 0; m < CPT; ++m) {
        const int jc = cg + CG * m;
        const int j = tj * TB + jc;
        if (i < j || i >= a.n || j >= a.n) continue;
        // exponents of the stationary leaves
        double e[MAXL];
#pragma unroll
        for (int l = 0; l < MAXL; ++l) e[l] = 0.0;
#pragma unroll
        for (int l = 0; l < MAXL; ++l) {
            if (l >= p.n_leaves) break;
            const int kind = p.leaf_kind[l];
            if (kind == kLinear || kind == kConstant) continue;
            const int doff = p.leaf_dim_off[l], poff = p.leaf_par_off[l];
            for (int q = 0; q < p.leaf_ndim[l]; ++q) {
                const int code = p.dims[doff + q];
                const int slot = code & 0xffff;
                double f;
                if (kind == kPeriodic) {
                    const double sn = row_f[slot * TB + r] * col_f[(slot + 1) * TB + jc] -
                                      row_f[(slot + 1) * TB + r] * col_f[slot * TB + jc];
                    f = sn * sn;
                } else {
                    double diff = row_f[slot * TB + r] - col_f[slot * TB + jc];
                    if (code >> 16) diff = fmod(fabs(diff) + kPi, kTwoPi) - kPi;
                    f = diff * diff;
                }
                e[l] = fma(h[poff + q], f, e[l]);
            }
        }
        // G_l (stationary) and H_l (linear / constant) sums over the terms
        double G[MAXL], H[MAXL];
#pragma unroll
        for (int l = 0; l < MAXL; ++l) { G[l] = 0.0; H[l] = 0.0; }
        for (int t = 0; t < p.n_terms; ++t) {
            const uint32_t mask = p.term_mask[t];
            double ex = 0.0;
#pragma unroll
            for (int l = 0; l < MAXL; ++l)
                if ((mask >> l) & 1u) ex += e[l];
            const double X = exp_nonpos(ex, a.expc);
            const double T = h[p.n_par_kernel + t] * X;
#pragma unroll
            for (int l = 0; l < MAXL; ++l) {
                if (!((mask >> l) & 1u)) continue;
                G[l] += T;
                H[l] = fma(cex[t][l], X, H[l]);
            }
        }
        const double W = scale * al[i] * al[j] + nk[(size_t)j * a.np + i];
        const double ww = (i == j ? 0.5 : 1.0) * W;
#pragma unroll
        for (int k = 0; k < kGradChunk; ++k) {
            const int l = pleaf[k];
            if (l == -2) continue;
            double d;
            if (l == -1) {
                d = (i == j) ? 1.0 : 0.0;
            } else {
                const int kind = p.leaf_kind[l];
                if (kind == kLinear || kind == kConstant) {
                    d = sel<MAXL>(H, l);
                } else {
                    const int code = p.dims[p.leaf_dim_off[l] + pq[k]];
                    const int slot = code & 0xffff;
                    double f;
                    if (kind == kPeriodic) {
                        const double sn = row_f[slot * TB + r] * col_f[(slot + 1) * TB + jc] -
                                          row_f[(slot + 1) * TB + r] * col_f[slot * TB + jc];
                        f = sn * sn;
                    } else {
                        double diff = row_f[slot * TB + r] - col_f[slot * TB + jc];
                        if (code >> 16) diff = fmod(fabs(diff) + kPi, kTwoPi) - kPi;
                        f = diff * diff;
                    }
                    d = sel<MAXL>(G, l) * f;
                }
            }
            acc[k] = fma(ww, d, acc[k]);
        }
    }
    // work-group sums, one partial per tile and parameter
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kGradChunk; ++k) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kGradChunk && p0 + (int)threadIdx.x < a.P) {
        const int k = threadIdx.x;
        a.partial[((size_t)s * a.ntiles + b) * a.P + p0 + k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
    }
}

// ---- 5. fixed-order sums over tiles, times dh/dtheta or du/dtheta ----------

__global__ void grad_reduce_kernel(const double *__restrict__ partial, int ntiles, int P,
                                   const KernelProgram *__restrict__ prog, const double *__restrict__ theta,
                                   const int *__restrict__ info, double *__restrict__ grad)
{
    const int s = blockIdx.x;
    const KernelProgram &p = *prog;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    // one wave per parameter: lane k sums tiles k, k + 64, ... in order, then
    // a fixed butterfly over the 64 lanes
    for (int q = wave; q < P; q += nw) {
        if (info[s] != 0) {
            if (lane == 0) grad[(size_t)s * P + q] = __builtin_nan("");
            continue;
        }
        double sum = 0.0;
        for (int t = lane; t < ntiles; t += 64) sum += partial[((size_t)s * ntiles + t) * P + q];
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
        if (lane != 0) continue;
        double mult = 1.0;   // nugget and constant leaves
        const double th = theta[(size_t)s * p.n_par + q];
        for (int l = 0; l < p.n_leaves; ++l) {
            const int off = p.leaf_par_off[l];
            const int np_l = (p.leaf_kind[l] == kConstant) ? 1 : p.leaf_ndim[l];
            if (q < off || q >= off + np_l || q >= p.n_par_kernel) continue;
            switch (p.leaf_kind[l]) {
            case kRbf:
            case kRbfCyclic: mult = -0.5 * exp_cr(th); break;                // dh/dtheta = h
            case kPeriodic: { const double r = 1.0 / exp_cr(th); mult = 4.0 * (r * r); break; }   // -2 h
            case kLinear: mult = 1.0 / exp_cr(-th); break;                  // du/dtheta = exp(theta)
            default: mult = 1.0;
            }
        }
        grad[(size_t)s * P + q] = mult * sum;
    }
}

// ---- launcher --------------------------------------------------------------

void launch_gradient(const SwarmBuffers &b, const GradWork &w, double *grad, hipStream_t st)
{
    const int nbn = ceil_div(b.n, 64);
    const int TB = gradient_tile(b.n_slots);
    const int nbt = ceil_div(b.n, TB);
    const size_t mat = (size_t)b.np * b.np;
    (void)mat;
    const int nbp = b.np / 64;
    hipLaunchKernelGGL(uinv_diag_kernel, dim3(nbp, b.S), dim3(kThreads), 0, st, b.K, w.U, w.X, b.np, b.dinv, b.info);
    for (int m = 64; m < b.np; m *= 2) {
        const int pairs = ceil_div(b.np, 2 * m), tiles = pairs * (m / 64) * (m / 64);
        hipLaunchKernelGGL(uinv_t_kernel, dim3(tiles, b.S), dim3(kThreads), 0, st, b.K, w.U, w.NK, b.np, m, b.info);
        hipLaunchKernelGGL(uinv_ux_kernel, dim3(2 * tiles, b.S), dim3(kThreads), 0, st, w.U, w.X, w.NK, b.np, m,
                           b.info);
    }
    if (b.np > b.n)
        hipLaunchKernelGGL(zero_tail_columns_kernel, dim3(1024), dim3(256), 0, st, w.U, b.np, b.n, b.S);
    hipLaunchKernelGGL(ualpha_kernel, dim3(nbn, b.S), dim3(kThreads), 0, st, w.U, b.np, b.n, b.zbuf, b.info,
                       w.alpha);
    hipLaunchKernelGGL(nkinv_kernel, dim3(nbn * (nbn + 1) / 2, b.S), dim3(kThreads), 0, st, w.U, w.NK, b.np, nbn,
                       b.info);
    const int ntiles = nbt * (nbt + 1) / 2;
    GradArgs g;
    g.n = b.n; g.np = b.np; g.S = b.S; g.n_slots = b.n_slots; g.P = w.P; g.ntiles = ntiles; g.pass0 = 0;
    g.slots = b.slots; g.hyp = b.hyp; g.theta = b.theta; g.prog = b.prog; g.alpha = w.alpha; g.NK = w.NK;
    g.extra = b.extra; g.info = b.info; g.likelihood = b.likelihood; g.partial = w.partial;
    exp_constants(g.expc);
    const int passes = ceil_div(w.P, kGradChunk);
    const size_t lds = 2 * (size_t)b.n_slots * TB * sizeof(double);
    if (TB == 64)
        hipLaunchKernelGGL((grad_pair_kernel<64, kMaxLeaves>), dim3(ntiles, b.S, passes), dim3(kThreads), lds, st, g);
    else
        hipLaunchKernelGGL((grad_pair_kernel<32, kMaxLeaves>), dim3(ntiles, b.S, passes), dim3(kThreads), lds, st, g);
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(b.S), dim3(256), 0, st, w.partial, ntiles, w.P, b.prog, b.theta,
                       b.info, grad);
}

// Tile order of the fused pass: 64, or 32 when the staged features of a
// 64-point tile pair would not fit in 64 KB of LDS (as the build chooses).
int gradient_tile(int n_slots) { return 2 * n_slots * 64 * (int)sizeof(double) <= 64 * 1024 ? 64 : 32; }

size_t gradient_partial_count(int n, int n_slots, int S, int P)
{
    const int nbt = ceil_div(n, gradient_tile(n_slots));
    return (size_t)S * (nbt * (nbt + 1) / 2) * P;
}

}  // namespace fb
