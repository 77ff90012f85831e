// fb_build.hip -- covariance build for every particle of the swarm.
//
// Replaces, per particle, compute_R (src/model/gpr/gpr.f90:185-189) with the
// kernel builders it calls: R_stationary (src/kernels/stationary_kernel.f90:101-122),
// R_per (src/kernels/periodic_kernel.f90:181-206), the cyclic R_stationary
// (src/kernels/stationary_cyclic_kernel.f90:104-127), R_interface for linear and
// constant kernels (src/kernels/kernel_module.f90:208-219), R_sum / R_prod
// (kernel_module.f90:309-325, 359-375) and the nugget K += nugget*I.
//
// One work-group owns a TB x TB tile of the lower triangle of the padded
// matrix and SP particles.  The tile's per-point features (raw x, or
// sin(x/2) / cos(x/2) for periodic dimensions) are staged once in LDS and
// shared by the SP particles, and so are the particles' hyperparameters
// (broadcast reads).  Lanes of a wave own consecutive rows, so every store
// of a column segment is one contiguous 8*TB-byte write.
#include "fb_common.h"

#include <cstdlib>
#include <string>

namespace fb {

// Per-particle hyperparameter record, written by hyper_kernel:
//   [0, n_par_kernel)            exponent weights of stationary leaves
//                                 rbf / rbf-cyclic: -1/2 exp(theta)
//                                 periodic: -2 (1/exp(theta))^2
//   [n_par_kernel, +kMaxTerms)   coefficient of each sum-of-products term
//   [n_par_kernel + kMaxTerms]   nugget
__host__ __device__ inline int hyper_stride(const KernelProgram &p) {
    return p.n_par_kernel + kMaxTerms + 1;
}

__global__ void __launch_bounds__(64) hyper_kernel(const KernelProgram *__restrict__ prog,
                                                  const double *__restrict__ theta, int S,
                                                  double *__restrict__ hyp, double *__restrict__ hypT, int ldT)
{
    const int s = blockIdx.x;
    if (s >= S) return;
    const KernelProgram &p = *prog;
    const double *th = theta + (size_t)s * p.n_par;
    double *h = hyp + (size_t)s * hyper_stride(p);
    for (int l = 0; l < p.n_leaves; ++l) {
        const int kind = p.leaf_kind[l];
        const int off = p.leaf_par_off[l];
        const int nd = p.leaf_ndim[l];
        if (kind == kRbf || kind == kRbfCyclic) {
            // rbf_kernel.f90:113-119, rbf_cyclic_kernel.f90:109-115: l = exp(theta)
            for (int q = threadIdx.x; q < nd; q += blockDim.x) h[off + q] = -0.5 * exp_cr(th[off + q]);
        } else if (kind == kPeriodic) {
            // periodic_kernel.f90:111-118: l = 1/exp(theta); l = l*l
            for (int q = threadIdx.x; q < nd; q += blockDim.x) {
                double r = 1.0 / exp_cr(th[off + q]);
                h[off + q] = -2.0 * (r * r);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double leafval[kMaxLeaves];
        for (int l = 0; l < kMaxLeaves; ++l) leafval[l] = 1.0;
        for (int l = 0; l < p.n_leaves; ++l) {
            const int kind = p.leaf_kind[l];
            const int off = p.leaf_par_off[l];
            if (kind == kLinear) {
                // linear_kernel.f90:70-76, order 0: sum(1/c), c = exp(-theta)
                double t = 0.0;
                for (int q = 0; q < p.leaf_ndim[l]; ++q) t += 1.0 / exp_cr(-th[off + q]);
                leafval[l] = t;
            } else if (kind == kConstant) {
                leafval[l] = th[off];  // constant.f90:86-90
            }
        }
        for (int t = 0; t < kMaxTerms; ++t) {
            double c = 1.0;
            if (t < p.n_terms)
                for (int l = 0; l < p.n_leaves; ++l)
                    if (((p.term_mask[t] >> l) & 1u) && (p.leaf_kind[l] == kLinear || p.leaf_kind[l] == kConstant))
                        c *= leafval[l];
            h[p.n_par_kernel + t] = c;
        }
        // gpr.f90:191-202: nugget from theta when optimise_noise
        h[p.n_par_kernel + kMaxTerms] = p.optimise_noise ? th[p.n_par_kernel] : p.noise;
    }
    if (hypT) {
        // the transposed copy (read by build_wide_kernel); the last particle
        // also fills the 16 padding columns
        __syncthreads();
        const int hs = hyper_stride(p);
        const int last = s == S - 1 ? 16 : 0;
        for (int k = threadIdx.x; k < hs; k += blockDim.x) {
            const double v = h[k];
            for (int c = 0; c <= last; ++c) hypT[(size_t)k * ldT + s + c] = v;
        }
    }
}

// Wrap of distance_matrix.f90:119-120; only its square is used, so the sign
// factor of the reference drops out.
__device__ inline double cyclic_wrap_abs(double diff) {
    return fmod(fabs(diff) + kPi, kTwoPi) - kPi;
}

// Right-hand sides bordered below the matrix: y - m, and the ones column of
// the concentrated likelihood.
constexpr int kMaxRhs = 2;

struct BuildArgs {
    int n, np, nrhs, S, n_slots;
    const double *slots;   // n_slots x n: feature value of point i at slots[slot*n + i]
    const double *rhs;     // nrhs x n
    const double *hyp;
    const double *hypT;    // transposed records (build_wide_kernel), leading dimension ldT
    int ldT;
    double *K;
    int *info;
    int *chain;            // words of the chained factorisation schedule, reset here (or none)
    // by value: in the kernel-argument segment every field is read with
    // scalar loads, so the loop bounds, dimension codes and hyperparameter
    // addresses derived from it stay uniform (scalar) in the inner loops
    KernelProgram prog;
    double expc[kExpConsts];   // constants of exp_nonpos
};

// e[s] += h[s] f for the SP particles of a work-group (h: SP consecutive
// doubles in LDS, one parameter of every particle).
template <int SP>
__device__ inline void accum_dim(double (&e)[SP], const double *__restrict__ h, double f)
{
#pragma unroll
    for (int s = 0; s < SP; s += 2) {
        const double2 hv = *reinterpret_cast<const double2 *>(h + s);
        e[s] = fma(hv.x, f, e[s]);
        e[s + 1] = fma(hv.y, f, e[s + 1]);
    }
}

// TB: tile order; MAXL: leaves held in registers (>= n_leaves); SP: particles
// per work-group.
template <int TB, int MAXL, int SP>
__global__ void __launch_bounds__(kThreads) build_kernel(BuildArgs a)
{
    static_assert(SP % 2 == 0, "particles are read in pairs");
    constexpr int CG = kThreads / TB;   // column groups
    constexpr int CPT = TB / CG;        // columns per thread
    extern __shared__ double lds[];
    double *row_f = lds;                        // [n_slots][TB]
    double *col_f = lds + (size_t)a.n_slots * TB;
    // hyperparameter k of particle s0 + s at hyp_t[k * SP + s] (broadcast
    // reads, two particles per read); particles past the swarm repeat the
    // last one, so every particle of the work-group is computed without a
    // branch and only the stores are masked
    double *hyp_t = lds + 2 * (size_t)a.n_slots * TB;

    // lower-triangular tile index: b = ti*(ti+1)/2 + tj, tj <= ti
    const int b = blockIdx.x;
    int ti = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
    while (ti * (ti + 1) / 2 > b) --ti;
    const int tj = b - ti * (ti + 1) / 2;
    const int s0 = blockIdx.y * SP;
    const int nsp = min(SP, a.S - s0);
    const KernelProgram &p = a.prog;
    const int hs = hyper_stride(p);
    // the right-hand sides of the tile's columns (bordered rows n ..
    // n + nrhs - 1), staged with the features: a global load inside the
    // column loop would make every column wait until the stores of the
    // columns before it had completed (loads and stores share one in-order
    // counter)
    double *rhs_l = hyp_t + (size_t)SP * hs;

    for (int e = threadIdx.x; e < a.n_slots * TB; e += kThreads) {
        const int slot = e / TB, r = e % TB;
        const int gi = ti * TB + r, gj = tj * TB + r;
        row_f[e] = gi < a.n ? a.slots[(size_t)slot * a.n + gi] : 0.0;
        col_f[e] = gj < a.n ? a.slots[(size_t)slot * a.n + gj] : 0.0;
    }
    for (int e = threadIdx.x; e < a.nrhs * TB; e += kThreads) {
        const int gj = tj * TB + e % TB;
        rhs_l[e] = gj < a.n ? a.rhs[(size_t)(e / TB) * a.n + gj] : 0.0;
    }
    for (int e = threadIdx.x; e < SP * hs; e += kThreads) {
        const int k = e / SP, sp = min(s0 + e % SP, a.S - 1);
        hyp_t[e] = a.hyp[(size_t)sp * hs + k];
    }
    if (b == 0 && threadIdx.x < nsp) {
        a.info[s0 + threadIdx.x] = 0;
        if (a.chain) {
            int *w = a.chain + (size_t)(s0 + threadIdx.x) * kChainStride;
            w[kChainF0] = 0; w[kChainF1] = 0; w[kChainXccB] = 0; w[kChainXccC] = 0; w[kChainF2] = 0;
            w[kChainS0] = 0; w[kChainS1] = 0;
        }
    }
    __syncthreads();

    const int r = threadIdx.x % TB;
    const int cg = threadIdx.x / TB;
    const int i = ti * TB + r;
    const size_t mstride = (size_t)a.np * a.np;
    const double *hterm = hyp_t + (size_t)p.n_par_kernel * SP;

    for (int m = 0; m < CPT; ++m) {
        const int jc = cg + CG * m;
        const int j = tj * TB + jc;
        if (i < j) continue;  // strictly upper part of a diagonal tile
        double *kcol = a.K + (size_t)j * a.np + i + (size_t)s0 * mstride;
        if (i < a.n && j < a.n) {
            double e[MAXL][SP];
#pragma unroll
            for (int l = 0; l < MAXL; ++l)
#pragma unroll
                for (int s = 0; s < SP; ++s) e[l][s] = 0.0;
            // exponents of the stationary leaves, dimension by dimension
            // (the loops are split by leaf kind so that none branches inside)
#pragma unroll
            for (int l = 0; l < MAXL; ++l) {
                if (l >= p.n_leaves) break;
                const int kind = p.leaf_kind[l];
                const int nd = p.leaf_ndim[l];
                const int *dims = p.dims + p.leaf_dim_off[l];
                const double *hl = hyp_t + (size_t)p.leaf_par_off[l] * SP;
                if (kind == kPeriodic) {
                    // sin((xi-xj)/2) = sin(xi/2)cos(xj/2) - cos(xi/2)sin(xj/2)
#pragma unroll 2
                    for (int q = 0; q < nd; ++q) {
                        const int slot = dims[q] & 0xffff;
                        const double sh_i = row_f[slot * TB + r], ch_i = row_f[(slot + 1) * TB + r];
                        const double sh_j = col_f[slot * TB + jc], ch_j = col_f[(slot + 1) * TB + jc];
                        const double sn = sh_i * ch_j - ch_i * sh_j;
                        accum_dim<SP>(e[l], hl + q * SP, sn * sn);
                    }
                } else if (kind == kRbf) {
#pragma unroll 4
                    for (int q = 0; q < nd; ++q) {
                        const int slot = dims[q] & 0xffff;
                        const double diff = row_f[slot * TB + r] - col_f[slot * TB + jc];
                        accum_dim<SP>(e[l], hl + q * SP, diff * diff);
                    }
                } else if (kind == kRbfCyclic) {
                    for (int q = 0; q < nd; ++q) {
                        const int code = dims[q];
                        const int slot = code & 0xffff;
                        double diff = row_f[slot * TB + r] - col_f[slot * TB + jc];
                        if (code >> 16) diff = cyclic_wrap_abs(diff);
                        accum_dim<SP>(e[l], hl + q * SP, diff * diff);
                    }
                }
            }
            // sum over terms of coef * exp(sum of the term's exponents); a
            // term of linear / constant leaves only has exponent 0
            double v[SP];
#pragma unroll
            for (int s = 0; s < SP; ++s) v[s] = 0.0;
            for (int t = 0; t < p.n_terms; ++t) {
                const uint32_t mask = p.term_mask[t];
                const double *ct = hterm + t * SP;
#pragma unroll
                for (int s = 0; s < SP; ++s) {
                    double ex = 0.0;
#pragma unroll
                    for (int l = 0; l < MAXL; ++l)
                        if ((mask >> l) & 1u) ex += e[l][s];
                    v[s] += ct[s] * exp_nonpos(ex, a.expc);
                }
            }
            if (i == j) {
                const double *nug = hterm + kMaxTerms * SP;
#pragma unroll
                for (int s = 0; s < SP; ++s) v[s] += nug[s];
            }
#pragma unroll
            for (int s = 0; s < SP; ++s)
                if (s < nsp) kcol[s * mstride] = v[s];
        } else {
            // bordered rows (right-hand sides) and the identity padding block
            double v = 0.0;
            if (j < a.n) {
                if (i < a.n + a.nrhs) v = rhs_l[(i - a.n) * TB + jc];
            } else if (i == j) {
                v = 1.0;
            }
#pragma unroll
            for (int s = 0; s < SP; ++s)
                if (s < nsp) kcol[s * mstride] = v;
        }
    }
}

// Read-only data loaded through the scalar unit: a pointer into the constant
// address space makes the compiler load values whose address is uniform
// across the wave with scalar loads, into scalar registers, where they are
// direct operands of the vector FMAs.  build_wide_kernel never writes what
// it reads this way.
typedef const __attribute__((address_space(4))) double sdouble;

// The covariance build for many feature dimensions (the staged feature tiles
// of build_kernel do not fit 64 x 64 tiles in LDS).  There the dimension
// loop is bound by LDS reads: every (pair, dimension) reads the SP
// particles' hyperparameters from LDS again, four times the traffic the LDS
// serves while the vector unit does the arithmetic.  Here the work-group
// owns a 64 x 64 tile (lane = row, each wave 16 columns, CB at a time) and
// SP particles; only the row features are staged in LDS (per-lane reads),
// while the column features (one value per column and dimension) and the
// hyperparameters (one per particle and dimension) are wave-uniform and
// come through scalar loads -- each hyperparameter read serves CB columns,
// each column value SP particles.  The arithmetic per element is that of
// build_kernel (dimensions in the same order, the same expressions), so both
// kernels give the same bits.
template <int MAXL, int SP, int CB>
__global__ void __launch_bounds__(kThreads) build_wide_kernel(BuildArgs a)
{
    constexpr int TB = 64, CPW = TB / 4;   // columns per wave
    static_assert(CPW % CB == 0, "column batches");
    extern __shared__ double row_f[];      // [n_slots][TB]

    const int b = blockIdx.x;
    int ti = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
    while (ti * (ti + 1) / 2 > b) --ti;
    const int tj = b - ti * (ti + 1) / 2;
    const int s0 = blockIdx.y * SP;
    const int nsp = min(SP, a.S - s0);
    const KernelProgram &p = a.prog;

    // the right-hand sides of the tile's columns, staged as in build_kernel
    double *rhs_l = row_f + (size_t)a.n_slots * TB;
    for (int e = threadIdx.x; e < a.n_slots * TB; e += kThreads) {
        const int slot = e / TB, r = e % TB;
        const int gi = ti * TB + r;
        row_f[e] = gi < a.n ? a.slots[(size_t)slot * a.n + gi] : 0.0;
    }
    for (int e = threadIdx.x; e < a.nrhs * TB; e += kThreads) {
        const int gj = tj * TB + e % TB;
        rhs_l[e] = gj < a.n ? a.rhs[(size_t)(e / TB) * a.n + gj] : 0.0;
    }
    if (b == 0 && threadIdx.x < nsp) {
        a.info[s0 + threadIdx.x] = 0;
        if (a.chain) {
            int *w = a.chain + (size_t)(s0 + threadIdx.x) * kChainStride;
            w[kChainF0] = 0; w[kChainF1] = 0; w[kChainXccB] = 0; w[kChainXccC] = 0; w[kChainF2] = 0;
            w[kChainS0] = 0; w[kChainS1] = 0;
        }
    }
    __syncthreads();

    // the wave index, known to the compiler as uniform (scalar column loads)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = ti * TB + lane;
    const size_t mstride = (size_t)a.np * a.np;
    // slots is padded past its end, so the column reads of a batch are
    // contiguous and need no clamp (columns beyond n only feed elements the
    // stores below replace)
    sdouble *sl = (sdouble *)a.slots;
    // entry k of the work-group's particles at ht[k * ldT + s] (past the
    // swarm: the last particle again, computed and not stored)
    sdouble *ht = (sdouble *)a.hypT + s0;
    const size_t ldT = a.ldT;

    for (int cb = 0; cb < CPW; cb += CB) {
        const int j0 = tj * TB + wave * CPW + cb;
        if (ti == tj && j0 > ti * TB + TB - 1) break;
        double e[MAXL][CB][SP];
#pragma unroll
        for (int l = 0; l < MAXL; ++l)
#pragma unroll
            for (int c = 0; c < CB; ++c)
#pragma unroll
                for (int s = 0; s < SP; ++s) e[l][c][s] = 0.0;
#pragma unroll
        for (int l = 0; l < MAXL; ++l) {
            if (l >= p.n_leaves) break;
            const int kind = p.leaf_kind[l];
            const int nd = p.leaf_ndim[l];
            const int *dims = p.dims + p.leaf_dim_off[l];
            const int off = p.leaf_par_off[l];
            if (kind == kPeriodic) {
#pragma unroll 2
                for (int q = 0; q < nd; ++q) {
                    const int slot = dims[q] & 0xffff;
                    const double sh_i = row_f[slot * TB + lane], ch_i = row_f[(slot + 1) * TB + lane];
                    double f[CB];
#pragma unroll
                    for (int c = 0; c < CB; ++c) {
                        const double sh_j = sl[(size_t)slot * a.n + j0 + c], ch_j = sl[(size_t)(slot + 1) * a.n + j0 + c];
                        const double sn = sh_i * ch_j - ch_i * sh_j;
                        f[c] = sn * sn;
                    }
#pragma unroll
                    for (int s = 0; s < SP; ++s) {
                        const double h = ht[(size_t)(off + q) * ldT + s];
#pragma unroll
                        for (int c = 0; c < CB; ++c) e[l][c][s] = fma(h, f[c], e[l][c][s]);
                    }
                }
            } else if (kind == kRbf) {
#pragma unroll 2
                for (int q = 0; q < nd; ++q) {
                    const int slot = dims[q] & 0xffff;
                    const double xi = row_f[slot * TB + lane];
                    double f[CB];
#pragma unroll
                    for (int c = 0; c < CB; ++c) {
                        const double diff = xi - sl[(size_t)slot * a.n + j0 + c];
                        f[c] = diff * diff;
                    }
#pragma unroll
                    for (int s = 0; s < SP; ++s) {
                        const double h = ht[(size_t)(off + q) * ldT + s];
#pragma unroll
                        for (int c = 0; c < CB; ++c) e[l][c][s] = fma(h, f[c], e[l][c][s]);
                    }
                }
            } else if (kind == kRbfCyclic) {
                for (int q = 0; q < nd; ++q) {
                    const int code = dims[q];
                    const int slot = code & 0xffff;
                    const double xi = row_f[slot * TB + lane];
                    double f[CB];
#pragma unroll
                    for (int c = 0; c < CB; ++c) {
                        double diff = xi - sl[(size_t)slot * a.n + j0 + c];
                        if (code >> 16) diff = cyclic_wrap_abs(diff);
                        f[c] = diff * diff;
                    }
#pragma unroll
                    for (int s = 0; s < SP; ++s) {
                        const double h = ht[(size_t)(off + q) * ldT + s];
#pragma unroll
                        for (int c = 0; c < CB; ++c) e[l][c][s] = fma(h, f[c], e[l][c][s]);
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            const int j = j0 + c;
            if (i < j) continue;   // strictly upper part of a diagonal tile
            double *kcol = a.K + (size_t)j * a.np + i + (size_t)s0 * mstride;
            if (i < a.n && j < a.n) {
                double v[SP];
#pragma unroll
                for (int s = 0; s < SP; ++s) v[s] = 0.0;
                for (int t = 0; t < p.n_terms; ++t) {
                    const uint32_t mask = p.term_mask[t];
#pragma unroll
                    for (int s = 0; s < SP; ++s) {
                        double ex = 0.0;
#pragma unroll
                        for (int l = 0; l < MAXL; ++l)
                            if ((mask >> l) & 1u) ex += e[l][c][s];
                        v[s] += ht[(size_t)(p.n_par_kernel + t) * ldT + s] * exp_nonpos(ex, a.expc);
                    }
                }
                if (i == j) {
#pragma unroll
                    for (int s = 0; s < SP; ++s) v[s] += ht[(size_t)(p.n_par_kernel + kMaxTerms) * ldT + s];
                }
#pragma unroll
                for (int s = 0; s < SP; ++s)
                    if (s < nsp) kcol[s * mstride] = v[s];
            } else {
                double v = 0.0;
                if (j < a.n) {
                    if (i < a.n + a.nrhs) v = rhs_l[(i - a.n) * TB + (j - tj * TB)];
                } else if (i == j) {
                    v = 1.0;
                }
#pragma unroll
                for (int s = 0; s < SP; ++s)
                    if (s < nsp) kcol[s * mstride] = v;
            }
        }
    }
}

// Dynamic LDS of a build work-group: features of a row and a column tile, and
// the hyperparameter records of its SP particles.
static size_t build_lds_bytes(int TB, int n_slots, int SP, const KernelProgram &p)
{
    return ((size_t)2 * n_slots * TB + (size_t)SP * hyper_stride(p) + (size_t)kMaxRhs * TB) * sizeof(double);
}

// Dynamic LDS of a build_wide_kernel work-group: the row features of its
// 64-row tile and the right-hand sides of its columns.
static size_t wide_lds_bytes(int n_slots)
{
    return ((size_t)n_slots * 64 + (size_t)kMaxRhs * 64) * sizeof(double);
}

template <int TB, int MAXL, int SP>
static void launch_tb(const BuildArgs &a, int n_slots, hipStream_t st)
{
    const int nt = a.np / TB;
    dim3 grid(nt * (nt + 1) / 2, ceil_div(a.S, SP));
    const size_t shm = build_lds_bytes(TB, n_slots, SP, a.prog);
    if (shm > 65536)   // beyond the default limit: gfx950 work-groups may use up to 160 KB
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&build_kernel<TB, MAXL, SP>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    hipLaunchKernelGGL((build_kernel<TB, MAXL, SP>), grid, dim3(kThreads), shm, st, a);
}

// Particles per work-group: 8, or 10 when that pads the swarm less (the
// padding particles are computed and not stored): 50 -> 10, 20 -> 10,
// 40 -> 8, 64 -> 8.
static int build_sp(int S)
{
    return round_up(S, 10) - S < round_up(S, 8) - S ? 10 : 8;
}

template <int TB, int SP>
static void launch_maxl(const BuildArgs &a, int n_leaves, hipStream_t st)
{
    if (n_leaves <= 1) launch_tb<TB, 1, SP>(a, a.n_slots, st);
    else if (n_leaves <= 2) launch_tb<TB, 2, SP>(a, a.n_slots, st);
    else if (n_leaves <= 4) launch_tb<TB, 4, SP>(a, a.n_slots, st);
    else launch_tb<TB, 8, SP>(a, a.n_slots, st);
}

template <int MAXL, int SP, int CB>
static void launch_wide_cb(const BuildArgs &a, hipStream_t st)
{
    const int nt = a.np / 64;
    dim3 grid(nt * (nt + 1) / 2, ceil_div(a.S, SP));
    const size_t shm = wide_lds_bytes(a.n_slots);
    if (shm > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&build_wide_kernel<MAXL, SP, CB>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    hipLaunchKernelGGL((build_wide_kernel<MAXL, SP, CB>), grid, dim3(kThreads), shm, st, a);
}

// Columns per batch: as many as keep the exponent accumulators
// (leaves x columns x particles) within 80 doubles.
template <int SP>
static void launch_wide(const BuildArgs &a, int n_leaves, hipStream_t st)
{
    if (n_leaves <= 1) launch_wide_cb<1, SP, 4>(a, st);
    else if (n_leaves <= 2) launch_wide_cb<2, SP, 4>(a, st);
    else if (n_leaves <= 4) launch_wide_cb<4, SP, 2>(a, st);
    else launch_wide_cb<8, SP, 1>(a, st);
}

// The build kernel for this program: the LDS-staged one (64 x 64 tiles)
// while a tile's features and hyperparameters fit 64 KB of LDS, else the
// wide one (row features in LDS, the rest through scalar loads).
// FB_BUILD=narrow / wide pins one (read once).
static bool use_wide_build(const BuildArgs &a)
{
    static int forced = [] {
        const char *env = std::getenv("FB_BUILD");
        if (!env) return -1;
        return std::string(env) == "wide" ? 1 : (std::string(env) == "narrow" ? 0 : -1);
    }();
    if (forced >= 0) return forced == 1 && wide_lds_bytes(a.n_slots) <= 160 * 1024;
    return build_lds_bytes(64, a.n_slots, 10, a.prog) > 65536 && wide_lds_bytes(a.n_slots) <= 160 * 1024;
}

void launch_build(const SwarmBuffers &b, hipStream_t st)
{
    BuildArgs a;
    a.n = b.n; a.np = b.np; a.nrhs = b.nrhs; a.S = b.S; a.n_slots = b.n_slots;
    a.slots = b.slots; a.rhs = b.RHS; a.hyp = b.hyp; a.prog = *b.prog_host; a.K = b.K; a.info = b.info; a.chain = b.chain;
    a.hypT = b.hypT; a.ldT = b.ldT;
    exp_constants(a.expc);
    const bool wide_build = b.hypT && b.ldT >= b.S + 16 && use_wide_build(a);
    hipLaunchKernelGGL(hyper_kernel, dim3(b.S), dim3(64), 0, st, b.prog, b.theta, b.S, b.hyp,
                       wide_build ? b.hypT : nullptr, b.ldT);
    if (wide_build) {
        if (build_sp(b.S) == 10) launch_wide<10>(a, b.n_leaves, st);
        else launch_wide<8>(a, b.n_leaves, st);
        return;
    }
    // 64 x 64 tiles while the staged features fit 64 KB of LDS, else 32 x 32
    const bool wide = build_lds_bytes(64, b.n_slots, 10, a.prog) <= 65536;
    if (build_sp(b.S) == 10) {
        if (wide) launch_maxl<64, 10>(a, b.n_leaves, st);
        else launch_maxl<32, 10>(a, b.n_leaves, st);
    } else {
        if (wide) launch_maxl<64, 8>(a, b.n_leaves, st);
        else launch_maxl<32, 8>(a, b.n_leaves, st);
    }
}

}  // namespace fb
