// fb_potrf.hip -- batched blocked FP64 Cholesky of every particle's matrix.
//
// Replaces dpotrf('L') (src/utils/linalg.f90:73-98, called from
// src/model/gpr/likelihood.f90:145) for all particles at once, with LAPACK's
// INFO convention, plus the log-determinant of src/utils/arith.f90:72-93.
//
// Left-looking blocked algorithm with two block levels.  Panels of width
// kNBO = 128 carry the dominant trailing update; inside a panel two blocks of
// kNBI = 64 are factored.  For the panel at column c0:
//   gemm_update  C(c0:, c0:c0+128) -= L(c0:, 0:c0) L(c0:c0+128, 0:c0)^T
//                (FP64 MFMA 16x16x4, LDS-staged, double-buffered; the wide
//                panel halves how often each row panel is re-read from HBM)
//   potrf64      factor the 64x64 block at c0 in LDS (INFO, partial log-det)
//   trsm64       L(c0+64:, c0:c0+64) = C L(c0, c0)^{-T} by forward
//                substitution (backward stable, no explicit inverse)
//   gemm_update  C(c0+64:, c0+64:c0+128) -= L(c0+64:, c0:c0+64) L(...)^T
//   potrf64, trsm64 for the block at c0 + 64.
// The whole swarm advances in lock step: every launch covers all particles.
#include "fb_tiles.h"

namespace fb {

// ---- lock-step kernels (one launch per step, every particle) ---------------

template <int MB, int NB, int KC>
__global__ void __launch_bounds__(kThreads) gemm_update_kernel(double *__restrict__ K, int np, int col0,
                                                               int row_start, int k0, int k1,
                                                               const int *__restrict__ info)
{
    __shared__ __attribute__((aligned(16))) double smem[TileShape<MB, NB, KC>::SMEM];
    const int rt_per = (np - row_start) / MB;
    const int s = blockIdx.x / rt_per;
    const int rt = blockIdx.x % rt_per;
    if (info[s] != 0) return;
    gemm_tile_update<MB, NB, KC>(K + (size_t)s * np * np, np, row_start + rt * MB, col0, k0, k1, smem);
}

__global__ void __launch_bounds__(kThreads) trsm64_kernel(double *__restrict__ K, int np, int c0,
                                                          int row_start, const int *__restrict__ info,
                                                          const double *__restrict__ dinv)
{
    const int rt_per = (np - row_start) / kNBI;
    const int s = blockIdx.x / rt_per;
    const int rt = blockIdx.x % rt_per;
    if (info[s] != 0) return;
    trsm64_mfma(K + (size_t)s * np * np, np, c0, row_start + rt * kNBI,
                dinv + ((size_t)s * (np / kNBI) + c0 / kNBI) * kDinv);
}

__global__ void __launch_bounds__(kThreads) potrf64_kernel(double *__restrict__ K, int np, int n, int c0,
                                                           int *__restrict__ info,
                                                           double *__restrict__ logdet_part,
                                                           double *__restrict__ dinv)
{
    __shared__ __attribute__((aligned(16))) double smem[kPotrfSmem];
    const int s = blockIdx.x;
    if (info[s] != 0) return;
    double lg = 0.0;
    const int fail = potrf64_block(K + (size_t)s * np * np, np, n, c0, &lg, smem,
                                   dinv + ((size_t)s * (np / kNBI) + c0 / kNBI) * kDinv);
    if (threadIdx.x == 0) {
        if (fail) info[s] = fail;
        else logdet_part[(size_t)s * (np / kNBI) + c0 / kNBI] = lg;
    }
}

void launch_gemm_update(const SwarmBuffers &b, int col0, int width, int row_start, int k0, int k1,
                        hipStream_t st)
{
    if (k1 <= k0) return;
    const int rt = (b.np - row_start) / kMB;
    // chunk depth 32 (one work-group per CU, 115 KB of LDS) or 16 (two per
    // CU): the same accumulation order, so the same bits, at different speed
    if (width == kNBO) {
        if (b.gemm_kc == 16)
            hipLaunchKernelGGL((gemm_update_kernel<kMB, kNBO, 16>), dim3(b.S * rt), dim3(kThreads), 0, st,
                               b.K, b.np, col0, row_start, k0, k1, b.info);
        else
            hipLaunchKernelGGL((gemm_update_kernel<kMB, kNBO, 32>), dim3(b.S * rt), dim3(kThreads), 0, st,
                               b.K, b.np, col0, row_start, k0, k1, b.info);
    } else {
        if (b.gemm_kc == 16)
            hipLaunchKernelGGL((gemm_update_kernel<kMB, kNBI, 16>), dim3(b.S * rt), dim3(kThreads), 0, st,
                               b.K, b.np, col0, row_start, k0, k1, b.info);
        else
            hipLaunchKernelGGL((gemm_update_kernel<kMB, kNBI, 32>), dim3(b.S * rt), dim3(kThreads), 0, st,
                               b.K, b.np, col0, row_start, k0, k1, b.info);
    }
}

void launch_potrf64(const SwarmBuffers &b, int c0, hipStream_t st)
{
    hipLaunchKernelGGL(potrf64_kernel, dim3(b.S), dim3(kThreads), 0, st, b.K, b.np, b.n, c0, b.info,
                       b.logdet_part, b.dinv);
}

void launch_trsm64(const SwarmBuffers &b, int c0, int row_start, hipStream_t st)
{
    const int rt = (b.np - row_start) / kNBI;
    if (rt <= 0) return;
    hipLaunchKernelGGL(trsm64_kernel, dim3(b.S * rt), dim3(kThreads), 0, st, b.K, b.np, c0, row_start, b.info,
                       b.dinv);
}

}  // namespace fb
