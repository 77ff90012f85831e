// fb_potrf.hip -- batched blocked FP64 Cholesky of every particle's matrix.
//
// Replaces dpotrf('L') (src/utils/linalg.f90:73-98, called from
// src/model/gpr/likelihood.f90:145) for all particles at once, with LAPACK's
// INFO convention, plus the log-determinant of src/utils/arith.f90:72-93.
//
// Left-looking blocked algorithm with two block levels.  Panels of width
// kNBO = 128 carry the dominant trailing update; inside a panel two blocks of
// kNBI = 64 (b0 = c0, b1 = c0 + 64) are factored.  Three launches per panel:
//   1. C(c0:, c0:c0+128) -= L(c0:, 0:c0) L(c0:c0+128, 0:c0)^T (FP64 MFMA
//      16x16x4, LDS-staged); the first work-groups update the diagonal block
//      b0 on chip and factor it (INFO, partial log-det, 16x16 inverses)
//      while the other tiles are still being updated;
//   2. L(c0+64:, b0) = C L(b0, b0)^{-T} (block forward substitution on the
//      matrix cores); the first work-groups also update block (b1, b1) with
//      their solved tile on chip and factor it;
//   3. C(c0+128:, b1) -= L(c0+128:, b0) L(b1, b0)^T on chip, then solved
//      with L(b1, b1).
// Every step runs the tile routines of fb_tiles.h in the order the
// persistent schedule does, so both paths give the same bits.  The whole
// swarm advances in lock step: every launch covers all particles.
#include "fb_tiles.h"

namespace fb {

// ---- lock-step kernels (three launches per 128-column panel) ---------------

template <int KC, int MB, int NB>
constexpr int panel_smem()
{
    constexpr int a = TileShape<MB, NB, KC>::SMEM, b = TileShape<kMB, kNBI, KC>::SMEM;
    constexpr int m = a > b ? a : b;
    return m > kPotrfSmem ? m : kPotrfSmem;
}

// Launch 1 of panel c0: the trailing update C(c0:, c0:c0+128) -= L(c0:, 0:c0)
// L(c0:c0+128, 0:c0)^T.  Work-groups 0 .. S-1 take the diagonal block of
// each matrix (its lower 64x64 only), keep the updated block on chip and
// factor it (potrf64_block); they come first, so the factorisation runs
// while the other row tiles are still being updated.  The other tiles are
// MB x NB (64 x 128, 32 x 128 or 32 x 64: smaller tiles give small swarms
// more work-groups; every shape accumulates each element in the same
// order, so all give the same bits).
template <int KC, int MB, int NB>
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(KC == 16 ? 2 : 1, KC == 16 ? 2 : 1))) panel_update_factor_kernel(double *__restrict__ K, int np, int n,
                                                                       int c0, int S, int *__restrict__ info,
                                                                       double *__restrict__ logdet_part,
                                                                       double *__restrict__ dinv)
{
    __shared__ __attribute__((aligned(16))) double smem[panel_smem<KC, MB, NB>()];
    const int b = blockIdx.x;
    if (b < S) {
        if (info[b] != 0) return;
        double *A = K + (size_t)b * np * np;
        gemm_tile_update<kMB, kNBI, KC, 2, true>(A, np, c0, c0, 0, c0, smem);
        __syncthreads();
        double lg = 0.0;
        const int fail = potrf64_block<true>(A, np, n, c0, &lg, smem,
                                             dinv + ((size_t)b * (np / kNBI) + c0 / kNBI) * kDinv);
        if (threadIdx.x == 0) {
            if (fail) info[b] = fail;
            else logdet_part[(size_t)b * (np / kNBI) + c0 / kNBI] = lg;
        }
        return;
    }
    constexpr int CT = kNBO / NB;            // column tiles of the panel
    const int rt = (np - c0 - kNBI) / MB;    // row tiles below the diagonal block
    const int t = b - S;
    const int s = t / (rt * CT), u = t % (rt * CT);
    if (info[s] != 0) return;
    gemm_tile_update<MB, NB, KC>(K + (size_t)s * np * np, np, c0 + kNBI + (u / CT) * MB, c0 + (u % CT) * NB, 0, c0,
                                 smem);
}

// Launch 2 of panel c0: every row tile below block b0 = c0 is solved with
// L(b0, b0) (trsm64).  Work-groups 0 .. S-1 take the tile of block
// b1 = c0 + 64, keep the solved tile on chip, apply it to the diagonal block
// (b1, b1) (syrk64_lds) and factor that block.
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
trsm_factor_kernel(double *__restrict__ K, int np, int n, int c0, int S, int *__restrict__ info,
                   double *__restrict__ logdet_part, double *__restrict__ dinv)
{
    __shared__ __attribute__((aligned(16))) double smem[kNBI * kPLD + kPotrfSmem];
    const int b = blockIdx.x;
    const int rt = (np - c0) / kNBI - 1;    // row tiles below block b0
    const int s = b < S ? b : (b - S) / (rt - 1);
    const int r = b < S ? 1 : 2 + (b - S) % (rt - 1);
    if (info[s] != 0) return;
    double *A = K + (size_t)s * np * np;
    const double *d0 = dinv + ((size_t)s * (np / kNBI) + c0 / kNBI) * kDinv;
    const int row0 = c0 + r * kNBI;
    double *tile = A + (size_t)c0 * np + row0;
    if (b >= S) {
        trsm64_solve<false>(tile, np, tile, np, A + (size_t)c0 * np + c0, np, d0);
        return;
    }
    double *Xs = smem;                       // L(b1, b0) on chip
    double *Ls = smem + kNBI * kPLD;         // block (b1, b1), then its factorisation
    trsm64_solve<false, true>(tile, np, tile, np, A + (size_t)c0 * np + c0, np, d0, Xs);
    const int c1 = c0 + kNBI;
    const double *Cg = A + (size_t)c1 * np + c1;
    for (int e = threadIdx.x; e < kNBI * kNBI / 2; e += kThreads) {
        const int col = e / (kNBI / 2), r2 = (e % (kNBI / 2)) * 2;
        *reinterpret_cast<double2 *>(Ls + col * kPLD + r2) = *reinterpret_cast<const double2 *>(Cg + (size_t)col * np + r2);
    }
    __syncthreads();
    syrk64_lds(Ls, Xs);
    __syncthreads();
    double lg = 0.0;
    const int fail = potrf64_block<true>(A, np, n, c1, &lg, Ls,
                                         dinv + ((size_t)s * (np / kNBI) + c1 / kNBI) * kDinv);
    if (threadIdx.x == 0) {
        if (fail) info[s] = fail;
        else logdet_part[(size_t)s * (np / kNBI) + c1 / kNBI] = lg;
    }
}

// Launch 3 of panel c0 (block b1 = c1 = c0 + 64): every row tile below b1
// takes the update C(i, b1) -= L(i, b0) L(b1, b0)^T (64 deep, kept on chip)
// and is solved with L(b1, b1).
template <int KC>
__global__ void __launch_bounds__(kThreads) update_trsm_kernel(double *__restrict__ K, int np, int c1,
                                                               const int *__restrict__ info,
                                                               const double *__restrict__ dinv)
{
    __shared__ __attribute__((aligned(16))) double smem[TileShape<kMB, kNBI, KC>::SMEM];
    const int rt = (np - c1) / kNBI - 1;
    const int s = blockIdx.x / rt;
    const int row0 = c1 + (1 + blockIdx.x % rt) * kNBI;
    if (info[s] != 0) return;
    double *A = K + (size_t)s * np * np;
    gemm_tile_update<kMB, kNBI, KC, 2, true>(A, np, row0, c1, c1 - kNBI, c1, smem);
    __syncthreads();
    trsm64_solve<false>(smem, kPLD, A + (size_t)c1 * np + row0, np, A + (size_t)c1 * np + c1, np,
                        dinv + ((size_t)s * (np / kNBI) + c1 / kNBI) * kDinv);
}

template <int KC, int MB, int NB>
static void panel_launch(const SwarmBuffers &b, int c0, hipStream_t st)
{
    // the first panel has nothing to update: only the diagonal work-groups
    const int tiles = c0 > 0 ? ((b.np - c0 - kNBI) / MB) * (kNBO / NB) : 0;
    hipLaunchKernelGGL((panel_update_factor_kernel<KC, MB, NB>), dim3(b.S * (1 + tiles)), dim3(kThreads), 0, st,
                       b.K, b.np, b.n, c0, b.S, b.info, b.logdet_part, b.dinv);
}

void launch_panel_update_factor(const SwarmBuffers &b, int c0, hipStream_t st)
{
    const int tile = b.panel_tile ? b.panel_tile[c0 / kNBO] : b.tile;
    if (b.gemm_kc == 32) panel_launch<32, 64, 128>(b, c0, st);
    else if (tile == 1) panel_launch<16, 32, 128>(b, c0, st);
    else if (tile == 2) panel_launch<16, 32, 64>(b, c0, st);
    else panel_launch<16, 64, 128>(b, c0, st);
}

void launch_trsm_factor(const SwarmBuffers &b, int c0, hipStream_t st)
{
    const int rt = (b.np - c0) / kNBI - 1;   // >= 1: the panel holds two blocks
    hipLaunchKernelGGL(trsm_factor_kernel, dim3(b.S * rt), dim3(kThreads), 0, st, b.K, b.np, b.n, c0, b.S,
                       b.info, b.logdet_part, b.dinv);
}

void launch_update_trsm(const SwarmBuffers &b, int c1, hipStream_t st)
{
    const int rt = (b.np - c1) / kNBI - 1;
    if (rt <= 0) return;
    if (b.gemm_kc == 16)
        hipLaunchKernelGGL((update_trsm_kernel<16>), dim3(b.S * rt), dim3(kThreads), 0, st, b.K, b.np, c1, b.info,
                           b.dinv);
    else
        hipLaunchKernelGGL((update_trsm_kernel<32>), dim3(b.S * rt), dim3(kThreads), 0, st, b.K, b.np, c1, b.info,
                           b.dinv);
}

}  // namespace fb
