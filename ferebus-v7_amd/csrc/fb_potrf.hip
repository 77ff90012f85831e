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
#include "fb_common.h"

namespace fb {

typedef double d4 __attribute__((ext_vector_type(4)));

// One 16x16x4 FP64 matrix-core product: D = A B + C where lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[l / 16 + 4 v][l % 16]
// in component v (rows interleaved across the four 16-lane groups; checked
// on gfx950 with a one-wave probe).
__device__ inline d4 mfma_f64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Work-group tile of MB rows x NB columns computed by 4 waves.
template <int MB, int NB>
struct TileShape {
    static constexpr int WAVES_N = (NB / 32 < 4) ? NB / 32 : 4;
    static constexpr int WAVES_M = 4 / WAVES_N;
    static constexpr int WM = MB / WAVES_M;
    static constexpr int WN = NB / WAVES_N;
    static constexpr int BM = WM / 16;
    static constexpr int BN = WN / 16;
    static constexpr int LDA = MB + 16;   // LDS strides in doubles; = 16 mod 32
    static constexpr int LDB = NB + 16;   // keeps the two k rows of a fragment read
    static constexpr int LDC = MB + 4;    // on different LDS banks
    static constexpr int PIPE = 2 * kKC * (LDA + LDB);
    static constexpr int EPI = NB * LDC;
    static constexpr int SMEM = PIPE > EPI ? PIPE : EPI;
    static_assert(WAVES_M * WAVES_N == 4, "4 waves per tile");
    static_assert(BM >= 1 && BN >= 1, "wave tile at least 16x16");
    static_assert((MB * kKC / 2) % kThreads == 0 && (NB * kKC / 2) % kThreads == 0, "loads");
};

// acc = sum_k A[r][k] B[c][k] over k < depth, A[r][k] = Ag[k*lda + r],
// B[c][k] = Bg[k*ldb + c].  depth is a multiple of kKC.  LDS holds two
// stages [A0 A1 B0 B1]; the next stage is fetched into registers while the
// current one feeds the matrix cores.
template <int MB, int NB>
__device__ inline void tile_gemm(const double *__restrict__ Ag, int lda,
                                 const double *__restrict__ Bg, int ldb, int depth,
                                 double *__restrict__ smem,
                                 d4 (&acc)[TileShape<MB, NB>::BM][TileShape<MB, NB>::BN])
{
    using T = TileShape<MB, NB>;
    constexpr int NA = MB * kKC / 2 / kThreads;   // double2 loads per thread (A)
    constexpr int NBL = NB * kKC / 2 / kThreads;  // double2 loads per thread (B)
    constexpr int ASZ = kKC * T::LDA;
    constexpr int BSZ = kKC * T::LDB;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wrow = (wave / T::WAVES_N) * T::WM;
    const int wcol = (wave % T::WAVES_N) * T::WN;

#pragma unroll
    for (int i = 0; i < T::BM; ++i)
#pragma unroll
        for (int j = 0; j < T::BN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    if (depth <= 0) return;

    // per-thread load coordinates (fixed for every stage)
    int a_off[NA], a_lds[NA], b_off[NBL], b_lds[NBL];
#pragma unroll
    for (int u = 0; u < NA; ++u) {
        const int q = tid + u * kThreads;
        const int col = q / (MB / 2), r2 = (q % (MB / 2)) * 2;
        a_off[u] = col * lda + r2;
        a_lds[u] = col * T::LDA + r2;
    }
#pragma unroll
    for (int u = 0; u < NBL; ++u) {
        const int q = tid + u * kThreads;
        const int col = q / (NB / 2), r2 = (q % (NB / 2)) * 2;
        b_off[u] = col * ldb + r2;
        b_lds[u] = col * T::LDB + r2;
    }
    // staging registers as scalar pairs: arrays of double2 are not promoted
    // to registers by the compiler and would live in scratch
    double rax[NA], ray[NA], rbx[NBL], rby[NBL];
#pragma unroll
    for (int u = 0; u < NA; ++u) {
        const double2 t = *reinterpret_cast<const double2 *>(Ag + a_off[u]);
        rax[u] = t.x; ray[u] = t.y;
    }
#pragma unroll
    for (int u = 0; u < NBL; ++u) {
        const double2 t = *reinterpret_cast<const double2 *>(Bg + b_off[u]);
        rbx[u] = t.x; rby[u] = t.y;
    }
#pragma unroll
    for (int u = 0; u < NA; ++u) *reinterpret_cast<double2 *>(smem + a_lds[u]) = double2{rax[u], ray[u]};
#pragma unroll
    for (int u = 0; u < NBL; ++u) *reinterpret_cast<double2 *>(smem + 2 * ASZ + b_lds[u]) = double2{rbx[u], rby[u]};
    __syncthreads();

    const int kq = lane >> 4, lr = lane & 15;
    const int nchunks = depth / kKC;
    for (int c = 0; c < nchunks; ++c) {
        const int cur = c & 1;
        const bool more = c + 1 < nchunks;
        if (more) {
            const size_t ka = (size_t)(c + 1) * kKC * lda, kb = (size_t)(c + 1) * kKC * ldb;
#pragma unroll
            for (int u = 0; u < NA; ++u) {
                const double2 t = *reinterpret_cast<const double2 *>(Ag + ka + a_off[u]);
                rax[u] = t.x; ray[u] = t.y;
            }
#pragma unroll
            for (int u = 0; u < NBL; ++u) {
                const double2 t = *reinterpret_cast<const double2 *>(Bg + kb + b_off[u]);
                rbx[u] = t.x; rby[u] = t.y;
            }
        }
        const double *as = smem + cur * ASZ;
        const double *bs = smem + 2 * ASZ + cur * BSZ;
#pragma unroll
        for (int kk = 0; kk < kKC / 4; ++kk) {
            double af[T::BM], bf[T::BN];
#pragma unroll
            for (int i = 0; i < T::BM; ++i) af[i] = as[(kk * 4 + kq) * T::LDA + wrow + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < T::BN; ++j) bf[j] = bs[(kk * 4 + kq) * T::LDB + wcol + j * 16 + lr];
#pragma unroll
            for (int i = 0; i < T::BM; ++i)
#pragma unroll
                for (int j = 0; j < T::BN; ++j) acc[i][j] = mfma_f64(af[i], bf[j], acc[i][j]);
        }
        if (more) {
            double *an = smem + (cur ^ 1) * ASZ;
            double *bn = smem + 2 * ASZ + (cur ^ 1) * BSZ;
#pragma unroll
            for (int u = 0; u < NA; ++u) *reinterpret_cast<double2 *>(an + a_lds[u]) = double2{rax[u], ray[u]};
#pragma unroll
            for (int u = 0; u < NBL; ++u) *reinterpret_cast<double2 *>(bn + b_lds[u]) = double2{rbx[u], rby[u]};
        }
        __syncthreads();
    }
}

// Writes acc into the LDS tile Cs[col * LDC + row] (tile-local coordinates).
template <int MB, int NB>
__device__ inline void acc_to_lds(double *Cs, d4 (&acc)[TileShape<MB, NB>::BM][TileShape<MB, NB>::BN])
{
    using T = TileShape<MB, NB>;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wrow = (wave / T::WAVES_N) * T::WM;
    const int wcol = (wave % T::WAVES_N) * T::WN;
#pragma unroll
    for (int i = 0; i < T::BM; ++i)
#pragma unroll
        for (int j = 0; j < T::BN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wrow + i * 16 + (lane >> 4) + 4 * v;
                const int col = wcol + j * 16 + (lane & 15);
                Cs[col * T::LDC + row] = acc[i][j][v];
            }
}

// ---- step kernels ---------------------------------------------------------

// C(rows >= row_start, col0:col0+NB) -= L(rows, k0:k1) L(col0:col0+NB, k0:k1)^T
// for every live matrix; one work-group per MB-row tile.
template <int MB, int NB>
__global__ void __launch_bounds__(kThreads) gemm_update_kernel(double *__restrict__ K, int np, int col0,
                                                               int row_start, int k0, int k1,
                                                               const int *__restrict__ info)
{
    using T = TileShape<MB, NB>;
    __shared__ __attribute__((aligned(16))) double smem[T::SMEM];
    const int rt_per = (np - row_start) / MB;
    const int s = blockIdx.x / rt_per;
    const int rt = blockIdx.x % rt_per;
    if (info[s] != 0) return;
    const int row0 = row_start + rt * MB;
    double *A = K + (size_t)s * np * np;
    d4 acc[T::BM][T::BN];
    tile_gemm<MB, NB>(A + (size_t)k0 * np + row0, np, A + (size_t)k0 * np + col0, np, k1 - k0, smem, acc);
    acc_to_lds<MB, NB>(smem, acc);
    __syncthreads();
    double *Ct = A + (size_t)col0 * np + row0;
    for (int q = threadIdx.x; q < MB * NB / 2; q += kThreads) {
        const int col = q / (MB / 2), r2 = (q % (MB / 2)) * 2;
        double2 *g = reinterpret_cast<double2 *>(Ct + (size_t)col * np + r2);
        const double2 c = *reinterpret_cast<const double2 *>(smem + col * T::LDC + r2);
        double2 v = *g;
        v.x -= c.x;
        v.y -= c.y;
        *g = v;
    }
}

// L(rows >= row_start, c0:c0+64) = C L(c0:c0+64, c0:c0+64)^{-T}.
// One wave per 64-row tile (four tiles of one matrix per work-group): lane r
// keeps row r of the tile in registers and solves x L^T = c right-looking,
// x_c = c_c / L(c,c) then c_k -= x_c L(k,c) for k > c, reading L from LDS as
// broadcasts; no barrier after the loads.
__global__ void __launch_bounds__(kThreads) trsm64_kernel(double *__restrict__ K, int np, int c0,
                                                          int row_start, int groups,
                                                          const int *__restrict__ info)
{
    constexpr int B = kNBI;
    __shared__ double Lc[B * B];      // Lc[c * B + k] = L(k, c) (column c of L)
    __shared__ double rdiag[B];
    const int s = blockIdx.x / groups;
    const int grp = blockIdx.x % groups;
    if (info[s] != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *A = K + (size_t)s * np * np;
    const double *Lg = A + (size_t)c0 * np + c0;
#pragma unroll
    for (int u = 0; u < B * B / kThreads; ++u) {
        const int e = tid + u * kThreads;
        const int c = e / B, k = e % B;
        Lc[e] = k >= c ? Lg[(size_t)c * np + k] : 0.0;
    }
    __syncthreads();
    if (tid < B) rdiag[tid] = 1.0 / Lc[tid * B + tid];
    const int ntiles = (np - row_start) / B;
    const int rt = grp * 4 + wave;
    __syncthreads();
    if (rt >= ntiles) return;
    double *Ct = A + (size_t)c0 * np + row_start + (size_t)rt * B + lane;
    double x[B];
#pragma unroll
    for (int k = 0; k < B; ++k) x[k] = Ct[(size_t)k * np];
#pragma unroll
    for (int c = 0; c < B; ++c) {
        x[c] *= rdiag[c];
#pragma unroll
        for (int k = c + 1; k < B; ++k) x[k] -= x[c] * Lc[c * B + k];
    }
#pragma unroll
    for (int k = 0; k < B; ++k) Ct[(size_t)k * np] = x[k];
}

// Factor the 64x64 diagonal block at c0 of every live matrix.  Thread
// (lane i, wave w) keeps row i of columns w, w+4, ..., w+60 in registers;
// the wave owning column c takes the pivot (LAPACK dpotf2 test), scales the
// column and publishes it through LDS, then every thread applies the rank-1
// update to its registers: one barrier per column.  Columns at or beyond n
// (bordered right-hand-side rows and padding) are not factored: they become
// the identity, leaving the right-hand-side rows holding L^{-1} rhs.
__global__ void __launch_bounds__(kThreads) potrf64_kernel(double *__restrict__ K, int np, int n, int c0,
                                                           int *__restrict__ info,
                                                           double *__restrict__ logdet_part)
{
    constexpr int B = kNBI;
    constexpr int M = B / 4;
    __shared__ double col[2][B];
    __shared__ int fail_col;
    __shared__ double red[kThreads / 64];
    const int s = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    if (info[s] != 0) return;
    double *Ag = K + (size_t)s * np * np + (size_t)c0 * np + c0 + lane;
    double a[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int k = wave + 4 * m;
        a[m] = lane >= k ? Ag[(size_t)k * np] : 0.0;
    }
    if (tid == 0) fail_col = 0;
    const int ncol = min(B, n - c0);
    __syncthreads();
    // every thread runs all B steps (the barrier is unconditional); steps at
    // or beyond ncol, or after a failed pivot, do no work
#pragma unroll
    for (int c = 0; c < B; ++c) {
        const bool live = c < ncol && fail_col == 0;
        if (live && wave == (c & 3)) {
            const double piv = __shfl(a[c >> 2], c, 64);
            if (!(piv > 0.0)) {
                // dpotf2: stop at the first pivot that is not positive or is NaN
                if (lane == 0) fail_col = c0 + c + 1;
            } else {
                const double d = sqrt(piv);
                const double l = lane > c ? a[c >> 2] * (1.0 / d) : (lane == c ? d : 0.0);
                if (lane >= c) a[c >> 2] = l;
                col[c & 1][lane] = l;
            }
        }
        __syncthreads();
        if (c < ncol && fail_col == 0) {
            const double li = col[c & 1][lane];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int k = wave + 4 * m;
                if (k > c) a[m] -= li * col[c & 1][k];
            }
        }
    }
    if (fail_col != 0) {
        if (tid == 0) info[s] = fail_col;
        return;
    }
    // identity on the unfactored (right-hand-side / padding) columns, and the
    // partial log-determinant of the real columns of this block
    double dv = 1.0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int k = wave + 4 * m;
        if (k >= ncol) a[m] = lane == k ? 1.0 : 0.0;
        if (k == lane) dv = a[m];
    }
    double lg = ((lane & 3) == wave && lane < ncol) ? log(dv) : 0.0;
    for (int off = 32; off > 0; off >>= 1) lg += __shfl_xor(lg, off, 64);
    if (lane == 0) red[wave] = lg;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int k = wave + 4 * m;
        if (lane >= k) Ag[(size_t)k * np] = a[m];
    }
    __syncthreads();
    if (tid == 0) logdet_part[(size_t)s * (np / B) + c0 / B] = (red[0] + red[1]) + (red[2] + red[3]);
}

void launch_gemm_update(const SwarmBuffers &b, int col0, int width, int row_start, int k0, int k1,
                        hipStream_t st)
{
    if (k1 <= k0) return;
    if (width == kNBO) {
        const int rt = (b.np - row_start) / kMB;
        hipLaunchKernelGGL((gemm_update_kernel<kMB, kNBO>), dim3(b.S * rt), dim3(kThreads), 0, st,
                           b.K, b.np, col0, row_start, k0, k1, b.info);
    } else {
        const int rt = (b.np - row_start) / kMB;
        hipLaunchKernelGGL((gemm_update_kernel<kMB, kNBI>), dim3(b.S * rt), dim3(kThreads), 0, st,
                           b.K, b.np, col0, row_start, k0, k1, b.info);
    }
}

void launch_potrf64(const SwarmBuffers &b, int c0, hipStream_t st)
{
    hipLaunchKernelGGL(potrf64_kernel, dim3(b.S), dim3(kThreads), 0, st, b.K, b.np, b.n, c0, b.info,
                       b.logdet_part);
}

void launch_trsm64(const SwarmBuffers &b, int c0, int row_start, hipStream_t st)
{
    const int rt = (b.np - row_start) / kNBI;
    if (rt <= 0) return;
    const int groups = ceil_div(rt, 4);
    hipLaunchKernelGGL(trsm64_kernel, dim3(b.S * groups), dim3(kThreads), 0, st, b.K, b.np, c0, row_start,
                       groups, b.info);
}

}  // namespace fb
