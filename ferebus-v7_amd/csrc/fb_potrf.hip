// fb_potrf.hip -- batched blocked FP64 Cholesky of every particle's matrix.
//
// Replaces dpotrf('L') (src/utils/linalg.f90:73-98, called from
// src/model/gpr/likelihood.f90:145) for all particles at once, with LAPACK's
// INFO convention, plus the log-determinant of src/utils/arith.f90:72-93.
//
// Left-looking blocked algorithm over block columns of width kNB.  Step j:
//   gemm_step   C(i, j) = A(i, j) - L(i, 0:j) L(j, 0:j)^T for every row tile
//               i >= j of every live matrix (FP64 MFMA 16x16x4, LDS-staged,
//               double-buffered); skipped at j = 0.
//   potrf_diag  factor the diagonal block in LDS (INFO, partial log-det) and
//               form its explicit inverse for the panel solve.
//   trsm_step   L(i, j) = C(i, j) L(j, j)^{-T} for the row tiles below, as one
//               more MFMA tile product with the inverse.
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
// B[c][k] = Bg[k*ldb + c].  depth is a multiple of kKC.
template <int MB, int NB>
__device__ inline void tile_gemm(const double *__restrict__ Ag, int lda,
                                 const double *__restrict__ Bg, int ldb, int depth,
                                 double *smem, d4 (&acc)[TileShape<MB, NB>::BM][TileShape<MB, NB>::BN])
{
    using T = TileShape<MB, NB>;
    constexpr int NA = MB * kKC / 2 / kThreads;   // double2 loads per thread (A)
    constexpr int NBL = NB * kKC / 2 / kThreads;  // double2 loads per thread (B)
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wrow = (wave / T::WAVES_N) * T::WM;
    const int wcol = (wave % T::WAVES_N) * T::WN;
    double *As[2] = {smem, smem + kKC * T::LDA};
    double *Bs[2] = {smem + 2 * kKC * T::LDA, smem + 2 * kKC * T::LDA + kKC * T::LDB};

#pragma unroll
    for (int i = 0; i < T::BM; ++i)
#pragma unroll
        for (int j = 0; j < T::BN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    if (depth <= 0) return;

    double2 ra[NA], rb[NBL];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int q = tid + u * kThreads;
            const int col = q / (MB / 2), r2 = (q % (MB / 2)) * 2;
            ra[u] = *reinterpret_cast<const double2 *>(Ag + (size_t)(k0 + col) * lda + r2);
        }
#pragma unroll
        for (int u = 0; u < NBL; ++u) {
            const int q = tid + u * kThreads;
            const int col = q / (NB / 2), r2 = (q % (NB / 2)) * 2;
            rb[u] = *reinterpret_cast<const double2 *>(Bg + (size_t)(k0 + col) * ldb + r2);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int q = tid + u * kThreads;
            const int col = q / (MB / 2), r2 = (q % (MB / 2)) * 2;
            *reinterpret_cast<double2 *>(As[buf] + col * T::LDA + r2) = ra[u];
        }
#pragma unroll
        for (int u = 0; u < NBL; ++u) {
            const int q = tid + u * kThreads;
            const int col = q / (NB / 2), r2 = (q % (NB / 2)) * 2;
            *reinterpret_cast<double2 *>(Bs[buf] + col * T::LDB + r2) = rb[u];
        }
    };

    const int nchunks = depth / kKC;
    gload(0);
    sstore(0);
    __syncthreads();
    const int kq = lane >> 4, lr = lane & 15;
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        if (c + 1 < nchunks) gload((c + 1) * kKC);
        const double *as = As[buf];
        const double *bs = Bs[buf];
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
        if (c + 1 < nchunks) sstore(buf ^ 1);
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

template <int MB, int NB>
__global__ void __launch_bounds__(kThreads) gemm_step_kernel(double *__restrict__ K, int np, int step,
                                                             const int *__restrict__ info)
{
    using T = TileShape<MB, NB>;
    __shared__ __attribute__((aligned(16))) double smem[T::SMEM];
    const int rt_per = (np - step * NB) / MB;
    const int s = blockIdx.x / rt_per;
    const int rt = blockIdx.x % rt_per;
    if (info[s] != 0) return;
    const int row0 = step * NB + rt * MB;
    const int col0 = step * NB;
    double *A = K + (size_t)s * np * np;
    d4 acc[T::BM][T::BN];
    tile_gemm<MB, NB>(A + row0, np, A + col0, np, step * NB, smem, acc);
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

template <int MB, int NB>
__global__ void __launch_bounds__(kThreads) trsm_step_kernel(double *__restrict__ K, int np, int step,
                                                             const int *__restrict__ info,
                                                             const double *__restrict__ inv_diag)
{
    using T = TileShape<MB, NB>;
    __shared__ __attribute__((aligned(16))) double smem[T::SMEM];
    const int rt_per = (np - (step + 1) * NB) / MB;
    const int s = blockIdx.x / rt_per;
    const int rt = blockIdx.x % rt_per;
    if (info[s] != 0) return;
    const int row0 = (step + 1) * NB + rt * MB;
    const int col0 = step * NB;
    double *A = K + (size_t)s * np * np;
    double *Ct = A + (size_t)col0 * np + row0;
    // X = C inv^T:  X[r][c] = sum_k C[r][k] inv[c][k], inv[c][k] at inv[k*NB + c]
    d4 acc[T::BM][T::BN];
    tile_gemm<MB, NB>(Ct, np, inv_diag + (size_t)s * NB * NB, NB, NB, smem, acc);
    acc_to_lds<MB, NB>(smem, acc);
    __syncthreads();
    for (int q = threadIdx.x; q < MB * NB / 2; q += kThreads) {
        const int col = q / (MB / 2), r2 = (q % (MB / 2)) * 2;
        *reinterpret_cast<double2 *>(Ct + (size_t)col * np + r2) =
            *reinterpret_cast<const double2 *>(smem + col * T::LDC + r2);
    }
}

// Factor the diagonal block of step `step` of every live matrix in LDS.
// Columns at or beyond n (bordered right-hand-side rows and padding) are not
// factored: they are set to the identity, which leaves the right-hand-side
// rows holding L^{-1} rhs for the real columns.
__global__ void __launch_bounds__(kThreads) potrf_diag_kernel(double *__restrict__ K, int np, int n,
                                                              int step, int nsteps,
                                                              int *__restrict__ info,
                                                              double *__restrict__ logdet_part,
                                                              double *__restrict__ inv_diag)
{
    constexpr int NB = kNB;
    constexpr int LDP = NB + 1;
    __shared__ double L[NB * LDP];
    __shared__ double X[NB * LDP];
    __shared__ double red[kThreads / 64];
    const int s = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    if (info[s] != 0) return;
    const int c0 = step * NB;
    double *Ag = K + (size_t)s * np * np + (size_t)c0 * np + c0;
    for (int e = tid; e < NB * NB; e += kThreads) {
        const int c = e / NB, i = e % NB;
        L[c * LDP + i] = i >= c ? Ag[(size_t)c * np + i] : 0.0;
    }
    __syncthreads();
    const int ncol = min(NB, n - c0);
    int failed = 0;
    for (int c = 0; c < ncol; ++c) {
        const double piv = L[c * LDP + c];
        // dpotf2: stop at the first pivot that is not positive or is NaN
        if (!(piv > 0.0)) {
            if (tid == 0) info[s] = c0 + c + 1;
            failed = 1;
            break;
        }
        const double d = sqrt(piv);
        const double r = 1.0 / d;
        for (int i = c + 1 + tid; i < NB; i += kThreads) L[c * LDP + i] *= r;
        __syncthreads();
        if (tid == 0) L[c * LDP + c] = d;
        for (int k = c + 1 + wave; k < NB; k += kThreads / 64) {
            const double lk = L[c * LDP + k];
            for (int i = k + lane; i < NB; i += 64) L[k * LDP + i] -= L[c * LDP + i] * lk;
        }
        __syncthreads();
    }
    if (failed) return;
    // identity on the unfactored (right-hand-side / padding) columns
    for (int e = tid; e < NB * NB; e += kThreads) {
        const int c = e / NB, i = e % NB;
        if (c >= ncol && i >= c) L[c * LDP + i] = (i == c) ? 1.0 : 0.0;
    }
    // partial log-determinant of the real columns of this block
    double lg = 0.0;
    for (int c = tid; c < ncol; c += kThreads) lg += log(L[c * LDP + c]);
    for (int off = 32; off > 0; off >>= 1) lg += __shfl_xor(lg, off, 64);
    if (lane == 0) red[wave] = lg;
    __syncthreads();
    if (tid == 0) logdet_part[(size_t)s * nsteps + step] = (red[0] + red[1]) + (red[2] + red[3]);
    for (int e = tid; e < NB * NB; e += kThreads) {
        const int c = e / NB, i = e % NB;
        if (i >= c) Ag[(size_t)c * np + i] = L[c * LDP + i];
    }
    // inverse of the lower-triangular block, one column per thread:
    // X(i, t) = (delta_it - sum_{t<=k<i} L(i, k) X(k, t)) / L(i, i)
    // The k loop is uniform across the wave so L(i, k) is a broadcast read;
    // X(k, t) = 0 for k < t makes the extra terms vanish.
    if (tid < NB) {
        const int t = tid;
        for (int i = 0; i < NB; ++i) {
            double acc = (i == t) ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) acc -= L[k * LDP + i] * X[t * LDP + k];
            X[t * LDP + i] = (i >= t) ? acc / L[i * LDP + i] : 0.0;
        }
    }
    __syncthreads();
    double *inv = inv_diag + (size_t)s * NB * NB;
    for (int e = tid; e < NB * NB; e += kThreads) {
        const int c = e / NB, i = e % NB;
        inv[e] = X[c * LDP + i];   // inv(i, c) at inv[c*NB + i]; zero above the diagonal
    }
}

void launch_gemm_step(const SwarmBuffers &b, int step, hipStream_t st)
{
    if (step == 0) return;
    const int rt_per = (b.np - step * kNB) / kMB;
    hipLaunchKernelGGL((gemm_step_kernel<kMB, kNB>), dim3(b.S * rt_per), dim3(kThreads), 0, st,
                       b.K, b.np, step, b.info);
}

void launch_potrf_diag(const SwarmBuffers &b, int step, hipStream_t st)
{
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(b.S), dim3(kThreads), 0, st, b.K, b.np, b.n, step,
                       b.np / kNB, b.info, b.logdet_part, b.inv_diag);
}

void launch_trsm_step(const SwarmBuffers &b, int step, hipStream_t st)
{
    const int rt_per = (b.np - (step + 1) * kNB) / kMB;
    if (rt_per <= 0) return;
    hipLaunchKernelGGL((trsm_step_kernel<kMB, kNB>), dim3(b.S * rt_per), dim3(kThreads), 0, st,
                       b.K, b.np, step, b.info, b.inv_diag);
}

}  // namespace fb
