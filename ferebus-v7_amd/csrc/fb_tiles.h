// fb_tiles.h -- tile-level device routines of the batched FP64 Cholesky,
// shared by the lock-step kernels (fb_potrf.hip) and the persistent task
// kernel (fb_dag.hip), so both paths perform identical arithmetic.
//
// Matrices are column major with leading dimension np; a "tile" is a 64-row
// slice of a 64- or 128-column block.
#pragma once

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

// Work-group tile of MB rows x NB columns computed by 4 waves, K streamed in
// chunks of KC through two LDS stages.
template <int MB, int NB, int KC = kKC>
struct TileShape {
    static constexpr int WAVES_N = (NB / 32 < 4) ? NB / 32 : 4;
    static constexpr int WAVES_M = 4 / WAVES_N;
    static constexpr int WM = MB / WAVES_M;
    static constexpr int WN = NB / WAVES_N;
    static constexpr int BM = WM / 16;
    static constexpr int BN = WN / 16;
    static constexpr int LDA = MB + 16;   // LDS strides in doubles; = 16 mod 32
    static constexpr int LDB = NB + 16;   // keeps the two k rows of a fragment read
    static constexpr int ASZ = KC * LDA;  // on different LDS banks
    static constexpr int BSZ = KC * LDB;
    static constexpr int SMEM = 2 * (ASZ + BSZ);
    static_assert(WAVES_M * WAVES_N == 4, "4 waves per tile");
    static_assert(BM >= 1 && BN >= 1, "wave tile at least 16x16");
    static_assert(KC % 4 == 0 && (MB * KC / 2) % kThreads == 0 && (NB * KC / 2) % kThreads == 0, "loads");
};

// C(row0:row0+MB, col0:col0+NB) -= A(row0.., 0:depth) B(col0.., 0:depth)^T
// with A[r][k] = Ag[k*lda + r], B[c][k] = Bg[k*ldb + c] and C column major
// (leading dimension ldc); depth is a multiple of KC (may be 0).
//
// The products are issued with the B panel as the first MFMA operand, so
// the accumulator of a 16x16 block holds C^T: lane l, component v is
// C[r0 + l%16][c0 + 4v + l/16] -- sixteen consecutive rows of four columns
// per load, which coalesces.  C is therefore read into registers before the
// K loop (its latency hidden behind the loop) and written back straight from
// the accumulators: no LDS epilogue, no barrier after the loop.  Each chunk
// of KC is fetched into registers while the previous one feeds the matrix
// cores from LDS.
template <int MB, int NB, int KC>
__device__ inline void tile_gemm_sub(double *__restrict__ Cg, int ldc,
                                     const double *__restrict__ Ag, int lda,
                                     const double *__restrict__ Bg, int ldb, int depth,
                                     double *__restrict__ smem)
{
    using T = TileShape<MB, NB, KC>;
    constexpr int NA = MB * KC / 2 / kThreads;   // double2 loads per thread (A)
    constexpr int NBL = NB * KC / 2 / kThreads;  // double2 loads per thread (B)
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wrow = (wave / T::WAVES_N) * T::WM;
    const int wcol = (wave % T::WAVES_N) * T::WN;
    const int lr = lane & 15, lq = lane >> 4;

    // C^T blocks: acc[i][j][v] = C[wrow + 16 i + lr][wcol + 16 j + 4 v + lq]
    d4 acc[T::BM][T::BN];
#pragma unroll
    for (int i = 0; i < T::BM; ++i)
#pragma unroll
        for (int j = 0; j < T::BN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                acc[i][j][v] = Cg[(size_t)(wcol + 16 * j + 4 * v + lq) * ldc + wrow + 16 * i + lr];

    if (depth > 0) {
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
        // staging registers as scalar pairs: arrays of double2 are not
        // promoted to registers by the compiler and would live in scratch
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
        for (int u = 0; u < NBL; ++u)
            *reinterpret_cast<double2 *>(smem + 2 * T::ASZ + b_lds[u]) = double2{rbx[u], rby[u]};
        __syncthreads();

        const int nchunks = depth / KC;
        for (int c = 0; c < nchunks; ++c) {
            const int cur = c & 1;
            const bool more = c + 1 < nchunks;
            if (more) {
                const size_t ka = (size_t)(c + 1) * KC * lda, kb = (size_t)(c + 1) * KC * ldb;
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
            const double *as = smem + cur * T::ASZ;
            const double *bs = smem + 2 * T::ASZ + cur * T::BSZ;
#pragma unroll
            for (int kk = 0; kk < KC / 4; ++kk) {
                double af[T::BM], bf[T::BN];
#pragma unroll
                for (int i = 0; i < T::BM; ++i) af[i] = -as[(kk * 4 + lq) * T::LDA + wrow + i * 16 + lr];
#pragma unroll
                for (int j = 0; j < T::BN; ++j) bf[j] = bs[(kk * 4 + lq) * T::LDB + wcol + j * 16 + lr];
#pragma unroll
                for (int i = 0; i < T::BM; ++i)
#pragma unroll
                    for (int j = 0; j < T::BN; ++j) acc[i][j] = mfma_f64(bf[j], af[i], acc[i][j]);
            }
            if (more) {
                double *an = smem + (cur ^ 1) * T::ASZ;
                double *bn = smem + 2 * T::ASZ + (cur ^ 1) * T::BSZ;
#pragma unroll
                for (int u = 0; u < NA; ++u) *reinterpret_cast<double2 *>(an + a_lds[u]) = double2{rax[u], ray[u]};
#pragma unroll
                for (int u = 0; u < NBL; ++u) *reinterpret_cast<double2 *>(bn + b_lds[u]) = double2{rbx[u], rby[u]};
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < T::BM; ++i)
#pragma unroll
        for (int j = 0; j < T::BN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                Cg[(size_t)(wcol + 16 * j + 4 * v + lq) * ldc + wrow + 16 * i + lr] = acc[i][j][v];
}

// C(row0:row0+MB, col0:col0+NB) -= L(row0.., k0:k1) L(col0.., k0:k1)^T
// for the matrix at A (trailing update of one tile).
template <int MB, int NB, int KC = kKC>
__device__ inline void gemm_tile_update(double *__restrict__ A, int np, int row0, int col0, int k0, int k1,
                                        double *__restrict__ smem)
{
    tile_gemm_sub<MB, NB, KC>(A + (size_t)col0 * np + row0, np, A + (size_t)k0 * np + row0, np,
                              A + (size_t)k0 * np + col0, np, k1 - k0, smem);
}

constexpr int kDiagLD = 17;                                         // LDS stride of a 16x16 block
constexpr int kPotrfSmem = 2 * kNBI + 8 + 4 * 16 * kDiagLD;       // doubles

// Factors the 64x64 diagonal block at (c0, c0) of the matrix at A.  Thread
// (lane i, wave w) keeps row i of columns w, w+4, ..., w+60 in registers; the
// wave owning column c takes the pivot (LAPACK dpotf2 test: not positive or
// NaN fails), scales the column and publishes it through LDS, then every
// thread applies the rank-1 update to its registers: one barrier per column.
// Columns at or beyond n (bordered right-hand-side rows and padding) are not
// factored: they become the identity, leaving the right-hand-side rows
// holding L^{-1} rhs.  Returns the 1-based global failing column (0 = success)
// and, on success, writes sum(log L_ii) of the block's real columns to
// *logdet (thread 0) and the inverses of the four 16x16 diagonal blocks of
// L to dinv (block q column major at dinv[256 q]), for trsm64_mfma.
__device__ inline int potrf64_block(double *__restrict__ A, int np, int n, int c0, double *logdet,
                                    double *__restrict__ smem, double *__restrict__ dinv)
{
    constexpr int B = kNBI;
    constexpr int M = B / 4;
    double *col0 = smem, *col1 = smem + B;
    double *red = smem + 2 * B;
    int *fail_col = reinterpret_cast<int *>(smem + 2 * B + kThreads / 64);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    double *Ag = A + (size_t)c0 * np + c0 + lane;
    double a[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int k = wave + 4 * m;
        a[m] = lane >= k ? Ag[(size_t)k * np] : 0.0;
    }
    if (tid == 0) *fail_col = 0;
    const int ncol = min(B, n - c0);
    __syncthreads();
    // every thread runs all B steps (the barrier is unconditional); steps at
    // or beyond ncol, or after a failed pivot, do no work
#pragma unroll
    for (int c = 0; c < B; ++c) {
        double *col = (c & 1) ? col1 : col0;
        const bool live = c < ncol && *fail_col == 0;
        if (live && wave == (c & 3)) {
            const double piv = __shfl(a[c >> 2], c, 64);
            if (!(piv > 0.0)) {
                if (lane == 0) *fail_col = c0 + c + 1;
            } else {
                const double d = sqrt(piv);
                const double l = lane > c ? a[c >> 2] * (1.0 / d) : (lane == c ? d : 0.0);
                if (lane >= c) a[c >> 2] = l;
                col[lane] = l;
            }
        }
        __syncthreads();
        if (c < ncol && *fail_col == 0) {
            // all reads first (they are independent), then the predicated update
            double lk[M];
#pragma unroll
            for (int m = 0; m < M; ++m) lk[m] = col[wave + 4 * m];
            const double li = col[lane];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const double t = a[m] - li * lk[m];
                a[m] = (wave + 4 * m > c) ? t : a[m];
            }
        }
    }
    const int failed = *fail_col;
    if (failed != 0) return failed;
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
    // diagonal 16x16 blocks to LDS: row lane of block q = lane / 16 holds
    // columns k = wave + 4 m with m in [4q, 4q + 4)
    double *Ds = smem + 2 * B + 8;
    const int q = lane >> 4, il = lane & 15;
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) {
        double v = 0.0;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) v = (q == qq) ? a[4 * qq + mm] : v;
        Ds[q * 16 * kDiagLD + (wave + 4 * mm) * kDiagLD + il] = v;   // column wave + 4 mm of block q
    }
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int k = wave + 4 * m;
        if (lane >= k) Ag[(size_t)k * np] = a[m];
    }
    __syncthreads();
    if (tid == 0) *logdet = (red[0] + red[1]) + (red[2] + red[3]);
    // wave w inverts diagonal block w, lane c < 16 one column (column-oriented
    // forward substitution of L_ww x = e_c)
    if (lane < 16) {
        const double *D = Ds + wave * 16 * kDiagLD;   // D(i, k) at D[k * kDiagLD + i]
        double x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = i == lane ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            x[k] = x[k] / D[k * kDiagLD + k];
#pragma unroll
            for (int i = k + 1; i < 16; ++i) x[i] -= D[k * kDiagLD + i] * x[k];
        }
        double *W = dinv + wave * 256 + lane * 16;
#pragma unroll
        for (int i = 0; i < 16; ++i) W[i] = x[i];
    }
    return 0;
}

// L(row0:row0+64, c0:c0+64) = C L_bb^{-T} in place, C the tile and L_bb the
// factored diagonal block at (c0, c0).  Each wave solves its 16 rows alone:
// with Y = X^T (X the wave's 16 x 64 slice of the result), L_bb Y = C^T is
// block forward substitution over four 16-row groups,
//     Y_j = W_j (C_j^T - sum_{k<j} L_jk Y_k),    W_j = L_jj^{-1} (dinv),
// every product on the matrix cores.  An MFMA result holds D[4v + l/16][l%16]
// in lane l, component v: the B-operand layout of four consecutive k-steps,
// so each Y_k feeds the later products straight from registers.  No LDS, no
// barrier; the tile is read completely before it is overwritten.
__device__ inline void trsm64_mfma(double *__restrict__ A, int np, int c0, int row0,
                                   const double *__restrict__ dinv)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    double *Ct = A + row0 + 16 * wave + lr;
    const double *Lb = A + (size_t)c0 * np + c0;
    d4 t[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) t[j][v] = Ct[(size_t)(c0 + 16 * j + 4 * v + lq) * np];
    // L_jk (j > k) in A layout: lane l holds L_jk[l%16][4kk + l/16]
    double lf[6][4];
#pragma unroll
    for (int j = 1, p = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < j; ++k, ++p)
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) lf[p][kk] = Lb[(size_t)(16 * k + 4 * kk + lq) * np + 16 * j + lr];
    double wf[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) wf[j][kk] = dinv[j * 256 + (4 * kk + lq) * 16 + lr];
    d4 y[4];
#pragma unroll
    for (int j = 0, p = 0; j < 4; ++j) {
        d4 r = t[j];
#pragma unroll
        for (int k = 0; k < j; ++k, ++p)
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) r = mfma_f64(-lf[p][kk], y[k][kk], r);
        d4 o = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) o = mfma_f64(wf[j][kk], r[kk], o);
        y[j] = o;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) Ct[(size_t)(c0 + 16 * j + 4 * v + lq) * np] = y[j][v];
}

}  // namespace fb
