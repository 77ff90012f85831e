// fb_tiles.h -- tile-level device routines of the batched FP64 Cholesky,
// shared by the lock-step kernels (fb_potrf.hip) and the persistent task
// kernel (fb_dag.hip), so both paths perform identical arithmetic.
//
// Matrices are column major with leading dimension np; a "tile" is a 64-row
// slice of a 64- or 128-column block.
#pragma once

#include <type_traits>

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

// D = C - A B: the same product with the first operand negated by the
// instruction's neg modifier (on the f64 products the last argument is
// neg[a, b, c]; bit for bit the product with -A, checked on gfx950 with a
// one-wave probe), so no operand is negated in registers or LDS.
__device__ inline d4 mfma_f64_sub(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 1);
}

constexpr int kPLD = 68;                        // LDS stride (doubles) of a 64x64 block kept on chip

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
// the accumulators: no LDS epilogue, no barrier after the loop.  K is
// streamed in chunks of KC through two LDS buffers; the global loads of a
// chunk are issued two chunks ahead (two register sets), so each has two
// chunks of matrix-core work to land.  For -= the products are subtracted by
// the matrix cores' neg modifier (mfma_f64_sub): C is loaded straight into
// the accumulators and stored straight back, the chunks are staged as
// loaded and the inner loop passes LDS reads straight to the matrix cores
// (no negation anywhere, so the loads of the first chunks are not held up
// behind the loads of C either).
// ZERO_C: C starts at zero (not loaded); PLUS: C += A B^T instead of -=.
// hook (Hook::kActive): called by every thread once, right before the
// global loads of chunk hook_chunk (0 <= hook_chunk < depth / KC) are
// issued -- a wait for operands that another work-group produces while this
// one accumulates the earlier chunks; the accumulators stay in registers.
struct NoHook {
    static constexpr bool kActive = false;
    __device__ void operator()() const {}
};

template <int MB, int NB, int KC, int DIST = 2, bool TO_LDS = false, bool ZERO_C = false, bool PLUS = false,
          class Hook = NoHook>
__device__ inline void tile_gemm_sub(double *__restrict__ Cg, int ldc,
                                     const double *__restrict__ Ag, int lda,
                                     const double *__restrict__ Bg, int ldb, int depth,
                                     double *__restrict__ smem, int hook_chunk = -1, const Hook &hook = Hook())
{
    using T = TileShape<MB, NB, KC>;
    // the software-pipelined K loop: two fragment sets cost 12-24 VGPRs, so
    // only the 16- and 32-deep tiles (two work-groups per CU) take it
    constexpr bool PIPE = DIST > 1 && KC >= 16;
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
                acc[i][j][v] = ZERO_C ? 0.0 : Cg[(size_t)(wcol + 16 * j + 4 * v + lq) * ldc + wrow + 16 * i + lr];

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
        // two register staging sets (scalar pairs: arrays of double2 would
        // live in scratch): the loads of chunk c + 2 are issued before chunk
        // c is computed, so every load has two chunks of MFMA work to land
        double r0ax[NA], r0ay[NA], r0bx[NBL], r0by[NBL];
        double r1ax[DIST > 1 ? NA : 1], r1ay[DIST > 1 ? NA : 1], r1bx[DIST > 1 ? NBL : 1], r1by[DIST > 1 ? NBL : 1];
#define FB_LOAD(SET, CH)                                                                        \
        do {                                                                                    \
            if constexpr (Hook::kActive) {                                                      \
                if ((CH) == hook_chunk) hook();                                                 \
            }                                                                                   \
            const size_t ka = (size_t)(CH) * KC * lda, kb = (size_t)(CH) * KC * ldb;            \
            _Pragma("unroll") for (int u = 0; u < NA; ++u) {                                    \
                const double2 t = *reinterpret_cast<const double2 *>(Ag + ka + a_off[u]);      \
                r##SET##ax[u] = t.x; r##SET##ay[u] = t.y;                                       \
            }                                                                                   \
            _Pragma("unroll") for (int u = 0; u < NBL; ++u) {                                   \
                const double2 t = *reinterpret_cast<const double2 *>(Bg + kb + b_off[u]);      \
                r##SET##bx[u] = t.x; r##SET##by[u] = t.y;                                       \
            }                                                                                   \
        } while (0)
#define FB_STORE(SET, BUF)                                                                      \
        do {                                                                                    \
            double *an = smem + (BUF) * T::ASZ;                                                 \
            double *bn = smem + 2 * T::ASZ + (BUF) * T::BSZ;                                    \
            _Pragma("unroll") for (int u = 0; u < NA; ++u)                                      \
                *reinterpret_cast<double2 *>(an + a_lds[u]) = double2{r##SET##ax[u], r##SET##ay[u]};  \
            _Pragma("unroll") for (int u = 0; u < NBL; ++u)                                     \
                *reinterpret_cast<double2 *>(bn + b_lds[u]) = double2{r##SET##bx[u], r##SET##by[u]}; \
        } while (0)
        auto compute = [&](int buf) {
            const double *as = smem + buf * T::ASZ;
            const double *bs = smem + 2 * T::ASZ + buf * T::BSZ;
#pragma unroll DIST > 1 ? KC / 4 : 2
            for (int kk = 0; kk < KC / 4; ++kk) {
                double af[T::BM], bf[T::BN];
#pragma unroll
                for (int i = 0; i < T::BM; ++i) af[i] = as[(kk * 4 + lq) * T::LDA + wrow + i * 16 + lr];
#pragma unroll
                for (int j = 0; j < T::BN; ++j) bf[j] = bs[(kk * 4 + lq) * T::LDB + wcol + j * 16 + lr];
#pragma unroll
                for (int i = 0; i < T::BM; ++i)
#pragma unroll
                    for (int j = 0; j < T::BN; ++j)
                        acc[i][j] = PLUS ? mfma_f64(bf[j], af[i], acc[i][j]) : mfma_f64_sub(bf[j], af[i], acc[i][j]);
            }
        };
        const int nchunks = depth / KC;
        FB_LOAD(0, 0);
        FB_STORE(0, 0);
        if constexpr (PIPE) {
            // fragments of k-step kk + 1 read from LDS before the products of
            // k-step kk are issued (two fragment register sets; sched_barrier
            // keeps the order, which the scheduler otherwise undoes); the
            // first step of the next chunk is read right after the barrier,
            // while the matrix cores still work through the last step
            double af[2][T::BM], bf[2][T::BN];
            auto rd = [&](int set, int buf, int kk) {
                const double *as = smem + buf * T::ASZ;
                const double *bs = smem + 2 * T::ASZ + buf * T::BSZ;
#pragma unroll
                for (int i = 0; i < T::BM; ++i) af[set][i] = as[(kk * 4 + lq) * T::LDA + wrow + i * 16 + lr];
#pragma unroll
                for (int j = 0; j < T::BN; ++j) bf[set][j] = bs[(kk * 4 + lq) * T::LDB + wcol + j * 16 + lr];
            };
            auto mm = [&](int set) {
#pragma unroll
                for (int i = 0; i < T::BM; ++i)
#pragma unroll
                    for (int j = 0; j < T::BN; ++j)
                        acc[i][j] = PLUS ? mfma_f64(bf[set][j], af[set][i], acc[i][j])
                                         : mfma_f64_sub(bf[set][j], af[set][i], acc[i][j]);
            };
            // k-steps 1 .. KC/4 - 1 of the chunk in buf (step 0's fragments
            // in set 0); the last step's fragments are left in set 1
            auto steps = [&](int buf) {
#pragma unroll
                for (int kk = 1; kk < KC / 4; ++kk) {
                    rd(kk & 1, buf, kk);
                    __builtin_amdgcn_sched_barrier(0);
                    mm((kk - 1) & 1);
                    __builtin_amdgcn_sched_barrier(0);
                }
            };
            if (nchunks > 1) FB_LOAD(1, 1);
            __syncthreads();
            rd(0, 0, 0);
            int c = 0;
            for (; c + 1 < nchunks; c += 2) {
                // chunk c in LDS buffer 0, chunk c + 1 in register set 1
                if (c + 2 < nchunks) FB_LOAD(0, c + 2);
                steps(0);
                FB_STORE(1, 1);
                __syncthreads();
                // chunk c + 1 in LDS buffer 1, chunk c + 2 in register set 0
                // (the loads of c + 3 go first: a branch after the fragment
                // reads would let the compiler sink them past the products)
                if (c + 3 < nchunks) FB_LOAD(1, c + 3);
                rd(0, 1, 0);
                __builtin_amdgcn_sched_barrier(0);
                mm(1);
                __builtin_amdgcn_sched_barrier(0);
                steps(1);
                if (c + 2 < nchunks) FB_STORE(0, 0);
                __syncthreads();
                if (c + 2 < nchunks) rd(0, 0, 0);   // no read past the last chunk
                __builtin_amdgcn_sched_barrier(0);
                mm(1);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (c < nchunks) {
                // an odd count: the last chunk is in LDS buffer 0
                steps(0);
                mm(1);
                __syncthreads();
            }
        } else if constexpr (DIST == 1) {
            // one register set: the loads of chunk c + 1 land during chunk c
            __syncthreads();
            for (int c = 0; c < nchunks; ++c) {
                if (c + 1 < nchunks) FB_LOAD(0, c + 1);
                compute(c & 1);
                if (c + 1 < nchunks) FB_STORE(0, (c + 1) & 1);
                __syncthreads();
            }
        } else {
        if (nchunks > 1) FB_LOAD(1, 1);
        __syncthreads();
        // pairs of chunks, one loop exit: an exit between the two halves
        // makes the compiler keep a second accumulator set and copy it
        int c = 0;
        for (; c + 1 < nchunks; c += 2) {
            // chunk c in LDS buffer 0, chunk c + 1 in register set 1
            if (c + 2 < nchunks) FB_LOAD(0, c + 2);
            compute(0);
            FB_STORE(1, 1);
            __syncthreads();
            // chunk c + 1 in LDS buffer 1, chunk c + 2 in register set 0
            if (c + 3 < nchunks) FB_LOAD(1, c + 3);
            compute(1);
            if (c + 2 < nchunks) FB_STORE(0, 0);
            __syncthreads();
        }
        if (c < nchunks) {
            // an odd count: the last chunk is in LDS buffer 0
            compute(0);
            __syncthreads();
        }
        }
#undef FB_LOAD
#undef FB_STORE
    }
    if constexpr (TO_LDS) {
        // the tile stays on chip for the caller: smem[c * kPLD + r], the
        // layout of the diagonal-block routines (the K loop ended with a
        // barrier, so the staging buffers are free)
        static_assert(MB <= 64 && NB <= 64, "LDS tile is at most 64 x 64");
#pragma unroll
        for (int i = 0; i < T::BM; ++i)
#pragma unroll
            for (int j = 0; j < T::BN; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    smem[(wcol + 16 * j + 4 * v + lq) * kPLD + wrow + 16 * i + lr] = acc[i][j][v];
    } else {
#pragma unroll
        for (int i = 0; i < T::BM; ++i)
#pragma unroll
            for (int j = 0; j < T::BN; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    Cg[(size_t)(wcol + 16 * j + 4 * v + lq) * ldc + wrow + 16 * i + lr] = acc[i][j][v];
    }
}

// C(row0:row0+MB, col0:col0+NB) -= L(row0.., k0:k1) L(col0.., k0:k1)^T
// for the matrix at A (trailing update of one tile); hook: called before
// the loads of columns kh .. (k0 <= kh < k1, kh - k0 a multiple of KC).
template <int MB, int NB, int KC = kKC, int DIST = 2, bool TO_LDS = false, class Hook = NoHook>
__device__ inline void gemm_tile_update(double *__restrict__ A, int np, int row0, int col0, int k0, int k1,
                                        double *__restrict__ smem, int kh = -1, const Hook &hook = Hook())
{
    tile_gemm_sub<MB, NB, KC, DIST, TO_LDS, false, false, Hook>(
        A + (size_t)col0 * np + row0, np, A + (size_t)k0 * np + row0, np, A + (size_t)k0 * np + col0, np, k1 - k0,
        smem, Hook::kActive ? (kh - k0) / KC : -1, hook);
}

constexpr int kPotrfSmem = kNBI * kPLD + kNBI + 8;   // doubles: block, 1 / L_ii, flag, log sum

// A store of v at p; WT: one that reaches memory (a device-scope atomic store
// goes past the XCD's L2), for results read by work-groups of other XCDs in
// the same launch.
template <bool WT>
__device__ inline void st_f64(double *p, double v)
{
    if constexpr (WT)
        __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), (unsigned long long)__double_as_longlong(v),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        *p = v;
}

// sqrt(x) and 1 / sqrt(x) for x > 0 from an estimate y of 1 / sqrt(x),
// refined by two steps of Goldschmidt's coupled iteration
// (g -> sqrt x, h -> 1 / (2 sqrt x); r = 1/2 - g h, g += g r, h += h r),
// each step roughly doubling the correct bits; a few dependent FMAs instead
// of the long sqrt and division sequences.  The square root comes out
// correctly rounded, its reciprocal within 2 ulp, both unbiased (a host
// build of this routine is checked against long double on 400,000 points).
__host__ __device__ inline void goldschmidt_sqrt_rsqrt(double x, double y, double &sq, double &rsq)
{
    double g = x * y, h = 0.5 * y;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double r = fma(-g, h, 0.5);
        g = fma(g, r, g);
        h = fma(h, r, h);
    }
    // the square root corrected by its exact residual x - g^2 (Markstein),
    // then h by one more step against the corrected root: the coupled steps
    // alone leave up to ~1.4 ulp in g and 2 ulp in h from their roundings
    const double e = fma(-g, g, x);
    g = fma(e, h, g);
    const double r = fma(-g, h, 0.5);
    h = fma(h, r, h);
    sq = g;
    rsq = 2.0 * h;
}

// The estimate on the device is the hardware reciprocal square root.
__device__ inline void sqrt_rsqrt(double x, double &sq, double &rsq)
{
    goldschmidt_sqrt_rsqrt(x, __builtin_amdgcn_rsq(x), sq, rsq);
}

// v, made where it is used: the compiler cannot hoist the value out of a
// loop (the identity column of a 16x16 inverse, hoisted out of the panel
// loop, stays live across the whole block factorisation and spills when the
// register budget is that of three waves per SIMD).
__device__ inline double opaque_f64(double v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// x from lane l (uniform l) of the wave.
__device__ inline double readlane_f64(double x, int l)
{
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)v, l);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Ratio of the largest to the smallest pivot of a 16x16 diagonal block above
// which the panel solves substitute inside the block instead of multiplying
// by its inverse: a product with the inverse is not backward stable for an
// ill-conditioned block, and at a 1e-10 nugget the first block of columns
// (pivots falling from ~1 to ~1e-5 as the smooth part of the kernel is
// exhausted) then perturbs every later pivot by ~1e-5 relative, log det K by
// ~1e-4 -- ten times what the substitution order of dpotf2 gives.
constexpr double kSubstRatio = 32.0;

// Wave-wide: lanes 0..15 hold the pivots L(c, c) of a 16x16 diagonal block
// (real = false for identity columns beyond n).  Lane 0 stores the block's
// flag at *F: 1 when its largest pivot exceeds kSubstRatio times its
// smallest, else 0; trsm16_rows reads it to choose substitution.
template <bool WT>
__device__ __forceinline__ void store_subst_flag(double piv, bool real, double *F, int lane)
{
    double lo = real ? piv : 1e300, hi = real ? piv : 0.0;
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 16));
        hi = fmax(hi, __shfl_xor(hi, off, 16));
    }
    if (lane == 0) st_f64<WT>(F, hi > kSubstRatio * lo ? 1.0 : 0.0);
}

// Factors the 64x64 diagonal block at (c0, c0) of the matrix at A,
// right-looking in four 16-column panels, the block staged in LDS:
//  * wave 0 factors panel j (lane = row) in groups of four columns: inside
//    a group the pivots and multipliers come from other lanes by
//    v_readlane, and a finished group's multipliers reach the panel's later
//    columns as LDS broadcast reads (one wave: its LDS operations complete
//    in order), so the 16 dependent steps need no barrier (LAPACK dpotf2
//    test: a pivot that is not positive, or NaN, fails; L = a * (1 /
//    sqrt(pivot)), the square root and its reciprocal from sqrt_rsqrt).  Every lane updates every column
//    of the panel: lanes above the diagonal hold upper-triangle entries,
//    which no lower entry ever reads, so no per-update lane mask is needed;
//  * the four waves then apply the panel to the trailing 16x16 blocks on
//    the matrix cores (accumulators hold C^T, as in tile_gemm_sub);
//  * while wave 0 factors panel j, the finished panel j - 1 is completed by
//    the other waves: logarithms of its diagonal (wave 1), the inverse of
//    its 16x16 diagonal block (wave 2), its store to global memory (wave 3).
// Columns at or beyond n (bordered right-hand-side rows and padding) are not
// factored: they become the identity, leaving the right-hand-side rows
// holding L^{-1} rhs (such a block is completed after the loop).  Returns
// the 1-based global failing column (0 = success); on success writes the
// block's lower triangle back, sum(log L_ii) of its real columns to *logdet
// (thread 0) and the inverses of its four 16x16 diagonal blocks to dinv
// (block q column major at dinv[256 q]), for trsm64_mfma.  STAGED: the
// block is already in smem (Ls layout), left there by the caller's update.
template <bool STAGED = false, bool WT = false>
__device__ inline int potrf64_block(double *__restrict__ A, int np, int n, int c0, double *logdet,
                                    double *__restrict__ smem, double *__restrict__ dinv)
{
    constexpr int B = kNBI;
    double *Ls = smem;                               // A(r, c) at Ls[c * kPLD + r]
    double *rdiag = smem + B * kPLD;                 // 1 / L_ii
    int *flag = reinterpret_cast<int *>(smem + B * kPLD + B);
    double *lgsum = smem + B * kPLD + B + 2;         // reduced log-determinant
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    double *Ag = A + (size_t)c0 * np + c0;
    if constexpr (!STAGED) {
#pragma unroll
        for (int u = 0; u < B * B / 2 / kThreads; ++u) {
            const int e = tid + u * kThreads;
            const int col = e / (B / 2), r2 = (e % (B / 2)) * 2;
            *reinterpret_cast<double2 *>(Ls + col * kPLD + r2) = *reinterpret_cast<const double2 *>(Ag + (size_t)col * np + r2);
        }
    }
    if (tid == 0) *flag = 0;
    const int ncol = min(B, n - c0);
    const bool full = ncol == B;                     // no identity columns in this block
    double lg = 0.0;                                 // wave 1, lanes < 16: logs of its diagonal entries
    // completion of panel q (columns 16q .. 16q + 15, all real): one wave each
    auto finish = [&](int q) {
        if (wave == 1) {
            const double piv = Ls[(16 * q + (lane & 15)) * kPLD + 16 * q + (lane & 15)];
            if (lane < 16) lg += log(piv);
            store_subst_flag<WT>(piv, true, dinv + kDinvFlags + q, lane);
        } else if (wave == 2) {
            // column-oriented forward substitution of L_qq x = e_c, lane c < 16
            if (lane < 16) {
                const double *D = Ls + 16 * q * kPLD + 16 * q;   // D(i, k) at D[k * kPLD + i]
                const double *rd = rdiag + 16 * q;
                double x[16];
                const double one = opaque_f64(1.0), zero = opaque_f64(0.0);
#pragma unroll
                for (int i = 0; i < 16; ++i) x[i] = i == lane ? one : zero;
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    x[k] = x[k] * rd[k];
#pragma unroll
                    for (int i = k + 1; i < 16; ++i) x[i] -= D[k * kPLD + i] * x[k];
                }
                double *W = dinv + q * 256 + lane * 16;
#pragma unroll
                for (int i = 0; i < 16; ++i)
                    st_f64<WT>(W + i, x[i]);
            }
        } else if (wave == 3) {
            for (int e = lane; e < 16 * B; e += 64) {
                const int col = 16 * q + e / B, r = e % B;
                if (r >= col) st_f64<WT>(Ag + (size_t)col * np + r, Ls[col * kPLD + r]);
            }
        }
    };
    __syncthreads();
    for (int j = 0; j < 4; ++j) {
        if (wave == 0) {
            double p[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) p[k] = Ls[(16 * j + k) * kPLD + lane];
            int fail = 0;
            // four columns at a time: inside a group each column first takes
            // the updates of the group's earlier columns (multipliers by
            // v_readlane) and is then factored; the finished group goes to
            // LDS once and the panel's later columns take its four updates,
            // the multipliers read back as broadcasts (one wave: its LDS
            // operations complete in order).  Every entry still receives
            // the updates of columns 0, 1, 2, ... in that order.  After a
            // failed pivot the remaining columns carry garbage; the block is
            // abandoned at the barrier.  REAL: every column of the panel is
            // below n (no per-column tests).
            auto panel = [&](auto real_tag) {
                constexpr bool REAL = decltype(real_tag)::value;
#pragma unroll
                for (int kb = 0; kb < 16; kb += 4) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const int k = kb + t, c = 16 * j + k;
                        if (REAL || c < ncol) {
#pragma unroll
                            for (int t2 = 0; t2 < t; ++t2)
                                p[k] = fma(-p[kb + t2], readlane_f64(p[kb + t2], c), p[k]);
                            const double piv = readlane_f64(p[k], c);
                            if (fail == 0 && !(piv > 0.0)) fail = c0 + c + 1;
                            double d, rinv;
                            sqrt_rsqrt(piv, d, rinv);
                            if (lane == 0) rdiag[c] = rinv;
                            p[k] = lane == c ? d : p[k] * rinv;
                        }
                    }
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        if (REAL || 16 * j + kb + t < ncol) Ls[(16 * j + kb + t) * kPLD + lane] = p[kb + t];
                    __builtin_amdgcn_wave_barrier();
                    if (REAL) {
#pragma unroll
                        for (int k2 = kb + 4; k2 < 16; ++k2)
#pragma unroll
                            for (int t = 0; t < 4; ++t)
                                p[k2] = fma(-p[kb + t], Ls[(16 * j + kb + t) * kPLD + 16 * j + k2], p[k2]);
                    } else {
                        const int tn = min(4, ncol - (16 * j + kb));   // real columns of the group
#pragma unroll
                        for (int k2 = kb + 4; k2 < 16; ++k2)
#pragma unroll
                            for (int t = 0; t < 4; ++t)
                                if (t < tn)
                                    p[k2] = fma(-p[kb + t], Ls[(16 * j + kb + t) * kPLD + 16 * j + k2], p[k2]);
                    }
                }
            };
            if (16 * j + 16 <= ncol) panel(std::true_type{});
            else panel(std::false_type{});
            if (lane == 0) *flag = fail;
        } else if (full && j > 0) {
            finish(j - 1);
        }
        __syncthreads();
        if (*flag != 0) return *flag;
        // trailing blocks (bi, bj), j < bj <= bi <= 3, round robin over the waves
        int t = 0;
        for (int bj = j + 1; bj < 4; ++bj)
            for (int bi = bj; bi < 4; ++bi, ++t) {
                if ((t & 3) != wave) continue;
                d4 acc;
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[v] = Ls[(16 * bj + 4 * v + lq) * kPLD + 16 * bi + lr];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const double af = Ls[(16 * j + 4 * kk + lq) * kPLD + 16 * bi + lr];
                    const double bf = Ls[(16 * j + 4 * kk + lq) * kPLD + 16 * bj + lr];
                    acc = mfma_f64_sub(bf, af, acc);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) Ls[(16 * bj + 4 * v + lq) * kPLD + 16 * bi + lr] = acc[v];
            }
        if (j < 3) __syncthreads();
    }
    if (full) {
        finish(3);
    } else {
        // columns at or beyond n become the identity; then every panel is
        // completed here
        for (int e = tid; e < B * B; e += kThreads) {
            const int col = e / B, r = e % B;
            if (col >= ncol) Ls[col * kPLD + r] = r == col ? 1.0 : 0.0;
        }
        if (tid >= ncol && tid < B) rdiag[tid] = 1.0;
        __syncthreads();
        if (wave == 1) {
            // logs of the real columns only
            for (int q = 0; q < 4; ++q) {
                const double piv = Ls[(16 * q + (lane & 15)) * kPLD + 16 * q + (lane & 15)];
                const bool real = 16 * q + (lane & 15) < ncol;
                if (lane < 16 && real) lg += log(piv);
                store_subst_flag<WT>(piv, real, dinv + kDinvFlags + q, lane);
            }
        } else {
            for (int q = 0; q < 4; ++q) {
                const int w = wave;   // waves 0, 2, 3 share the stores and inverses
                if (w == 2) {
                    if (lane < 16) {
                        const double *D = Ls + 16 * q * kPLD + 16 * q;
                        const double *rd = rdiag + 16 * q;
                        double x[16];
                        const double one = opaque_f64(1.0), zero = opaque_f64(0.0);
#pragma unroll
                        for (int i = 0; i < 16; ++i) x[i] = i == lane ? one : zero;
#pragma unroll
                        for (int k = 0; k < 16; ++k) {
                            x[k] = x[k] * rd[k];
#pragma unroll
                            for (int i = k + 1; i < 16; ++i) x[i] -= D[k * kPLD + i] * x[k];
                        }
                        double *W = dinv + q * 256 + lane * 16;
#pragma unroll
                        for (int i = 0; i < 16; ++i)
                    st_f64<WT>(W + i, x[i]);
                    }
                } else {
                    for (int e = lane + (w == 3 ? 64 : 0); e < 16 * B; e += 128) {
                        const int col = 16 * q + e / B, r = e % B;
                        if (r >= col) st_f64<WT>(Ag + (size_t)col * np + r, Ls[col * kPLD + r]);
                    }
                }
            }
        }
    }
    // sum of the logs: wave 1, lanes 0..15
    if (wave == 1) {
        for (int off = 8; off > 0; off >>= 1) lg += __shfl_xor(lg, off, 16);
        if (lane == 0) *lgsum = lg;
    }
    __syncthreads();
    if (tid == 0) *logdet = *lgsum;
    return 0;
}

// Forward substitution inside one 16x16 diagonal block L_jj (L_jj(i, k) at
// Ljj[k * ldl + i], W(c, c) = 1 / L(c, c) in Wj) of the 16 rows whose
// right-hand sides r hold column 4v + l/16 of row l % 16 in lane l,
// component v: y_c = (r_c - sum_{k<c} y_k L(c, k)) / L(c, c), column after
// column as in dpotf2, the division a product with W(c, c).  It runs only
// for flagged (ill-conditioned) blocks, as a call: one out-of-line copy
// keeps its registers and code out of the solves' hot path (inlined into
// every solve it cost the factorisation 2-4 %).
__device__ __attribute__((noinline)) d4 subst16(d4 r, const double *__restrict__ Wj, const double *__restrict__ Ljj,
                                                int ldl)
{
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll 1
        for (int q = 0; q < 4; ++q) {
            const int c = 4 * v + q;
            const double yc = __shfl(r[v] * Wj[c * 16 + c], lr + 16 * q, 64);
            if (lq == q) r[v] = yc;
#pragma unroll
            for (int v2 = v; v2 < 4; ++v2) {
                const int col = 4 * v2 + lq;
                if (col > c) r[v2] = fma(-yc, Ljj[(size_t)c * ldl + col], r[v2]);
            }
        }
    }
    return r;
}

// X = C L_bb^{-T} in place on the 64x64 tile at Ct (leading dimension ldc),
// L_bb the factored diagonal block at Lb (leading dimension ldl) with the
// inverses of its 16x16 diagonal blocks in dinv; IDENT: C is the identity
// (the tile is only written), giving L_bb^{-T} itself.  Each wave solves its
// 16 rows alone: with Y = X^T (X the wave's 16 x 64 slice of the result),
// L_bb Y = C^T is block forward substitution over four 16-row groups,
//     Y_j = W_j (C_j^T - sum_{k<j} L_jk Y_k),    W_j = L_jj^{-1} (dinv),
// every product on the matrix cores.  An MFMA result holds D[4v + l/16][l%16]
// in lane l, component v: the B-operand layout of four consecutive k-steps,
// so each Y_k feeds the later products straight from registers.  No LDS, no
// barrier; the tile is read completely before it is overwritten.
// The solve of one wave's 16 rows: t[j][v] = C(lr, 16j + 4v + lq) in lane
// (lr, lq) = (l % 16, l / 16) gives y = X = C L_bb^{-T} in the same layout.
// L_jk (j > k) in A layout (lane l holds L_jk[l%16][4kk + l/16]) and the
// 16x16 inverses likewise, each read where it is used (held all at once
// they take 80 registers).
__device__ inline void trsm16_rows(const d4 t[4], const double *__restrict__ Lb, int ldl,
                                   const double *__restrict__ dinv, d4 y[4])
{
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15, lq = lane >> 4;
    // the four blocks' flags up front: one load latency, not four in the chain
    double flag[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) flag[j] = dinv[kDinvFlags + j];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        d4 r = t[j];
#pragma unroll
        for (int k = 0; k < j; ++k) {
            double lf[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) lf[kk] = Lb[(size_t)(16 * k + 4 * kk + lq) * ldl + 16 * j + lr];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) r = mfma_f64_sub(lf[kk], y[k][kk], r);
        }
        const double *Wj = dinv + j * 256;
        if (flag[j] != 0.0) {
            // an ill-conditioned diagonal block (store_subst_flag): forward
            // substitution inside it, column c of the block after column c - 1
            // as in dpotf2, y_c = (r_c - sum_{k<c} y_k L(c, k)) / L(c, c),
            // the division a product with W(c, c) = 1 / L(c, c); column c of
            // row lr sits in lane lr + 16 (c % 4), component c / 4
            y[j] = subst16(r, Wj, Lb + (size_t)(16 * j) * ldl + 16 * j, ldl);
            continue;
        }
        double wf[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) wf[kk] = Wj[(4 * kk + lq) * 16 + lr];
        d4 o = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) o = mfma_f64(wf[kk], r[kk], o);
        y[j] = o;
    }
}

template <bool IDENT, bool COPY = false, bool WT = false>
__device__ inline void trsm64_solve(const double *Cin, int ldi, double *Ct, int ldc,
                                    const double *__restrict__ Lb, int ldl, const double *__restrict__ dinv,
                                    double *__restrict__ copy = nullptr)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const double *Ciw = Cin + 16 * wave + lr;
    double *Cw = Ct + 16 * wave + lr;
    d4 t[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            t[j][v] = IDENT ? ((16 * j + 4 * v + lq == 16 * wave + lr) ? 1.0 : 0.0)
                            : Ciw[(size_t)(16 * j + 4 * v + lq) * ldi];
    d4 y[4];
    trsm16_rows(t, Lb, ldl, dinv, y);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) st_f64<WT>(Cw + (size_t)(16 * j + 4 * v + lq) * ldc, y[j][v]);
    if constexpr (COPY) {
        // the solved tile also on chip: copy[c * kPLD + r]
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) copy[(16 * j + 4 * v + lq) * kPLD + 16 * wave + lr] = y[j][v];
    }
}

// The row solve of the 16 rows at `row` (a multiple of 16, below the
// 128-column panel c0) of the matrix at A by one wave, with no LDS and no
// barrier (blocks b0 = c0, b1 = c0 + 64; dinv_s: the matrix's 16x16
// inverses):
//     X0 = C(row, b0) L(b0, b0)^{-T}
//     C(row, b1) -= X0 L(b1, b0)^T
//     X1 = C(row, b1) L(b1, b1)^{-T}.
// The middle product runs as tile_gemm_sub runs it (C^T accumulators, the
// L(b1, b0) fragment as the first operand, k in steps of four from c0), its
// A operand -X0 taken from the solve's registers: an MFMA result holds four
// consecutive k-steps of an operand.  Each element sees the operations of
// the 64-row tile routines in the same order, so the bits are theirs.
// WT: the results are stored past the XCD's L2 (st_f64<true>).
template <bool WT = false>
__device__ inline void solve_rows16(double *A, int np, int c0, int row, const double *dinv_s)
{
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15, lq = lane >> 4;
    const int c1 = c0 + kNBI;
    double *C0 = A + (size_t)c0 * np + row + lr;   // C(row + lr, c0 + c) at C0[c * np]
    double *C1 = A + (size_t)c1 * np + row + lr;
    d4 t[4], x0[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) t[j][v] = C0[(size_t)(16 * j + 4 * v + lq) * np];
    trsm16_rows(t, A + (size_t)c0 * np + c0, np, dinv_s + (size_t)(c0 / kNBI) * kDinv, x0);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) st_f64<WT>(C0 + (size_t)(16 * j + 4 * v + lq) * np, x0[j][v]);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) t[j][v] = C1[(size_t)(16 * j + 4 * v + lq) * np];
    const double *Bg = A + (size_t)c0 * np + c1 + lr;   // L(c1 + r, c0 + k) at Bg[k * np + r]
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
        const double af = -x0[kk / 4][kk % 4];
#pragma unroll
        for (int j = 0; j < 4; ++j) t[j] = mfma_f64(Bg[(size_t)(4 * kk + lq) * np + 16 * j], af, t[j]);
    }
    d4 x1[4];
    trsm16_rows(t, A + (size_t)c1 * np + c1, np, dinv_s + (size_t)(c1 / kNBI) * kDinv, x1);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) st_f64<WT>(C1 + (size_t)(16 * j + 4 * v + lq) * np, x1[j][v]);
}

// Ls -= X X^T on the lower 16x16 blocks of a 64x64 block on chip
// (Ls[c * kPLD + r]), X(r, k) at X[k * ldx + r] on chip or in global memory,
// 64 deep on the matrix cores with the accumulation order of tile_gemm_sub
// (so the same bits as that routine's depth-64 update of the same block).
__device__ inline void syrk64(double *__restrict__ Ls, const double *__restrict__ X, int ldx)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    // lower blocks t = 0 .. 9 in column order, block t on wave t mod 4; one
    // block at a time (a wave's blocks are not unrolled together, which would
    // hold all their operand loads in registers at once)
#pragma unroll 1
    for (int t = wave; t < 10; t += 4) {
        const int bj = t < 4 ? 0 : t < 7 ? 1 : t < 9 ? 2 : 3;
        const int bi = t - (bj == 0 ? 0 : bj == 1 ? 3 : bj == 2 ? 5 : 6);
        d4 acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = Ls[(16 * bj + 4 * v + lq) * kPLD + 16 * bi + lr];
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            const double af = X[(size_t)(4 * kk + lq) * ldx + 16 * bi + lr];
            const double bf = X[(size_t)(4 * kk + lq) * ldx + 16 * bj + lr];
            acc = mfma_f64_sub(bf, af, acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) Ls[(16 * bj + 4 * v + lq) * kPLD + 16 * bi + lr] = acc[v];
    }
}

// L(row0:row0+64, c0:c0+64) = C L_bb^{-T} in place, C the tile and L_bb the
// factored diagonal block at (c0, c0) of the matrix at A (the panel solve).
__device__ inline void trsm64_mfma(double *__restrict__ A, int np, int c0, int row0,
                                   const double *__restrict__ dinv)
{
    trsm64_solve<false>(A + (size_t)c0 * np + row0, np, A + (size_t)c0 * np + row0, np, A + (size_t)c0 * np + c0, np,
                        dinv);
}

}  // namespace fb
