// fb_potrf.hip -- batched blocked FP64 Cholesky of every particle's matrix.
//
// Replaces dpotrf('L') (src/utils/linalg.f90:73-98, called from
// src/model/gpr/likelihood.f90:145) for all particles at once, with LAPACK's
// INFO convention, plus the log-determinant of src/utils/arith.f90:72-93.
//
// Left-looking blocked algorithm with two block levels.  Panels of width
// kNBO = 128 carry the dominant trailing update; inside a panel two blocks of
// kNBI = 64 (b0 = c0, b1 = c0 + 64) are factored.  Two schedules: the chained
// one (two launches per panel, or one with the row solve fused; below, the
// default for 16-deep chunks and swarms up to kChainMaxSwarm), and three
// launches per panel:
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
// (b1, b1) (syrk64) and factor that block.
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
    syrk64(Ls, Xs, kPLD);
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

// ---- chained lock-step schedule (two launches per 128-column panel) ------
//
// The diagonal chain of a panel -- factor (b0, b0), solve (b1, b0) with it,
// apply that to (b1, b1) and factor it -- runs inside the update launch, on
// three role work-groups per particle whose updates proceed in parallel:
//   D updates (b0, b0) on chip and factors it;
//   B updates (b1, b0) on chip, waits for D, solves the tile with L(b0, b0);
//   C updates (b1, b1) on chip, waits for B, applies L(b1, b0) L(b1, b0)^T
//     and factors the block.
// The remaining row tiles of the panel are then solved in one row-local
// launch (solve_rows_kernel: solve with L(b0, b0), the 64-deep update of
// their b1 part, solve with L(b1, b1)).  Same tile routines, same order: the
// bits of the three-launch schedule.
//
// Hand-over.  A producer publishes a flag word (the panel's tag, increasing
// within an evaluation) after its stores have reached its XCD's L2; the
// consumer announces at its start which XCD it runs on (tagged), and the
// producer writes its L2 back to memory before publishing only when the
// consumer is elsewhere or not yet known.  The role work-groups of particle
// s get block ids congruent to s modulo 8, which the dispatcher places on one
// XCD, so the write-back is normally skipped.  The consumer polls the flag
// with device-scope atomics (they go to memory), then drops its CU's L1 and
// the clean non-coherent L2 lines before reading.  Role work-groups come
// first in the grid; at most two per particle wait, fewer than the launch's
// resident slots (kChainMaxSwarm), and D never waits, so every wait ends.
// In the fused variant every row tile waits too (for F2): there the
// argument is the dispatch order -- work-groups are dispatched in
// increasing id on each XCD, every role has a lower id than every row tile
// of its XCD, so each role holds a slot (or has finished) before any row
// tile can occupy one, and the roles' own waits end as above.  Every wait is
// still bounded at 1 s of the 100 MHz clock (INFO kInfoWaitTimeout).  The
// fused roles write INFO past their L2 (set_info<true>) as they write their
// results, since the row tiles reading it after F2 may run on other XCDs.

// INFO of a particle; WT: a device-scope store that goes past the XCD's L2
// (read by work-groups of other XCDs after a flag published without an L2
// write-back).
template <bool WT>
__device__ inline void set_info(int *p, int v)
{
    if constexpr (WT) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

__device__ inline int chain_xcc()
{
    // hwreg(HW_REG_XCC_ID, 0, 4): id 20, offset 0, 4 bits
    return (int)(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 0xf);
}

// Consumer, at its start: tag and XCD.
__device__ inline void chain_announce(int *w, int tag)
{
    if (threadIdx.x == 0) __hip_atomic_store(w, (tag << 4) | chain_xcc(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Producer: every thread's stores are done; publish `tag` at *flag, writing
// the L2 back first unless the consumer announced this XCD for this tag.
__device__ inline void chain_publish(int *flag, const int *consumer, int tag)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        if (__hip_atomic_load(consumer, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != ((tag << 4) | chain_xcc()))
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(flag, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Consumer: wait until *flag >= tag (thread 0 polls, bounded at `limit`
// ticks of the 100 MHz clock: 1 s normally), then drop this CU's L1 and the
// clean non-coherent L2 lines.  False on timeout.
__device__ inline bool chain_wait(const int *flag, int tag, int *ok, unsigned long long limit)
{
    if (threadIdx.x == 0) {
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        int good = 1;
        while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < tag) {
            __builtin_amdgcn_s_sleep(2);
            if (__builtin_amdgcn_s_memrealtime() - t0 >= limit) {
                good = 0;
                break;
            }
        }
        asm volatile("buffer_inv sc1\n\ts_waitcnt vmcnt(0)" ::: "memory");
        *ok = good;
    }
    __syncthreads();
    return *ok != 0;
}

// Producer whose results were written to memory (st_f64<true>): publish
// `tag` once every thread's stores are done.
__device__ inline void chain_publish_mem(int *flag, int tag)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The row solve of one 64-row tile at row0 below panel c0 of the matrix at
// A: each wave solves its 16 rows alone (solve_rows16: solve with
// L(b0, b0), the 64-deep update of the b1 part, solve with L(b1, b1);
// dinv_s: the particle's 16x16 inverses).  WT: stored past the XCD's L2.
// Rows at or beyond nreal are padding: their entries are zero left of the
// diagonal and stay zero (every product they take has a zero factor), so a
// wave whose 16 rows are all padding skips its solve.
template <bool WT = false>
__device__ inline void solve_rows_tile(double *A, int np, int c0, int row0, const double *dinv_s, int nreal)
{
    const int row = row0 + 16 * (threadIdx.x >> 6);
    if (row < nreal) solve_rows16<WT>(A, np, c0, row, dinv_s);
}

// The update of the MB x NB tile at (row0, col0) over columns k0 .. k1 of
// the matrix at A, with rows at or beyond nreal (padding, zero left of the
// diagonal: their products are all zero) left out where that saves whole
// 32-row halves of a 64-row tile; each element's products in the same
// order, so the same bits.  Returns without work when every row is padding.
template <int MB, int NB, int KC, int DIST, class Hook = NoHook>
__device__ __forceinline__ void row_tile_update(double *A, int np, int row0, int col0, int k0, int k1, int nreal,
                                                double *smem, int kh = -1, const Hook &hook = Hook())
{
    if constexpr (MB == 64 && (32 * KC / 2) % kThreads == 0) {   // the 32-row tile's loads fit its chunk
        if (row0 + 32 >= nreal) {
            gemm_tile_update<32, NB, KC, DIST, false, Hook>(A, np, row0, col0, k0, k1, smem, kh, hook);
            return;
        }
    }
    if (row0 >= nreal) {
        // nothing to update; a hook (a wait) still runs
        if constexpr (Hook::kActive) hook();
        return;
    }
    gemm_tile_update<MB, NB, KC, DIST, false, Hook>(A, np, row0, col0, k0, k1, smem, kh, hook);
}

// Hook of an update that takes the previous panel's columns last
// (kTileAhead): waits for the particle's flag(s) (rows that role
// work-groups solved against the previous panel in this launch); a wait
// that reaches its bound is recorded in INFO.
struct FlagWait {
    static constexpr bool kActive = true;
    const int *f1, *f2;    // f2 may be null
    int tag;
    int *info_s;
    int *ok;
    unsigned long long limit;
    __device__ void operator()() const
    {
        bool good = chain_wait(f1, tag, ok, limit);
        if (f2) good = chain_wait(f2, tag, ok, limit) && good;
        if (!good && threadIdx.x == 0) set_info<true>(info_s, kInfoWaitTimeout);
    }
};

// LDS of the chained update launch: its row tiles, the 64 x 64 and 64 x 128
// tiles of the roles and of the tail pre-update (whatever the row-tile
// shape), the diagonal factorisation.
template <int KC, int MB, int NB>
constexpr int chain_smem()
{
    constexpr int a = panel_smem<KC, MB, NB>(), c = TileShape<kMB, kNBO, KC>::SMEM;
    return a > c ? a : c;
}

// Role work-groups per 8 particles: D, B, C (TWO = false), or D and BC
// (TWO = true: one work-group updates the whole 64 x 128 row tile of b1,
// then waits for D, solves (b1, b0), applies it to (b1, b1) and factors it
// -- fewer work-groups holding slots, a longer chain after the update).
template <bool TWO>
constexpr int chain_roles() { return TWO ? 16 : 24; }

// XCD-aware order of T tiles numbered particle-major: the dispatcher places
// block b on XCD b mod 8, so block base + t (base a multiple of 8) takes
// tile (t mod 8) * ceil(T / 8) + t / 8 -- each XCD gets one contiguous
// eighth of the tiles, and a particle's tiles share one or two L2s (its B
// panel and diagonal blocks are fetched into those only).  -1: padding.
__device__ inline int xcd_tile(int t, int T)
{
    const int c = ceil_div(T, 8);
    const int tp = (t % 8) * c + t / 8;
    return tp < T ? tp : -1;
}

// Tail pre-update: the update launch of panel c0 may also apply columns
// k0 .. c0 to the next `tail` panels (their diagonal 64 x 64 block and their
// 64 x 128 row tiles below it), so that those panels' own launches, with few
// rows left and deep updates, only apply columns from c0 on.  Each element
// still takes its products in increasing k (the accumulator goes through
// memory exactly), so the bits do not change.
__host__ __device__ inline int tail_tiles(int np, int c0, int tail)
{
    int t = 0;
    for (int j = 1; j <= tail; ++j) t += 1 + (np - c0 - kNBO * j - kNBI) / kNBI;
    return t;
}

// Launch 1 of panel c0, chained: role work-groups first (block R g + 8 r + l
// is role r of particle 8 g + l, R = chain_roles), then the MB x NB update
// tiles of the rows below the panel, then the tail pre-update tiles (both in
// XCD-aware order).  The panel's own updates apply columns k0 .. c0 (k0 > 0:
// an earlier launch applied the first k0).  FUSED (64 x 128 row tiles): the
// roles write their results to memory and the last one publishes F2; each
// row tile then solves its rows (solve_rows_tile) -- the panel in one launch.
// Role work-group b (< R ceil(S / 8)) of the chained update launch of panel
// c0 for the S matrices at K (see panel_chain_kernel).
// AHEAD (kTileAhead, c0 >= 128): D and B (BC) first solve rows b0 and b1
// against the previous panel (stored past the L2: every row tile of the
// particle reads them) and publish S0 / S1; the updates that read rows
// solved by another role take the previous panel's columns last, after its
// flag.
template <int KC, bool TWO, bool FUSED, bool AHEAD = false, int DIST = 2>
__device__ __forceinline__ void chain_role(int b, double *__restrict__ K, int np, int n, int c0, int k0, int S,
                                           int *__restrict__ info, double *__restrict__ logdet_part,
                                           double *__restrict__ dinv, int *__restrict__ chain,
                                           unsigned long long wait_limit, double *smem, int &ok)
{
    constexpr int R = chain_roles<TWO>();
    const int role = (b % R) / 8;
    const int s = 8 * (b / R) + b % 8;
    if (s >= S) return;
    const int tag = c0 / kNBO + 1, c1 = c0 + kNBI;
    const size_t nbk = (size_t)np / kNBI;
    double *A = K + (size_t)s * np * np;
    int *cw = chain + (size_t)s * kChainStride;
    double *dinv0 = dinv + ((size_t)s * nbk + c0 / kNBI) * kDinv;
    const int cp = c0 - kNBO;   // the previous panel (AHEAD)
    if constexpr (AHEAD) {
        // the role's own 64 rows against the previous panel (D: b0, B / BC:
        // b1; C reads B's)
        if (role <= 1) {
            if (info[s] == 0)
                solve_rows_tile<true>(A, np, cp, role == 0 ? c0 : c1, dinv + (size_t)s * nbk * kDinv, n + 2);
            chain_publish_mem(cw + (role == 0 ? kChainS0 : kChainS1), tag);
        }
    }
    if (role == 0) {
        if (info[s] == 0) {
            gemm_tile_update<kMB, kNBI, KC, DIST, true>(A, np, c0, c0, k0, c0, smem);
            __syncthreads();
            double lg = 0.0;
            const int fail = potrf64_block<true, FUSED>(A, np, n, c0, &lg, smem, dinv0);
            if (threadIdx.x == 0) {
                if (fail) set_info<FUSED>(info + s, fail);
                else logdet_part[(size_t)s * nbk + c0 / kNBI] = lg;
            }
        }
        chain_publish(cw + kChainF0, cw + kChainXccB, tag);
    } else if (TWO) {
        // BC: the b1 row tile (b1, b0:b1), then the rest of the chain
        chain_announce(cw + kChainXccB, tag);
        if constexpr (AHEAD)
            gemm_tile_update<kMB, kNBO, KC, DIST, false, FlagWait>(
                A, np, c1, c0, k0, c0, smem, cp, FlagWait{cw + kChainS0, nullptr, tag, info + s, &ok, wait_limit});
        else
            gemm_tile_update<kMB, kNBO, KC, DIST>(A, np, c1, c0, k0, c0, smem);
        if (!chain_wait(cw + kChainF0, tag, &ok, wait_limit)) {
            if (threadIdx.x == 0) set_info<FUSED>(info + s, kInfoWaitTimeout);
            if constexpr (FUSED) chain_publish_mem(cw + kChainF2, tag);
            return;
        }
        if (info[s] != 0) {
            if constexpr (FUSED) chain_publish_mem(cw + kChainF2, tag);
            return;
        }
        // L(b1, b0) solved in place, then re-read by every wave (global
        // stores of a work-group are visible to it after the barrier)
        double *Ls = smem;                       // block (b1, b1), then its factorisation
        double *tile = A + (size_t)c0 * np + c1;
        trsm64_solve<false, false, FUSED>(tile, np, tile, np, A + (size_t)c0 * np + c0, np, dinv0);
        const double *Cg = A + (size_t)c1 * np + c1;
        for (int e = threadIdx.x; e < kNBI * kNBI / 2; e += kThreads) {
            const int col = e / (kNBI / 2), r2 = (e % (kNBI / 2)) * 2;
            *reinterpret_cast<double2 *>(Ls + col * kPLD + r2) = *reinterpret_cast<const double2 *>(Cg + (size_t)col * np + r2);
        }
        __syncthreads();
        syrk64(Ls, tile, np);
        __syncthreads();
        double lg = 0.0;
        const int fail = potrf64_block<true, FUSED>(A, np, n, c1, &lg, Ls,
                                                    dinv + ((size_t)s * nbk + c1 / kNBI) * kDinv);
        if (threadIdx.x == 0) {
            if (fail) set_info<FUSED>(info + s, fail);
            else logdet_part[(size_t)s * nbk + c1 / kNBI] = lg;
        }
        if constexpr (FUSED) chain_publish_mem(cw + kChainF2, tag);
    } else if (role == 1) {
        chain_announce(cw + kChainXccB, tag);
        if constexpr (AHEAD)
            gemm_tile_update<kMB, kNBI, KC, DIST, true, FlagWait>(
                A, np, c1, c0, k0, c0, smem, cp, FlagWait{cw + kChainS0, nullptr, tag, info + s, &ok, wait_limit});
        else
            gemm_tile_update<kMB, kNBI, KC, DIST, true>(A, np, c1, c0, k0, c0, smem);
        if (!chain_wait(cw + kChainF0, tag, &ok, wait_limit)) {
            if (threadIdx.x == 0) set_info<FUSED>(info + s, kInfoWaitTimeout);
        } else if (info[s] == 0) {
            trsm64_solve<false, false, FUSED>(smem, kPLD, A + (size_t)c0 * np + c1, np, A + (size_t)c0 * np + c0,
                                              np, dinv0);
        }
        chain_publish(cw + kChainF1, cw + kChainXccC, tag);
    } else {
        chain_announce(cw + kChainXccC, tag);
        if constexpr (AHEAD)
            gemm_tile_update<kMB, kNBI, KC, DIST, true, FlagWait>(
                A, np, c1, c1, k0, c0, smem, cp, FlagWait{cw + kChainS1, nullptr, tag, info + s, &ok, wait_limit});
        else
            gemm_tile_update<kMB, kNBI, KC, DIST, true>(A, np, c1, c1, k0, c0, smem);
        if (!chain_wait(cw + kChainF1, tag, &ok, wait_limit)) {
            if (threadIdx.x == 0) set_info<FUSED>(info + s, kInfoWaitTimeout);
            if constexpr (FUSED) chain_publish_mem(cw + kChainF2, tag);
            return;
        }
        if (info[s] != 0) {
            if constexpr (FUSED) chain_publish_mem(cw + kChainF2, tag);
            return;
        }
        // L(b1, b0), written by role B, read from memory
        syrk64(smem, A + (size_t)c0 * np + c1, np);
        __syncthreads();
        double lg = 0.0;
        const int fail = potrf64_block<true, FUSED>(A, np, n, c1, &lg, smem,
                                                    dinv + ((size_t)s * nbk + c1 / kNBI) * kDinv);
        if (threadIdx.x == 0) {
            if (fail) set_info<FUSED>(info + s, fail);
            else logdet_part[(size_t)s * nbk + c1 / kNBI] = lg;
        }
        if constexpr (FUSED) chain_publish_mem(cw + kChainF2, tag);
    }}

// Row tile lt (< 8 ceil(S rt CT / 8), XCD-aware order) of the chained update
// launch of panel c0: the MB x NB update of the rows below the panel (and,
// FUSED, their row solve once the panel's chain is published).
// AHEAD (64 x 128 tiles): the tile first solves its rows against the
// previous panel, then updates with that panel's columns last, after the
// flags of the two diagonal row blocks.
template <int KC, int MB, int NB, bool FUSED, bool AHEAD = false, int DIST = 2>
__device__ __forceinline__ void chain_rowtile(int lt, double *__restrict__ K, int np, int nreal, int c0, int k0,
                                              int S, int *__restrict__ info, double *__restrict__ dinv,
                                              int *__restrict__ chain, unsigned long long wait_limit,
                                              double *smem, int &ok)
{
    constexpr int CT = kNBO / NB;
    const int rt = (np - c0 - kNBO) / MB;
    const int t = xcd_tile(lt, S * rt * CT);
    if (t < 0) return;
    const int s = t / (rt * CT), u = t % (rt * CT);
    if (info[s] != 0) return;
    double *A = K + (size_t)s * np * np;
    const int row0 = c0 + kNBO + (u / CT) * MB;
    if constexpr (AHEAD) {
        static_assert(MB == kNBI && NB == kNBO && !FUSED, "the ahead row solve needs 64 x 128 row tiles");
        solve_rows_tile(A, np, c0 - kNBO, row0, dinv + (size_t)s * (np / kNBI) * kDinv, nreal);
        __syncthreads();   // each wave solved 16 rows; the update reads all 64
        const int *cw = chain + (size_t)s * kChainStride, tag = c0 / kNBO + 1;
        row_tile_update<MB, NB, KC, DIST, FlagWait>(
            A, np, row0, c0, k0, c0, nreal, smem, c0 - kNBO,
            FlagWait{cw + kChainS0, cw + kChainS1, tag, info + s, &ok, wait_limit});
        return;
    }
    row_tile_update<MB, NB, KC, DIST>(A, np, row0, c0 + (u % CT) * NB, k0, c0, nreal, smem);
    if constexpr (FUSED) {
        static_assert(MB == kNBI && NB == kNBO, "the fused row solve needs 64 x 128 row tiles");
        const int tag = c0 / kNBO + 1;
        if (!chain_wait(chain + (size_t)s * kChainStride + kChainF2, tag, &ok, wait_limit)) {
            if (threadIdx.x == 0) set_info<FUSED>(info + s, kInfoWaitTimeout);
            return;
        }
        if (info[s] != 0) return;
        solve_rows_tile(A, np, c0, row0, dinv + (size_t)s * (np / kNBI) * kDinv, nreal);
    }
}

// Tail pre-update tile lt (< 8 ceil(S tail_tiles / 8), XCD-aware order) of
// the chained update launch of panel c0: columns kt .. c0 applied to the
// diagonal 64 x 64 block or a 64 x 128 row tile of one of the next `tail`
// panels.
template <int KC, int DIST = 2>
__device__ __forceinline__ void chain_tailtile(int lt, double *__restrict__ K, int np, int nreal, int c0, int kt,
                                               int tail, int S, const int *__restrict__ info, double *smem)
{
    const int ct = tail_tiles(np, c0, tail);
    const int t = xcd_tile(lt, S * ct);
    if (t < 0) return;
    const int s = t / ct;
    int v = t % ct, cp = c0;
    for (int j = 1; j <= tail; ++j) {
        cp = c0 + kNBO * j;
        const int cnt = 1 + (np - cp - kNBI) / kNBI;
        if (v < cnt) break;
        v -= cnt;
    }
    if (info[s] != 0) return;
    double *A = K + (size_t)s * np * np;
    if (v == 0) gemm_tile_update<kMB, kNBI, KC, DIST>(A, np, cp, cp, kt, c0, smem);
    else row_tile_update<kMB, kNBO, KC, DIST>(A, np, cp + kNBI * v, cp, kt, c0, nreal, smem);
}

// WPE work-groups per CU (waves per SIMD): 2 (16-deep chunks), 3 or 4
// (8-deep; at 4 the update loads run one chunk ahead instead of two, which
// fits its registers, and the diagonal factorisation spills a little).
template <int KC, int MB, int NB, bool TWO, bool FUSED, bool AHEAD, int WPE>
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(WPE, WPE))) panel_chain_kernel(double *__restrict__ K, int np, int n,
                                                                       int c0, int k0, int kt, int tail, int S,
                                                                       int *__restrict__ info,
                                                                       double *__restrict__ logdet_part,
                                                                       double *__restrict__ dinv,
                                                                       int *__restrict__ chain,
                                                                       unsigned long long wait_limit)
{
    __shared__ __attribute__((aligned(16))) double smem[chain_smem<KC, MB, NB>()];
    __shared__ int ok;
    const int b = blockIdx.x;
    constexpr int R = chain_roles<TWO>();
    const int nrole = R * ceil_div(S, 8);
    if (b < nrole) {
        chain_role<KC, TWO, FUSED, AHEAD, (WPE >= 4 ? 1 : 2)>(b, K, np, n, c0, k0, S, info, logdet_part, dinv, chain,
                                                            wait_limit, smem, ok);
        return;
    }
    constexpr int CT = kNBO / NB;            // column tiles of the panel
    const int rt = (np - c0 - kNBO) / MB;    // row tiles below the panel
    const int nreg = 8 * ceil_div(S * rt * CT, 8);
    if (b - nrole < nreg) {
        chain_rowtile<KC, MB, NB, FUSED, AHEAD, (WPE >= 4 ? 1 : 2)>(b - nrole, K, np, n + 2, c0, k0, S, info, dinv,
                                                                   chain, wait_limit, smem, ok);
        return;
    }
    chain_tailtile<KC, (WPE >= 4 ? 1 : 2)>(b - nrole - nreg, K, np, n + 2, c0, kt, tail, S, info, smem);
}

// Row-solve tile lt (< 8 ceil(S rt / 8), XCD-aware order) of panel c0.
__device__ __forceinline__ void solve_rows_block(int lt, double *__restrict__ K, int np, int nreal, int c0, int S,
                                                 const int *__restrict__ info, const double *__restrict__ dinv)
{
    const int rt = (np - c0 - kNBO) / kNBI;
    const int t = xcd_tile(lt, S * rt);
    if (t < 0) return;
    const int s = t / rt;
    const int row0 = c0 + kNBO + (t % rt) * kNBI;
    if (info[s] != 0) return;
    solve_rows_tile(K + (size_t)s * np * np, np, c0, row0, dinv + (size_t)s * (np / kNBI) * kDinv, nreal);
}

// Launch 2 of panel c0, chained: every row tile below the panel is solved
// (solve_rows_tile: launches 2 and 3 of the three-launch schedule for one
// tile, each wave on its own 16 rows; no LDS).
// (three waves per SIMD: at most 168 VGPRs)
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3)))
solve_rows_kernel(double *__restrict__ K, int np, int nreal, int c0, int S, const int *__restrict__ info,
                  const double *__restrict__ dinv)
{
    solve_rows_block(blockIdx.x, K, np, nreal, c0, S, info, dinv);
}

// ---- several swarms in one chained launch ----------------------------------
//
// The swarms of several training sets (different N) share each launch of the
// chained schedule: segment g of a launch is one panel (column c0_g) of
// swarm g.  Block ids: the role work-groups of every segment first (segment
// g from role_end[g-1], a multiple of 8, so a particle's three roles keep one
// XCD and the order D < B < C), then the row tiles of every segment, each
// segment padded to a multiple of 8 in XCD-aware order.  Per matrix the
// tiles, their order and the hand-over are those of panel_chain_kernel for
// the same tile code, so each matrix gets the bits of its own evaluation.
//
// Progress: the waits no longer rest on a bound of the swarm size (all the
// segments' particles together exceed kChainMaxSwarm) but on the dispatch
// order, as in the fused variant: on each XCD work-groups are dispatched in
// increasing id; B and C of a particle are dispatched after its D (lower id,
// same XCD), which never waits, and C after its B, which waits for D only,
// so each producer a role waits for holds a slot or has finished; the fused
// row tiles come after every role.  Every wait is still bounded (INFO
// kInfoWaitTimeout).
template <int KC, int MB, int NB, bool TWO, bool FUSED>
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(KC == 8 ? 3 : 2, KC == 8 ? 3 : 2))) panel_chain_multi_kernel(ChainMulti m)
{
    __shared__ __attribute__((aligned(16))) double smem[chain_smem<KC, MB, NB>()];
    __shared__ int ok;
    const int b = blockIdx.x;
    int g = 0;
    if (b < m.seg[m.G - 1].role_end) {
        while (b >= m.seg[g].role_end) ++g;
        const ChainSegment &q = m.seg[g];
        chain_role<KC, TWO, FUSED>(b - (g ? m.seg[g - 1].role_end : 0), q.K, q.np, q.n, q.c0, q.k0, q.S, q.info,
                                   q.logdet_part, q.dinv, q.chain, m.wait_limit, smem, ok);
        return;
    }
    if (b < m.seg[m.G - 1].tile_end) {
        while (b >= m.seg[g].tile_end) ++g;
        const ChainSegment &q = m.seg[g];
        chain_rowtile<KC, MB, NB, FUSED>(b - (g ? m.seg[g - 1].tile_end : m.seg[m.G - 1].role_end), q.K, q.np,
                                         q.n + 2, q.c0, q.k0, q.S, q.info, q.dinv, q.chain, m.wait_limit, smem, ok);
        return;
    }
    while (b >= m.seg[g].tail_end) ++g;
    const ChainSegment &q = m.seg[g];
    chain_tailtile<KC>(b - (g ? m.seg[g - 1].tail_end : m.seg[m.G - 1].tile_end), q.K, q.np, q.n + 2, q.c0, q.kt,
                       q.tail, q.S, q.info, smem);
}

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3)))
solve_rows_multi_kernel(ChainMulti m)
{
    const int b = blockIdx.x;
    int g = 0;
    while (b >= m.seg[g].solve_end) ++g;
    const ChainSegment &q = m.seg[g];
    solve_rows_block(b - (g ? m.seg[g - 1].solve_end : 0), q.K, q.np, q.n + 2, q.c0, q.S, q.info, q.dinv);
}

template <int KC, int MB, int NB, bool TWO, bool FUSED = false, bool AHEAD = false, int WPE = (KC == 8 ? 3 : 2)>
static void chain_launch(const SwarmBuffers &b, int c0, int k0, int kt, int tail, hipStream_t st)
{
    const int tiles = ((b.np - c0 - kNBO) / MB) * (kNBO / NB);
    const int nrole = chain_roles<TWO>() * ceil_div(b.S, 8);
    const int grid = nrole + 8 * ceil_div(b.S * tiles, 8) + 8 * ceil_div(b.S * tail_tiles(b.np, c0, tail), 8);
    hipLaunchKernelGGL((panel_chain_kernel<KC, MB, NB, TWO, FUSED, AHEAD, WPE>), dim3(grid), dim3(kThreads), 0, st, b.K,
                       b.np, b.n, c0, k0, kt, tail, b.S, b.info, b.logdet_part, b.dinv, b.chain, b.wait_limit);
}

template <bool TWO>
static void chain_tiles(const SwarmBuffers &b, int tile, int c0, int k0, int kt, int tail, hipStream_t st)
{
    if ((tile & kTileAhead) && (tile & kTileFourPerCU)) chain_launch<8, 64, 128, TWO, false, true, 4>(b, c0, k0, kt, tail, st);
    else if ((tile & kTileAhead) && (tile & kTileThreePerCU)) chain_launch<8, 64, 128, TWO, false, true>(b, c0, k0, kt, tail, st);
    else if (tile & kTileAhead) chain_launch<16, 64, 128, TWO, false, true>(b, c0, k0, kt, tail, st);
    else if (tile & kTileFourPerCU) chain_launch<8, 64, 128, TWO, false, false, 4>(b, c0, k0, kt, tail, st);
    else if (tile & kTileFused) chain_launch<16, 64, 128, TWO, true>(b, c0, k0, kt, tail, st);
    else if (tile & kTileThreePerCU) chain_launch<8, 64, 128, TWO>(b, c0, k0, kt, tail, st);
    else if (tile == 1) chain_launch<16, 32, 128, TWO>(b, c0, k0, kt, tail, st);
    else if (tile == 2) chain_launch<16, 32, 64, TWO>(b, c0, k0, kt, tail, st);
    else chain_launch<16, 64, 128, TWO>(b, c0, k0, kt, tail, st);
}

// Pre-update schedule of the chained update launches: launch p applies
// columns k0 .. c0p to its own panel (the first k0 were applied by earlier
// launches) and columns kt .. c0p to the next `tail` panels.  Launch
// npan - 1 - tail_set pre-updates the last tail_set panels; with lookahead
// every other launch pre-updates the next panel.  Each panel's columns are
// applied in increasing k across launches, so the bits do not change.
// A launch that also solves the previous panel's rows (panel_ahead) takes
// no pre-updates: its tail tiles would read rows solved in the same launch.
static int panel_code(const SwarmBuffers &b, int c0)
{
    return b.panel_tile ? b.panel_tile[c0 / kNBO] : b.tile;
}

bool panel_ahead(const SwarmBuffers &b, int c0)
{
    if (!b.chain || c0 < kNBO || c0 >= b.np) return false;
    const int code = panel_code(b, c0);
    return (code & kTileAhead) && (code & 3) == 0 && !(code & kTileFused) && !(panel_code(b, c0 - kNBO) & kTileFused);
}

static void chain_schedule(int np, int tail_set, int lookahead, int p, int &k0, int &kt, int &tail,
                           const SwarmBuffers *b = nullptr)
{
    const int npan = np / kNBO, q = npan - 1 - tail_set;
    int k[1024];
    const int np_ = npan < 1024 ? npan : 1024;
    for (int r = 0; r < np_; ++r) k[r] = 0;
    for (int r = 0; r <= p && r < np_; ++r) {
        int t = 0;
        if (b && panel_ahead(*b, r * kNBO)) t = 0;
        else if (tail_set > 0 && q >= 1 && r == q) t = tail_set;
        else if (lookahead && r >= 1 && r + 1 < npan) t = 1;
        if (r == p) {
            k0 = k[r];
            kt = r + 1 < np_ ? k[r + 1] : 0;
            tail = t;
        }
        for (int j = 1; j <= t && r + j < np_; ++j) k[r + j] = r * kNBO;
    }
}

// 16- or 8-deep chunks (the 32-deep update tile leaves no room for the
// roles).  Tile code: 0 64x128, 1 32x128, 2 32x64 row tiles, plus
// kTileTwoRoles for the two-role variant; kTileFused and kTileThreePerCU
// select 64x128 variants.
void launch_panel_chain(const SwarmBuffers &b, int c0, hipStream_t st)
{
    const int code = panel_code(b, c0);
    int k0 = 0, kt = 0, tail = 0;
    chain_schedule(b.np, b.tail, b.lookahead, c0 / kNBO, k0, kt, tail, &b);
    const int cu = code & (kTileThreePerCU | kTileFourPerCU);
    const int tile = panel_ahead(b, c0) ? kTileAhead | cu
                     : (code & kTileFused) ? kTileFused : cu ? cu : (code & 3);
    if (code & kTileTwoRoles) chain_tiles<true>(b, tile, c0, k0, kt, tail, st);
    else chain_tiles<false>(b, tile, c0, k0, kt, tail, st);
}

// The panel's update launch also solved its rows (kTileFused): no row-solve
// launch for it.
bool panel_fused(const SwarmBuffers &b, int c0)
{
    return b.chain && (panel_code(b, c0) & kTileFused);
}

// The rows below panel c0 need their own solve launch: neither the panel's
// update launch (fused) nor the next panel's (ahead) solves them.
bool panel_rows_launch(const SwarmBuffers &b, int c0)
{
    return c0 + kNBO < b.np && !panel_fused(b, c0) && !panel_ahead(b, c0 + kNBO);
}

void launch_solve_rows(const SwarmBuffers &b, int c0, hipStream_t st)
{
    const int rt = (b.np - c0 - kNBO) / kNBI;
    if (rt <= 0) return;
    hipLaunchKernelGGL(solve_rows_kernel, dim3(8 * ceil_div(b.S * rt, 8)), dim3(kThreads), 0, st, b.K, b.np,
                       b.n + 2, c0, b.S, b.info, b.dinv);
}

template <int KC, int MB, int NB, bool TWO, bool FUSED>
static void chain_multi_launch(ChainMulti &m, hipStream_t st)
{
    int r = 0;
    for (int g = 0; g < m.G; ++g) m.seg[g].role_end = r += chain_roles<TWO>() * ceil_div(m.seg[g].S, 8);
    int t = r, s = 0;
    for (int g = 0; g < m.G; ++g) {
        ChainSegment &q = m.seg[g];
        chain_schedule(q.np, q.tail_set, 0, q.c0 / kNBO, q.k0, q.kt, q.tail);
        const int rows = q.np - q.c0 - kNBO;
        q.tile_end = t += 8 * ceil_div(q.S * (rows / MB) * (kNBO / NB), 8);
        q.solve_end = s += FUSED ? 0 : 8 * ceil_div(q.S * (rows / kNBI), 8);
    }
    for (int g = 0; g < m.G; ++g) {
        ChainSegment &q = m.seg[g];
        q.tail_end = t += 8 * ceil_div(q.S * tail_tiles(q.np, q.c0, q.tail), 8);
    }
    hipLaunchKernelGGL((panel_chain_multi_kernel<KC, MB, NB, TWO, FUSED>), dim3(t), dim3(kThreads), 0, st, m);
    if (s > 0) hipLaunchKernelGGL(solve_rows_multi_kernel, dim3(s), dim3(kThreads), 0, st, m);
}

template <bool TWO>
static void chain_multi_tiles(ChainMulti &m, hipStream_t st)
{
    if (m.tile & kTileFused) chain_multi_launch<16, 64, 128, TWO, true>(m, st);
    else if (m.tile & kTileThreePerCU) chain_multi_launch<8, 64, 128, TWO, false>(m, st);
    else if ((m.tile & 3) == 1) chain_multi_launch<16, 32, 128, TWO, false>(m, st);
    else if ((m.tile & 3) == 2) chain_multi_launch<16, 32, 64, TWO, false>(m, st);
    else chain_multi_launch<16, 64, 128, TWO, false>(m, st);
}

void launch_chain_multi(ChainMulti &m, hipStream_t st)
{
    if (m.G < 1 || m.G > kMultiMaxSegments) return;
    if (m.tile & kTileTwoRoles) chain_multi_tiles<true>(m, st);
    else chain_multi_tiles<false>(m, st);
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
    // the shape only: the role, fused and three-per-CU variants belong to
    // the chained schedule
    const int tile = (b.panel_tile ? b.panel_tile[c0 / kNBO] : b.tile) & 3;
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
