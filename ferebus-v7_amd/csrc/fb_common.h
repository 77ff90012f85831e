// fb_common.h -- shared definitions of the gfx950 swarm log-likelihood library.
//
// Data layout in HBM (one training set = one handle):
//   slots    n_slots x n FP64 per-point feature values used by the build:
//            the raw feature for rbf / rbf-cyclic dimensions, and the pair
//            sin(x/2), cos(x/2) for periodic ones, so that
//            sin((xi - xj)/2) = sin(xi/2)cos(xj/2) - cos(xi/2)sin(xj/2)
//            needs no per-pair transcendental in the covariance build.
//   RHS      nrhs x n: the right-hand sides the likelihood needs
//            (y - m for the marginal likelihood, y and 1 for the concentrated).
//   K        S matrices of np x np FP64, column major, ld = np, matrix s at
//            K + s*np*np.  np = roundup(n + nrhs, 128).  Rows n .. n+nrhs-1 hold
//            the right-hand sides ("bordered" matrix), so that the Cholesky
//            factorisation of the leading (n+nrhs) block writes
//            z = L^{-1} rhs into those rows; the remaining padding rows and
//            columns form an identity block that factors to itself.
//            Only the lower triangle is ever read or written.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>
#include <stdint.h>

namespace fb {

// exp(x) for the exponent of a stationary kernel term (x <= 0, or NaN).
// Argument reduction after Cody and Waite, as in Tang (ACM TOMS 15, 1989):
// x = k ln2 + r, |r| <= ln2/2, with ln2 split into a high part of 32
// significant bits (k ln2_hi and x - k ln2_hi exact) and the rounded
// remainder; exp(r) by the degree-13 Taylor polynomial in Horner form
// (truncation below 5e-18 relative on |r| <= ln2/2; a host build of this
// routine is within 0.85 ulp of long double exp on [-708, 0], mean signed
// error -0.001 ulp, correctly rounded in 96 % of 400,000 points); 2^k by ldexp, which
// flushes k < -1074 to zero.  x below -1000 is clamped there (exp = 0
// either way), so -inf does not become NaN; NaN stays NaN.  The constants
// c (kExpConsts of them, exp_constants() on the host) come through the
// kernel arguments: in scalar registers they are FMA operands, where
// literals would each need a vector-register copy per use.
constexpr int kExpConsts = 15;   // 1/ln2, ln2 hi, ln2 lo, 1/13! .. 1/2!

__host__ __device__ inline void exp_constants(double *c)
{
    c[0] = 1.4426950408889634074;
    c[1] = 6.93147180369123816490e-01;
    c[2] = 1.90821492927058770002e-10;
    double f = 1.0;
    for (int j = 2; j <= 13; ++j) f *= j;               // 13! (exact in a double)
    for (int j = 13; j >= 2; --j) {
        c[3 + 13 - j] = 1.0 / f;                       // correctly rounded 1/j!
        f /= j;
    }
}

__host__ __device__ inline double exp_nonpos(double x, const double *c)
{
    x = x < -1000.0 ? -1000.0 : x;
    const double k = __builtin_rint(x * c[0]);
    double r = __builtin_fma(-k, c[1], x);
    r = __builtin_fma(-k, c[2], r);
    double p = c[3];
#pragma unroll
    for (int j = 4; j < kExpConsts; ++j) p = __builtin_fma(p, r, c[j]);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)k);
}

// exp(x) for the hyperparameter transforms (l = exp(theta) and its
// relatives, a few per particle): the same reduction as exp_nonpos with ln2
// in three parts, r = x - k ln2 kept as a double-double, and
// exp(r) = 1 + r + r^2 (1/2 + r (1/6 + r R(r))) with the outer three levels
// in double-double arithmetic (R, the tail from 1/24, in double), rounded
// once at the end.  The result before that rounding is within ~2^-63 of
// exp(x), so it is the correctly rounded value except in about one case in
// a thousand -- the value the reference's FP64 host exp gives, so both builds
// start every covariance element from the same lengthscales (a one-ulp
// difference in a lengthscale biases every element it touches).  |x| above
// 700 falls back to the library exp.
__host__ __device__ inline double exp_cr(double x)
{
    if (!(x > -700.0 && x < 700.0)) return exp(x);
    const double L1 = 6.93147180369123816490e-01;     // ln2, 32 significant bits
    const double L2 = 1.90821492927058770002e-10;
    const double L3 = 1.1612227229362532e-26;         // ln2 - L1 - L2 (rounded)
    const double C6H = 1.0 / 6.0, C6L = 9.25185853854297e-18;   // 1/6 = C6H + C6L
    const double k = __builtin_rint(x * 1.4426950408889634074);
    const double t = __builtin_fma(-k, L1, x);        // exact
    const double p = k * L2, pe = __builtin_fma(k, L2, -p);
    const double rh = t - p;
    const double rl = ((t - rh) - p) - pe - k * L3;   // r = rh + rl, |rh| >> |p|
    // R(r) = sum_{j = 4..16} r^(j-4) / j! (the first omitted term, r^17/17!,
    // is below 5e-23 for |r| <= ln2/2)
    double R = 1.0 / 20922789888000.0;                 // 1/16!
    R = __builtin_fma(R, rh, 1.0 / 1307674368000.0);
    R = __builtin_fma(R, rh, 1.0 / 87178291200.0);
    R = __builtin_fma(R, rh, 1.0 / 6227020800.0);
    R = __builtin_fma(R, rh, 1.0 / 479001600.0);
    R = __builtin_fma(R, rh, 1.0 / 39916800.0);
    R = __builtin_fma(R, rh, 1.0 / 3628800.0);
    R = __builtin_fma(R, rh, 1.0 / 362880.0);
    R = __builtin_fma(R, rh, 1.0 / 40320.0);
    R = __builtin_fma(R, rh, 1.0 / 5040.0);
    R = __builtin_fma(R, rh, 1.0 / 720.0);
    R = __builtin_fma(R, rh, 1.0 / 120.0);
    R = __builtin_fma(R, rh, 1.0 / 24.0);
    // U = 1/6 + r R
    const double m = rh * R, me = __builtin_fma(rh, R, -m);
    const double uh = C6H + m, ue = (C6H - uh) + m;   // |C6H| > |m|
    const double ul = ue + me + C6L + rl * R;
    // V = 1/2 + r U
    const double v1 = rh * uh, v1e = __builtin_fma(rh, uh, -v1);
    const double v1l = v1e + rh * ul + rl * uh;
    const double vh = 0.5 + v1, ve = (0.5 - vh) + v1;
    const double vl = ve + v1l;
    // W = r^2 V
    const double sh = rh * rh, sl = __builtin_fma(rh, rh, -sh) + 2.0 * rh * rl;
    const double wh = sh * vh, wl = __builtin_fma(sh, vh, -wh) + sh * vl + sl * vh;
    // 1 + r + W
    const double s = rh + wh, se = (rh - s) + wh;     // |rh| >= |wh|
    const double s_l = se + rl + wl;
    const double a = 1.0 + s, ae = (1.0 - a) + s;     // 1 > |s|
    return __builtin_ldexp(a + (ae + s_l), (int)k);
}

constexpr int kMaxLeaves = 8;     // stationary / constant leaves of the kernel tree
constexpr int kMaxTerms = 16;     // sum-of-products terms after expansion
constexpr int kMaxDims = 512;     // total active dimensions over all leaves

// Leaf kinds on the device (same numbering as FB_LEAF_*).
enum LeafKind : int { kRbf = 1, kPeriodic = 2, kRbfCyclic = 3, kLinear = 4, kConstant = 5 };

// The kernel expression expanded into a sum of products of leaves:
//   K_ij = sum_t  coef_t(s) * exp( sum_{leaf in t} e_leaf(i, j, s) )
// where stationary leaves contribute exponents e_leaf and linear / constant
// leaves contribute the factor coef_t (their values do not depend on x).
struct KernelProgram {
    int n_leaves;
    int n_terms;
    int n_par_kernel;                 // kernel parameters (without the nugget)
    int n_par;                        // theta entries per particle
    int optimise_noise;
    double noise;
    int leaf_kind[kMaxLeaves];
    int leaf_ndim[kMaxLeaves];
    int leaf_dim_off[kMaxLeaves];     // offset into dims[]
    int leaf_par_off[kMaxLeaves];     // offset into theta
    uint32_t term_mask[kMaxTerms];    // bit l set: leaf l is a factor of the term
    int n_dims;
    int dims[kMaxDims];               // 0-based feature indices, leaf after leaf
};

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;

// Tile sizes of the batched factorisation.
constexpr int kNBO = 128;     // panel width of the trailing update
constexpr int kNBI = 64;      // diagonal block order (potrf / trsm)
constexpr int kMB = 64;       // rows of one update tile
constexpr int kKC = 32;       // depth of one LDS stage of the update GEMM
constexpr int kThreads = 256; // work-group size of every kernel (4 wave64)
// doubles per 64x64 diagonal block: four 16x16 inverses, then one flag per
// 16x16 block (nonzero: ill-conditioned, the panel solves substitute)
constexpr int kDinvFlags = 4 * 16 * 16;
constexpr int kDinv = kDinvFlags + 4;

// Per-particle words of the chained lock-step schedule (two launches per
// panel, or one with the row solves fused; fb_potrf.hip): flags and XCD
// announcements, each in its own 32-byte sector.
constexpr int kChainStride = 64;
constexpr int kChainF0 = 0, kChainF1 = 8, kChainXccB = 16, kChainXccC = 24, kChainF2 = 32;
// (kTileAhead) rows b0 / b1 of the panel solved against the previous panel
constexpr int kChainS0 = 40, kChainS1 = 48;
// Swarms up to this size: the waiting work-groups of the update launch (two
// per particle) hold fewer than its 512 resident slots (two per CU; the
// three-per-CU variant has 768), so every producer they wait for finds a
// slot.
constexpr int kChainMaxSwarm = 160;
// INFO of a particle whose factorisation stopped at a bounded wait (1 s, or
// FB_CHAIN_WAIT_TICKS of the 100 MHz clock) instead of hanging; the host
// reports an error for it.
constexpr int kInfoWaitTimeout = -9;
// Panel tile code bit: the chained update launch with two role work-groups
// per particle instead of three.
constexpr int kTileTwoRoles = 4;
// Panel tile code bit (64 x 128 row tiles only): each row tile also solves
// its rows once the panel's diagonal chain is published -- one launch for
// the panel instead of two.
constexpr int kTileFused = 8;
// Panel tile code bit (64 x 128 row tiles, not fused): the chained update
// launch in 8-deep chunks, whose smaller register and LDS footprint holds
// three work-groups per CU instead of two (more resident row tiles for
// launches of more than 512 work-groups; the same bits).
constexpr int kTileThreePerCU = 16;
// Panel tile code bit (64 x 128 row tiles, not fused; ignored on the first
// panel and after a fused one): the update launch of the panel also solves
// the previous panel's rows -- each row tile its own rows before its update,
// the role work-groups the panel's two diagonal row blocks first -- so the
// previous panel needs no row-solve launch.  The update takes the previous
// panel's 128 columns last, after waiting for the two diagonal row blocks
// (flags kChainS0 / kChainS1), with the accumulators kept in registers: the
// same operations in the same order, so the same bits.
constexpr int kTileAhead = 32;
// Panel tile code bit (64 x 128 row tiles, not fused; alone or with
// kTileAhead): 8-deep chunks at four work-groups per CU (1024 resident
// slots instead of 512: launches of a few hundred work-groups more than a
// multiple of 512 no longer leave a nearly empty last round; the same bits).
constexpr int kTileFourPerCU = 64;

inline __host__ __device__ int ceil_div(int a, int b) { return (a + b - 1) / b; }
inline __host__ __device__ int round_up(int a, int b) { return ceil_div(a, b) * b; }

// Launchers (each defined next to its kernels).  All enqueue on `stream`.
struct SwarmBuffers {
    int n, d, np, nrhs, likelihood;
    int S;                 // particles of this launch
    int n_slots, n_leaves;
    const double *slots;   // n_slots x n
    const double *RHS;     // nrhs x n
    double *hyp;           // S x hyper record (see fb_build.hip)
    double *hypT = nullptr;   // the same records transposed: entry k of particle s at hypT[k * ldT + s],
    int ldT = 0;              // particles S .. S + 15 repeating particle S - 1 (ldT >= S + 16)
    const double *theta;   // n_par x S
    const KernelProgram *prog;   // device copy
    const KernelProgram *prog_host;   // host copy (passed by value to kernels whose loops read it)
    int gemm_kc = 32;      // chunk depth of the lock-step trailing update (16 or 32)
    int tile = 0;          // lock-step panel-update tiles: 0 64x128, 1 32x128, 2 32x64,
                           // | kTileTwoRoles: two role work-groups in the chained update launch,
                           // | kTileFused, | kTileThreePerCU (see their definitions)
    const signed char *panel_tile = nullptr;   // host: per-panel tile shape (overrides `tile`), or none
    double *K;             // S x np x np
    int *info;             // S
    double *logdet_part;   // S x (np / kNBI)
    double *dinv;          // S x (np / kNBI) x kDinv: inverses of the 16x16 diagonal blocks of L
    double *ll;            // S
    double *extra;         // S x 4: ll, logdet, mu_hat, sigma2_hat
    double *zbuf;          // S x n: z = L^{-1} rhs_0 (or the concentrated residual)
    int *chain = nullptr;  // S x kChainStride: chained schedule (two launches per panel), or none
    int tail = 0;          // chained: the last `tail` panels are pre-updated in the launch of the one before
    int lookahead = 0;     // chained: every update launch also pre-updates the next panel
    unsigned long long wait_limit = 100000000ull;   // chained: bound of a hand-over wait (100 MHz ticks)
};

void launch_build(const SwarmBuffers &b, hipStream_t st);
void launch_panel_update_factor(const SwarmBuffers &b, int c0, hipStream_t st);
void launch_trsm_factor(const SwarmBuffers &b, int c0, hipStream_t st);
void launch_update_trsm(const SwarmBuffers &b, int c1, hipStream_t st);
void launch_panel_chain(const SwarmBuffers &b, int c0, hipStream_t st);
void launch_solve_rows(const SwarmBuffers &b, int c0, hipStream_t st);
bool panel_fused(const SwarmBuffers &b, int c0);
bool panel_ahead(const SwarmBuffers &b, int c0);
bool panel_rows_launch(const SwarmBuffers &b, int c0);
void launch_likelihood(const SwarmBuffers &b, hipStream_t st);

// Several swarms (training sets of different N) in one launch of the
// chained schedule: one segment per swarm with the panel column c0 it
// factors in this launch.  The launcher fills the block-id ends.
constexpr int kMultiMaxSegments = 16;
struct ChainSegment {
    double *K;             // S x np x np
    int *info;             // S
    double *logdet_part;   // S x (np / kNBI)
    double *dinv;          // S x (np / kNBI) x kDinv
    int *chain;            // S x kChainStride, reset by the swarm's build
    int np, n, S, c0;
    int tail_set;          // the last tail_set panels are pre-updated in the launch of the one before
    int k0, kt, tail;      // pre-update schedule of this launch (set by the launcher)
    int role_end, tile_end, tail_end, solve_end;   // exclusive ends of the segment's block ids
};
struct ChainMulti {
    int G = 0;             // segments
    int tile = 0;          // panel tile code (0, 1, 2, | kTileTwoRoles, kTileFused, kTileThreePerCU)
    unsigned long long wait_limit = 100000000ull;
    ChainSegment seg[kMultiMaxSegments];
};
// The update launch (and, unless fused, the row-solve launch) of one panel
// of every segment.
void launch_chain_multi(ChainMulti &m, hipStream_t st);

// Adjoint gradient (fb_grad.hip).
struct GradWork {
    double *U = nullptr;        // S x np x np: L^{-T}
    double *NK = nullptr;       // S x np x np: -K^{-1} (lower tiles); scratch of the inverse before that
    double *X = nullptr;        // S x np x np: L^{-1} (lower blocks), built alongside U
    double *alpha = nullptr;    // S x n
    double *partial = nullptr;  // gradient_partial_count(n, n_slots, S, P)
    int P = 0;
};
int gradient_tile(int n_slots);
size_t gradient_partial_count(int n, int n_slots, int S, int P);
void launch_gradient(const SwarmBuffers &b, const GradWork &w, double *grad, hipStream_t st);

// Persistent task-graph factorisation (fb_dag.hip).
constexpr int kDagMaxQueues = 16;
constexpr int kDagMaxQueueMats = 128;   // matrices per XCD queue in one launch
constexpr int kDagCtl = (6 * kDagMaxQueues + 2) * 32;   // control words, one per 128-byte line
constexpr int kDagErrorWord = kDagCtl - 32;              // 0 ok, 1 no ready task for 2 s, 2 queue not drained
// One matrix of a persistent factorisation.
struct DagMatrix {
    double *K;               // np x np, column major
    double *dinv;            // nb x kDinv
    double *logdet_part;     // nb
    int *info;               // INFO of this matrix
    int np, n, nb;
    int dep_off;             // first counter (dag_plan)
    int queue, rank;         // XCD queue and position among its matrices (dag_plan)
};
struct DagPlan {
    bool ok = true;
    int Q = 1, n_mats = 0, n_dep = 0, n_slots = 0;
    int n_hi[kDagMaxQueues], n_lo[kDagMaxQueues], n_first[kDagMaxQueues];
    int slot_off[kDagMaxQueues], slot_cap[kDagMaxQueues];
};
struct DagWork {
    int *ctl = nullptr;                  // kDagCtl control words
    int *dep = nullptr;                  // DagPlan::n_dep counters
    unsigned long long *slots = nullptr; // DagPlan::n_slots ready-list slots
    int *dead = nullptr;                 // one per matrix
};
int dag_grid_size(int *n_xcd);
int dag_index_count(int np);
int dag_tasks_per_matrix(int np);
// Fills queue, rank, dep_off; plan.ok = false when a queue would hold more
// than kDagMaxQueueMats matrices.
DagPlan dag_plan(std::vector<DagMatrix> &mats, int n_xcd);
void launch_dag(const DagPlan &plan, const DagMatrix *d_mats, const DagWork &w, int grid, hipStream_t st);
void launch_backsolve(const SwarmBuffers &b, int particle, double *alpha, hipStream_t st);

// Particle swarm state on the device (fb_pso.hip).  Particle k's parameters
// at x[k * P + j]; state[0] = gbest cost.
struct PsoState {
    int S = 0, P = 0;
    double *x = nullptr, *v = nullptr, *pbest = nullptr, *pcost = nullptr, *gbest = nullptr;
    double *cost = nullptr, *state = nullptr;
    int *info = nullptr, *pending = nullptr, *gidx = nullptr;
};
void launch_pso_cost(const PsoState &p, const double *ll, const int *info_src, hipStream_t st);
void launch_pso_select(const PsoState &p, hipStream_t st);
void launch_pso_update(const PsoState &p, const double *r1, const double *r2, const double *coef,
                       hipStream_t st);
void launch_pso_scatter(const PsoState &p, int k, const int *idx, const double *xs, const double *vs,
                        hipStream_t st);
void launch_pso_gather(const PsoState &p, int k, const int *idx, double *out, hipStream_t st);
void launch_pso_put_cost(const PsoState &p, int k, const int *idx, const double *ll, const int *info_k,
                         hipStream_t st);

}  // namespace fb
