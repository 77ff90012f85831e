// fb_dag.hip -- the swarm's whole Cholesky factorisation in one persistent launch.
//
// Same arithmetic as the lock-step kernels (fb_tiles.h), scheduled as a task
// graph instead of one launch per step: a fixed grid of resident work-groups
// takes ready tasks from per-XCD queues, runs them, and makes their
// dependants ready.  Every matrix of the swarm advances at its own pace, so
// the latency-bound diagonal factorisations and panel solves of one matrix
// overlap the matrix-core trailing updates of the others, with no launch
// boundaries in between.
//
// Tasks of matrix s (64-row tiles i, 64-column blocks b, 128-column panels p
// made of blocks b0 = 2p, b1 = 2p + 1, nb = np / 64 tiles, P = nb / 2):
//   TG (p, i)   C(i, panel p) -= L(i, 0:128p) L(panel rows, 0:128p)^T   1 <= p, 2p <= i
//               (block b0 only for the diagonal row i = b0)
//   TP (b)      factor the diagonal block b (INFO, partial log-det, the
//               inverses of its 16x16 diagonal blocks)
//   TT (b, i)   L(i, b) = C(i, b) L(b, b)^{-T}                           b < i
//   TG2(p, i)   C(i, b1) -= L(i, b0) L(b1, b0)^T                         b1 <= i
// Data-flow dependencies (each task has at most three):
//   TG (p, i)  <- TT(b0-1, i), TT(b0-1, b0), TT(b0-1, b1)
//   TP (b0)    <- TG(p, b0) (p >= 1);          TP(b1) <- TG2(p, b1)
//   TT (b, i)  <- TP(b), and TG(p, i) (b = b0, p >= 1) or TG2(p, i) (b = b1)
//   TG2(p, i)  <- TT(b0, i), TT(b0, b1)
//
// Scheduling.  Each task of each matrix has a counter holding its number of
// unfinished predecessors.  A work-group that finishes a task decrements the
// counters of the task's dependants (one task, or one row of tasks) and pushes
// every dependant that reaches zero onto its queue's ready list -- the
// high-priority list for the chain that bounds a matrix (diagonal blocks and
// the rows of the next panel's diagonal blocks), the normal list otherwise.
// Every task is pushed exactly once and the split between the lists is
// static, so the lists are flat arrays (no wrap) whose final lengths are
// known.  A work-group claims the filled head slot of the high list with a
// compare-and-swap whenever there is one; otherwise it holds one ticket on
// the normal list (fetch-and-add, always succeeds) and runs that slot's task
// once the slot is filled.
// Device-scope atomics go to memory on this chip, so every control word
// sits on its own cache line and polls back off.  TP(0) of every matrix is
// ready at the start.
//
// XCD pinning.  gfx950 has 8 accelerator dies (XCDs), each with its own L2;
// the L2s are not coherent with each other inside a kernel.  Matrix s belongs
// to queue s mod Q, and a queue is served by the work-groups of exactly one
// XCD (the first XCD whose work-group claims it), so producer and consumer of
// every tile share one L2.  Publishing a tile then needs only its stores to
// have reached that L2 (s_waitcnt vmcnt(0); the vector L1 is write-through),
// and taking one drops the consumer CU's L1 and the L2's clean non-coherent
// lines (buffer_inv sc1; dirty lines stay), instead of writing the whole L2
// back to memory after every task, which costs 20-30 us per task on a busy
// XCD.  (buffer_inv sc0 is not enough: at work-group scope the L1 is kept and
// stale tiles are read.)  Kernel boundaries write L2 back and invalidate it
// as usual, so the build before and the likelihood after see coherent data.
//
// Termination: a work-group leaves once the high list is claimed to its final
// length and its ticket is at or beyond the final length of the normal list.  Every slot
// below the final length is filled eventually (by induction over the task
// graph a ready task always exists while tasks remain and none runs), so no
// wait is endless; waits are nevertheless bounded in wall-clock time (the
// error flag is raised instead of hanging).  A matrix whose diagonal block is
// not positive definite is marked dead: its remaining tasks skip their work
// but still release their dependants.  The last work-group out checks that
// every queue was drained (error 2 otherwise, e.g. a queue no XCD could
// claim), and the host then evaluates the swarm on the lock-step path.
#include <algorithm>
#include <vector>

#include "fb_tiles.h"

namespace fb {

enum DagTaskType : int { kTaskTG = 0, kTaskTP = 1, kTaskTT = 2, kTaskTG2 = 3 };

// Control words, one per 128-byte line (kCtlLine ints): per queue q the
// words ctl_word(q, k) for k = high head, high tail, normal head, normal
// tail, owner XCD (one spare line); then finished work-groups and error.
constexpr int kCtlLine = 32;
enum : int { kHiHead = 0, kHiTail = 1, kLoHead = 2, kLoTail = 3, kOwner = 4, kCtlPerQueue = 6 };
__host__ __device__ constexpr int ctl_word(int q, int k) { return (q * kCtlPerQueue + k) * kCtlLine; }
constexpr int kCtlFinished = ctl_word(kDagMaxQueues, 0);
constexpr int kCtlError = kCtlFinished + kCtlLine;
static_assert(kCtlError + kCtlLine == kDagCtl && kCtlError == kDagErrorWord, "control word layout");

struct DagArgs {
    double *K;
    int np, n, nb, S, Q;
    int tpm;            // tasks per matrix
    int hpm;            // of them on the high list
    int ntm;            // task index space per matrix (dag_index)
    int slot_cap;       // slots per list (tasks of the largest queue)
    int *ctl;
    int *dep;           // S x ntm unfinished-predecessor counters
    const int *dep0;    // ntm initial counters
    unsigned long long *slots;   // Q x 2 lists x slot_cap
    int *dead;          // S
    int *info;
    double *logdet_part;
    double *dinv;       // S x nb x kDinv diagonal-block inverses
};

// Dense index of a task within its matrix.
__host__ __device__ inline int dag_index(int type, int a, int i, int nb)
{
    const int P = nb / 2;
    switch (type) {
    case kTaskTG: return a * nb + i;
    case kTaskTP: return P * nb + a;
    case kTaskTT: return P * nb + nb + a * nb + i;
    default: return P * nb + nb + nb * nb + a * nb + i;
    }
}

__host__ __device__ inline int dag_index_space(int nb) { return 2 * nb * nb + nb; }

// A ready-list entry: type + 1 (never zero), matrix, a, i.
__device__ inline unsigned long long dag_encode(int type, int s, int a, int i)
{
    return (unsigned long long)(type + 1) | ((unsigned long long)s << 8) | ((unsigned long long)a << 24) |
           ((unsigned long long)i << 40);
}

__device__ inline int ld_relaxed(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ inline void st_relaxed(int *p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ inline unsigned long long ld_slot(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// This wave's global stores have reached the XCD's L2.
__device__ inline void xcd_release() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Drops this CU's vector L1 (and clean non-coherent L2 lines), so later loads
// read the XCD's L2 or memory.
__device__ inline void xcd_acquire() { asm volatile("s_waitcnt vmcnt(0)\n\tbuffer_inv sc1" ::: "memory"); }

__device__ inline int xcc_id()
{
    // hwreg(HW_REG_XCC_ID, 0, 4): id 20, offset 0, 4 bits
    return (int)(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 0xf);
}

constexpr unsigned long long kWaitTicks = 200000000ull;   // 2 s of the 100 MHz real-time clock

__host__ __device__ inline bool dag_is_high(int type, int a, int i)
{
    const int panel = type == kTaskTT ? a / 2 : a;
    return type == kTaskTP || i <= 2 * panel + 3;
}

// Sleeps between polls, doubling up to ~2 us.
__device__ inline void backoff(int &sl)
{
    if (sl <= 1) __builtin_amdgcn_s_sleep(1);
    else if (sl <= 2) __builtin_amdgcn_s_sleep(2);
    else if (sl <= 4) __builtin_amdgcn_s_sleep(4);
    else if (sl <= 8) __builtin_amdgcn_s_sleep(8);
    else if (sl <= 16) __builtin_amdgcn_s_sleep(16);
    else __builtin_amdgcn_s_sleep(32);
    sl = sl < 32 ? 2 * sl : 32;
}

// Thread 0: the next task of queue q -- the high list's head slot when it
// is filled (claimed by compare-and-swap), else the slot of this
// work-group's ticket on the normal list once it is filled (*ticket < 0: no
// ticket held yet).  Returns 0 when both lists are exhausted or on error.
__device__ inline unsigned long long dag_pop(const DagArgs &a, int q, int n_hi, int n_lo, int *ticket)
{
    int *error = a.ctl + kCtlError;
    int *hi_head = a.ctl + ctl_word(q, kHiHead);
    const unsigned long long *hi = a.slots + (size_t)q * 2 * a.slot_cap;
    const unsigned long long *lo = hi + a.slot_cap;
    unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    int sl = 1;
    for (int k = 0;; ++k) {
        const int hh = ld_relaxed(hi_head);
        if (hh < n_hi) {
            const unsigned long long v = ld_slot(hi + hh);
            if (v != 0) {
                int expected = hh;
                if (__hip_atomic_compare_exchange_strong(hi_head, &expected, hh + 1, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    return v;
                t0 = __builtin_amdgcn_s_memrealtime();   // another work-group made progress
                sl = 1;
                continue;
            }
        }
        if (*ticket < 0) *ticket = atomicAdd(a.ctl + ctl_word(q, kLoHead), 1);
        if (*ticket < n_lo) {
            const unsigned long long v = ld_slot(lo + *ticket);
            if (v != 0) {
                *ticket = -1;
                return v;
            }
        } else if (hh >= n_hi) {
            return 0;
        }
        if ((k & 15) == 15) {
            if (ld_relaxed(error)) return 0;
            if (__builtin_amdgcn_s_memrealtime() - t0 > kWaitTicks) {
                st_relaxed(error, 1);
                return 0;
            }
        }
        backoff(sl);
    }
}

// Decrements the counters of dependants (type, a, i) for i in [lo, hi) of
// matrix s and pushes those that become ready.  Work-group cooperative.
__device__ inline void dag_release(const DagArgs &a, int q, int s, int type, int da, int lo, int hi)
{
    int *dep = a.dep + (size_t)s * a.ntm;
    unsigned long long *lists = a.slots + (size_t)q * 2 * a.slot_cap;
    for (int i = lo + (int)threadIdx.x; i < hi; i += kThreads) {
        if (atomicSub(dep + dag_index(type, da, i, a.nb), 1) == 1) {
            const int l = dag_is_high(type, da, i) ? 0 : 1;
            const int slot = atomicAdd(a.ctl + ctl_word(q, l == 0 ? kHiTail : kLoTail), 1);
            __hip_atomic_store(lists + (size_t)l * a.slot_cap + slot, dag_encode(type, s, da, i), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

constexpr int kDagSmem = TileShape<kMB, kNBO>::SMEM > kPotrfSmem ? TileShape<kMB, kNBO>::SMEM : kPotrfSmem;

__global__ void __launch_bounds__(kThreads) dag_factor_kernel(DagArgs a)
{
    __shared__ __attribute__((aligned(16))) double smem[kDagSmem];
    __shared__ unsigned long long sh_task;
    __shared__ int sh_queue;
    const int tid = threadIdx.x;
    const int P = a.nb / 2;
    if (tid == 0) {
        const int x = xcc_id();
        const int q = x % a.Q;
        const int own = atomicCAS(a.ctl + ctl_word(q, kOwner), -1, x);
        sh_queue = (own == -1 || own == x) ? q : -1;
    }
    __syncthreads();
    const int q = sh_queue;
    const int nmat = q >= 0 ? (a.S - q + a.Q - 1) / a.Q : 0;
    const int n_hi = nmat * a.hpm, n_lo = nmat * (a.tpm - a.hpm);
    int ticket = -1;   // thread 0's ticket on the normal list
    while (q >= 0) {
        if (tid == 0) sh_task = dag_pop(a, q, n_hi, n_lo, &ticket);
        __syncthreads();
        const unsigned long long v = sh_task;
        if (v == 0) break;
        const int type = (int)(v & 0xff) - 1;
        const int s = (int)((v >> 8) & 0xffff);
        const int ta = (int)((v >> 24) & 0xffff);
        const int ti = (int)((v >> 40) & 0xffff);
        // every wave drops stale L1 copies of the producers' tiles; all of the
        // task's loads come after the barrier
        xcd_acquire();
        __syncthreads();
        double *A = a.K + (size_t)s * a.np * a.np;
        const bool live = ld_relaxed(a.dead + s) == 0;
        int rtype = kTaskTT, ra = 0, rlo = 0, rhi = 0;   // dependants to release
        if (type == kTaskTG) {
            const int p = ta, i = ti;
            if (live) {
                if (i == 2 * p)   // only the lower 64 x 64 block of the diagonal row is used
                    gemm_tile_update<kMB, kNBI>(A, a.np, i * kNBI, p * kNBO, 0, p * kNBO, smem);
                else
                    gemm_tile_update<kMB, kNBO>(A, a.np, i * kNBI, p * kNBO, 0, p * kNBO, smem);
            }
            if (i == 2 * p) { rtype = kTaskTP; ra = i; rlo = 0; rhi = 1; }
            else { rtype = kTaskTT; ra = 2 * p; rlo = i; rhi = i + 1; }
        } else if (type == kTaskTP) {
            const int b = ta;
            if (live) {
                double lg = 0.0;
                const int fail = potrf64_block(A, a.np, a.n, b * kNBI, &lg, smem,
                                               a.dinv + ((size_t)s * a.nb + b) * kDinv);
                if (tid == 0) {
                    if (fail) {
                        a.info[s] = fail;
                        st_relaxed(a.dead + s, 1);
                    } else {
                        a.logdet_part[(size_t)s * a.nb + b] = lg;
                    }
                }
            }
            rtype = kTaskTT; ra = b; rlo = b + 1; rhi = a.nb;
        } else if (type == kTaskTT) {
            const int b = ta, i = ti, p = b / 2;
            if (live) trsm64_mfma(A, a.np, b * kNBI, i * kNBI, a.dinv + ((size_t)s * a.nb + b) * kDinv);
            if ((b & 1) == 0) {
                rtype = kTaskTG2; ra = p;
                if (i == b + 1) { rlo = b + 1; rhi = a.nb; } else { rlo = i; rhi = i + 1; }
            } else if (p + 1 < P) {
                rtype = kTaskTG; ra = p + 1;
                if (i == b + 1 || i == b + 2) { rlo = b + 1; rhi = a.nb; } else { rlo = i; rhi = i + 1; }
            }
        } else {
            const int p = ta, i = ti, b0 = 2 * p, b1 = 2 * p + 1;
            if (live) gemm_tile_update<kMB, kNBI>(A, a.np, i * kNBI, b1 * kNBI, b0 * kNBI, b1 * kNBI, smem);
            if (i == b1) { rtype = kTaskTP; ra = b1; rlo = 0; rhi = 1; }
            else { rtype = kTaskTT; ra = b1; rlo = i; rhi = i + 1; }
        }
        // publish: every wave's stores reach L2, then the dependants are released
        xcd_release();
        __syncthreads();
        dag_release(a, q, s, rtype, ra, rlo, rhi);
    }
    // the last work-group out checks that every queue was drained
    if (tid == 0) {
        if (atomicAdd(a.ctl + kCtlFinished, 1) == (int)gridDim.x - 1) {
            for (int k = 0; k < a.Q; ++k) {
                const int m = (a.S - k + a.Q - 1) / a.Q;
                if (ld_relaxed(a.ctl + ctl_word(k, kHiHead)) < m * a.hpm ||
                    ld_relaxed(a.ctl + ctl_word(k, kLoHead)) < m * (a.tpm - a.hpm))
                    st_relaxed(a.ctl + kCtlError, 2);
            }
        }
    }
}

// Fresh counters, empty lists holding only TP(0) of every matrix, control words.
__global__ void dag_reset_kernel(DagArgs a)
{
    const size_t n_dep = (size_t)a.S * a.ntm;
    const size_t n_slot = (size_t)a.Q * 2 * a.slot_cap;
    const size_t total = n_dep + n_slot + a.S + kDagCtl;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        if (e < n_dep) {
            a.dep[e] = a.dep0[e % a.ntm];
        } else if (e < n_dep + n_slot) {
            const size_t k = e - n_dep;
            const int q = (int)(k / (2 * (size_t)a.slot_cap));
            const int l = (int)((k / a.slot_cap) % 2);
            const int idx = (int)(k % a.slot_cap);
            const int s = q + idx * a.Q;
            a.slots[k] = (l == 0 && s < a.S) ? dag_encode(kTaskTP, s, 0, 0) : 0ull;
        } else if (e < n_dep + n_slot + a.S) {
            a.dead[e - n_dep - n_slot] = 0;
        } else {
            const int w = (int)(e - n_dep - n_slot - a.S);
            const int q = w / (kCtlPerQueue * kCtlLine), k = (w / kCtlLine) % kCtlPerQueue;
            int v = 0;
            if (q < a.Q && k == kHiTail) v = (a.S - q + a.Q - 1) / a.Q;   // the TP(0)s
            if (q < kDagMaxQueues && k == kOwner) v = -1;
            a.ctl[w] = v;
        }
    }
}

int dag_high_tasks_per_matrix(int np)
{
    const int nb = np / kNBI, P = nb / 2;
    int t = 0;
    for (int p = 1; p < P; ++p)
        for (int i = 2 * p; i < nb; ++i) t += dag_is_high(kTaskTG, p, i);
    for (int b = 0; b < nb; ++b) t += dag_is_high(kTaskTP, b, 0);
    for (int b = 0; b < nb; ++b)
        for (int i = b + 1; i < nb; ++i) t += dag_is_high(kTaskTT, b, i);
    for (int p = 0; p < P; ++p)
        for (int i = 2 * p + 1; i < nb; ++i) t += dag_is_high(kTaskTG2, p, i);
    return t;
}

int dag_tasks_per_matrix(int np)
{
    const int nb = np / kNBI, P = nb / 2;
    int t = nb + nb * (nb - 1) / 2;   // TP, TT
    for (int p = 1; p < P; ++p) t += nb - 2 * p;        // TG
    for (int p = 0; p < P; ++p) t += nb - 2 * p - 1;    // TG2
    return t;
}

std::vector<int> dag_initial_counts(int np)
{
    const int nb = np / kNBI, P = nb / 2;
    std::vector<int> d(dag_index_space(nb), 0);
    for (int p = 1; p < P; ++p)
        for (int i = 2 * p; i < nb; ++i) d[dag_index(kTaskTG, p, i, nb)] = (i <= 2 * p + 1) ? 2 : 3;
    for (int b = 0; b < nb; ++b) d[dag_index(kTaskTP, b, 0, nb)] = b == 0 ? 0 : 1;
    for (int b = 0; b < nb; ++b)
        for (int i = b + 1; i < nb; ++i) d[dag_index(kTaskTT, b, i, nb)] = b == 0 ? 1 : 2;
    for (int p = 0; p < P; ++p)
        for (int i = 2 * p + 1; i < nb; ++i) d[dag_index(kTaskTG2, p, i, nb)] = (i == 2 * p + 1) ? 1 : 2;
    return d;
}

int dag_index_count(int np) { return dag_index_space(np / kNBI); }

int dag_slot_capacity(int S, int np, int Q) { return ((S + Q - 1) / Q) * dag_tasks_per_matrix(np); }

int dag_grid_size(int *n_xcd)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dag_factor_kernel, kThreads, 0) != hipSuccess)
        return 0;
    // gfx950: 32 CUs per XCD (8 XCDs on a 256-CU device)
    *n_xcd = (cus % 32 == 0) ? std::max(1, std::min(cus / 32, kDagMaxQueues)) : 1;
    return cus * std::max(per_cu, 1);
}

int dag_queues(int S, int n_xcd) { return std::max(1, std::min(std::min(n_xcd, kDagMaxQueues), S)); }

void launch_dag(const SwarmBuffers &b, const DagWork &w, int n_queues, int grid, hipStream_t st)
{
    DagArgs a;
    a.K = b.K; a.np = b.np; a.n = b.n; a.nb = b.np / kNBI; a.S = b.S;
    a.Q = dag_queues(b.S, n_queues);
    a.tpm = dag_tasks_per_matrix(b.np);
    a.hpm = dag_high_tasks_per_matrix(b.np);
    a.ntm = dag_index_space(a.nb);
    a.slot_cap = dag_slot_capacity(b.S, b.np, a.Q);
    a.ctl = w.ctl; a.dep = w.dep; a.dep0 = w.dep0; a.slots = w.slots; a.dead = w.dead;
    a.info = b.info; a.logdet_part = b.logdet_part; a.dinv = b.dinv;
    const size_t total = (size_t)a.S * a.ntm + (size_t)a.Q * 2 * a.slot_cap + a.S + kDagCtl;
    const int rgrid = (int)std::min<size_t>(1024, (total + 255) / 256);
    hipLaunchKernelGGL(dag_reset_kernel, dim3(rgrid), dim3(256), 0, st, a);
    hipLaunchKernelGGL(dag_factor_kernel, dim3(std::max(1, grid)), dim3(kThreads), 0, st, a);
}

}  // namespace fb
