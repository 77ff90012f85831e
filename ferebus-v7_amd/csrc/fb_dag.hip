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
// The matrices are described one by one (DagMatrix: storage, order, INFO,
// log-det and inverse slots), so one launch factors the swarms of several
// training sets of different N together -- the batched uneven case -- as
// well as the equal-size swarm of one training set.
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
// the L2s are not coherent with each other inside a kernel.  Every matrix
// belongs to one queue (chosen on the host: round robin for equal sizes,
// longest-processing-time on N^3 for uneven ones), and a queue is served by
// the work-groups of exactly one XCD (the first XCD whose work-group claims
// it), so producer and consumer of every tile share one L2.  Publishing a tile then needs only its stores to
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
#include <cmath>
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
    const DagMatrix *mats;   // device copy of the matrix descriptors
    int n_mats, Q;
    int n_hi[kDagMaxQueues], n_lo[kDagMaxQueues];   // final list lengths per queue
    int n_first[kDagMaxQueues];                     // matrices per queue (TP(0) ready at the start)
    int slot_off[kDagMaxQueues];                    // the queue's high list; normal list after it
    int slot_cap[kDagMaxQueues];                    // slots per list of the queue
    int n_dep, n_slots;
    int *ctl;
    int *dep;                  // unfinished-predecessor counters, mats[s].dep_off onwards
    unsigned long long *slots;
    int *dead;                 // n_mats
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

// Number of predecessors of the task with dense index idx (0 for indices
// that name no task).
__host__ __device__ inline int dag_initial_count(int idx, int nb)
{
    const int P = nb / 2;
    if (idx < P * nb) {                                   // TG(a, i)
        const int a = idx / nb, i = idx % nb;
        return (a >= 1 && i >= 2 * a) ? (i <= 2 * a + 1 ? 2 : 3) : 0;
    }
    idx -= P * nb;
    if (idx < nb) return idx == 0 ? 0 : 1;               // TP(b)
    idx -= nb;
    if (idx < nb * nb) {                                  // TT(b, i)
        const int b = idx / nb, i = idx % nb;
        return i > b ? (b == 0 ? 1 : 2) : 0;
    }
    idx -= nb * nb;                                       // TG2(a, i)
    const int a = idx / nb, i = idx % nb;
    return i >= 2 * a + 1 ? (i == 2 * a + 1 ? 1 : 2) : 0;
}

// A ready-list entry: type + 1 (never zero), the matrix's rank in its queue, a, i.
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

// Values every lane holds alike, made visibly uniform to the compiler so
// that strides and addresses derived from them stay in scalar registers.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <typename T>
__device__ inline T *uni(T *p)
{
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return (T *)(((unsigned long long)hi << 32) | lo);
}

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
    const unsigned long long *hi = a.slots + a.slot_off[q];
    const unsigned long long *lo = hi + a.slot_cap[q];
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
__device__ inline void dag_release(const DagArgs &a, int q, int rank, int nb, int dep_off, int type, int da,
                                   int lo, int hi)
{
    int *dep = a.dep + dep_off;
    unsigned long long *lists = a.slots + a.slot_off[q];
    for (int i = lo + (int)threadIdx.x; i < hi; i += kThreads) {
        if (atomicSub(dep + dag_index(type, da, i, nb), 1) == 1) {
            const int l = dag_is_high(type, da, i) ? 0 : 1;
            const int slot = atomicAdd(a.ctl + ctl_word(q, l == 0 ? kHiTail : kLoTail), 1);
            __hip_atomic_store(lists + (size_t)l * a.slot_cap[q] + slot, dag_encode(type, rank, da, i),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

constexpr int kDagKC = 16;
constexpr int kDagSmem = TileShape<kMB, kNBO, kDagKC>::SMEM > kPotrfSmem ? TileShape<kMB, kNBO, kDagKC>::SMEM : kPotrfSmem;

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2))) dag_factor_kernel(DagArgs a)
{
    __shared__ __attribute__((aligned(16))) double smem[kDagSmem];
    __shared__ unsigned long long sh_task;
    __shared__ int sh_queue;
    // the descriptors of this queue's matrices, by rank (read once: a global
    // read per task would come after the acquire's cache invalidation)
    __shared__ DagMatrix sh_mats[kDagMaxQueueMats];
    __shared__ int sh_index[kDagMaxQueueMats];
    const int tid = threadIdx.x;
    if (tid == 0) {
        const int x = xcc_id();
        const int q = x % a.Q;
        const int own = atomicCAS(a.ctl + ctl_word(q, kOwner), -1, x);
        sh_queue = (own == -1 || own == x) ? q : -1;
    }
    __syncthreads();
    const int q = sh_queue;
    if (q >= 0)
        for (int m = tid; m < a.n_mats; m += kThreads) {
            const DagMatrix d = a.mats[m];
            if (d.queue == q) {
                sh_mats[d.rank] = d;
                sh_index[d.rank] = m;
            }
        }
    __syncthreads();
    const int n_hi = q >= 0 ? a.n_hi[q] : 0, n_lo = q >= 0 ? a.n_lo[q] : 0;
    int ticket = -1;   // thread 0's ticket on the normal list
    while (q >= 0) {
        if (tid == 0) sh_task = dag_pop(a, q, n_hi, n_lo, &ticket);
        __syncthreads();
        const unsigned long long v = sh_task;
        if (v == 0) break;
        const int type = uni((int)(v & 0xff) - 1);
        const int rank = uni((int)((v >> 8) & 0xffff));
        const int ta = uni((int)((v >> 24) & 0xffff));
        const int ti = uni((int)((v >> 40) & 0xffff));
        // every wave drops stale L1 copies of the producers' tiles; all of the
        // task's loads come after the barrier
        xcd_acquire();
        __syncthreads();
        const DagMatrix &Ml = sh_mats[rank];
        const int s = uni(sh_index[rank]);
        double *A = uni(Ml.K);
        double *Mdinv = uni(Ml.dinv);
        const int np = uni(Ml.np), nb = uni(Ml.nb), Mn = uni(Ml.n), dep_off = uni(Ml.dep_off), P = nb / 2;
        const bool live = ld_relaxed(a.dead + s) == 0;
        int rtype = kTaskTT, ra = 0, rlo = 0, rhi = 0;   // dependants to release
        if (type == kTaskTG) {
            const int p = ta, i = ti;
            if (live) {
                if (i == 2 * p)   // only the lower 64 x 64 block of the diagonal row is used
                    gemm_tile_update<kMB, kNBI, kDagKC, 1>(A, np, i * kNBI, p * kNBO, 0, p * kNBO, smem);
                else
                    gemm_tile_update<kMB, kNBO, kDagKC, 1>(A, np, i * kNBI, p * kNBO, 0, p * kNBO, smem);
            }
            if (i == 2 * p) { rtype = kTaskTP; ra = i; rlo = 0; rhi = 1; }
            else { rtype = kTaskTT; ra = 2 * p; rlo = i; rhi = i + 1; }
        } else if (type == kTaskTP) {
            const int b = ta;
            if (live) {
                double lg = 0.0;
                const int fail = potrf64_block(A, np, Mn, b * kNBI, &lg, smem, Mdinv + (size_t)b * kDinv);
                if (tid == 0) {
                    if (fail) {
                        *Ml.info = fail;
                        st_relaxed(a.dead + s, 1);
                    } else {
                        Ml.logdet_part[b] = lg;
                    }
                }
            }
            rtype = kTaskTT; ra = b; rlo = b + 1; rhi = nb;
        } else if (type == kTaskTT) {
            const int b = ta, i = ti, p = b / 2;
            if (live) trsm64_mfma(A, np, b * kNBI, i * kNBI, Mdinv + (size_t)b * kDinv);
            if ((b & 1) == 0) {
                rtype = kTaskTG2; ra = p;
                if (i == b + 1) { rlo = b + 1; rhi = nb; } else { rlo = i; rhi = i + 1; }
            } else if (p + 1 < P) {
                rtype = kTaskTG; ra = p + 1;
                if (i == b + 1 || i == b + 2) { rlo = b + 1; rhi = nb; } else { rlo = i; rhi = i + 1; }
            }
        } else {
            const int p = ta, i = ti, b0 = 2 * p, b1 = 2 * p + 1;
            if (live) gemm_tile_update<kMB, kNBI, kDagKC, 1>(A, np, i * kNBI, b1 * kNBI, b0 * kNBI, b1 * kNBI, smem);
            if (i == b1) { rtype = kTaskTP; ra = b1; rlo = 0; rhi = 1; }
            else { rtype = kTaskTT; ra = b1; rlo = i; rhi = i + 1; }
        }
        // publish: every wave's stores reach L2, then the dependants are released
        xcd_release();
        __syncthreads();
        dag_release(a, q, rank, nb, dep_off, rtype, ra, rlo, rhi);
    }
    // the last work-group out checks that every queue was drained
    if (tid == 0) {
        if (atomicAdd(a.ctl + kCtlFinished, 1) == (int)gridDim.x - 1) {
            for (int k = 0; k < a.Q; ++k)
                if (ld_relaxed(a.ctl + ctl_word(k, kHiHead)) < a.n_hi[k] ||
                    ld_relaxed(a.ctl + ctl_word(k, kLoHead)) < a.n_lo[k])
                    st_relaxed(a.ctl + kCtlError, 2);
        }
    }
}

// Fresh counters (matrix blockIdx.y), empty lists holding only TP(0) of
// every matrix, control words (blockIdx.y = n_mats).
__global__ void dag_reset_kernel(DagArgs a)
{
    const int y = blockIdx.y;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t e0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (y < a.n_mats) {
        const DagMatrix M = a.mats[y];
        const int ntm = dag_index_space(M.nb);
        for (size_t e = e0; e < (size_t)ntm; e += stride) a.dep[M.dep_off + e] = dag_initial_count((int)e, M.nb);
        if (e0 == 0) {
            a.dead[y] = 0;
            a.slots[a.slot_off[M.queue] + M.rank] = dag_encode(kTaskTP, M.rank, 0, 0);
        }
        return;
    }
    for (size_t e = e0; e < (size_t)a.n_slots; e += stride) {
        int q = 0;
        while (q + 1 < a.Q && (size_t)a.slot_off[q + 1] <= e) ++q;
        if (e - a.slot_off[q] >= (size_t)a.n_first[q]) a.slots[e] = 0ull;   // the TP(0) slots are the matrices'
    }
    for (size_t w = e0; w < (size_t)kDagCtl; w += stride) {
        const int q = (int)(w / (kCtlPerQueue * kCtlLine)), k = (int)((w / kCtlLine) % kCtlPerQueue);
        int v = 0;
        if (q < a.Q && k == kHiTail) v = a.n_first[q];
        if (q < kDagMaxQueues && k == kOwner) v = -1;
        a.ctl[w] = v;
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

int dag_index_count(int np) { return dag_index_space(np / kNBI); }

DagPlan dag_plan(std::vector<DagMatrix> &mats, int n_xcd)
{
    DagPlan pl;
    const int n = (int)mats.size();
    pl.Q = std::max(1, std::min(std::min(n_xcd, kDagMaxQueues), n));
    bool uniform = true;
    for (const DagMatrix &m : mats) uniform = uniform && m.np == mats[0].np;
    std::vector<double> load(pl.Q, 0.0);
    if (uniform) {
        for (int s = 0; s < n; ++s) mats[s].queue = s % pl.Q;
    } else {
        // longest processing time first on the factorisation work ~ N^3
        std::vector<int> order(n);
        for (int s = 0; s < n; ++s) order[s] = s;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return mats[x].n > mats[y].n; });
        for (int s : order) {
            const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            mats[s].queue = q;
            load[q] += std::pow((double)mats[s].n, 3.0);
        }
    }
    int dep = 0;
    std::vector<int> tasks(pl.Q, 0);
    for (int q = 0; q < pl.Q; ++q) pl.n_hi[q] = pl.n_lo[q] = pl.n_first[q] = 0;
    for (DagMatrix &m : mats) {
        m.dep_off = dep;
        dep += dag_index_space(m.nb);
        m.rank = pl.n_first[m.queue]++;
        const int t = dag_tasks_per_matrix(m.np), h = dag_high_tasks_per_matrix(m.np);
        pl.n_hi[m.queue] += h;
        pl.n_lo[m.queue] += t - h;
        tasks[m.queue] += t;
    }
    int off = 0;
    for (int q = 0; q < kDagMaxQueues; ++q) {
        pl.slot_cap[q] = q < pl.Q ? tasks[q] : 0;
        pl.slot_off[q] = off;
        off += 2 * pl.slot_cap[q];
        if (q >= pl.Q) pl.n_hi[q] = pl.n_lo[q] = pl.n_first[q] = 0;
    }
    pl.n_dep = dep;
    pl.n_slots = off;
    pl.n_mats = n;
    for (int q = 0; q < pl.Q; ++q) pl.ok = pl.ok && pl.n_first[q] <= kDagMaxQueueMats;
    return pl;
}

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

void launch_dag(const DagPlan &pl, const DagMatrix *d_mats, const DagWork &w, int grid, hipStream_t st)
{
    DagArgs a;
    a.mats = d_mats;
    a.n_mats = pl.n_mats;
    a.Q = pl.Q;
    for (int q = 0; q < kDagMaxQueues; ++q) {
        a.n_hi[q] = pl.n_hi[q]; a.n_lo[q] = pl.n_lo[q]; a.n_first[q] = pl.n_first[q];
        a.slot_off[q] = pl.slot_off[q]; a.slot_cap[q] = pl.slot_cap[q];
    }
    a.n_dep = pl.n_dep; a.n_slots = pl.n_slots;
    a.ctl = w.ctl; a.dep = w.dep; a.slots = w.slots; a.dead = w.dead;
    const int rgrid = std::max(1, std::min(64, (std::max(pl.n_slots, kDagCtl) + 255) / 256));
    hipLaunchKernelGGL(dag_reset_kernel, dim3(rgrid, pl.n_mats + 1), dim3(256), 0, st, a);
    hipLaunchKernelGGL(dag_factor_kernel, dim3(std::max(1, grid)), dim3(kThreads), 0, st, a);
}

}  // namespace fb
