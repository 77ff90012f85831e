// fb_dag.hip -- the swarm's whole Cholesky factorisation in one persistent launch.
//
// Same arithmetic as the lock-step kernels (fb_tiles.h), scheduled as a task
// graph instead of one launch per step: a fixed grid of resident work-groups
// pulls tasks from per-XCD queues in topological order and waits, per task,
// on the progress counters of the one matrix the task belongs to.  Every
// matrix of the swarm advances at its own pace, so the latency-bound diagonal
// factorisations and panel solves of one matrix overlap the matrix-core
// trailing updates of the others, with no launch boundaries in between.
//
// Tasks of matrix s (64-row tiles i, 64-column blocks b, 128-column panels p
// made of blocks b0 = 2p, b1 = 2p + 1):
//   TG (p, i)   C(i, panel p) -= L(i, 0:128p) L(panel rows, 0:128p)^T    p >= 1
//   TP (b)      factor the diagonal block b (INFO, partial log-det)
//   TT (b, i)   L(i, b) = C(i, b) L(b, b)^{-T}                            i > b
//   TG2(p, i)   C(i, b1) -= L(i, b0) L(b1, b0)^T                          i >= b1
// Progress counters of matrix s: rowfin[i] = blocks of row tile i final,
// gdone[i] / g2done[i] = last panel whose TG / TG2 of row tile i finished,
// dead = a diagonal block was not positive definite (dependants skip).
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
// Deadlock freedom: a task only waits on tasks of its own matrix that precede
// it in its queue, and a task is taken only by a running work-group, so by
// induction over the queue every taken task finishes.  Waits are bounded in
// wall-clock time; a wait that exceeds the bound raises the error flag
// instead of hanging.  The last work-group to leave checks that every queue
// was drained (error 2 otherwise, e.g. a queue no XCD could claim), and the
// host then evaluates the swarm on the lock-step path.
#include <algorithm>
#include <vector>

#include "fb_tiles.h"

namespace fb {

enum DagTaskType : short { kTaskTG = 0, kTaskTP = 1, kTaskTT = 2, kTaskTG2 = 3 };

struct DagArgs {
    double *K;
    int np, n, nb, S;
    const DagTask *tasks;
    int n_queues;
    int qoff[kDagMaxQueues + 1];   // queue q = tasks[qoff[q], qoff[q + 1])
    int *ctl;           // heads[kDagMaxQueues], owners[kDagMaxQueues], finished work-groups, error
    int *cnt;           // S x (3 nb + 1) progress counters
    int *info;
    double *logdet_part;
    double *dinv;       // S x nb x kDinv diagonal-block inverses
};

constexpr unsigned long long kWaitTicks = 200000000ull;   // 2 s of the 100 MHz real-time clock

// Counters are read and written coherently (device-scope atomics); the tile
// data they guard stays in the XCD's L2 (see the XCD pinning note above).
__device__ inline int ld_relaxed(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void st_relaxed(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// Waits until c1 >= t1 and c2 >= t2 and c3 >= t3 (null pointers are
// satisfied).  Returns false when the matrix is dead or the wait timed out.
__device__ inline bool dag_wait(const DagArgs &a, const int *dead, const int *c1, int t1, const int *c2,
                                int t2, const int *c3, int t3, int *flag)
{
    int *error = a.ctl + 2 * kDagMaxQueues + 1;
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        for (;;) {
            if (ld_relaxed(dead) || ld_relaxed(error)) {
                ok = 0;
                break;
            }
            if ((!c1 || ld_relaxed(c1) >= t1) && (!c2 || ld_relaxed(c2) >= t2) && (!c3 || ld_relaxed(c3) >= t3))
                break;
            if (__builtin_amdgcn_s_memrealtime() - t0 > kWaitTicks) {
                // abort every other wait quickly: the launch reports failure
                st_relaxed(error, 1);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        *flag = ok;
    }
    // every wave drops stale L1 copies of the producers' tiles before the
    // barrier; all of the task's loads come after it
    xcd_acquire();
    __syncthreads();
    return *flag != 0;
}

// Publishes this work-group's results: every wave's stores reach L2, then one
// thread advances the counter.
__device__ inline void dag_signal(int *c, int v)
{
    xcd_release();
    __syncthreads();
    if (threadIdx.x == 0) st_relaxed(c, v);
}

constexpr int kDagSmem = TileShape<kMB, kNBO>::SMEM > kPotrfSmem ? TileShape<kMB, kNBO>::SMEM : kPotrfSmem;

__global__ void __launch_bounds__(kThreads) dag_factor_kernel(DagArgs a)
{
    __shared__ __attribute__((aligned(16))) double smem[kDagSmem];
    __shared__ int sh_task, sh_flag, sh_queue;
    const int tid = threadIdx.x;
    const int stride = 3 * a.nb + 1;
    int *heads = a.ctl, *owners = a.ctl + kDagMaxQueues;
    if (tid == 0) {
        const int x = xcc_id();
        const int q = x % a.n_queues;
        const int own = atomicCAS(owners + q, -1, x);
        sh_queue = (own == -1 || own == x) ? q : -1;
    }
    __syncthreads();
    const int q = sh_queue;
    const int q_tasks = q >= 0 ? a.qoff[q + 1] - a.qoff[q] : 0;
    for (;;) {
        if (tid == 0) sh_task = q >= 0 ? atomicAdd(heads + q, 1) : q_tasks;
        __syncthreads();
        const int t = sh_task;
        __syncthreads();
        if (t >= q_tasks) break;
        const DagTask task = a.tasks[a.qoff[q] + t];
        const int s = task.s;
        double *A = a.K + (size_t)s * a.np * a.np;
        int *rowfin = a.cnt + (size_t)s * stride;
        int *gdone = rowfin + a.nb;
        int *g2done = gdone + a.nb;
        int *dead = g2done + a.nb;
        if (task.type == kTaskTG) {
            const int p = task.a, i = task.i, b0 = 2 * p, b1 = 2 * p + 1;
            if (!dag_wait(a, dead, rowfin + i, b0, rowfin + b0, b0, rowfin + b1, b0, &sh_flag)) continue;
            gemm_tile_update<kMB, kNBO>(A, a.np, i * kNBI, p * kNBO, 0, p * kNBO, smem);
            dag_signal(gdone + i, p + 1);
        } else if (task.type == kTaskTP) {
            const int b = task.a, p = b / 2;
            const int *dep = (b & 1) ? g2done + b : (p > 0 ? gdone + b : nullptr);
            if (!dag_wait(a, dead, dep, p + 1, nullptr, 0, nullptr, 0, &sh_flag)) continue;
            double lg = 0.0;
            const int fail = potrf64_block(A, a.np, a.n, b * kNBI, &lg, smem,
                                           a.dinv + ((size_t)s * a.nb + b) * kDinv);
            if (fail) {
                if (tid == 0) {
                    a.info[s] = fail;
                    st_relaxed(dead, 1);
                }
                __syncthreads();
                continue;
            }
            if (tid == 0) a.logdet_part[(size_t)s * a.nb + b] = lg;
            dag_signal(rowfin + b, b + 1);
        } else if (task.type == kTaskTT) {
            const int b = task.a, i = task.i, p = b / 2;
            const int *dep = (b & 1) ? g2done + i : (p > 0 ? gdone + i : nullptr);
            if (!dag_wait(a, dead, rowfin + b, b + 1, dep, p + 1, nullptr, 0, &sh_flag)) continue;
            trsm64_mfma(A, a.np, b * kNBI, i * kNBI, a.dinv + ((size_t)s * a.nb + b) * kDinv);
            dag_signal(rowfin + i, b + 1);
        } else {
            const int p = task.a, i = task.i, b0 = 2 * p, b1 = 2 * p + 1;
            if (!dag_wait(a, dead, rowfin + i, b0 + 1, rowfin + b1, b0 + 1, nullptr, 0, &sh_flag)) continue;
            gemm_tile_update<kMB, kNBI>(A, a.np, i * kNBI, b1 * kNBI, b0 * kNBI, b1 * kNBI, smem);
            dag_signal(g2done + i, p + 1);
        }
    }
    // the last work-group out checks that every queue was drained
    if (tid == 0) {
        int *finished = a.ctl + 2 * kDagMaxQueues;
        if (atomicAdd(finished, 1) == (int)gridDim.x - 1) {
            for (int k = 0; k < a.n_queues; ++k)
                if (ld_relaxed(heads + k) < a.qoff[k + 1] - a.qoff[k]) st_relaxed(a.ctl + 2 * kDagMaxQueues + 1, 2);
        }
    }
}

__global__ void dag_reset_kernel(int *cnt, int n, int *ctl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = 0;
    if (i < kDagCtl) ctl[i] = (i >= kDagMaxQueues && i < 2 * kDagMaxQueues) ? -1 : 0;
}

// Task table of queue q (matrices s = q, q + Q, ...) of np x np in
// topological order.  The chain that limits one matrix (diagonal tile update
// -> factor -> next diagonal row) is listed first in every panel, the bulk of
// the panel after it.
static void dag_queue_tasks(int S, int np, int q, int Q, std::vector<DagTask> &t)
{
    const int nb = np / kNBI;
    auto add = [&](short type, int a, int i) {
        for (int s = q; s < S; s += Q) t.push_back(DagTask{type, (short)s, (short)a, (short)i});
    };
    for (int p = 0; p < nb / 2; ++p) {
        const int b0 = 2 * p, b1 = 2 * p + 1;
        if (p > 0) add(kTaskTG, p, b0);
        add(kTaskTP, b0, 0);
        if (p > 0) add(kTaskTG, p, b1);
        add(kTaskTT, b0, b1);
        add(kTaskTG2, p, b1);
        add(kTaskTP, b1, 0);
        for (int i = b1 + 1; i < nb; ++i) {
            if (p > 0) add(kTaskTG, p, i);
        }
        for (int i = b1 + 1; i < nb; ++i) add(kTaskTT, b0, i);
        for (int i = b1 + 1; i < nb; ++i) add(kTaskTG2, p, i);
        for (int i = b1 + 1; i < nb; ++i) add(kTaskTT, b1, i);
    }
}

DagSchedule dag_schedule(int S, int np, int n_queues)
{
    DagSchedule d;
    d.n_queues = std::max(1, std::min(std::min(n_queues, kDagMaxQueues), S));
    d.qoff.push_back(0);
    for (int q = 0; q < d.n_queues; ++q) {
        dag_queue_tasks(S, np, q, d.n_queues, d.tasks);
        d.qoff.push_back((int)d.tasks.size());
    }
    return d;
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

void launch_dag(const SwarmBuffers &b, const DagTask *tasks, const std::vector<int> &qoff, int *cnt, int *ctl,
                int grid, hipStream_t st)
{
    const int nb = b.np / kNBI;
    const int ncnt = b.S * (3 * nb + 1);
    hipLaunchKernelGGL(dag_reset_kernel, dim3(ceil_div(std::max(ncnt, kDagCtl), 256)), dim3(256), 0, st, cnt, ncnt,
                       ctl);
    DagArgs a;
    a.K = b.K; a.np = b.np; a.n = b.n; a.nb = nb; a.S = b.S;
    a.tasks = tasks;
    a.n_queues = (int)qoff.size() - 1;
    for (int q = 0; q <= kDagMaxQueues; ++q) a.qoff[q] = q < (int)qoff.size() ? qoff[q] : qoff.back();
    a.ctl = ctl; a.cnt = cnt;
    a.info = b.info; a.logdet_part = b.logdet_part; a.dinv = b.dinv;
    hipLaunchKernelGGL(dag_factor_kernel, dim3(std::max(1, grid)), dim3(kThreads), 0, st, a);
}

}  // namespace fb
