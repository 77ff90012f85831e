// fb_api.hip -- C ABI of the swarm log-likelihood library (include/ferebus_hip.h).
//
// Host side of the drop-in boundary: validates the kernel specification,
// expands it into the device KernelProgram, owns every device buffer of a
// training set, and enqueues the swarm pipeline
//   hyper + build -> factorisation (per 128-wide panel: the chained update
//   and row-solve launches, or the three-launch schedule, or one persistent
//   launch for the whole swarm) -> likelihood
// on one HIP stream; the autotuner (choose_path) picks the factorisation per
// swarm size.  The batched path replays a captured hipGraph of that pipeline;
// the plain path launches it eagerly, one particle at a time.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "fb_common.h"
#include "../../include/ferebus_hip.h"

using fb::KernelProgram;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define FB_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(FB_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

// Expands the postfix program of the kernel expression into a sum of
// products of leaves (the terms), and assigns feature slots.
int make_program(const fb_kernel_spec *sp, int d, int optimise_noise, double noise,
                 KernelProgram &prog, std::vector<int> &slot_feature, std::vector<int> &slot_kind)
{
    if (!sp || sp->n_nodes < 1 || sp->n_nodes > 64 || !sp->node_type)
        return fail(FB_E_SPEC, "kernel spec: empty or too long program");
    std::memset(&prog, 0, sizeof(prog));
    std::vector<std::vector<uint32_t>> stack;
    std::map<std::pair<int, int>, int> slot_of;  // (feature, kind) -> slot
    int par = 0, doff = 0, ndims = 0;
    for (int k = 0; k < sp->n_nodes; ++k) {
        const int t = sp->node_type[k];
        const int nd = sp->node_ndim ? sp->node_ndim[k] : 0;
        if (t == FB_OP_SUM || t == FB_OP_PROD) {
            if (stack.size() < 2) return fail(FB_E_SPEC, "kernel spec: operator without two operands");
            std::vector<uint32_t> b = stack.back();
            stack.pop_back();
            std::vector<uint32_t> a = stack.back();
            stack.pop_back();
            std::vector<uint32_t> r;
            if (t == FB_OP_SUM) {
                r = a;
                r.insert(r.end(), b.begin(), b.end());
            } else {
                for (uint32_t x : a)
                    for (uint32_t y : b) r.push_back(x | y);
            }
            if ((int)r.size() > fb::kMaxTerms)
                return fail(FB_E_SPEC, "kernel spec: more than 16 product terms");
            stack.push_back(r);
            continue;
        }
        const int l = prog.n_leaves;
        if (l >= fb::kMaxLeaves) return fail(FB_E_SPEC, "kernel spec: more than 8 leaves");
        int npar = 0;
        switch (t) {
        case FB_LEAF_RBF: case FB_LEAF_PERIODIC: case FB_LEAF_LINEAR:
            if (nd < 1) return fail(FB_E_SPEC, "kernel spec: leaf without active dimensions");
            npar = nd;
            break;
        case FB_LEAF_RBF_CYCLIC:
            if (nd != d) return fail(FB_E_SPEC, "kernel spec: rbf-cyclic must span every feature");
            npar = nd;
            break;
        case FB_LEAF_CONSTANT:
            if (nd != 0) return fail(FB_E_SPEC, "kernel spec: constant leaf takes no dimensions");
            npar = 1;
            break;
        default:
            return fail(FB_E_SPEC, "kernel spec: unknown node type");
        }
        prog.leaf_kind[l] = t;
        prog.leaf_ndim[l] = (t == FB_LEAF_CONSTANT) ? 0 : nd;
        prog.leaf_par_off[l] = par;
        prog.leaf_dim_off[l] = ndims;
        for (int q = 0; q < prog.leaf_ndim[l]; ++q) {
            if (!sp->active_dims) return fail(FB_E_SPEC, "kernel spec: missing active_dims");
            const int f = sp->active_dims[doff + q];
            if (f < 1 || f > d) return fail(FB_E_SPEC, "kernel spec: active dimension out of range");
            if (t == FB_LEAF_RBF_CYCLIC && f != q + 1)
                return fail(FB_E_SPEC, "kernel spec: rbf-cyclic dimensions must be 1..d in order");
            if (ndims >= fb::kMaxDims) return fail(FB_E_SPEC, "kernel spec: too many dimensions");
            int code = 0;
            if (t == FB_LEAF_RBF || t == FB_LEAF_RBF_CYCLIC || t == FB_LEAF_PERIODIC) {
                const int skind = (t == FB_LEAF_PERIODIC) ? 1 : 0;
                auto key = std::make_pair(f - 1, skind);
                auto it = slot_of.find(key);
                int slot;
                if (it == slot_of.end()) {
                    slot = (int)slot_feature.size();
                    slot_of[key] = slot;
                    slot_feature.push_back(f - 1);
                    slot_kind.push_back(skind == 1 ? 1 : 0);
                    if (skind == 1) {  // cos(x/2) right after sin(x/2)
                        slot_feature.push_back(f - 1);
                        slot_kind.push_back(2);
                    }
                } else {
                    slot = it->second;
                }
                // distance_matrix.f90:119-120 wraps features 3, 6, 9, ...
                const int wrap = (t == FB_LEAF_RBF_CYCLIC && f % 3 == 0) ? 1 : 0;
                code = slot | (wrap << 16);
            }
            prog.dims[ndims++] = code;
        }
        doff += (t == FB_LEAF_CONSTANT) ? 0 : nd;
        par += npar;
        prog.n_leaves++;
        stack.push_back(std::vector<uint32_t>{1u << l});
    }
    if (stack.size() != 1) return fail(FB_E_SPEC, "kernel spec: program does not reduce to one kernel");
    const std::vector<uint32_t> &terms = stack.back();
    prog.n_terms = (int)terms.size();
    for (int t = 0; t < prog.n_terms; ++t) prog.term_mask[t] = terms[t];
    prog.n_dims = ndims;
    prog.n_par_kernel = par;
    prog.optimise_noise = optimise_noise ? 1 : 0;
    prog.n_par = par + prog.optimise_noise;
    prog.noise = noise;
    return FB_OK;
}

struct EventPair {
    int kind;
    hipEvent_t a, b;
};

}  // namespace

// factorisation paths the autotuner chooses from: 0 persistent single launch;
// lock-step launches with 1: 32-deep GEMM chunks, 64x128 update tiles,
// 2: 16-deep, 64x128, 3: 16-deep, 32x128, 4: 16-deep, 32x64 (kNumModes, each
// can be pinned); 5: each panel's fastest of those three tile shapes and,
// chained, of their role, fused and three-per-CU variants (measured per
// panel: reachable by tuning or FB_PANEL_TILES)
constexpr int kNumModes = 5;
// zeros after the slot values: np - n < 130 columns past the last slot row
constexpr size_t kSlotPad = 256;
constexpr int kModePerPanel = 5;

struct fb_handle {
    int n = 0, d = 0, np = 0, nrhs = 0, nsteps = 0, likelihood = 0, max_swarm = 0;
    int n_slots = 0;
    KernelProgram prog;
    hipStream_t own_stream = nullptr, stream = nullptr;
    // extra streams: the swarm is split into groups whose block pipelines run
    // concurrently, so one group's latency-bound potrf / trsm steps overlap the
    // other groups' trailing updates
    static constexpr int kMaxGroups = 8;
    static constexpr int kTuneGroups = 4;   // the autotuner tries 1 .. kTuneGroups stream groups
    int n_groups = 1;                // streams created for the groups
    int forced_groups = -1;          // FB_STREAMS pins the group count
    std::map<int, int> groups;       // tuned stream-group count per swarm size
    bool use_graph = true;
    bool use_dag = true;             // persistent factorisation for the batched path
    int gemm_kc = 32;                // lock-step trailing-update chunk depth
    bool chain = true;               // lock-step with 16-deep chunks: two launches per panel (FB_CHAIN=0: three)
    int tail = 3;                    // chained: trailing panels pre-updated by the launch before them (FB_TAIL)
    int lookahead = 0;               // chained: each update launch also pre-updates the next panel (FB_LOOKAHEAD)
    unsigned long long wait_limit = 100000000ull;   // chained: hand-over wait bound (FB_CHAIN_WAIT_TICKS)
    int *d_chain = nullptr;          // per-particle words of the two-launch schedule
    int tile = 0;                    // lock-step panel-update tile shape (16-deep chunks)
    const signed char *panel_tile = nullptr;   // per-panel shapes of mode kModePerPanel
    std::map<int, std::vector<signed char>> panel_tiles;   // per swarm size
    std::vector<signed char> pinned_tiles;   // FB_PANEL_TILES with FB_FACTOR_PATH=5: every swarm size
    std::vector<std::pair<hipEvent_t, hipEvent_t>> *k1_events = nullptr;   // tuning: times each panel update
    int mode = -1;                   // factorisation path in use (kNumModes below)
    // autotuner: (np, swarm size) -> 0 persistent, 1 lock-step KC 32, 2 lock-step KC 16
    bool autotune = true;
    int forced_mode = -1;            // FB_PERSISTENT=0/1 pins the path
    std::map<int, int> tuned;        // per swarm size (np is fixed per handle)
    std::string tune_file;           // FB_TUNE_CACHE: choices kept across processes
    std::string device_key;          // device the choices were timed on
    int dag_grid = 0, n_xcd = 1;
    int *d_ctl = nullptr, *d_dep = nullptr, *d_dead = nullptr, *h_err = nullptr;
    // swarm size -> schedule plan and device matrix descriptors
    std::map<int, std::pair<fb::DagPlan, fb::DagMatrix *>> dag_plans;
    // adjoint gradient workspace, allocated on first use for up to grad_cap particles
    int grad_cap = 0;
    fb::GradWork grad;
    double *d_grad = nullptr, *h_grad = nullptr;
    unsigned long long *d_ready = nullptr;
    // swarm size -> device task table and its queue offsets
    hipStream_t group_stream[kMaxGroups] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kMaxGroups] = {};
    int device = 0;
    // device buffers
    KernelProgram *d_prog = nullptr;
    double *d_slots = nullptr, *d_rhs = nullptr, *d_theta = nullptr, *d_hyp = nullptr;
    double *d_K = nullptr, *d_logdet = nullptr, *d_dinv = nullptr, *d_ll = nullptr;
    double *d_extra = nullptr, *d_z = nullptr, *d_alpha = nullptr;
    int *d_info = nullptr;
    // pinned staging
    double *h_theta = nullptr, *h_ll = nullptr, *h_extra = nullptr;
    int *h_info = nullptr;
    // captured pipelines of the batched path, by swarm size
    std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
    int graph_epoch = 0;             // bumped whenever captured work goes stale
    // profiling
    int profile = 0;                 // 1: an event pair per launch, 2: per phase only (kinds 0, 4, 5, 6, 8)
    std::vector<EventPair> pending;
    std::vector<hipEvent_t> event_pool;
    double prof_ms[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    long long prof_n[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    long long tune_evals = 0;        // evaluations the autotuner has timed (read as profile kind 7)
    long long dispatch_fallbacks = 0;   // persistent evaluations re-run on lock-step (profile kind 9)
};

namespace {

// Buffers of particles [g0, g0 + S) of a swarm whose theta / ll / info
// arrays start at particle 0.
fb::SwarmBuffers buffers(fb_handle *h, int g0, int S, const double *theta, double *ll, int *info)
{
    fb::SwarmBuffers b;
    b.n = h->n; b.d = h->d; b.np = h->np; b.nrhs = h->nrhs; b.likelihood = h->likelihood;
    b.S = S;
    b.n_slots = h->n_slots; b.n_leaves = h->prog.n_leaves;
    b.slots = h->d_slots; b.RHS = h->d_rhs;
    b.hyp = h->d_hyp + (size_t)g0 * (h->prog.n_par_kernel + fb::kMaxTerms + 1);
    // the transposed hyperparameter table follows the records of max_swarm particles
    b.hypT = h->d_hyp + (size_t)h->max_swarm * (h->prog.n_par_kernel + fb::kMaxTerms + 1);
    b.ldT = h->max_swarm + 16;
    b.theta = theta + (size_t)g0 * h->prog.n_par; b.prog = h->d_prog; b.prog_host = &h->prog;
    b.gemm_kc = h->gemm_kc;
    b.tile = h->tile;
    b.panel_tile = h->panel_tile;
    b.K = h->d_K + (size_t)g0 * h->np * h->np;
    b.info = info + g0; b.ll = ll + g0;
    b.logdet_part = h->d_logdet + (size_t)g0 * h->nsteps;
    b.dinv = h->d_dinv + (size_t)g0 * h->nsteps * fb::kDinv;
    b.extra = h->d_extra + (size_t)g0 * 4; b.zbuf = h->d_z + (size_t)g0 * h->n;
    b.chain = (h->chain && h->gemm_kc == 16 && S <= fb::kChainMaxSwarm)
                  ? h->d_chain + (size_t)g0 * fb::kChainStride : nullptr;
    b.tail = b.chain ? h->tail : 0;
    b.lookahead = b.chain ? h->lookahead : 0;
    b.wait_limit = h->wait_limit;
    return b;
}

hipEvent_t take_event(fb_handle *h)
{
    if (!h->event_pool.empty()) {
        hipEvent_t e = h->event_pool.back();
        h->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Records an event pair around `fn` when profiling.
template <class F>
void timed(fb_handle *h, int kind, hipStream_t st, F fn)
{
    if (!h->profile || (h->profile == 2 && kind >= 1 && kind <= 3)) { fn(); return; }
    EventPair p{kind, take_event(h), take_event(h)};
    if (p.a) (void)hipEventRecord(p.a, st);
    fn();
    if (p.b) (void)hipEventRecord(p.b, st);
    if (p.a && p.b) h->pending.push_back(p);
}

// The lock-step factorisation of the particles of b: one launch per step.
// The launches of panel c0 (column c0, a multiple of 128) of the lock-step
// factorisation of the particles of b: update + first block's
// factorisation (kind 1), panel solve + second block's factorisation (kind
// 2), second block's update + panel solve (kind 3); chained: the update
// launch also solves and factors the second block (kind 1) and the rows
// below are solved in one launch (kind 2).
void enqueue_panel(fb_handle *h, const fb::SwarmBuffers &b, int c0, hipStream_t st)
{
    {
        if (b.chain) {
            const bool rows = fb::panel_rows_launch(b, c0);
            if (h->k1_events) {
                // tuning: the panel's whole time, its row-solve launch included
                std::pair<hipEvent_t, hipEvent_t> ev{take_event(h), take_event(h)};
                (void)hipEventRecord(ev.first, st);
                fb::launch_panel_chain(b, c0, st);
                if (rows) fb::launch_solve_rows(b, c0, st);
                (void)hipEventRecord(ev.second, st);
                h->k1_events->push_back(ev);
            } else {
                timed(h, 1, st, [&] { fb::launch_panel_chain(b, c0, st); });
                if (rows) timed(h, 2, st, [&] { fb::launch_solve_rows(b, c0, st); });
            }
            return;
        }
        const int c1 = c0 + fb::kNBI;
        if (h->k1_events) {
            std::pair<hipEvent_t, hipEvent_t> ev{take_event(h), take_event(h)};
            (void)hipEventRecord(ev.first, st);
            fb::launch_panel_update_factor(b, c0, st);
            (void)hipEventRecord(ev.second, st);
            h->k1_events->push_back(ev);
        } else {
            timed(h, 1, st, [&] { fb::launch_panel_update_factor(b, c0, st); });
        }
        timed(h, 2, st, [&] { fb::launch_trsm_factor(b, c0, st); });
        if (c1 + fb::kNBI < h->np) timed(h, 3, st, [&] { fb::launch_update_trsm(b, c1, st); });
    }
}

// The lock-step factorisation of the particles of b: one launch per step.
void enqueue_factor(fb_handle *h, const fb::SwarmBuffers &b, hipStream_t st)
{
    for (int c0 = 0; c0 < h->np; c0 += fb::kNBO) enqueue_panel(h, b, c0, st);
}

// Lock-step pipeline: build, factorisation (its span: kind 8), likelihood.
void enqueue_group(fb_handle *h, const fb::SwarmBuffers &b, hipStream_t st)
{
    timed(h, 0, st, [&] { fb::launch_build(b, st); });
    timed(h, 8, st, [&] { enqueue_factor(h, b, st); });
    timed(h, 4, st, [&] { fb::launch_likelihood(b, st); });
}

// Persistent pipeline: build, the whole factorisation as one task-graph
// launch, likelihood.
void enqueue_dag(fb_handle *h, const fb::SwarmBuffers &b, hipStream_t st)
{
    fb::DagWork w;
    w.ctl = h->d_ctl; w.dep = h->d_dep; w.slots = h->d_ready; w.dead = h->d_dead;
    const auto &pl = h->dag_plans.at(b.S);
    timed(h, 0, st, [&] { fb::launch_build(b, st); });
    timed(h, 6, st, [&] { fb::launch_dag(pl.first, pl.second, w, h->dag_grid, st); });
    timed(h, 4, st, [&] { fb::launch_likelihood(b, st); });
}

// Schedule plan and device descriptors of a swarm of S (outside any graph
// capture; cached per S).
int dag_plan_for(fb_handle *h, int S)
{
    if (h->dag_plans.count(S)) return FB_OK;
    const fb::SwarmBuffers b = buffers(h, 0, S, h->d_theta, h->d_ll, h->d_info);
    std::vector<fb::DagMatrix> mats(S);
    for (int s = 0; s < S; ++s) {
        fb::DagMatrix &m = mats[s];
        m.K = b.K + (size_t)s * h->np * h->np;
        m.dinv = b.dinv + (size_t)s * h->nsteps * fb::kDinv;
        m.logdet_part = b.logdet_part + (size_t)s * h->nsteps;
        m.info = b.info + s;
        m.np = h->np; m.n = h->n; m.nb = h->nsteps;
    }
    fb::DagPlan pl = fb::dag_plan(mats, h->n_xcd);
    if (!pl.ok) return fail(FB_E_CAPACITY, "swarm too large for one persistent launch (matrices per XCD queue)");
    fb::DagMatrix *d = nullptr;
    FB_HIP(hipMalloc((void **)&d, sizeof(fb::DagMatrix) * S));
    hipError_t e = hipMemcpy(d, mats.data(), sizeof(fb::DagMatrix) * S, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(FB_E_HIP, std::string("schedule upload: ") + hipGetErrorString(e));
    }
    h->dag_plans[S] = std::make_pair(pl, d);
    return FB_OK;
}

// Stream groups of a swarm of S on the lock-step path: FB_STREAMS, else the
// autotuner's choice for S, else one.
int groups_for(const fb_handle *h, int S)
{
    int G = h->forced_groups;
    if (G < 1) {
        auto it = h->groups.find(S);
        G = it != h->groups.end() ? it->second : 1;
    }
    return std::max(1, std::min({G, S, h->n_groups}));
}

// The whole swarm [0, S): the persistent path, or the lock-step one -- with
// G > 1 stream groups, one build of the swarm, the factorisations of G
// groups of particles on concurrent streams forked from and joined back into
// `st` (also inside a graph capture: a launch of one group that leaves the
// chip partly idle in its last round overlaps with the others'), then one
// likelihood launch.
void enqueue_pipeline(fb_handle *h, int S, const double *theta, double *ll, int *info, hipStream_t st)
{
    if (h->use_dag) {
        enqueue_dag(h, buffers(h, 0, S, theta, ll, info), st);
        return;
    }
    const int G = groups_for(h, S);
    if (G == 1) {
        enqueue_group(h, buffers(h, 0, S, theta, ll, info), st);
        return;
    }
    fb::SwarmBuffers all = buffers(h, 0, S, theta, ll, info);
    if (h->chain) all.chain = h->d_chain;   // the build resets every group's chain words
    timed(h, 0, st, [&] { fb::launch_build(all, st); });
    timed(h, 8, st, [&] {
        (void)hipEventRecord(h->fork_ev, st);
        const int per = (S + G - 1) / G;
        fb::SwarmBuffers bg[fb_handle::kMaxGroups];
        hipStream_t gs[fb_handle::kMaxGroups];
        int ng = 0;
        for (int g = 0; g < G; ++g) {
            const int g0 = g * per;
            const int cnt = std::min(per, S - g0);
            if (cnt <= 0) break;
            gs[g] = g == 0 ? st : h->group_stream[g];
            if (g > 0) (void)hipStreamWaitEvent(gs[g], h->fork_ev, 0);
            bg[g] = buffers(h, g0, cnt, theta, ll, info);
            ng = g + 1;
        }
        // panel by panel, group after group: every group's first launches
        // are submitted at once (group after group, a group's launches
        // would wait for the host to submit all of the previous group's);
        // per-panel tuning times the first group's launches
        std::vector<std::pair<hipEvent_t, hipEvent_t>> *k1 = h->k1_events;
        for (int c0 = 0; c0 < h->np; c0 += fb::kNBO)
            for (int g = 0; g < ng; ++g) {
                h->k1_events = g == 0 ? k1 : nullptr;
                enqueue_panel(h, bg[g], c0, gs[g]);
            }
        h->k1_events = k1;
        for (int g = 1; g < ng; ++g) {
            (void)hipEventRecord(h->join_ev[g], gs[g]);
            (void)hipStreamWaitEvent(st, h->join_ev[g], 0);
        }
    });
    timed(h, 4, st, [&] { fb::launch_likelihood(all, st); });
}

int drain_profile(fb_handle *h)
{
    if (h->pending.empty()) return FB_OK;
    FB_HIP(hipStreamSynchronize(h->stream));
    for (const EventPair &p : h->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            h->prof_ms[p.kind] += ms;
            h->prof_n[p.kind] += 1;
        }
        h->event_pool.push_back(p.a);
        h->event_pool.push_back(p.b);
    }
    h->pending.clear();
    return FB_OK;
}

int check_gfx950(int *dev)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
        return fail(FB_E_NOGPU, "no HIP device visible");
    FB_HIP(hipGetDevice(dev));
    hipDeviceProp_t prop;
    FB_HIP(hipGetDeviceProperties(&prop, *dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(FB_E_NOGPU, std::string("device is ") + prop.gcnArchName +
                                    ", this library is built for gfx950 only");
    return FB_OK;
}

void drop_graphs(fb_handle *h)
{
    for (auto &g : h->graphs) {
        (void)hipGraphExecDestroy(g.second.second);
        (void)hipGraphDestroy(g.second.first);
    }
    h->graphs.clear();
    ++h->graph_epoch;
}

void free_grad(fb_handle *h)
{
    for (void *p : {(void *)h->grad.U, (void *)h->grad.NK, (void *)h->grad.X, (void *)h->grad.alpha,
                    (void *)h->grad.partial, (void *)h->d_grad})
        if (p) (void)hipFree(p);
    if (h->h_grad) (void)hipHostFree(h->h_grad);
    h->grad = fb::GradWork();
    h->d_grad = h->h_grad = nullptr;
    h->grad_cap = 0;
}

void free_all(fb_handle *h)
{
    drop_graphs(h);
    free_grad(h);
    for (void *p : {(void *)h->d_ctl, (void *)h->d_dep, (void *)h->d_dead, (void *)h->d_ready})
        if (p) (void)hipFree(p);
    for (auto &pl : h->dag_plans) (void)hipFree(pl.second.second);
    h->dag_plans.clear();
    if (h->h_err) (void)hipHostFree(h->h_err);
    for (const EventPair &p : h->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (hipEvent_t e : h->event_pool) (void)hipEventDestroy(e);
    void *dev[] = {h->d_prog, h->d_slots, h->d_rhs, h->d_theta, h->d_hyp, h->d_chain, h->d_K, h->d_logdet, h->d_dinv,
                   h->d_ll, h->d_extra, h->d_z, h->d_alpha, h->d_info};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    void *host[] = {h->h_theta, h->h_ll, h->h_extra, h->h_info};
    for (void *p : host)
        if (p) (void)hipHostFree(p);
    for (int g = 0; g < fb_handle::kMaxGroups; ++g) {
        if (h->group_stream[g]) (void)hipStreamDestroy(h->group_stream[g]);
        if (h->join_ev[g]) (void)hipEventDestroy(h->join_ev[g]);
    }
    if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
}

}  // namespace

extern "C" {

const char *fb_last_error(void) { return g_err.c_str(); }

int fb_device_name(char *buf, int len)
{
    if (!buf || len < 1) return fail(FB_E_ARG, "fb_device_name: null buffer");
    int dev = 0, count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return fail(FB_E_NOGPU, "no HIP device visible");
    FB_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    FB_HIP(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, (size_t)len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return FB_OK;
}

// Tuning cache file: a JSON array of {"device", "np", "swarm", "path", "ms"}
// records, one per line; read at fb_create, rewritten (temp file + rename)
// after each new choice.  Records of other devices or orders are kept.
struct TuneRecord {
    std::string device;
    int np = 0, swarm = 0, path = -1;
    double ms = 0.0;
    std::string tiles;     // per-panel shapes of the per-panel path, "0,2,1,..."
    int groups = 1;        // stream groups of the lock-step paths
};

// Per-panel tile codes the lock-step launchers accept: a shape (0 64x128,
// 1 32x128, 2 32x64), + kTileTwoRoles; kTileFused, kTileThreePerCU and
// kTileAhead (alone or with kTileThreePerCU) are 64x128 only.
static bool valid_tile_code(int t)
{
    const int shape = t & 3, var = t & ~(3 | fb::kTileTwoRoles);
    if (t < 0 || shape == 3) return false;
    if (var == 0) return true;
    return shape == 0 && (var == fb::kTileFused || var == fb::kTileThreePerCU || var == fb::kTileFourPerCU ||
                          var == fb::kTileAhead || var == (fb::kTileAhead | fb::kTileThreePerCU) ||
                          var == (fb::kTileAhead | fb::kTileFourPerCU));
}

static std::vector<TuneRecord> read_tune_file(const std::string &path)
{
    std::vector<TuneRecord> out;
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return out;
    char line[1024];
    while (std::fgets(line, sizeof line, f)) {
        TuneRecord r;
        const char *d = std::strstr(line, "\"device\": \"");
        const char *a = std::strstr(line, "\"np\": ");
        const char *b = std::strstr(line, "\"swarm\": ");
        const char *c = std::strstr(line, "\"path\": ");
        const char *m = std::strstr(line, "\"ms\": ");
        if (!d || !a || !b || !c) continue;
        d += 11;
        const char *e = std::strchr(d, '"');
        if (!e) continue;
        r.device.assign(d, e);
        r.np = std::atoi(a + 6);
        r.swarm = std::atoi(b + 9);
        r.path = std::atoi(c + 8);
        r.ms = m ? std::atof(m + 6) : 0.0;
        if (const char *g = std::strstr(line, "\"groups\": ")) r.groups = std::max(1, std::atoi(g + 10));
        if (const char *t = std::strstr(line, "\"tiles\": \"")) {
            t += 10;
            const char *te = std::strchr(t, '"');
            if (te) r.tiles.assign(t, te);
        }
        if (r.path >= 0 && (r.path < kNumModes || (r.path == kModePerPanel && !r.tiles.empty()))) out.push_back(r);
    }
    std::fclose(f);
    return out;
}

static void write_tune_file(const fb_handle *h, int s, int path, double ms)
{
    std::vector<TuneRecord> recs = read_tune_file(h->tune_file);
    std::string tiles;
    if (path == kModePerPanel && h->panel_tiles.count(s))
        for (signed char t : h->panel_tiles.at(s)) tiles += (tiles.empty() ? "" : ",") + std::to_string((int)t);
    const int groups = h->groups.count(s) ? h->groups.at(s) : 1;
    bool found = false;
    for (auto &r : recs)
        if (r.device == h->device_key && r.np == h->np && r.swarm == s) {
            r.path = path; r.ms = ms; r.tiles = tiles; r.groups = groups;
            found = true;
        }
    if (!found) recs.push_back({h->device_key, h->np, s, path, ms, tiles, groups});
    const std::string tmp = h->tune_file + ".tmp" + std::to_string((long long)std::rand());
    FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) return;   // a cache that cannot be written only costs re-tuning
    std::fprintf(f, "[\n");
    for (size_t k = 0; k < recs.size(); ++k)
        std::fprintf(f,
                     "  {\"device\": \"%s\", \"np\": %d, \"swarm\": %d, \"path\": %d, \"ms\": %.4f, \"tiles\": \"%s\", "
                     "\"groups\": %d}%s\n",
                     recs[k].device.c_str(), recs[k].np, recs[k].swarm, recs[k].path, recs[k].ms,
                     recs[k].tiles.c_str(), recs[k].groups, k + 1 < recs.size() ? "," : "");
    std::fprintf(f, "]\n");
    const bool ok = std::fclose(f) == 0;
    if (!ok || std::rename(tmp.c_str(), h->tune_file.c_str()) != 0) std::remove(tmp.c_str());
}

int fb_create(fb_handle **out, int n, int d, const double *x, const double *y, const double *mean,
              const fb_kernel_spec *spec, double noise, int optimise_noise, int likelihood,
              int max_swarm)
{
    if (!out) return fail(FB_E_ARG, "fb_create: null handle pointer");
    *out = nullptr;
    if (n < 2 || d < 1 || !x || !y || max_swarm < 1)
        return fail(FB_E_ARG, "fb_create: need n >= 2, d >= 1, x, y and max_swarm >= 1");
    if (likelihood != FB_LIKELIHOOD_MARGINAL && likelihood != FB_LIKELIHOOD_CONCENTRATED)
        return fail(FB_E_ARG, "fb_create: likelihood must be 0 (marginal) or 1 (concentrated)");
    fb_handle *h = new fb_handle();
    std::vector<int> slot_feature, slot_kind;
    int rc = make_program(spec, d, optimise_noise, noise, h->prog, slot_feature, slot_kind);
    if (rc) { delete h; return rc; }
    rc = check_gfx950(&h->device);
    if (rc) { delete h; return rc; }
    h->n = n; h->d = d; h->likelihood = likelihood; h->max_swarm = max_swarm;
    h->nrhs = likelihood == FB_LIKELIHOOD_CONCENTRATED ? 2 : 1;
    h->np = fb::round_up(n + h->nrhs, fb::kNBO);
    h->nsteps = h->np / fb::kNBI;
    h->n_slots = (int)slot_feature.size();

    // per-point slot values: raw feature, or sin(x/2) / cos(x/2)
    std::vector<double> slots((size_t)std::max(h->n_slots, 1) * n, 0.0);
    for (int s = 0; s < h->n_slots; ++s)
        for (int i = 0; i < n; ++i) {
            const double v = x[i + (size_t)n * slot_feature[s]];
            slots[(size_t)s * n + i] = slot_kind[s] == 0 ? v : (slot_kind[s] == 1 ? std::sin(v / 2.0) : std::cos(v / 2.0));
        }
    std::vector<double> rhs((size_t)h->nrhs * n);
    for (int i = 0; i < n; ++i) {
        if (h->nrhs == 1) {
            rhs[i] = y[i] - (mean ? mean[i] : 0.0);   // likelihood.f90:178
        } else {
            rhs[i] = y[i];
            rhs[(size_t)n + i] = 1.0;
        }
    }

    const size_t S = (size_t)max_swarm, np2 = (size_t)h->np * h->np;
    const int hs = h->prog.n_par_kernel + fb::kMaxTerms + 1;
#define FB_ALLOC(ptr, bytes)                                                              \
    do {                                                                                  \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) {                          \
            free_all(h);                                                                  \
            delete h;                                                                     \
            return fail(FB_E_NOMEM, std::string("fb_create: hipMalloc of ") + #ptr + " failed"); \
        }                                                                                 \
    } while (0)
    FB_ALLOC(h->d_prog, sizeof(KernelProgram));
    // slots padded with kSlotPad zeros: the build reads columns up to np - 1
    // of every slot row without a clamp
    FB_ALLOC(h->d_slots, sizeof(double) * (slots.size() + kSlotPad));
    FB_ALLOC(h->d_rhs, sizeof(double) * rhs.size());
    FB_ALLOC(h->d_theta, sizeof(double) * std::max(h->prog.n_par, 1) * S);
    FB_ALLOC(h->d_hyp, sizeof(double) * hs * (2 * S + 16));   // records, then their transpose
    FB_ALLOC(h->d_chain, sizeof(int) * fb::kChainStride * S);
    FB_ALLOC(h->d_K, sizeof(double) * np2 * S);
    FB_ALLOC(h->d_logdet, sizeof(double) * h->nsteps * S);
    FB_ALLOC(h->d_dinv, sizeof(double) * h->nsteps * S * fb::kDinv);
    FB_ALLOC(h->d_ll, sizeof(double) * S);
    FB_ALLOC(h->d_extra, sizeof(double) * 4 * S);
    FB_ALLOC(h->d_z, sizeof(double) * n * S);
    FB_ALLOC(h->d_alpha, sizeof(double) * n);
    FB_ALLOC(h->d_info, sizeof(int) * S);
    // persistent scheduler: counters per task of every matrix, ready lists
    // sized for the largest queue of any swarm size up to S
    h->dag_grid = fb::dag_grid_size(&h->n_xcd);
    FB_ALLOC(h->d_ctl, sizeof(int) * fb::kDagCtl);
    FB_ALLOC(h->d_dep, sizeof(int) * fb::dag_index_count(h->np) * S);
    FB_ALLOC(h->d_dead, sizeof(int) * S);
    FB_ALLOC(h->d_ready, sizeof(unsigned long long) * 2 * fb::dag_tasks_per_matrix(h->np) * S);
#undef FB_ALLOC
    if (hipHostMalloc((void **)&h->h_theta, sizeof(double) * std::max(h->prog.n_par, 1) * S) != hipSuccess ||
        hipHostMalloc((void **)&h->h_ll, sizeof(double) * S) != hipSuccess ||
        hipHostMalloc((void **)&h->h_extra, sizeof(double) * 4 * S) != hipSuccess ||
        hipHostMalloc((void **)&h->h_info, sizeof(int) * S) != hipSuccess ||
        hipHostMalloc((void **)&h->h_err, sizeof(int)) != hipSuccess) {
        free_all(h);
        delete h;
        return fail(FB_E_NOMEM, "fb_create: pinned host allocation failed");
    }
    if (hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking) != hipSuccess) {
        free_all(h);
        delete h;
        return fail(FB_E_HIP, "fb_create: stream creation failed");
    }
    h->stream = h->own_stream;
    if (const char *env = std::getenv("FB_STREAMS")) h->forced_groups = std::atoi(env);
    if (const char *env = std::getenv("FB_GRAPH")) h->use_graph = std::atoi(env) != 0;
    if (const char *env = std::getenv("FB_PERSISTENT")) {
        h->use_dag = std::atoi(env) != 0;
        h->forced_mode = h->use_dag ? 0 : 1;
    }
    if (const char *env = std::getenv("FB_AUTOTUNE")) h->autotune = std::atoi(env) != 0;
    if (const char *env = std::getenv("FB_CHAIN")) h->chain = std::atoi(env) != 0;
    if (const char *env = std::getenv("FB_TAIL")) h->tail = std::max(0, std::atoi(env));
    if (const char *env = std::getenv("FB_LOOKAHEAD")) h->lookahead = std::atoi(env) != 0;
    // a shorter bound for the hand-over waits (a test of the timeout path:
    // 0 gives up at the first poll)
    if (const char *env = std::getenv("FB_CHAIN_WAIT_TICKS")) h->wait_limit = std::strtoull(env, nullptr, 10);
    if (const char *env = std::getenv("FB_FACTOR_PATH")) {
        const int m = std::atoi(env);
        if (m >= 0 && m < kNumModes) h->forced_mode = m;
        // the per-panel path with its shapes given: FB_PANEL_TILES=t0,t1,...
        // (one of 0 64x128, 1 32x128, 2 32x64 per 128-column panel, + 4 for
        // two role work-groups in the chained update launch; 8 fused, 16
        // three work-groups per CU, each 64x128, + 4)
        const char *tl = std::getenv("FB_PANEL_TILES");
        if (m == kModePerPanel && tl) {
            std::vector<signed char> v;
            for (const char *c = tl; *c;) {
                v.push_back((signed char)std::atoi(c));
                while (*c && *c != ',') ++c;
                if (*c == ',') ++c;
            }
            bool ok = (int)v.size() == h->np / fb::kNBO;
            for (signed char t : v) ok = ok && valid_tile_code(t);
            if (ok) {
                h->pinned_tiles = v;
                h->forced_mode = kModePerPanel;
            }
        }
    }
    if (const char *env = std::getenv("FB_TUNE_CACHE")) {
        h->tune_file = env;
        hipDeviceProp_t prop;
        int dev = 0;
        if (!h->tune_file.empty() && hipGetDevice(&dev) == hipSuccess &&
            hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            h->device_key = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount) + "cu";
            for (const auto &r : read_tune_file(h->tune_file)) {
                if (r.device != h->device_key || r.np != h->np || (r.path == 0 && h->dag_grid <= 0)) continue;
                if (r.path == kModePerPanel) {
                    std::vector<signed char> v;
                    for (const char *c = r.tiles.c_str(); *c;) {
                        v.push_back((signed char)std::atoi(c));
                        while (*c && *c != ',') ++c;
                        if (*c == ',') ++c;
                    }
                    bool ok = (int)v.size() == h->np / fb::kNBO;
                    for (signed char t : v) ok = ok && valid_tile_code(t);
                    if (!ok) continue;
                    h->panel_tiles[r.swarm] = v;
                }
                h->tuned[r.swarm] = r.path;
                if (r.groups > 1) h->groups[r.swarm] = std::min(r.groups, (int)fb_handle::kTuneGroups);
            }
        } else {
            h->tune_file.clear();
        }
    }
    if (h->dag_grid <= 0) {
        h->use_dag = false;
        if (h->forced_mode == 0) h->forced_mode = 1;
    }
    h->n_groups = std::max(1, std::min(std::max(h->forced_groups, (int)fb_handle::kTuneGroups), fb_handle::kMaxGroups));
    bool ev_ok = hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming) == hipSuccess;
    for (int g = 1; g < h->n_groups && ev_ok; ++g)
        ev_ok = hipStreamCreateWithFlags(&h->group_stream[g], hipStreamNonBlocking) == hipSuccess &&
                hipEventCreateWithFlags(&h->join_ev[g], hipEventDisableTiming) == hipSuccess;
    if (!ev_ok) {
        free_all(h);
        delete h;
        return fail(FB_E_HIP, "fb_create: stream / event creation failed");
    }
    hipError_t e = hipMemcpy(h->d_prog, &h->prog, sizeof(KernelProgram), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->d_slots, slots.data(), sizeof(double) * slots.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(h->d_slots + slots.size(), 0, sizeof(double) * kSlotPad);
    if (e == hipSuccess) e = hipMemcpy(h->d_rhs, rhs.data(), sizeof(double) * rhs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(h->d_info, 0, sizeof(int) * S);
    if (e == hipSuccess) e = hipMemset(h->d_ctl, 0, sizeof(int) * fb::kDagCtl);
    if (e != hipSuccess) {
        free_all(h);
        delete h;
        return fail(FB_E_HIP, std::string("fb_create: upload failed: ") + hipGetErrorString(e));
    }
    *out = h;
    return FB_OK;
}

int fb_destroy(fb_handle *h)
{
    if (!h) return FB_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_all(h);
    delete h;
    return FB_OK;
}

int fb_nparams(const fb_handle *h, int *p)
{
    if (!h || !p) return fail(FB_E_ARG, "fb_nparams: null argument");
    *p = h->prog.n_par;
    return FB_OK;
}

int fb_set_stream(fb_handle *h, void *stream)
{
    if (!h) return fail(FB_E_ARG, "fb_set_stream: null handle");
    FB_HIP(hipStreamSynchronize(h->stream));
    h->stream = stream ? (hipStream_t)stream : h->own_stream;
    return FB_OK;
}

static void set_mode(fb_handle *h, int mode, int s = 0)
{
    h->mode = mode;
    h->use_dag = mode == 0;
    h->gemm_kc = mode == 1 ? 32 : 16;
    h->tile = (mode >= 3 && mode <= 4) ? mode - 2 : 0;
    h->panel_tile = nullptr;
    if (mode == kModePerPanel) {
        auto it = h->panel_tiles.find(s);
        if (it != h->panel_tiles.end()) h->panel_tile = it->second.data();
        else if (!h->pinned_tiles.empty()) h->panel_tile = h->pinned_tiles.data();
    }
}

int fb_swarm_loglik_async(fb_handle *h, int s, const double *d_theta, double *d_ll, int *d_info)
{
    if (!h || !d_theta || !d_ll || !d_info) return fail(FB_E_ARG, "fb_swarm_loglik_async: null argument");
    if (s < 1 || s > h->max_swarm) return fail(FB_E_CAPACITY, "fb_swarm_loglik_async: swarm size out of range");
    {
        auto it = h->tuned.find(s);
        set_mode(h, h->forced_mode >= 0 ? h->forced_mode : it != h->tuned.end() ? it->second : (h->dag_grid > 0 ? 0 : 1),
                 s);
    }
    if (h->use_dag) {
        // the persistent schedule's descriptors name the handle's INFO buffer
        const int rc = dag_plan_for(h, s);
        if (rc) return rc;
        enqueue_pipeline(h, s, d_theta, d_ll, h->d_info, h->stream);
        FB_HIP(hipMemcpyAsync(d_info, h->d_info, sizeof(int) * s, hipMemcpyDeviceToDevice, h->stream));
    } else {
        enqueue_pipeline(h, s, d_theta, d_ll, d_info, h->stream);
    }
    FB_HIP(hipGetLastError());
    return FB_OK;
}

// Picks the factorisation path for a swarm of s (theta already staged): the
// persistent launch, or the lock-step launches with their GEMM chunk depth
// and panel-update tile shape (kNumModes), each timed twice on the real
// inputs, the fastest kept per swarm size.  All of them compute the same
// bits (the same tile routines and per-element accumulation order), so the
// choice changes only speed.  Not run while a stream capture is active (the
// cached or default choice is used).
static int choose_path(fb_handle *h, int s)
{
    if (h->forced_mode >= 0) { set_mode(h, h->forced_mode); return FB_OK; }
    auto it = h->tuned.find(s);
    if (it != h->tuned.end()) { set_mode(h, it->second, s); return FB_OK; }
    if (!h->autotune || h->dag_grid <= 0) { set_mode(h, h->dag_grid > 0 ? 0 : 1); return FB_OK; }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(h->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
        set_mode(h, 0);
        return FB_OK;
    }
    hipStream_t st = h->stream;
    hipEvent_t e0, e1;
    FB_HIP(hipEventCreate(&e0));
    FB_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int best_mode = 0;
    std::vector<std::pair<float, int>> lockstep_ms;   // (time, mode) of the lock-step paths, one group
    // every lock-step configuration timed (mode, stream groups, per-panel
    // codes), for the run-off at the end
    struct Cand {
        float ms;
        int mode, groups;
        std::vector<signed char> tiles;
    };
    std::vector<Cand> cands;
    for (int mode = 0; mode < kNumModes; ++mode) {
        set_mode(h, mode);
        if (mode == 0 && dag_plan_for(h, s) != FB_OK) continue;
        float ms_best = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            (void)hipEventRecord(e0, st);
            enqueue_pipeline(h, s, h->d_theta, h->d_ll, h->d_info, st);
            ++h->tune_evals;
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms_best = std::min(ms_best, ms);
        }
        if (mode == 0) {   // a persistent run that could not be served does not count
            int err = 0;
            (void)hipMemcpy(&err, h->d_ctl + fb::kDagErrorWord, sizeof(int), hipMemcpyDeviceToHost);
            if (err != 0) ms_best = 1e30f;
        }
        if (mode != 0) {
            lockstep_ms.push_back({ms_best, mode});
            cands.push_back({ms_best, mode, 1, {}});
        }
        if (ms_best < best) { best = ms_best; best_mode = mode; }
    }
    // the panel updates of the three 16-deep shapes (chained: each with three
    // or two role work-groups, and the fused, ahead, three- and four-per-CU
    // 64x128 variants), timed panel by panel: a mix of the fastest variant
    // of each panel is one more candidate
    const int npan = h->np / fb::kNBO;
    auto set_groups = [&](int G) {
        if (G > 1) h->groups[s] = G;
        else h->groups.erase(s);
    };
    auto time_eval = [&]() -> float {
        float ms_best = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            (void)hipEventRecord(e0, st);
            enqueue_pipeline(h, s, h->d_theta, h->d_ll, h->d_info, st);
            ++h->tune_evals;
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms_best = std::min(ms_best, ms);
        }
        return ms_best;
    };
    // the fastest code of each panel with the swarm in G stream groups (timed
    // on the first group's launches, the others running beside them); true
    // when the mix is worth timing as a whole
    auto panel_mix = [&](int G, std::vector<signed char> &mix) -> bool {
        std::vector<int> codes = {0, 1, 2};
        if (h->chain && fb::ceil_div(s, G) <= fb::kChainMaxSwarm) {
            for (int t = 0; t < 3; ++t) codes.push_back(t | fb::kTileTwoRoles);
            codes.push_back(fb::kTileFused);
            codes.push_back(fb::kTileFused | fb::kTileTwoRoles);
            codes.push_back(fb::kTileThreePerCU);
            codes.push_back(fb::kTileThreePerCU | fb::kTileTwoRoles);
            for (int v : {0, fb::kTileTwoRoles, fb::kTileThreePerCU, fb::kTileThreePerCU | fb::kTileTwoRoles})
                codes.push_back(fb::kTileAhead | v);
            for (int v : {0, fb::kTileTwoRoles, fb::kTileAhead, fb::kTileAhead | fb::kTileTwoRoles})
                codes.push_back(fb::kTileFourPerCU | v);
        }
        set_groups(G);
        const int nc = (int)codes.size();
        std::vector<std::vector<float>> tpan(nc, std::vector<float>(npan, 1e30f));
        std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
        // three passes over the candidates, the fastest time of each panel
        // kept (with two, the choice between variants a few per cent apart
        // changed from run to run)
        for (int pass = 0; pass < 3; ++pass)
            for (int t = 0; t < nc; ++t) {
                set_mode(h, 2 + ((codes[t] & fb::kTileFused) ? 0 : (codes[t] & 3)), s);
                h->tile = codes[t];
                evs.clear();
                h->k1_events = &evs;
                enqueue_pipeline(h, s, h->d_theta, h->d_ll, h->d_info, st);
                ++h->tune_evals;
                h->k1_events = nullptr;
                if (hipStreamSynchronize(st) != hipSuccess) break;
                for (size_t k = 0; k < evs.size() && (int)k < npan; ++k) {
                    float ms = 1e30f;
                    (void)hipEventElapsedTime(&ms, evs[k].first, evs[k].second);
                    tpan[t][k] = std::min(tpan[t][k], ms);
                }
                for (auto &e : evs) {
                    h->event_pool.push_back(e.first);
                    h->event_pool.push_back(e.second);
                }
            }
        mix.assign(npan, 0);
        bool mixed = false;
        for (int k = 0; k < npan; ++k) {
            int bt = 0;
            for (int t = 1; t < nc; ++t)
                if (tpan[t][k] < tpan[bt][k]) bt = t;
            mix[k] = (signed char)codes[bt];
            mixed = mixed || mix[k] != mix[0];
        }
        return mixed || (mix[0] & (fb::kTileTwoRoles | fb::kTileFused | fb::kTileThreePerCU | fb::kTileAhead |
                                   fb::kTileFourPerCU));
    };
    std::vector<signed char> mix;
    if (panel_mix(1, mix)) {
        h->panel_tiles[s] = mix;
        set_mode(h, kModePerPanel, s);
        const float ms = time_eval();
        lockstep_ms.push_back({ms, kModePerPanel});
        cands.push_back({ms, kModePerPanel, 1, mix});
        if (ms < best) { best = ms; best_mode = kModePerPanel; }
    }
    // stream groups: the swarm's factorisations split over 2 .. kTuneGroups
    // concurrent streams (the groups' launches fill each other's last
    // rounds), timed with the three fastest lock-step paths of one group,
    // then the per-panel mix re-timed with the best group count
    int best_g = 1;
    if (h->forced_groups < 1 && !lockstep_ms.empty()) {
        std::sort(lockstep_ms.begin(), lockstep_ms.end());
        const int nc = std::min<int>(3, (int)lockstep_ms.size());
        for (int G = 2; G <= std::min((int)fb_handle::kTuneGroups, std::min(s, h->n_groups)); ++G)
            for (int c = 0; c < nc; ++c) {
                const int mode = lockstep_ms[c].second;
                set_mode(h, mode, s);
                set_groups(G);
                const float ms = time_eval();
                cands.push_back({ms, mode, G, mode == kModePerPanel ? h->panel_tiles[s] : std::vector<signed char>()});
                // more groups must win by 1 %: timed eagerly, three or four
                // groups can look a little faster than they replay, and
                // near-ties would flip the choice from run to run
                if (ms < best * (G > best_g ? 0.99f : 1.0f)) { best = ms; best_mode = mode; best_g = G; }
            }
        std::vector<signed char> mixg;
        if (best_g > 1 && panel_mix(best_g, mixg)) {
            const bool had = h->panel_tiles.count(s) != 0;
            const std::vector<signed char> keep = had ? h->panel_tiles[s] : std::vector<signed char>();
            h->panel_tiles[s] = mixg;
            set_mode(h, kModePerPanel, s);
            set_groups(best_g);
            const float ms = time_eval();
            cands.push_back({ms, kModePerPanel, best_g, mixg});
            if (ms < best) {
                best = ms;
                best_mode = kModePerPanel;
            } else if (had) {
                h->panel_tiles[s] = keep;
            } else {
                h->panel_tiles.erase(s);
            }
        }
    }
    // two uniform candidates the per-panel mix can miss (it times each
    // panel beside the other groups' launches): every panel before the tail
    // pre-update solves the previous panel's rows in its update launch, so
    // no row-solve launch runs until the tail; and the same with the panels
    // from the tail pre-update on (but the last) solving their own rows in
    // their update launch, with two role work-groups
    const int q_tail = npan - 1 - h->tail;
    for (int variant = 0; variant < 2; ++variant) {
        if (!(h->chain && q_tail >= 2 && fb::ceil_div(s, std::max(best_g, 1)) <= fb::kChainMaxSwarm)) break;
        std::vector<signed char> ahead(npan, 0);
        for (int k = 1; k < q_tail; ++k) ahead[k] = (signed char)fb::kTileAhead;
        if (variant == 1)
            for (int k = q_tail; k < npan - 1; ++k) ahead[k] = (signed char)(fb::kTileFused | fb::kTileTwoRoles);
        const bool had = h->panel_tiles.count(s) != 0;
        const std::vector<signed char> keep = had ? h->panel_tiles[s] : std::vector<signed char>();
        h->panel_tiles[s] = ahead;
        set_mode(h, kModePerPanel, s);
        set_groups(best_g);
        const float ms = time_eval();
        cands.push_back({ms, kModePerPanel, best_g, ahead});
        if (ms < best) {
            best = ms;
            best_mode = kModePerPanel;
        } else if (had) {
            h->panel_tiles[s] = keep;
        } else {
            h->panel_tiles.erase(s);
        }
    }
    // run-off: single timings of configurations a few per cent apart swap
    // places from run to run, so the best and the runner-up lock-step
    // configurations within 3 % are timed again, alternately, three times
    // each, and the faster kept
    if (best_mode != 0 && cands.size() >= 2) {
        std::sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
        const Cand &a = cands[0];
        const Cand *b = nullptr;
        for (size_t k = 1; k < cands.size() && !b; ++k)
            if (cands[k].mode != a.mode || cands[k].groups != a.groups || cands[k].tiles != a.tiles) b = &cands[k];
        if (b && b->ms < a.ms * 1.03f) {
            auto apply = [&](const Cand &c) {
                set_groups(c.groups);
                if (c.mode == kModePerPanel) h->panel_tiles[s] = c.tiles;
                else h->panel_tiles.erase(s);
                set_mode(h, c.mode, s);
            };
            float ta = 1e30f, tb = 1e30f;
            for (int r = 0; r < 3; ++r) {
                apply(a);
                ta = std::min(ta, time_eval());
                apply(*b);
                tb = std::min(tb, time_eval());
            }
            // as in the group loop, more stream groups must win by 1 %
            const float pa = a.groups > b->groups ? 0.99f : 1.0f;
            const float pb = b->groups > a.groups ? 0.99f : 1.0f;
            const Cand &w = tb * pa < ta * pb ? *b : a;
            apply(w);
            best = std::min(ta, tb);
            best_mode = w.mode;
            best_g = w.groups;
        }
    }
    if (best_g > 1) h->groups[s] = best_g;
    else h->groups.erase(s);
    if (best_mode != kModePerPanel) h->panel_tiles.erase(s);
    h->tuned[s] = best_mode;
    set_mode(h, best_mode, s);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    FB_HIP(hipGetLastError());
    if (!h->tune_file.empty()) write_tune_file(h, s, best_mode, best);
    return FB_OK;
}

// The path for a swarm of s without timing anything (the re-evaluation of a
// PSO iteration's failed particles, whose count changes from one iteration
// to the next): the size's own choice when it has one, else the choice of
// the nearest tuned size -- its per-panel shapes included -- on the
// lock-step launches (the persistent launch would need a schedule
// allocated per size), else 32 x 64 tiles for a few particles and 64 x 128
// ones beyond.  All paths give the same bits.
static void choose_path_untimed(fb_handle *h, int s)
{
    if (h->forced_mode >= 0) { set_mode(h, h->forced_mode, s); return; }
    auto it = h->tuned.find(s);
    if (it != h->tuned.end()) { set_mode(h, it->second, s); return; }
    int best_s = -1;
    for (const auto &t : h->tuned)
        if (best_s < 0 || std::abs(t.first - s) < std::abs(best_s - s)) best_s = t.first;
    int mode = s <= 4 ? 4 : 2;
    if (best_s >= 0 && h->tuned[best_s] != 0) {
        mode = h->tuned[best_s];
        if (mode == kModePerPanel) {
            auto pt = h->panel_tiles.find(best_s);
            if (pt != h->panel_tiles.end()) h->panel_tiles[s] = pt->second;
            else mode = 2;
        }
    }
    set_mode(h, mode, s);
}

// INFO kInfoWaitTimeout: a hand-over wait of the chained factorisation
// schedule reached its bound (1 s by default) instead of hanging.
static int check_waits(const int *info, int s, const char *where)
{
    for (int q = 0; q < s; ++q)
        if (info[q] == fb::kInfoWaitTimeout)
            return fail(FB_E_HIP, std::string(where) + ": a hand-over wait of the factorisation reached its bound");
    return FB_OK;
}

// One evaluation of the swarm already staged in d_theta; leaves ll / INFO in
// the pinned host buffers.  *dag_error: the persistent factorisation's error
// word (0 ok, 1 no ready task for 2 s, 2 a queue was not drained).
static int run_swarm(fb_handle *h, int s, int *dag_error)
{
    hipStream_t st = h->stream;
    if (h->profile) {
        // eager, with an event pair around each launch (with stream groups,
        // launches of different groups overlap; kind 8 is the factorisation's
        // span)
        timed(h, 5, st, [&] { enqueue_pipeline(h, s, h->d_theta, h->d_ll, h->d_info, st); });
    } else if (!h->use_graph) {
        enqueue_pipeline(h, s, h->d_theta, h->d_ll, h->d_info, st);
    } else {
        auto it = h->graphs.find(s);
        if (it == h->graphs.end()) {
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            FB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            enqueue_pipeline(h, s, h->d_theta, h->d_ll, h->d_info, st);
            FB_HIP(hipStreamEndCapture(st, &g));
            FB_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            it = h->graphs.emplace(s, std::make_pair(g, ge)).first;
        }
        FB_HIP(hipGraphLaunch(it->second.second, st));
    }
    FB_HIP(hipMemcpyAsync(h->h_ll, h->d_ll, sizeof(double) * s, hipMemcpyDeviceToHost, st));
    FB_HIP(hipMemcpyAsync(h->h_info, h->d_info, sizeof(int) * s, hipMemcpyDeviceToHost, st));
    *h->h_err = 0;
    if (h->use_dag)
        FB_HIP(hipMemcpyAsync(h->h_err, h->d_ctl + fb::kDagErrorWord, sizeof(int), hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    FB_HIP(hipGetLastError());
    *dag_error = *h->h_err;
    return check_waits(h->h_info, s, "swarm evaluation");
}

int fb_swarm_loglik(fb_handle *h, int s, const double *theta, double *ll, int *info)
{
    if (!h || !theta || !ll || !info) return fail(FB_E_ARG, "fb_swarm_loglik: null argument");
    if (s < 1 || s > h->max_swarm) return fail(FB_E_CAPACITY, "fb_swarm_loglik: swarm size out of range");
    const int P = h->prog.n_par;
    std::memcpy(h->h_theta, theta, sizeof(double) * P * s);
    FB_HIP(hipMemcpyAsync(h->d_theta, h->h_theta, sizeof(double) * P * s, hipMemcpyHostToDevice, h->stream));
    int rc = choose_path(h, s);
    if (rc) return rc;
    if (h->use_dag) {
        rc = dag_plan_for(h, s);
        if (rc) return rc;
    }
    int err = 0;
    rc = run_swarm(h, s, &err);
    if (rc) return rc;
    if (err == 1) return fail(FB_E_HIP, "fb_swarm_loglik: the persistent factorisation found no ready task for 2 s");
    if (err == 2) {
        // some XCD queue found no work-group of its own: this device does not
        // dispatch the way the persistent schedule assumes; use lock-step
        // (counted: profile kind 9)
        ++h->dispatch_fallbacks;
        h->forced_mode = 1;
        set_mode(h, 1);
        drop_graphs(h);
        rc = run_swarm(h, s, &err);
        if (rc) return rc;
    }
    std::memcpy(ll, h->h_ll, sizeof(double) * s);
    std::memcpy(info, h->h_info, sizeof(int) * s);
    return h->profile ? drain_profile(h) : FB_OK;
}

int fb_particle_loglik(fb_handle *h, int s, const double *theta, double *ll, int *info)
{
    if (!h || !theta || !ll || !info) return fail(FB_E_ARG, "fb_particle_loglik: null argument");
    if (s < 1) return fail(FB_E_ARG, "fb_particle_loglik: empty swarm");
    const int P = h->prog.n_par;
    hipStream_t st = h->stream;
    for (int q = 0; q < s; ++q) {
        std::memcpy(h->h_theta, theta + (size_t)P * q, sizeof(double) * P);
        FB_HIP(hipMemcpyAsync(h->d_theta, h->h_theta, sizeof(double) * P, hipMemcpyHostToDevice, st));
        enqueue_group(h, buffers(h, 0, 1, h->d_theta, h->d_ll, h->d_info), st);
        FB_HIP(hipMemcpyAsync(h->h_ll, h->d_ll, sizeof(double), hipMemcpyDeviceToHost, st));
        FB_HIP(hipMemcpyAsync(h->h_info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, st));
        FB_HIP(hipStreamSynchronize(st));
        FB_HIP(hipGetLastError());
        if (int rc = check_waits(h->h_info, 1, "fb_particle_loglik")) return rc;
        ll[q] = h->h_ll[0];
        info[q] = h->h_info[0];
    }
    return h->profile ? drain_profile(h) : FB_OK;
}

int fb_get_alpha(fb_handle *h, const double *theta, double *alpha, double *out4, int *info)
{
    if (!h || !theta || !alpha || !out4 || !info) return fail(FB_E_ARG, "fb_get_alpha: null argument");
    const int P = h->prog.n_par;
    hipStream_t st = h->stream;
    std::memcpy(h->h_theta, theta, sizeof(double) * P);
    FB_HIP(hipMemcpyAsync(h->d_theta, h->h_theta, sizeof(double) * P, hipMemcpyHostToDevice, st));
    fb::SwarmBuffers b = buffers(h, 0, 1, h->d_theta, h->d_ll, h->d_info);
    enqueue_group(h, b, st);
    FB_HIP(hipMemcpyAsync(h->h_info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    if (int rc = check_waits(h->h_info, 1, "fb_get_alpha")) return rc;
    *info = h->h_info[0];
    if (*info != 0) return h->profile ? drain_profile(h) : FB_OK;
    fb::launch_backsolve(b, 0, h->d_alpha, st);
    FB_HIP(hipMemcpyAsync(alpha, h->d_alpha, sizeof(double) * h->n, hipMemcpyDeviceToHost, st));
    FB_HIP(hipMemcpyAsync(h->h_extra, h->d_extra, sizeof(double) * 4, hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    FB_HIP(hipGetLastError());
    std::memcpy(out4, h->h_extra, sizeof(double) * 4);
    return h->profile ? drain_profile(h) : FB_OK;
}

int fb_debug_matrix(fb_handle *h, const double *theta, int stage, double *out, int *info)
{
    if (!h || !theta || !out || !info) return fail(FB_E_ARG, "fb_debug_matrix: null argument");
    if (stage != 0 && stage != 1) return fail(FB_E_ARG, "fb_debug_matrix: stage must be 0 (K) or 1 (L)");
    const int P = h->prog.n_par, n = h->n;
    hipStream_t st = h->stream;
    std::memcpy(h->h_theta, theta, sizeof(double) * P);
    FB_HIP(hipMemcpyAsync(h->d_theta, h->h_theta, sizeof(double) * P, hipMemcpyHostToDevice, st));
    const fb::SwarmBuffers b = buffers(h, 0, 1, h->d_theta, h->d_ll, h->d_info);
    if (stage == 0) fb::launch_build(b, st);
    else enqueue_group(h, b, st);   // the plain path's kernels: the bits of every batched path
    FB_HIP(hipMemcpyAsync(h->h_info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, st));
    FB_HIP(hipMemcpy2DAsync(out, sizeof(double) * n, h->d_K, sizeof(double) * h->np, sizeof(double) * n, n,
                            hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    FB_HIP(hipGetLastError());
    if (int rc = check_waits(h->h_info, 1, "fb_debug_matrix")) return rc;
    *info = h->h_info[0];
    for (int j = 1; j < n; ++j)   // the strictly upper part is never written on the device
        for (int i = 0; i < j; ++i) out[i + (size_t)n * j] = 0.0;
    return h->profile ? drain_profile(h) : FB_OK;
}

int fb_swarm_loglik_grad(fb_handle *h, int s, const double *theta, double *ll, double *grad, int *info)
{
    if (!h || !theta || !ll || !grad || !info) return fail(FB_E_ARG, "fb_swarm_loglik_grad: null argument");
    if (s < 1 || s > h->max_swarm) return fail(FB_E_CAPACITY, "fb_swarm_loglik_grad: swarm size out of range");
    const int P = h->prog.n_par;
    if (s > h->grad_cap) {
        free_grad(h);
        const size_t mat = (size_t)h->np * h->np * s;
        const size_t npart = fb::gradient_partial_count(h->n, h->n_slots, s, P);
        if (hipMalloc((void **)&h->grad.U, sizeof(double) * mat) != hipSuccess ||
            hipMalloc((void **)&h->grad.NK, sizeof(double) * mat) != hipSuccess ||
            hipMalloc((void **)&h->grad.X, sizeof(double) * mat) != hipSuccess ||
            hipMalloc((void **)&h->grad.alpha, sizeof(double) * (size_t)h->n * s) != hipSuccess ||
            hipMalloc((void **)&h->grad.partial, sizeof(double) * npart) != hipSuccess ||
            hipMalloc((void **)&h->d_grad, sizeof(double) * (size_t)P * s) != hipSuccess ||
            hipHostMalloc((void **)&h->h_grad, sizeof(double) * (size_t)P * s) != hipSuccess) {
            free_grad(h);
            (void)hipGetLastError();
            return fail(FB_E_NOMEM, "fb_swarm_loglik_grad: device memory for the gradient workspace");
        }
        h->grad_cap = s;
    }
    h->grad.P = P;
    std::memcpy(h->h_theta, theta, sizeof(double) * P * s);
    FB_HIP(hipMemcpyAsync(h->d_theta, h->h_theta, sizeof(double) * P * s, hipMemcpyHostToDevice, h->stream));
    {
        const int prc = choose_path(h, s);
        if (prc) return prc;
        if (h->use_dag) {
            const int qrc = dag_plan_for(h, s);
            if (qrc) return qrc;
        }
    }
    int err = 0;
    int rc = run_swarm(h, s, &err);
    if (rc) return rc;
    if (err == 1) return fail(FB_E_HIP, "fb_swarm_loglik_grad: the persistent factorisation found no ready task for 2 s");
    if (err == 2) {
        ++h->dispatch_fallbacks;
        h->forced_mode = 1;
        set_mode(h, 1);
        drop_graphs(h);
        rc = run_swarm(h, s, &err);
        if (rc) return rc;
    }
    hipStream_t st = h->stream;
    fb::launch_gradient(buffers(h, 0, s, h->d_theta, h->d_ll, h->d_info), h->grad, h->d_grad, st);
    FB_HIP(hipGetLastError());
    FB_HIP(hipMemcpyAsync(h->h_grad, h->d_grad, sizeof(double) * P * s, hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    std::memcpy(ll, h->h_ll, sizeof(double) * s);
    std::memcpy(info, h->h_info, sizeof(int) * s);
    std::memcpy(grad, h->h_grad, sizeof(double) * P * s);
    return h->profile ? drain_profile(h) : FB_OK;
}

// ---- evaluation with no allocation: the library's own scratch array ---------

}  // extern "C"

namespace {

// Work buffers of fb_swarm_loglik_static: hyperparameters, chain words,
// the swarm's K matrices, partial log-determinants, 16x16 inverses and
// flags, likelihood extras and residuals, for every call on this device.
// In host code the name is a stand-in, not a device address: the address
// comes from hipGetSymbolAddress, once per device.
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// The scratch array's address on the current device, looked up on the
// device's first call (with that device current) and kept.
int scratch_address(void **out)
{
    int dev = 0;
    FB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(FB_E_NOGPU, "fb_swarm_loglik_static: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(FB_E_HIP, std::string("fb_swarm_loglik_static: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FB_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" {

int fb_swarm_loglik_static(fb_handle *h, int s, const double *d_theta, double *d_ll, int *d_info, void *stream)
{
    if (!h || !d_theta || !d_ll || !d_info) return fail(FB_E_ARG, "fb_swarm_loglik_static: null argument");
    if (s < 1) return fail(FB_E_ARG, "fb_swarm_loglik_static: empty swarm");
    const size_t S = (size_t)s, np2 = (size_t)h->np * h->np;
    const size_t hs = (size_t)h->prog.n_par_kernel + fb::kMaxTerms + 1;
    const size_t off_hyp = 0;
    const size_t off_chain = align256(off_hyp + sizeof(double) * hs * (2 * S + 16));   // records, transpose
    const size_t off_K = align256(off_chain + sizeof(int) * fb::kChainStride * S);
    const size_t off_logdet = align256(off_K + sizeof(double) * np2 * S);
    const size_t off_dinv = align256(off_logdet + sizeof(double) * h->nsteps * S);
    const size_t off_extra = align256(off_dinv + sizeof(double) * h->nsteps * fb::kDinv * S);
    const size_t off_z = align256(off_extra + sizeof(double) * 4 * S);
    const size_t total = align256(off_z + sizeof(double) * h->n * S);
    if (total > kStaticScratchBytes)
        return fail(FB_E_CAPACITY, "fb_swarm_loglik_static: swarm needs " + std::to_string(total >> 20) +
                                       " MiB of scratch, the library holds " +
                                       std::to_string(kStaticScratchBytes >> 20) + " MiB per device");
    void *base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    unsigned char *w = static_cast<unsigned char *>(base);
    fb::SwarmBuffers b;
    b.n = h->n; b.d = h->d; b.np = h->np; b.nrhs = h->nrhs; b.likelihood = h->likelihood;
    b.S = s;
    b.n_slots = h->n_slots; b.n_leaves = h->prog.n_leaves;
    b.slots = h->d_slots; b.RHS = h->d_rhs; b.prog = h->d_prog; b.prog_host = &h->prog;
    b.theta = d_theta; b.ll = d_ll; b.info = d_info;
    b.hyp = reinterpret_cast<double *>(w + off_hyp);
    b.hypT = b.hyp + hs * S;
    b.ldT = s + 16;
    b.K = reinterpret_cast<double *>(w + off_K);
    b.logdet_part = reinterpret_cast<double *>(w + off_logdet);
    b.dinv = reinterpret_cast<double *>(w + off_dinv);
    b.extra = reinterpret_cast<double *>(w + off_extra);
    b.zbuf = reinterpret_cast<double *>(w + off_z);
    // one fixed lock-step schedule (16-deep chunks, 64 x 128 tiles; chained
    // up to kChainMaxSwarm particles, three launches per panel beyond): the
    // same tile routines in the same order as every other path, so the same
    // bits; no tuning, no graph, no persistent-schedule descriptors
    b.gemm_kc = 16;
    b.tile = 0;
    b.panel_tile = nullptr;
    b.chain = (h->chain && s <= fb::kChainMaxSwarm) ? reinterpret_cast<int *>(w + off_chain) : nullptr;
    b.tail = b.chain ? h->tail : 0;
    b.lookahead = 0;
    b.wait_limit = h->wait_limit;
    hipStream_t st = stream ? (hipStream_t)stream : h->stream;
    fb::launch_build(b, st);
    for (int c0 = 0; c0 < h->np; c0 += fb::kNBO) {
        if (b.chain) {
            fb::launch_panel_chain(b, c0, st);
            if (fb::panel_rows_launch(b, c0)) fb::launch_solve_rows(b, c0, st);
            continue;
        }
        fb::launch_panel_update_factor(b, c0, st);
        fb::launch_trsm_factor(b, c0, st);
        if (c0 + fb::kNBI + fb::kNBI < h->np) fb::launch_update_trsm(b, c0 + fb::kNBI, st);
    }
    fb::launch_likelihood(b, st);
    FB_HIP(hipGetLastError());
    return FB_OK;
}

// ---- several handles in one persistent factorisation ------------------------

// Schedules of fb_multi_swarm_loglik.
enum { kMultiSequential = 0, kMultiPersistent = 1, kMultiMerged = 2, kMultiSchedules = 3 };

struct fb_multi {
    std::vector<fb_handle *> gps;
    int dag_grid = 0, n_xcd = 1;
    fb::DagWork w;
    int *h_err = nullptr;
    std::map<std::vector<int>, std::pair<fb::DagPlan, fb::DagMatrix *>> plans;   // per particle counts
    // per particle counts: the schedule -- each handle's own tuned batched
    // path one after another (kMultiSequential), all factorisations in one
    // persistent launch (kMultiPersistent), or the chained lock-step schedule
    // with every handle's matrices in each launch (kMultiMerged) -- timed on
    // the first call; FB_MULTI_SCHEDULE=0/1/2 (or FB_MULTI_PERSISTENT=0/1)
    // pins it
    std::map<std::vector<int>, int> schedule;
    std::map<std::vector<int>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;   // merged schedule, per counts
    // fb_multi_pso_step: one captured iteration per schedule, set of swarms
    // (their ids) and the handles' graph epochs
    std::map<std::vector<long long>, std::pair<hipGraph_t, hipGraphExec_t>> pso_graphs;
    int forced = -1;
    int merged_tile = fb::kTileFused;   // panel tile code of the merged launches (FB_MULTI_TILE)
    int merged_tail = 3;            // panels pre-updated at the end of each matrix (FB_MULTI_TAIL)
};

namespace {

void free_multi(fb_multi *m)
{
    for (void *p : {(void *)m->w.ctl, (void *)m->w.dep, (void *)m->w.slots, (void *)m->w.dead})
        if (p) (void)hipFree(p);
    for (auto &pl : m->plans) (void)hipFree(pl.second.second);
    for (auto &g : m->graphs) {
        (void)hipGraphExecDestroy(g.second.second);
        (void)hipGraphDestroy(g.second.first);
    }
    for (auto &g : m->pso_graphs) {
        (void)hipGraphExecDestroy(g.second.second);
        (void)hipGraphDestroy(g.second.first);
    }
    if (m->h_err) (void)hipHostFree(m->h_err);
}

// Buffers of a handle's swarm in the merged schedule at theta (device): the
// chained lock-step variant (16-deep chunks, the chain words reset by the
// build; the pre-updates of the last panels are set per launch segment).
fb::SwarmBuffers merged_buffers(fb_handle *h, int s, const double *theta)
{
    fb::SwarmBuffers b = buffers(h, 0, s, theta, h->d_ll, h->d_info);
    b.gemm_kc = 16;
    b.tile = 0;
    b.panel_tile = nullptr;
    b.chain = h->d_chain;
    b.tail = 0;
    b.lookahead = 0;
    return b;
}

// The merged schedule: every handle's build; then launch p of the chained
// schedule holds panel p - (P - P_g) of handle g (P_g panels of 128
// columns, P the most), so that the last, smallest panels of all the
// handles share the last launches, where one swarm alone fills the chip
// least; then every handle's likelihood.
void enqueue_merged(fb_multi *m, const std::vector<int> &sv, const double *const *theta, hipStream_t st)
{
    const int G = (int)m->gps.size();
    std::vector<fb::SwarmBuffers> bs(G);
    int P = 0;
    for (int g = 0; g < G; ++g) {
        if (sv[g] == 0) continue;
        bs[g] = merged_buffers(m->gps[g], sv[g], theta[g]);
        fb::launch_build(bs[g], st);
        P = std::max(P, bs[g].np / fb::kNBO);
    }
    for (int p = 0; p < P; ++p) {
        fb::ChainMulti cm;
        cm.tile = m->merged_tile;
        cm.wait_limit = ~0ull;
        for (int g = 0; g < G; ++g) {
            if (sv[g] == 0) continue;
            const fb::SwarmBuffers &b = bs[g];
            const int q = p - (P - b.np / fb::kNBO);
            if (q < 0) continue;
            fb::ChainSegment &sg = cm.seg[cm.G++];
            sg.K = b.K; sg.info = b.info; sg.logdet_part = b.logdet_part; sg.dinv = b.dinv; sg.chain = b.chain;
            sg.np = b.np; sg.n = b.n; sg.S = b.S; sg.c0 = q * fb::kNBO;
            sg.tail_set = m->merged_tail;
            cm.wait_limit = std::min(cm.wait_limit, b.wait_limit);
            if (cm.G == fb::kMultiMaxSegments) {
                fb::launch_chain_multi(cm, st);
                cm.G = 0;
            }
        }
        if (cm.G) fb::launch_chain_multi(cm, st);
    }
    for (int g = 0; g < G; ++g)
        if (sv[g] > 0) fb::launch_likelihood(bs[g], st);
}

// The merged schedule needs the chained schedule's words in every handle.
bool merged_possible(const fb_multi *m, const std::vector<int> &sv)
{
    for (size_t g = 0; g < m->gps.size(); ++g)
        if (sv[g] > 0 && !m->gps[g]->chain) return false;
    return true;
}

int multi_plan(fb_multi *m, const std::vector<int> &s)
{
    if (m->plans.count(s)) return FB_OK;
    std::vector<fb::DagMatrix> mats;
    for (size_t g = 0; g < m->gps.size(); ++g) {
        fb_handle *h = m->gps[g];
        const fb::SwarmBuffers b = buffers(h, 0, std::max(s[g], 1), h->d_theta, h->d_ll, h->d_info);
        for (int k = 0; k < s[g]; ++k) {
            fb::DagMatrix d;
            d.K = b.K + (size_t)k * h->np * h->np;
            d.dinv = b.dinv + (size_t)k * h->nsteps * fb::kDinv;
            d.logdet_part = b.logdet_part + (size_t)k * h->nsteps;
            d.info = b.info + k;
            d.np = h->np; d.n = h->n; d.nb = h->nsteps;
            mats.push_back(d);
        }
    }
    if (mats.empty()) return FB_OK;
    fb::DagPlan pl = fb::dag_plan(mats, m->n_xcd);
    if (!pl.ok) return fail(FB_E_CAPACITY, "fb_multi: too many matrices for one persistent launch");
    fb::DagMatrix *d = nullptr;
    FB_HIP(hipMalloc((void **)&d, sizeof(fb::DagMatrix) * mats.size()));
    hipError_t e = hipMemcpy(d, mats.data(), sizeof(fb::DagMatrix) * mats.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(FB_E_HIP, std::string("fb_multi: schedule upload: ") + hipGetErrorString(e));
    }
    m->plans[s] = std::make_pair(pl, d);
    return FB_OK;
}

}  // namespace

int fb_multi_create(fb_multi **out, int n_gps, fb_handle *const *gps)
{
    if (!out || n_gps < 1 || !gps) return fail(FB_E_ARG, "fb_multi_create: bad argument");
    *out = nullptr;
    fb_multi *m = new fb_multi();
    size_t dep = 0, slots = 0, mats = 0;
    for (int g = 0; g < n_gps; ++g) {
        if (!gps[g]) {
            delete m;
            return fail(FB_E_ARG, "fb_multi_create: null handle");
        }
        fb_handle *h = gps[g];
        m->gps.push_back(h);
        dep += (size_t)fb::dag_index_count(h->np) * h->max_swarm;
        slots += 2 * (size_t)fb::dag_tasks_per_matrix(h->np) * h->max_swarm;
        mats += h->max_swarm;
    }
    m->dag_grid = fb::dag_grid_size(&m->n_xcd);
    if (m->dag_grid <= 0 || hipMalloc((void **)&m->w.ctl, sizeof(int) * fb::kDagCtl) != hipSuccess ||
        hipMalloc((void **)&m->w.dep, sizeof(int) * dep) != hipSuccess ||
        hipMalloc((void **)&m->w.slots, sizeof(unsigned long long) * slots) != hipSuccess ||
        hipMalloc((void **)&m->w.dead, sizeof(int) * mats) != hipSuccess ||
        hipHostMalloc((void **)&m->h_err, sizeof(int)) != hipSuccess) {
        free_multi(m);
        delete m;
        (void)hipGetLastError();
        return fail(FB_E_NOMEM, "fb_multi_create: scheduler workspace");
    }
    if (const char *env = std::getenv("FB_MULTI_PERSISTENT"))
        m->forced = std::atoi(env) != 0 ? kMultiPersistent : kMultiSequential;
    if (const char *env = std::getenv("FB_MULTI_SCHEDULE")) {
        const int v = std::atoi(env);
        if (v >= 0 && v < kMultiSchedules) m->forced = v;
    }
    if (const char *env = std::getenv("FB_MULTI_TILE")) m->merged_tile = std::atoi(env);
    if (const char *env = std::getenv("FB_MULTI_TAIL")) m->merged_tail = std::max(0, std::atoi(env));
    *out = m;
    return FB_OK;
}

int fb_multi_schedule(const fb_multi *m, const int *s)
{
    if (!m || !s) return -1;
    auto it = m->schedule.find(std::vector<int>(s, s + m->gps.size()));
    return it == m->schedule.end() ? -1 : it->second;
}

int fb_multi_destroy(fb_multi *m)
{
    if (!m) return FB_OK;
    if (!m->gps.empty() && m->gps[0]->stream) (void)hipStreamSynchronize(m->gps[0]->stream);
    free_multi(m);
    delete m;
    return FB_OK;
}

namespace {

// One evaluation of every swarm (s[g] particles at theta[g], device) on
// schedule `mode`, enqueued on st; ll / INFO land in each handle's buffers.
// The sequential schedule runs each handle's tuned batched path for its
// swarm size, the persistent one needs the plan of multi_plan.
void multi_enqueue(fb_multi *m, const std::vector<int> &sv, const double *const *theta, int mode, hipStream_t st)
{
    const int G = (int)m->gps.size();
    if (mode == kMultiMerged) {
        enqueue_merged(m, sv, theta, st);
        return;
    }
    const bool persistent = mode == kMultiPersistent;
    for (int g = 0; g < G; ++g) {
        if (sv[g] == 0) continue;
        fb_handle *h = m->gps[g];
        if (!persistent) {
            auto it = h->tuned.find(sv[g]);
            set_mode(h, it != h->tuned.end() && it->second != 0 ? it->second : 2, sv[g]);
        }
        const fb::SwarmBuffers b = buffers(h, 0, sv[g], theta[g], h->d_ll, h->d_info);
        if (persistent) fb::launch_build(b, st);
        else enqueue_group(h, b, st);
    }
    if (!persistent) return;
    if (m->plans.at(sv).second)
        fb::launch_dag(m->plans.at(sv).first, m->plans.at(sv).second, m->w, m->dag_grid, st);
    for (int g = 0; g < G; ++g)
        if (sv[g] > 0) fb::launch_likelihood(buffers(m->gps[g], 0, sv[g], theta[g], m->gps[g]->d_ll,
                                                     m->gps[g]->d_info), st);
}

// Evaluates the swarms staged in the handles' theta buffers on schedule
// `mode` (the merged one as a captured graph, one per set of swarm sizes)
// and reads ll / INFO back into the handles' pinned buffers; *m->h_err: the
// persistent launch's error word.
int multi_run(fb_multi *m, const std::vector<int> &sv, int mode)
{
    const int G = (int)m->gps.size();
    hipStream_t st = m->gps[0]->stream;
    std::vector<const double *> th(G);
    for (int g = 0; g < G; ++g) th[g] = m->gps[g]->d_theta;
    if (mode == kMultiMerged) {
        auto it = m->graphs.find(sv);
        if (it == m->graphs.end()) {
            hipGraph_t gr = nullptr;
            hipGraphExec_t ge = nullptr;
            FB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            enqueue_merged(m, sv, th.data(), st);
            FB_HIP(hipStreamEndCapture(st, &gr));
            if (hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGraphDestroy(gr);
                return fail(FB_E_HIP, "fb_multi: graph instantiation failed");
            }
            it = m->graphs.emplace(sv, std::make_pair(gr, ge)).first;
        }
        FB_HIP(hipGraphLaunch(it->second.second, st));
    } else {
        multi_enqueue(m, sv, th.data(), mode, st);
    }
    for (int g = 0; g < G; ++g) {
        if (sv[g] == 0) continue;
        fb_handle *h = m->gps[g];
        FB_HIP(hipMemcpyAsync(h->h_ll, h->d_ll, sizeof(double) * sv[g], hipMemcpyDeviceToHost, st));
        FB_HIP(hipMemcpyAsync(h->h_info, h->d_info, sizeof(int) * sv[g], hipMemcpyDeviceToHost, st));
    }
    *m->h_err = 0;
    if (mode == kMultiPersistent)
        FB_HIP(hipMemcpyAsync(m->h_err, m->w.ctl + fb::kDagErrorWord, sizeof(int), hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    FB_HIP(hipGetLastError());
    return FB_OK;
}

// The plan of the persistent schedule for these swarm sizes (FB_E_CAPACITY
// when it does not fit one launch).
int multi_plan_for(fb_multi *m, const std::vector<int> &sv)
{
    return m->plans.count(sv) ? FB_OK : multi_plan(m, sv);
}

// Chooses the schedule for the swarm sizes sv (the pinned one, or the
// fastest of two timed evaluations each on the swarms staged in the
// handles' theta buffers, after tuning every handle for its size).
int multi_choose(fb_multi *m, const std::vector<int> &sv)
{
    const int G = (int)m->gps.size();
    const bool merged = merged_possible(m, sv);
    if (m->forced >= 0) {
        if (m->forced == kMultiPersistent) {
            if (int rc = multi_plan_for(m, sv)) return rc;
        }
        if (m->forced == kMultiMerged && !merged)
            return fail(FB_E_ARG, "fb_multi: the merged schedule needs the chained schedule (FB_CHAIN=0 is set)");
        m->schedule[sv] = m->forced;
        return FB_OK;
    }
    hipStream_t st = m->gps[0]->stream;
    FB_HIP(hipStreamSynchronize(st));
    for (int g = 0; g < G; ++g)
        if (sv[g] > 0) {
            if (int rc = choose_path(m->gps[g], sv[g])) return rc;
        }
    // the persistent schedule is a candidate only where its plan fits
    const bool can[kMultiSchedules] = {true, multi_plan_for(m, sv) == FB_OK, merged};
    hipEvent_t e0, e1;
    FB_HIP(hipEventCreate(&e0));
    FB_HIP(hipEventCreate(&e1));
    float best[kMultiSchedules] = {1e30f, 1e30f, 1e30f};
    int rc = FB_OK;
    for (int rep = 0; rep < 2 && rc == FB_OK; ++rep)
        for (int p = 0; p < kMultiSchedules; ++p) {
            if (!can[p]) continue;
            (void)hipEventRecord(e0, st);
            rc = multi_run(m, sv, p);
            (void)hipEventRecord(e1, st);
            if (rc || hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (p == kMultiPersistent && *m->h_err != 0) ms = 1e30f;   // could not be served
            best[p] = std::min(best[p], ms);
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    int pick = kMultiSequential;
    for (int p = 1; p < kMultiSchedules; ++p)
        if (best[p] < best[pick]) pick = p;
    m->schedule[sv] = pick;
    return FB_OK;
}

}  // namespace

int fb_multi_swarm_loglik(fb_multi *m, const int *s, const double *const *theta, double *const *ll,
                          int *const *info)
{
    if (!m || !s || !theta || !ll || !info) return fail(FB_E_ARG, "fb_multi_swarm_loglik: null argument");
    const int G = (int)m->gps.size();
    std::vector<int> sv(s, s + G);
    for (int g = 0; g < G; ++g) {
        if (sv[g] < 0 || sv[g] > m->gps[g]->max_swarm)
            return fail(FB_E_CAPACITY, "fb_multi_swarm_loglik: swarm size out of range");
        if (sv[g] > 0 && (!theta[g] || !ll[g] || !info[g]))
            return fail(FB_E_ARG, "fb_multi_swarm_loglik: null argument");
    }
    if (std::all_of(sv.begin(), sv.end(), [](int v) { return v == 0; })) return FB_OK;
    hipStream_t st = m->gps[0]->stream;
    for (int g = 0; g < G; ++g) {
        if (sv[g] == 0) continue;
        fb_handle *h = m->gps[g];
        const int P = h->prog.n_par;
        std::memcpy(h->h_theta, theta[g], sizeof(double) * P * sv[g]);
        FB_HIP(hipMemcpyAsync(h->d_theta, h->h_theta, sizeof(double) * P * sv[g], hipMemcpyHostToDevice, st));
    }
    int rc = FB_OK;
    if (!m->schedule.count(sv)) {
        rc = multi_choose(m, sv);
        if (rc) return rc;
    }
    const int mode = m->schedule.at(sv);
    if (mode == kMultiPersistent) {
        rc = multi_plan_for(m, sv);
        if (rc) return rc;
    }
    rc = multi_run(m, sv, mode);
    if (rc) return rc;
    if (mode == kMultiPersistent && *m->h_err == 1)
        return fail(FB_E_HIP, "fb_multi_swarm_loglik: the persistent factorisation found no ready task for 2 s");
    if (mode == kMultiPersistent && *m->h_err == 2) {
        for (int g = 0; g < G; ++g) ++m->gps[g]->dispatch_fallbacks;
        m->schedule[sv] = kMultiSequential;
        rc = multi_run(m, sv, kMultiSequential);   // some XCD queue was not served: lock-step path
        if (rc) return rc;
    }
    for (int g = 0; g < G; ++g)
        if (sv[g] > 0) {
            rc = check_waits(m->gps[g]->h_info, sv[g], "fb_multi_swarm_loglik");
            if (rc) return rc;
        }
    for (int g = 0; g < G; ++g) {
        if (sv[g] == 0) continue;
        std::memcpy(ll[g], m->gps[g]->h_ll, sizeof(double) * sv[g]);
        std::memcpy(info[g], m->gps[g]->h_info, sizeof(int) * sv[g]);
    }
    return FB_OK;
}


// ---- particle swarm on the device ------------------------------------------

struct fb_pso {
    fb_handle *h = nullptr;
    long long id = 0;                                 // creation number (keys of captured multi steps)
    fb::PsoState st;
    double coef[3] = {0.0, 0.0, 0.0};                 // w, c1, c2 of the next step
    double *d_r = nullptr, *h_r = nullptr;            // r1 | r2 | w c1 c2, device and pinned staging
    double *h_cost = nullptr, *h_state = nullptr;     // pinned read-back
    int *h_info = nullptr, *h_gidx = nullptr;
    int *d_idx = nullptr, *h_idx = nullptr;           // retry lists
    double *d_xs = nullptr, *h_xs = nullptr;          // retry positions | velocities
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_epoch = -1;
    int graph_mode = -1;                              // factorisation path captured
};

namespace {

void pso_drop_graph(fb_pso *p)
{
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

void free_pso(fb_pso *p)
{
    pso_drop_graph(p);
    for (void *q : {(void *)p->st.x, (void *)p->st.v, (void *)p->st.pbest, (void *)p->st.pcost,
                    (void *)p->st.gbest, (void *)p->st.cost, (void *)p->st.state, (void *)p->st.info,
                    (void *)p->st.pending, (void *)p->st.gidx, (void *)p->d_r, (void *)p->d_idx, (void *)p->d_xs})
        if (q) (void)hipFree(q);
    for (void *q : {(void *)p->h_r, (void *)p->h_cost, (void *)p->h_state, (void *)p->h_info, (void *)p->h_gidx,
                    (void *)p->h_idx, (void *)p->h_xs})
        if (q) (void)hipHostFree(q);
}

// The swarm's evaluation at its current positions (ll / INFO of the handle,
// then costs), on the handle's chosen path for S.
void pso_enqueue_eval(fb_pso *p, hipStream_t st)
{
    fb_handle *h = p->h;
    enqueue_pipeline(h, p->st.S, p->st.x, h->d_ll, h->d_info, st);
    fb::launch_pso_cost(p->st, h->d_ll, h->d_info, st);
}

// Path for the swarm size and its schedule; the positions are copied to the
// handle's theta buffer first, since a first call times the candidates there.
int pso_prepare(fb_pso *p)
{
    fb_handle *h = p->h;
    const int S = p->st.S;
    if (!h->tuned.count(S) && h->forced_mode < 0)
        FB_HIP(hipMemcpyAsync(h->d_theta, p->st.x, sizeof(double) * S * p->st.P, hipMemcpyDeviceToDevice,
                              h->stream));
    int rc = choose_path(h, S);
    if (rc) return rc;
    if (h->use_dag) {
        rc = dag_plan_for(h, S);
        if (rc) return rc;
    }
    return FB_OK;
}

// Reads back costs / INFO (and the persistent path's error word); on a
// device whose dispatch the persistent schedule cannot use, re-evaluates the
// current positions on the lock-step path.
int pso_finish(fb_pso *p, double *cost, int *info)
{
    fb_handle *h = p->h;
    const int S = p->st.S;
    hipStream_t st = h->stream;
    FB_HIP(hipMemcpyAsync(p->h_cost, p->st.cost, sizeof(double) * S, hipMemcpyDeviceToHost, st));
    FB_HIP(hipMemcpyAsync(p->h_info, p->st.info, sizeof(int) * S, hipMemcpyDeviceToHost, st));
    *h->h_err = 0;
    if (h->use_dag)
        FB_HIP(hipMemcpyAsync(h->h_err, h->d_ctl + fb::kDagErrorWord, sizeof(int), hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    FB_HIP(hipGetLastError());
    if (int rc = check_waits(p->h_info, S, "PSO")) return rc;
    if (*h->h_err == 1) return fail(FB_E_HIP, "PSO: the persistent factorisation found no ready task for 2 s");
    if (*h->h_err == 2) {
        ++h->dispatch_fallbacks;
        h->forced_mode = 1;
        set_mode(h, 1);
        drop_graphs(h);
        pso_drop_graph(p);
        pso_enqueue_eval(p, st);
        FB_HIP(hipMemcpyAsync(p->h_cost, p->st.cost, sizeof(double) * S, hipMemcpyDeviceToHost, st));
        FB_HIP(hipMemcpyAsync(p->h_info, p->st.info, sizeof(int) * S, hipMemcpyDeviceToHost, st));
        FB_HIP(hipStreamSynchronize(st));
        FB_HIP(hipGetLastError());
    }
    if (cost) std::memcpy(cost, p->h_cost, sizeof(double) * S);
    if (info) std::memcpy(info, p->h_info, sizeof(int) * S);
    return FB_OK;
}

}  // namespace

int fb_pso_create(fb_pso **out, fb_handle *h, int s, double w, double c1, double c2)
{
    if (!out || !h) return fail(FB_E_ARG, "fb_pso_create: null argument");
    *out = nullptr;
    if (s < 1 || s > h->max_swarm) return fail(FB_E_CAPACITY, "fb_pso_create: swarm size out of range");
    const int P = h->prog.n_par;
    if (P < 1) return fail(FB_E_ARG, "fb_pso_create: the model has no parameters to optimise");
    static std::atomic<long long> next_id{1};
    fb_pso *p = new fb_pso;
    p->h = h;
    p->id = next_id++;
    p->st.S = s; p->st.P = P;
    p->coef[0] = w; p->coef[1] = c1; p->coef[2] = c2;
    const size_t SP = (size_t)s * P;
    bool ok = true;
    auto dev = [&](void **q, size_t bytes) { ok = ok && hipMalloc(q, bytes) == hipSuccess; };
    auto pin = [&](void **q, size_t bytes) { ok = ok && hipHostMalloc(q, bytes) == hipSuccess; };
    dev((void **)&p->st.x, sizeof(double) * SP);
    dev((void **)&p->st.v, sizeof(double) * SP);
    dev((void **)&p->st.pbest, sizeof(double) * SP);
    dev((void **)&p->st.pcost, sizeof(double) * s);
    dev((void **)&p->st.gbest, sizeof(double) * P);
    dev((void **)&p->st.cost, sizeof(double) * s);
    dev((void **)&p->st.state, sizeof(double) * 2);
    dev((void **)&p->st.info, sizeof(int) * s);
    dev((void **)&p->st.pending, sizeof(int));
    dev((void **)&p->st.gidx, sizeof(int));
    dev((void **)&p->d_r, sizeof(double) * (2 * SP + 3));
    dev((void **)&p->d_idx, sizeof(int) * s);
    dev((void **)&p->d_xs, sizeof(double) * 2 * SP);
    pin((void **)&p->h_r, sizeof(double) * (2 * SP + 3));
    pin((void **)&p->h_cost, sizeof(double) * s);
    pin((void **)&p->h_state, sizeof(double) * (2 + P));
    pin((void **)&p->h_info, sizeof(int) * s);
    pin((void **)&p->h_gidx, sizeof(int));
    pin((void **)&p->h_idx, sizeof(int) * s);
    pin((void **)&p->h_xs, sizeof(double) * 2 * SP);
    if (!ok) {
        free_pso(p);
        delete p;
        return fail(FB_E_NOMEM, "fb_pso_create: allocation failed");
    }
    *out = p;
    return FB_OK;
}

int fb_pso_destroy(fb_pso *p)
{
    if (!p) return FB_OK;
    (void)hipStreamSynchronize(p->h->stream);
    free_pso(p);
    delete p;
    return FB_OK;
}

int fb_pso_init(fb_pso *p, const double *x0, const double *v0, double *cost, int *info)
{
    if (!p || !x0 || !v0) return fail(FB_E_ARG, "fb_pso_init: null argument");
    fb_handle *h = p->h;
    const int S = p->st.S, P = p->st.P;
    const size_t SP = (size_t)S * P;
    hipStream_t st = h->stream;
    std::memcpy(p->h_xs, x0, sizeof(double) * SP);
    std::memcpy(p->h_xs + SP, v0, sizeof(double) * SP);
    FB_HIP(hipMemcpyAsync(p->st.x, p->h_xs, sizeof(double) * SP, hipMemcpyHostToDevice, st));
    FB_HIP(hipMemcpyAsync(p->st.v, p->h_xs + SP, sizeof(double) * SP, hipMemcpyHostToDevice, st));
    int rc = pso_prepare(p);
    if (rc) return rc;
    // initial swarm (pso.f90:291-308): evaluated once; pbest = x0, pbest
    // cost = cost; gbest from +Inf in particle order
    pso_enqueue_eval(p, st);
    FB_HIP(hipMemcpyAsync(p->st.pbest, p->st.x, sizeof(double) * SP, hipMemcpyDeviceToDevice, st));
    FB_HIP(hipMemcpyAsync(p->st.pcost, p->st.cost, sizeof(double) * S, hipMemcpyDeviceToDevice, st));
    p->h_state[0] = std::numeric_limits<double>::infinity();
    p->h_state[1] = 0.0;
    FB_HIP(hipMemcpyAsync(p->st.state, p->h_state, sizeof(double) * 2, hipMemcpyHostToDevice, st));
    rc = pso_finish(p, cost, info);
    if (rc) return rc;
    // a lock-step re-evaluation (if any) left fresh costs: seed pcost again
    FB_HIP(hipMemcpyAsync(p->st.pcost, p->st.cost, sizeof(double) * S, hipMemcpyDeviceToDevice, st));
    fb::launch_pso_select(p->st, st);
    FB_HIP(hipStreamSynchronize(st));
    FB_HIP(hipGetLastError());
    return FB_OK;
}

int fb_pso_step(fb_pso *p, const double *r1, const double *r2, double *cost, int *info)
{
    if (!p || !r1 || !r2) return fail(FB_E_ARG, "fb_pso_step: null argument");
    fb_handle *h = p->h;
    const int S = p->st.S, P = p->st.P;
    const size_t SP = (size_t)S * P;
    hipStream_t st = h->stream;
    int rc = pso_prepare(p);
    if (rc) return rc;
    std::memcpy(p->h_r, r1, sizeof(double) * SP);
    std::memcpy(p->h_r + SP, r2, sizeof(double) * SP);
    std::memcpy(p->h_r + 2 * SP, p->coef, sizeof(double) * 3);
    auto enqueue = [&](hipStream_t s) {
        (void)hipMemcpyAsync(p->d_r, p->h_r, sizeof(double) * (2 * SP + 3), hipMemcpyHostToDevice, s);
        fb::launch_pso_select(p->st, s);
        fb::launch_pso_update(p->st, p->d_r, p->d_r + SP, p->d_r + 2 * SP, s);
        pso_enqueue_eval(p, s);
    };
    if (h->profile || !h->use_graph) {
        enqueue(st);
    } else {
        const int mode = h->mode;
        if (p->exec && (p->graph_epoch != h->graph_epoch || p->graph_mode != mode)) pso_drop_graph(p);
        if (!p->exec) {
            FB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            enqueue(st);
            FB_HIP(hipStreamEndCapture(st, &p->graph));
            FB_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
            p->graph_epoch = h->graph_epoch;
            p->graph_mode = mode;
        }
        FB_HIP(hipGraphLaunch(p->exec, st));
    }
    rc = pso_finish(p, cost, info);
    if (rc) return rc;
    return h->profile ? drain_profile(h) : FB_OK;
}

int fb_pso_retry(fb_pso *p, int k, const int *idx, const double *x, const double *v, double *cost, int *info)
{
    if (!p || (k > 0 && (!idx || !x || !v))) return fail(FB_E_ARG, "fb_pso_retry: null argument");
    fb_handle *h = p->h;
    const int S = p->st.S, P = p->st.P;
    if (k < 0 || k > S) return fail(FB_E_CAPACITY, "fb_pso_retry: particle count out of range");
    if (k == 0) return FB_OK;
    for (int r = 0; r < k; ++r)
        if (idx[r] < 0 || idx[r] >= S) return fail(FB_E_ARG, "fb_pso_retry: particle index out of range");
    hipStream_t st = h->stream;
    const size_t kP = (size_t)k * P;
    std::memcpy(p->h_idx, idx, sizeof(int) * k);
    std::memcpy(p->h_xs, x, sizeof(double) * kP);
    std::memcpy(p->h_xs + kP, v, sizeof(double) * kP);
    FB_HIP(hipMemcpyAsync(p->d_idx, p->h_idx, sizeof(int) * k, hipMemcpyHostToDevice, st));
    FB_HIP(hipMemcpyAsync(p->d_xs, p->h_xs, sizeof(double) * 2 * kP, hipMemcpyHostToDevice, st));
    fb::launch_pso_scatter(p->st, k, p->d_idx, p->d_xs, p->d_xs + kP, st);
    // the listed particles as a swarm of k on the ordinary batched path
    fb::launch_pso_gather(p->st, k, p->d_idx, h->d_theta, st);
    choose_path_untimed(h, k);
    int rc = FB_OK;
    if (h->use_dag) {
        rc = dag_plan_for(h, k);
        if (rc) return rc;
    }
    int err = 0;
    rc = run_swarm(h, k, &err);
    if (rc) return rc;
    if (err == 1) return fail(FB_E_HIP, "fb_pso_retry: the persistent factorisation found no ready task for 2 s");
    if (err == 2) {
        ++h->dispatch_fallbacks;
        h->forced_mode = 1;
        set_mode(h, 1);
        drop_graphs(h);
        rc = run_swarm(h, k, &err);
        if (rc) return rc;
    }
    fb::launch_pso_put_cost(p->st, k, p->d_idx, h->d_ll, h->d_info, st);
    FB_HIP(hipGetLastError());
    for (int r = 0; r < k; ++r) {
        if (cost) cost[r] = h->h_info[r] == 0 ? -h->h_ll[r] : std::numeric_limits<double>::infinity();
        if (info) info[r] = h->h_info[r];
    }
    return h->profile ? drain_profile(h) : FB_OK;
}

int fb_pso_set_coefficients(fb_pso *p, double w, double c1, double c2)
{
    if (!p) return fail(FB_E_ARG, "fb_pso_set_coefficients: null argument");
    p->coef[0] = w; p->coef[1] = c1; p->coef[2] = c2;
    return FB_OK;
}

int fb_pso_select(fb_pso *p, double *gcost, double *gbest, int *gidx)
{
    if (!p) return fail(FB_E_ARG, "fb_pso_select: null argument");
    hipStream_t st = p->h->stream;
    const int P = p->st.P;
    fb::launch_pso_select(p->st, st);
    FB_HIP(hipMemcpyAsync(p->h_state, p->st.state, sizeof(double), hipMemcpyDeviceToHost, st));
    FB_HIP(hipMemcpyAsync(p->h_state + 2, p->st.gbest, sizeof(double) * P, hipMemcpyDeviceToHost, st));
    FB_HIP(hipMemcpyAsync(p->h_gidx, p->st.gidx, sizeof(int), hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    FB_HIP(hipGetLastError());
    if (gcost) *gcost = p->h_state[0];
    if (gbest) std::memcpy(gbest, p->h_state + 2, sizeof(double) * P);
    if (gidx) *gidx = *p->h_gidx;
    return FB_OK;
}

int fb_pso_set_gbest(fb_pso *p, const double *gbest, double gcost)
{
    if (!p || !gbest) return fail(FB_E_ARG, "fb_pso_set_gbest: null argument");
    hipStream_t st = p->h->stream;
    const int P = p->st.P;
    fb::launch_pso_select(p->st, st);   // pending costs are selected first
    p->h_state[0] = gcost;
    std::memcpy(p->h_state + 2, gbest, sizeof(double) * P);
    FB_HIP(hipMemcpyAsync(p->st.state, p->h_state, sizeof(double), hipMemcpyHostToDevice, st));
    FB_HIP(hipMemcpyAsync(p->st.gbest, p->h_state + 2, sizeof(double) * P, hipMemcpyHostToDevice, st));
    FB_HIP(hipStreamSynchronize(st));
    FB_HIP(hipGetLastError());
    return FB_OK;
}

int fb_pso_positions(fb_pso *p, double *x, double *v)
{
    if (!p || (!x && !v)) return fail(FB_E_ARG, "fb_pso_positions: null argument");
    hipStream_t st = p->h->stream;
    const size_t SP = (size_t)p->st.S * p->st.P;
    if (x) FB_HIP(hipMemcpyAsync(p->h_xs, p->st.x, sizeof(double) * SP, hipMemcpyDeviceToHost, st));
    if (v) FB_HIP(hipMemcpyAsync(p->h_xs + SP, p->st.v, sizeof(double) * SP, hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    if (x) std::memcpy(x, p->h_xs, sizeof(double) * SP);
    if (v) std::memcpy(v, p->h_xs + SP, sizeof(double) * SP);
    return FB_OK;
}

int fb_multi_pso_step(fb_multi *m, fb_pso *const *ps, const double *const *r1, const double *const *r2,
                      double *const *cost, int *const *info)
{
    if (!m || !ps || !r1 || !r2) return fail(FB_E_ARG, "fb_multi_pso_step: null argument");
    const int G = (int)m->gps.size();
    std::vector<int> sv(G, 0);
    for (int g = 0; g < G; ++g) {
        if (!ps[g]) continue;
        if (ps[g]->h != m->gps[g])
            return fail(FB_E_ARG, "fb_multi_pso_step: swarm " + std::to_string(g) + " is not on handle " +
                                      std::to_string(g) + " of the set");
        if (!r1[g] || !r2[g]) return fail(FB_E_ARG, "fb_multi_pso_step: null argument");
        sv[g] = ps[g]->st.S;
    }
    if (std::all_of(sv.begin(), sv.end(), [](int v) { return v == 0; })) return FB_OK;
    hipStream_t st = m->gps[0]->stream;
    int rc = FB_OK;
    if (!m->schedule.count(sv)) {
        // chosen on the swarms' current positions, staged where the
        // evaluation call stages its particles
        for (int g = 0; g < G; ++g)
            if (sv[g] > 0)
                FB_HIP(hipMemcpyAsync(m->gps[g]->d_theta, ps[g]->st.x, sizeof(double) * sv[g] * ps[g]->st.P,
                                      hipMemcpyDeviceToDevice, st));
        rc = multi_choose(m, sv);
        if (rc) return rc;
    }
    // the persistent schedule reports its errors after the fact: steps use
    // the merged (or sequential) schedule instead
    int mode = m->schedule.at(sv);
    if (mode == kMultiPersistent) mode = merged_possible(m, sv) ? kMultiMerged : kMultiSequential;
    std::vector<long long> key{mode};
    std::vector<const double *> th(G, nullptr);
    for (int g = 0; g < G; ++g) {
        key.push_back(ps[g] ? ps[g]->id : 0);
        key.push_back(m->gps[g]->graph_epoch);
        if (!ps[g]) continue;
        fb_pso *p = ps[g];
        const size_t SP = (size_t)p->st.S * p->st.P;
        std::memcpy(p->h_r, r1[g], sizeof(double) * SP);
        std::memcpy(p->h_r + SP, r2[g], sizeof(double) * SP);
        std::memcpy(p->h_r + 2 * SP, p->coef, sizeof(double) * 3);
        th[g] = p->st.x;
    }
    auto enqueue = [&](hipStream_t s) {
        for (int g = 0; g < G; ++g) {
            if (!ps[g]) continue;
            fb_pso *p = ps[g];
            const size_t SP = (size_t)p->st.S * p->st.P;
            (void)hipMemcpyAsync(p->d_r, p->h_r, sizeof(double) * (2 * SP + 3), hipMemcpyHostToDevice, s);
            fb::launch_pso_select(p->st, s);
            fb::launch_pso_update(p->st, p->d_r, p->d_r + SP, p->d_r + 2 * SP, s);
        }
        multi_enqueue(m, sv, th.data(), mode, s);
        for (int g = 0; g < G; ++g)
            if (ps[g]) fb::launch_pso_cost(ps[g]->st, m->gps[g]->d_ll, m->gps[g]->d_info, s);
    };
    auto it = m->pso_graphs.find(key);
    if (it == m->pso_graphs.end()) {
        hipGraph_t gr = nullptr;
        hipGraphExec_t ge = nullptr;
        FB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue(st);
        FB_HIP(hipStreamEndCapture(st, &gr));
        if (hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(gr);
            return fail(FB_E_HIP, "fb_multi_pso_step: graph instantiation failed");
        }
        it = m->pso_graphs.emplace(key, std::make_pair(gr, ge)).first;
    }
    FB_HIP(hipGraphLaunch(it->second.second, st));
    for (int g = 0; g < G; ++g) {
        if (!ps[g]) continue;
        fb_pso *p = ps[g];
        FB_HIP(hipMemcpyAsync(p->h_cost, p->st.cost, sizeof(double) * sv[g], hipMemcpyDeviceToHost, st));
        FB_HIP(hipMemcpyAsync(p->h_info, p->st.info, sizeof(int) * sv[g], hipMemcpyDeviceToHost, st));
    }
    FB_HIP(hipStreamSynchronize(st));
    FB_HIP(hipGetLastError());
    for (int g = 0; g < G; ++g) {
        if (!ps[g]) continue;
        rc = check_waits(ps[g]->h_info, sv[g], "fb_multi_pso_step");
        if (rc) return rc;
        if (cost && cost[g]) std::memcpy(cost[g], ps[g]->h_cost, sizeof(double) * sv[g]);
        if (info && info[g]) std::memcpy(info[g], ps[g]->h_info, sizeof(int) * sv[g]);
    }
    return FB_OK;
}

int fb_factor_groups(const fb_handle *h, int s, int *groups)
{
    if (!h || !groups) return fail(FB_E_ARG, "fb_factor_groups: null argument");
    *groups = s >= 1 ? groups_for(h, s) : 1;
    return FB_OK;
}

int fb_factor_path(const fb_handle *h, int s, int *path)
{
    if (!h || !path) return fail(FB_E_ARG, "fb_factor_path: null argument");
    if (h->forced_mode >= 0) { *path = h->forced_mode; return FB_OK; }
    auto it = h->tuned.find(s);
    *path = it != h->tuned.end() ? it->second : -1;
    return FB_OK;
}

int fb_profile_enable(fb_handle *h, int enable)
{
    if (!h) return fail(FB_E_ARG, "fb_profile_enable: null handle");
    int rc = drain_profile(h);
    h->profile = enable == 2 ? 2 : enable != 0 ? 1 : 0;
    return rc;
}

int fb_profile_read(fb_handle *h, int kind, double *total_ms, long long *launches)
{
    if (!h || kind < 0 || kind > 9 || !total_ms || !launches) return fail(FB_E_ARG, "fb_profile_read: bad argument");
    if (kind == 7 || kind == 9) {   // counters, not reset
        *total_ms = 0.0;
        *launches = kind == 7 ? h->tune_evals : h->dispatch_fallbacks;
        return FB_OK;
    }
    int rc = drain_profile(h);
    *total_ms = h->prof_ms[kind];
    *launches = h->prof_n[kind];
    h->prof_ms[kind] = 0.0;
    h->prof_n[kind] = 0;
    return rc;
}

}  // extern "C"
