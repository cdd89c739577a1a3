// Batched frames: many independent simulations of any sizes stepped as one frame.
//
// A frame enqueues every simulation's step (advection, forces, MGPCG projection, or the learned
// flexible CG with its MGPCG fallback for a simulation with a coarse-correction network) on one of
// a few HIP streams forked from the caller's stream and joined back into it, either directly or as
// one captured hipGraph that is replayed frame after frame.  Every projection is sequenced on the
// device with a PCG iteration budget per simulation (the kernels after the device's stop test are
// no-ops, the device stops by itself at the budget), so a frame never waits on the host.  Each
// simulation's last kernel is followed by a timestamp (the constant-rate wall clock of the GPU),
// which gives per-simulation completion times inside the frame.  After a frame the device gathers
// every simulation's PCG state into a pinned buffer; vfx_batch_wait reads it.  A budget starts at
// min(P_MAXIT, 32); after each waited frame it is at least the largest iteration count of the
// simulation's last 8 frames plus 1, and a simulation that stopped at its budget unconverged gets
// twice the budget for the next frame (a restart run that stopped at its budget: twice its own).
// Budgets grow at once; once 8 frames have been seen they also shrink to that count when they sit 2 or more above it (iterations past the stop are
// launched and return at once, so a loose budget costs launches).
// The autotuner picks, per grid size, the cells per thread of the cell kernels and the size from
// which the multigrid coarse levels run in one workgroup, and the number of streams, by timing
// them on snapshots of the simulations' own states.
//
// FP32 build only.  No reference counterpart: the reference steps one Solver3D per call
// (vfx/sim3d/solver.py:146-155) and captures one solver's step as a CUDA graph (solver.py:388-459).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "ctx.hpp"

#if !VFX_F64

namespace vfx {

constexpr int BATCH_MAX = 4096;
constexpr int BATCH_MAX_STREAMS = 16;
constexpr int BUDGET_MARGIN = 1;   // iterations kept beyond the largest count seen
constexpr int BUDGET_FIRST = 32;   // first budget of a simulation (capped by its P_MAXIT)
constexpr int BUDGET_SHRINK = 2;   // a budget shrinks when it is this far above the window's need
constexpr int FRAME_STATS = 9;     // per simulation: iterations, stop_at, r.r, b.b, max|r|, threshold,
                                   // true residual, budget-stop bits, first run's iterations

// parameters of every simulation, staged in one device buffer, to each context's own slot
__global__ void __launch_bounds__(256) k_scatter_prm(const float* __restrict__ src, float* const* __restrict__ dst) {
    const float* s = src + (size_t)blockIdx.x * P_SIZE;
    float* d = dst[blockIdx.x];
    for (int i = threadIdx.x; i < P_SIZE; i += blockDim.x) d[i] = s[i];
}

__global__ void k_stamp(unsigned long long* ts, int i) {
    if (threadIdx.x == 0) ts[i] = wall_clock64();
}

__global__ void __launch_bounds__(256) k_gather_state(DevState* const* __restrict__ st, int n,
                                                      double* __restrict__ out) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const DevState* s = st[i];
        double* o = out + (size_t)i * FRAME_STATS;
        o[0] = (double)(s->it_total + s->it);
        o[1] = (double)s->stop_at;
        o[2] = s->s[S_RR];
        o[3] = s->s[S_BB];
        o[4] = s->s[S_MAXR];
        o[5] = s->s[S_THR];
        o[6] = s->s[S_TRUE_RR];
        o[7] = (double)s->hit;
        o[8] = (double)(s->restarts > 0 ? s->it_total : s->it);
    }
}

}  // namespace vfx

struct vfx_batch {
    int n = 0;
    std::vector<vfx_ctx*> sims;
    int nstreams = 0;
    std::vector<hipStream_t> streams;
    std::vector<int> stream_of;   // simulation -> stream
    std::vector<int> order;       // enqueue order: largest grids first
    hipStream_t cap = nullptr;   // capture stream (the caller's may be the null stream, which cannot capture)
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<hipEvent_t> join;
    // device: staged parameters, tables, timestamps, gathered PCG state
    void* dpool = nullptr;
    float* dprm = nullptr;
    float** dprm_dst = nullptr;
    DevState** dstate = nullptr;
    unsigned long long* ts = nullptr;   // n completion stamps, then the frame start
    double* stats = nullptr;
    // pinned host
    float* pin_prm[2] = {nullptr, nullptr};
    hipEvent_t pin_ev[2] = {nullptr, nullptr};
    bool pin_used[2] = {false, false};
    int pin_i = 0;
    unsigned long long* pin_ts = nullptr;
    double* pin_stats = nullptr;
    // per simulation
    std::vector<int> budget;       // 0 until the first frame sets it from P_MAXIT
    std::vector<BudgetWindow> bwin;   // iteration counts of the last frames
    std::vector<float> user_maxit, rel_tol;
    std::vector<char> judged;      // the frame computes the simulation's true residual
    bool absorbed = true;          // the last frame's gathered state has been read
    long long n_hit = 0;           // simulation-frames that stopped at their budget unconverged
    double ticks_per_ms = 0.0;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<float*> g_data;
    std::vector<const float*> g_solid;
    std::vector<char> g_warm;   // warm-start flags the capture was made with (they pick kernel variants)
    std::vector<const vfx_net*> g_net;   // each simulation's learned preconditioner at capture
    bool dirty = true;
    bool frames = false;   // a frame has been enqueued
};

namespace {

void drop_graph(vfx_batch* b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->dirty = true;
}

void destroy_streams(vfx_batch* b) {
    for (hipStream_t s : b->streams) hipStreamDestroy(s);
    for (hipEvent_t e : b->join) hipEventDestroy(e);
    b->streams.clear();
    b->join.clear();
}

// ns streams (default priority); simulations assigned largest first to the least-loaded stream
int make_streams(vfx_batch* b, int ns) {
    if (ns < 1 || ns > BATCH_MAX_STREAMS) return arg_error("stream count must be 1..16");
    if (ns == b->nstreams && (int)b->streams.size() == ns) return VFX_OK;
    hipDeviceSynchronize();
    destroy_streams(b);
    drop_graph(b);
    for (int j = 0; j < ns; ++j) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        if (!hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "batch stream")) return VFX_E_HIP;
        b->streams.push_back(s);
        if (!hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "batch event")) return VFX_E_HIP;
        b->join.push_back(e);
    }
    b->nstreams = ns;
    std::vector<double> load(ns, 0.0);
    b->stream_of.assign(b->n, 0);
    for (int i : b->order) {
        const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->stream_of[i] = j;
        load[j] += (double)b->sims[i]->g.n3;
    }
    return VFX_OK;
}

// Every kernel of one frame on `main` and the batch streams (also what a capture records).
int enqueue_frame(vfx_batch* b, float* const* data, const float* const* solid, hipStream_t main) {
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, main, b->ts, b->n);
    hipLaunchKernelGGL(k_scatter_prm, dim3(b->n), dim3(256), 0, main, b->dprm, b->dprm_dst);
    if (!launch_ok("frame start")) return VFX_E_HIP;
    if (!hip_ok(hipEventRecord(b->fork, main), "frame fork")) return VFX_E_HIP;
    for (hipStream_t s : b->streams)
        if (!hip_ok(hipStreamWaitEvent(s, b->fork, 0), "frame fork")) return VFX_E_HIP;
    for (int i : b->order) {
        vfx_ctx* c = b->sims[i];
        hipStream_t s = b->streams[b->stream_of[i]];
        const int own = c->cg_budget;   // the context's own setting applies again outside frames
        c->cg_budget = b->budget[i];
        int rc = stage_advect(c, data[i], solid[i], s);
        if (!rc) rc = stage_forces(c, data[i], s);
        if (!rc) rc = stage_project(c, data[i], s, true);
        c->cg_budget = own;
        if (rc) return rc;
        hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, b->ts, i);
    }
    for (int j = 0; j < b->nstreams; ++j) {
        if (!hip_ok(hipEventRecord(b->join[j], b->streams[j]), "frame join")) return VFX_E_HIP;
        if (!hip_ok(hipStreamWaitEvent(main, b->join[j], 0), "frame join")) return VFX_E_HIP;
    }
    hipLaunchKernelGGL(k_gather_state, dim3(1), dim3(256), 0, main, b->dstate, b->n, b->stats);
    return launch_ok("frame end") ? VFX_OK : VFX_E_HIP;
}

// stage the frame's parameters: pinned slot -> device staging buffer, on `main`
int stage_params(vfx_batch* b, const float* prm_host, hipStream_t main) {
    const int k = b->pin_i;
    b->pin_i ^= 1;
    if (b->pin_used[k] && !hip_ok(hipEventSynchronize(b->pin_ev[k]), "parameter slot wait")) return VFX_E_HIP;
    float* slot = b->pin_prm[k];
    std::memcpy(slot, prm_host, sizeof(float) * P_SIZE * b->n);
    for (int i = 0; i < b->n; ++i) {
        float* p = slot + (size_t)i * P_SIZE;
        b->user_maxit[i] = p[P_MAXIT];
        b->rel_tol[i] = p[P_REL_TOL];
        if (b->budget[i] == 0) {
            b->budget[i] = std::max(1, std::min((int)p[P_MAXIT], BUDGET_FIRST));
            b->dirty = true;
        }
        std::memcpy(b->sims[i]->prm_last, p, sizeof(float) * P_SIZE);
    }
    if (!hip_ok(hipMemcpyAsync(b->dprm, slot, sizeof(float) * P_SIZE * b->n, hipMemcpyHostToDevice, main),
                "parameter upload"))
        return VFX_E_HIP;
    if (!hip_ok(hipEventRecord(b->pin_ev[k], main), "parameter slot event")) return VFX_E_HIP;
    b->pin_used[k] = true;
    return VFX_OK;
}

int capture(vfx_batch* b, float* const* data, const float* const* solid) {
    drop_graph(b);
    if (!hip_ok(hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed), "frame capture")) return VFX_E_HIP;
    int rc = enqueue_frame(b, data, solid, b->cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (!hip_ok(e, "frame capture")) return VFX_E_HIP;
    b->graph = g;
    if (!hip_ok(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "frame instantiate")) {
        drop_graph(b);
        return VFX_E_HIP;
    }
    b->g_data.assign(data, data + b->n);
    b->g_solid.assign(solid, solid + b->n);
    b->g_warm.resize(b->n);
    b->g_net.resize(b->n);
    for (int i = 0; i < b->n; ++i) {
        b->g_warm[i] = b->sims[i]->prm_last[P_WARM] > 0.5f;
        b->g_net[i] = b->sims[i]->net;
    }
    b->dirty = false;
    return VFX_OK;
}

int frame(vfx_batch* b, float* const* data, const float* const* solid, const float* prm_host, int use_graph,
          hipStream_t main) {
    if (int rc = stage_params(b, prm_host, main)) return rc;
    bool setup_pending = false;
    for (vfx_ctx* c : b->sims) setup_pending = setup_pending || c->P.force_setup;
    if (use_graph && !setup_pending) {
        // a forced operator rebuild is a one-off: it runs eagerly and stays out of the graph
        bool same = !b->dirty && b->exec && std::equal(data, data + b->n, b->g_data.begin()) &&
                    std::equal(solid, solid + b->n, b->g_solid.begin());
        for (int i = 0; same && i < b->n; ++i)
            same = b->g_warm[i] == (b->sims[i]->prm_last[P_WARM] > 0.5f) && b->g_net[i] == b->sims[i]->net;
        if (!same)
            if (int rc = capture(b, data, solid)) return rc;
        if (!hip_ok(hipGraphLaunch(b->exec, main), "frame launch")) return VFX_E_HIP;
    } else {
        if (int rc = enqueue_frame(b, data, solid, main)) return rc;
    }
    if (!hip_ok(hipMemcpyAsync(b->pin_ts, b->ts, sizeof(unsigned long long) * (b->n + 1), hipMemcpyDeviceToHost,
                               main),
                "stamp read"))
        return VFX_E_HIP;
    if (!hip_ok(hipMemcpyAsync(b->pin_stats, b->stats, sizeof(double) * FRAME_STATS * b->n, hipMemcpyDeviceToHost,
                               main),
                "state read"))
        return VFX_E_HIP;
    for (vfx_ctx* c : b->sims) {
        c->last_iters = 0;
        c->last_restarts = 0;
        if (!hip_ok(hipEventRecord(c->done_ev, main), "done event")) return VFX_E_HIP;
    }
    if (!hip_ok(hipEventRecord(b->done, main), "frame event")) return VFX_E_HIP;
    for (int i = 0; i < b->n; ++i) b->judged[i] = b->sims[i]->last_true_valid ? 1 : 0;
    b->frames = true;
    b->absorbed = false;
    return VFX_OK;
}

// converged: the true residual judges when the frame computed it
bool sim_converged(const vfx_batch* b, int i) {
    const double* o = b->pin_stats + (size_t)i * FRAME_STATS;
    const double rt = (double)b->rel_tol[i];
    if (b->judged[i] && rt > 0.0) return std::sqrt(o[6]) <= rt * std::sqrt(o[3]);
    return rt > 0.0 ? std::sqrt(o[2]) <= rt * std::sqrt(o[3]) : o[4] <= o[5];
}

// the finished frame's gathered state -> the next budgets (a budget stop doubles the budget)
void absorb_frame(vfx_batch* b) {
    if (b->absorbed) return;
    b->absorbed = true;
    for (int i = 0; i < b->n; ++i) {
        const double* o = b->pin_stats + (size_t)i * FRAME_STATS;
        const int hit = (int)o[7];
        const int first = (int)o[8];
        if (hit) b->n_hit += 1;
        const int cap = std::max(1, (int)b->user_maxit[i]);
        BudgetWindow& w = b->bwin[i];
        w.push(first);
        int want = w.max() + BUDGET_MARGIN;
        if (hit & 1) want = std::max(want, 2 * b->budget[i]);
        want = std::max(1, std::min(want, cap));
        // grow at once; shrink once the window is full (the first frames of a flow from rest
        // keep the start budget while their counts climb)
        if (want > b->budget[i] || (w.n == BUDGET_WINDOW && want + BUDGET_SHRINK <= b->budget[i])) {
            b->budget[i] = want;
            b->dirty = true;
        }
        // the restart run stopped at its budget: the context's restart budget doubles (after a
        // learned run it is the MGPCG fallback, a whole solve: at least the first run's budget)
        if (hit & 2) {
            vfx_ctx* c = b->sims[i];
            const int br = cg_budget_restart(c, cap);
            int nb = 2 * br;
            if (c->net) nb = std::max(nb, std::max(b->budget[i], 16));
            c->bw_r.budget = std::min(nb, cap);
            b->dirty = true;
        }
    }
}

// wait for the last frame; per-simulation results; grow the budgets that were too small
int wait_frame(vfx_batch* b, float* latency_ms, int* iters, int* converged, double* rel_res, float* frame_ms) {
    if (!b->frames) return arg_error("no frame was enqueued");
    if (!hip_ok(hipEventSynchronize(b->done), "frame wait")) return VFX_E_HIP;
    absorb_frame(b);
    const unsigned long long t0 = b->pin_ts[b->n];
    unsigned long long tmax = t0;
    for (int i = 0; i < b->n; ++i) {
        const unsigned long long te = b->pin_ts[i];
        tmax = std::max(tmax, te);
        if (latency_ms) latency_ms[i] = (float)((double)(te - t0) / b->ticks_per_ms);
        const double* o = b->pin_stats + (size_t)i * FRAME_STATS;
        if (iters) iters[i] = (int)o[0];
        if (converged) converged[i] = sim_converged(b, i) ? 1 : 0;
        if (rel_res) rel_res[i] = o[3] > 0.0 ? std::sqrt((b->judged[i] ? o[6] : o[2]) / o[3]) : 0.0;
    }
    if (frame_ms) *frame_ms = (float)((double)(tmax - t0) / b->ticks_per_ms);
    return VFX_OK;
}

// device snapshot of every simulation's data and pressure (the autotuner restores it)
struct Snapshot {
    float* buf = nullptr;
    std::vector<size_t> off;
    int save(vfx_batch* b, float* const* data, hipStream_t s) {
        size_t tot = 0;
        for (vfx_ctx* c : b->sims) {
            off.push_back(tot);
            tot += 7 * c->g.n3;
        }
        if (!hip_ok(hipMalloc(&buf, tot * sizeof(float)), "tuning snapshot")) return VFX_E_HIP;
        for (int i = 0; i < b->n; ++i) {
            const size_t n3 = b->sims[i]->g.n3;
            if (!hip_ok(hipMemcpyAsync(buf + off[i], data[i], 6 * n3 * sizeof(float), hipMemcpyDeviceToDevice, s),
                        "snapshot") ||
                !hip_ok(hipMemcpyAsync(buf + off[i] + 6 * n3, b->sims[i]->P.phi, n3 * sizeof(float),
                                       hipMemcpyDeviceToDevice, s),
                        "snapshot"))
                return VFX_E_HIP;
        }
        return VFX_OK;
    }
    int restore(vfx_batch* b, int i, float* const* data, hipStream_t s) const {
        const size_t n3 = b->sims[i]->g.n3;
        if (!hip_ok(hipMemcpyAsync(data[i], buf + off[i], 6 * n3 * sizeof(float), hipMemcpyDeviceToDevice, s),
                    "restore") ||
            !hip_ok(hipMemcpyAsync(b->sims[i]->P.phi, buf + off[i] + 6 * n3, n3 * sizeof(float),
                                   hipMemcpyDeviceToDevice, s),
                    "restore"))
            return VFX_E_HIP;
        return VFX_OK;
    }
    ~Snapshot() {
        if (buf) {
            hipDeviceSynchronize();
            hipFree(buf);
        }
    }
};

float median(std::vector<float> v) {
    std::sort(v.begin(), v.end());
    return v.empty() ? 0.0f : v[v.size() / 2];
}

// ms of one step of simulation i alone on stream s (reps timed runs from the snapshot)
int time_sim(vfx_batch* b, int i, float* const* data, const float* const* solid, const float* prm_host,
             const Snapshot& snap, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1, float* ms) {
    vfx_ctx* c = b->sims[i];
    float p[P_SIZE];
    std::memcpy(p, prm_host + (size_t)i * P_SIZE, sizeof(p));
    std::memcpy(c->prm_last, p, sizeof(p));
    if (int rc = upload(c, p, c->prm, s)) return rc;
    std::vector<float> t;
    for (int r = 0; r <= reps; ++r) {
        if (int rc = snap.restore(b, i, data, s)) return rc;
        if (!hip_ok(hipEventRecord(e0, s), "tune event")) return VFX_E_HIP;
        const int own = c->cg_budget;
        c->cg_budget = b->budget[i];
        int rc = stage_advect(c, data[i], solid[i], s);
        if (!rc) rc = stage_forces(c, data[i], s);
        if (!rc) rc = stage_project(c, data[i], s, true);
        c->cg_budget = own;
        if (rc) return rc;
        if (!hip_ok(hipEventRecord(e1, s), "tune event") || !hip_ok(hipEventSynchronize(e1), "tune wait"))
            return VFX_E_HIP;
        float v = 0.0f;
        if (!hip_ok(hipEventElapsedTime(&v, e0, e1), "tune time")) return VFX_E_HIP;
        if (r > 0) t.push_back(v);   // the first run warms the caches and the code objects
    }
    *ms = median(t);
    return VFX_OK;
}

}  // namespace

extern "C" {

int vfx_batch_create(vfx_ctx* const* ctxs, int n, int streams, vfx_batch** out) {
    if (out == nullptr || ctxs == nullptr) return arg_error("null pointer");
    *out = nullptr;
    if (n < 1 || n > BATCH_MAX) return arg_error("batch size must be 1..4096");
    for (int i = 0; i < n; ++i) {
        if (ctxs[i] == nullptr) return arg_error("null context in the batch");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return arg_error("a context appears twice in the batch");
    }
    vfx_batch* b = new (std::nothrow) vfx_batch();
    if (!b) return arg_error("out of host memory");
    b->n = n;
    b->sims.assign(ctxs, ctxs + n);
    b->order.resize(n);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](int a, int c) { return b->sims[a]->g.n3 > b->sims[c]->g.n3; });
    b->budget.assign(n, 0);
    b->bwin.assign(n, BudgetWindow{});
    b->user_maxit.assign(n, 0.0f);
    b->rel_tol.assign(n, 0.0f);
    b->judged.assign(n, 0);
    int dev = 0, khz = 0;
    bool ok = hip_ok(hipGetDevice(&dev), "device") &&
              hip_ok(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate");
    b->ticks_per_ms = khz > 0 ? (double)khz : 100000.0;
    // device: parameters, pointer tables, stamps, gathered state (each 256-byte aligned)
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_prm = 0, o_dst = up(o_prm + sizeof(float) * P_SIZE * n), o_st = up(o_dst + sizeof(void*) * n),
                 o_ts = up(o_st + sizeof(void*) * n), o_stats = up(o_ts + 8 * (size_t)(n + 1)),
                 total = up(o_stats + sizeof(double) * FRAME_STATS * n);
    ok = ok && hip_ok(hipMalloc(&b->dpool, total), "batch allocation");
    if (ok) {
        char* base = (char*)b->dpool;
        b->dprm = (float*)(base + o_prm);
        b->dprm_dst = (float**)(base + o_dst);
        b->dstate = (DevState**)(base + o_st);
        b->ts = (unsigned long long*)(base + o_ts);
        b->stats = (double*)(base + o_stats);
        std::vector<float*> pd(n);
        std::vector<DevState*> sd(n);
        for (int i = 0; i < n; ++i) {
            pd[i] = ctxs[i]->prm;
            sd[i] = ctxs[i]->st;
        }
        ok = hip_ok(hipMemset(b->dpool, 0, total), "batch clear") &&
             hip_ok(hipMemcpy(b->dprm_dst, pd.data(), sizeof(void*) * n, hipMemcpyHostToDevice), "batch tables") &&
             hip_ok(hipMemcpy(b->dstate, sd.data(), sizeof(void*) * n, hipMemcpyHostToDevice), "batch tables");
    }
    for (int k = 0; ok && k < 2; ++k)
        ok = hip_ok(hipHostMalloc((void**)&b->pin_prm[k], sizeof(float) * P_SIZE * n, hipHostMallocDefault),
                    "pinned allocation") &&
             hip_ok(hipEventCreateWithFlags(&b->pin_ev[k], hipEventDisableTiming), "event");
    ok = ok &&
         hip_ok(hipHostMalloc((void**)&b->pin_ts, 8 * (size_t)(n + 1), hipHostMallocDefault), "pinned allocation") &&
         hip_ok(hipHostMalloc((void**)&b->pin_stats, sizeof(double) * FRAME_STATS * n, hipHostMallocDefault),
                "pinned allocation") &&
         hip_ok(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking), "capture stream") &&
         hip_ok(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "event") &&
         hip_ok(hipEventCreateWithFlags(&b->done, hipEventDisableTiming), "event");
    if (ok) {
        if (int rc = make_streams(b, streams)) {
            vfx_batch_destroy(b);
            return rc;
        }
    }
    if (!ok) {
        vfx_batch_destroy(b);
        return VFX_E_HIP;
    }
    *out = b;
    return VFX_OK;
}

int vfx_batch_destroy(vfx_batch* b) {
    if (!b) return VFX_OK;
    hipDeviceSynchronize();
    drop_graph(b);
    destroy_streams(b);
    for (int k = 0; k < 2; ++k) {
        if (b->pin_ev[k]) hipEventDestroy(b->pin_ev[k]);
        if (b->pin_prm[k]) hipHostFree(b->pin_prm[k]);
    }
    if (b->pin_ts) hipHostFree(b->pin_ts);
    if (b->pin_stats) hipHostFree(b->pin_stats);
    if (b->cap) hipStreamDestroy(b->cap);
    if (b->fork) hipEventDestroy(b->fork);
    if (b->done) hipEventDestroy(b->done);
    if (b->dpool) hipFree(b->dpool);
    for (vfx_ctx* c : b->sims) c->cg_budget = 0;
    delete b;
    return VFX_OK;
}

int vfx_batch_set_budget(vfx_batch* b, int sim, int budget) {
    if (!b) return arg_error("null batch");
    if (budget < 1 || budget > 1000) return arg_error("budget must be 1..1000");
    b->absorbed = true;   // a frame enqueued before this call does not override the new budget
    for (int i = 0; i < b->n; ++i)
        if (sim < 0 || sim == i) {
            if (b->budget[i] != budget) b->dirty = true;
            b->budget[i] = budget;
        }
    return VFX_OK;
}

int vfx_batch_get_budget(vfx_batch* b, int* budget) {
    if (!b || !budget) return arg_error("null pointer");
    for (int i = 0; i < b->n; ++i) budget[i] = b->budget[i];
    return VFX_OK;
}

int vfx_batch_truncated(vfx_batch* b, long long* count) {
    if (!b || !count) return arg_error("null pointer");
    *count = b->n_hit;
    return VFX_OK;
}

int vfx_batch_step(vfx_batch* b, float* const* data, const float* const* solid, const float* prm_host, int graph,
                   void* stream) {
    if (!b || !data || !solid || !prm_host) return arg_error("null pointer");
    for (int i = 0; i < b->n; ++i) {
        if (!data[i] || !solid[i]) return arg_error("null field pointer");
        if (int rc = validate_prm(prm_host + (size_t)i * P_SIZE)) return rc;
        vfx_ctx* c = b->sims[i];
        if ((c->net && !learned_on_device(c)) || c->fcg_mode) {
            set_error("a batched frame runs MGPCG or the coarse-correction network only (detach the other "
                      "learned preconditioners)");
            return VFX_E_VALUE;
        }
    }
    return frame(b, data, solid, prm_host, graph, reinterpret_cast<hipStream_t>(stream));
}

int vfx_batch_wait(vfx_batch* b, float* latency_ms, int* iters, int* converged, double* rel_res, float* frame_ms) {
    if (!b) return arg_error("null batch");
    return wait_frame(b, latency_ms, iters, converged, rel_res, frame_ms);
}

int vfx_batch_tune(vfx_batch* b, float* const* data, const float* const* solid, const float* prm_host, int reps,
                   void* stream, double* report, int report_len) {
    if (!b || !data || !solid || !prm_host) return arg_error("null pointer");
    if (reps < 1 || reps > 50) return arg_error("reps must be 1..50");
    if (report && report_len < VFX_TUNE_REPORT) return arg_error("report needs VFX_TUNE_REPORT doubles");
    for (int i = 0; i < b->n; ++i) {
        if (!data[i] || !solid[i]) return arg_error("null field pointer");
        if (int rc = validate_prm(prm_host + (size_t)i * P_SIZE)) return rc;
        if ((b->sims[i]->net && !learned_on_device(b->sims[i])) || b->sims[i]->fcg_mode) {
            set_error("a batched frame runs MGPCG or the coarse-correction network only (detach the other "
                      "learned preconditioners)");
            return VFX_E_VALUE;
        }
    }
    hipStream_t main = reinterpret_cast<hipStream_t>(stream);
    std::vector<double> rep(VFX_TUNE_REPORT, 0.0);
    Snapshot snap;
    if (int rc = snap.save(b, data, main)) return rc;
    if (!hip_ok(hipStreamSynchronize(main), "snapshot")) return VFX_E_HIP;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!hip_ok(hipEventCreate(&e0), "event") || !hip_ok(hipEventCreate(&e1), "event")) return VFX_E_HIP;
    struct Ev {
        hipEvent_t a, b;
        ~Ev() {
            if (a) hipEventDestroy(a);
            if (b) hipEventDestroy(b);
        }
    } ev{e0, e1};
    // 1. per grid size: cells per thread x coarse size, timed on the size's first simulation
    std::vector<int> sizes;
    for (vfx_ctx* c : b->sims)
        if (std::find(sizes.begin(), sizes.end(), c->N) == sizes.end()) sizes.push_back(c->N);
    std::sort(sizes.begin(), sizes.end());
    if ((int)sizes.size() > VFX_TUNE_MAX_SIZES) return arg_error("the autotuner handles at most 8 grid sizes");
    rep[0] = (double)sizes.size();
    const int zcs[] = {1, 2, 4};
    const int cms[] = {4, 8, 16};
    for (size_t k = 0; k < sizes.size(); ++k) {
        int i0 = 0;
        while (b->sims[i0]->N != sizes[k]) ++i0;
        vfx_ctx* c = b->sims[i0];
        const int zc0 = c->g.zc, cm0 = c->P.coarse_m;
        float best = 0.0f, dflt = 0.0f;
        int bzc = zc0, bcm = cm0;
        for (int zc : zcs)
            for (int cm : cms) {
                if (int rc = vfx_set_option(c, VFX_OPT_CELLS_PER_THREAD, zc)) return rc;
                if (int rc = vfx_set_option(c, VFX_OPT_COARSE_M, cm)) return rc;
                float ms = 0.0f;
                if (int rc = time_sim(b, i0, data, solid, prm_host, snap, reps, b->streams[0], e0, e1, &ms)) return rc;
                if (zc == zc0 && cm == cm0) dflt = ms;
                if (best == 0.0f || ms < best) {
                    best = ms;
                    bzc = zc;
                    bcm = cm;
                }
            }
        for (vfx_ctx* s : b->sims)
            if (s->N == sizes[k]) {
                if (int rc = vfx_set_option(s, VFX_OPT_CELLS_PER_THREAD, bzc)) return rc;
                if (int rc = vfx_set_option(s, VFX_OPT_COARSE_M, bcm)) return rc;
            }
        if (int rc = snap.restore(b, i0, data, b->streams[0])) return rc;
        double* r = rep.data() + 1 + 5 * k;
        r[0] = sizes[k];
        r[1] = bzc;
        r[2] = bcm;
        r[3] = best;
        r[4] = dflt;
    }
    if (!hip_ok(hipDeviceSynchronize(), "tuning")) return VFX_E_HIP;
    b->dirty = true;   // launch shapes changed: the next graph frame captures again
    // 2. stream count, on whole captured frames
    const int nss[] = {1, 2, 4, 8};
    float best = 0.0f;
    int bns = b->nstreams;
    double* rs = rep.data() + 1 + 5 * VFX_TUNE_MAX_SIZES;
    for (int q = 0; q < 4; ++q) {
        if (int rc = make_streams(b, nss[q])) return rc;
        std::vector<float> t;
        for (int r = 0; r <= reps; ++r) {
            for (int i = 0; i < b->n; ++i)
                if (int rc = snap.restore(b, i, data, main)) return rc;
            if (int rc = frame(b, data, solid, prm_host, 1, main)) return rc;
            float fm = 0.0f;
            if (int rc = wait_frame(b, nullptr, nullptr, nullptr, nullptr, &fm)) return rc;
            if (r > 0) t.push_back(fm);
        }
        const float ms = median(t);
        rs[1 + q] = ms;
        if (best == 0.0f || ms < best) {
            best = ms;
            bns = nss[q];
        }
    }
    rs[0] = bns;
    if (int rc = make_streams(b, bns)) return rc;
    for (int i = 0; i < b->n; ++i)
        if (int rc = snap.restore(b, i, data, main)) return rc;
    if (!hip_ok(hipStreamSynchronize(main), "restore")) return VFX_E_HIP;
    b->dirty = true;
    if (report) std::memcpy(report, rep.data(), sizeof(double) * VFX_TUNE_REPORT);
    return VFX_OK;
}

}  // extern "C"

#endif  // !VFX_F64
