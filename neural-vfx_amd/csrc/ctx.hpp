// The per-simulation context of the library (vfx_ctx of include/vfxsim3d.h) and the step stages
// that the single-simulation entry points (capi.hip) and the batched frame (batch.hip) share.
#pragma once

#include <string>
#include <vector>

#include "../../include/vfxsim3d.h"
#include "pressure_ops.hpp"
#include "sim3d_kernels.hpp"
#if !VFX_F64
#include "unet.hpp"
struct vfx_net {
    vfx::Net net;
};
#endif

#if VFX_F64
#define vfx_ctx vfx_ctx_f64   // the context of the FP64 build (include/vfxsim3d.h)
#define vfx_batch vfx_batch_f64
#endif

using namespace vfx;

constexpr int HOST_RING = 16;  // pinned parameter slots: how many steps the host may run ahead (solver.py:331-359)
constexpr int FCG_BLOCKS = 1024;   // blocks of the flexible-CG reductions (4 partial arrays)
constexpr int STOP_RING = 4;   // pinned read-back slots of the CG stop flag (host-polled mode)
constexpr int READBACK_RING = HOST_RING;   // pinned DevState copies of finished projections
constexpr int BUDGET_WINDOW = 8;   // projections whose iteration counts set the next budget
constexpr int GRAPH_SLOTS = 4;     // captured step graphs kept per context (least recently used goes)

// projection sequencing (VFX_OPT_SYNC)
enum : int { SYNC_DEVICE = 0, SYNC_POLL = 1 };

// what the host decides while enqueueing a step: a captured step graph is replayed only while
// all of it is unchanged
struct StepSig {
    const void* data;
    const void* solid;
    const void* net;
    int v[16];
};

// per projection in flight: what the host enqueued (the read-back adapts the budgets from it)
struct ReadbackInfo {
    int device;     // enqueued with device-side budgets (not host-polled)
    int fused;      // the fused projection (budget = tail cycles)
    int budget;     // iterations of the first CG run (fused: tail cycles)
    int budget_r;   // iterations of the gated restart run (0: none)
    int cycles;     // fused: V-cycles that always run
    int judged;     // the true residual was computed
    float rel_tol;
    int learned;    // the learned flexible CG ran on the device (its restart run is the MGPCG fallback)
};

struct StepGraph {
    StepSig sig;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;     // recorded after each launch: the exec is destroyed only after it
    bool launched = false;
    unsigned long long used = 0;   // replay clock of the last use
    ReadbackInfo info{};           // what the captured enqueue recorded on the host
    bool true_valid = false;
};

// Iteration counts of recent projections -> the next budget (the largest recent count plus a
// margin, grown at once after a projection that ran out of it).
struct BudgetWindow {
    int w[BUDGET_WINDOW] = {};
    int n = 0, i = 0;
    int budget = 0;   // 0: no count seen yet
    void push(int its) {
        w[i] = its;
        i = (i + 1) % BUDGET_WINDOW;
        if (n < BUDGET_WINDOW) ++n;
    }
    int max() const {
        int m = 0;
        for (int k = 0; k < n; ++k) m = w[k] > m ? w[k] : m;
        return m;
    }
    void reset() { *this = BudgetWindow{}; }
};

// The finer scalar grid (Solver3D(res, scalar_res=S), hires.py; solver.py:217-264): the caller's
// fine fields and the library's fine scratch.
struct HiresGrid {
    int S = 0;                     // 0: off
    real* hi = nullptr;            // (3, S, S, S) density, temperature, fuel (caller's)
    const real* hsolid = nullptr;  // (S, S, S), 1 = obstacle (caller's)
    void* pool = nullptr;
    real *hat = nullptr, *adv = nullptr;   // 3 S^3 each
    int* qidx = nullptr;           // S^3 backward-trace base corners
    uint8_t* ffm = nullptr;        // S^3 fluid flags (hsolid <= 0.5), written by the first pass
    double* part = nullptr;        // mass partials
    Geom g;                        // fine launch geometry
    real a = RL(0.0), b = RL(0.0); // fine index -> coarse index: p_c = a p_f + b
    real ia = RL(0.0);             // S / res (exactly 1 / a when pow2)
    real dxf = RL(0.0);            // fine cell size
    real ratio = RL(0.0);          // fine cells per coarse cell, S / res
    int span = 0;                  // fine cells a coarse cell can overlap per axis
    int nd = 1;                    // diffusion substeps on the fine grid
    bool pow2 = false;             // S / res is a power of two (a = 2^-k)
};

struct vfx_ctx {
    int N = 0;
    Geom g;
    real inv_dx = RL(0.0), inv2dx = RL(0.0);
    // device buffers
    void* pool = nullptr;
    real *hat = nullptr, *adv = nullptr, *omega = nullptr, *omega_s = nullptr, *force = nullptr;
    // the advection's trace corners (N^3 ints) live in `force`, which is free until the forces stage
    int* qidx() const { return reinterpret_cast<int*>(force); }
    float* prm = nullptr;      // parameters of the step
    float* prm_adv = nullptr;  // parameters of the standalone advect()
    double* mass_part = nullptr;
    DevState* st = nullptr;
    Pressure P;
    // host side
    float* ring = nullptr;   // HOST_RING * P_SIZE pinned floats
    hipEvent_t ring_ev[HOST_RING] = {};
    bool ring_used[HOST_RING] = {};
    int ring_i = 0;
    float prm_last[P_SIZE];
    int* stop_pin = nullptr;  // STOP_RING pinned ints
    hipEvent_t stop_ev[STOP_RING] = {};
    hipEvent_t done_ev = nullptr;
    bool projected = false;
    // options
    bool opt_true_res = false;
    int opt_restarts = 0;
    int cg_budget = 0;   // > 0: fixed budget, exactly this many PCG iterations are enqueued
    int sync_mode = SYNC_DEVICE;
    bool use_graph = true;   // replay whole steps as captured hipGraphs (device mode)
    BudgetWindow bw;         // first CG run (fused: tail cycles)
    BudgetWindow bw_r;       // gated restart run
    // lazy read-back of finished projections (device mode): FIFO of pinned DevState slots
    DevState* rb_pin = nullptr;
    hipEvent_t rb_ev[READBACK_RING] = {};
    ReadbackInfo rb_info[READBACK_RING] = {};
    int rb_head = 0, rb_count = 0;
    ReadbackInfo last_info{};   // of the last enqueued projection
    long long n_seen = 0, n_hit = 0, n_unconverged = 0;   // read-back projections, budget stops, unconverged
    // captured step graphs
    hipStream_t cap = nullptr;
    StepGraph graphs[GRAPH_SLOTS];
    unsigned long long graph_clock = 0;
    int graph_captures = 0, graph_replays = 0;
    // pressure_mode "fused" (stationary multigrid, no CG): cycles, sweeps, smoother (0 Chebyshev, 1 Jacobi)
    bool fused = false;
    int fused_cycles = 2, fused_smooth = 4, fused_jacobi = 0;
#if !VFX_F64
    vfx_net* net = nullptr;     // learned preconditioner of flexible CG (not owned)
    int fcg_mode = 0;           // flexible CG without a network: 1 MG V-cycle, 2 identity (plain CG)
    int learned_maxit = 100;
    int learned_sweeps = 2;     // damped-Jacobi corrections after each network application
    int last_learned = 0;       // flexible-CG iterations of the last projection
    bool last_fallback = false; // the last projection fell back to MGPCG
    double last_fcg_rn = 0.0;   // ||r|| at which flexible CG converged
    std::vector<double> fcg_hist;   // ||r_k|| of the last flexible-CG run
#endif
    HiresGrid hr;
    // diagnostics of the last projection
    int last_iters = 0, last_restarts = 0;
    bool last_true_valid = false;
};

namespace vfx {

int arg_error(const std::string& m);
int upload(vfx_ctx* c, const float* host, float* dst, hipStream_t s);
int validate_prm(const float* prm);
// the three stages of Solver3D.step (solver.py:146-155) on stream s, parameters already in c->prm
int stage_advect(vfx_ctx* c, real* data, const real* solid, hipStream_t s);
int stage_forces(vfx_ctx* c, real* data, hipStream_t s);
// in_frame: part of a batched frame (the frame's CG budget, device-side sequencing; the caller
// records done_ev).  Outside frames the caller follows it with the completion event and read-back.
int stage_project(vfx_ctx* c, real* data, hipStream_t s, bool in_frame = false);
// destroy the context's captured step graphs (after a change they cannot see, e.g. new fine buffers)
void drop_graphs(vfx_ctx* c);
// iteration budget of the first CG run / of the gated restart run for P_MAXIT = maxit
int cg_budget_main(const vfx_ctx* c, int maxit);
int cg_budget_restart(const vfx_ctx* c, int maxit);
// the fused projection of the velocity vel (fused.hip); phi is the context's warm start
// tail_budget >= 0: at most that many adaptive tail cycles are enqueued, with no host polling;
// -1: the host polls the stop flag after each tail cycle
int run_fused(vfx_ctx* c, real* vel, hipStream_t s, int tail_budget);
#if !VFX_F64
// the context's learned preconditioner runs its flexible CG on the device (the coarse correction),
// so its projections need no host synchronisation (single steps and batched frames)
bool learned_on_device(const vfx_ctx* c);
// the learned flexible CG on the device (cg.hip): iteration k's p update and matvec with the
// coarse correction cz; after the V-cycle, the network input (block means into cin, scale into
// scal) and the block sums of r and r_prev; after the network, beta and the stop tests
void launch_upd_matvec_learned(const Pressure& P, int k, const real* cz, hipStream_t s);
unsigned resid_pool_blocks(int N);
void launch_resid_pool(const Pressure& P, int k, real* cin, real* sr, real* srp, double* lpart, double* scal,
                       hipStream_t s);
void launch_fcg_fin(const Pressure& P, int k, const float* out, const double* scal, const uint8_t* cmask,
                    const real* sr, const real* srp, real* cz, const double* lpart, double* fpart, hipStream_t s);
#endif
// the finer scalar grid (hires.hip): the fine stage (advection, mass ratio, sources, diffusion) by
// the velocity of `data`, and the restriction of the fine scalars into data's density,
// temperature and fuel channels
int stage_scalars(vfx_ctx* c, const real* data, hipStream_t s);
void launch_restrict(vfx_ctx* c, real* out, hipStream_t s);
void hires_free(vfx_ctx* c);
// one launch of fine kernel VFX_K_F* for vfx_time_kernel; false without a fine grid or for another id
bool hires_launch_single(vfx_ctx* c, int which, hipStream_t s);

}  // namespace vfx
