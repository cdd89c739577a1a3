// C ABI of the sim3d library (include/vfxsim3d.h): context, buffers, parameter upload ring and
// the host-side step / projection sequencing (Solver3D, vfx/sim3d/solver.py:79-459).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "ctx.hpp"

namespace vfx {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

bool launch_ok(const char* what) { return hip_ok(hipGetLastError(), what); }

Geom make_geom(int N, int zc) {
    Geom g;
    g.N = N;
    g.n3 = (size_t)N * N * N;
    g.bx = (N % 64 == 0 || N > 32) ? 64 : 32;
    g.by = 256 / g.bx;
    if (zc <= 0) {
        // bound the block count (and so the reduction partials) to about 4096
        const size_t per = (size_t)256 * 4096;
        zc = (int)((g.n3 + per - 1) / per);
        if (zc < 1) zc = 1;
    }
    g.zc = zc;
    g.block = dim3(g.bx, g.by, 1);
    g.grid = dim3((N + g.bx - 1) / g.bx, (N + g.by - 1) / g.by, (N + zc - 1) / zc);
    return g;
}

// the level-1 scale of a standalone V-cycle: 2^-e with 2^e <= max|r| over fluid cells < 2^(e+1)
// (the solver takes it from the CG residual; it matters only for 16-bit level storage): block
// maxima, then the scale in one workgroup
constexpr int LSCALE_BLOCKS = 1024;
__global__ void __launch_bounds__(256) k_lscale_part(size_t n, const real* __restrict__ r,
                                                     const uint8_t* __restrict__ fm, double* __restrict__ part) {
    __shared__ double sh[4];
    double m = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fm[i] ? fmax(m, fabs((double)r[i])) : m;
    const double v = block_max(m, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(FIN_THREADS) k_lscale_fin(const double* __restrict__ part, DevState* __restrict__ st) {
    __shared__ double sh[FIN_THREADS / 64];
    const double v = reduce_parts_max(part, LSCALE_BLOCKS, sh);
    if (threadIdx.x == 0) st->s[S_LSCALE] = (v > 0.0 && v < 1e300) ? ldexp(1.0, -ilogb(v)) : 1.0;
}

static void launch_lscale(vfx_ctx* c, const real* r, hipStream_t s) {
    hipLaunchKernelGGL(k_lscale_part, dim3(LSCALE_BLOCKS), dim3(256), 0, s, c->g.n3, r, c->P.fm, c->P.part);
    hipLaunchKernelGGL(k_lscale_fin, dim3(1), dim3(FIN_THREADS), 0, s, c->P.part, c->st);
}

__global__ void k_reset_cg(DevState* st) {
    if (threadIdx.x == 0) {
        st->it = 0;
        st->stop_at = INT_MAX;
        st->early = 0;
    }
}

}  // namespace vfx

using namespace vfx;

static inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

#define CHECK_HIP(expr, what)                                    \
    do {                                                         \
        if (!hip_ok((expr), what)) return VFX_E_HIP;             \
    } while (0)
#define CHECK_LAUNCH(what)                                       \
    do {                                                         \
        if (!launch_ok(what)) return VFX_E_HIP;                  \
    } while (0)

int vfx::arg_error(const std::string& m) {
    set_error(m);
    return VFX_E_ARG;
}

int vfx::upload(vfx_ctx* c, const float* host, float* dst, hipStream_t s) {
    const int i = c->ring_i;
    c->ring_i = (i + 1) % HOST_RING;
    if (c->ring_used[i]) CHECK_HIP(hipEventSynchronize(c->ring_ev[i]), "parameter ring wait");
    float* slot = c->ring + (size_t)i * P_SIZE;
    std::memcpy(slot, host, sizeof(float) * P_SIZE);
    CHECK_HIP(hipMemcpyAsync(dst, slot, sizeof(float) * P_SIZE, hipMemcpyHostToDevice, s), "parameter upload");
    CHECK_HIP(hipEventRecord(c->ring_ev[i], s), "parameter ring event");
    c->ring_used[i] = true;
    return VFX_OK;
}

int vfx::validate_prm(const float* prm) {
    if (prm == nullptr) return arg_error("prm_host is NULL");
    if (prm[P_N_EM] < 0 || prm[P_N_EM] > MAX_EMITTERS || prm[P_N_IMP] < 0 || prm[P_N_IMP] > MAX_IMPULSES) {
        set_error("at most 32 emitters / 32 impulses");
        return VFX_E_VALUE;
    }
    return VFX_OK;
}

int vfx::stage_advect(vfx_ctx* c, real* data, const real* solid, hipStream_t s) {
    c->P.setup(solid, s);
    CHECK_LAUNCH("obstacle setup");
    const size_t n3 = c->g.n3;
    if (c->hr.S > 0) {
        // with the finer scalar grid only the velocity is advected here: the coarse density,
        // temperature and fuel after the step are the restriction of the fine ones, and nothing
        // reads their advected values (solver.py:443-445)
        launch_advect(c->g, 3, data + n3, data + n3, c->prm, c->inv_dx, c->hat + n3, c->qidx(), c->adv + n3,
                      nullptr, nullptr, s);
    } else {
        launch_advect(c->g, 6, data, data + n3, c->prm, c->inv_dx, c->hat, c->qidx(), c->adv, c->P.fm,
                      c->mass_part, s);
        launch_mass_fin(c->g, c->mass_part, c->st, s);
    }
    CHECK_LAUNCH("advection");
    return VFX_OK;
}

int vfx::stage_forces(vfx_ctx* c, real* data, hipStream_t s) {
    if (c->hr.S > 0) {
        // the coarse scalars the forces read = the fine scalars, box-restricted (solver.py:443-445)
        launch_sources_vel(c->g, c->adv, c->P.fm, c->prm, data, s);
        launch_restrict(c, data, s);
    } else {
        launch_sources_diffuse(c->g, c->adv, c->P.fm, c->prm, c->st, c->hat, data, s);
    }
    launch_vort_forces(c->g, data, c->P.fm, c->prm, c->inv2dx, c->omega_s, s);
    CHECK_LAUNCH("forces");
    return VFX_OK;
}

// ---------------------------------------------------------------- projection sequencing
// The reference runs its PCG loop on the device as a CUDA-graph conditional (pressure.py:940-950);
// HIP has no conditional graph nodes.  Here a projection is sequenced in one of two ways:
//  - device (default): the host enqueues an iteration budget; every kernel of an iteration after
//    the device's own stop test returns at once.  With restarts on, the true residual, a restart
//    gate and one gated restart run follow, all decided on the device.  The host never waits: the
//    budgets come from the iteration counts of earlier projections, which the host reads back
//    lazily (pinned copies of the device state, polled without blocking).
//  - polled: the host enqueues iteration k, then waits for iteration k-1's stop flag (the learned
//    path's MGPCG fallback and VFX_OPT_SYNC = 1).
constexpr int BUDGET_START = 32;        // first budget of a context, before any read-back arrived
constexpr int RESTART_BUDGET_START = 1;   // a restart from a true residual just above tolerance needs ~1
constexpr int FUSED_TAIL_START = 1 << 20;   // fused projection: the whole tail until a count is known
constexpr int FALLBACK_BUDGET_MIN = 16;   // restart budget after a learned run's MGPCG fallback ran out of it
constexpr int BUDGET_MARGIN = 1;        // iterations enqueued beyond the largest recent count
constexpr int BUDGET_CAP = 1000;

int vfx::cg_budget_main(const vfx_ctx* c, int maxit) {
    const int b = c->cg_budget > 0 ? c->cg_budget : (c->bw.budget > 0 ? c->bw.budget : BUDGET_START);
    return std::max(1, std::min(b, maxit));
}

int vfx::cg_budget_restart(const vfx_ctx* c, int maxit) {
    const int b = c->bw_r.budget > 0 ? c->bw_r.budget : RESTART_BUDGET_START;
    return std::max(1, std::min(b, maxit));
}

static int fused_tail_budget(const vfx_ctx* c) {
    return c->cg_budget > 0 ? c->cg_budget : (c->bw.budget > 0 ? c->bw.budget : FUSED_TAIL_START);
}

// The host-polled PCG loop: iteration k is enqueued before the stop flag of iteration k-1 is
// read, so the device always has the next iteration queued and stopped iterations are no-ops.
static int run_cg_polled(vfx_ctx* c, hipStream_t s) {
    Pressure& P = c->P;
    const int maxit = (int)c->prm_last[P_MAXIT];
    P.iteration(0, c->prm, s);
    CHECK_LAUNCH("pcg init");
    CHECK_HIP(hipMemcpyAsync(&c->stop_pin[0], &c->st->stop_at, sizeof(int), hipMemcpyDeviceToHost, s), "stop read");
    CHECK_HIP(hipEventRecord(c->stop_ev[0], s), "stop event");
    for (int k = 1; k <= maxit; ++k) {
        P.iteration(k, c->prm, s);
        CHECK_LAUNCH("pcg iteration");
        const int slot = k % STOP_RING;
        CHECK_HIP(hipMemcpyAsync(&c->stop_pin[slot], &c->st->stop_at, sizeof(int), hipMemcpyDeviceToHost, s),
                  "stop read");
        CHECK_HIP(hipEventRecord(c->stop_ev[slot], s), "stop event");
        const int prev = (k - 1) % STOP_RING;
        CHECK_HIP(hipEventSynchronize(c->stop_ev[prev]), "stop wait");
        if (c->stop_pin[prev] <= k - 1) break;
    }
    return VFX_OK;
}

// Device-sequenced projection of vel (after rhs): the first CG run within its budget, then with
// restarts on the true residual, the restart gate, one gated restart run and its true residual.
static int project_device(vfx_ctx* c, real* vel, hipStream_t s) {
    Pressure& P = c->P;
    const int maxit = (int)c->prm_last[P_MAXIT];
    const double rel_tol = (double)c->prm_last[P_REL_TOL];
    const int B = cg_budget_main(c, maxit);
    P.rhs(vel, c->prm, c->prm_last[P_WARM] > 0.5f, s);
    for (int k = 0; k <= B; ++k) P.iteration(k, c->prm, s, B, 1);
    P.finish(s);
    const bool restart = c->opt_restarts > 0 && rel_tol > 0.0;
    const bool judged = c->opt_true_res || restart;
    int Br = 0;
    if (judged) {
        P.true_residual(vel, s);
        if (restart) {
            Br = cg_budget_restart(c, maxit);
            P.restart_gate(c->prm, s);
            P.rhs(vel, c->prm, true, s, true, true);
            for (int k = 0; k <= Br; ++k) P.iteration(k, c->prm, s, Br, 2);
            P.finish(s, true);
            P.true_residual(vel, s, true, true);
        }
    }
    CHECK_LAUNCH("pressure solve");
    c->last_true_valid = judged;
    c->last_info = ReadbackInfo{1, 0, B, Br, 0, judged ? 1 : 0, (float)rel_tol};
    return VFX_OK;
}

// A read-back projection's device state -> convergence counters and the next budgets.
static void absorb(vfx_ctx* c, const DevState& d, const ReadbackInfo& in) {
    c->n_seen += 1;
    if (in.fused) {
        if (!in.device) return;
        const bool hit = d.stop_at == INT_MAX;
        c->bw.push(std::max(0, d.it - in.cycles));
        int want = c->bw.max() + BUDGET_MARGIN;
        if (hit) {
            c->n_hit += 1;
            want = std::max(want, 2 * std::max(in.budget, 1));
        }
        c->bw.budget = std::min(want, BUDGET_CAP);
        return;
    }
    bool conv = d.conv != 0;
    if (in.judged && in.rel_tol > 0.0f) conv = std::sqrt(d.s[S_TRUE_RR]) <= (double)in.rel_tol * std::sqrt(d.s[S_BB]);
    if (!conv) c->n_unconverged += 1;
    if (!in.device) return;
    if (d.hit) c->n_hit += 1;
    c->bw.push(d.restarts > 0 ? d.it_total : d.it);
    int want = c->bw.max() + BUDGET_MARGIN;
    if (d.hit & 1) want = std::max(want, 2 * in.budget);
    c->bw.budget = std::min(want, BUDGET_CAP);
    if (d.restarts > 0) {
        c->bw_r.push(d.it);
        int wr = std::max(RESTART_BUDGET_START, c->bw_r.max() + BUDGET_MARGIN);
        if (d.hit & 2) wr = std::max(wr, 2 * in.budget_r);
        // after a learned run, the restart run is the MGPCG fallback: a whole solve, not a touch-up
        if ((d.hit & 2) && in.learned) wr = std::max(wr, std::max(in.budget, FALLBACK_BUDGET_MIN));
        c->bw_r.budget = std::min(wr, BUDGET_CAP);
    }
}

// Absorb every finished read-back in order without blocking (wait_oldest: wait for the oldest
// pending one first).
static int drain_readbacks(vfx_ctx* c, bool wait_oldest) {
    while (c->rb_count > 0) {
        const int slot = c->rb_head;
        const hipError_t e = wait_oldest ? hipEventSynchronize(c->rb_ev[slot]) : hipEventQuery(c->rb_ev[slot]);
        wait_oldest = false;
        if (e == hipErrorNotReady) {
            (void)hipGetLastError();   // a pending event is not an error
            break;
        }
        CHECK_HIP(e, "state read-back");
        absorb(c, c->rb_pin[slot], c->rb_info[slot]);
        c->rb_head = (slot + 1) % READBACK_RING;
        c->rb_count -= 1;
    }
    return VFX_OK;
}

// After a projection was enqueued on s: its completion event and the read-back of its state.
static int after_projection(vfx_ctx* c, hipStream_t s) {
    CHECK_HIP(hipEventRecord(c->done_ev, s), "done event");
    if (c->rb_count == READBACK_RING)
        if (int rc = drain_readbacks(c, true)) return rc;
    const int slot = (c->rb_head + c->rb_count) % READBACK_RING;
    CHECK_HIP(hipMemcpyAsync(&c->rb_pin[slot], c->st, sizeof(DevState), hipMemcpyDeviceToHost, s), "state read-back");
    CHECK_HIP(hipEventRecord(c->rb_ev[slot], s), "read-back event");
    c->rb_info[slot] = c->last_info;
    c->rb_count += 1;
    return VFX_OK;
}

#if !VFX_F64
// ---------------------------------------------------------------- flexible CG
// r.z, z.r_prev, r.r and max|r| block partials (FP64)
__global__ void __launch_bounds__(256) k_fcg_dots(size_t n, const real* __restrict__ r, const real* __restrict__ z,
                                                  const real* __restrict__ rp, double* __restrict__ part, int nb) {
    __shared__ double sh[4];
    double a = 0.0, b = 0.0, c = 0.0, m = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double ri = r[i], zi = z[i];
        a += ri * zi;
        b += zi * (double)rp[i];
        c += ri * ri;
        m = fmax(m, fabs(ri));
    }
    const double a1 = block_sum(a, sh), b1 = block_sum(b, sh), c1 = block_sum(c, sh), m1 = block_max(m, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a1;
        part[nb + blockIdx.x] = b1;
        part[2 * nb + blockIdx.x] = c1;
        part[3 * nb + blockIdx.x] = m1;
    }
}

__global__ void __launch_bounds__(256) k_fcg_dot(size_t n, const real* __restrict__ a, const real* __restrict__ b,
                                                 double* __restrict__ part) {
    __shared__ double sh[4];
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += (double)a[i] * (double)b[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ordered sums (max for slot 3) of nk partial arrays of nb blocks -> out[0..nk)
__global__ void __launch_bounds__(FIN_THREADS) k_fcg_fin(const double* __restrict__ part, int nb, int nk,
                                                         double* __restrict__ out) {
    __shared__ double sh[FIN_THREADS / 64];
    for (int j = 0; j < nk; ++j) {
        const double v = j == 3 ? reduce_parts_max(part + (size_t)j * nb, nb, sh) : reduce_parts_sum(part + (size_t)j * nb, nb, sh);
        if (threadIdx.x == 0) out[j] = v;
        __syncthreads();
    }
}

// z += omega D^-1 (r - t), t = K z (a damped-Jacobi correction of the network's output)
__global__ void __launch_bounds__(256) k_fcg_jacobi(size_t n, const real* __restrict__ r, const real* __restrict__ t,
                                                    const uint8_t* __restrict__ fm, real* __restrict__ z) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = fm[i];
    if (kopen(m)) z[i] = z[i] + OMEGA_JAC * (r[i] - t[i]) / kdiag(m);
}

__global__ void __launch_bounds__(256) k_fcg_sub(size_t n, const real* __restrict__ a, const real* __restrict__ b,
                                                 real* __restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

__global__ void __launch_bounds__(256) k_fcg_p(size_t n, const real* __restrict__ z, real beta, real* __restrict__ p) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

__global__ void __launch_bounds__(256) k_fcg_update(size_t n, real alpha, const real* __restrict__ p,
                                                    const real* __restrict__ q, const real* __restrict__ r,
                                                    real* __restrict__ phi, real* __restrict__ r_new) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    phi[i] = phi[i] + alpha * p[i];
    r_new[i] = r[i] - alpha * q[i];
}

static int fcg_reduce(vfx_ctx* c, int nk, double* host, hipStream_t s) {
    double* part = c->P.part;
    double* out = c->P.part + (size_t)4 * FCG_BLOCKS;
    hipLaunchKernelGGL(k_fcg_fin, dim3(1), dim3(FIN_THREADS), 0, s, part, FCG_BLOCKS, nk, out);
    CHECK_HIP(hipMemcpyAsync(host, out, sizeof(double) * nk, hipMemcpyDeviceToHost, s), "fcg scalars");
    CHECK_HIP(hipStreamSynchronize(s), "fcg scalars");
    return VFX_OK;
}

// Flexible CG (Polak-Ribiere form: beta = z_k.(r_k - r_{k-1}) / z_{k-1}.r_{k-1}, Golub & Ye 1999)
// preconditioned by the network (or, as a check, by the MG V-cycle), FP64 scalars.  Falls back to
// MGPCG from the current phi when ||r|| falls less than 2x over 3 iterations, turns non-finite or
// the learned iteration budget runs out.  Returns with phi (and r in P.R[0] semantics undefined).
static int run_fcg(vfx_ctx* c, const real* vel, hipStream_t s, bool* converged) {
    Pressure& P = c->P;
    const size_t n3 = c->g.n3;
    real* r = c->hat;
    real* rp = c->hat + n3;
    real* p = c->hat + 2 * n3;
    real* q = c->hat + 3 * n3;
    real* z = c->hat + 4 * n3;
    const unsigned nb = (unsigned)((n3 + 255) / 256);
    const double rel_tol = (double)c->prm_last[P_REL_TOL];
    const float tol = c->prm_last[P_TOL] > 1.0e-6f ? c->prm_last[P_TOL] : 1.0e-6f;
    const int maxit = std::min(c->learned_maxit, (int)c->prm_last[P_MAXIT]);
    DevState hs;
    CHECK_HIP(hipMemcpyAsync(r, P.R[0], n3 * sizeof(real), hipMemcpyDeviceToDevice, s), "fcg r0");
    CHECK_HIP(hipMemcpyAsync(&hs, c->st, sizeof(DevState), hipMemcpyDeviceToHost, s), "fcg state");
    CHECK_HIP(hipStreamSynchronize(s), "fcg state");
    const double bnorm = std::sqrt(hs.s[S_BB]);
    const double thr = hs.s[S_THR];
    double rz_prev = 0.0;
    std::vector<double>& hist = c->fcg_hist;
    hist.clear();
    *converged = false;
    c->last_learned = 0;
    for (int k = 0; k < maxit; ++k) {
        // z = M(r)
        if (c->fcg_mode == 2) {
            CHECK_HIP(hipMemcpyAsync(z, r, n3 * sizeof(real), hipMemcpyDeviceToDevice, s), "fcg identity");
        } else if (c->fcg_mode == 1 || c->net->net.N * 4 == c->N) {
            // the MG V-cycle; with a coarse network, plus its correction of the residual the
            // V-cycle leaves: z = V(r) + P net(R (r - K V(r)))
            CHECK_HIP(hipMemcpyAsync(P.R[0], r, n3 * sizeof(real), hipMemcpyDeviceToDevice, s), "fcg mg input");
            hipLaunchKernelGGL(k_reset_cg, dim3(1), dim3(64), 0, s, c->st);
            launch_lscale(c, P.R[0], s);
            P.precond(0, s);
            CHECK_HIP(hipMemcpyAsync(z, P.z, n3 * sizeof(real), hipMemcpyDeviceToDevice, s), "fcg mg output");
            if (c->fcg_mode != 1) {
                real* rv = c->hat + 5 * n3;
                launch_kmul(c->g, z, P.fm, q, s);
                hipLaunchKernelGGL(k_fcg_sub, dim3(nb), dim3(256), 0, s, n3, r, q, rv);
                c->net->net.apply_coarse(c->N, rv, P.fm, z, s);
            }
        } else {
            c->net->net.apply(r, P.fm, z, s);
            for (int sw = 0; sw < c->learned_sweeps; ++sw) {
                launch_kmul(c->g, z, P.fm, q, s);
                hipLaunchKernelGGL(k_fcg_jacobi, dim3(nb), dim3(256), 0, s, n3, r, q, P.fm, z);
            }
        }
        hipLaunchKernelGGL(k_fcg_dots, dim3(FCG_BLOCKS), dim3(256), 0, s, n3, r, z, rp, P.part, FCG_BLOCKS);
        double d[4];
        if (int rc = fcg_reduce(c, 4, d, s)) return rc;
        const double rz = d[0], zrp = d[1], rr = d[2], mr = d[3];
        const double rn = std::sqrt(rr);
        hist.push_back(rn);
        c->last_learned = k;
        const bool conv = rel_tol > 0.0 ? rn <= rel_tol * bnorm : mr <= thr;
        if (conv) {
            *converged = true;
            c->last_fcg_rn = rn;
            return VFX_OK;
        }
        if (!std::isfinite(rz) || !std::isfinite(rn) || rz <= 0.0) return VFX_OK;   // fallback
        if (k >= 3 && rn > 0.5 * hist[k - 3]) return VFX_OK;                           // stalled: fallback
        const double beta = k == 0 ? 0.0 : (rz - zrp) / rz_prev;
        rz_prev = rz;
        hipLaunchKernelGGL(k_fcg_p, dim3(nb), dim3(256), 0, s, n3, z, (real)beta, p);
        launch_kmul(c->g, p, P.fm, q, s);
        hipLaunchKernelGGL(k_fcg_dot, dim3(FCG_BLOCKS), dim3(256), 0, s, n3, p, q, P.part);
        double pq;
        if (int rc = fcg_reduce(c, 1, &pq, s)) return rc;
        if (!(pq > 0.0) || !std::isfinite(pq)) return VFX_OK;   // fallback
        const double alpha = rz / pq;
        hipLaunchKernelGGL(k_fcg_update, dim3(nb), dim3(256), 0, s, n3, (real)alpha, p, q, r, P.phi, rp);
        std::swap(r, rp);   // r <- r_new, rp <- r_old
        CHECK_LAUNCH("flexible cg");
        c->last_learned = k + 1;
    }
    return VFX_OK;
}

// The coarse-correction network runs its flexible CG on the device (no host synchronisation);
// the full-resolution network and the flexible-CG checks without a network are host-polled.
bool vfx::learned_on_device(const vfx_ctx* c) {
    return c->net != nullptr && c->fcg_mode == 0 && c->net->net.N * 4 == c->N;
}

// Iteration k of the learned flexible CG (run budget kmax): the MGPCG iteration's kernels with
// the search direction formed from z_V + P c, then the network's coarse correction and beta.
static void learned_iteration(vfx_ctx* c, int k, int kmax, int hitbit, hipStream_t s) {
    Pressure& P = c->P;
    Net& n = c->net->net;
    if (k > 0) launch_upd_matvec_learned(P, k, n.ccz, s);
    if (k >= kmax) {
        P.precond_front(k, s, c->prm, 2, hitbit);   // the stop test on r decides
        return;
    }
    P.precond_front(k, s, c->prm, 1);
    P.precond_back(k, s);   // z_V = V(r_k) and the partials r.z_V, max|r|, r.r (no finaliser)
    launch_resid_pool(P, k, n.work2, n.csr, n.csrp, n.lpart, n.scal, s);
    n.coarse_forward(c->st, k, s);
    launch_fcg_fin(P, k, n.out, n.scal, n.cmask, n.csr, n.csrp, n.ccz, n.lpart, n.fpart, s);
}

// Device-sequenced learned projection of vel: the learned run within its budget (at most
// learned_maxit iterations), the true residual, the gate and the gated MGPCG run that finishes
// the solve from phi when the learned run stalled, broke down, stopped unconverged or left a true
// residual above the relative tolerance.
static int project_learned_device(vfx_ctx* c, real* vel, hipStream_t s) {
    Pressure& P = c->P;
    const int maxit = (int)c->prm_last[P_MAXIT];
    const int kl = std::max(1, std::min(c->learned_maxit, maxit));
    const double rel_tol = (double)c->prm_last[P_REL_TOL];
    const int B = std::min(cg_budget_main(c, maxit), kl);
    c->fcg_hist.clear();
    P.rhs(vel, c->prm, c->prm_last[P_WARM] > 0.5f, s);
    c->net->net.coarse_mask(c->N, P.fm, s);
    const int hb = B < kl ? 1 : 0;   // a stop at the learned limit is not a budget stop
    for (int k = 0; k <= B; ++k) learned_iteration(c, k, B, hb, s);
    P.finish(s);
    P.true_residual(vel, s);
    const int Br = cg_budget_restart(c, maxit);
    P.restart_gate(c->prm, s, true);
    P.rhs(vel, c->prm, true, s, true, true);
    for (int k = 0; k <= Br; ++k) P.iteration(k, c->prm, s, Br, 2);
    P.finish(s, true);
    P.true_residual(vel, s, true, true);
    CHECK_LAUNCH("learned pressure solve");
    c->last_true_valid = true;
    c->last_info = ReadbackInfo{1, 0, B, Br, 0, 1, (float)rel_tol, 1};
    return VFX_OK;
}
#endif

int vfx::stage_project(vfx_ctx* c, real* data, hipStream_t s, bool in_frame) {
    Pressure& P = c->P;
    real* vel = data + c->g.n3;
    const bool device = in_frame || c->sync_mode == SYNC_DEVICE;
    c->last_iters = 0;
    c->last_restarts = 0;
    c->last_true_valid = false;
#if !VFX_F64
    c->last_learned = 0;
    c->last_fallback = false;
    const bool learned = c->net != nullptr || c->fcg_mode != 0;
#else
    const bool learned = false;
#endif
    if (c->fused) {
        const int tb = device ? fused_tail_budget(c) : -1;
        if (int rc = run_fused(c, vel, s, tb)) return rc;
        c->last_info = ReadbackInfo{device ? 1 : 0, 1, tb, 0, c->fused_cycles, 0, 0.0f};
        c->projected = true;
        return VFX_OK;
    }
#if !VFX_F64
    if (device && learned_on_device(c)) {
        if (int rc = project_learned_device(c, vel, s)) return rc;
        P.apply(vel, s);
        CHECK_LAUNCH("pressure apply");
        c->projected = true;
        return VFX_OK;
    }
#endif
    if (in_frame && learned) {
        set_error("a batched frame runs MGPCG or the coarse-correction network only (the other learned "
                  "preconditioners wait on the host)");
        return VFX_E_VALUE;
    }
    if (device && !learned) {
        if (int rc = project_device(c, vel, s)) return rc;
        P.apply(vel, s);
        CHECK_LAUNCH("pressure apply");
        c->projected = true;
        return VFX_OK;
    }
    // host-polled: the learned path (flexible CG, then the MGPCG fallback) and VFX_OPT_SYNC = 1
    P.rhs(vel, c->prm, c->prm_last[P_WARM] > 0.5f, s);
    CHECK_LAUNCH("pressure rhs");
    int rc = VFX_OK;
    bool need_pcg = true;
#if !VFX_F64
    if (learned) {
        bool conv = false;
        rc = run_fcg(c, vel, s, &conv);
        if (rc) return rc;
        need_pcg = !conv;
        c->last_fallback = !conv;
        if (!conv) {
            // plain MGPCG from the current phi
            P.rhs(vel, c->prm, true, s);
            CHECK_LAUNCH("fallback rhs");
        } else {
            hipLaunchKernelGGL(k_reset_cg, dim3(1), dim3(64), 0, s, c->st);   // no MGPCG iterations
        }
    }
#endif
    if (need_pcg) {
        rc = run_cg_polled(c, s);
        if (rc) return rc;
        P.finish(s);
    }
    int total = 0;
    const double rel_tol = (double)c->prm_last[P_REL_TOL];
    const bool judged = c->opt_true_res || (c->opt_restarts > 0 && rel_tol > 0.0);
    if (judged) {
        for (;;) {
            P.true_residual(vel, s, false, c->last_restarts > 0);
            CHECK_LAUNCH("true residual");
            c->last_true_valid = true;
            if (!(c->opt_restarts > c->last_restarts && rel_tol > 0.0)) break;
            DevState hs;
            CHECK_HIP(hipMemcpyAsync(&hs, c->st, sizeof(DevState), hipMemcpyDeviceToHost, s), "state read");
            CHECK_HIP(hipStreamSynchronize(s), "state sync");
            if (std::sqrt(hs.s[S_TRUE_RR]) <= rel_tol * std::sqrt(hs.s[S_BB])) break;
            total += hs.it;
            c->last_restarts += 1;
            // restart from the true residual: warm start from the current phi
            P.rhs(vel, c->prm, true, s, false, true);
            CHECK_LAUNCH("pressure restart rhs");
            rc = run_cg_polled(c, s);
            if (rc) return rc;
            P.finish(s, true);
        }
    }
    c->last_iters = total;   // plus DevState::it of the final run, read lazily
    c->last_info = ReadbackInfo{0, 0, 0, 0, 0, judged ? 1 : 0, (float)rel_tol};
    P.apply(vel, s);
    CHECK_LAUNCH("pressure apply");
    c->projected = true;
    return VFX_OK;
}

// ---------------------------------------------------------------- whole steps as graphs
static int enqueue_step(vfx_ctx* c, real* data, const real* solid, hipStream_t s) {
    if (int rc = stage_advect(c, data, solid, s)) return rc;
    if (int rc = stage_scalars(c, data, s)) return rc;
    if (int rc = stage_forces(c, data, s)) return rc;
    return stage_project(c, data, s);
}

// destroy a slot's exec once its last launch has finished (the host waits for that launch if it is
// still running; the least recently used slot's launch has normally finished long before)
static void release_exec(StepGraph& g) {
    if (g.exec) {
        if (g.launched) (void)hipEventSynchronize(g.done);
        (void)hipGraphExecDestroy(g.exec);
    }
    g.exec = nullptr;
    g.launched = false;
}

void vfx::drop_graphs(vfx_ctx* c) {
    for (StepGraph& g : c->graphs) release_exec(g);
}

// A step is replayed as a graph in device mode, except when an operator rebuild was forced (a
// one-off that runs eagerly) and on the learned path (its loop waits on the host).
static bool graph_capable(const vfx_ctx* c) {
    if (!c->use_graph || c->sync_mode != SYNC_DEVICE || c->P.force_setup) return false;
#if !VFX_F64
    if ((c->net != nullptr && !learned_on_device(c)) || c->fcg_mode != 0) return false;
#endif
    return true;
}

// every host-side decision that enqueue_step makes from the context and the parameters
static StepSig step_sig(const vfx_ctx* c, const real* data, const real* solid) {
    StepSig g;
    std::memset(&g, 0, sizeof(g));
    g.data = data;
    g.solid = solid;
#if !VFX_F64
    g.net = c->net;
#endif
    const int maxit = (int)c->prm_last[P_MAXIT];
    const double rel_tol = (double)c->prm_last[P_REL_TOL];
    int* v = g.v;
    v[0] = c->prm_last[P_WARM] > 0.5f;
    v[1] = c->fused ? 1 : 0;
    v[2] = c->hr.S;
    v[3] = c->g.zc;
    v[4] = c->P.mg_kind;
    v[5] = c->P.S;
    v[6] = c->P.use_std ? 1 : 0;
    if (c->fused) {
        v[7] = fused_tail_budget(c);
        v[8] = c->fused_cycles;
        v[9] = c->fused_smooth;
        v[10] = c->fused_jacobi;
        v[11] = c->prm_last[P_TOL] > 0.0f;
        v[12] = std::max(0, maxit - std::max(0, c->fused_cycles));
    } else {
        const bool restart = c->opt_restarts > 0 && rel_tol > 0.0;
        v[7] = cg_budget_main(c, maxit);
        v[8] = restart ? cg_budget_restart(c, maxit) : 0;
        v[9] = (c->opt_true_res || restart) ? 1 : 0;
#if !VFX_F64
        if (c->net != nullptr) {
            v[8] = cg_budget_restart(c, maxit);
            v[10] = c->learned_maxit;
        }
#endif
    }
    return g;
}

// Enqueue one step by replaying the captured graph of its signature (capturing it first if none
// of the context's graphs has it; the least recently used one is replaced).
static int graph_step(vfx_ctx* c, real* data, const real* solid, hipStream_t s) {
    const StepSig sig = step_sig(c, data, solid);
    StepGraph* hit = nullptr;
    StepGraph* victim = nullptr;   // an empty slot, else the least recently used
    for (StepGraph& g : c->graphs) {
        if (g.exec && std::memcmp(&g.sig, &sig, sizeof(sig)) == 0) {
            hit = &g;
            break;
        }
        if (victim == nullptr || (victim->exec && (!g.exec || g.used < victim->used))) victim = &g;
    }
    if (hit == nullptr) {
        if (!c->cap) CHECK_HIP(hipStreamCreateWithFlags(&c->cap, hipStreamNonBlocking), "capture stream");
        CHECK_HIP(hipStreamBeginCapture(c->cap, hipStreamCaptureModeRelaxed), "step capture");
        const int rc = enqueue_step(c, data, solid, c->cap);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->cap, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        CHECK_HIP(e, "step capture");
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        CHECK_HIP(ei, "step graph instantiate");
        release_exec(*victim);
        if (!victim->done) CHECK_HIP(hipEventCreateWithFlags(&victim->done, hipEventDisableTiming), "graph event");
        victim->exec = exec;
        victim->sig = sig;
        victim->info = c->last_info;
        victim->true_valid = c->last_true_valid;
        hit = victim;
        c->graph_captures += 1;
    }
    hit->used = ++c->graph_clock;
    CHECK_HIP(hipGraphLaunch(hit->exec, s), "step graph launch");
    CHECK_HIP(hipEventRecord(hit->done, s), "graph event");
    hit->launched = true;
    // the host-side results of the enqueue that the capture made
    c->last_info = hit->info;
    c->last_true_valid = hit->true_valid;
    c->last_iters = 0;
    c->last_restarts = 0;
    c->projected = true;
    c->graph_replays += 1;
    return VFX_OK;
}

static int check_ctx(vfx_ctx* c) {
    if (c == nullptr) return arg_error("null context");
    return VFX_OK;
}

extern "C" {

int VFX_FN(vfx_p_size)(void) { return P_SIZE; }

int VFX_FN(vfx_last_error)(char* buf, int n) {
    if (buf && n > 0) {
        std::snprintf(buf, (size_t)n, "%s", g_err.c_str());
    }
    return VFX_OK;
}

int VFX_FN(vfx_ctx_create)(int res, vfx_ctx** out) {
    if (out == nullptr) return arg_error("out is NULL");
    *out = nullptr;
    if (res < 2 || res % 2 != 0) return arg_error("even grid size >= 2 required");
    // the kernels index channel c of a field as c N^3 + i in 32-bit integers: 6 N^3 < 2^31 up to 512
    if (res > 512) return arg_error("grid size above 512 (channel offsets are 32-bit)");
    vfx_ctx* c = new (std::nothrow) vfx_ctx();
    if (!c) return arg_error("out of host memory");
    const int N = res;
    c->N = N;
    c->g = make_geom(N, 1);   // one cell per thread: every load of a thread is in flight at once
    const double dx = 1.0 / (double)N;   // DOMAIN_SIZE / res (fields.py:25)
    const double inv = 1.0 / dx;
    c->inv_dx = (real)inv;
    c->inv2dx = (real)(0.5 * inv);
    Pressure& P = c->P;
    P.init_levels(N);
    P.geom = c->g;
    P.geom_r = make_geom(N, 0);
    P.init_tiles();
    P.nparts = std::max({nblocks(c->g), P.ntiles, 2 * FCG_BLOCKS});   // flexible CG uses 4 * FCG_BLOCKS + 4
    const size_t n3 = c->g.n3;
    // one pool: floats first (every array a multiple of 8 floats, so 32-byte aligned), then
    // doubles, then the mask
    const size_t nf = n3 * (6 + 6 + 3 + 4 + 3) + 2 * (size_t)P_SIZE + n3 * 7 + P.level_floats();
    const size_t nd = mass_part_doubles(c->g) + (size_t)P.nparts * (3 + 1) + sizeof(DevState) / sizeof(double) + 8;
    const size_t bytes = nf * sizeof(real) + 64 + nd * 8 + 64 + n3 + 64;
    if (!hip_ok(hipMalloc(&c->pool, bytes), "device allocation")) { delete c; return VFX_E_HIP; }
    if (!hip_ok(hipMemset(c->pool, 0, bytes), "device clear")) { hipFree(c->pool); delete c; return VFX_E_HIP; }
    real* f = (real*)c->pool;
    c->hat = f; f += 6 * n3;
    c->adv = f; f += 6 * n3;
    c->omega = f; f += 3 * n3;
    c->omega_s = f; f += 4 * n3;   // smoothed vorticity and its magnitude
    c->force = f; f += 3 * n3;
    c->prm = reinterpret_cast<float*>(f); f += P_SIZE;
    c->prm_adv = reinterpret_cast<float*>(f); f += P_SIZE;
    P.R[0] = f; f += n3;
    P.R[1] = f; f += n3;
    P.P[0] = f; f += n3;
    P.P[1] = f; f += n3;
    P.q = f; f += n3;
    P.z = f; f += n3;
    P.phi = f; f += n3;
    P.bind_levels(f); f += P.level_floats();
    uintptr_t a = ((uintptr_t)f + 63) & ~(uintptr_t)63;
    double* d = (double*)a;
    c->mass_part = d; d += mass_part_doubles(c->g);
    P.part = d; d += 3 * (size_t)P.nparts;
    P.part_pq = d; d += (size_t)P.nparts;
    a = ((uintptr_t)d + 63) & ~(uintptr_t)63;
    c->st = (DevState*)a;
    P.st = c->st;
    a = (a + sizeof(DevState) + 63) & ~(uintptr_t)63;
    P.fm = (uint8_t*)a;   // n3 mask bytes, then zeros: fm[n3] == 0 is read for cells outside the box
    bool ok = hip_ok(hipHostMalloc((void**)&c->ring, sizeof(float) * P_SIZE * HOST_RING, hipHostMallocDefault),
                     "pinned allocation") &&
              hip_ok(hipHostMalloc((void**)&c->stop_pin, sizeof(int) * STOP_RING, hipHostMallocDefault),
                     "pinned allocation") &&
              hip_ok(hipHostMalloc((void**)&c->rb_pin, sizeof(DevState) * READBACK_RING, hipHostMallocDefault),
                     "pinned allocation");
    for (int i = 0; ok && i < HOST_RING; ++i)
        ok = hip_ok(hipEventCreateWithFlags(&c->ring_ev[i], hipEventDisableTiming), "event");
    for (int i = 0; ok && i < STOP_RING; ++i)
        ok = hip_ok(hipEventCreateWithFlags(&c->stop_ev[i], hipEventDisableTiming), "event");
    for (int i = 0; ok && i < READBACK_RING; ++i)
        ok = hip_ok(hipEventCreateWithFlags(&c->rb_ev[i], hipEventDisableTiming), "event");
    ok = ok && hip_ok(hipEventCreateWithFlags(&c->done_ev, hipEventDisableTiming), "event");
    std::memset(c->prm_last, 0, sizeof(c->prm_last));
    if (!ok) { VFX_FN(vfx_ctx_destroy)(c); return VFX_E_HIP; }
    // The cell kernels' default from 128^3 is two cells per thread (the 256^3 fire step 6.87 ->
    // 6.79 ms; four: 6.82 ms).  Every buffer above is sized for one, the geometry with the most blocks.
    if (N >= 128) {
        c->g = make_geom(N, 2);
        P.geom = c->g;
    }
    *out = c;
    return VFX_OK;
}

int VFX_FN(vfx_ctx_destroy)(vfx_ctx* c) {
    if (!c) return VFX_OK;
    hipDeviceSynchronize();
    for (int i = 0; i < HOST_RING; ++i) if (c->ring_ev[i]) hipEventDestroy(c->ring_ev[i]);
    for (int i = 0; i < STOP_RING; ++i) if (c->stop_ev[i]) hipEventDestroy(c->stop_ev[i]);
    for (int i = 0; i < READBACK_RING; ++i) if (c->rb_ev[i]) hipEventDestroy(c->rb_ev[i]);
    drop_graphs(c);
    for (StepGraph& g : c->graphs)
        if (g.done) (void)hipEventDestroy(g.done);
    if (c->cap) hipStreamDestroy(c->cap);
    if (c->rb_pin) hipHostFree(c->rb_pin);
    if (c->done_ev) hipEventDestroy(c->done_ev);
    if (c->ring) hipHostFree(c->ring);
    if (c->stop_pin) hipHostFree(c->stop_pin);
    if (c->pool) hipFree(c->pool);
    hires_free(c);
    delete c;
    return VFX_OK;
}

int VFX_FN(vfx_set_option)(vfx_ctx* c, int key, double value) {
    if (int rc = check_ctx(c)) return rc;
    switch (key) {
        case VFX_OPT_TRUE_RESIDUAL: c->opt_true_res = value != 0.0; return VFX_OK;
        case VFX_OPT_RESTARTS: c->opt_restarts = value < 0 ? 0 : (int)value; return VFX_OK;
        case VFX_OPT_COARSE_M: {
            const int cm = (int)value;
            if (cm < 1) return arg_error("coarse_m must be >= 1");
            c->P.coarse_m = cm;
            c->P.S = c->P.nlev - 1;
            for (int l = 0; l < c->P.nlev; ++l)
                if (c->P.lv[l].m <= cm) { c->P.S = l; break; }
            return VFX_OK;
        }
#if !VFX_F64
        case VFX_OPT_FCG_MG:
            if (value < 0 || value > 2) return arg_error("flexible-CG mode must be 0, 1 or 2");
            c->fcg_mode = (int)value;
            return VFX_OK;
        case VFX_OPT_LEARNED_SWEEPS:
            if (value < 0 || value > 8) return arg_error("learned smoothing sweeps must be 0..8");
            c->learned_sweeps = (int)value;
            return VFX_OK;
        case VFX_OPT_LEARNED_MAXIT:
            if (value < 1) return arg_error("learned iteration budget must be >= 1");
            c->learned_maxit = (int)value;
            return VFX_OK;
#endif
        case VFX_OPT_PRESSURE_MODE:
            if (value != 0.0 && value != 1.0) return arg_error("pressure mode must be 0 (pcg) or 1 (fused)");
            if (c->fused != (value != 0.0)) {
                c->bw.reset();   // the budget counts CG iterations or fused tail cycles
                c->bw_r.reset();
            }
            c->fused = value != 0.0;
            return VFX_OK;
        case VFX_OPT_FUSED_CYCLES:
            if (value < 0 || value > 1000) return arg_error("fused cycles must be 0..1000");
            c->fused_cycles = (int)value;
            return VFX_OK;
        case VFX_OPT_FUSED_SMOOTH:
            if (value < 1 || value > 64) return arg_error("fused sweeps must be 1..64");
            c->fused_smooth = (int)value;
            return VFX_OK;
        case VFX_OPT_FUSED_SMOOTHER:
            if (value != 0.0 && value != 1.0) return arg_error("fused smoother must be 0 (chebyshev) or 1 (jacobi)");
            c->fused_jacobi = (int)value;
            return VFX_OK;
        case VFX_OPT_CG_BUDGET:
            if (value < 0 || value > 1000) return arg_error("CG budget must be 0 (adaptive) or 1..1000");
            c->cg_budget = (int)value;
            return VFX_OK;
        case VFX_OPT_SYNC:
            if (value != 0.0 && value != 1.0) return arg_error("projection sync must be 0 (device) or 1 (polled)");
            c->sync_mode = (int)value;
            return VFX_OK;
        case VFX_OPT_GRAPH:
            c->use_graph = value != 0.0;
            if (!c->use_graph) drop_graphs(c);
            return VFX_OK;
        case VFX_OPT_CELLS_PER_THREAD: {
            const int zc = (int)value;
            if (zc != 1 && zc != 2 && zc != 4 && zc != 8) return arg_error("cells per thread must be 1, 2, 4 or 8");
            c->g = make_geom(c->N, zc);   // fewer blocks than zc = 1: every partial buffer still fits
            c->P.geom = c->g;
            return VFX_OK;
        }
        case VFX_OPT_MG_STORAGE: {
            const int kind = (int)value;
            if (kind != MG_REAL && kind != MG_FP16 && kind != MG_BF16) {
                set_error("multigrid storage must be 0 (FP32), 1 (FP16) or 2 (BF16)");
                return VFX_E_VALUE;
            }
            if (VFX_F64 && kind != MG_REAL) {
                set_error("the FP64 build stores its multigrid levels in FP64");
                return VFX_E_VALUE;
            }
            if (kind != c->P.mg_kind) {
                c->P.mg_kind = kind;
                c->P.force_setup = true;   // the operators are rebuilt in the new storage
            }
            return VFX_OK;
        }
        case VFX_OPT_STD_STENCIL: {
            const bool on = value != 0.0;
            if (on != c->P.use_std) {
                c->P.use_std = on;
                c->P.force_setup = true;   // the flags are rebuilt by the next setup
            }
            return VFX_OK;
        }
        default: set_error("unknown option"); return VFX_E_VALUE;
    }
}

int VFX_FN(vfx_step_stage)(vfx_ctx* c, int stage, real* data, const real* solid, const float* prm_host, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!data || !solid) return arg_error("null field pointer");
    hipStream_t s = S(stream);
    if (stage == VFX_STAGE_ADVECT || prm_host != nullptr) {
        if (int rc = validate_prm(prm_host)) return rc;
        std::memcpy(c->prm_last, prm_host, sizeof(c->prm_last));
        if (int rc = upload(c, prm_host, c->prm, s)) return rc;
    }
    if (int rc = drain_readbacks(c, false)) return rc;
    switch (stage) {
        case VFX_STAGE_ADVECT: return stage_advect(c, data, solid, s);
        case VFX_STAGE_SCALARS: return stage_scalars(c, data, s);
        case VFX_STAGE_FORCES: return stage_forces(c, data, s);
        case VFX_STAGE_PROJECT:
            c->P.setup(solid, s);
            CHECK_LAUNCH("obstacle setup");
            if (int rc = stage_project(c, data, s)) return rc;
            return after_projection(c, s);
        default: set_error("unknown stage"); return VFX_E_VALUE;
    }
}

int VFX_FN(vfx_step)(vfx_ctx* c, real* data, const real* solid, const float* prm_host, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!data || !solid) return arg_error("null field pointer");
    if (int rc = validate_prm(prm_host)) return rc;
    hipStream_t s = S(stream);
    std::memcpy(c->prm_last, prm_host, sizeof(c->prm_last));
    if (int rc = upload(c, prm_host, c->prm, s)) return rc;
    if (int rc = drain_readbacks(c, false)) return rc;
    const int rc = graph_capable(c) ? graph_step(c, data, solid, s) : enqueue_step(c, data, solid, s);
    if (rc) return rc;
    return after_projection(c, s);
}

int VFX_FN(vfx_project)(vfx_ctx* c, real* data, const real* solid, const float* prm_host, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!data || !solid) return arg_error("null field pointer");
    if (int rc = validate_prm(prm_host)) return rc;
    hipStream_t s = S(stream);
    std::memcpy(c->prm_last, prm_host, sizeof(c->prm_last));
    if (int rc = upload(c, prm_host, c->prm, s)) return rc;
    if (int rc = drain_readbacks(c, false)) return rc;
    c->P.setup(solid, s);
    CHECK_LAUNCH("obstacle setup");
    if (int rc = stage_project(c, data, s)) return rc;
    return after_projection(c, s);
}

int VFX_FN(vfx_advect)(vfx_ctx* c, const real* field, int C, const real* vel, real* out, float dt, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!field || !vel || !out) return arg_error("null pointer");
    if (C < 1) return arg_error("channel count must be >= 1");
    hipStream_t s = S(stream);
    float h[P_SIZE];
    std::memset(h, 0, sizeof(h));
    h[P_DT] = dt;
    if (int rc = upload(c, h, c->prm_adv, s)) return rc;
    const size_t n3 = c->g.n3;
    for (int c0 = 0; c0 < C; c0 += 6) {
        const int cc = C - c0 < 6 ? C - c0 : 6;
        launch_advect(c->g, cc, field + c0 * n3, vel, c->prm_adv, c->inv_dx, c->hat, c->qidx(), out + c0 * n3,
                      nullptr, nullptr, s);
    }
    CHECK_LAUNCH("advect");
    return VFX_OK;
}

int VFX_FN(vfx_div_abs)(vfx_ctx* c, const real* data, const real* solid, real* out, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!data || !solid || !out) return arg_error("null pointer");
    launch_div_abs(c->g, data + c->g.n3, solid, out, S(stream));
    CHECK_LAUNCH("divergence");
    return VFX_OK;
}

int VFX_FN(vfx_pressure_iters)(vfx_ctx* c, int* out) {
    if (int rc = check_ctx(c)) return rc;
    if (!out) return arg_error("null pointer");
    if (!c->projected) { *out = 0; return VFX_OK; }
    CHECK_HIP(hipEventSynchronize(c->done_ev), "projection wait");
    DevState hs;
    CHECK_HIP(hipMemcpy(&hs, c->st, sizeof(DevState), hipMemcpyDeviceToHost), "state read");
    *out = c->last_iters + hs.it_total + hs.it;
    return VFX_OK;
}

int VFX_FN(vfx_pressure_stats)(vfx_ctx* c, double* out, int n) {
    if (int rc = check_ctx(c)) return rc;
    if (!out || n < 1) return arg_error("null pointer");
    double v[VFX_STATS] = {0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (c->projected) {
        CHECK_HIP(hipEventSynchronize(c->done_ev), "projection wait");
        while (c->rb_count > 0)   // every read-back up to the last projection's (they are done or nearly)
            if (int rc = drain_readbacks(c, true)) return rc;
        DevState hs;
        CHECK_HIP(hipMemcpy(&hs, c->st, sizeof(DevState), hipMemcpyDeviceToHost), "state read");
        const ReadbackInfo& in = c->last_info;
        v[0] = c->last_iters + hs.it_total + hs.it;
        v[1] = hs.s[S_MAXU];
        v[2] = std::sqrt(hs.s[S_BB]);
        v[3] = std::sqrt(hs.s[S_RR]);
        v[4] = hs.s[S_MAXR];
        v[5] = c->last_true_valid ? std::sqrt(hs.s[S_TRUE_RR]) : -1.0;
        v[6] = c->last_restarts + hs.restarts;
        bool conv = in.fused ? hs.stop_at != INT_MAX : hs.conv != 0;
#if !VFX_F64
        if (in.learned) {
            // the learned run, then the gated MGPCG run (counted as MGPCG iterations)
            v[0] = hs.restarts > 0 ? hs.it : 0;
            v[7] = hs.restarts > 0 ? hs.it_total : hs.it;
            v[8] = hs.fallback ? 1.0 : 0.0;
        } else {
            if ((c->net || c->fcg_mode) && !c->last_fallback && c->last_restarts == 0) {
                v[3] = c->last_fcg_rn;
                conv = true;
            }
            v[7] = c->last_learned;
            v[8] = c->last_fallback ? 1.0 : 0.0;
        }
#endif
        if (c->last_true_valid && in.rel_tol > 0.0f)
            conv = std::sqrt(hs.s[S_TRUE_RR]) <= (double)in.rel_tol * std::sqrt(hs.s[S_BB]);
        v[9] = conv ? 1.0 : 0.0;
        v[10] = in.device ? in.budget : 0;
    }
    v[11] = (double)c->n_hit;
    v[12] = (double)c->n_unconverged;
    v[13] = c->graph_captures;
    v[14] = c->graph_replays;
    for (int i = 0; i < n && i < VFX_STATS; ++i) out[i] = v[i];
    return VFX_OK;
}

int VFX_FN(vfx_get_phi)(vfx_ctx* c, real* dst, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!dst) return arg_error("null pointer");
    CHECK_HIP(hipMemcpyAsync(dst, c->P.phi, c->g.n3 * sizeof(real), hipMemcpyDeviceToDevice, S(stream)), "phi copy");
    return VFX_OK;
}

int VFX_FN(vfx_set_phi)(vfx_ctx* c, const real* src, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!src) return arg_error("null pointer");
    CHECK_HIP(hipMemcpyAsync(c->P.phi, src, c->g.n3 * sizeof(real), hipMemcpyDeviceToDevice, S(stream)), "phi copy");
    return VFX_OK;
}

int VFX_FN(vfx_get_advected)(vfx_ctx* c, real* dst, double* mass_ratio, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    hipStream_t s = S(stream);
    if (dst) CHECK_HIP(hipMemcpyAsync(dst, c->adv, 6 * c->g.n3 * sizeof(real), hipMemcpyDeviceToDevice, s), "adv copy");
    if (mass_ratio) {
        DevState hs;
        CHECK_HIP(hipMemcpyAsync(&hs, c->st, sizeof(DevState), hipMemcpyDeviceToHost, s), "state read");
        CHECK_HIP(hipStreamSynchronize(s), "state sync");
        *mass_ratio = hs.s[S_MASS_RATIO];
    }
    return VFX_OK;
}

int VFX_FN(vfx_kmul)(vfx_ctx* c, const real* solid, const real* p, real* out, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!solid || !p || !out) return arg_error("null pointer");
    hipStream_t s = S(stream);
    c->P.setup(solid, s);
    launch_kmul(c->g, p, c->P.fm, out, s);
    CHECK_LAUNCH("kmul");
    return VFX_OK;
}

int VFX_FN(vfx_precond)(vfx_ctx* c, const real* solid, const real* r, real* z, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!solid || !r || !z) return arg_error("null pointer");
    hipStream_t s = S(stream);
    const size_t n3 = c->g.n3;
    c->P.setup(solid, s);
    CHECK_HIP(hipMemcpyAsync(c->P.R[0], r, n3 * sizeof(real), hipMemcpyDeviceToDevice, s), "r copy");
    hipLaunchKernelGGL(k_reset_cg, dim3(1), dim3(64), 0, s, c->st);
    launch_lscale(c, c->P.R[0], s);
    c->P.precond(0, s);
    CHECK_LAUNCH("precond");
    CHECK_HIP(hipMemcpyAsync(z, c->P.z, n3 * sizeof(real), hipMemcpyDeviceToDevice, s), "z copy");
    return VFX_OK;
}

int VFX_FN(vfx_time_kernel)(vfx_ctx* c, int which, int reps, void* stream, float* ms_avg) {
    if (int rc = check_ctx(c)) return rc;
    if (!ms_avg || reps < 1) return arg_error("bad arguments");
    hipStream_t s = S(stream);
    DevState saved;
    CHECK_HIP(hipMemcpyAsync(&saved, c->st, sizeof(DevState), hipMemcpyDeviceToHost, s), "state save");
    CHECK_HIP(hipStreamSynchronize(s), "state save");
    hipLaunchKernelGGL(k_reset_cg, dim3(1), dim3(64), 0, s, c->st);
    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0), "event");
    CHECK_HIP(hipEventCreate(&e1), "event");
    // advection is timed on the advected copy of the last step (adv -> hat -> omega..omega_s, which
    // are contiguous 6-channel scratch); the projection kernels on the solver's own buffers
    auto one = [&](void) -> bool {
        const size_t n3 = c->g.n3;
        switch (which) {
            case VFX_K_ADV_SL:
                launch_adv_pass(0, c->g, 6, c->adv, c->adv + n3, c->prm, c->inv_dx, c->hat, c->qidx(), nullptr,
                                nullptr, nullptr, s);
                return true;
            case VFX_K_ADV_MC:
                launch_adv_pass(1, c->g, 6, c->adv, c->adv + n3, c->prm, c->inv_dx, c->hat, c->qidx(), c->omega,
                                c->P.fm, c->mass_part, s);
                return true;
            case VFX_K_SOURCES: launch_sources_pass(0, c->g, c->adv, c->P.fm, c->prm, c->st, c->hat, s); return true;
            case VFX_K_DIFFUSE: launch_sources_pass(1, c->g, c->hat, c->P.fm, c->prm, c->st, c->omega, s); return true;
            default:
                if (which >= VFX_K_FADV_SL) return hires_launch_single(c, which, s);
                return which >= VFX_K_MATVEC && c->P.launch_single(which - VFX_K_MATVEC, c->hat, s);
        }
    };
    // pass 2 of the advection reads pass 1's trace corners: pass 1 runs first on the same input
    if (which == VFX_K_FADV_MC) hires_launch_single(c, VFX_K_FADV_SL, s);
    if (which == VFX_K_ADV_MC)
        launch_adv_pass(0, c->g, 6, c->adv, c->adv + c->g.n3, c->prm, c->inv_dx, c->hat, c->qidx(), nullptr, nullptr,
                        nullptr, s);
    if (!one()) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        set_error("unknown kernel id (or a level the grid does not have)");
        return VFX_E_VALUE;
    }
    CHECK_HIP(hipEventRecord(e0, s), "event");
    for (int i = 0; i < reps; ++i) one();
    CHECK_HIP(hipEventRecord(e1, s), "event");
    CHECK_LAUNCH("timed kernel");
    CHECK_HIP(hipEventSynchronize(e1), "event");
    float ms = 0.0f;
    CHECK_HIP(hipEventElapsedTime(&ms, e0, e1), "event");
    *ms_avg = ms / (float)reps;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    CHECK_HIP(hipMemcpy(c->st, &saved, sizeof(DevState), hipMemcpyHostToDevice), "state restore");
    return VFX_OK;
}

int VFX_FN(vfx_get_ops)(vfx_ctx* c, const real* solid, int level, real* diag, real* off, int* nlev, int* m,
                void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!solid) return arg_error("null pointer");
    if (nlev) *nlev = c->P.nlev;
    if (level < 0 || level >= c->P.nlev) return arg_error("level out of range");
    if (c->P.mg_kind != MG_REAL) {
        set_error("operators are exported in FP32 level storage only");
        return VFX_E_VALUE;
    }
    const Lvl& L = c->P.lv[level];
    if (m) *m = L.m;
    hipStream_t s = S(stream);
    c->P.setup(solid, s);
    CHECK_LAUNCH("setup");
    const size_t m3 = (size_t)L.m * L.m * L.m;
    if (diag) CHECK_HIP(hipMemcpyAsync(diag, L.diag, m3 * 8 * sizeof(real), hipMemcpyDeviceToDevice, s), "diag copy");
    if (off) CHECK_HIP(hipMemcpyAsync(off, L.off, m3 * 24 * sizeof(real), hipMemcpyDeviceToDevice, s), "off copy");
    return VFX_OK;
}

#if !VFX_F64
int vfx_fcg_history(vfx_ctx* c, double* out, int n, int* count) {
    if (int rc = check_ctx(c)) return rc;
    if (count) *count = (int)c->fcg_hist.size();
    for (int i = 0; out && i < n && i < (int)c->fcg_hist.size(); ++i) out[i] = c->fcg_hist[i];
    return VFX_OK;
}

int vfx_ctx_set_net(vfx_ctx* c, vfx_net* n) {
    if (int rc = check_ctx(c)) return rc;
    if (n && n->net.N != c->N && n->net.N * 4 != c->N)
        return arg_error("the network's grid must be the solver's (the preconditioner) or a quarter of it (the "
                         "coarse correction of the V-cycle)");
    c->net = n;
    return VFX_OK;
}

int vfx_net_train_step(vfx_net* n, vfx_ctx* c, const real* solid, const real* b, const real* phi, float lr,
                       double* loss, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!n || !solid || !b || !phi) return arg_error("null pointer");
    if (n->net.N != c->N) return arg_error("network and solver grid sizes differ");
    hipStream_t s = S(stream);
    c->P.setup(solid, s);
    net_train_step(n->net, c->g, b, phi, c->P.fm, lr, s);
    CHECK_LAUNCH("network training step");
    if (loss) {
        CHECK_HIP(hipMemcpyAsync(loss, n->net.scal + 4, sizeof(double), hipMemcpyDeviceToHost, s), "loss copy");
        CHECK_HIP(hipStreamSynchronize(s), "loss copy");
    }
    return VFX_OK;
}

int vfx_net_train_step_coarse(vfx_net* n, vfx_ctx* c, const real* solid, const real* rv, const real* ev, float lr,
                              double* loss, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!n || !solid || !rv || !ev) return arg_error("null pointer");
    if (n->net.N * 4 != c->N) return arg_error("the coarse correction's grid must be a quarter of the solver's");
    hipStream_t s = S(stream);
    c->P.setup(solid, s);
    net_train_step_coarse(n->net, c->g, rv, ev, c->P.fm, lr, c->hat, s);
    CHECK_LAUNCH("coarse correction training step");
    if (loss) {
        CHECK_HIP(hipMemcpyAsync(loss, n->net.scal + 4, sizeof(double), hipMemcpyDeviceToHost, s), "loss copy");
        CHECK_HIP(hipStreamSynchronize(s), "loss copy");
    }
    return VFX_OK;
}

int vfx_net_apply(vfx_net* n, vfx_ctx* c, const real* solid, const real* r, real* z, void* stream) {
    if (int rc = check_ctx(c)) return rc;
    if (!n || !solid || !r || !z) return arg_error("null pointer");
    if (n->net.N != c->N) return arg_error("network and solver grid sizes differ");
    hipStream_t s = S(stream);
    c->P.setup(solid, s);
    n->net.apply(r, c->P.fm, z, s);
    CHECK_LAUNCH("network preconditioner");
    return VFX_OK;
}
#endif

}  // extern "C"
