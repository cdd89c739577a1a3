// The fused pressure projection (SolverConfig3D.pressure_mode = "fused"): multigrid as a
// stationary iteration on phi, no CG (vfx/sim3d/pressure_fused.py).  `smooth` weighted Jacobi
// sweeps on the exact K (Chebyshev or 0.8 weights), then `cycles` x (phi += T^T V(b - K phi) with
// the same V(1,1) as the MGPCG preconditioner, + `smooth` sweeps); with pressure_tol > 0 more
// cycles while 0.5 max|b - K phi| >= tol max|u*| and fewer than pressure_max_iters cycles have
// run; then u -= g(phi).  The reference's default kernel boundaries (FUSE_TOP off, no level
// fusion, pressure_fused.py:491-499) are the ones restated; the level kernels, the LDS coarse
// solve and the operators are the MGPCG path's.  The adaptive tail enqueues cycle j before the
// test of cycle j - 1 is read back, like the PCG loop; the kernels of cycles after the device's
// stop test are no-ops.  FP32 / FP64 level storage (the 16-bit storage scales the level-1
// right-hand side inside the PCG kernels, which this path does not run).
#include <climits>
#include <cmath>
#include <vector>

#include "ctx.hpp"

namespace vfx {

#define FCELL_LOOP_BEGIN                                                \
    const int x = blockIdx.x * blockDim.x + threadIdx.x;                \
    const int y = blockIdx.y * blockDim.y + threadIdx.y;                \
    const bool live = x < N && y < N;                                   \
    const int z1 = min((int)(blockIdx.z + 1) * zc, N);                  \
    if (live)                                                           \
        for (int cz = blockIdx.z * zc; cz < z1; ++cz) {                 \
            const int i = (cz * N + y) * N + x;
#define FCELL_LOOP_END }

// k_sweep (pressure_fused.py:41-57): dst = src + w D^-1 (b - K src), b = -d(u*); COLD (first sweep
// without warm start): dst = w b / kdiag.  0 in solids; a sealed fluid cell (no open face,
// kdiag = 0) keeps src (0 when COLD) where the reference divides 0 / 0.  Loads unconditional.
template <bool COLD>
__global__ void __launch_bounds__(256) k_fsweep(int N, int zc, const real* __restrict__ src, real* __restrict__ dst,
                                                const real* __restrict__ vel, const uint8_t* __restrict__ fm, real w,
                                                const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    const int n3 = N * N * N;
    FCELL_LOOP_BEGIN
        const int m = fm[i];
        const real b = -div2_xyz(vel, n3, m, N, x, y, cz, i);
        const real kd = kopen(m) ? kdiag(m) : RL(1.0);
        real v;
        if (COLD) {
            v = w * b / kd;
        } else {
            const real s0 = src[i];
            v = s0 + w * (b - kmul_xyz(src, fm, N, x, y, cz, i, m)) / kd;
            v = kopen(m) ? v : s0;
        }
        dst[i] = m == 0 ? RL(0.0) : (kopen(m) ? v : (COLD ? RL(0.0) : v));
    FCELL_LOOP_END
}

// k_resid_phi / resid_max (pressure_fused.py:77-85, :312-330): rr = b - K phi (0 in solids) and
// the block max |rr|; WITHU also the block max |u*| over fluid (k_maxu :296-309).
template <bool WITHU>
__global__ void __launch_bounds__(256) k_fresid(int N, int zc, const real* __restrict__ phi,
                                                const real* __restrict__ vel, const uint8_t* __restrict__ fm,
                                                real* __restrict__ rr, double* __restrict__ part, int nparts,
                                                const DevState* __restrict__ st, int k) {
    __shared__ double sh[8];
    if (cg_inactive(st, k)) return;
    const int n3 = N * N * N;
    double mr = 0.0, mu = 0.0;
    FCELL_LOOP_BEGIN
        const int m = fm[i];
        const real v = -div2_xyz(vel, n3, m, N, x, y, cz, i) - kmul_xyz(phi, fm, N, x, y, cz, i, m);
        const real vo = m != 0 ? v : RL(0.0);
        rr[i] = vo;
        mr = fmax(mr, fabs((double)vo));
        if (WITHU) {
            const real ux = vel[i], uy = vel[n3 + i], uz = vel[2 * n3 + i];
            const real sp = sqrt(ux * ux + uy * uy + uz * uz);
            mu = m != 0 ? fmax(mu, (double)sp) : mu;
        }
    FCELL_LOOP_END
    const double a = block_max(mr, sh);
    const double b = WITHU ? block_max(mu, sh) : 0.0;
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        part[block_linear()] = a;
        if (WITHU) part[nparts + block_linear()] = b;
    }
}

// k_fwd_red (pressure.py:427-441): Hadamard transform of rr to level 1 + the zero-guess red sweep
__global__ void __launch_bounds__(256) k_ffwd_red(int N, const real* __restrict__ rr, Lvl L,
                                                  const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    const int m = L.m;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint32_t)m * m * m) return;
    const uint32_t q = t / (uint32_t)m, I = t - q * (uint32_t)m, K = q / (uint32_t)m, J = q - K * (uint32_t)m;
    real v[8];
#pragma unroll
    for (int s = 0; s < 8; ++s)
        v[s] = rr[((size_t)(2 * K + (s >> 2)) * N + 2 * J + ((s >> 1) & 1)) * N + 2 * I + (s & 1)];
    const bool red = ((I + J + K) & 1) == 0;
    const uint32_t o = lcell(m, (int)K, (int)J, (int)I), xs = lsto(m, (int)K, (int)J, (int)I);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        real acc = RL(0.0);
#pragma unroll
        for (int s = 0; s < 8; ++s) acc += hsign(c, s >> 2, (s >> 1) & 1, s & 1) * v[s];
        L.r[(size_t)xs * 8 + c] = acc;
        if (red) {
            const real d = L.diag[(size_t)o * 8 + c];
            L.x[(size_t)xs * 8 + c] = d <= L.dmin ? RL(0.0) : acc / d;
        }
    }
}

// k_inv_transform / k_inv_add (pressure.py:353-365, pressure_fused.py:271-280): dst = base + T^T x1,
// 0 in solids (per cell, so base == dst is safe)
__global__ void __launch_bounds__(256) k_finv_add(int N, int zc, Lvl L, const uint8_t* __restrict__ fm,
                                                  const real* base, real* dst, const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    const int m = L.m;
    FCELL_LOOP_BEGIN
        const int mk = fm[i];
        const uint32_t xs = lsto(m, cz >> 1, y >> 1, x >> 1);
        real acc = base[i];
#pragma unroll
        for (int c = 0; c < 8; ++c) acc += hsign(c, cz & 1, y & 1, x & 1) * L.x[(size_t)xs * 8 + c];
        dst[i] = mk != 0 ? acc : RL(0.0);
    FCELL_LOOP_END
}

__global__ void __launch_bounds__(256) k_fcopy(size_t n, const real* __restrict__ src, real* __restrict__ dst,
                                               const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// k_test (pressure_fused.py:335-347): n = n_fixed at the first test, else it + 1; continue while
// n < P_MAXIT, 0.5 max|rr| >= P_TOL max|u*| and max|u*| > 0 (float arithmetic, as the reference).
// A stop makes the cycles after k no-ops.  INIT also reduces max|u*| (kept in S_MAXU).
__global__ void __launch_bounds__(FIN_THREADS) k_ftest(const double* __restrict__ part, int nparts,
                                                       const float* __restrict__ prm, DevState* __restrict__ st,
                                                       int n_fixed, int init, int k) {
    __shared__ double sh[FIN_THREADS / 64];
    if (cg_inactive(st, k)) return;
    const double rmax = reduce_parts_max(part, nparts, sh);
    const double umax = init ? reduce_parts_max(part + nparts, nparts, sh) : 0.0;
    if (threadIdx.x != 0) return;
    if (init) st->s[S_MAXU] = umax;
    const float um = (float)st->s[S_MAXU], rm = (float)rmax;
    const int n = init ? n_fixed : st->it + 1;
    st->it = n;
    st->s[S_MAXR] = rmax;
    const bool go = n < (int)prm[P_MAXIT] && 0.5f * rm >= prm[P_TOL] * um && um > 0.0f;
    if (!go) st->stop_at = k;
}

__global__ void k_fstart(DevState* st, int cycles) {
    if (threadIdx.x == 0) {
        st->it = cycles;
        st->stop_at = INT_MAX;
        st->early = 0;
        st->s[S_MAXR] = -1.0;
        st->it_total = 0;
        st->restarts = 0;
        st->hit = 0;
        st->fallback = 0;
    }
}

// sweep weights (chebyshev_weights / weights, pressure_fused.py:348-352, :388-395), as floats
static std::vector<float> fused_weights(int smooth, int jacobi) {
    const int n = smooth < 1 ? 1 : smooth;
    std::vector<float> w(n);
    const double lo = 2.0 / 4.0, hi = 2.0;
    for (int i = 0; i < n; ++i)
        w[i] = jacobi ? 0.8f
                      : (float)(1.0 / ((hi + lo) / 2 + (hi - lo) / 2 * std::cos(M_PI * (2 * i + 1) / (2.0 * n))));
    return w;
}

int run_fused(vfx_ctx* c, real* vel, hipStream_t s, int tail_budget) {
    const bool polled = tail_budget < 0;
    Pressure& P = c->P;
    if (P.mg_kind != MG_REAL) {
        set_error("the fused projection runs with FP32 (or FP64) multigrid levels only");
        return VFX_E_VALUE;
    }
    const Geom& g = c->g;
    const size_t n3 = g.n3;
    const bool warm = c->prm_last[P_WARM] > 0.5f;
    const float tol = c->prm_last[P_TOL];
    const int maxit = (int)c->prm_last[P_MAXIT];
    const int cycles = c->fused_cycles < 0 ? 0 : c->fused_cycles;
    const std::vector<float> ws = fused_weights(c->fused_smooth, c->fused_jacobi);
    real* const home = P.phi;
    real* const alt = P.z;
    real* const rr = P.q;
    const int nb = nblocks(g);
    real* cur = home;
    bool first = true;
    hipLaunchKernelGGL(k_fstart, dim3(1), dim3(64), 0, s, P.st, cycles);
    auto sweeps = [&](int k) {
        for (float w : ws) {
            real* dst = cur == home ? alt : home;
            if (first && !warm)
                hipLaunchKernelGGL(k_fsweep<true>, g.grid, g.block, 0, s, g.N, g.zc, cur, dst, vel, P.fm, (real)w,
                                   P.st, k);
            else
                hipLaunchKernelGGL(k_fsweep<false>, g.grid, g.block, 0, s, g.N, g.zc, cur, dst, vel, P.fm, (real)w,
                                   P.st, k);
            cur = dst;
            first = false;
        }
    };
    const Lvl& L1 = P.lv[0];
    const unsigned fwd_blocks = (unsigned)(((size_t)L1.m * L1.m * L1.m + 255) / 256);
    // V-cycle correction of phi = base -> dst: residual, transform + red sweep, V(1,1), T^T add
    auto vcycle = [&](real* base, real* dst, int k, bool rr_ready) {
        if (!rr_ready)
            hipLaunchKernelGGL(k_fresid<false>, g.grid, g.block, 0, s, g.N, g.zc, base, vel, P.fm, rr, P.part, nb,
                               P.st, k);
        hipLaunchKernelGGL(k_ffwd_red, dim3(fwd_blocks), dim3(256), 0, s, g.N, rr, L1, P.st, k);
        P.vcycle_levels(k, s);
        hipLaunchKernelGGL(k_finv_add, g.grid, g.block, 0, s, g.N, g.zc, L1, P.fm, base, dst, P.st, k);
    };
    sweeps(0);
    for (int cy = 0; cy < cycles; ++cy) {
        real* dst = cur == home ? alt : home;
        vcycle(cur, dst, 0, false);
        cur = dst;
        sweeps(0);
    }
    if (!launch_ok("fused cycles")) return VFX_E_HIP;
    if (tol > 0.0f && maxit > cycles) {
        // adaptive tail: tail cycle j runs as iteration index j; the test after it stops the rest
        hipLaunchKernelGGL(k_fresid<true>, g.grid, g.block, 0, s, g.N, g.zc, cur, vel, P.fm, rr, P.part, nb, P.st, 0);
        hipLaunchKernelGGL(k_ftest, dim3(1), dim3(FIN_THREADS), 0, s, P.part, nb, c->prm, P.st, cycles, 1, 0);
        real* const loop = cur;
        const int tail = polled ? maxit - cycles : std::min(maxit - cycles, tail_budget);
        if (polled && !hip_ok(hipMemcpyAsync(&c->stop_pin[0], &c->st->stop_at, sizeof(int), hipMemcpyDeviceToHost,
                                                s), "stop read"))
            return VFX_E_HIP;
        if (polled && !hip_ok(hipEventRecord(c->stop_ev[0], s), "stop event")) return VFX_E_HIP;
        for (int j = 1; j <= tail; ++j) {
            vcycle(loop, loop, j, true);
            cur = loop;
            sweeps(j);
            if (cur != loop)
                hipLaunchKernelGGL(k_fcopy, dim3((unsigned)std::min<size_t>((n3 + 255) / 256, 65535)), dim3(256), 0, s,
                                   n3, cur, loop, P.st, j);
            cur = loop;
            hipLaunchKernelGGL(k_fresid<false>, g.grid, g.block, 0, s, g.N, g.zc, loop, vel, P.fm, rr, P.part, nb,
                               P.st, j);
            hipLaunchKernelGGL(k_ftest, dim3(1), dim3(FIN_THREADS), 0, s, P.part, nb, c->prm, P.st, cycles, 0, j);
            if (!launch_ok("fused tail")) return VFX_E_HIP;
            if (!polled) continue;   // the whole budget is enqueued (no-ops after the stop)
            const int slot = j % STOP_RING, prev = (j - 1) % STOP_RING;
            if (!hip_ok(hipMemcpyAsync(&c->stop_pin[slot], &c->st->stop_at, sizeof(int), hipMemcpyDeviceToHost, s),
                        "stop read") ||
                !hip_ok(hipEventRecord(c->stop_ev[slot], s), "stop event") ||
                !hip_ok(hipEventSynchronize(c->stop_ev[prev]), "stop wait"))
                return VFX_E_HIP;
            if (c->stop_pin[prev] <= j - 1) break;   // cycle j was enqueued as a no-op
        }
        cur = loop;
    }
    if (cur != home && !hip_ok(hipMemcpyAsync(home, cur, n3 * sizeof(real), hipMemcpyDeviceToDevice, s), "phi copy"))
        return VFX_E_HIP;
    P.apply(vel, s);
    return launch_ok("fused apply") ? VFX_OK : VFX_E_HIP;
}

}  // namespace vfx
