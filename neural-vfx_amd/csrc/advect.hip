// MacCormack advection (K-ADV) and the density mass correction.
//
// Restates vfx/sim3d/kernels.py:66-178 (trilinear helpers, RK2 midpoint trace, advect_sl,
// advect_mc with the 8-corner clamp of Selle et al. 2008) and solver.py:40-76 (fluid density
// sums before / after advection and their ratio).  All channels share one backtrace.
//
// One cell per thread, channel count a template parameter: every gather of a trace step (3
// channels x 8 corners) and of the final samples (C channels x 8 corners) is loaded before any
// of it is used, so each RK2 trace costs three dependent memory rounds, not one per load.
//
// Both passes are bound by the address rate of their gathers (one vector memory instruction costs
// the same for 4 to 16 bytes per lane), so the backward values travel between the passes as
// channel pairs: plane k of hat holds (channel 2k, channel 2k + 1) of every cell as two adjacent
// values, and the MacCormack pass reads both channels of a corner pair with one 16-byte load.
#include "sim3d_kernels.hpp"
#include "trilinear.hpp"

#ifndef MC_PLANES
#define MC_PLANES 1
#endif

namespace vfx {

// Pass 1: semi-Lagrangian backward value of C channels -> hat (channel-pair planes, the second
// value of an odd C's last pair 0), and the base corner of the backward trace -> qidx (pass 2's
// clamp reads the same corners; the trace is not repeated).  hat holds 2 ceil(C / 2) n3 values.
template <int C>
__global__ void __launch_bounds__(256) k_adv_sl(int N, const real* __restrict__ src, const real* __restrict__ vel,
                                                const float* __restrict__ prm, real inv_dx, real* __restrict__ hat,
                                                int* __restrict__ qidx) {
    const int n3 = N * N * N;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= N || y >= N) return;
    const real s = (real)prm[P_DT] * inv_dx;
    const int i = (z * N + y) * N + x;
    real q[3];
    trace1(vel, N, n3, i, (real)x, (real)y, (real)z, -s, q);
    const Corner c = corner(N, q[0], q[1], q[2]);
    qidx[i] = c.o;
    Oct g[C];
#pragma unroll
    for (int ch = 0; ch < C; ++ch) g[ch] = gather(src + ch * n3, N, c.o);
    real v[C + 1];
#pragma unroll
    for (int ch = 0; ch < C; ++ch) v[ch] = lerp3(g[ch], c);
    v[C] = RL(0.0);
#pragma unroll
    for (int k = 0; 2 * k < C; ++k) {
        real2_e w;
        w.x = v[2 * k];
        w.y = v[2 * k + 1];
        *reinterpret_cast<real2_e*>(hat + 2 * ((size_t)k * n3 + i)) = w;
    }
}

// Pass 2: MacCormack correction with the 8-corner clamp around the backward trace's base corner
// (qidx, from pass 1; clamped into the box so that any content stays in bounds); optionally the
// block partial sums of fluid density before (src[0]) and after (out[0]) advection
// (solver.py:46-65), in double.
template <int C>
__global__ void __launch_bounds__(256) k_adv_mc(int N, const real* __restrict__ src, const real* __restrict__ vel,
                                                const float* __restrict__ prm, real inv_dx,
                                                const real* __restrict__ hat, const int* __restrict__ qidx,
                                                real* __restrict__ out, const uint8_t* __restrict__ fm,
                                                double* __restrict__ part, int nparts) {
    __shared__ double sh[2 * 8];
    const int n3 = N * N * N;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    const bool live = x < N && y < N;
    double sa = 0.0, sb = 0.0;
    if (live) {
        const real s = (real)prm[P_DT] * inv_dx;
        const int i = (z * N + y) * N + x;
        const int qo = min(max(qidx[i], 0), n3 - N * N - N - 2);
        real r[3];
        trace1(vel, N, n3, i, (real)x, (real)y, (real)z, s, r);
        const Corner cr = corner(N, r[0], r[1], r[2]);
        const int fmi = part != nullptr ? (int)fm[i] : 0;
        // channel pairs in groups of MC_PLANES planes: the loads of one group are in flight
        // together, and a group's registers are free before the next group loads
        constexpr int NP = (C + 1) / 2;
#pragma unroll
        for (int p0 = 0; p0 < NP; p0 += MC_PLANES) {
        Oct gs[2 * MC_PLANES], gh[2 * MC_PLANES];
        real h[2 * MC_PLANES], s0[2 * MC_PLANES];
#pragma unroll
        for (int k = 0; k < MC_PLANES; ++k) {
            const int pl = p0 + k < NP ? p0 + k : NP - 1;
            const real* hp = hat + 2 * (size_t)pl * n3;
            gather_pair(hp, N, cr.o, gh[2 * k], gh[2 * k + 1]);
            const real2_e hv = *reinterpret_cast<const real2_e*>(hp + 2 * i);
            h[2 * k] = hv.x;
            h[2 * k + 1] = hv.y;
            const int c0 = 2 * pl, c1 = 2 * pl + 1 < C ? 2 * pl + 1 : 2 * pl;
            gs[2 * k] = gather(src + c0 * n3, N, qo);
            s0[2 * k] = src[c0 * n3 + i];
            if (2 * pl + 1 < C) {
                gs[2 * k + 1] = gather(src + c1 * n3, N, qo);
                s0[2 * k + 1] = src[c1 * n3 + i];
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * MC_PLANES; ++k) {
            const int ch = 2 * p0 + k;
            if (ch >= C) break;
            const real bar = lerp3(gh[k], cr);
            real val = h[k] + RL(0.5) * (s0[k] - bar);
            // wp.min / wp.max over the 8 source values (kernels.py:171-177); lo and hi are only
            // compared with val, so the sign of a zero bound does not matter
            real lo = gs[k].v[0], hi = lo;
#pragma unroll
            for (int j = 1; j < 8; ++j) {
                lo = fmin(lo, gs[k].v[j]);
                hi = fmax(hi, gs[k].v[j]);
            }
            if (val < lo || val > hi) val = h[k];
            out[ch * n3 + i] = val;
            if (ch == 0 && fmi != 0) {
                sa = (double)s0[0];
                sb = (double)val;
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        }
    }
    if (part != nullptr) {
        double v2[2] = {sa, sb};
        block_reduce<2, 0u>(v2, sh);
        const double a = v2[0], b = v2[1];
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            const int bl = block_linear();
            part[bl] = a;
            part[nparts + bl] = b;
        }
    }
}

// First reduction level: block b sums partials [b*per, (b+1)*per) of both halves.
__global__ void __launch_bounds__(256) k_mass_part2(const double* __restrict__ part, int nparts, int per,
                                                    double* __restrict__ part2, int n2) {
    __shared__ double sh[2 * 4];
    const int b = blockIdx.x;
    const int lo = b * per, hi = min(lo + per, nparts);
    double a = 0.0, c = 0.0;
    for (int i = lo + (int)threadIdx.x; i < hi; i += 256) {
        a += part[i];
        c += part[nparts + i];
    }
    double v2[2] = {a, c};
    block_reduce<2, 0u>(v2, sh);
    const double ra = v2[0], rc = v2[1];
    if (threadIdx.x == 0) {
        part2[b] = ra;
        part2[n2 + b] = rc;
    }
}

// ratio = a / b if a > 0 and b > 1e-30, else 1 (solver.py:67-74).
__global__ void __launch_bounds__(FIN_THREADS) k_mass_fin(const double* __restrict__ part, int nparts,
                                                          DevState* __restrict__ st, int slot) {
    __shared__ double sh[2 * FIN_THREADS / 64];
    // both sums in one pass, each in reduce_parts_sum's order
    double v2[2] = {0.0, 0.0};
    for_parts<2, 16>(part, nparts, {0, nparts}, [&](const double (&x)[2]) {
        v2[0] += x[0];
        v2[1] += x[1];
    });
    block_reduce<2, 0u>(v2, sh);
    const double a = v2[0], b = v2[1];
    if (threadIdx.x == 0) st->s[slot] = (a > 0.0 && b > 1e-30) ? a / b : 1.0;
}

static dim3 cell_grid(const Geom& g) { return dim3(g.grid.x, g.grid.y, (unsigned)g.N); }

int advect_parts(const Geom& g) { return (int)(g.grid.x * g.grid.y) * g.N; }

template <int C>
static void launch_pass_t(int pass, const Geom& g, const real* src, const real* vel, const float* prm, real inv_dx,
                          real* hat, int* qidx, real* out, const uint8_t* fm, double* part, hipStream_t s) {
    if (pass == 0)
        hipLaunchKernelGGL(k_adv_sl<C>, cell_grid(g), g.block, 0, s, g.N, src, vel, prm, inv_dx, hat, qidx);
    else
        hipLaunchKernelGGL(k_adv_mc<C>, cell_grid(g), g.block, 0, s, g.N, src, vel, prm, inv_dx, hat, qidx, out, fm,
                           part, advect_parts(g));
}

void launch_adv_pass(int pass, const Geom& g, int C, const real* src, const real* vel, const float* prm,
                     real inv_dx, real* hat, int* qidx, real* out, const uint8_t* fm, double* part, hipStream_t s) {
    switch (C) {
        case 1: launch_pass_t<1>(pass, g, src, vel, prm, inv_dx, hat, qidx, out, fm, part, s); break;
        case 2: launch_pass_t<2>(pass, g, src, vel, prm, inv_dx, hat, qidx, out, fm, part, s); break;
        case 3: launch_pass_t<3>(pass, g, src, vel, prm, inv_dx, hat, qidx, out, fm, part, s); break;
        case 4: launch_pass_t<4>(pass, g, src, vel, prm, inv_dx, hat, qidx, out, fm, part, s); break;
        case 5: launch_pass_t<5>(pass, g, src, vel, prm, inv_dx, hat, qidx, out, fm, part, s); break;
        default: launch_pass_t<6>(pass, g, src, vel, prm, inv_dx, hat, qidx, out, fm, part, s); break;
    }
}

void launch_advect(const Geom& g, int C, const real* src, const real* vel, const float* prm, real inv_dx,
                   real* hat, int* qidx, real* out, const uint8_t* fm, double* part, hipStream_t s) {
    launch_adv_pass(0, g, C, src, vel, prm, inv_dx, hat, qidx, out, fm, part, s);
    launch_adv_pass(1, g, C, src, vel, prm, inv_dx, hat, qidx, out, fm, part, s);
}

// part holds 2 * advect_parts(g) partials; part2 (after them) 2 * MASS_PART2 partials.
constexpr int MASS_PART2 = 256;

void launch_mass_fin(const Geom& g, double* part, DevState* st, hipStream_t s, int slot) {
    const int n = advect_parts(g);
    const int per = (n + MASS_PART2 - 1) / MASS_PART2;
    double* part2 = part + 2 * (size_t)n;
    hipLaunchKernelGGL(k_mass_part2, dim3(MASS_PART2), dim3(256), 0, s, part, n, per, part2, MASS_PART2);
    hipLaunchKernelGGL(k_mass_fin, dim3(1), dim3(FIN_THREADS), 0, s, part2, MASS_PART2, st, slot);
}

size_t mass_part_doubles(const Geom& g) { return 2 * (size_t)advect_parts(g) + 2 * (size_t)MASS_PART2; }

}  // namespace vfx
