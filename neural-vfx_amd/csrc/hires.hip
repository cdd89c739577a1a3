// The finer scalar grid (Solver3D(res, scalar_res=S)): density, temperature and fuel carried on
// an S^3 grid (same world box, fine cell size 1/S) by the res^3 velocity, then box-restricted to
// res^3 before the force chain reads them.  Velocity, forces and the projection stay at res^3.
//
// Restates vfx/sim3d/hires.py (f2c :23-26, advect_sl_fine :29-48, advect_mc_fine :51-89,
// sources_fine :92-139, nbv / diffuse_fine :142-170, ovl / restrict_scalars :173-219) in the
// launch order of solver.py:241-264, and Solver3D.init_hires_from_state (solver.py:204-215,
// a trilinear resize with half-pixel centres).
//
// Fine fields: (3, S, S, S) `real`, channels density, temperature, fuel, owned by the caller; the
// fine solid (S, S, S) `real`, 1 = obstacle, also the caller's.  Traces run in the coarse grid's
// cell-centre index space (p_c = a p_f + b, a = res / S, b = 0.5 res / S - 0.5), so the coarse and
// fine backtraces are one curve in world space.  Every fine kernel is one cell per thread;
// indices are 32-bit (S <= 1024, S^3 <= 2^30) and channel offsets 64-bit.
#include <algorithm>
#include <cmath>

#include "ctx.hpp"
#include "sources_ops.hpp"
#include "trilinear.hpp"

namespace vfx {

constexpr int HIRES_MAX = 1024;   // S^3 indices stay 32-bit

// coarse index -> fine index, (p_c - b) / a (hires.py:36, :61-62).  POW2: a = res / S is a power of
// two, so the quotient is exactly (p_c - b) * ia with ia = 1 / a = S / res, and the multiply
// replaces the division.
template <bool POW2>
__device__ __forceinline__ real to_fine(real pc, real a, real ia, real b) {
    if constexpr (POW2) return (pc - b) * ia;
    else return (pc - b) / a;
}

// The first sample of a fine cell's trace is at a fixed coarse point (p_c = a p_f + b), so the
// coarse velocity a workgroup's first samples read is known before the trace: its corner range
// per axis runs from the first cell's base corner to the last cell's plus one (base_idx is
// monotonic).  It is staged in LDS with coalesced row loads, and the first sample reads its
// 8 corners x 3 channels there instead of with 12 gathers.  VT_*: the largest range a <= 1
// (S >= res) needs for the fine workgroups (make_geom: 64 x 4 or 32 x 8 cells, one plane): at
// most blockDim + 1 corners per axis.
constexpr int VT_X = 66, VT_Y = 10, VT_Z = 2, VT_N = VT_X * VT_Y * VT_Z;

struct VelTile {
    int x0, y0, z0, nx, ny;
};

// every thread of the workgroup calls this (it ends with a barrier)
__device__ __forceinline__ VelTile stage_vel(const real* __restrict__ vel, int N, int S, real a, real b,
                                             real* __restrict__ vt) {
    const int fx0 = blockIdx.x * blockDim.x, fy0 = blockIdx.y * blockDim.y, fz = blockIdx.z;
    const int fx1 = min(fx0 + (int)blockDim.x, S) - 1, fy1 = min(fy0 + (int)blockDim.y, S) - 1;
    VelTile t;
    t.x0 = base_idx(a * (real)fx0 + b, N);
    t.y0 = base_idx(a * (real)fy0 + b, N);
    t.z0 = base_idx(a * (real)fz + b, N);
    t.nx = base_idx(a * (real)fx1 + b, N) + 2 - t.x0;
    t.ny = base_idx(a * (real)fy1 + b, N) + 2 - t.y0;
    const int n3 = N * N * N, per = t.nx * t.ny * VT_Z;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x, nt = blockDim.x * blockDim.y;
    for (int k = tid; k < 3 * per; k += nt) {
        const int ch = k / per, r = k - ch * per;
        const int lz = r / (t.nx * t.ny), r2 = r - lz * t.nx * t.ny, ly = r2 / t.nx, lx = r2 - ly * t.nx;
        vt[ch * VT_N + (lz * VT_Y + ly) * VT_X + lx] = vel[ch * n3 + ((t.z0 + lz) * N + t.y0 + ly) * N + t.x0 + lx];
    }
    __syncthreads();
    return t;
}

// trace_pt (kernels.py:115-120) with the first velocity sample from the staged tile (the same
// corner values as sample_vel gathers, so the same result)
__device__ __forceinline__ void trace_pt_tile(const real* __restrict__ vel, int N, int n3, const real* __restrict__ vt,
                                              const VelTile& t, real px, real py, real pz, real s, real* q) {
    const Corner c0 = corner(N, px, py, pz);
    const int x0 = c0.o % N, r = c0.o / N, y0 = r % N, z0 = r / N;
    const int l = ((z0 - t.z0) * VT_Y + (y0 - t.y0)) * VT_X + (x0 - t.x0);
    real v1[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const real* p = vt + ch * VT_N + l;
        Oct g;
        g.v[0] = p[0]; g.v[1] = p[1];
        g.v[2] = p[VT_X]; g.v[3] = p[VT_X + 1];
        g.v[4] = p[VT_X * VT_Y]; g.v[5] = p[VT_X * VT_Y + 1];
        g.v[6] = p[VT_X * VT_Y + VT_X]; g.v[7] = p[VT_X * VT_Y + VT_X + 1];
        v1[ch] = lerp3(g, c0);
    }
    const real h = RL(0.5) * s;
    const Corner c1 = corner(N, px + h * v1[0], py + h * v1[1], pz + h * v1[2]);
    real v2[3];
    sample_vel(vel, N, n3, c1, v2);
    q[0] = px + s * v2[0];
    q[1] = py + s * v2[1];
    q[2] = pz + s * v2[2];
}

// Pass 1 (advect_sl_fine): the backward trace through the coarse velocity from the fine cell's
// centre, mapped to fine index space; the semi-Lagrangian value of the 3 channels -> hat (channels
// 0 and 1 as one pair plane, channel 2 after it, for the pair gathers of pass 2), the
// base corner -> qidx (pass 2's clamp), and the fine fluid flag (solid <= 0.5) -> ffm for the
// later passes.
template <bool POW2>
__global__ void __launch_bounds__(256) k_fadv_sl(int S, int N, const real* __restrict__ src,
                                                 const real* __restrict__ vel, const real* __restrict__ hsolid,
                                                 const float* __restrict__ prm, real inv_dx, real a, real ia, real b,
                                                 real* __restrict__ hat, int* __restrict__ qidx,
                                                 uint8_t* __restrict__ ffm) {
    __shared__ real vt[3 * VT_N];
    const size_t n3 = (size_t)S * S * S;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    const VelTile t = stage_vel(vel, N, S, a, b, vt);
    if (x >= S || y >= S) return;
    const int i = (z * S + y) * S + x;
    const real s = (real)prm[P_DT] * inv_dx;
    const real so = hsolid[i];
    real qc[3];
    trace_pt_tile(vel, N, N * N * N, vt, t, a * (real)x + b, a * (real)y + b, a * (real)z + b, -s, qc);
    const Corner c = corner(S, to_fine<POW2>(qc[0], a, ia, b), to_fine<POW2>(qc[1], a, ia, b),
                            to_fine<POW2>(qc[2], a, ia, b));
    Oct g[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) g[ch] = gather(src + ch * n3, S, c.o);
    qidx[i] = c.o;
    ffm[i] = so > RL(0.5) ? 0 : 1;
    // hat: channels 0 and 1 as one pair plane (2 n3 values), channel 2 alone after it
    real2_e w;
    w.x = lerp3(g[0], c);
    w.y = lerp3(g[1], c);
    *reinterpret_cast<real2_e*>(hat + 2 * (size_t)i) = w;
    hat[2 * n3 + i] = lerp3(g[2], c);
}

// Pass 2 (advect_mc_fine): MacCormack correction with the 8-corner clamp around the backward
// trace's base corner (qidx, clamped into the box so that any content stays in bounds), and the
// block partial sums of fine fluid density before (src) and after (out) the advection, in double
// (the fine grid's own _make_mass_kernels pair, solver.py:250-253).
template <bool POW2>
__global__ void __launch_bounds__(256) k_fadv_mc(int S, int N, const real* __restrict__ src,
                                                 const real* __restrict__ vel, const float* __restrict__ prm,
                                                 real inv_dx, real a, real ia, real b, const real* __restrict__ hat,
                                                 const int* __restrict__ qidx, const uint8_t* __restrict__ ffm,
                                                 real* __restrict__ out, double* __restrict__ part, int nparts) {
    __shared__ double sh[2 * 8];
    __shared__ real vt[3 * VT_N];
    const size_t n3 = (size_t)S * S * S;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    const VelTile t = stage_vel(vel, N, S, a, b, vt);
    double sa = 0.0, sb = 0.0;
    if (x < S && y < S) {
        const int i = (z * S + y) * S + x;
        const real s = (real)prm[P_DT] * inv_dx;
        const int qo = min(max(qidx[i], 0), S * S * S - S * S - S - 2);
        const int fl = ffm[i];
        real rc[3];
        trace_pt_tile(vel, N, N * N * N, vt, t, a * (real)x + b, a * (real)y + b, a * (real)z + b, s, rc);
        const Corner cr = corner(S, to_fine<POW2>(rc[0], a, ia, b), to_fine<POW2>(rc[1], a, ia, b),
                                    to_fine<POW2>(rc[2], a, ia, b));
        Oct gs[3], gh[3];
        real h[3], s0[3];
        gather_pair(hat, S, cr.o, gh[0], gh[1]);
        gh[2] = gather(hat + 2 * n3, S, cr.o);
        const real2_e hp = *reinterpret_cast<const real2_e*>(hat + 2 * (size_t)i);
        h[0] = hp.x;
        h[1] = hp.y;
        h[2] = hat[2 * n3 + i];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            gs[ch] = gather(src + ch * n3, S, qo);
            s0[ch] = src[ch * n3 + i];
        }
        real v0 = RL(0.0);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const real bar = lerp3(gh[ch], cr);
            real val = h[ch] + RL(0.5) * (s0[ch] - bar);
            real lo = gs[ch].v[0], hi = lo;
#pragma unroll
            for (int j = 1; j < 8; ++j) {
                lo = fmin(lo, gs[ch].v[j]);
                hi = fmax(hi, gs[ch].v[j]);
            }
            if (val < lo || val > hi) val = h[ch];
            out[ch * n3 + i] = val;
            if (ch == 0) v0 = val;
        }
        if (fl != 0) {
            sa = (double)s0[0];
            sb = (double)v0;
        }
    }
    double v2[2] = {sa, sb};
    block_reduce<2, 0u>(v2, sh);
    const double ra = v2[0], rb = v2[1];
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const int bl = block_linear();
        part[bl] = ra;
        part[nparts + bl] = rb;
    }
}

// sources_fine: density (times the fine mass ratio), temperature and fuel at the fine cell
// centre; 0 / ambient / 0 in fine solids.  Loads are issued before the solid test selects.
// Density goes to out_d (its final place: the diffusion that follows changes only temperature and
// fuel), temperature and fuel to channels 1, 2 of `out`.
__global__ void __launch_bounds__(256) k_fsources(int S, const real* __restrict__ adv, const uint8_t* __restrict__ ffm,
                                                  const float* __restrict__ prm, const DevState* __restrict__ st,
                                                  real dxf, real* __restrict__ out_d, real* __restrict__ out) {
    const size_t n3 = (size_t)S * S * S;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= S || y >= S) return;
    const int i = (z * S + y) * S + x;
    const int fl = ffm[i];
    const real a0 = adv[i], a1 = adv[n3 + i], a2 = adv[2 * n3 + i];
    const SrcConst k(prm);
    if (fl == 0) {
        out_d[i] = RL(0.0);
        out[n3 + i] = k.amb;
        out[2 * n3 + i] = RL(0.0);
        return;
    }
    const real px = ((real)x + RL(0.5)) * dxf, py = ((real)y + RL(0.5)) * dxf, pz = ((real)z + RL(0.5)) * dxf;
    real d = a0 * (real)st->s[S_MASS_RATIO_F], T = a1, Fu = a2, vx = RL(0.0), vy = RL(0.0), vz = RL(0.0);
    cell_sources<true, false>(prm, k, px, py, pz, (uint32_t)i, d, T, Fu, vx, vy, vz);
    out_d[i] = d;
    out[n3 + i] = T;
    out[2 * n3 + i] = Fu;
}

// diffuse_fine: temperature and fuel take one explicit 7-point substep (a missing or solid
// neighbour contributes the cell's own value, nbv); density, which the reference copies through,
// is already in place (k_fsources).
__global__ void __launch_bounds__(256) k_fdiffuse(int S, const real* __restrict__ src, const uint8_t* __restrict__ ffm,
                                                  const float* __restrict__ prm, real nsub, real dxf,
                                                  real* __restrict__ dst) {
    const size_t n3 = (size_t)S * S * S;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= S || y >= S) return;
    real lam = (real)prm[P_KAPPA] * (real)prm[P_DT] / (nsub * dxf * dxf);
    lam = lam < LAM_CAP ? lam : LAM_CAP;
    const int i = (z * S + y) * S + x;
    const Nb6 n = nb6(S, x, y, z, i);
    const int m0 = ffm[i], mxm = ffm[n.xm], mxp = ffm[n.xp], mym = ffm[n.ym], myp = ffm[n.yp], mzm = ffm[n.zm],
              mzp = ffm[n.zp];
    const bool sxm = !n.oxm || mxm == 0, sxp = !n.oxp || mxp == 0, sym = !n.oym || mym == 0;
    const bool syp = !n.oyp || myp == 0, szm = !n.ozm || mzm == 0, szp = !n.ozp || mzp == 0;
    real o[2];
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
        const real* f = src + (1 + cc) * n3;
        const real v = f[i], a0 = f[n.xm], a1 = f[n.xp], a2 = f[n.ym], a3 = f[n.yp], a4 = f[n.zm], a5 = f[n.zp];
        const real sum = (sxm ? v : a0) + (sxp ? v : a1) + (sym ? v : a2) + (syp ? v : a3) + (szm ? v : a4) +
                         (szp ? v : a5);
        o[cc] = (m0 == 0 || lam <= RL(0.0)) ? v : v + lam * (sum - RL(6.0) * v);
    }
    dst[n3 + i] = o[0];
    dst[2 * n3 + i] = o[1];
}

// ovl (hires.py:173-176): length of [j, j+1) intersected with [i ratio, (i+1) ratio), in fine cells
__device__ __forceinline__ real ovl(int i, int j, real ratio) {
    const real hi_a = (real)(j + 1), hi_b = (real)(i + 1) * ratio;
    const real lo_a = (real)j, lo_b = (real)i * ratio;
    const real len = (hi_a < hi_b ? hi_a : hi_b) - (lo_a > lo_b ? lo_a : lo_b);
    return len > RL(0.0) ? len : RL(0.0);
}

// restrict_scalars: coarse density / temperature / fuel = the overlap-weighted mean of the fine
// cells a coarse cell covers (fine solids count with their 0 / ambient / 0); coarse solids get
// 0 / ambient / 0.  Written into channels 0, 4, 5 of the coarse (6, N, N, N) array `out`.
// R > 0: S = R N, where the overlap weights are exactly 1 on the R^3 fine cells of the box and 0
// elsewhere in the span, so the sums over those cells in the same (z, y, x) order are the
// reference's sums bit for bit; R = 0: any ratio, the reference's weighted loop over the span.
template <int R>
__global__ void __launch_bounds__(256) k_frestrict(int N, int S, const real* __restrict__ fine,
                                                   const uint8_t* __restrict__ fm, const float* __restrict__ prm,
                                                   real ratio, int span, real* __restrict__ out) {
    const size_t fn3 = (size_t)S * S * S, cn3 = (size_t)N * N * N;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= N || y >= N) return;
    const int i = (z * N + y) * N + x;
    const int m = fm[i];
    real sw = RL(0.0), sd = RL(0.0), sT = RL(0.0), sf = RL(0.0);
    if constexpr (R > 0) {
        // row by row in (z, y) order, x innermost: the reference's summation order
#pragma unroll
        for (int k = 0; k < R; ++k)
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int o = ((z * R + k) * S + (y * R + j)) * S + x * R;
                real v[3][R];
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    const real* f = fine + ch * fn3 + o;
                    if constexpr (R == 2) {
                        const real2_e p = *reinterpret_cast<const real2_e*>(f);
                        v[ch][0] = p.x;
                        v[ch][1] = p.y;
                    } else {
#pragma unroll
                        for (int ii = 0; ii < R; ++ii) v[ch][ii] = f[ii];
                    }
                }
#pragma unroll
                for (int ii = 0; ii < R; ++ii) {
                    sw += RL(1.0);
                    sd += v[0][ii];
                    sT += v[1][ii];
                    sf += v[2][ii];
                }
            }
    } else {
        const int z0 = (int)floor((real)z * ratio), y0 = (int)floor((real)y * ratio), x0 = (int)floor((real)x * ratio);
        for (int k = 0; k < span; ++k) {
            const int zz = z0 + k;
            const real wz = zz < S ? ovl(z, zz, ratio) : RL(0.0);
            for (int j = 0; j < span; ++j) {
                const int yy = y0 + j;
                const real wy = yy < S ? ovl(y, yy, ratio) : RL(0.0);
                for (int ii = 0; ii < span; ++ii) {
                    const int xx = x0 + ii;
                    const real wx = xx < S ? ovl(x, xx, ratio) : RL(0.0);
                    const real w = wz * wy * wx;
                    if (w > RL(0.0)) {
                        const size_t o = ((size_t)zz * S + yy) * S + xx;
                        sw += w;
                        sd += w * fine[o];
                        sT += w * fine[fn3 + o];
                        sf += w * fine[2 * fn3 + o];
                    }
                }
            }
        }
    }
    const real amb = (real)prm[P_AMB];
    const bool keep = m != 0 && sw > RL(0.0);
    out[i] = keep ? sd / sw : RL(0.0);
    out[4 * cn3 + i] = keep ? sT / sw : amb;
    out[5 * cn3 + i] = keep ? sf / sw : RL(0.0);
}

// init_hires_from_state (solver.py:204-215): the coarse density / temperature / fuel resized to
// the fine grid trilinearly with half-pixel centres (source index max(scale (d + 0.5) - 0.5, 0),
// scale = res / S, upper neighbour clamped to the last cell), then 0 / ambient / 0 where the fine
// solid is >= 0.5.
__global__ void __launch_bounds__(256) k_fupsample(int S, int N, const real* __restrict__ data,
                                                   const real* __restrict__ hsolid, real amb, real scale,
                                                   real* __restrict__ hi) {
    const size_t fn3 = (size_t)S * S * S, cn3 = (size_t)N * N * N;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= S || y >= S) return;
    int i0[3], i1[3];
    real l1[3];
    const int d[3] = {x, y, z};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        real sidx = scale * ((real)d[ax] + RL(0.5)) - RL(0.5);
        sidx = sidx < RL(0.0) ? RL(0.0) : sidx;
        i0[ax] = (int)sidx;
        i0[ax] = i0[ax] < N - 1 ? i0[ax] : N - 1;
        i1[ax] = i0[ax] + (i0[ax] < N - 1 ? 1 : 0);
        l1[ax] = sidx - (real)i0[ax];
    }
    const int fi = (z * S + y) * S + x;
    const bool fluid = hsolid[fi] < RL(0.5);
    const int chans[3] = {0, 4, 5};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const real* f = data + chans[c] * cn3;
        auto at = [&](int zz, int yy, int xx) { return f[((size_t)zz * N + yy) * N + xx]; };
        const real lx0 = RL(1.0) - l1[0], ly0 = RL(1.0) - l1[1], lz0 = RL(1.0) - l1[2];
        const real r00 = lx0 * at(i0[2], i0[1], i0[0]) + l1[0] * at(i0[2], i0[1], i1[0]);
        const real r01 = lx0 * at(i0[2], i1[1], i0[0]) + l1[0] * at(i0[2], i1[1], i1[0]);
        const real r10 = lx0 * at(i1[2], i0[1], i0[0]) + l1[0] * at(i1[2], i0[1], i1[0]);
        const real r11 = lx0 * at(i1[2], i1[1], i0[0]) + l1[0] * at(i1[2], i1[1], i1[0]);
        const real v = lz0 * (ly0 * r00 + l1[1] * r01) + l1[2] * (ly0 * r10 + l1[1] * r11);
        hi[c * fn3 + fi] = fluid ? v : (c == 1 ? amb : RL(0.0));
    }
}

static dim3 fine_grid(const Geom& g) { return dim3(g.grid.x, g.grid.y, (unsigned)g.N); }

// the restriction's launch: one coarse cell per thread, z = blockIdx.z
static void restrict_launch(const HiresGrid& h, int N, const uint8_t* fm, const float* prm, real* out, hipStream_t s) {
    const Geom gc = make_geom(N, 1);
    const dim3 grid(gc.grid.x, gc.grid.y, (unsigned)N);
    const int R = h.S % N == 0 ? h.S / N : 0;
    switch (R) {
        case 1: hipLaunchKernelGGL(k_frestrict<1>, grid, gc.block, 0, s, N, h.S, h.hi, fm, prm, h.ratio, h.span, out); break;
        case 2: hipLaunchKernelGGL(k_frestrict<2>, grid, gc.block, 0, s, N, h.S, h.hi, fm, prm, h.ratio, h.span, out); break;
        case 3: hipLaunchKernelGGL(k_frestrict<3>, grid, gc.block, 0, s, N, h.S, h.hi, fm, prm, h.ratio, h.span, out); break;
        case 4: hipLaunchKernelGGL(k_frestrict<4>, grid, gc.block, 0, s, N, h.S, h.hi, fm, prm, h.ratio, h.span, out); break;
        default: hipLaunchKernelGGL(k_frestrict<0>, grid, gc.block, 0, s, N, h.S, h.hi, fm, prm, h.ratio, h.span, out); break;
    }
}

void launch_restrict(vfx_ctx* c, real* out, hipStream_t s) { restrict_launch(c->hr, c->N, c->P.fm, c->prm, out, s); }

// one launch of a fine-stage pass (0 SL, 1 MC, 2 sources, 3 one diffusion substep) from `src`
// with the coarse velocity `vel` into the pass's usual destination `dst` (the sources' density
// into dst_d when given)
static void fine_pass(vfx_ctx* c, int pass, const real* src, const real* vel, real* dst, hipStream_t s,
                      real* dst_d = nullptr) {
    const HiresGrid& h = c->hr;
    const Geom& g = h.g;
    switch (pass) {
        case 0:
            hipLaunchKernelGGL(h.pow2 ? k_fadv_sl<true> : k_fadv_sl<false>, fine_grid(g), g.block, 0, s, h.S, c->N,
                               src, vel, h.hsolid, c->prm, c->inv_dx, h.a, h.ia, h.b, dst, h.qidx, h.ffm);
            break;
        case 1:
            hipLaunchKernelGGL(h.pow2 ? k_fadv_mc<true> : k_fadv_mc<false>, fine_grid(g), g.block, 0, s, h.S, c->N,
                               src, vel, c->prm, c->inv_dx, h.a, h.ia, h.b, h.hat, h.qidx, h.ffm, dst, h.part,
                               advect_parts(g));
            break;
        case 2:
            hipLaunchKernelGGL(k_fsources, fine_grid(g), g.block, 0, s, h.S, src, h.ffm, c->prm, c->st, h.dxf,
                               dst_d ? dst_d : dst, dst);
            break;
        default:
            hipLaunchKernelGGL(k_fdiffuse, fine_grid(g), g.block, 0, s, h.S, src, h.ffm, c->prm, (real)h.nd, h.dxf,
                               dst);
            break;
    }
}

// _record_hires (solver.py:241-259): SL and MC passes by the coarse velocity, the fine mass
// ratio, sources, then nd diffusion substeps ping-ponging so that the last one ends in hi.
int stage_scalars(vfx_ctx* c, const real* data, hipStream_t s) {
    HiresGrid& h = c->hr;
    if (h.S == 0) return VFX_OK;
    const real* vel = data + c->g.n3;
    fine_pass(c, 0, h.hi, vel, h.hat, s);
    fine_pass(c, 1, h.hi, vel, h.adv, s);
    launch_mass_fin(h.g, h.part, c->st, s, S_MASS_RATIO_F);
    real* b0 = h.nd % 2 == 0 ? h.hi : h.hat;
    real* b1 = h.nd % 2 == 0 ? h.hat : h.hi;
    fine_pass(c, 2, h.adv, nullptr, b0, s, h.hi);
    for (int k = 0; k < h.nd; ++k) fine_pass(c, 3, k % 2 == 0 ? b0 : b1, nullptr, k % 2 == 0 ? b1 : b0, s);
    if (!launch_ok("fine scalars")) return VFX_E_HIP;
    return VFX_OK;
}

void hires_free(vfx_ctx* c) {
    if (c->hr.pool) {
        (void)hipDeviceSynchronize();
        (void)hipFree(c->hr.pool);
    }
    c->hr = HiresGrid{};
}

// Timed launches of the fine kernels (vfx_time_kernel): on the fine scratch and the fine field of
// the last step, with the advected coarse velocity; nothing the next step reads is changed (the
// SL pass rewrites the fluid flags with their own values, the restriction writes into `hat`).
bool hires_launch_single(vfx_ctx* c, int which, hipStream_t s) {
    const HiresGrid& h = c->hr;
    if (h.S == 0) return false;
    const real* vel = c->adv + c->g.n3;
    switch (which) {
        case VFX_K_FADV_SL: fine_pass(c, 0, h.hi, vel, h.hat, s); return true;
        case VFX_K_FADV_MC: fine_pass(c, 1, h.hi, vel, h.adv, s); return true;
        case VFX_K_FSOURCES: fine_pass(c, 2, h.adv, nullptr, h.hat, s); return true;
        case VFX_K_FDIFFUSE: fine_pass(c, 3, h.hat, nullptr, h.adv, s); return true;
        case VFX_K_FRESTRICT: launch_restrict(c, c->hat, s); return true;
        default: return false;
    }
}

}  // namespace vfx

extern "C" {

int VFX_FN(vfx_hires_attach)(vfx_ctx* c, int S, real* hi, const real* hsolid) {
    if (c == nullptr) return arg_error("null context");
    if (S == 0) {
        drop_graphs(c);
        hires_free(c);
        return VFX_OK;
    }
    if (S < c->N) {
        set_error("scalar_res must be >= res");
        return VFX_E_VALUE;
    }
    if (S > HIRES_MAX) {
        set_error("scalar_res must be <= " + std::to_string(HIRES_MAX));
        return VFX_E_VALUE;
    }
    if (!hi || !hsolid) return arg_error("null fine field pointer");
    drop_graphs(c);   // captured steps hold the old fine buffers
    HiresGrid& h = c->hr;
    if (h.S != S) {
        hires_free(c);
        const Geom g = make_geom(S, 1);
        // the fine passes stage the coarse velocity of one workgroup's first samples in a fixed
        // LDS tile, which holds blockDim + 1 corners per axis (S >= res)
        if ((int)g.block.x + 1 > VT_X || (int)g.block.y + 1 > VT_Y || (int)g.block.z != 1)
            return arg_error("fine workgroup shape exceeds the velocity tile");
        const size_t n3 = g.n3;
        const size_t nd = mass_part_doubles(g);
        const size_t bytes = nd * sizeof(double) + 6 * n3 * sizeof(real) + n3 * sizeof(int) + n3;
        if (!hip_ok(hipMalloc(&h.pool, bytes), "fine scalar buffers")) {
            h = HiresGrid{};
            return VFX_E_HIP;
        }
        char* p = static_cast<char*>(h.pool);
        h.part = reinterpret_cast<double*>(p);
        p += nd * sizeof(double);
        h.hat = reinterpret_cast<real*>(p);
        h.adv = h.hat + 3 * n3;
        p += 6 * n3 * sizeof(real);
        h.qidx = reinterpret_cast<int*>(p);
        p += n3 * sizeof(int);
        h.ffm = reinterpret_cast<uint8_t*>(p);
        h.g = g;
        h.S = S;
        // the reference's Python doubles, rounded once to the kernels' real (solver.py:217-239)
        const int N = c->N;
        h.dxf = (real)(1.0 / (double)S);
        h.a = (real)((double)N / (double)S);
        h.b = (real)(0.5 * (double)N / (double)S - 0.5);
        h.ratio = (real)((double)S / (double)N);
        h.span = (S + N - 1) / N + 1;
        h.pow2 = S % N == 0 && ((S / N) & (S / N - 1)) == 0;   // a = 2^-k
        h.ia = (real)((double)S / (double)N);
        h.nd = std::max(1, (S + 255) / 256);
    }
    h.hi = hi;
    h.hsolid = hsolid;
    return VFX_OK;
}

int VFX_FN(vfx_hires_init)(vfx_ctx* c, const real* data, float ambient, void* stream) {
    if (c == nullptr) return arg_error("null context");
    if (!data) return arg_error("null pointer");
    const HiresGrid& h = c->hr;
    if (h.S == 0) return arg_error("no fine scalar grid attached");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_fupsample, fine_grid(h.g), h.g.block, 0, s, h.S, c->N, data, h.hsolid, (real)ambient,
                       (real)((double)c->N / (double)h.S), h.hi);
    if (!launch_ok("fine scalar init")) return VFX_E_HIP;
    return VFX_OK;
}

}  // extern "C"
