// Trilinear sampling and the RK2 midpoint trace shared by the advection kernels of the
// simulation grid (advect.hip) and of the finer scalar grid (hires.hip).
//
// Restates vfx/sim3d/kernels.py:66-120 (base_idx, frac, tri, sample_vel, trace).
#pragma once

#include "sim3d_common.hpp"

namespace vfx {

// clamp (kernels.py:67-74): min / max instructions (equal to the branching form for every
// non-NaN g, including the sign of a zero result used only through subtraction of an integer)
__device__ __forceinline__ real clamp_mm(real g, real lo, real hi) { return fmin(fmax(g, lo), hi); }

// base_idx / frac (kernels.py:67-74): clamp to [0, n-1], floor, keep one cell of room.  The
// clamp uses fmaxf/fminf so that a NaN coordinate also lands on cell 0 (equal to the reference's
// clamp for every other input).
__device__ __forceinline__ int base_idx(real g, int n) {
    int i = (int)floor(fmin(fmax(g, RL(0.0)), (real)(n - 1)));
    return i < n - 2 ? i : n - 2;
}

struct Corner {
    int o;      // linear index of the (z0, y0, x0) corner
    real fx, fy, fz;
};

__device__ __forceinline__ Corner corner(int N, real px, real py, real pz) {
    const int x0 = base_idx(px, N), y0 = base_idx(py, N), z0 = base_idx(pz, N);
    Corner c;
    c.o = (z0 * N + y0) * N + x0;
    c.fx = clamp_mm(px, RL(0.0), (real)(N - 1)) - (real)x0;
    c.fy = clamp_mm(py, RL(0.0), (real)(N - 1)) - (real)y0;
    c.fz = clamp_mm(pz, RL(0.0), (real)(N - 1)) - (real)z0;
    return c;
}

// the 8 corner values of a trilinear sample, in (z, y, x) bit order 000 .. 111
struct Oct {
    real v[8];
};

// The two x corners of a row are adjacent (x0 <= N - 2, so the pair never leaves the row): one
// 2-element load at element alignment, which gfx950 issues as a single load instruction.
typedef real real2_e __attribute__((ext_vector_type(2), aligned(sizeof(real))));

__device__ __forceinline__ Oct gather(const real* __restrict__ f, int N, int o) {
    const int sy = N, sz = N * N;
    const real2_e a = *reinterpret_cast<const real2_e*>(f + o);
    const real2_e b = *reinterpret_cast<const real2_e*>(f + o + sy);
    const real2_e c = *reinterpret_cast<const real2_e*>(f + o + sz);
    const real2_e d = *reinterpret_cast<const real2_e*>(f + o + sz + sy);
    Oct g;
    g.v[0] = a.x; g.v[1] = a.y;
    g.v[2] = b.x; g.v[3] = b.y;
    g.v[4] = c.x; g.v[5] = c.y;
    g.v[6] = d.x; g.v[7] = d.y;
    return g;
}

// four values at 2-value alignment (two cells of one channel-pair plane)
typedef real real4_e __attribute__((ext_vector_type(4), aligned(2 * sizeof(real))));

// The advection passes keep their backward values as channel pairs: plane `pl` holds channels
// 2k and 2k + 1 of every cell side by side, so one 16-byte load reads both channels of a corner
// pair.  gather_pair: the 8 corners of both channels at corner o.
__device__ __forceinline__ void gather_pair(const real* __restrict__ pl, int N, int o, Oct& a, Oct& b) {
    const int sy = N, sz = N * N;
    const real4_e r0 = *reinterpret_cast<const real4_e*>(pl + 2 * o);
    const real4_e r1 = *reinterpret_cast<const real4_e*>(pl + 2 * (o + sy));
    const real4_e r2 = *reinterpret_cast<const real4_e*>(pl + 2 * (o + sz));
    const real4_e r3 = *reinterpret_cast<const real4_e*>(pl + 2 * (o + sz + sy));
    a.v[0] = r0.x; a.v[1] = r0.z; b.v[0] = r0.y; b.v[1] = r0.w;
    a.v[2] = r1.x; a.v[3] = r1.z; b.v[2] = r1.y; b.v[3] = r1.w;
    a.v[4] = r2.x; a.v[5] = r2.z; b.v[4] = r2.y; b.v[5] = r2.w;
    a.v[6] = r3.x; a.v[7] = r3.z; b.v[6] = r3.y; b.v[7] = r3.w;
}

// tri (kernels.py:77-85): x first, then y, then z, in the reference's operation order.
__device__ __forceinline__ real lerp3(const Oct& g, const Corner& c) {
    const real a00 = g.v[0] * (RL(1.0) - c.fx) + g.v[1] * c.fx;
    const real a01 = g.v[2] * (RL(1.0) - c.fx) + g.v[3] * c.fx;
    const real a10 = g.v[4] * (RL(1.0) - c.fx) + g.v[5] * c.fx;
    const real a11 = g.v[6] * (RL(1.0) - c.fx) + g.v[7] * c.fx;
    const real a0 = a00 * (RL(1.0) - c.fy) + a01 * c.fy;
    const real a1 = a10 * (RL(1.0) - c.fy) + a11 * c.fy;
    return a0 * (RL(1.0) - c.fz) + a1 * c.fz;
}

// velocity at p (sample_vel, kernels.py:100-112)
__device__ __forceinline__ void sample_vel(const real* __restrict__ vel, int N, int n3, const Corner& c, real* v) {
    const Oct gx = gather(vel, N, c.o), gy = gather(vel + n3, N, c.o), gz = gather(vel + 2 * n3, N, c.o);
    v[0] = lerp3(gx, c);
    v[1] = lerp3(gy, c);
    v[2] = lerp3(gz, c);
}

// velocity at the centre of cell i: the trilinear sample there has weights exactly 0 and 1, so
// it is the cell's own value (up to the sign of a zero, which no traced coordinate keeps)
__device__ __forceinline__ void cell_vel(const real* __restrict__ vel, int n3, int i, real* v) {
    v[0] = vel[i];
    v[1] = vel[n3 + i];
    v[2] = vel[2 * n3 + i];
}

// trace (kernels.py:115-120): RK2 midpoint over signed s (cells per m/s) from the centre of
// cell i = (px, py, pz).
__device__ __forceinline__ void trace1(const real* __restrict__ vel, int N, int n3, int i, real px, real py, real pz,
                                       real s, real* q) {
    real v1[3];
    cell_vel(vel, n3, i, v1);
    const real h = RL(0.5) * s;
    const Corner c1 = corner(N, px + h * v1[0], py + h * v1[1], pz + h * v1[2]);
    real v2[3];
    sample_vel(vel, N, n3, c1, v2);
    q[0] = px + s * v2[0];
    q[1] = py + s * v2[1];
    q[2] = pz + s * v2[2];
}

// trace (kernels.py:115-120) from an arbitrary point p of the index space (both velocity
// samples trilinear), as the finer scalar grid needs (hires.py:30-48).
__device__ __forceinline__ void trace_pt(const real* __restrict__ vel, int N, int n3, real px, real py, real pz,
                                         real s, real* q) {
    real v1[3];
    sample_vel(vel, N, n3, corner(N, px, py, pz), v1);
    const real h = RL(0.5) * s;
    const Corner c1 = corner(N, px + h * v1[0], py + h * v1[1], pz + h * v1[2]);
    real v2[3];
    sample_vel(vel, N, n3, c1, v2);
    q[0] = px + s * v2[0];
    q[1] = py + s * v2[1];
    q[2] = pz + s * v2[2];
}

}  // namespace vfx
