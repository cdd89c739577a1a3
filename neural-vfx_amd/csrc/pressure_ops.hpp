// Device-side collocated operators shared by the multigrid (pressure.hip) and CG (cg.hip)
// kernels: the face mask, g = -d^T, K = d d^T, diag(K) and d = 2dx div (pressure.py:36-111).
//
// Cell indices are 32-bit (N <= 1024).  Mask bit INTERIOR marks a fluid cell whose neighbours at
// distance 1 and 2 along every axis exist and are fluid: there every face that kmul / div2 / g_ax
// test is open, so the fast paths run the reference's operations in the reference's order
// without reading a mask bit (same rounding as the general path).
#pragma once

#include <climits>

#include "sim3d_kernels.hpp"

namespace vfx {

constexpr int INTERIOR = 128;

__device__ __forceinline__ int axstride(int N, int ax) { return ax == 0 ? 1 : (ax == 1 ? N : N * N); }

// g_ax (pressure.py:54-65) at linear cell i
__device__ __forceinline__ float g_ax(const float* __restrict__ p, const uint8_t* __restrict__ fm, int i, int st,
                                      int ax) {
    const int m = fm[i];
    const float c = p[i];
    float g = 0.0f;
    if ((m >> (2 * ax)) & 1) g += p[i + st] - c;
    if ((m >> (2 * ax + 1)) & 1) g += c - p[i - st];
    return g;
}

// kmul along one axis with every face open: d += gi + g+; d -= gi + g-
__device__ __forceinline__ float kmul_axis_open(const float* __restrict__ p, int i, int st, float d) {
    const float m2 = p[i - 2 * st], m1 = p[i - st], c = p[i], p1 = p[i + st], p2 = p[i + 2 * st];
    const float gi = (p1 - c) + (c - m1);
    const float gp = (p2 - p1) + (p1 - c);
    const float gm = (c - m1) + (m1 - m2);
    d += gi + gp;
    d -= gi + gm;
    return d;
}

// kmul (pressure.py:68-80): (K p)_i at a fluid cell with mask m = fm[i]
__device__ __forceinline__ float kmul_m(const float* __restrict__ p, const uint8_t* __restrict__ fm, int N, int i,
                                        int m) {
    float d = 0.0f;
    if (m & INTERIOR) {
        d = kmul_axis_open(p, i, 1, d);
        d = kmul_axis_open(p, i, N, d);
        d = kmul_axis_open(p, i, N * N, d);
        return -d;
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int st = axstride(N, ax);
        const float gi = g_ax(p, fm, i, st, ax);
        if ((m >> (2 * ax)) & 1) d += gi + g_ax(p, fm, i + st, st, ax);
        if ((m >> (2 * ax + 1)) & 1) d -= gi + g_ax(p, fm, i - st, st, ax);
    }
    return -d;
}

__device__ __forceinline__ float kmul(const float* __restrict__ p, const uint8_t* __restrict__ fm, int N, int i) {
    return kmul_m(p, fm, N, i, fm[i]);
}

// kdiag (pressure.py:83-91)
__device__ __forceinline__ float kdiag(int m) {
    float s = 0.0f;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float oR = (float)((m >> (2 * ax)) & 1), oL = (float)((m >> (2 * ax + 1)) & 1);
        s += (oL - oR) * (oL - oR) + oL + oR;
    }
    return s;
}

// div2 (pressure.py:94-106) of the velocity channels vel[0..2]
__device__ __forceinline__ float div2(const float* __restrict__ vel, int n3, int m, int N, int i) {
    float d = 0.0f;
    if (m & INTERIOR) {
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const float* va = vel + ax * n3;
            const int st = axstride(N, ax);
            const float u = va[i];
            d += u + va[i + st];
            d -= u + va[i - st];
        }
        return d;
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float* va = vel + ax * n3;
        const int st = axstride(N, ax);
        const float u = va[i];
        if ((m >> (2 * ax)) & 1) d += u + va[i + st];
        if ((m >> (2 * ax + 1)) & 1) d -= u + va[i - st];
    }
    return d;
}

__device__ __forceinline__ bool fluid_at(const float* __restrict__ solid, int N, int z, int y, int x) {
    return x >= 0 && y >= 0 && z >= 0 && x < N && y < N && z < N && solid[((z * N) + y) * N + x] <= 0.5f;
}

// face mask of one cell from `solid` (k_mask pressure.py:115-143) plus the INTERIOR bit
__device__ __forceinline__ int mask_from_solid(const float* __restrict__ solid, int N, int z, int y, int x) {
    const int i = (z * N + y) * N + x;
    if (solid[i] > 0.5f) return 0;
    const int sz = N * N;
    int m = FLUID;
    if (x + 1 < N && solid[i + 1] <= 0.5f) m |= 1;
    if (x > 0 && solid[i - 1] <= 0.5f) m |= 2;
    if (y + 1 < N && solid[i + N] <= 0.5f) m |= 4;
    if (y > 0 && solid[i - N] <= 0.5f) m |= 8;
    if (z + 1 < N && solid[i + sz] <= 0.5f) m |= 16;
    if (z > 0 && solid[i - sz] <= 0.5f) m |= 32;
    if (m == (FLUID | 63) && fluid_at(solid, N, z, y, x + 2) && fluid_at(solid, N, z, y, x - 2) &&
        fluid_at(solid, N, z, y + 2, x) && fluid_at(solid, N, z, y - 2, x) && fluid_at(solid, N, z + 2, y, x) &&
        fluid_at(solid, N, z - 2, y, x))
        m |= INTERIOR;
    return m;
}

// 2x2x2 Hadamard sign of component c at in-block offset (dz, dy, dx) (hsign, pressure.py:179-184)
__device__ __forceinline__ float hsign(int c, int dz, int dy, int dx) {
    const int e = (c & 1) * dx + ((c >> 1) & 1) * dy + ((c >> 2) & 1) * dz;
    return (e & 1) ? -HS : HS;
}

__device__ __forceinline__ size_t lcell(int m, int K, int J, int I) { return ((size_t)K * m + J) * m + I; }

__device__ __forceinline__ bool cg_inactive(const DevState* st, int k) { return st->stop_at < k; }

// ---------------------------------------------------------------- LDS tiles of the fine grid
// A workgroup owns TX x TY x TZ fine cells (whole 2x2x2 blocks) and stages them with a 2-cell
// halo, the reach of K = d d^T.  Cells outside the box are staged with mask 0 and value 0.
constexpr int TX = 32, TY = 8, TZ = 8;
constexpr int HX = TX + 4, HY = TY + 4, HZ = TZ + 4;
constexpr int HN = HX * HY * HZ;
constexpr int TILE_THREADS = 256;
constexpr int BTX = TX / 2, BTY = TY / 2, BTZ = TZ / 2;     // blocks per tile
constexpr int HBX = BTX + 2, HBY = BTY + 2, HBZ = BTZ + 2;  // blocks with a 1-block halo
constexpr int HBN = HBX * HBY * HBZ;

// kmul (pressure.py:68-80) on an LDS tile: open-face tests become 0/1 multipliers (x * 1 == x,
// d + (+-0) == d), so the result is bit-identical to the branching form for finite values.
__device__ __forceinline__ float kmul_tile(const float* P, const uint8_t* M, int h) {
    float d = 0.0f;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int st = ax == 0 ? 1 : (ax == 1 ? HX : HX * HY);
        const int m0 = M[h], mp = M[h + st], mm = M[h - st];
        const float op = (float)((m0 >> (2 * ax)) & 1), om = (float)((m0 >> (2 * ax + 1)) & 1);
        const float opp = (float)((mp >> (2 * ax)) & 1), omp = (float)((mp >> (2 * ax + 1)) & 1);
        const float opm = (float)((mm >> (2 * ax)) & 1), omm = (float)((mm >> (2 * ax + 1)) & 1);
        const float m2 = P[h - 2 * st], m1 = P[h - st], c = P[h], p1 = P[h + st], p2 = P[h + 2 * st];
        const float gi = (0.0f + op * (p1 - c)) + om * (c - m1);
        const float gp = (0.0f + opp * (p2 - p1)) + omp * (p1 - c);
        const float gm = (0.0f + opm * (c - m1)) + omm * (m1 - m2);
        d = d + op * (gi + gp);
        d = d - om * (gi + gm);
    }
    return -d;
}

// halo index of fine cell (gx, gy, gz) relative to the tile origin (ox, oy, oz) (halo excluded)
__device__ __forceinline__ int tile_h(int lx, int ly, int lz) { return ((lz + 2) * HY + (ly + 2)) * HX + (lx + 2); }

// offsum (pressure.py:271-287) in the component-fastest layout
__device__ __forceinline__ float offsum(const Lvl& L, int c, int K, int J, int I) {
    const int m = L.m;
    const size_t o = lcell(m, K, J, I);
    const size_t sy = (size_t)m, sz = (size_t)m * m;
    float a = 0.0f;
    if (I + 1 < m) a += L.off[o * 24 + c] * L.x[(o + 1) * 8 + c];
    if (I > 0) a += L.off[(o - 1) * 24 + c] * L.x[(o - 1) * 8 + c];
    if (J + 1 < m) a += L.off[o * 24 + 8 + c] * L.x[(o + sy) * 8 + c];
    if (J > 0) a += L.off[(o - sy) * 24 + 8 + c] * L.x[(o - sy) * 8 + c];
    if (K + 1 < m) a += L.off[o * 24 + 16 + c] * L.x[(o + sz) * 8 + c];
    if (K > 0) a += L.off[(o - sz) * 24 + 16 + c] * L.x[(o - sz) * 8 + c];
    return a;
}

// one GS update (k_smooth pressure.py:290-303); returns the new value
__device__ __forceinline__ float smooth_item(const Lvl& L, int c, int K, int J, int I, bool zero_init) {
    const size_t o = lcell(L.m, K, J, I) * 8 + c;
    const float d = L.diag[o];
    float v = 0.0f;
    if (d > L.dmin) {
        float acc = L.r[o];
        if (!zero_init) acc -= offsum(L, c, K, J, I);
        v = acc / d;
    }
    L.x[o] = v;
    return v;
}

}  // namespace vfx
