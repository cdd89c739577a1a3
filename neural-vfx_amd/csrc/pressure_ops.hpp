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

// Branch-free forms.  A load guarded by a bounds or mask test is issued only after the test's own
// load returns, so each guarded load costs a full memory round trip; instead every load below
// goes to an in-bounds address (clamped to the cell itself when the neighbour is missing) and the
// open-face tests become 0/1 multipliers: x * 1 == x and d + (+-0) == d, so the value equals the
// reference's branching form bit for bit (for finite inputs).

// neighbour indices along axis ax of cell i at coordinate co: +1, +2, -1, -2 (clamped to i)
struct Nb {
    int p1, p2, m1, m2;
};

__device__ __forceinline__ Nb neighbours(int N, int i, int co, int st) {
    Nb n;
    n.p1 = co + 1 < N ? i + st : i;
    n.p2 = co + 2 < N ? i + 2 * st : n.p1;
    n.m1 = co >= 1 ? i - st : i;
    n.m2 = co >= 2 ? i - 2 * st : n.m1;
    return n;
}

__device__ __forceinline__ real bitf(int m, int b) { return (real)((m >> b) & 1); }

// g_ax (pressure.py:54-65) at cell i (coordinate co along ax)
__device__ __forceinline__ real g_ax_xyz(const real* __restrict__ p, int m, int N, int i, int co, int ax) {
    const Nb n = neighbours(N, i, co, axstride(N, ax));
    const real c = p[i], p1 = p[n.p1], m1 = p[n.m1];
    return (RL(0.0) + bitf(m, 2 * ax) * (p1 - c)) + bitf(m, 2 * ax + 1) * (c - m1);
}

// kmul (pressure.py:68-80): (K p)_i at a fluid cell (x, y, z) with mask m = fm[i]
__device__ __forceinline__ real kmul_xyz(const real* __restrict__ p, const uint8_t* __restrict__ fm, int N, int x,
                                          int y, int z, int i, int m) {
    real d = RL(0.0);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int co = ax == 0 ? x : (ax == 1 ? y : z);
        const Nb n = neighbours(N, i, co, axstride(N, ax));
        const int mp = fm[n.p1], mm = fm[n.m1];
        const real c = p[i], p1 = p[n.p1], p2 = p[n.p2], m1 = p[n.m1], m2 = p[n.m2];
        const real op = bitf(m, 2 * ax), om = bitf(m, 2 * ax + 1);
        const real gi = (RL(0.0) + op * (p1 - c)) + om * (c - m1);
        const real gp = (RL(0.0) + bitf(mp, 2 * ax) * (p2 - p1)) + bitf(mp, 2 * ax + 1) * (p1 - c);
        const real gm = (RL(0.0) + bitf(mm, 2 * ax) * (c - m1)) + bitf(mm, 2 * ax + 1) * (m1 - m2);
        d = d + op * (gi + gp);
        d = d - om * (gi + gm);
    }
    return -d;
}

// kdiag (pressure.py:83-91): sum over axes of (oL - oR)^2 + oL + oR, which is 2 for an axis with
// any open face and 0 otherwise (small integers, exact in real)
__device__ __forceinline__ real kdiag(int m) { return (real)(2 * __popc((m | (m >> 1)) & 0x15)); }

// x / 6 correctly rounded, without the division sequence: q = x c with c = 1/6 rounded, then one
// residual correction q + (x - 6 q) c with fused multiply-adds (Markstein's method for a constant
// divisor).  Over all 2^32 single-precision inputs this equals IEEE x / 6 bit for bit except where
// x / 6 is subnormal and x is not zero (|x| <= 2^-125); those, the infinities and NaN take the
// division itself, and the sign of a zero result is x's.  diag(K) is 6 at every INTERIOR cell.
__device__ __forceinline__ real div6(real x) {
    if constexpr (sizeof(real) == 4) {
        const float c6 = 1.0f / 6.0f;
        const float q = x * c6;
        const float r = __builtin_fmaf(-q, 6.0f, x);
        float y = __builtin_copysignf(__builtin_fmaf(r, c6, q), x);
        const float ax = __builtin_fabsf(x);
        if (__builtin_expect(!(ax >= 0x1p-124f && ax <= 3.40282347e38f) && x != 0.0f, 0)) y = x / 6.0f;
        return y;
    } else {
        return x / RL(6.0);
    }
}

// A fluid cell with at least one open face.  A fluid cell sealed in by solids on all six sides has
// diag(K) = 0 and a zero row and column of K (its right-hand side is 0 too): its Jacobi terms are
// 0 here.  The reference divides 0 / 0 there (pressure.py:671, :702), and the NaN stops its PCG at
// iteration 0 with the velocity left unprojected.
__device__ __forceinline__ bool kopen(int m) { return (m & 63) != 0; }

// div2 (pressure.py:94-106) of the velocity channels vel[0..2] at cell (x, y, z), mask m
__device__ __forceinline__ real div2_xyz(const real* __restrict__ vel, int n3, int m, int N, int x, int y, int z,
                                          int i) {
    real d = RL(0.0);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const real* va = vel + ax * n3;
        const int co = ax == 0 ? x : (ax == 1 ? y : z);
        const Nb n = neighbours(N, i, co, axstride(N, ax));
        const real u = va[i], up = va[n.p1], um = va[n.m1];
        d = d + bitf(m, 2 * ax) * (u + up);
        d = d - bitf(m, 2 * ax + 1) * (u + um);
    }
    return d;
}

__device__ __forceinline__ bool fluid_at(const real* __restrict__ solid, int N, int z, int y, int x) {
    return x >= 0 && y >= 0 && z >= 0 && x < N && y < N && z < N && solid[((z * N) + y) * N + x] <= RL(0.5);
}

// face mask of one cell from `solid` (k_mask pressure.py:115-143) plus the INTERIOR bit
__device__ __forceinline__ int mask_from_solid(const real* __restrict__ solid, int N, int z, int y, int x) {
    const int i = (z * N + y) * N + x;
    if (solid[i] > RL(0.5)) return 0;
    const int sz = N * N;
    int m = FLUID;
    if (x + 1 < N && solid[i + 1] <= RL(0.5)) m |= 1;
    if (x > 0 && solid[i - 1] <= RL(0.5)) m |= 2;
    if (y + 1 < N && solid[i + N] <= RL(0.5)) m |= 4;
    if (y > 0 && solid[i - N] <= RL(0.5)) m |= 8;
    if (z + 1 < N && solid[i + sz] <= RL(0.5)) m |= 16;
    if (z > 0 && solid[i - sz] <= RL(0.5)) m |= 32;
    if (m == (FLUID | 63) && fluid_at(solid, N, z, y, x + 2) && fluid_at(solid, N, z, y, x - 2) &&
        fluid_at(solid, N, z, y + 2, x) && fluid_at(solid, N, z, y - 2, x) && fluid_at(solid, N, z + 2, y, x) &&
        fluid_at(solid, N, z - 2, y, x))
        m |= INTERIOR;
    return m;
}

// 2x2x2 Hadamard sign of component c at in-block offset (dz, dy, dx) (hsign, pressure.py:179-184)
__device__ __forceinline__ real hsign(int c, int dz, int dy, int dx) {
    const int e = (c & 1) * dx + ((c >> 1) & 1) * dy + ((c >> 2) & 1) * dz;
    return (e & 1) ? -HS : HS;
}

// Level cell indices are 32-bit: for N <= 1024 the largest level offset, 24 m^3 with m = 512, is
// below 2^32 (the context rejects larger grids).
__device__ __forceinline__ uint32_t lcell(int m, int K, int J, int I) {
    return ((uint32_t)K * (uint32_t)m + (uint32_t)J) * (uint32_t)m + (uint32_t)I;
}

// Storage of the level unknowns x and r (8 components per cell).  For even m they are split by
// colour: cell o = lcell(K, J, I) of colour (I + J + K) & 1 is stored at colour * m^3/2 + o / 2,
// so one colour's cells form one contiguous half (cells I and I + 1 of a row share o / 2 and have
// opposite colours) and a red-black sweep reads and writes only the lines it uses.  Odd m (the
// coarsest level of some grids) keeps the natural order.  Operator coefficients (diag, off, sf)
// stay in the natural order.
__device__ __forceinline__ uint32_t lsto(int m, uint32_t o, int color) {
    const uint32_t um = (uint32_t)m;
    return (m & 1) ? o : (uint32_t)color * ((um * um * um) >> 1) + (o >> 1);
}
__device__ __forceinline__ uint32_t lsto(int m, int K, int J, int I) {
    return lsto(m, lcell(m, K, J, I), (I + J + K) & 1);
}

// ---------------------------------------------------------------- LDS tiles of the fine grid
// A workgroup owns TX x TY x TZ fine cells (whole 2x2x2 blocks) and stages them with a 2-cell
// halo, the reach of K = d d^T.  Cells outside the box are staged with mask 0 and value 0.
constexpr int TX = 32, TY = 8, TZ = VFX_F64 ? 4 : 16;   // FP64 tiles: a quarter as deep (LDS)
constexpr int HX = TX + 4, HY = TY + 4, HZ = TZ + 4;
constexpr int HN = HX * HY * HZ;
constexpr int TILE_THREADS = (TX / 2) * (TY / 2) * (TZ / 2);   // one thread per own 2x2x2 block
constexpr int BTX = TX / 2, BTY = TY / 2, BTZ = TZ / 2;     // blocks per tile
constexpr int HBX = BTX + 2, HBY = BTY + 2, HBZ = BTZ + 2;  // blocks with a 1-block halo
constexpr int HBN = HBX * HBY * HBZ;

// kmul (pressure.py:68-80) on an LDS tile: open-face tests become 0/1 multipliers (x * 1 == x,
// d + (+-0) == d), so the result is bit-identical to the branching form for finite values.
__device__ __forceinline__ real kmul_tile(const real* P, const uint8_t* M, int h) {
    real d = RL(0.0);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int st = ax == 0 ? 1 : (ax == 1 ? HX : HX * HY);
        const int m0 = M[h], mp = M[h + st], mm = M[h - st];
        const real op = (real)((m0 >> (2 * ax)) & 1), om = (real)((m0 >> (2 * ax + 1)) & 1);
        const real opp = (real)((mp >> (2 * ax)) & 1), omp = (real)((mp >> (2 * ax + 1)) & 1);
        const real opm = (real)((mm >> (2 * ax)) & 1), omm = (real)((mm >> (2 * ax + 1)) & 1);
        const real m2 = P[h - 2 * st], m1 = P[h - st], c = P[h], p1 = P[h + st], p2 = P[h + 2 * st];
        const real gi = (RL(0.0) + op * (p1 - c)) + om * (c - m1);
        const real gp = (RL(0.0) + opp * (p2 - p1)) + omp * (p1 - c);
        const real gm = (RL(0.0) + opm * (c - m1)) + omm * (m1 - m2);
        d = d + op * (gi + gp);
        d = d - om * (gi + gm);
    }
    return -d;
}

// kmul_tile at an INTERIOR cell: every open-face multiplier is 1 (x * 1 == x), so the same
// operations without the mask, and without the additions of +0 that start each face sum
// (0 + a == a except that it turns -0 into +0: the values are equal, only the sign of a zero
// result can differ)
__device__ __forceinline__ real kmul_tile_interior(const real* P, int h) {
    real d = RL(0.0);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int st = ax == 0 ? 1 : (ax == 1 ? HX : HX * HY);
        const real m2 = P[h - 2 * st], m1 = P[h - st], c = P[h], p1 = P[h + st], p2 = P[h + 2 * st];
        const real gi = (p1 - c) + (c - m1);
        const real gp = (p2 - p1) + (p1 - c);
        const real gm = (c - m1) + (m1 - m2);
        d = d + (gi + gp);
        d = d - (gi + gm);
    }
    return -d;
}

// halo index of fine cell (gx, gy, gz) relative to the tile origin (ox, oy, oz) (halo excluded)
__device__ __forceinline__ int tile_h(int lx, int ly, int lz) { return ((lz + 2) * HY + (ly + 2)) * HX + (lx + 2); }

// The six neighbour cells of a level cell, clamped to the cell itself where missing, with the
// existence flags (the tests of offsum, pressure.py:275-286).
struct LNb {
    uint32_t c[6];   // +I, -I, +J, -J, +K, -K
    bool e[6];
};

__device__ __forceinline__ LNb level_neighbours(int m, uint32_t o, int K, int J, int I) {
    LNb n;
    const uint32_t sy = (uint32_t)m, sz = (uint32_t)m * (uint32_t)m;
    n.e[0] = I + 1 < m; n.c[0] = n.e[0] ? o + 1 : o;
    n.e[1] = I > 0;     n.c[1] = n.e[1] ? o - 1 : o;
    n.e[2] = J + 1 < m; n.c[2] = n.e[2] ? o + sy : o;
    n.e[3] = J > 0;     n.c[3] = n.e[3] ? o - sy : o;
    n.e[4] = K + 1 < m; n.c[4] = n.e[4] ? o + sz : o;
    n.e[5] = K > 0;     n.c[5] = n.e[5] ? o - sz : o;
    return n;
}

// coupling coefficient of link k (0..5) of component c: +ax links are stored at the cell, -ax
// links at the -ax neighbour
template <class LT>
__device__ __forceinline__ real link_coef(const LvlT<LT>& L, const LNb& n, uint32_t o, int k, int c) {
    const int ax = k >> 1;
    return ldl(L.off + ((k & 1) ? n.c[k] : o) * 24 + ax * 8 + c);
}

// Standard cells.  Away from walls and solids every cell of a level has the same diag and link
// coefficients (tab: the same arithmetic on the same values).  A cell whose flag sf is set has
// all 6 neighbours, its own diag and links and its -ax neighbours' links equal to tab bit for bit,
// and neighbours with diag > dmin.  When every lane of a wave is at a standard cell, the
// coefficients come from tab (a 128-byte table) instead of the level arrays; the arithmetic is
// unchanged, so the result is the same as the general path's.
struct Tab {
    real d, o[3];   // diag and the x, y, z link coefficients of component c
};

template <class LT>
__device__ __forceinline__ Tab load_tab(const LvlT<LT>& L, int c) {
    Tab t;
    t.d = L.tab[c];
    t.o[0] = L.tab[8 + c];
    t.o[1] = L.tab[16 + c];
    t.o[2] = L.tab[24 + c];
    return t;
}

// offsum (pressure.py:271-287) from the link coefficients cf and neighbour values xn
__device__ __forceinline__ real offsum_v(const real* cf, const real* xn, const bool* e) {
    real a = RL(0.0);
#pragma unroll
    for (int k = 0; k < 6; ++k) a = a + (e[k] ? cf[k] * xn[k] : RL(0.0));
    return a;
}

// one GS update (k_smooth pressure.py:290-303); returns the new value.  STD: waves of standard
// cells use the table (the single-workgroup coarse levels have few standard cells and skip it).
template <bool STD, class LT>
__device__ __forceinline__ real smooth_value(const LvlT<LT>& L, int c, int K, int J, int I, bool zero_init) {
    const uint32_t o = lcell(L.m, K, J, I);
    const int col = (I + J + K) & 1;
    const real r = ldl(L.r + lsto(L.m, o, col) * 8 + c);
    if (zero_init) {
        const real d = ldl(L.diag + o * 8 + c);
        return d <= L.dmin ? RL(0.0) : r / d;
    }
    const LNb n = level_neighbours(L.m, o, K, J, I);
    const int f = STD ? L.sf[o] : 0;
    real xn[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) xn[k] = ldl(L.x + lsto(L.m, n.c[k], n.e[k] ? col ^ 1 : col) * 8 + c);
    const Tab tb = STD ? load_tab(L, c) : Tab{};
    real d, cf[6];
    if (STD && __all(f != 0)) {
        d = tb.d;
#pragma unroll
        for (int k = 0; k < 6; ++k) cf[k] = tb.o[k >> 1];
    } else {
        d = ldl(L.diag + o * 8 + c);
#pragma unroll
        for (int k = 0; k < 6; ++k) cf[k] = link_coef(L, n, o, k, c);
    }
    const real acc = r - offsum_v(cf, xn, n.e);
    return d <= L.dmin ? RL(0.0) : acc / d;
}

template <bool STD, class LT>
__device__ __forceinline__ real smooth_item(const LvlT<LT>& L, int c, int K, int J, int I, bool zero_init) {
    const real v = smooth_value<STD>(L, c, K, J, I, zero_init);
    stl(L.x + lsto(L.m, K, J, I) * 8 + c, v);
    return v;
}

}  // namespace vfx
