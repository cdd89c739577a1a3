// Collocated exact projection: face masks, Hadamard-basis multigrid operators, the V(1,1)
// preconditioner and the PCG loop (K-MASK, K-PROBE, K-GALERKIN, K-MG-*, K-RHS, K-MATVEC,
// K-CG-*, K-APPLY, K-DIV).
//
// Restates vfx/sim3d/pressure.py.  The operators (pressure.py:36-111): d(u) = 2dx div u with
// face velocity the mean of the two cells and closed walls / solid faces; g = -d^T; K = d d^T.
// The preconditioner (pressure.py:10-15): one damped-Jacobi sweep on exact K, a V(1,1)
// red-black Gauss-Seidel cycle on the 8-component Hadamard level-1 system with Galerkin coarse
// levels (scale 1/16, constant prolongation), one damped-Jacobi sweep again.
//
// Multigrid level storage is component-fastest: x, r, diag are [K][J][I][8] and off is
// [K][J][I][24] (off[ax*8 + c] couples to the +ax neighbour block), so the 8 lanes that treat
// the 8 components of one block cell read one contiguous 32-byte record.
//
// CG stepping: every kernel of CG iteration k takes k and returns at once when the device stop
// test has fired at an earlier iteration (DevState::stop_at < k).  This lets the host enqueue
// iterations ahead of the device without reading anything back per iteration.
#include "pressure_ops.hpp"

namespace vfx {

// ---------------------------------------------------------------- setup (solid changes only)
// k_solid_changed (pressure.py:146-153) against the cached face mask: the operators depend on
// solid only through (solid > 0.5), which the mask's fluid bit records.
__global__ void __launch_bounds__(256) k_solid_check(size_t n3, const real* __restrict__ solid,
                                                     const uint8_t* __restrict__ fm, DevState* __restrict__ st) {
    bool diff = false;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x)
        diff |= (solid[i] > RL(0.5)) != (fm[i] == 0);
    if (__any(diff) && (threadIdx.x & 63) == 0) st->changed = 1;
}

__global__ void __launch_bounds__(256) k_mask(int N, const real* __restrict__ solid, uint8_t* __restrict__ fm,
                                              const DevState* __restrict__ st) {
    if (!st->changed) return;
    const size_t n3 = (size_t)N * N * N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / ((size_t)N * N));
        fm[i] = (uint8_t)mask_from_solid(solid, N, z, y, x);
    }
}

// tv / gt / axcell (pressure.py:179-217)
__device__ __forceinline__ real tv(int c, int BK, int BJ, int BI, int z, int y, int x) {
    if ((z >> 1) == BK && (y >> 1) == BJ && (x >> 1) == BI) return hsign(c, z & 1, y & 1, x & 1);
    return RL(0.0);
}

__device__ __forceinline__ int mask_at(const uint8_t* __restrict__ fm, int N, int z, int y, int x) {
    if (x < 0 || y < 0 || z < 0 || z >= N || y >= N || x >= N) return 0;
    return fm[idx3(N, z, y, x)];
}

// ALL: every cell fluid with every face open (the neighbourhood of a standard cell)
template <bool ALL>
__device__ __forceinline__ real gt(const uint8_t* __restrict__ fm, int N, int c, int BK, int BJ, int BI, int z,
                                    int y, int x, int ax) {
    const int m = ALL ? (FLUID | 63) : mask_at(fm, N, z, y, x);
    const int ex = ax == 0, ey = ax == 1, ez = ax == 2;
    const real t0 = tv(c, BK, BJ, BI, z, y, x);
    real g = RL(0.0);
    if ((m >> (2 * ax)) & 1) g += tv(c, BK, BJ, BI, z + ez, y + ey, x + ex) - t0;
    if ((m >> (2 * ax + 1)) & 1) g += t0 - tv(c, BK, BJ, BI, z - ez, y - ey, x - ex);
    return g;
}

__device__ __forceinline__ void axcell(int ax, int oz, int oy, int ox, int a, int b, int c, int& z, int& y, int& x) {
    if (ax == 0) { z = oz + c; y = oy + b; x = ox + a; }
    else if (ax == 1) { z = oz + c; y = oy + a; x = ox + b; }
    else { z = oz + a; y = oy + c; x = ox + b; }
}

// k_probe1 (pressure.py:220-246): exact component-diagonal of T K T^T for (block, c).  ALL
// evaluates it for a block in an all-fluid neighbourhood (the standard-cell table).
template <bool ALL>
__device__ void probe_item(const uint8_t* __restrict__ fm, int N, int m, int c, int K, int J, int I, real& dsum,
                           real* off3) {
    const int oz = 2 * K, oy = 2 * J, ox = 2 * I;
    dsum = RL(0.0);
    for (int ax = 0; ax < 3; ++ax)
        for (int b = 0; b < 2; ++b)
            for (int cc = 0; cc < 2; ++cc)
                for (int a = 0; a < 4; ++a) {
                    int z, y, x;
                    axcell(ax, oz, oy, ox, a - 1, b, cc, z, y, x);
                    const real v = gt<ALL>(fm, N, c, K, J, I, z, y, x, ax);
                    dsum += v * v;
                }
    for (int ax = 0; ax < 3; ++ax) {
        const int nK = K + (ax == 2), nJ = J + (ax == 1), nI = I + (ax == 0);
        real s = RL(0.0);
        if (ALL || (nK < m && nJ < m && nI < m)) {
            for (int b = 0; b < 2; ++b)
                for (int cc = 0; cc < 2; ++cc)
                    for (int a = 0; a < 2; ++a) {
                        int z, y, x;
                        axcell(ax, oz, oy, ox, 1 + a, b, cc, z, y, x);
                        s += gt<ALL>(fm, N, c, K, J, I, z, y, x, ax) * gt<ALL>(fm, N, c, nK, nJ, nI, z, y, x, ax);
                    }
        }
        off3[ax] = s;
    }
}

template <class LT>
__global__ void __launch_bounds__(256) k_probe(int N, const uint8_t* __restrict__ fm, LvlT<LT> L,
                                               const DevState* __restrict__ st) {
    if (!st->changed) return;
    const int m = L.m;
    const size_t items = (size_t)m * m * m * 8;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < items; t += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(t & 7);
        const size_t cell = t >> 3;
        const int I = (int)(cell % m), J = (int)((cell / m) % m), K = (int)(cell / ((size_t)m * m));
        real d, o3[3];
        probe_item<false>(fm, N, m, c, K, J, I, d, o3);
        stl(L.diag + cell * 8 + c, d);
        for (int ax = 0; ax < 3; ++ax) stl(L.off + cell * 24 + ax * 8 + c, o3[ax]);
    }
}

// k_galerkin (pressure.py:249-268)
template <class LT>
__global__ void __launch_bounds__(256) k_galerkin(LvlT<LT> F, LvlT<LT> C, const DevState* __restrict__ st) {
    if (!st->changed) return;
    const int m = C.m, mf = F.m;
    const size_t items = (size_t)m * m * m * 8;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < items; t += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(t & 7);
        const size_t cell = t >> 3;
        const int I = (int)(cell % m), J = (int)((cell / m) % m), K = (int)(cell / ((size_t)m * m));
        const int oz = 2 * K, oy = 2 * J, ox = 2 * I;
        real d = RL(0.0);
        for (int k = 0; k < 8; ++k) d += ldl(F.diag + lcell(mf, oz + (k >> 2), oy + ((k >> 1) & 1), ox + (k & 1)) * 8 + c);
        for (int ax = 0; ax < 3; ++ax) {
            real oi = RL(0.0), oo = RL(0.0);
            for (int b = 0; b < 2; ++b)
                for (int cc = 0; cc < 2; ++cc) {
                    int z, y, x;
                    axcell(ax, oz, oy, ox, 0, b, cc, z, y, x);
                    oi += ldl(F.off + lcell(mf, z, y, x) * 24 + ax * 8 + c);
                    axcell(ax, oz, oy, ox, 1, b, cc, z, y, x);
                    oo += ldl(F.off + lcell(mf, z, y, x) * 24 + ax * 8 + c);
                }
            d += RL(2.0) * oi;
            stl(C.off + cell * 24 + ax * 8 + c, oo * GALERKIN_SCALE);
        }
        stl(C.diag + cell * 8 + c, d * GALERKIN_SCALE);
    }
}

__global__ void k_clear_changed(DevState* st) {
    if (threadIdx.x == 0) st->changed = 0;
}

// Standard-cell tables of every level (thread c: component c): level 1 from the probe of a block
// in an all-fluid neighbourhood, each coarser level from k_galerkin's sums over standard fine
// values, in k_galerkin's order; every value as stored (rounded to LT), like the level arrays.
template <class LT>
__global__ void __launch_bounds__(64) k_std_tab(CoarseLevelsT<LT> all, int N, const DevState* __restrict__ st) {
    if (!st->changed) return;
    const int c = (int)threadIdx.x;
    if (c >= 8) return;
    real d, o3[3];
    probe_item<true>(nullptr, N, all.lv[0].m, c, 1, 1, 1, d, o3);
    all.lv[0].tab[c] = as_stored<LT>(d);
    for (int ax = 0; ax < 3; ++ax) all.lv[0].tab[8 + ax * 8 + c] = as_stored<LT>(o3[ax]);
    for (int l = 0; l + 1 < all.n; ++l) {
        const real* ft = all.lv[l].tab;
        real* ct = all.lv[l + 1].tab;
        real dd = RL(0.0);
        for (int k = 0; k < 8; ++k) dd += ft[c];
        for (int ax = 0; ax < 3; ++ax) {
            real oi = RL(0.0), oo = RL(0.0);
            for (int b = 0; b < 4; ++b) {
                oi += ft[8 + ax * 8 + c];
                oo += ft[8 + ax * 8 + c];
            }
            dd += RL(2.0) * oi;
            ct[8 + ax * 8 + c] = as_stored<LT>(oo * GALERKIN_SCALE);
        }
        ct[c] = as_stored<LT>(dd * GALERKIN_SCALE);
    }
}

// sf = 1 where the cell has all 6 neighbours, its diag and links and its -ax neighbours' links
// equal tab bit for bit, and every neighbour has diag > dmin
template <class LT>
__global__ void __launch_bounds__(256) k_std_flags(LvlT<LT> L, int use_std, const DevState* __restrict__ st) {
    if (!st->changed) return;
    const int m = L.m;
    const size_t m3 = (size_t)m * m * m;
    for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < m3; o += (size_t)gridDim.x * blockDim.x) {
        const int I = (int)(o % m), J = (int)((o / m) % m), K = (int)(o / ((size_t)m * m));
        bool s = use_std && I > 0 && J > 0 && K > 0 && I + 1 < m && J + 1 < m && K + 1 < m;
        if (s) {
            const size_t sy = (size_t)m, sz = (size_t)m * m;
            const size_t nb[6] = {o + 1, o - 1, o + sy, o - sy, o + sz, o - sz};
            for (int c = 0; c < 8; ++c) {
                s = s && same_bits(ldl(L.diag + o * 8 + c), L.tab[c]);
                for (int ax = 0; ax < 3; ++ax) {
                    s = s && same_bits(ldl(L.off + o * 24 + ax * 8 + c), L.tab[8 + ax * 8 + c]);
                    s = s && same_bits(ldl(L.off + nb[2 * ax + 1] * 24 + ax * 8 + c), L.tab[8 + ax * 8 + c]);
                }
                for (int k = 0; k < 6; ++k) s = s && ldl(L.diag + nb[k] * 8 + c) > L.dmin;
            }
        }
        L.sf[o] = s ? 1 : 0;
    }
}

// ---------------------------------------------------------------- multigrid items
// residual restriction (k_resrestrict :306-316) + optional zero-guess red sweep at the coarse
// level (k_restrict_red :444-459).  A child's neighbour inside the 2x2x2 block is another child
// (the same address, loaded once); one outside is clamped to the child itself where missing.
template <bool STD, class LT>
__device__ __forceinline__ void restrict_item(const LvlT<LT>& F, const LvlT<LT>& C, int c, int K, int J, int I,
                                              bool do_red) {
    const int m = F.m;
    uint32_t co[8], nb[8][6];
    bool e[8][6];
    real r[8], x0[8], xn[8][6];
    int f = 1;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int dz = k >> 2, dy = (k >> 1) & 1, dx = k & 1;
        const int z = 2 * K + dz, y = 2 * J + dy, x = 2 * I + dx;
        e[k][0] = dx == 0 || x + 1 < m;
        e[k][1] = dx == 1 || x > 0;
        e[k][2] = dy == 0 || y + 1 < m;
        e[k][3] = dy == 1 || y > 0;
        e[k][4] = dz == 0 || z + 1 < m;
        e[k][5] = dz == 1 || z > 0;
        co[k] = lcell(m, z, y, x);
        nb[k][0] = e[k][0] ? lcell(m, z, y, x + 1) : co[k];
        nb[k][1] = e[k][1] ? lcell(m, z, y, x - 1) : co[k];
        nb[k][2] = e[k][2] ? lcell(m, z, y + 1, x) : co[k];
        nb[k][3] = e[k][3] ? lcell(m, z, y - 1, x) : co[k];
        nb[k][4] = e[k][4] ? lcell(m, z + 1, y, x) : co[k];
        nb[k][5] = e[k][5] ? lcell(m, z - 1, y, x) : co[k];
        if (STD) f &= (int)F.sf[co[k]];
        const int col = (x + y + z) & 1;
        r[k] = ldl(F.r + lsto(m, co[k], col) * 8 + c);
        x0[k] = ldl(F.x + lsto(m, co[k], col) * 8 + c);
#pragma unroll
        for (int j = 0; j < 6; ++j) xn[k][j] = ldl(F.x + lsto(m, nb[k][j], e[k][j] ? col ^ 1 : col) * 8 + c);
    }
    const Tab tb = STD ? load_tab(F, c) : Tab{};
    real res[8];
    if (STD && __all(f != 0)) {
        const real cf[6] = {tb.o[0], tb.o[0], tb.o[1], tb.o[1], tb.o[2], tb.o[2]};
#pragma unroll
        for (int k = 0; k < 8; ++k) res[k] = r[k] - tb.d * x0[k] - offsum_v(cf, xn[k], e[k]);
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            real cf[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) cf[j] = ldl(F.off + ((j & 1) ? nb[k][j] : co[k]) * 24 + (j >> 1) * 8 + c);
            res[k] = r[k] - ldl(F.diag + co[k] * 8 + c) * x0[k] - offsum_v(cf, xn[k], e[k]);
        }
    }
    real acc = RL(0.0);
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += res[k];
    acc = RL(0.125) * acc;
    const uint32_t oc = lcell(C.m, K, J, I) * 8 + c, xc = lsto(C.m, K, J, I) * 8 + c;
    stl(C.r + xc, acc);
    if (do_red && ((I + J + K) & 1) == 0) {
        // the coarse red sweep reads the right-hand side as stored
        const real d = ldl(C.diag + oc), ra = as_stored<LT>(acc);
        stl(C.x + xc, d <= C.dmin ? RL(0.0) : ra / d);
    }
}

// constant prolongation (k_prolong :319-325)
template <class LT>
__device__ __forceinline__ void prolong_item(const LvlT<LT>& F, const LvlT<LT>& C, int c, int K, int J, int I) {
    const uint32_t o = lcell(F.m, K, J, I) * 8 + c, ox = lsto(F.m, K, J, I) * 8 + c;
    const real x = ldl(F.x + ox), xc = ldl(C.x + lsto(C.m, K >> 1, J >> 1, I >> 1) * 8 + c);
    if (ldl(F.diag + o) <= F.dmin) return;
    stl(F.x + ox, x + xc);
}

// The black post-sweep with the prolongation fused in (prolong_black_even below): equal to
// prolonging every cell, then sweeping black (a GS update ignores the cell's own old value; red
// cells are rewritten by the red sweep that follows).  Same rounding as the unfused pair
// (pressure.py:396-423 vs :902-931).

// ---------------------------------------------------------------- even-level fast paths
// Every level launched on its own has even m (the hierarchy stops at the first odd size, which
// the single-workgroup coarse kernel handles), so its x and r are colour-split: cell
// o = lcell(K, J, I) of colour col = (I + J + K) & 1 is stored at col m^3/2 + h, h = o >> 1, and
// o = 2h + b with b = I & 1.  Its six neighbours, all of the other colour, are stored at
//   +I: sb + b,   -I: sb + b - 1,   +-J: sb +- m/2,   +-K: sb +- m^2/2,   sb = (col ^ 1) m^3/2 + h,
// and item t of a one-colour launch has h = t >> 3, component c = t & 7.  The items below load
// exactly the values the generic items load (a missing neighbour is read at sb instead of at the
// cell itself, and either value is discarded), in the same order of arithmetic, so their results
// are the same bit for bit; they spend fewer instructions on addresses, and for power-of-two m
// the item decomposition is shifts instead of divisions.
struct LevelShape {
    uint32_t m, mh, sz, half;   // m, m/2, m^2/2, m^3/2
    int lm;                     // log2 m for a power of two, else -1
};

__device__ __forceinline__ LevelShape level_shape(int m) {
    LevelShape s;
    s.m = (uint32_t)m;
    s.mh = s.m >> 1;
    s.sz = s.m * s.mh;
    s.half = s.m * s.sz;
    s.lm = (m & (m - 1)) == 0 ? (int)__builtin_ctz((unsigned)m) : -1;
    return s;
}

// q = h / (m/2), Ih = h % (m/2), K = q / m, J = q % m
__device__ __forceinline__ void split_half(const LevelShape& s, uint32_t h, int& K, int& J, uint32_t& Ih) {
    uint32_t q;
    if (s.lm > 0) {
        Ih = h & (s.mh - 1u);
        q = h >> (s.lm - 1);
        J = (int)(q & (s.m - 1u));
        K = (int)(q >> s.lm);
    } else {
        q = h / s.mh;
        Ih = h - q * s.mh;
        K = (int)(q / s.m);
        J = (int)(q - (uint32_t)K * s.m);
    }
}

// cell -> (K, J, I) of a level of size m (natural order)
__device__ __forceinline__ void split_cell(const LevelShape& s, uint32_t cell, int& K, int& J, int& I) {
    if (s.lm >= 0) {
        I = (int)(cell & (s.m - 1u));
        const uint32_t q = cell >> s.lm;
        J = (int)(q & (s.m - 1u));
        K = (int)(q >> s.lm);
    } else {
        const uint32_t q = cell / s.m;
        I = (int)(cell - q * s.m);
        K = (int)(q / s.m);
        J = (int)(q - (uint32_t)K * s.m);
    }
}

struct HalfItem {
    int c, K, J, I;
    uint32_t h, o;   // offset in the colour half, natural cell index
};

// Components per lane of the even-level sweeps: 2 with 16-bit storage (the two components of a
// lane are one 4-byte access, so a wave's access is one contiguous 256-byte run, as with one
// 4-byte component per lane at FP32), 1 otherwise.  The arithmetic of each component is the
// one-component item's, so the results do not depend on CP.  The restriction keeps one component
// per lane: with two, its 40 loads per item need 167 VGPRs, and at 3 waves per SIMD it ran slower
// (30.9 against 28.3 us at level 1 of 256^3; the prolongation went from 39.3 to 34.7 us).
template <class LT>
constexpr int level_cp() {
    return sizeof(LT) == 2 ? 2 : 1;
}
template <int CP>
constexpr uint32_t cp_shift() {
    return CP == 2 ? 2u : 3u;   // log2 of the lanes per cell (8 / CP)
}

// item t of a one-colour launch (m even), CP components from it.c; false past the last item
// (4 m^3 < 2^32 for m <= 512)
template <int CP>
__device__ __forceinline__ bool half_item_even(const LevelShape& s, uint32_t t, int color, HalfItem& it) {
    if (t >= s.half * (8u / CP)) return false;
    it.c = (int)(t & ((8u / CP) - 1u)) * CP;
    it.h = t >> cp_shift<CP>();
    uint32_t Ih;
    split_half(s, it.h, it.K, it.J, Ih);
    const int b = (color + it.J + it.K) & 1;
    it.I = (int)(2u * Ih) + b;
    it.o = 2u * it.h + (uint32_t)b;
    return true;
}

// storage offsets of the six neighbours (sb where one is missing) and their existence
struct EvenNb {
    uint32_t s[6];
    bool e[6];
};

__device__ __forceinline__ EvenNb even_neighbours(const LevelShape& s, int color, const HalfItem& it) {
    EvenNb n;
    const uint32_t sb = (uint32_t)(color ^ 1) * s.half + it.h, b = (uint32_t)(it.I & 1);
    n.e[0] = (uint32_t)it.I + 1u < s.m; n.e[1] = it.I > 0;
    n.e[2] = (uint32_t)it.J + 1u < s.m; n.e[3] = it.J > 0;
    n.e[4] = (uint32_t)it.K + 1u < s.m; n.e[5] = it.K > 0;
    n.s[0] = n.e[0] ? sb + b : sb;
    n.s[1] = n.e[1] ? sb + b - 1u : sb;
    n.s[2] = n.e[2] ? sb + s.mh : sb;
    n.s[3] = n.e[3] ? sb - s.mh : sb;
    n.s[4] = n.e[4] ? sb + s.sz : sb;
    n.s[5] = n.e[5] ? sb - s.sz : sb;
    return n;
}

// value of component c of storage cell s: the byte offset 2 (8 s + c) or 4 (8 s + c) stays below
// 2^32 (8 m^3 <= 2^30 values per array), so it is a 32-bit offset from the array's base (one
// address register per load, the base in scalar registers)
template <class LT>
__device__ __forceinline__ real ldc(const LT* base, uint32_t s, int c) {
    return ldl(reinterpret_cast<const LT*>(reinterpret_cast<const char*>(base) + (s * 8u + (uint32_t)c) * (uint32_t)sizeof(LT)));
}
template <class LT>
__device__ __forceinline__ void stc(LT* base, uint32_t s, int c, real v) {
    stl(reinterpret_cast<LT*>(reinterpret_cast<char*>(base) + (s * 8u + (uint32_t)c) * (uint32_t)sizeof(LT)), v);
}

// CP consecutive values from p (CP values aligned to CP elements): one load or store of them all
template <class LT>
struct alignas(2 * sizeof(LT)) LPair {
    LT a, b;
};
template <int CP, class LT>
__device__ __forceinline__ void ldpv(const LT* p, real (&v)[CP]) {
    if constexpr (CP == 1) {
        v[0] = ldl(p);
    } else {
        const LPair<LT> w = *reinterpret_cast<const LPair<LT>*>(p);
        v[0] = ldl(&w.a);
        v[1] = ldl(&w.b);
    }
}
template <int CP, class LT>
__device__ __forceinline__ void stpv(LT* p, const real (&v)[CP]) {
    if constexpr (CP == 1) {
        stl(p, v[0]);
    } else {
        LPair<LT> w;
        stl(&w.a, v[0]);
        stl(&w.b, v[1]);
        *reinterpret_cast<LPair<LT>*>(p) = w;
    }
}
// components c .. c + CP - 1 of storage cell s (ldc / stc)
template <int CP, class LT>
__device__ __forceinline__ void ldcv(const LT* base, uint32_t s, int c, real (&v)[CP]) {
    ldpv<CP>(reinterpret_cast<const LT*>(reinterpret_cast<const char*>(base) + (s * 8u + (uint32_t)c) * (uint32_t)sizeof(LT)), v);
}
template <int CP, class LT>
__device__ __forceinline__ void stcv(LT* base, uint32_t s, int c, const real (&v)[CP]) {
    stpv<CP>(reinterpret_cast<LT*>(reinterpret_cast<char*>(base) + (s * 8u + (uint32_t)c) * (uint32_t)sizeof(LT)), v);
}

// offsum_v with every neighbour present (a standard cell): the same sums without the tests
__device__ __forceinline__ real offsum_all(const real* cf, const real* xn) {
    real a = RL(0.0);
#pragma unroll
    for (int k = 0; k < 6; ++k) a = a + cf[k] * xn[k];
    return a;
}

// smooth_item<true> (non-zero guess) of a one-colour item, components it.c .. it.c + CP - 1; tb:
// the level table of those components (the same for every item of a thread)
template <int CP, class LT>
__device__ __forceinline__ void smooth_even(const LvlT<LT>& L, const LevelShape& s, int color, const HalfItem& it,
                                            const Tab (&tb)[CP]) {
    const int c = it.c;
    const uint32_t own = (uint32_t)color * s.half + it.h;
    const EvenNb n = even_neighbours(s, color, it);
    real r[CP];
    ldcv<CP>(L.r, own, c, r);
    const int f = L.sf[it.o];
    real xn[6][CP];
#pragma unroll
    for (int k = 0; k < 6; ++k) ldcv<CP>(L.x, n.s[k], c, xn[k]);
    real out[CP];
    if (__all(f != 0)) {
        // a standard cell has all six neighbours
#pragma unroll
        for (int q = 0; q < CP; ++q) {
            const real d = tb[q].d;
            const real cf[6] = {tb[q].o[0], tb[q].o[0], tb[q].o[1], tb[q].o[1], tb[q].o[2], tb[q].o[2]};
            const real xq[6] = {xn[0][q], xn[1][q], xn[2][q], xn[3][q], xn[4][q], xn[5][q]};
            const real acc = r[q] - offsum_all(cf, xq);
            out[q] = d <= L.dmin ? RL(0.0) : acc / d;
        }
    } else {
        const LNb nn = level_neighbours(L.m, it.o, it.K, it.J, it.I);
        real dg[CP], cf[6][CP];
        ldpv<CP>(L.diag + it.o * 8u + c, dg);
#pragma unroll
        for (int k = 0; k < 6; ++k) ldpv<CP>(L.off + ((k & 1) ? nn.c[k] : it.o) * 24u + (k >> 1) * 8 + c, cf[k]);
#pragma unroll
        for (int q = 0; q < CP; ++q) {
            const real cq[6] = {cf[0][q], cf[1][q], cf[2][q], cf[3][q], cf[4][q], cf[5][q]};
            const real xq[6] = {xn[0][q], xn[1][q], xn[2][q], xn[3][q], xn[4][q], xn[5][q]};
            const real acc = r[q] - offsum_v(cq, xq, n.e);
            out[q] = dg[q] <= L.dmin ? RL(0.0) : acc / dg[q];
        }
    }
    stcv<CP>(L.x, own, c, out);
}

// prolong_black_item of a black item, components it.c .. it.c + CP - 1.  The coarse values: the
// cell's own coarse cell and, along each axis, the coarse cell across the face of the 2x2x2 block
// the cell touches (the +ax one for an odd coordinate, the -ax one for an even), which holds one
// of its two neighbours.
template <int CP, class LT>
__device__ __forceinline__ void prolong_black_even(const LvlT<LT>& F, const LvlT<LT>& C, const LevelShape& s,
                                                   const HalfItem& it, const Tab (&tb)[CP]) {
    const int c = it.c, I = it.I, J = it.J, K = it.K;
    const uint32_t own = s.half + it.h;
    const EvenNb n = even_neighbours(s, 1, it);
    const int f = F.sf[it.o];
    real r[CP];
    ldcv<CP>(F.r, own, c, r);
    const int Ic = I >> 1, Jc = J >> 1, Kc = K >> 1, bI = I & 1, bJ = J & 1, bK = K & 1;
    const bool ox = bI ? n.e[0] : n.e[1], oy = bJ ? n.e[2] : n.e[3], oz = bK ? n.e[4] : n.e[5];
    real xn[6][CP];
#pragma unroll
    for (int k = 0; k < 6; ++k) ldcv<CP>(F.x, n.s[k], c, xn[k]);
    // storage of the coarse cells: colour-split like the fine level (its neighbours are of the
    // other colour) unless the coarse size is odd
    uint32_t s0, sx, sy, sz;
    const uint32_t mc = (uint32_t)C.m;
    if ((mc & 1u) == 0) {
        const uint32_t hc = mc >> 1, half_c = mc * mc * hc, oc = ((uint32_t)Kc * mc + (uint32_t)Jc) * mc + (uint32_t)Ic;
        const uint32_t colc = (uint32_t)((Ic + Jc + Kc) & 1), sbc = (colc ^ 1u) * half_c + (oc >> 1);
        const uint32_t bc = (uint32_t)(Ic & 1);
        s0 = colc * half_c + (oc >> 1);
        sx = ox ? (bI ? sbc + bc : sbc + bc - 1u) : s0;
        sy = oy ? (bJ ? sbc + hc : sbc - hc) : s0;
        sz = oz ? (bK ? sbc + mc * hc : sbc - mc * hc) : s0;
    } else {
        const int Io = ox ? Ic + 2 * bI - 1 : Ic, Jo = oy ? Jc + 2 * bJ - 1 : Jc, Ko = oz ? Kc + 2 * bK - 1 : Kc;
        s0 = lsto(C.m, Kc, Jc, Ic);
        sx = lsto(C.m, Kc, Jc, Io);
        sy = lsto(C.m, Kc, Jo, Ic);
        sz = lsto(C.m, Ko, Jc, Ic);
    }
    real xc0[CP], xcx[CP], xcy[CP], xcz[CP];
    ldcv<CP>(C.x, s0, c, xc0);
    ldcv<CP>(C.x, sx, c, xcx);
    ldcv<CP>(C.x, sy, c, xcy);
    ldcv<CP>(C.x, sz, c, xcz);
    real out[CP];
    if (__all(f != 0)) {
        // a standard cell has all six neighbours, each with diag > dmin
#pragma unroll
        for (int q = 0; q < CP; ++q) {
            const real xcn[6] = {bI ? xcx[q] : xc0[q], bI ? xc0[q] : xcx[q], bJ ? xcy[q] : xc0[q],
                                 bJ ? xc0[q] : xcy[q], bK ? xcz[q] : xc0[q], bK ? xc0[q] : xcz[q]};
            const real d = tb[q].d;
            const real cf[6] = {tb[q].o[0], tb[q].o[0], tb[q].o[1], tb[q].o[1], tb[q].o[2], tb[q].o[2]};
            real xpv[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) xpv[k] = as_stored<LT>(xn[k][q] + xcn[k]);
            const real a = offsum_all(cf, xpv);
            out[q] = d <= F.dmin ? RL(0.0) : (r[q] - a) / d;
        }
    } else {
        const LNb nn = level_neighbours(F.m, it.o, K, J, I);
        real dg[CP], cf[6][CP], dn[6][CP];
        ldpv<CP>(F.diag + it.o * 8u + c, dg);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            ldpv<CP>(F.off + ((k & 1) ? nn.c[k] : it.o) * 24u + (k >> 1) * 8 + c, cf[k]);
            ldpv<CP>(F.diag + nn.c[k] * 8u + c, dn[k]);
        }
#pragma unroll
        for (int q = 0; q < CP; ++q) {
            const real xcn[6] = {bI ? xcx[q] : xc0[q], bI ? xc0[q] : xcx[q], bJ ? xcy[q] : xc0[q],
                                 bJ ? xc0[q] : xcy[q], bK ? xcz[q] : xc0[q], bK ? xc0[q] : xcz[q]};
            real cq[6], xpv[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                cq[k] = cf[k][q];
                // x of the neighbour after the constant prolongation (which skips diag <= dmin)
                xpv[k] = dn[k][q] <= F.dmin ? xn[k][q] : as_stored<LT>(xn[k][q] + xcn[k]);
            }
            const real a = offsum_v(cq, xpv, n.e);
            out[q] = dg[q] <= F.dmin ? RL(0.0) : (r[q] - a) / dg[q];
        }
    }
    stcv<CP>(F.x, own, c, out);
}

// restrict_item<true> onto coarse cell (K, J, I) from an even fine level, components c ..
// c + CP - 1.  Child k = 4 dz + 2 dy + dx is fine cell (2K + dz, 2J + dy, 2I + dx) of colour
// (dx + dy + dz) & 1, stored at colour m^3/2 + base(dz, dy) with base = ((2K + dz) m + 2J + dy)
// m/2 + I.  A child's neighbour inside the block is a sibling (its x0 is that value); one outside
// is loaded from the other colour at the offsets of even_neighbours.
template <int CP, class LT>
__device__ __forceinline__ void restrict_even(const LvlT<LT>& F, const LvlT<LT>& C, const LevelShape& s, int c, int K,
                                              int J, int I, bool do_red, const Tab (&tb)[CP]) {
    const uint32_t m = s.m, mc = s.mh;
    uint32_t base[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) base[q] = ((2u * K + (q >> 1)) * m + 2u * J + (q & 1)) * s.mh + (uint32_t)I;
    const bool in_x[2] = {I > 0, (uint32_t)I + 1u < mc}, in_y[2] = {J > 0, (uint32_t)J + 1u < mc},
               in_z[2] = {K > 0, (uint32_t)K + 1u < mc};
    real r[8][CP], x0[8][CP], xo[8][3][CP];
    int f = 1;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int dz = k >> 2, dy = (k >> 1) & 1, dx = k & 1, col = (dx + dy + dz) & 1;
        const uint32_t bs = base[k >> 1], st = (uint32_t)col * s.half + bs, sb = (uint32_t)(col ^ 1) * s.half + bs;
        f &= (int)F.sf[2u * bs + (uint32_t)dx];
        ldcv<CP>(F.r, st, c, r[k]);
        ldcv<CP>(F.x, st, c, x0[k]);
        // the outer neighbour along each axis: -ax for an even offset, +ax for an odd one
        const uint32_t ax_x = in_x[dx] ? (dx ? sb + 1u : sb - 1u) : sb;
        const uint32_t ax_y = in_y[dy] ? (dy ? sb + s.mh : sb - s.mh) : sb;
        const uint32_t ax_z = in_z[dz] ? (dz ? sb + s.sz : sb - s.sz) : sb;
        ldcv<CP>(F.x, ax_x, c, xo[k][0]);
        ldcv<CP>(F.x, ax_y, c, xo[k][1]);
        ldcv<CP>(F.x, ax_z, c, xo[k][2]);
    }
    bool e[8][6];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int dz = k >> 2, dy = (k >> 1) & 1, dx = k & 1;
        // +ax, -ax along x, y, z (the order of level_neighbours)
        e[k][0] = dx ? in_x[1] : true;
        e[k][1] = dx ? true : in_x[0];
        e[k][2] = dy ? in_y[1] : true;
        e[k][3] = dy ? true : in_y[0];
        e[k][4] = dz ? in_z[1] : true;
        e[k][5] = dz ? true : in_z[0];
    }
    // neighbour j of child k, component q: a sibling's x0 or the loaded outer value
    auto xnb = [&](int k, int j, int q) -> real {
        const int dz = k >> 2, dy = (k >> 1) & 1, dx = k & 1;
        switch (j) {
            case 0: return dx ? xo[k][0][q] : x0[k | 1][q];
            case 1: return dx ? x0[k & 6][q] : xo[k][0][q];
            case 2: return dy ? xo[k][1][q] : x0[k | 2][q];
            case 3: return dy ? x0[k & 5][q] : xo[k][1][q];
            case 4: return dz ? xo[k][2][q] : x0[k | 4][q];
            default: return dz ? x0[k & 3][q] : xo[k][2][q];
        }
    };
    real res[CP][8];
    if (__all(f != 0)) {
        // standard children have all six neighbours
#pragma unroll
        for (int q = 0; q < CP; ++q) {
            const real cf[6] = {tb[q].o[0], tb[q].o[0], tb[q].o[1], tb[q].o[1], tb[q].o[2], tb[q].o[2]};
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                real xq[6];
#pragma unroll
                for (int j = 0; j < 6; ++j) xq[j] = xnb(k, j, q);
                res[q][k] = r[k][q] - tb[q].d * x0[k][q] - offsum_all(cf, xq);
            }
        }
    } else {
        const uint32_t sy = m, sz = m * m;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t co = 2u * base[k >> 1] + (uint32_t)(k & 1);
            const uint32_t nb[6] = {e[k][0] ? co + 1u : co, e[k][1] ? co - 1u : co, e[k][2] ? co + sy : co,
                                    e[k][3] ? co - sy : co, e[k][4] ? co + sz : co, e[k][5] ? co - sz : co};
            real cf[6][CP], dg[CP];
#pragma unroll
            for (int j = 0; j < 6; ++j) ldpv<CP>(F.off + ((j & 1) ? nb[j] : co) * 24u + (j >> 1) * 8 + c, cf[j]);
            ldpv<CP>(F.diag + co * 8u + c, dg);
#pragma unroll
            for (int q = 0; q < CP; ++q) {
                real cq[6], xq[6];
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    cq[j] = cf[j][q];
                    xq[j] = xnb(k, j, q);
                }
                res[q][k] = r[k][q] - dg[q] * x0[k][q] - offsum_v(cq, xq, e[k]);
            }
        }
    }
    real acc[CP];
#pragma unroll
    for (int q = 0; q < CP; ++q) {
        real a = RL(0.0);
#pragma unroll
        for (int k = 0; k < 8; ++k) a += res[q][k];
        acc[q] = RL(0.125) * a;
    }
    const uint32_t oc = lcell(C.m, K, J, I) * 8 + c, xc = lsto(C.m, K, J, I) * 8 + c;
    stpv<CP>(C.r + xc, acc);
    if (do_red && ((I + J + K) & 1) == 0) {
        // the coarse red sweep reads the right-hand side as stored
        real dg[CP], xr[CP];
        ldpv<CP>(C.diag + oc, dg);
#pragma unroll
        for (int q = 0; q < CP; ++q) xr[q] = dg[q] <= C.dmin ? RL(0.0) : as_stored<LT>(acc[q]) / dg[q];
        stpv<CP>(C.x + xc, xr);
    }
}

// item of this thread in a 1D launch, blocks renumbered per XCD (neighbouring rows in one L2)
__device__ __forceinline__ size_t xcd_item() {
    return (size_t)xcd_block((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
}

// full-size launch: item t -> (c, K, J, I).  Item counts stay below 2^32 (8 m^3 <= 2^30 for
// N <= 1024), so the decomposition is 32-bit unsigned arithmetic (a 64-bit division is a long
// emulated sequence per lane).
__device__ __forceinline__ void full_item(uint32_t t, uint32_t m, int& c, int& K, int& J, int& I) {
    c = (int)(t & 7u);
    const uint32_t cell = t >> 3, q = cell / m;
    I = (int)(cell - q * m);
    K = (int)(q / m);
    J = (int)(q - (uint32_t)K * m);
}

// Level launches use at most LEVEL_GRID / CP workgroups (the restriction restrict_grid), each
// striding over the items of the launch in passes of that many * 256 (blocks renumbered per XCD
// within a pass): two items per thread at level 1 of 256^3 for either storage.  A launch of a CG
// iteration that has already stopped then costs the dispatch of that many workgroups, not of one
// per 256 items.  Level 1 of 256^3 with FP16 storage (CP = 2): 8192 workgroups ran the sweep in
// 19.8 us and the prolongation in 28.2 us, against 24.6 and 34.0 us with 4096 (four items per
// thread) and 20.9 and 30.1 us with 16384 (one); with FP32 storage (CP = 1) 16384 workgroups took
// 29.0 and 38.1 us against 35.6 and 44.5 us with 8192.  The restriction (one component per lane)
// took 28.0 us with 4096 workgroups and 28.8 us with 8192 at FP16.
constexpr unsigned LEVEL_GRID = 16384;
// the restriction's workgroup cap: 4096 for 16-bit levels (28.0 us at level 1 of 256^3 against 28.8
// with 8192), 8192 for FP32 (42.1 us against 48.2 with 4096)
template <class LT>
constexpr unsigned restrict_grid() {
    return sizeof(LT) == 2 ? 4096u : 8192u;
}

// The logical blocks [0, T) of a striding level launch, split so that each XCD (workgroup b runs
// on XCD b % 8) sweeps one contiguous eighth of them: the neighbour planes an item reads are then
// mostly the ones its own XCD reads, in its own L2.  A launch of one pass (or a grid that is not a
// multiple of 8) keeps the per-pass XCD renumbering.
struct XcdRange {
    uint32_t first, end, step;
};
__device__ __forceinline__ XcdRange xcd_range(uint32_t T) {
    const uint32_t G = gridDim.x;
    XcdRange r;
    if (G >= T || G % NUM_XCD != 0) {
        r.first = (uint32_t)xcd_block((int)blockIdx.x, (int)G);
        r.end = T;
        r.step = G;
        return r;
    }
    const uint32_t x = blockIdx.x % NUM_XCD, per = (T + NUM_XCD - 1) / NUM_XCD;
    r.first = x * per + blockIdx.x / NUM_XCD;
    r.end = min((x + 1) * per, T);
    r.step = G / NUM_XCD;
    return r;
}

static unsigned level_grid(size_t items, int cp = 1) {
    const size_t b = (items + 255) / 256, cap = LEVEL_GRID / (unsigned)cp;
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

// the level tables of the CP components of this thread (blockDim and every stride are multiples
// of 8 / CP, so a thread's components are fixed)
template <int CP, class LT>
__device__ __forceinline__ void load_tabs(const LvlT<LT>& L, Tab (&tb)[CP]) {
    const int c = (int)(threadIdx.x & ((8u / CP) - 1u)) * CP;
#pragma unroll
    for (int q = 0; q < CP; ++q) tb[q] = load_tab(L, c + q);
}

template <class LT>
__global__ void __launch_bounds__(256) k_smooth_half(LvlT<LT> L, int color, const DevState* __restrict__ st, int k) {
    constexpr int CP = level_cp<LT>();
    if (cg_inactive(st, k)) return;
    const LevelShape sh = level_shape(L.m);
    Tab tb[CP];
    load_tabs<CP>(L, tb);
    const XcdRange rg = xcd_range((sh.half * (8u / CP) + blockDim.x - 1) / blockDim.x);
    for (uint32_t lb = rg.first; lb < rg.end; lb += rg.step) {
        HalfItem it;
        if (!half_item_even<CP>(sh, lb * blockDim.x + threadIdx.x, color, it)) return;
        smooth_even<CP>(L, sh, color, it, tb);
    }
}

// Restriction of the level below C.  For m >= RESTRICT_ROW the coarse cells go by columns: logical
// block b (XCD-renumbered) owns RESTRICT_ROW cells of one row (8 components each) and consecutive
// logical blocks walk up K, then J, then along I, so the fine planes two coarse planes share are
// read again while they are still in the L2.  Smaller levels use the natural order.
constexpr int RESTRICT_CP = 1;
constexpr int RESTRICT_ROW = 32 * RESTRICT_CP;   // coarse cells of a 256-thread workgroup

template <class LT>
__global__ void __launch_bounds__(256) k_restrict(LvlT<LT> F, LvlT<LT> C, int do_red, const DevState* __restrict__ st,
                                                  int k) {
    constexpr int CP = RESTRICT_CP;
    constexpr uint32_t ROW = RESTRICT_ROW, SH = cp_shift<CP>();
    if (cg_inactive(st, k)) return;
    const LevelShape fs = level_shape(F.m), cs = level_shape(C.m);
    const int m = C.m;
    const int c = (int)(threadIdx.x & ((8u / CP) - 1u)) * CP;   // blockDim and the stride are multiples of 8 / CP
    Tab tb[CP];
    load_tabs<CP>(F, tb);
    if (m >= (int)ROW) {
        const uint32_t um = (uint32_t)m, total = um * um * ((um + ROW - 1u) / ROW);
        const XcdRange rg = xcd_range(total);
        for (uint32_t lb = rg.first; lb < rg.end; lb += rg.step) {
            uint32_t col, seg;
            if (cs.lm >= 0) {
                col = lb >> cs.lm;
                seg = col >> cs.lm;
            } else {
                col = lb / um;
                seg = col / um;
            }
            const int K = (int)(lb - col * um);
            const int J = (int)(col - seg * um);
            const int I = (int)(seg * ROW + (threadIdx.x >> SH));
            if (I < m) restrict_even<CP>(F, C, fs, c, K, J, I, do_red != 0, tb);
        }
    } else {
        const uint32_t items = (uint32_t)m * m * m * (8u / CP);
        const XcdRange rg = xcd_range((items + blockDim.x - 1) / blockDim.x);
        for (uint32_t lb = rg.first; lb < rg.end; lb += rg.step) {
            const uint32_t t = lb * blockDim.x + threadIdx.x;
            if (t >= items) break;
            int K, J, I;
            split_cell(cs, t >> SH, K, J, I);
            restrict_even<CP>(F, C, fs, c, K, J, I, do_red != 0, tb);
        }
    }
}

// workgroups of a k_restrict launch onto coarse size m
template <class LT>
static unsigned restrict_blocks(int m) {
    constexpr int ROW = RESTRICT_ROW;
    constexpr unsigned G = restrict_grid<LT>();
    if (m >= ROW) return std::min((unsigned)m * (unsigned)m * (unsigned)((m + ROW - 1) / ROW), G);
    return std::min(level_grid((size_t)m * m * m * (8 / RESTRICT_CP)), G);
}

template <class LT>
__global__ void __launch_bounds__(256) k_prolong_black(LvlT<LT> F, LvlT<LT> C, const DevState* __restrict__ st, int k) {
    constexpr int CP = level_cp<LT>();
    if (cg_inactive(st, k)) return;
    const LevelShape sh = level_shape(F.m);
    Tab tb[CP];
    load_tabs<CP>(F, tb);
    const XcdRange rg = xcd_range((sh.half * (8u / CP) + blockDim.x - 1) / blockDim.x);
    for (uint32_t lb = rg.first; lb < rg.end; lb += rg.step) {
        HalfItem it;
        if (!half_item_even<CP>(sh, lb * blockDim.x + threadIdx.x, 1, it)) return;
        prolong_black_even<CP>(F, C, sh, it, tb);
    }
}

// The bottom of the V-cycle in ONE workgroup (k_coarse_block :550-591 generalised to any
// number of levels and to odd coarsest sizes): the V(1,1) of _vcycle_base (:902-931) on
// levels lv[0..n-1], starting with the zero-guess red sweep at lv[0].
constexpr int CB_THREADS = 1024;

template <class LT>
__device__ void cb_smooth(const LvlT<LT>& L, int color, bool zero_init) {
    const uint32_t m = (uint32_t)L.m;
    const uint32_t items = m * m * m * 8;
    for (uint32_t t = threadIdx.x; t < items; t += CB_THREADS) {
        int c, K, J, I;
        full_item(t, m, c, K, J, I);
        if (((I + J + K) & 1) == color) smooth_item<false>(L, c, K, J, I, zero_init);
    }
    __syncthreads();
}

// the V-cycle over levels lvl(0 .. n-1); lvl(j) gives level j (in global memory or in LDS)
template <class LT, class Get>
__device__ __forceinline__ void coarse_vcycle(int n, const Get& lvl) {
    for (int j = 0; j + 1 < n; ++j) {
        const auto& F = lvl(j);
        const auto& C = lvl(j + 1);
        cb_smooth(F, 0, true);
        cb_smooth(F, 1, false);
        const uint32_t items = (uint32_t)C.m * C.m * C.m * 8;
        for (uint32_t t = threadIdx.x; t < items; t += CB_THREADS) {
            int c, K, J, I;
            full_item(t, (uint32_t)C.m, c, K, J, I);
            restrict_item<false>(F, C, c, K, J, I, false);
        }
        __syncthreads();
    }
    const auto& B = lvl(n - 1);
    if (B.m == 1) {
        if (threadIdx.x < 8) {
            const real d = ldl(B.diag + threadIdx.x);
            stl(B.x + threadIdx.x, d <= B.dmin ? RL(0.0) : ldl(B.r + threadIdx.x) / d);
        }
        __syncthreads();
    } else {
        for (int s = 0; s < 2 * B.m; ++s) {
            cb_smooth(B, 0, s == 0);
            cb_smooth(B, 1, false);
        }
    }
    for (int j = n - 2; j >= 0; --j) {
        const auto& F = lvl(j);
        const auto& C = lvl(j + 1);
        const uint32_t items = (uint32_t)F.m * F.m * F.m * 8;
        for (uint32_t t = threadIdx.x; t < items; t += CB_THREADS) {
            int c, K, J, I;
            full_item(t, (uint32_t)F.m, c, K, J, I);
            prolong_item(F, C, c, K, J, I);
        }
        __syncthreads();
        cb_smooth(F, 1, false);
        cb_smooth(F, 0, false);
    }
}

template <class LT>
__global__ void __launch_bounds__(CB_THREADS) k_coarse(CoarseLevelsT<LT> cl, const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    coarse_vcycle<LT>(cl.n, [&](int j) -> const LvlT<LT>& { return cl.lv[j]; });
}

// The same V-cycle with every coarse level resident in LDS: x, r, diag and off of each level in
// the global layout and storage type (48 values per cell), so every load and store rounds as the
// global kernel's does and the results are the same bit for bit.  The operators and the
// right-hand side of the top level are read once, and only the top level's x is written back
// (the next finer level prolongs it); every sweep then waits on LDS instead of the L2.
template <class LT>
__host__ __device__ inline size_t coarse_lds_vals(const CoarseLevelsT<LT>& cl, int upto) {
    size_t off = 0;
    for (int i = 0; i < upto; ++i) off += (size_t)48 * cl.lv[i].m * cl.lv[i].m * cl.lv[i].m;
    return off;
}

template <class LT>
__global__ void __launch_bounds__(CB_THREADS) k_coarse_lds(CoarseLevelsT<LT> cl, const DevState* __restrict__ st,
                                                           int k) {
    extern __shared__ __align__(16) unsigned char cb_raw[];
    LT* cb_lds = reinterpret_cast<LT*>(cb_raw);
    if (cg_inactive(st, k)) return;
    auto lvl = [&](int j) {
        LvlT<LT> L = cl.lv[j];
        const size_t u = (size_t)L.m * L.m * L.m * 8;
        L.x = cb_lds + coarse_lds_vals(cl, j);
        L.r = L.x + u;
        L.diag = L.r + u;
        L.off = L.diag + u;
        return L;
    };
    for (int j = 0; j < cl.n; ++j) {
        const LvlT<LT> G = cl.lv[j], L = lvl(j);
        const uint32_t u = (uint32_t)G.m * G.m * G.m * 8;
        for (uint32_t t = threadIdx.x; t < u; t += CB_THREADS) {
            L.diag[t] = G.diag[t];
            if (j == 0) L.r[t] = G.r[t];
        }
        for (uint32_t t = threadIdx.x; t < 3 * u; t += CB_THREADS) L.off[t] = G.off[t];
    }
    __syncthreads();
    coarse_vcycle<LT>(cl.n, lvl);
    const LvlT<LT> L0 = lvl(0);
    const uint32_t u0 = (uint32_t)L0.m * L0.m * L0.m * 8;
    for (uint32_t t = threadIdx.x; t < u0; t += CB_THREADS) cl.lv[0].x[t] = L0.x[t];
}

constexpr size_t COARSE_LDS_MAX = 64 * 1024;   // bytes of dynamic LDS the resident kernel may use

// the coarse V-cycle of levels S.. on stream s: LDS-resident when it fits
template <class LT>
static void launch_coarse(const Pressure& P, hipStream_t s, int k) {
    CoarseLevelsT<LT> cl{};
    cl.n = P.nlev - P.S;
    for (int j = 0; j < cl.n; ++j) cl.lv[j] = lvl_as<LT>(P.lv[P.S + j]);
    const size_t bytes = coarse_lds_vals(cl, cl.n) * sizeof(LT);
    if (bytes <= COARSE_LDS_MAX) {
        hipLaunchKernelGGL(k_coarse_lds<LT>, dim3(1), dim3(CB_THREADS), bytes, s, cl, P.st, k);
        return;
    }
    hipLaunchKernelGGL(k_coarse<LT>, dim3(1), dim3(CB_THREADS), 0, s, cl, P.st, k);
}

// ================================================================= host side
static int grid_for(size_t items, int cap) {
    size_t b = (items + 255) / 256;
    if (b > (size_t)cap) b = (size_t)cap;
    return (int)(b < 1 ? 1 : b);
}

void Pressure::init_levels(int N_) {
    N = N_;
    nlev = 0;
    int m = N / 2, level = 1;
    for (;;) {
        Lvl L{};
        L.m = m;
        L.dmin = (real)(1.0e-3 * 6.0 * std::pow(0.25, level - 1));
        lv[nlev++] = L;
        if (m % 2 == 1 || m == 1) break;
        m /= 2;
        level += 1;
    }
    // levels from the first with m <= coarse_m down run in the single-workgroup kernel
    S = nlev - 1;
    for (int l = 0; l < nlev; ++l)
        if (lv[l].m <= coarse_m) { S = l; break; }
}

// floats that hold m3 flag bytes, a multiple of 8 (every level array stays 32-byte aligned)
static size_t flag_floats(size_t m3) { return ((m3 + 3) / 4 + 7) / 8 * 8; }

size_t Pressure::level_floats() const {
    size_t t = 0;
    for (int l = 0; l < nlev; ++l) {
        const size_t m3 = (size_t)lv[l].m * lv[l].m * lv[l].m;
        t += m3 * (8 + 8 + 8 + 24) + 32 + flag_floats(m3);
    }
    return t;
}

void Pressure::bind_levels(real* base) {
    for (int l = 0; l < nlev; ++l) {
        const size_t m3 = (size_t)lv[l].m * lv[l].m * lv[l].m;
        lv[l].x = base; base += 8 * m3;
        lv[l].r = base; base += 8 * m3;
        lv[l].diag = base; base += 8 * m3;
        lv[l].off = base; base += 24 * m3;
        lv[l].tab = base; base += 32;
        lv[l].sf = reinterpret_cast<uint8_t*>(base); base += flag_floats(m3);
    }
}

__global__ void k_set_changed(DevState* st) {
    if (threadIdx.x == 0) st->changed = 1;
}

void Pressure::setup(const real* solid, hipStream_t s) {
    const size_t n3 = (size_t)N * N * N;
    if (force_setup) {
        hipLaunchKernelGGL(k_set_changed, dim3(1), dim3(64), 0, s, st);
        force_setup = false;
    }
    hipLaunchKernelGGL(k_solid_check, dim3(grid_for(n3, 4096)), dim3(256), 0, s, n3, solid, fm, st);
    hipLaunchKernelGGL(k_mask, dim3(grid_for(n3, 4096)), dim3(256), 0, s, N, solid, fm, st);
    with_mg_type(mg_kind, [&](auto tag) {
        using LT = decltype(tag);
        const size_t m3 = (size_t)lv[0].m * lv[0].m * lv[0].m;
        hipLaunchKernelGGL(k_probe<LT>, dim3(grid_for(m3 * 8, 4096)), dim3(256), 0, s, N, fm, lvl_as<LT>(lv[0]), st);
        for (int l = 0; l + 1 < nlev; ++l) {
            const size_t mc3 = (size_t)lv[l + 1].m * lv[l + 1].m * lv[l + 1].m;
            hipLaunchKernelGGL(k_galerkin<LT>, dim3(grid_for(mc3 * 8, 4096)), dim3(256), 0, s, lvl_as<LT>(lv[l]),
                               lvl_as<LT>(lv[l + 1]), st);
        }
        CoarseLevelsT<LT> all{};
        all.n = nlev;
        for (int l = 0; l < nlev; ++l) all.lv[l] = lvl_as<LT>(lv[l]);
        hipLaunchKernelGGL(k_std_tab<LT>, dim3(1), dim3(64), 0, s, all, N, st);
        for (int l = 0; l < nlev; ++l) {
            const size_t ml3 = (size_t)lv[l].m * lv[l].m * lv[l].m;
            hipLaunchKernelGGL(k_std_flags<LT>, dim3(grid_for(ml3, 4096)), dim3(256), 0, s, lvl_as<LT>(lv[l]),
                               use_std ? 1 : 0, st);
        }
    });
    hipLaunchKernelGGL(k_clear_changed, dim3(1), dim3(64), 0, s, st);
}

}  // namespace vfx

namespace vfx {

// The multigrid part of one V(1,1) between the fused fine-grid ends (k_pre_resid_fwd before,
// k_inv_post after): black sweep + restriction on every level above the coarse kernel, the
// single-workgroup coarse cycle, then prolongation + black sweep and the red sweep on the way up.
void Pressure::vcycle_levels(int k, hipStream_t s) {
    with_mg_type(mg_kind, [&](auto tag) {
        using LT = decltype(tag);
        constexpr int CP = level_cp<LT>();
        for (int l = 0; l < S; ++l) {
            const size_t half = (size_t)lv[l].m * lv[l].m * (lv[l].m / 2) * (8 / CP);
            hipLaunchKernelGGL(k_smooth_half<LT>, dim3(level_grid(half, CP)), dim3(256), 0, s, lvl_as<LT>(lv[l]), 1, st, k);
            hipLaunchKernelGGL(k_restrict<LT>, dim3(restrict_blocks<LT>(lv[l + 1].m)), dim3(256), 0, s,
                               lvl_as<LT>(lv[l]), lvl_as<LT>(lv[l + 1]), (l + 1 < S) ? 1 : 0, st, k);
        }
        launch_coarse<LT>(*this, s, k);
        for (int l = S - 1; l >= 0; --l) {
            const size_t half = (size_t)lv[l].m * lv[l].m * (lv[l].m / 2) * (8 / CP);
            hipLaunchKernelGGL(k_prolong_black<LT>, dim3(level_grid(half, CP)), dim3(256), 0, s, lvl_as<LT>(lv[l]),
                               lvl_as<LT>(lv[l + 1]), st, k);
            hipLaunchKernelGGL(k_smooth_half<LT>, dim3(level_grid(half, CP)), dim3(256), 0, s, lvl_as<LT>(lv[l]), 0, st, k);
        }
    });
}

// Multigrid kernels for timing (ids relative to VFX_K_MATVEC).
bool Pressure::launch_single_mg(int which, hipStream_t s) {
    const size_t m0 = (size_t)lv[0].m;
    if (which != 10 && (which < 2 || which > 4)) return false;
    if (which != 10 && S < 1) return false;
    with_mg_type(mg_kind, [&](auto tag) {
        using LT = decltype(tag);
        const size_t half0 = m0 * m0 * (m0 / 2) * (8 / level_cp<LT>());
        switch (which) {
            case 2:
                hipLaunchKernelGGL(k_smooth_half<LT>, dim3(level_grid(half0, level_cp<LT>())), dim3(256), 0, s,
                                   lvl_as<LT>(lv[0]), 1, st, 0);
                break;
            case 3: {
                hipLaunchKernelGGL(k_restrict<LT>, dim3(restrict_blocks<LT>(lv[1].m)), dim3(256), 0, s,
                                   lvl_as<LT>(lv[0]), lvl_as<LT>(lv[1]), S > 1 ? 1 : 0, st, 0);
                break;
            }
            case 4:
                hipLaunchKernelGGL(k_prolong_black<LT>, dim3(level_grid(half0, level_cp<LT>())), dim3(256), 0, s,
                                   lvl_as<LT>(lv[0]), lvl_as<LT>(lv[1]), st, 0);
                break;
            default: launch_coarse<LT>(*this, s, 0); break;
        }
    });
    return true;
}

}  // namespace vfx
