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
#include <climits>

#include "sim3d_kernels.hpp"

namespace vfx {

// ---------------------------------------------------------------- fine-grid operators
__device__ __forceinline__ size_t axstride(int N, int ax) {
    return ax == 0 ? (size_t)1 : (ax == 1 ? (size_t)N : (size_t)N * N);
}

// g_ax (pressure.py:54-65) at linear cell i
__device__ __forceinline__ float g_ax(const float* __restrict__ p, const uint8_t* __restrict__ fm, size_t i,
                                      size_t st, int ax) {
    const int m = fm[i];
    const float c = p[i];
    float g = 0.0f;
    if ((m >> (2 * ax)) & 1) g += p[i + st] - c;
    if ((m >> (2 * ax + 1)) & 1) g += c - p[i - st];
    return g;
}

// kmul (pressure.py:68-80): (K p)_i at a fluid cell
__device__ __forceinline__ float kmul(const float* __restrict__ p, const uint8_t* __restrict__ fm, int N, size_t i) {
    const int m = fm[i];
    float d = 0.0f;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const size_t st = axstride(N, ax);
        const float gi = g_ax(p, fm, i, st, ax);
        if ((m >> (2 * ax)) & 1) d += gi + g_ax(p, fm, i + st, st, ax);
        if ((m >> (2 * ax + 1)) & 1) d -= gi + g_ax(p, fm, i - st, st, ax);
    }
    return -d;
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
__device__ __forceinline__ float div2(const float* __restrict__ vel, size_t n3, int m, int N, size_t i) {
    float d = 0.0f;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float* va = vel + ax * n3;
        const size_t st = axstride(N, ax);
        const float u = va[i];
        if ((m >> (2 * ax)) & 1) d += u + va[i + st];
        if ((m >> (2 * ax + 1)) & 1) d -= u + va[i - st];
    }
    return d;
}

// face mask of one cell from `solid` (k_mask pressure.py:115-143)
__device__ __forceinline__ int mask_from_solid(const float* __restrict__ solid, int N, int z, int y, int x) {
    const size_t i = idx3(N, z, y, x);
    if (solid[i] > 0.5f) return 0;
    const size_t sz = (size_t)N * N;
    int m = FLUID;
    if (x + 1 < N && solid[i + 1] <= 0.5f) m |= 1;
    if (x > 0 && solid[i - 1] <= 0.5f) m |= 2;
    if (y + 1 < N && solid[i + N] <= 0.5f) m |= 4;
    if (y > 0 && solid[i - N] <= 0.5f) m |= 8;
    if (z + 1 < N && solid[i + sz] <= 0.5f) m |= 16;
    if (z > 0 && solid[i - sz] <= 0.5f) m |= 32;
    return m;
}

// ---------------------------------------------------------------- setup (solid changes only)
// k_solid_changed (pressure.py:146-153) against the cached face mask: the operators depend on
// solid only through (solid > 0.5), which the mask's fluid bit records.
__global__ void __launch_bounds__(256) k_solid_check(size_t n3, const float* __restrict__ solid,
                                                     const uint8_t* __restrict__ fm, DevState* __restrict__ st) {
    bool diff = false;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x)
        diff |= (solid[i] > 0.5f) != (fm[i] == 0);
    if (__any(diff) && (threadIdx.x & 63) == 0) st->changed = 1;
}

__global__ void __launch_bounds__(256) k_mask(int N, const float* __restrict__ solid, uint8_t* __restrict__ fm,
                                              const DevState* __restrict__ st) {
    if (!st->changed) return;
    const size_t n3 = (size_t)N * N * N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / ((size_t)N * N));
        fm[i] = (uint8_t)mask_from_solid(solid, N, z, y, x);
    }
}

// hsign / tv / gt / axcell (pressure.py:179-217)
__device__ __forceinline__ float hsign(int c, int dz, int dy, int dx) {
    const int e = (c & 1) * dx + ((c >> 1) & 1) * dy + ((c >> 2) & 1) * dz;
    return (e & 1) ? -HS : HS;
}

__device__ __forceinline__ float tv(int c, int BK, int BJ, int BI, int z, int y, int x) {
    if ((z >> 1) == BK && (y >> 1) == BJ && (x >> 1) == BI) return hsign(c, z & 1, y & 1, x & 1);
    return 0.0f;
}

__device__ __forceinline__ int mask_at(const uint8_t* __restrict__ fm, int N, int z, int y, int x) {
    if (x < 0 || y < 0 || z < 0 || z >= N || y >= N || x >= N) return 0;
    return fm[idx3(N, z, y, x)];
}

__device__ __forceinline__ float gt(const uint8_t* __restrict__ fm, int N, int c, int BK, int BJ, int BI, int z,
                                    int y, int x, int ax) {
    const int m = mask_at(fm, N, z, y, x);
    const int ex = ax == 0, ey = ax == 1, ez = ax == 2;
    const float t0 = tv(c, BK, BJ, BI, z, y, x);
    float g = 0.0f;
    if ((m >> (2 * ax)) & 1) g += tv(c, BK, BJ, BI, z + ez, y + ey, x + ex) - t0;
    if ((m >> (2 * ax + 1)) & 1) g += t0 - tv(c, BK, BJ, BI, z - ez, y - ey, x - ex);
    return g;
}

__device__ __forceinline__ void axcell(int ax, int oz, int oy, int ox, int a, int b, int c, int& z, int& y, int& x) {
    if (ax == 0) { z = oz + c; y = oy + b; x = ox + a; }
    else if (ax == 1) { z = oz + c; y = oy + a; x = ox + b; }
    else { z = oz + a; y = oy + c; x = ox + b; }
}

__device__ __forceinline__ size_t lcell(int m, int K, int J, int I) { return ((size_t)K * m + J) * m + I; }

// k_probe1 (pressure.py:220-246): exact component-diagonal of T K T^T, one (cell, c) per item.
__global__ void __launch_bounds__(256) k_probe(int N, const uint8_t* __restrict__ fm, Lvl L,
                                               const DevState* __restrict__ st) {
    if (!st->changed) return;
    const int m = L.m;
    const size_t items = (size_t)m * m * m * 8;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < items; t += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(t & 7);
        const size_t cell = t >> 3;
        const int I = (int)(cell % m), J = (int)((cell / m) % m), K = (int)(cell / ((size_t)m * m));
        const int oz = 2 * K, oy = 2 * J, ox = 2 * I;
        float dsum = 0.0f;
        for (int ax = 0; ax < 3; ++ax)
            for (int b = 0; b < 2; ++b)
                for (int cc = 0; cc < 2; ++cc)
                    for (int a = 0; a < 4; ++a) {
                        int z, y, x;
                        axcell(ax, oz, oy, ox, a - 1, b, cc, z, y, x);
                        const float v = gt(fm, N, c, K, J, I, z, y, x, ax);
                        dsum += v * v;
                    }
        L.diag[cell * 8 + c] = dsum;
        for (int ax = 0; ax < 3; ++ax) {
            const int nK = K + (ax == 2), nJ = J + (ax == 1), nI = I + (ax == 0);
            float s = 0.0f;
            if (nK < m && nJ < m && nI < m) {
                for (int b = 0; b < 2; ++b)
                    for (int cc = 0; cc < 2; ++cc)
                        for (int a = 0; a < 2; ++a) {
                            int z, y, x;
                            axcell(ax, oz, oy, ox, 1 + a, b, cc, z, y, x);
                            s += gt(fm, N, c, K, J, I, z, y, x, ax) * gt(fm, N, c, nK, nJ, nI, z, y, x, ax);
                        }
            }
            L.off[cell * 24 + ax * 8 + c] = s;
        }
    }
}

// k_galerkin (pressure.py:249-268)
__global__ void __launch_bounds__(256) k_galerkin(Lvl F, Lvl C, const DevState* __restrict__ st) {
    if (!st->changed) return;
    const int m = C.m, mf = F.m;
    const size_t items = (size_t)m * m * m * 8;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < items; t += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(t & 7);
        const size_t cell = t >> 3;
        const int I = (int)(cell % m), J = (int)((cell / m) % m), K = (int)(cell / ((size_t)m * m));
        const int oz = 2 * K, oy = 2 * J, ox = 2 * I;
        float d = 0.0f;
        for (int k = 0; k < 8; ++k) d += F.diag[lcell(mf, oz + (k >> 2), oy + ((k >> 1) & 1), ox + (k & 1)) * 8 + c];
        for (int ax = 0; ax < 3; ++ax) {
            float oi = 0.0f, oo = 0.0f;
            for (int b = 0; b < 2; ++b)
                for (int cc = 0; cc < 2; ++cc) {
                    int z, y, x;
                    axcell(ax, oz, oy, ox, 0, b, cc, z, y, x);
                    oi += F.off[lcell(mf, z, y, x) * 24 + ax * 8 + c];
                    axcell(ax, oz, oy, ox, 1, b, cc, z, y, x);
                    oo += F.off[lcell(mf, z, y, x) * 24 + ax * 8 + c];
                }
            d += 2.0f * oi;
            C.off[cell * 24 + ax * 8 + c] = oo * GALERKIN_SCALE;
        }
        C.diag[cell * 8 + c] = d * GALERKIN_SCALE;
    }
}

__global__ void k_clear_changed(DevState* st) {
    if (threadIdx.x == 0) st->changed = 0;
}

// ---------------------------------------------------------------- multigrid items
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

// one GS update (k_smooth pressure.py:290-303)
__device__ __forceinline__ void smooth_item(const Lvl& L, int c, int K, int J, int I, bool zero_init) {
    const size_t o = lcell(L.m, K, J, I) * 8 + c;
    const float d = L.diag[o];
    if (d <= L.dmin) { L.x[o] = 0.0f; return; }
    float acc = L.r[o];
    if (!zero_init) acc -= offsum(L, c, K, J, I);
    L.x[o] = acc / d;
}

// residual restriction (k_resrestrict :306-316) + optional zero-guess red sweep at the coarse
// level (k_restrict_red :444-459)
__device__ __forceinline__ void restrict_item(const Lvl& F, const Lvl& C, int c, int K, int J, int I, bool do_red) {
    float acc = 0.0f;
    for (int k = 0; k < 8; ++k) {
        const int z = 2 * K + (k >> 2), y = 2 * J + ((k >> 1) & 1), x = 2 * I + (k & 1);
        const size_t o = lcell(F.m, z, y, x) * 8 + c;
        acc += F.r[o] - F.diag[o] * F.x[o] - offsum(F, c, z, y, x);
    }
    acc = 0.125f * acc;
    const size_t oc = lcell(C.m, K, J, I) * 8 + c;
    C.r[oc] = acc;
    if (do_red && ((I + J + K) & 1) == 0) {
        const float d = C.diag[oc];
        C.x[oc] = d <= C.dmin ? 0.0f : acc / d;
    }
}

// constant prolongation (k_prolong :319-325)
__device__ __forceinline__ void prolong_item(const Lvl& F, const Lvl& C, int c, int K, int J, int I) {
    const size_t o = lcell(F.m, K, J, I) * 8 + c;
    if (F.diag[o] <= F.dmin) return;
    F.x[o] = F.x[o] + C.x[lcell(C.m, K >> 1, J >> 1, I >> 1) * 8 + c];
}

// x value of a fine cell after the constant prolongation, without storing it
__device__ __forceinline__ float xp(const Lvl& F, const Lvl& C, int c, int K, int J, int I) {
    const size_t o = lcell(F.m, K, J, I) * 8 + c;
    const float x = F.x[o];
    if (F.diag[o] <= F.dmin) return x;
    return x + C.x[lcell(C.m, K >> 1, J >> 1, I >> 1) * 8 + c];
}

// prolongation fused into the black post-sweep: equal to prolonging every cell, then sweeping
// black (a GS update ignores the cell's own old value; red cells are rewritten by the red sweep
// that follows).  Same rounding as the unfused pair (pressure.py:396-423 vs :902-931).
__device__ __forceinline__ void prolong_black_item(const Lvl& F, const Lvl& C, int c, int K, int J, int I) {
    const int m = F.m;
    const size_t o = lcell(m, K, J, I);
    const float d = F.diag[o * 8 + c];
    if (d <= F.dmin) { F.x[o * 8 + c] = 0.0f; return; }
    const size_t sy = (size_t)m, sz = (size_t)m * m;
    float a = 0.0f;
    if (I + 1 < m) a += F.off[o * 24 + c] * xp(F, C, c, K, J, I + 1);
    if (I > 0) a += F.off[(o - 1) * 24 + c] * xp(F, C, c, K, J, I - 1);
    if (J + 1 < m) a += F.off[o * 24 + 8 + c] * xp(F, C, c, K, J + 1, I);
    if (J > 0) a += F.off[(o - sy) * 24 + 8 + c] * xp(F, C, c, K, J - 1, I);
    if (K + 1 < m) a += F.off[o * 24 + 16 + c] * xp(F, C, c, K + 1, J, I);
    if (K > 0) a += F.off[(o - sz) * 24 + 16 + c] * xp(F, C, c, K - 1, J, I);
    F.x[o * 8 + c] = (F.r[o * 8 + c] - a) / d;
}

// half-size colour launch: item t -> (c, K, J, I) with (I + J + K) & 1 == color (m even)
__device__ __forceinline__ bool half_item(size_t t, int m, int color, int& c, int& K, int& J, int& I) {
    const int mh = m >> 1;
    const size_t items = (size_t)m * m * mh * 8;
    if (t >= items) return false;
    c = (int)(t & 7);
    const size_t h = t >> 3;
    const int Ih = (int)(h % mh);
    J = (int)((h / mh) % m);
    K = (int)(h / ((size_t)mh * m));
    I = 2 * Ih + ((color + J + K) & 1);
    return true;
}

__device__ __forceinline__ bool cg_inactive(const DevState* st, int k) { return st->stop_at < k; }

// Hadamard forward transform of the Jacobi residual to level 1 (k_fwd_transform :339-349),
// optionally with the zero-guess red sweep (k_fwd_red :426-441).
__global__ void __launch_bounds__(256) k_fwd(int N, const float* __restrict__ rr, Lvl L, int do_red,
                                             const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    const int m = L.m;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)m * m * m * 8) return;
    const int c = (int)(t & 7);
    const size_t cell = t >> 3;
    const int I = (int)(cell % m), J = (int)((cell / m) % m), K = (int)(cell / ((size_t)m * m));
    float acc = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float v = rr[idx3(N, 2 * K + (q >> 2), 2 * J + ((q >> 1) & 1), 2 * I + (q & 1))];
        acc += hsign(c, q >> 2, (q >> 1) & 1, q & 1) * v;
    }
    L.r[cell * 8 + c] = acc;
    if (do_red && ((I + J + K) & 1) == 0) {
        const float d = L.diag[cell * 8 + c];
        L.x[cell * 8 + c] = d <= L.dmin ? 0.0f : acc / d;
    }
}

__global__ void __launch_bounds__(256) k_smooth_half(Lvl L, int color, const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    int c, K, J, I;
    if (!half_item(blockIdx.x * (size_t)blockDim.x + threadIdx.x, L.m, color, c, K, J, I)) return;
    smooth_item(L, c, K, J, I, false);
}

__global__ void __launch_bounds__(256) k_restrict(Lvl F, Lvl C, int do_red, const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    const int m = C.m;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)m * m * m * 8) return;
    const int c = (int)(t & 7);
    const size_t cell = t >> 3;
    restrict_item(F, C, c, (int)(cell / ((size_t)m * m)), (int)((cell / m) % m), (int)(cell % m), do_red != 0);
}

__global__ void __launch_bounds__(256) k_prolong_black(Lvl F, Lvl C, const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    int c, K, J, I;
    if (!half_item(blockIdx.x * (size_t)blockDim.x + threadIdx.x, F.m, 1, c, K, J, I)) return;
    prolong_black_item(F, C, c, K, J, I);
}

// The bottom of the V-cycle in ONE workgroup (k_coarse_block :550-591 generalised to any
// number of levels and to odd coarsest sizes): the V(1,1) of _vcycle_base (:902-931) on
// levels lv[0..n-1], starting with the zero-guess red sweep at lv[0].
constexpr int CB_THREADS = 1024;

__device__ void cb_smooth(const Lvl& L, int color, bool zero_init) {
    const int m = L.m;
    const size_t items = (size_t)m * m * m * 8;
    for (size_t t = threadIdx.x; t < items; t += CB_THREADS) {
        const int c = (int)(t & 7);
        const size_t cell = t >> 3;
        const int I = (int)(cell % m), J = (int)((cell / m) % m), K = (int)(cell / ((size_t)m * m));
        if (((I + J + K) & 1) == color) smooth_item(L, c, K, J, I, zero_init);
    }
    __syncthreads();
}

__global__ void __launch_bounds__(CB_THREADS) k_coarse(CoarseLevels cl, const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    const int n = cl.n;
    for (int j = 0; j + 1 < n; ++j) {
        cb_smooth(cl.lv[j], 0, true);
        cb_smooth(cl.lv[j], 1, false);
        const Lvl& F = cl.lv[j];
        const Lvl& C = cl.lv[j + 1];
        const size_t items = (size_t)C.m * C.m * C.m * 8;
        for (size_t t = threadIdx.x; t < items; t += CB_THREADS) {
            const int c = (int)(t & 7);
            const size_t cell = t >> 3;
            restrict_item(F, C, c, (int)(cell / ((size_t)C.m * C.m)), (int)((cell / C.m) % C.m), (int)(cell % C.m),
                          false);
        }
        __syncthreads();
    }
    const Lvl& B = cl.lv[n - 1];
    if (B.m == 1) {
        if (threadIdx.x < 8) {
            const float d = B.diag[threadIdx.x];
            B.x[threadIdx.x] = d <= B.dmin ? 0.0f : B.r[threadIdx.x] / d;
        }
        __syncthreads();
    } else {
        for (int s = 0; s < 2 * B.m; ++s) {
            cb_smooth(B, 0, s == 0);
            cb_smooth(B, 1, false);
        }
    }
    for (int j = n - 2; j >= 0; --j) {
        const Lvl& F = cl.lv[j];
        const Lvl& C = cl.lv[j + 1];
        const size_t items = (size_t)F.m * F.m * F.m * 8;
        for (size_t t = threadIdx.x; t < items; t += CB_THREADS) {
            const int c = (int)(t & 7);
            const size_t cell = t >> 3;
            prolong_item(F, C, c, (int)(cell / ((size_t)F.m * F.m)), (int)((cell / F.m) % F.m), (int)(cell % F.m));
        }
        __syncthreads();
        cb_smooth(F, 1, false);
        cb_smooth(F, 0, false);
    }
}

// zt = z0 + T^T x1, 0 in solids (k_inv_transform :352-365)
__global__ void __launch_bounds__(256) k_inv(int N, int zc, Lvl L, const uint8_t* __restrict__ fm,
                                             const float* __restrict__ z0, float* __restrict__ zt,
                                             const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= N || y >= N) return;
    const int m = L.m;
    const int z1 = min((int)(blockIdx.z + 1) * zc, N);
    for (int z = blockIdx.z * zc; z < z1; ++z) {
        const size_t i = idx3(N, z, y, x);
        if (fm[i] == 0) { zt[i] = 0.0f; continue; }
        const float* xr = L.x + lcell(m, z >> 1, y >> 1, x >> 1) * 8;
        float acc = z0[i];
#pragma unroll
        for (int c = 0; c < 8; ++c) acc += hsign(c, z & 1, y & 1, x & 1) * xr[c];
        zt[i] = acc;
    }
}

// ---------------------------------------------------------------- CG kernels (pressure.py:595-775)
#define CELL_LOOP_BEGIN                                                 \
    const int x = blockIdx.x * blockDim.x + threadIdx.x;                \
    const int y = blockIdx.y * blockDim.y + threadIdx.y;                \
    const bool live = x < N && y < N;                                   \
    const int z1 = min((int)(blockIdx.z + 1) * zc, N);                  \
    if (live)                                                           \
        for (int cz = blockIdx.z * zc; cz < z1; ++cz) {                 \
            const size_t i = idx3(N, cz, y, x);
#define CELL_LOOP_END }

// k_rhs (:602-627): r = -d(u*) - K phi0 (phi = 0 without warm start); block max |u*| over fluid
// and block sum of b^2 (b = -d(u*)) for the relative-residual test.
__global__ void __launch_bounds__(256) k_rhs(int N, int zc, const float* __restrict__ vel,
                                             const uint8_t* __restrict__ fm, float* __restrict__ r,
                                             float* __restrict__ phi, const float* __restrict__ prm,
                                             double* __restrict__ part, int nparts) {
    __shared__ double sh[8];
    const size_t n3 = (size_t)N * N * N;
    const bool warm = prm[P_WARM] > 0.5f;
    double mx = 0.0, bb = 0.0;
    CELL_LOOP_BEGIN
        if (!warm) phi[i] = 0.0f;
        const int m = fm[i];
        if (m == 0) { r[i] = 0.0f; continue; }
        const float b = -div2(vel, n3, m, N, i);
        float rv = b;
        if (warm) rv -= kmul(phi, fm, N, i);
        r[i] = rv;
        bb += (double)b * (double)b;
        const float ux = vel[i], uy = vel[n3 + i], uz = vel[2 * n3 + i];
        mx = fmax(mx, (double)sqrtf(ux * ux + uy * uy + uz * uz));
    CELL_LOOP_END
    const double a = block_max(mx, sh);
    const double b2 = block_sum(bb, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        part[block_linear()] = a;
        part[nparts + block_linear()] = b2;
    }
}

// k_fin_maxu (:730-735) + the CG state reset.
__global__ void __launch_bounds__(256) k_fin_rhs(const double* __restrict__ part, int nparts,
                                                 const float* __restrict__ prm, DevState* __restrict__ st) {
    __shared__ double sh[4];
    const double mx = reduce_parts_max(part, nparts, sh);
    const double bb = reduce_parts_sum(part + nparts, nparts, sh);
    if (threadIdx.x == 0) {
        const float tol = prm[P_TOL] > 1.0e-6f ? prm[P_TOL] : 1.0e-6f;
        st->s[S_MAXU] = mx;
        st->s[S_THR] = (double)(2.0f * tol * (float)mx);
        st->s[S_BB] = bb;
        st->it = 0;
        st->stop_at = INT_MAX;
    }
}

// k_matvec_dot (:629-646): q = K p, block partial p.q (double)
__global__ void __launch_bounds__(256) k_matvec_dot(int N, int zc, const float* __restrict__ p,
                                                    const uint8_t* __restrict__ fm, float* __restrict__ q,
                                                    double* __restrict__ part, const DevState* __restrict__ st, int k) {
    __shared__ double sh[8];
    if (cg_inactive(st, k)) return;
    double s = 0.0;
    CELL_LOOP_BEGIN
        const float v = fm[i] ? kmul(p, fm, N, i) : 0.0f;
        q[i] = v;
        s += (double)v * (double)p[i];
    CELL_LOOP_END
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = t;
}

// alpha = rz / p.q (the finalize the reference folds into k_pre, :652-659)
__global__ void __launch_bounds__(256) k_fin_pq(const double* __restrict__ part, int nparts,
                                                DevState* __restrict__ st, int k) {
    __shared__ double sh[4];
    if (cg_inactive(st, k)) return;
    const double pq = reduce_parts_sum(part, nparts, sh);
    if (threadIdx.x == 0) st->s[S_ALPHA] = pq > 0.0 ? st->s[S_RZ] / pq : 0.0;
}

// k_pre (:648-672): r -= alpha q (not at k = 0); z0 = omega D^-1 r
__global__ void __launch_bounds__(256) k_pre(int N, int zc, float* __restrict__ r, const float* __restrict__ q,
                                             const uint8_t* __restrict__ fm, float* __restrict__ z0,
                                             const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    const float a = (float)st->s[S_ALPHA];
    CELL_LOOP_BEGIN
        const int m = fm[i];
        float rv = r[i];
        if (k > 0) {
            rv = rv - a * q[i];
            r[i] = rv;
        }
        z0[i] = m ? OMEGA_JAC * rv / kdiag(m) : 0.0f;
    CELL_LOOP_END
}

// k_resid (:674-685): rr = r - K z0
__global__ void __launch_bounds__(256) k_resid(int N, int zc, const float* __restrict__ r,
                                               const float* __restrict__ z0, const uint8_t* __restrict__ fm,
                                               float* __restrict__ rr, const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    CELL_LOOP_BEGIN
        rr[i] = fm[i] ? r[i] - kmul(z0, fm, N, i) : 0.0f;
    CELL_LOOP_END
}

// k_post_rz (:687-710): z = zt + omega D^-1 (r - K zt); block partials r.z, max|r|, r.r
__global__ void __launch_bounds__(256) k_post_rz(int N, int zc, const float* __restrict__ zt,
                                                 const float* __restrict__ r, const uint8_t* __restrict__ fm,
                                                 float* __restrict__ z, double* __restrict__ part, int nparts,
                                                 const DevState* __restrict__ st, int k) {
    __shared__ double sh[8];
    if (cg_inactive(st, k)) return;
    double s = 0.0, mx = 0.0, rr = 0.0;
    CELL_LOOP_BEGIN
        const int m = fm[i];
        const float rv = r[i];
        float zv = 0.0f;
        if (m) zv = zt[i] + OMEGA_JAC * (rv - kmul(zt, fm, N, i)) / kdiag(m);
        z[i] = zv;
        s += (double)rv * (double)zv;
        mx = fmax(mx, fabs((double)rv));
        rr += (double)rv * (double)rv;
    CELL_LOOP_END
    const double a = block_sum(s, sh);
    const double b = block_max(mx, sh);
    const double c = block_sum(rr, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const int bl = block_linear();
        part[bl] = a;
        part[nparts + bl] = b;
        part[2 * nparts + bl] = c;
    }
}

// k_fin_rz (:737-771) with FP64 scalars, plus the optional relative-residual criterion.
__global__ void __launch_bounds__(256) k_fin_rz(const double* __restrict__ part, int nparts,
                                                const float* __restrict__ prm, DevState* __restrict__ st, int k) {
    __shared__ double sh[4];
    if (cg_inactive(st, k)) return;
    const double rz_new = reduce_parts_sum(part, nparts, sh);
    const double mr = reduce_parts_max(part + nparts, nparts, sh);
    const double rr = reduce_parts_sum(part + 2 * nparts, nparts, sh);
    if (threadIdx.x != 0) return;
    double best = mr;
    if (k == 0) {
        st->s[S_BETA] = 0.0;
    } else {
        const double rz = st->s[S_RZ];
        st->s[S_BETA] = rz > 0.0 ? rz_new / rz : 0.0;
        best = fmin(st->s[S_BEST], mr);
    }
    st->it = k;
    st->s[S_RZ] = rz_new;
    st->s[S_MAXR] = mr;
    st->s[S_RR] = rr;
    st->s[S_BEST] = best;
    const double rel_tol = (double)prm[P_REL_TOL];
    const bool not_conv = rel_tol > 0.0 ? sqrt(rr) > rel_tol * sqrt(st->s[S_BB]) : mr > st->s[S_THR];
    const bool go = k < (int)prm[P_MAXIT] && rz_new > 0.0 && mr <= 4.0 * best && not_conv;
    if (!go) st->stop_at = k;
}

// k_upd_p (:712-728): phi += alpha p_old (not at k = 0); p = z + beta p_old
__global__ void __launch_bounds__(256) k_upd_p(int N, int zc, float* __restrict__ phi, float* __restrict__ p,
                                               const float* __restrict__ z, const DevState* __restrict__ st, int k) {
    if (cg_inactive(st, k)) return;
    const float a = (float)st->s[S_ALPHA], b = (float)st->s[S_BETA];
    CELL_LOOP_BEGIN
        if (k == 0) {
            p[i] = z[i];
        } else {
            const float po = p[i];
            phi[i] = phi[i] + a * po;
            p[i] = z[i] + b * po;
        }
    CELL_LOOP_END
}

// k_apply (:166-175): u -= g(phi) on fluid, u = 0 in solids
__global__ void __launch_bounds__(256) k_apply(int N, int zc, float* __restrict__ vel,
                                               const uint8_t* __restrict__ fm, const float* __restrict__ phi) {
    const size_t n3 = (size_t)N * N * N;
    CELL_LOOP_BEGIN
        const int m = fm[i];
        if (m == 0) {
            vel[i] = 0.0f; vel[n3 + i] = 0.0f; vel[2 * n3 + i] = 0.0f;
            continue;
        }
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) vel[ax * n3 + i] = vel[ax * n3 + i] - g_ax(phi, fm, i, axstride(N, ax), ax);
    CELL_LOOP_END
}

// True residual ||b - K phi||^2 with b = -d(u*) (fresh matvec, FP64 accumulation)
__global__ void __launch_bounds__(256) k_true_res(int N, int zc, const float* __restrict__ vel,
                                                  const uint8_t* __restrict__ fm, const float* __restrict__ phi,
                                                  double* __restrict__ part) {
    __shared__ double sh[8];
    const size_t n3 = (size_t)N * N * N;
    double s = 0.0;
    CELL_LOOP_BEGIN
        const int m = fm[i];
        if (m == 0) continue;
        const double v = -(double)div2(vel, n3, m, N, i) - (double)kmul(phi, fm, N, i);
        s += v * v;
    CELL_LOOP_END
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = t;
}

__global__ void __launch_bounds__(256) k_fin_true(const double* __restrict__ part, int nparts,
                                                  DevState* __restrict__ st) {
    __shared__ double sh[4];
    const double s = reduce_parts_sum(part, nparts, sh);
    if (threadIdx.x == 0) st->s[S_TRUE_RR] = s;
}

// |div u| dx with the mask computed from solid (k_div_abs :156-163 + k_mask)
__global__ void __launch_bounds__(256) k_div_abs(int N, int zc, const float* __restrict__ vel,
                                                 const float* __restrict__ solid, float* __restrict__ out) {
    const size_t n3 = (size_t)N * N * N;
    CELL_LOOP_BEGIN
        const int m = mask_from_solid(solid, N, cz, y, x);
        out[i] = m == 0 ? 0.0f : 0.5f * fabsf(div2(vel, n3, m, N, i));
    CELL_LOOP_END
}

// K p at fluid cells (parity tests of the operator)
__global__ void __launch_bounds__(256) k_kmul(int N, int zc, const float* __restrict__ p,
                                              const uint8_t* __restrict__ fm, float* __restrict__ out) {
    CELL_LOOP_BEGIN
        out[i] = fm[i] ? kmul(p, fm, N, i) : 0.0f;
    CELL_LOOP_END
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
        L.dmin = (float)(1.0e-3 * 6.0 * std::pow(0.25, level - 1));
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

size_t Pressure::level_floats() const {
    size_t t = 0;
    for (int l = 0; l < nlev; ++l) {
        const size_t m3 = (size_t)lv[l].m * lv[l].m * lv[l].m;
        t += m3 * (8 + 8 + 8 + 24);
    }
    return t;
}

void Pressure::bind_levels(float* base) {
    for (int l = 0; l < nlev; ++l) {
        const size_t m3 = (size_t)lv[l].m * lv[l].m * lv[l].m;
        lv[l].x = base; base += 8 * m3;
        lv[l].r = base; base += 8 * m3;
        lv[l].diag = base; base += 8 * m3;
        lv[l].off = base; base += 24 * m3;
    }
}

void Pressure::setup(const float* solid, hipStream_t s) {
    const size_t n3 = (size_t)N * N * N;
    hipLaunchKernelGGL(k_solid_check, dim3(grid_for(n3, 4096)), dim3(256), 0, s, n3, solid, fm, st);
    hipLaunchKernelGGL(k_mask, dim3(grid_for(n3, 4096)), dim3(256), 0, s, N, solid, fm, st);
    const size_t m3 = (size_t)lv[0].m * lv[0].m * lv[0].m;
    hipLaunchKernelGGL(k_probe, dim3(grid_for(m3 * 8, 4096)), dim3(256), 0, s, N, fm, lv[0], st);
    for (int l = 0; l + 1 < nlev; ++l) {
        const size_t mc3 = (size_t)lv[l + 1].m * lv[l + 1].m * lv[l + 1].m;
        hipLaunchKernelGGL(k_galerkin, dim3(grid_for(mc3 * 8, 4096)), dim3(256), 0, s, lv[l], lv[l + 1], st);
    }
    hipLaunchKernelGGL(k_clear_changed, dim3(1), dim3(64), 0, s, st);
}

void Pressure::precond(int k, hipStream_t s) {
    const Geom& g = geom;
    hipLaunchKernelGGL(k_pre, g.grid, g.block, 0, s, N, g.zc, r, q, fm, z0, st, k);
    hipLaunchKernelGGL(k_resid, g.grid, g.block, 0, s, N, g.zc, r, z0, fm, rr, st, k);
    {
        const size_t items = (size_t)lv[0].m * lv[0].m * lv[0].m * 8;
        hipLaunchKernelGGL(k_fwd, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, s, N, rr, lv[0], S > 0 ? 1 : 0,
                           st, k);
    }
    for (int l = 0; l < S; ++l) {
        const size_t half = (size_t)lv[l].m * lv[l].m * (lv[l].m / 2) * 8;
        hipLaunchKernelGGL(k_smooth_half, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, s, lv[l], 1, st, k);
        const size_t items = (size_t)lv[l + 1].m * lv[l + 1].m * lv[l + 1].m * 8;
        hipLaunchKernelGGL(k_restrict, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, s, lv[l], lv[l + 1],
                           (l + 1 < S) ? 1 : 0, st, k);
    }
    CoarseLevels cl{};
    cl.n = nlev - S;
    for (int j = 0; j < cl.n; ++j) cl.lv[j] = lv[S + j];
    hipLaunchKernelGGL(k_coarse, dim3(1), dim3(CB_THREADS), 0, s, cl, st, k);
    for (int l = S - 1; l >= 0; --l) {
        const size_t half = (size_t)lv[l].m * lv[l].m * (lv[l].m / 2) * 8;
        hipLaunchKernelGGL(k_prolong_black, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, s, lv[l], lv[l + 1],
                           st, k);
        hipLaunchKernelGGL(k_smooth_half, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, s, lv[l], 0, st, k);
    }
    hipLaunchKernelGGL(k_inv, g.grid, g.block, 0, s, N, g.zc, lv[0], fm, z0, zt, st, k);
    hipLaunchKernelGGL(k_post_rz, g.grid, g.block, 0, s, N, g.zc, zt, r, fm, z, part, nparts, st, k);
}

void Pressure::rhs(const float* vel, const float* prm, hipStream_t s) {
    const Geom& g = geom;
    hipLaunchKernelGGL(k_rhs, g.grid, g.block, 0, s, N, g.zc, vel, fm, r, phi, prm, part, nparts);
    hipLaunchKernelGGL(k_fin_rhs, dim3(1), dim3(256), 0, s, part, nparts, prm, st);
}

void Pressure::iteration(int k, const float* prm, hipStream_t s) {
    const Geom& g = geom;
    if (k > 0) {
        hipLaunchKernelGGL(k_matvec_dot, g.grid, g.block, 0, s, N, g.zc, p, fm, q, part_pq, st, k);
        hipLaunchKernelGGL(k_fin_pq, dim3(1), dim3(256), 0, s, part_pq, nparts, st, k);
    }
    precond(k, s);
    hipLaunchKernelGGL(k_fin_rz, dim3(1), dim3(256), 0, s, part, nparts, prm, st, k);
    hipLaunchKernelGGL(k_upd_p, g.grid, g.block, 0, s, N, g.zc, phi, p, z, st, k);
}

void Pressure::apply(float* vel, hipStream_t s) {
    const Geom& g = geom;
    hipLaunchKernelGGL(k_apply, g.grid, g.block, 0, s, N, g.zc, vel, fm, phi);
}

void Pressure::true_residual(const float* vel, hipStream_t s) {
    const Geom& g = geom;
    hipLaunchKernelGGL(k_true_res, g.grid, g.block, 0, s, N, g.zc, vel, fm, phi, part);
    hipLaunchKernelGGL(k_fin_true, dim3(1), dim3(256), 0, s, part, nparts, st);
}

// One kernel of the projection on the current buffers (iteration index 0 semantics, so no CG
// state is advanced), for timing.  Returns false for an unknown id.
bool Pressure::launch_single(int which, hipStream_t s) {
    const Geom& g = geom;
    const size_t m0 = (size_t)lv[0].m;
    const size_t half0 = m0 * m0 * (m0 / 2) * 8;
    switch (which) {
        case 0: hipLaunchKernelGGL(k_matvec_dot, g.grid, g.block, 0, s, N, g.zc, p, fm, q, part_pq, st, 1); break;
        case 1: hipLaunchKernelGGL(k_post_rz, g.grid, g.block, 0, s, N, g.zc, zt, r, fm, z, part, nparts, st, 0); break;
        case 2:
            if (S < 1) return false;
            hipLaunchKernelGGL(k_smooth_half, dim3((unsigned)((half0 + 255) / 256)), dim3(256), 0, s, lv[0], 1, st, 0);
            break;
        case 3: {
            if (S < 1) return false;
            const size_t items = (size_t)lv[1].m * lv[1].m * lv[1].m * 8;
            hipLaunchKernelGGL(k_restrict, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, s, lv[0], lv[1],
                               S > 1 ? 1 : 0, st, 0);
            break;
        }
        case 4:
            if (S < 1) return false;
            hipLaunchKernelGGL(k_prolong_black, dim3((unsigned)((half0 + 255) / 256)), dim3(256), 0, s, lv[0], lv[1],
                               st, 0);
            break;
        case 5: hipLaunchKernelGGL(k_pre, g.grid, g.block, 0, s, N, g.zc, r, q, fm, z0, st, 0); break;
        case 6: hipLaunchKernelGGL(k_resid, g.grid, g.block, 0, s, N, g.zc, r, z0, fm, rr, st, 0); break;
        case 7:
            hipLaunchKernelGGL(k_fwd, dim3((unsigned)((m0 * m0 * m0 * 8 + 255) / 256)), dim3(256), 0, s, N, rr, lv[0],
                               S > 0 ? 1 : 0, st, 0);
            break;
        case 8: hipLaunchKernelGGL(k_inv, g.grid, g.block, 0, s, N, g.zc, lv[0], fm, z0, zt, st, 0); break;
        case 9: hipLaunchKernelGGL(k_upd_p, g.grid, g.block, 0, s, N, g.zc, phi, p, z, st, 0); break;
        case 10: {
            CoarseLevels cl{};
            cl.n = nlev - S;
            for (int j = 0; j < cl.n; ++j) cl.lv[j] = lv[S + j];
            hipLaunchKernelGGL(k_coarse, dim3(1), dim3(CB_THREADS), 0, s, cl, st, 0);
            break;
        }
        default: return false;
    }
    return true;
}

void launch_div_abs(const Geom& g, const float* vel, const float* solid, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_div_abs, g.grid, g.block, 0, s, g.N, g.zc, vel, solid, out);
}

void launch_kmul(const Geom& g, const float* p, const uint8_t* fm, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_kmul, g.grid, g.block, 0, s, g.N, g.zc, p, fm, out);
}

}  // namespace vfx
