// Shared definitions of the MI355X (gfx950) sim3d library: parameter-buffer slots, device
// scalar slots, launch geometry, the collocated-grid helpers and block reductions.
//
// Layout of every field: (C, D, H, W) `real`, index [c][z][y][x], x fastest (the reference's
// FieldState3D, vfx/core/fields3d.py:1-10).  Masks are uint8 open-face bit sets.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

// Field precision of this build.  The sources are compiled twice: VFX_F64 = 0 is the product
// path (FP32 fields and levels, FP64 reductions); VFX_F64 = 1 is an FP64 run of the same solver
// (every field, level and constant in double) whose C entry points carry the suffix _f64 and
// whose namespace is vfx_f64, so both builds link into one library.
#ifndef VFX_F64
#define VFX_F64 0
#endif
#if VFX_F64
#define vfx vfx_f64
#define VFX_FN(name) name##_f64
#define RL(x) (x)
#else
#define VFX_FN(name) name
#define RL(x) (x##f)
#endif

namespace vfx {

#if VFX_F64
using real = double;
using real2 = double2;
using real4 = double4;
__host__ __device__ inline real2 make_real2(real a, real b) { return make_double2(a, b); }
__host__ __device__ inline real4 make_real4(real a, real b, real c, real d) { return make_double4(a, b, c, d); }
__device__ __forceinline__ bool same_bits(real a, real b) { return __double_as_longlong(a) == __double_as_longlong(b); }
#else
using real = float;
using real2 = float2;
using real4 = float4;
__host__ __device__ inline real2 make_real2(real a, real b) { return make_float2(a, b); }
__host__ __device__ inline real4 make_real4(real a, real b, real c, real d) { return make_float4(a, b, c, d); }
__device__ __forceinline__ bool same_bits(real a, real b) { return __float_as_uint(a) == __float_as_uint(b); }
#endif

// ---------------------------------------------------------------- parameter buffer (floats)
// Slot numbers of the reference (vfx/sim3d/kernels.py:11-42) plus extension slots in the
// ranges the reference leaves unused.
enum : int {
    P_DT = 0, P_WIND_X = 1, P_WIND_Y = 2, P_WIND_Z = 3, P_N_EM = 4, P_N_IMP = 5, P_DX = 6,
    P_AMB = 7, P_KAPPA = 8, P_CFLOOR = 9, P_WARM = 10,
    P_IGN_RAMP = 11, P_NOISE_LO = 12, P_NOISE_HI = 13, P_SEED = 14, P_STEPNO = 15,
    P_ALPHA = 16, P_BETA = 17, P_EPS = 18, P_DDISS = 19, P_COOL = 20, P_BURN = 21, P_IGN = 22,
    P_BHEAT = 23, P_SMOKE = 24, P_TOL = 25, P_MAXIT = 26, P_TMAX = 27, P_WMARK = 28,
    P_REL_TOL = 29,
    P_EM0 = 32, EM_STRIDE = 10, MAX_EMITTERS = 32,
    P_IMP0 = P_EM0 + MAX_EMITTERS * EM_STRIDE, IMP_STRIDE = 7, MAX_IMPULSES = 32,
    P_SIZE = P_IMP0 + MAX_IMPULSES * IMP_STRIDE,
};

// ---------------------------------------------------------------- device scalar state
// doubles written by the finalize kernels (the reference keeps the same quantities as real
// slots, kernels.py:44-53); the CG inner products, alpha and beta are FP64 here.
enum : int {
    S_MASS_RATIO = 0, S_MAXU = 1, S_RZ = 2, S_ALPHA = 3, S_BETA = 4, S_MAXR = 5, S_THR = 6,
    S_BEST = 7, S_BB = 8, S_RR = 9, S_TRUE_RR = 10, S_LSCALE = 11,
    S_MASS_RATIO_F = 12,   // mass ratio of the finer scalar grid (hires.py, solver.py:250-253)
    S_FCG_H0 = 13,         // 13..15: ||r|| of the last three learned flexible-CG iterations (slot k % 3)
    S_COUNT = 16,
};

struct DevState {
    double s[S_COUNT];
    int it;        // CG iterations of the current / last CG run
    int stop_at;   // iteration at which the stop test fired (INT_MAX while running; -1 after a
                   // restart gate that found the solve converged: every gated kernel is a no-op)
    int changed;   // solid differs from the cached face mask -> rebuild operators
    int conv;      // the last stop test found the residual at tolerance
    int it_total;  // iterations of the earlier CG runs of this projection (device-side restarts)
    int restarts;  // device-side restarts of this projection
    int live;      // 1 while the projection's current CG run is live (gates the restart kernels)
    int hit;       // bit 0: the first CG run stopped at its iteration budget unconverged; bit 1: a restart did
    int early;     // 1 when the stop test of iteration stop_at fired on r before that iteration's
                   // V-cycle, whose kernels (and the rest of the iteration) are then no-ops
    int fallback;  // the learned flexible-CG run of this projection stalled or broke down
    unsigned done;   // workgroups of the current finalising launch that have written their partials
                     // (0 between launches)
};

// the kernels of CG iteration k do nothing after the run stopped at an earlier iteration, or at k
// itself on r before the V-cycle
__device__ __forceinline__ bool cg_inactive(const DevState* st, int k) {
    return st->stop_at < k || (st->stop_at == k && st->early != 0);
}
// kernels that also run outside a CG iteration (the network): gated only when st is given
__device__ __forceinline__ bool gated_off(const DevState* st, int k) { return st != nullptr && cg_inactive(st, k); }

constexpr int FLUID = 64;          // face-mask fluid bit (pressure.py:32-33)
constexpr real LAM_CAP = (real)0.16f;   // kernels.py:63 (the reference's float constant)
constexpr real HS = (real)0.35355339059327373;   // 1/sqrt(8), pressure.py:30
constexpr real OMEGA_JAC = (real)0.8f;  // pressure.py:785 (the reference's float constant)
constexpr real GALERKIN_SCALE = (real)0.0625;   // 1/16, pressure.py:783

// ---------------------------------------------------------------- launch geometry
// Cell kernels use 256-thread blocks of BX x (256/BX) threads over (x, y) and walk ZC z-planes
// per thread, so a wavefront reads whole x-rows (coalesced) and a block reuses y-neighbours.
struct Geom {
    int N;
    int bx, by, zc;
    dim3 grid, block;
    size_t n3;
};

Geom make_geom(int N, int zc = 1);
inline int nblocks(const Geom& g) { return (int)(g.grid.x * g.grid.y * g.grid.z); }

__device__ __forceinline__ size_t idx3(int N, int z, int y, int x) {
    return ((size_t)z * N + y) * N + x;
}

// Neighbour loads go to in-bounds addresses (clamped to the cell itself outside the box) and are
// issued unconditionally; the solid / range tests only select values.  A load guarded by a test
// would wait for the test's own load and serialise the memory traffic of every cell.
struct Nb6 {
    int xm, xp, ym, yp, zm, zp;   // clamped neighbour indices
    bool oxm, oxp, oym, oyp, ozm, ozp;   // neighbour inside the box
};

__device__ __forceinline__ Nb6 nb6(int N, int x, int y, int z, int i) {
    Nb6 n;
    const int sz = N * N;
    n.oxm = x > 0; n.oxp = x + 1 < N; n.oym = y > 0; n.oyp = y + 1 < N; n.ozm = z > 0; n.ozp = z + 1 < N;
    n.xm = n.oxm ? i - 1 : i; n.xp = n.oxp ? i + 1 : i;
    n.ym = n.oym ? i - N : i; n.yp = n.oyp ? i + N : i;
    n.zm = n.ozm ? i - sz : i; n.zp = n.ozp ? i + sz : i;
    return n;
}

// Workgroups are dispatched to the 8 XCDs round robin (workgroup b runs on XCD b % 8), and each
// XCD has its own L2.  xcd_block renumbers them so that every XCD works through one contiguous
// range of logical blocks: neighbouring blocks, which share halo and neighbour cache lines, then
// run on the same XCD at about the same time.  A bijection of [0, nb) for any nb.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb / NUM_XCD, r = nb % NUM_XCD, x = b % NUM_XCD, k = b / NUM_XCD;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

__device__ __forceinline__ int block_linear() {
    return (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}

// ---------------------------------------------------------------- reductions
// Fixed-order block reductions: wavefront (64 lanes) butterfly via DPP/shuffles, then LDS across
// the waves of the block.  Deterministic for a fixed launch shape.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Sum over the block; result valid in thread 0.  `sh` needs blockDim/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nw = (blockDim.x * blockDim.y * blockDim.z) >> 6;
    v = wave_sum(v);
    __syncthreads();
    if ((t & 63) == 0) sh[t >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (t == 0)
        for (int w = 0; w < nw; ++w) r += sh[w];
    return r;
}
__device__ __forceinline__ double block_max(double v, double* sh) {
    const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nw = (blockDim.x * blockDim.y * blockDim.z) >> 6;
    v = wave_max(v);
    __syncthreads();
    if ((t & 63) == 0) sh[t >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (t == 0)
        for (int w = 0; w < nw; ++w) r = fmax(r, sh[w]);
    return r;
}

// NV block reductions in one pass: value j is a sum (bit j of MAXMASK clear) or a max (set), each
// with block_sum's / block_max's butterfly and wave order, so the results are theirs bit for bit,
// behind two barriers instead of two per value.  Results valid in thread 0; `sh` needs NV
// doubles per wave.
template <int NV, unsigned MAXMASK>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double* sh) {
    const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nw = (blockDim.x * blockDim.y * blockDim.z) >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const double u = __shfl_xor(v[j], o, 64);
            v[j] = ((MAXMASK >> j) & 1u) ? fmax(v[j], u) : v[j] + u;
        }
    __syncthreads();
    if ((t & 63) == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j) sh[j * nw + (t >> 6)] = v[j];
    __syncthreads();
    if (t == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            double r = 0.0;
            for (int w = 0; w < nw; ++w) r = ((MAXMASK >> j) & 1u) ? fmax(r, sh[j * nw + w]) : r + sh[j * nw + w];
            v[j] = r;
        }
}

// Final reductions of block partials run in one workgroup of FIN_THREADS threads; the strided
// loop is unrolled so that a thread's loads are in flight together.  Result in thread 0, fixed
// order for a fixed n.
constexpr int FIN_THREADS = 1024;

// Block partials of NA arrays (array a at p + off[a]), n each: thread t uses i = t, t + blockDim,
// t + 2 blockDim, ... in that order, and issues the reads of CH of them before it uses any (each
// read is a round trip to the L2 or HBM; a single workgroup reducing 65536 partials one round of
// 8 at a time took 20-30 us).  Indices past n are read at n - 1 and not used.
template <int NA, int CH, class F>
__device__ __forceinline__ void for_parts(const double* __restrict__ p, int n, const int (&off)[NA], F&& use) {
    const int B = (int)blockDim.x;
    for (int i0 = (int)threadIdx.x; i0 < n; i0 += CH * B) {
        double v[CH][NA];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const int i = min(i0 + j * B, n - 1);
#pragma unroll
            for (int a = 0; a < NA; ++a) v[j][a] = p[off[a] + i];
        }
        __builtin_amdgcn_sched_barrier(0);   // every read above is issued before the first use
#pragma unroll
        for (int j = 0; j < CH; ++j)
            if (i0 + j * B < n) use(v[j]);
    }
}

__device__ __forceinline__ double reduce_parts_sum(const double* p, int n, double* sh) {
    double v = 0.0;
    for_parts<1, 32>(p, n, {0}, [&](const double (&x)[1]) { v += x[0]; });
    return block_sum(v, sh);
}
__device__ __forceinline__ double reduce_parts_max(const double* p, int n, double* sh) {
    double v = 0.0;
    for_parts<1, 32>(p, n, {0}, [&](const double (&x)[1]) { v = fmax(v, x[0]); });
    return block_max(v, sh);
}

// ---------------------------------------------------------------- error plumbing
void set_error(const std::string& msg);
bool hip_ok(hipError_t e, const char* what);
bool launch_ok(const char* what);

}  // namespace vfx
