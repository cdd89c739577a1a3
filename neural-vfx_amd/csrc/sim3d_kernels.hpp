// Host-side launchers and the per-resolution solver objects of the sim3d library.
#pragma once

#include <cmath>

#include "sim3d_common.hpp"

namespace vfx {

// One multigrid level: m^3 blocks x 8 Hadamard components (component-fastest storage).  LT is
// the storage type of the level arrays: `real`, or a 16-bit type (FP16 / BF16) whose values are
// converted to `real` on load and rounded to nearest on store; all arithmetic is in `real`.
template <class LT>
struct LvlT {
    int m;
    real dmin;   // diag <= dmin -> x = 0 (pressure.py:803)
    LT* x;       // [m^3][8]
    LT* r;       // [m^3][8]
    LT* diag;    // [m^3][8]
    LT* off;     // [m^3][24]: ax*8 + c couples to the +ax neighbour
    real* tab;   // [32]: diag[8], off[24] of a cell in an all-fluid neighbourhood (as stored)
    uint8_t* sf;  // [m^3]: 1 where a cell's whole stencil equals tab (see k_std_flags)
};
using Lvl = LvlT<real>;

// the same level arrays viewed with storage type LT (each array has room for m^3 x n reals)
template <class LT>
inline LvlT<LT> lvl_as(const Lvl& L) {
    return LvlT<LT>{L.m, L.dmin, reinterpret_cast<LT*>(L.x), reinterpret_cast<LT*>(L.r),
                    reinterpret_cast<LT*>(L.diag), reinterpret_cast<LT*>(L.off), L.tab, L.sf};
}

constexpr int MAX_LEVELS = 16;

template <class LT>
struct CoarseLevelsT {
    int n;
    LvlT<LT> lv[MAX_LEVELS];
};
using CoarseLevels = CoarseLevelsT<real>;

// storage kinds of the multigrid levels (VFX_OPT_MG_STORAGE)
enum : int { MG_REAL = 0, MG_FP16 = 1, MG_BF16 = 2 };

// element load / store of level arrays: convert to `real`, round to nearest on store
__device__ __forceinline__ real ldl(const real* p) { return *p; }
__device__ __forceinline__ void stl(real* p, real v) { *p = v; }
#if !VFX_F64
__device__ __forceinline__ real ldl(const __half* p) { return __half2float(*p); }
__device__ __forceinline__ void stl(__half* p, real v) { *p = __float2half_rn(v); }
__device__ __forceinline__ real ldl(const __hip_bfloat16* p) { return __bfloat162float(*p); }
__device__ __forceinline__ void stl(__hip_bfloat16* p, real v) { *p = __float2bfloat16(v); }
#endif

// v rounded to the storage type LT
template <class LT>
__device__ __forceinline__ real as_stored(real v) {
    LT t;
    stl(&t, v);
    return ldl(&t);
}

// the 8 components of one level cell (p = array + cell * 8): two real4 or one 16-byte load
template <class LT>
__device__ __forceinline__ void ld8(const LT* p, real* v) {
    if constexpr (sizeof(LT) == sizeof(real)) {
        const real4 a = reinterpret_cast<const real4*>(p)[0], b = reinterpret_cast<const real4*>(p)[1];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
        union { uint4 u; LT h[8]; } w;
        w.u = *reinterpret_cast<const uint4*>(p);
#pragma unroll
        for (int c = 0; c < 8; ++c) v[c] = ldl(&w.h[c]);
    }
}
template <class LT>
__device__ __forceinline__ void st8(LT* p, const real* v) {
    if constexpr (sizeof(LT) == sizeof(real)) {
        reinterpret_cast<real4*>(p)[0] = make_real4(v[0], v[1], v[2], v[3]);
        reinterpret_cast<real4*>(p)[1] = make_real4(v[4], v[5], v[6], v[7]);
    } else {
        union { uint4 u; LT h[8]; } w;
#pragma unroll
        for (int c = 0; c < 8; ++c) stl(&w.h[c], v[c]);
        *reinterpret_cast<uint4*>(p) = w.u;
    }
}

// call f(LT{}) with the level storage type of `kind` (the FP64 build stores levels in double)
template <class Fn>
void with_mg_type(int kind, Fn&& f) {
#if VFX_F64
    (void)kind;
    f(real{});
#else
    switch (kind) {
        case MG_FP16: f(__half{}); break;
        case MG_BF16: f(__hip_bfloat16{}); break;
        default: f(real{}); break;
    }
#endif
}

// scale of the level-1 right-hand side of a V-cycle with 16-bit storage: a power of two that
// brings the current residual's magnitude near 1 (exact in `real`; FP16 keeps no subnormal
// residuals), 1 for `real` storage
template <class LT>
__device__ __forceinline__ real level_scale(const DevState* st) {
    if constexpr (sizeof(LT) == sizeof(real)) {
        return RL(1.0);
    } else {
        const double v = st->s[S_LSCALE];
        return v > 0.0 ? (real)v : RL(1.0);
    }
}

// ---------------------------------------------------------------- advection / forces
// qidx: N^3 ints of scratch (pass 1 writes the backward trace's base corner, pass 2 reads it)
void launch_advect(const Geom& g, int C, const real* src, const real* vel, const float* prm, real inv_dx,
                   real* hat, int* qidx, real* out, const uint8_t* fm, double* part, hipStream_t s);
void launch_adv_pass(int pass, const Geom& g, int C, const real* src, const real* vel, const float* prm,
                     real inv_dx, real* hat, int* qidx, real* out, const uint8_t* fm, double* part, hipStream_t s);
void launch_sources_pass(int pass, const Geom& g, const real* src, const uint8_t* fm, const float* prm,
                         const DevState* st, real* dst, hipStream_t s);
// ratio -> st->s[slot] (S_MASS_RATIO of the simulation grid, S_MASS_RATIO_F of the fine scalar grid)
void launch_mass_fin(const Geom& g, double* part, DevState* st, hipStream_t s, int slot = S_MASS_RATIO);
size_t mass_part_doubles(const Geom& g);
int advect_parts(const Geom& g);           // blocks of the advection launch (one mass partial each)   // partial-sum slots of the advection mass sums
void launch_sources_diffuse(const Geom& g, const real* adv, const uint8_t* fm, const float* prm,
                            const DevState* st, real* hat, real* data, hipStream_t s);
// sources of the velocity channels only, adv -> data (with the finer scalar grid, whose restriction
// supplies density, temperature and fuel; no diffusion, which leaves the velocity as it is)
void launch_sources_vel(const Geom& g, const real* adv, const uint8_t* fm, const float* prm, real* data,
                        hipStream_t s);
void launch_vort_forces(const Geom& g, real* data, const uint8_t* fm, const float* prm, real inv2dx, real* omega_s,
                        hipStream_t s);

// ---------------------------------------------------------------- projection
void launch_div_abs(const Geom& g, const real* vel, const real* solid, real* out, hipStream_t s);
void launch_kmul(const Geom& g, const real* p, const uint8_t* fm, real* out, hipStream_t s);

// Buffers and launch sequence of the MG-preconditioned CG projection for one resolution
// (PressureSolver3D, pressure.py:779-956).
struct Pressure {
    int N = 0;
    int nlev = 0;
    int S = 0;            // first level handled by the single-workgroup coarse kernel
    int coarse_m = 4;     // levels with m <= coarse_m run inside that kernel
    bool use_std = true;  // standard-cell flags on (off: every cell reads the level arrays)
    int mg_kind = MG_REAL;   // storage of the multigrid levels (MG_REAL / MG_FP16 / MG_BF16)
    bool force_setup = false;   // rebuild the operators at the next setup
    Lvl lv[MAX_LEVELS];
    Geom geom;
    Geom geom_r;   // the restart run's cell kernels: about 2048 blocks, so a gated launch that does not run is cheap
    uint8_t* fm = nullptr;       // face mask (n3 bytes) followed by at least one zero byte
    real* R[2] = {nullptr, nullptr};   // residual, ping-pong by iteration parity
    real* P[2] = {nullptr, nullptr};   // search direction, ping-pong
    real *q = nullptr, *z = nullptr, *phi = nullptr;
    double* part = nullptr;      // 3 * nparts
    double* part_pq = nullptr;   // nparts
    int nparts = 0;              // partial-sum slots: max(cell-kernel blocks, tiles)
    dim3 tgrid;                  // LDS tiles of the fine grid
    int ntiles = 0;
    DevState* st = nullptr;

    void init_levels(int N);
    void init_tiles();
    real* r_buf(int k) const;
    size_t level_floats() const;
    void bind_levels(real* base);
    void setup(const real* solid, hipStream_t s);
    // warm: the host's copy of the P_WARM parameter (the kernel variant is chosen on the host);
    // gated: the rhs of a device-side restart (runs only when the restart gate opened);
    // restart: the rhs of a restart run (geom_r)
    void rhs(const real* vel, const float* prm, bool warm, hipStream_t s, bool gated = false, bool restart = false);
    // prm: with the parameters, iteration k's stop test also runs on r before the V-cycle (fin_r)
    void precond(int k, hipStream_t s, const float* prm = nullptr);
    // r update, z0 = omega D^-1 r, residual, level-1 transform; fin 1: the stop test on r, 2: the
    // budget's last iteration (its stop, bit hitbit of DevState::hit)
    void precond_front(int k, hipStream_t s, const float* prm = nullptr, int fin = 0, int hitbit = 0);
    // the multigrid levels, the inverse transform and post-smoothing; with prm, also iteration k's
    // CG scalars and stop test (fin_rz, budget kmax)
    void precond_back(int k, hipStream_t s, const float* prm = nullptr, int kmax = 0, int hitbit = 0);
    // CG iteration k of a run whose host enqueued iterations 0..kmax (hitbit: DevState::hit bit
    // set when that budget, not the stop test, ends the run)
    void iteration(int k, const float* prm, hipStream_t s, int kmax = 1 << 30, int hitbit = 1);
    void finish(hipStream_t s, bool restart = false);
    void apply(real* vel, hipStream_t s);
    void true_residual(const real* vel, hipStream_t s, bool gated = false, bool restart = false);
    // decide on the device whether the gated restart run after it runs (true residual above the
    // relative tolerance)
    void restart_gate(const float* prm, hipStream_t s, bool learned = false);
    void vcycle_levels(int k, hipStream_t s);
    bool launch_single(int which, real* scratch, hipStream_t s);
    bool launch_single_mg(int which, hipStream_t s);
};

}  // namespace vfx
