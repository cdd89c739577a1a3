// Host-side launchers and the per-resolution solver objects of the sim3d library.
#pragma once

#include <cmath>

#include "sim3d_common.hpp"

namespace vfx {

// One multigrid level: m^3 blocks x 8 Hadamard components (component-fastest storage).
struct Lvl {
    int m;
    real dmin;   // diag <= dmin -> x = 0 (pressure.py:803)
    real* x;     // [m^3][8]
    real* r;     // [m^3][8]
    real* diag;  // [m^3][8]
    real* off;   // [m^3][24]: ax*8 + c couples to the +ax neighbour
    real* tab;   // [32]: diag[8], off[24] of a cell in an all-fluid neighbourhood
    uint8_t* sf;  // [m^3]: 1 where a cell's whole stencil equals tab (see k_std_flags)
};

constexpr int MAX_LEVELS = 16;

struct CoarseLevels {
    int n;
    Lvl lv[MAX_LEVELS];
};

// ---------------------------------------------------------------- advection / forces
void launch_advect(const Geom& g, int C, const real* src, const real* vel, const float* prm, real inv_dx,
                   real* hat, real* out, const uint8_t* fm, double* part, hipStream_t s);
void launch_adv_pass(int pass, const Geom& g, int C, const real* src, const real* vel, const float* prm,
                     real inv_dx, real* hat, real* out, const uint8_t* fm, double* part, hipStream_t s);
void launch_sources_pass(int pass, const Geom& g, const real* src, const uint8_t* fm, const float* prm,
                         const DevState* st, real* dst, hipStream_t s);
void launch_mass_fin(const Geom& g, double* part, DevState* st, hipStream_t s);
size_t mass_part_doubles(const Geom& g);   // partial-sum slots of the advection mass sums
void launch_sources_diffuse(const Geom& g, const real* adv, const uint8_t* fm, const float* prm,
                            const DevState* st, real* hat, real* data, hipStream_t s);
void launch_vort_forces(const Geom& g, real* data, const uint8_t* fm, const float* prm, real inv2dx,
                        real* omega, real* omega_s, real* force, real* tmp_a, real* tmp_b, hipStream_t s);

// ---------------------------------------------------------------- projection
void launch_div_abs(const Geom& g, const real* vel, const real* solid, real* out, hipStream_t s);
void launch_kmul(const Geom& g, const real* p, const uint8_t* fm, real* out, hipStream_t s);

// Buffers and launch sequence of the MG-preconditioned CG projection for one resolution
// (PressureSolver3D, pressure.py:779-956).
struct Pressure {
    int N = 0;
    int nlev = 0;
    int S = 0;            // first level handled by the single-workgroup coarse kernel
    int coarse_m = 8;     // levels with m <= coarse_m run inside that kernel
    bool use_std = true;  // standard-cell flags on (off: every cell reads the level arrays)
    bool force_setup = false;   // rebuild the operators at the next setup
    Lvl lv[MAX_LEVELS];
    Geom geom;
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
    void rhs(const real* vel, const float* prm, hipStream_t s);
    void precond(int k, hipStream_t s);
    void iteration(int k, const float* prm, hipStream_t s);
    void finish(hipStream_t s);
    void apply(real* vel, hipStream_t s);
    void true_residual(const real* vel, hipStream_t s);
    void vcycle_levels(int k, hipStream_t s);
    bool launch_single(int which, real* scratch, hipStream_t s);
    bool launch_single_mg(int which, hipStream_t s);
};

}  // namespace vfx
