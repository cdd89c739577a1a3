// The per-cell sources update (emitters, impulses, combustion, cooling, dissipation) shared by the
// simulation grid (k_sources, forces.hip) and the finer scalar grid (k_fsources, hires.hip).
//
// Restates vfx/sim3d/kernels.py:182-250 (sources) and hires.py:92-139 (sources_fine, the scalar
// channels at fine-cell centres), with the ignition ramp width and per-cell emitter noise as
// parameters (reference behaviour: ramp 100 K, noise off).
#pragma once

#include "sim3d_common.hpp"

namespace vfx {

__device__ __forceinline__ real clampf_(real g, real lo, real hi) {
    return g < lo ? lo : (g > hi ? hi : g);
}

// Philox-4x32-10 (Salmon et al., SC'11), first output word; key (seed, emitter), counter
// (cell, step, 'EMIT', 0).
__device__ __forceinline__ uint32_t philox_u32(uint32_t seed, uint32_t em, uint32_t cell, uint32_t step) {
    uint32_t c0 = cell, c1 = step, c2 = 0x454D4954u, c3 = 0u, k0 = seed, k1 = em;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    return c0;
}

// step constants of the sources, read once per thread
struct SrcConst {
    real amb, dt, ramp, nlo, nhi, tmax, e_burn, e_cool, e_diss;
    bool noisy;
    uint32_t seed, stepno;
    int n_em, n_imp;

    __device__ __forceinline__ explicit SrcConst(const float* __restrict__ prm) {
        amb = (real)prm[P_AMB];
        dt = (real)prm[P_DT];
        n_em = (int)prm[P_N_EM];
        n_imp = (int)prm[P_N_IMP];
        ramp = (real)prm[P_IGN_RAMP];
        nlo = (real)prm[P_NOISE_LO];
        nhi = (real)prm[P_NOISE_HI];
        noisy = !(nlo == RL(1.0) && nhi == RL(1.0));
        seed = (uint32_t)prm[P_SEED];
        stepno = (uint32_t)prm[P_STEPNO];
        tmax = (real)prm[P_TMAX];
        e_burn = exp(-(real)prm[P_BURN] * dt);
        e_cool = exp(-(real)prm[P_COOL] * dt);
        e_diss = exp(-(real)prm[P_DDISS] * dt);
    }
};

// Sources at the cell centre (px, py, pz) (world units).  SCAL updates density d (already times
// the mass ratio), temperature T and fuel Fu and returns them clamped as the reference stores
// them; VEL updates the velocity (emitter velocity blend and impulses), which no scalar term
// reads.  `cell` is the noise counter (the cell's linear index on its own grid).
template <bool SCAL, bool VEL>
__device__ __forceinline__ void cell_sources(const float* __restrict__ prm, const SrcConst& k, real px, real py,
                                             real pz, uint32_t cell, real& d, real& T, real& Fu, real& vx, real& vy,
                                             real& vz) {
    const real amb = k.amb, dt = k.dt;
    for (int e = 0; e < k.n_em; ++e) {
        const float* b = prm + P_EM0 + e * EM_STRIDE;
        const real rad = b[3];
        const real ddx = px - b[0], ddy = py - b[1], ddz = pz - b[2];
        const real r2 = (ddx * ddx + ddy * ddy + ddz * ddz) / (rad * rad);
        if (r2 < RL(1.0)) {
            const real w = (RL(1.0) - r2) * (RL(1.0) - r2);
            if (SCAL) {
                real drate = b[4], frate = b[6];
                const real te = b[5];
                real nu = RL(1.0);
                if (k.noisy) {
                    const uint32_t u = philox_u32(k.seed, (uint32_t)e, cell, k.stepno);
                    nu = k.nlo + (k.nhi - k.nlo) * ((real)(u >> 8) * (RL(1.0) / RL(16777216.0)));
                    drate = drate * nu;
                    frate = frate * nu;
                }
                d += drate * w * dt;
                Fu += frate * w * dt;
                if (te > RL(0.0)) {
                    const real tex = k.noisy ? (te - amb) * nu : (te - amb);
                    const real tt = amb + tex * w;
                    T = T > tt ? T : tt;
                }
            }
            if (VEL) {
                const real evx = b[7], evy = b[8], evz = b[9];
                if (evx * evx + evy * evy + evz * evz > RL(0.0)) {
                    vx = vx + (evx - vx) * w;
                    vy = vy + (evy - vy) * w;
                    vz = vz + (evz - vz) * w;
                }
            }
        }
    }
    if (VEL) {
        for (int m = 0; m < k.n_imp; ++m) {
            const float* b = prm + P_IMP0 + m * IMP_STRIDE;
            const real rad = b[3];
            const real ddx = px - b[0], ddy = py - b[1], ddz = pz - b[2];
            const real r2 = (ddx * ddx + ddy * ddy + ddz * ddz) / (rad * rad);
            if (r2 < RL(9.0)) {
                const real gk = exp(-r2);
                vx = vx + gk * b[4];
                vy = vy + gk * b[5];
                vz = vz + gk * b[6];
            }
        }
    }
    if (SCAL) {
        const real tmax = k.tmax;
        const real ign = clampf_((T - (real)prm[P_IGN]) / k.ramp, RL(0.0), RL(1.0));
        const real burned = Fu * (RL(1.0) - k.e_burn) * ign;
        Fu -= burned;
        const real room = clampf_((tmax - T) / (tmax - amb), RL(0.0), RL(1.0));
        T += prm[P_BHEAT] * burned * room;
        d += prm[P_SMOKE] * burned;
        T = amb + (T - amb) * k.e_cool;
        d = d * k.e_diss;
        d = d > RL(0.0) ? d : RL(0.0);
        T = T < tmax ? T : tmax;
        Fu = Fu > RL(0.0) ? Fu : RL(0.0);
    }
}

}  // namespace vfx
