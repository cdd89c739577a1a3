/* Oracle for the vfx.sim3d step: a plain-C (+OpenMP) restatement of the reference Warp solver
 * (vfx/sim3d/kernels.py, pressure.py, solver.py), in FP32 (_f32) and FP64 (_f64).
 *
 * TEST INFRASTRUCTURE: used only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg as the checker (and as the timed CPU path).  The product library does not use it.
 *
 * Parameter-buffer layout: kernels.py:11-53 (P_* slots, emitter stride 10, impulse stride 7),
 * plus the extension slots below which the reference leaves unused (11-15, 29).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define P_DT 0
#define P_WIND_X 1
#define P_WIND_Y 2
#define P_WIND_Z 3
#define P_N_EM 4
#define P_N_IMP 5
#define P_DX 6
#define P_AMB 7
#define P_KAPPA 8
#define P_CFLOOR 9
#define P_WARM 10
#define P_IGN_RAMP 11   /* extension: ignition ramp width in K (reference: 100, kernels.py:234) */
#define P_NOISE_LO 12   /* extension: emitter noise factor range [lo, hi]; 1,1 = off */
#define P_NOISE_HI 13
#define P_SEED 14       /* extension: noise seed */
#define P_STEPNO 15     /* extension: step counter fed to the noise */
#define P_ALPHA 16
#define P_BETA 17
#define P_EPS 18
#define P_DDISS 19
#define P_COOL 20
#define P_BURN 21
#define P_IGN 22
#define P_BHEAT 23
#define P_SMOKE 24
#define P_TOL 25
#define P_MAXIT 26
#define P_TMAX 27
#define P_WMARK 28
#define P_REL_TOL 29    /* extension: relative-residual stop ||r||/||b|| (0 = reference test) */
#define P_EM0 32
#define EM_STRIDE 10
#define MAX_EMITTERS 32
#define P_IMP0 (P_EM0 + MAX_EMITTERS * EM_STRIDE)
#define IMP_STRIDE 7
#define MAX_IMPULSES 32
#define P_SIZE (P_IMP0 + MAX_IMPULSES * IMP_STRIDE)

#define LAM_CAP 0.16f
#define FLUID_BIT 64
#define HS 0.35355339059327373
#define GALERKIN_SCALE 0.0625
#define OMEGA_JAC 0.8f

#define IDX(z, y, x) (((size_t)(z) * (size_t)N + (size_t)(y)) * (size_t)N + (size_t)(x))

/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11): 10 rounds of the 4x32 bijection with multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl
 * key increments 0x9E3779B9 / 0xBB67AE85. */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Emitter noise draw: key (seed, emitter), counter (cell, step, 'EMIT', 0); first word. */
static inline uint32_t or_philox_u32(uint32_t seed, uint32_t em, uint32_t cell, uint32_t step) {
    uint32_t c[4] = {cell, step, 0x454D4954u, 0u}, k[2] = {seed, em}, o[4];
    or_philox4x32_10(c, k, o);
    return o[0];
}

int or_p_size(void) { return P_SIZE; }

/* ---------------------------------------------------------------- 16-bit level storage
 * The multigrid levels may be stored in IEEE binary16 or in bfloat16 with FP32 arithmetic.  A
 * stored value is the FP32 value rounded to the nearest representable 16-bit value, ties to even
 * (binary16: 11 significant bits, exponent range down to the subnormal quantum 2^-24, overflow to
 * infinity from 65520 = 65504 + half an ulp on; bfloat16: the upper 16 bits of the float after
 * round-to-nearest-even on the lower 16).  Written from the formats' definitions (IEEE 754-2008
 * binary16; bfloat16 = float32 with a 7-bit fraction). */
static inline float or_round_fp16(float f) {
    if (!(f == f) || f == INFINITY || f == -INFINITY) return f;
    const float a = fabsf(f);
    float q;
    if (a >= 65520.0f) {
        q = INFINITY;
    } else if (a < 6.103515625e-05f) {          /* below 2^-14: subnormal binary16, quantum 2^-24 */
        q = nearbyintf(a * 16777216.0f) / 16777216.0f;
    } else {
        int e;
        frexpf(a, &e);                            /* a = m 2^e, m in [0.5, 1): quantum 2^(e - 11) */
        const float quantum = ldexpf(1.0f, e - 11);
        q = nearbyintf(a / quantum) * quantum;
        if (q > 65504.0f) q = INFINITY;
    }
    return copysignf(q, f);
}

static inline float or_round_bf16(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    if ((b & 0x7fffffffu) > 0x7f800000u) return f;   /* NaN stays NaN */
    b = (b + 0x7fffu + ((b >> 16) & 1u)) & 0xffff0000u;
    float r;
    memcpy(&r, &b, 4);
    return r;
}

/* exported for the format pins (tests/test_oracle_pressure.py) */
float or_round16(float f, int kind) { return kind == 1 ? or_round_fp16(f) : kind == 2 ? or_round_bf16(f) : f; }

#define R float
#define SFX _f32
#define EXPR expf
#define SQRTR sqrtf
#define FLOORR floorf
#include "sim3d_oracle_impl.h"
#undef R
#undef SFX
#undef EXPR
#undef SQRTR
#undef FLOORR

#define R double
#define SFX _f64
#define EXPR exp
#define SQRTR sqrt
#define FLOORR floor
#include "sim3d_oracle_impl.h"
#undef R
#undef SFX
#undef EXPR
#undef SQRTR
#undef FLOORR
