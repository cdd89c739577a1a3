/* CPU restatement of the reference sim3d step, instantiated twice by sim3d_oracle.c
 * (R = float with suffix _f32, R = double with suffix _f64).
 *
 * TEST INFRASTRUCTURE ONLY: this is the oracle the HIP path is checked against (and the CPU
 * baseline that bench.py times).  The product library never links or calls it.
 *
 * Every function names the reference lines it restates.  Arithmetic is written in the same
 * operation order as the reference so that the FP32 instantiation, built with
 * -ffp-contract=off, reproduces an FP32 device build rounding for rounding wherever no
 * transcendental or reduction order is involved.  Reductions accumulate in double.
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define F(name) CAT(name, SFX)

/* ---------------------------------------------------------------- trilinear helpers
 * kernels.py:67-85 (base_idx, frac, tri), :100-120 (sample_vel, trace) */
static inline R F(clampr)(R g, R lo, R hi) { return g < lo ? lo : (g > hi ? hi : g); }

static inline int F(base_idx)(R g, int n) {
    /* NaN lands on cell 0 (same as the clamp for every other input) */
    int i = (int)FLOORR(g >= (R)0 ? (g > (R)(n - 1) ? (R)(n - 1) : g) : (R)0);
    return i < n - 2 ? i : n - 2;
}

static inline R F(fracr)(R g, int n, int i0) { return F(clampr)(g, (R)0, (R)(n - 1)) - (R)i0; }

static inline R F(tri)(const R* f, int N, int z0, int y0, int x0, R fz, R fy, R fx) {
    R a00 = f[IDX(z0, y0, x0)] * ((R)1 - fx) + f[IDX(z0, y0, x0 + 1)] * fx;
    R a01 = f[IDX(z0, y0 + 1, x0)] * ((R)1 - fx) + f[IDX(z0, y0 + 1, x0 + 1)] * fx;
    R a10 = f[IDX(z0 + 1, y0, x0)] * ((R)1 - fx) + f[IDX(z0 + 1, y0, x0 + 1)] * fx;
    R a11 = f[IDX(z0 + 1, y0 + 1, x0)] * ((R)1 - fx) + f[IDX(z0 + 1, y0 + 1, x0 + 1)] * fx;
    R a0 = a00 * ((R)1 - fy) + a01 * fy;
    R a1 = a10 * ((R)1 - fy) + a11 * fy;
    return a0 * ((R)1 - fz) + a1 * fz;
}

typedef struct { int x0, y0, z0; R fx, fy, fz; } F(corner_t);

static inline F(corner_t) F(corner)(int N, R px, R py, R pz) {
    F(corner_t) c;
    c.x0 = F(base_idx)(px, N);
    c.y0 = F(base_idx)(py, N);
    c.z0 = F(base_idx)(pz, N);
    c.fx = F(fracr)(px, N, c.x0);
    c.fy = F(fracr)(py, N, c.y0);
    c.fz = F(fracr)(pz, N, c.z0);
    return c;
}

static inline void F(sample_vel)(const R* vel, int N, R px, R py, R pz, R* v) {
    F(corner_t) c = F(corner)(N, px, py, pz);
    size_t n3 = (size_t)N * N * N;
    for (int k = 0; k < 3; ++k) v[k] = F(tri)(vel + k * n3, N, c.z0, c.y0, c.x0, c.fz, c.fy, c.fx);
}

/* RK2 midpoint trace, s in cells per (m/s) (kernels.py:115-120) */
static inline void F(trace)(const R* vel, int N, R px, R py, R pz, R s, R* q) {
    R v1[3], v2[3];
    F(sample_vel)(vel, N, px, py, pz, v1);
    R h = (R)0.5 * s;
    F(sample_vel)(vel, N, px + h * v1[0], py + h * v1[1], pz + h * v1[2], v2);
    q[0] = px + s * v2[0];
    q[1] = py + s * v2[1];
    q[2] = pz + s * v2[2];
}

/* ---------------------------------------------------------------- MacCormack advection
 * advect_sl kernels.py:124-140 and advect_mc kernels.py:143-178, over C channels of src,
 * velocity `vel` (3 channels).  s = dt * inv_dx (the sign is applied here, as in the reference). */
void F(or_advect)(int N, int C, const R* src, const R* vel, R s, R* hat, R* out) {
    size_t n3 = (size_t)N * N * N;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                R q[3];
                F(trace)(vel, N, (R)x, (R)y, (R)z, -s, q);
                F(corner_t) c = F(corner)(N, q[0], q[1], q[2]);
                for (int ch = 0; ch < C; ++ch)
                    hat[ch * n3 + IDX(z, y, x)] = F(tri)(src + ch * n3, N, c.z0, c.y0, c.x0, c.fz, c.fy, c.fx);
            }
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                R q[3], r[3];
                F(trace)(vel, N, (R)x, (R)y, (R)z, -s, q);
                F(trace)(vel, N, (R)x, (R)y, (R)z, s, r);
                int qx = F(base_idx)(q[0], N), qy = F(base_idx)(q[1], N), qz = F(base_idx)(q[2], N);
                F(corner_t) cr = F(corner)(N, r[0], r[1], r[2]);
                for (int ch = 0; ch < C; ++ch) {
                    const R* sc = src + ch * n3;
                    R h = hat[ch * n3 + IDX(z, y, x)];
                    R bar = F(tri)(hat + ch * n3, N, cr.z0, cr.y0, cr.x0, cr.fz, cr.fy, cr.fx);
                    R val = h + (R)0.5 * (sc[IDX(z, y, x)] - bar);
                    R lo = sc[IDX(qz, qy, qx)], hi = lo;
                    for (int k = 1; k < 8; ++k) {
                        R v = sc[IDX(qz + (k >> 2), qy + ((k >> 1) & 1), qx + (k & 1))];
                        lo = v < lo ? v : lo;
                        hi = v > hi ? v : hi;
                    }
                    if (val < lo || val > hi) val = h;
                    out[ch * n3 + IDX(z, y, x)] = val;
                }
            }
}

/* Global fluid density ratio before / after advection (solver.py:46-74).  Sums in double,
 * summed per z-plane then over planes in order, so the result does not depend on threads. */
double F(or_mass_ratio)(int N, const R* pre, const R* post, const R* solid) {
    double* pa = (double*)calloc((size_t)N, sizeof(double));
    double* pb = (double*)calloc((size_t)N, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z) {
        double sa = 0.0, sb = 0.0;
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x)
                if (!(solid[IDX(z, y, x)] > (R)0.5)) {
                    sa += (double)pre[IDX(z, y, x)];
                    sb += (double)post[IDX(z, y, x)];
                }
        pa[z] = sa;
        pb[z] = sb;
    }
    double a = 0.0, b = 0.0;
    for (int z = 0; z < N; ++z) { a += pa[z]; b += pb[z]; }
    free(pa);
    free(pb);
    return (a > 0.0 && b > 1e-30) ? a / b : 1.0;
}

/* ---------------------------------------------------------------- sources
 * kernels.py:182-250 (mass correction, emitters, impulses, combustion, cooling, dissipation).
 * Extensions with the reference behaviour as default: the ignition ramp width (P_IGN_RAMP,
 * reference hard-codes 100 K at :234) and per-cell emitter noise (P_NOISE_LO/HI, off when
 * lo == hi == 1), drawn from Philox-4x32-10 keyed by (seed, emitter) with counter
 * (cell, step). */
void F(or_sources)(int N, const R* adv, const R* solid, const float* prm, R mass_ratio, R* out) {
    size_t n3 = (size_t)N * N * N;
    const R amb = (R)prm[P_AMB], dt = (R)prm[P_DT], dx = (R)prm[P_DX];
    const int n_em = (int)prm[P_N_EM], n_imp = (int)prm[P_N_IMP];
    const R ramp = (R)prm[P_IGN_RAMP];
    const R nlo = (R)prm[P_NOISE_LO], nhi = (R)prm[P_NOISE_HI];
    const int noisy = !(prm[P_NOISE_LO] == 1.0f && prm[P_NOISE_HI] == 1.0f);
    const uint32_t seed = (uint32_t)prm[P_SEED], stepno = (uint32_t)prm[P_STEPNO];
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                if (solid[i] > (R)0.5) {
                    out[0 * n3 + i] = 0; out[1 * n3 + i] = 0; out[2 * n3 + i] = 0;
                    out[3 * n3 + i] = 0; out[4 * n3 + i] = amb; out[5 * n3 + i] = 0;
                    continue;
                }
                R px = ((R)x + (R)0.5) * dx, py = ((R)y + (R)0.5) * dx, pz = ((R)z + (R)0.5) * dx;
                R d = adv[0 * n3 + i] * mass_ratio;
                R vx = adv[1 * n3 + i], vy = adv[2 * n3 + i], vz = adv[3 * n3 + i];
                R T = adv[4 * n3 + i], Fu = adv[5 * n3 + i];
                for (int e = 0; e < n_em; ++e) {
                    const float* b = prm + P_EM0 + e * EM_STRIDE;
                    R rad = (R)b[3];
                    R ddx = px - (R)b[0], ddy = py - (R)b[1], ddz = pz - (R)b[2];
                    R r2 = (ddx * ddx + ddy * ddy + ddz * ddz) / (rad * rad);
                    if (r2 < (R)1) {
                        R w = ((R)1 - r2) * ((R)1 - r2);
                        R nu = (R)1;
                        if (noisy) {
                            uint32_t u = or_philox_u32(seed, (uint32_t)e, (uint32_t)i, stepno);
                            nu = nlo + (nhi - nlo) * ((R)(u >> 8) * (R)(1.0 / 16777216.0));
                        }
                        R drate = (R)b[4], frate = (R)b[6], te = (R)b[5];
                        if (noisy) { drate = drate * nu; frate = frate * nu; }
                        d += drate * w * dt;
                        Fu += frate * w * dt;
                        if (te > (R)0) {
                            R tex = noisy ? (te - amb) * nu : (te - amb);
                            R tt = amb + tex * w;
                            T = T > tt ? T : tt;
                        }
                        R evx = (R)b[7], evy = (R)b[8], evz = (R)b[9];
                        if (evx * evx + evy * evy + evz * evz > (R)0) {
                            vx = vx + (evx - vx) * w;
                            vy = vy + (evy - vy) * w;
                            vz = vz + (evz - vz) * w;
                        }
                    }
                }
                for (int m = 0; m < n_imp; ++m) {
                    const float* b = prm + P_IMP0 + m * IMP_STRIDE;
                    R rad = (R)b[3];
                    R ddx = px - (R)b[0], ddy = py - (R)b[1], ddz = pz - (R)b[2];
                    R r2 = (ddx * ddx + ddy * ddy + ddz * ddz) / (rad * rad);
                    if (r2 < (R)9) {
                        R g = EXPR(-r2);
                        vx = vx + g * (R)b[4];
                        vy = vy + g * (R)b[5];
                        vz = vz + g * (R)b[6];
                    }
                }
                R ign = F(clampr)((T - (R)prm[P_IGN]) / ramp, (R)0, (R)1);
                R burned = Fu * ((R)1 - EXPR(-(R)prm[P_BURN] * dt)) * ign;
                Fu -= burned;
                R tmax = (R)prm[P_TMAX];
                R room = F(clampr)((tmax - T) / (tmax - amb), (R)0, (R)1);
                T += (R)prm[P_BHEAT] * burned * room;
                d += (R)prm[P_SMOKE] * burned;
                T = amb + (T - amb) * EXPR(-(R)prm[P_COOL] * dt);
                d = d * EXPR(-(R)prm[P_DDISS] * dt);
                out[0 * n3 + i] = d > (R)0 ? d : (R)0;
                out[1 * n3 + i] = vx;
                out[2 * n3 + i] = vy;
                out[3 * n3 + i] = vz;
                out[4 * n3 + i] = T < tmax ? T : tmax;
                out[5 * n3 + i] = Fu > (R)0 ? Fu : (R)0;
            }
}

/* ---------------------------------------------------------------- diffusion
 * kernels.py:253-286: one explicit 7-point substep of T and F, zero flux at walls and solids. */
static inline int F(solid_at)(const R* solid, int N, int z, int y, int x) {
    if (x < 0 || y < 0 || z < 0 || x >= N || y >= N || z >= N) return 1;
    return solid[IDX(z, y, x)] > (R)0.5;
}

void F(or_diffuse)(int N, const R* src, const R* solid, const float* prm, R nsub, R* dst) {
    size_t n3 = (size_t)N * N * N;
    R dx = (R)prm[P_DX];
    R lam = (R)prm[P_KAPPA] * (R)prm[P_DT] / (nsub * dx * dx);
    lam = lam < (R)LAM_CAP ? lam : (R)LAM_CAP;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                for (int c = 0; c <= 3; ++c) dst[c * n3 + i] = src[c * n3 + i];
                for (int c = 4; c <= 5; ++c) {
                    const R* f = src + c * n3;
                    R v = f[i];
                    if (solid[i] > (R)0.5 || lam <= (R)0) {
                        dst[c * n3 + i] = v;
                    } else {
#define NB(zz, yy, xx) (F(solid_at)(solid, N, zz, yy, xx) ? v : f[IDX(zz, yy, xx)])
                        R s = NB(z, y, x - 1) + NB(z, y, x + 1) + NB(z, y - 1, x) + NB(z, y + 1, x)
                              + NB(z - 1, y, x) + NB(z + 1, y, x);
#undef NB
                        dst[c * n3 + i] = v + lam * (s - (R)6 * v);
                    }
                }
            }
}

/* ---------------------------------------------------------------- vorticity
 * curl kernels.py:290-310 (walls and solids count as zero velocity) */
void F(or_curl)(int N, const R* data, const R* solid, R inv2dx, R* w) {
    size_t n3 = (size_t)N * N * N;
    const R *vx = data + 1 * n3, *vy = data + 2 * n3, *vz = data + 3 * n3;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
#define V0(f, zz, yy, xx) (F(solid_at)(solid, N, zz, yy, xx) ? (R)0 : f[IDX(zz, yy, xx)])
                R dwdy = V0(vz, z, y + 1, x) - V0(vz, z, y - 1, x);
                R dvdz = V0(vy, z + 1, y, x) - V0(vy, z - 1, y, x);
                R dudz = V0(vx, z + 1, y, x) - V0(vx, z - 1, y, x);
                R dwdx = V0(vz, z, y, x + 1) - V0(vz, z, y, x - 1);
                R dvdx = V0(vy, z, y, x + 1) - V0(vy, z, y, x - 1);
                R dudy = V0(vx, z, y + 1, x) - V0(vx, z, y - 1, x);
#undef V0
                size_t i = IDX(z, y, x);
                w[0 * n3 + i] = (dwdy - dvdz) * inv2dx;
                w[1 * n3 + i] = (dudz - dwdx) * inv2dx;
                w[2 * n3 + i] = (dvdx - dudy) * inv2dx;
            }
}

/* Separable fluid-renormalised [1,2,1]^3 smoothing, sep_x/sep_y/sep_z_ratio
 * kernels.py:390-435.  tmp_a, tmp_b are (4,N,N,N).  mode 0: out = ratio (sep_z_store :438-444);
 * mode 1: out (the 6-channel state) velocity += dt * ratio on fluid cells (sep_z_apply :447-457). */
void F(or_smooth3)(int N, const R* a, const R* solid, R* tmp_a, R* tmp_b, int mode, R dt, R* out) {
    size_t n3 = (size_t)N * N * N;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                R s0 = 0, s1 = 0, s2 = 0, ws = 0;
                for (int i = -1; i <= 1; ++i) {
                    if (!F(solid_at)(solid, N, z, y, x + i)) {
                        R q = (R)(2 - (i < 0 ? -i : i));
                        s0 += q * a[0 * n3 + IDX(z, y, x + i)];
                        s1 += q * a[1 * n3 + IDX(z, y, x + i)];
                        s2 += q * a[2 * n3 + IDX(z, y, x + i)];
                        ws += q;
                    }
                }
                size_t i0 = IDX(z, y, x);
                tmp_a[0 * n3 + i0] = s0; tmp_a[1 * n3 + i0] = s1; tmp_a[2 * n3 + i0] = s2; tmp_a[3 * n3 + i0] = ws;
            }
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x)
                for (int c = 0; c < 4; ++c) {
                    R v = (R)2 * tmp_a[c * n3 + IDX(z, y, x)];
                    if (y > 0) v += tmp_a[c * n3 + IDX(z, y - 1, x)];
                    if (y + 1 < N) v += tmp_a[c * n3 + IDX(z, y + 1, x)];
                    tmp_b[c * n3 + IDX(z, y, x)] = v;
                }
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i0 = IDX(z, y, x);
                if (mode == 1 && solid[i0] > (R)0.5) continue;
                R s0 = 0, s1 = 0, s2 = 0, ws = 0;
                for (int k = -1; k <= 1; ++k) {
                    int zz = z + k;
                    if (zz >= 0 && zz < N) {
                        R q = (R)(2 - (k < 0 ? -k : k));
                        s0 += q * tmp_b[0 * n3 + IDX(zz, y, x)];
                        s1 += q * tmp_b[1 * n3 + IDX(zz, y, x)];
                        s2 += q * tmp_b[2 * n3 + IDX(zz, y, x)];
                        ws += q * tmp_b[3 * n3 + IDX(zz, y, x)];
                    }
                }
                R r0 = 0, r1 = 0, r2 = 0;
                if (ws > (R)0) { r0 = s0 / ws; r1 = s1 / ws; r2 = s2 / ws; }
                if (mode == 0) {
                    out[0 * n3 + i0] = r0; out[1 * n3 + i0] = r1; out[2 * n3 + i0] = r2;
                } else {
                    out[1 * n3 + i0] = out[1 * n3 + i0] + dt * r0;
                    out[2 * n3 + i0] = out[2 * n3 + i0] + dt * r1;
                    out[3 * n3 + i0] = out[3 * n3 + i0] + dt * r2;
                }
            }
}

/* 27-tap form binom3 (kernels.py:313-329): the reference states it equals the separable
 * passes; kept only as an independent check of or_smooth3. */
void F(or_binom3)(int N, const R* a, const R* solid, R* out) {
    size_t n3 = (size_t)N * N * N;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                R s0 = 0, s1 = 0, s2 = 0, ws = 0;
                for (int k = -1; k <= 1; ++k)
                    for (int j = -1; j <= 1; ++j)
                        for (int i = -1; i <= 1; ++i)
                            if (!F(solid_at)(solid, N, z + k, y + j, x + i)) {
                                R q = (R)((2 - abs(i)) * (2 - abs(j)) * (2 - abs(k)));
                                size_t t = IDX(z + k, y + j, x + i);
                                s0 += q * a[0 * n3 + t]; s1 += q * a[1 * n3 + t]; s2 += q * a[2 * n3 + t];
                                ws += q;
                            }
                size_t i0 = IDX(z, y, x);
                if (ws > (R)0) { out[i0] = s0 / ws; out[n3 + i0] = s1 / ws; out[2 * n3 + i0] = s2 / ws; }
                else { out[i0] = 0; out[n3 + i0] = 0; out[2 * n3 + i0] = 0; }
            }
}

/* body_forces kernels.py:341-387: buoyancy + vorticity confinement + wind marker force. */
static inline R F(wmag)(const R* o, size_t n3, size_t i) {
    R a = o[i], b = o[n3 + i], c = o[2 * n3 + i];
    return SQRTR(a * a + b * b + c * c);
}

void F(or_body_forces)(int N, const R* f, const R* solid, const R* om, const float* prm, R* force) {
    size_t n3 = (size_t)N * N * N;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                if (solid[i] > (R)0.5) { force[i] = 0; force[n3 + i] = 0; force[2 * n3 + i] = 0; continue; }
                R amb = (R)prm[P_AMB];
                R d = f[0 * n3 + i], T = f[4 * n3 + i];
                R fx = 0, fy = -(R)prm[P_ALPHA] * d + (R)prm[P_BETA] * (T - amb), fz = 0;
                R eps = (R)prm[P_EPS];
                if (eps != (R)0) {
                    int xr = x + 1 < N - 1 ? x + 1 : N - 1, xl = x - 1 > 0 ? x - 1 : 0;
                    int yt = y + 1 < N - 1 ? y + 1 : N - 1, yb = y - 1 > 0 ? y - 1 : 0;
                    int zf = z + 1 < N - 1 ? z + 1 : N - 1, zb = z - 1 > 0 ? z - 1 : 0;
                    R gx = (F(wmag)(om, n3, IDX(z, y, xr)) - F(wmag)(om, n3, IDX(z, y, xl))) / (R)(xr - xl > 1 ? xr - xl : 1);
                    R gy = (F(wmag)(om, n3, IDX(z, yt, x)) - F(wmag)(om, n3, IDX(z, yb, x))) / (R)(yt - yb > 1 ? yt - yb : 1);
                    R gz = (F(wmag)(om, n3, IDX(zf, y, x)) - F(wmag)(om, n3, IDX(zb, y, x))) / (R)(zf - zb > 1 ? zf - zb : 1);
                    R wx = om[i], wy = om[n3 + i], wz = om[2 * n3 + i];
                    R fl = (R)prm[P_CFLOOR] * SQRTR(wx * wx + wy * wy + wz * wz);
                    R mag = SQRTR(gx * gx + gy * gy + gz * gz + fl * fl) + (R)1.0e-20;
                    R k = eps * (R)prm[P_DX] / mag;
                    fx += k * (gy * wz - gz * wy);
                    fy += k * (gz * wx - gx * wz);
                    fz += k * (gx * wy - gy * wx);
                }
                R wdx = (R)prm[P_WIND_X], wdy = (R)prm[P_WIND_Y], wdz = (R)prm[P_WIND_Z];
                if (wdx * wdx + wdy * wdy + wdz * wdz > (R)0) {
                    R m = F(clampr)(d + f[5 * n3 + i] + (T - amb) / (R)prm[P_WMARK], (R)0, (R)1);
                    fx += wdx * m; fy += wdy * m; fz += wdz * m;
                }
                force[i] = fx; force[n3 + i] = fy; force[2 * n3 + i] = fz;
            }
}

/* ---------------------------------------------------------------- collocated operators
 * pressure.py:36-111 (mask bits, g_ax, kmul, kdiag, div2) and k_mask :115-143 */
void F(or_mask)(int N, const R* solid, int* fm) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                if (solid[i] > (R)0.5) { fm[i] = 0; continue; }
                int m = FLUID_BIT;
                if (x + 1 < N && solid[IDX(z, y, x + 1)] <= (R)0.5) m |= 1;
                if (x > 0 && solid[IDX(z, y, x - 1)] <= (R)0.5) m |= 2;
                if (y + 1 < N && solid[IDX(z, y + 1, x)] <= (R)0.5) m |= 4;
                if (y > 0 && solid[IDX(z, y - 1, x)] <= (R)0.5) m |= 8;
                if (z + 1 < N && solid[IDX(z + 1, y, x)] <= (R)0.5) m |= 16;
                if (z > 0 && solid[IDX(z - 1, y, x)] <= (R)0.5) m |= 32;
                fm[i] = m;
            }
}

static inline size_t F(nbi)(int N, int z, int y, int x, int ax, int sgn) {
    if (ax == 0) return IDX(z, y, x + sgn);
    if (ax == 1) return IDX(z, y + sgn, x);
    return IDX(z + sgn, y, x);
}

static inline R F(g_ax)(const R* p, const int* fm, int N, int z, int y, int x, int ax) {
    size_t i = IDX(z, y, x);
    int m = fm[i];
    R g = 0, c = p[i];
    if ((m >> (2 * ax)) & 1) g += p[F(nbi)(N, z, y, x, ax, 1)] - c;
    if ((m >> (2 * ax + 1)) & 1) g += c - p[F(nbi)(N, z, y, x, ax, -1)];
    return g;
}

static inline R F(kmul)(const R* p, const int* fm, int N, int z, int y, int x) {
    int m = fm[IDX(z, y, x)];
    R d = 0;
    for (int ax = 0; ax < 3; ++ax) {
        int ex = ax == 0, ey = ax == 1, ez = ax == 2;
        R gi = F(g_ax)(p, fm, N, z, y, x, ax);
        if ((m >> (2 * ax)) & 1) d += gi + F(g_ax)(p, fm, N, z + ez, y + ey, x + ex, ax);
        if ((m >> (2 * ax + 1)) & 1) d -= gi + F(g_ax)(p, fm, N, z - ez, y - ey, x - ex, ax);
    }
    return -d;
}

static inline R F(kdiag)(int m) {
    R s = 0;
    for (int ax = 0; ax < 3; ++ax) {
        R oR = (R)((m >> (2 * ax)) & 1), oL = (R)((m >> (2 * ax + 1)) & 1);
        s += (oL - oR) * (oL - oR) + oL + oR;
    }
    return s;
}

static inline R F(div2)(const R* v, const int* fm, int N, int z, int y, int x) {
    size_t n3 = (size_t)N * N * N;
    int m = fm[IDX(z, y, x)];
    R d = 0;
    for (int ax = 0; ax < 3; ++ax) {
        const R* va = v + ax * n3;
        R u = va[IDX(z, y, x)];
        if ((m >> (2 * ax)) & 1) d += u + va[F(nbi)(N, z, y, x, ax, 1)];
        if ((m >> (2 * ax + 1)) & 1) d -= u + va[F(nbi)(N, z, y, x, ax, -1)];
    }
    return d;
}

/* out = K p at fluid cells, 0 in solids (k_matvec_dot pressure.py:629-646 without the dot) */
void F(or_kmul)(int N, const int* fm, const R* p, R* out) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x)
                out[IDX(z, y, x)] = fm[IDX(z, y, x)] ? F(kmul)(p, fm, N, z, y, x) : (R)0;
}

/* out = d(u) (2 dx div u) at fluid cells, 0 in solids (div2 pressure.py:94-106) */
void F(or_div2)(int N, const int* fm, const R* vel3, R* out) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x)
                out[IDX(z, y, x)] = fm[IDX(z, y, x)] ? F(div2)(vel3, fm, N, z, y, x) : (R)0;
}

/* u -= g(phi) on fluid, u = 0 in solids (k_apply pressure.py:166-175) */
void F(or_apply)(int N, const int* fm, const R* phi, R* vel3) {
    size_t n3 = (size_t)N * N * N;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                if (fm[i] == 0) { vel3[i] = 0; vel3[n3 + i] = 0; vel3[2 * n3 + i] = 0; continue; }
                for (int ax = 0; ax < 3; ++ax)
                    vel3[ax * n3 + i] = vel3[ax * n3 + i] - F(g_ax)(phi, fm, N, z, y, x, ax);
            }
}

/* |div u| * dx = |d(u)| / 2, 0 in solids (k_div_abs pressure.py:156-163) */
void F(or_div_abs)(int N, const int* fm, const R* vel3, R* out) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                R d = fm[i] ? F(div2)(vel3, fm, N, z, y, x) : (R)0;
                out[i] = fm[i] ? (R)0.5 * (d < 0 ? -d : d) : (R)0;
            }
}

/* ---------------------------------------------------------------- multigrid (Hadamard basis)
 * pressure.py:178-268 (hsign, tv, gt, axcell, k_probe1, k_galerkin).  Level arrays use the
 * reference layout: x, r, diag (8,m,m,m); off (24,m,m,m) with off[ax*8+c] the +ax link.
 *
 * Level storage (an extension: the reference stores FP32): F(g_mg16) = 1 or 2 stores every level
 * value (x, r, diag, off) rounded to binary16 / bfloat16 (F(qst)), with FP32 arithmetic, and the
 * level-1 right-hand side scaled by F(g_lscale) (a power of two; the inverse transform divides it
 * out).  0 (the default, and always in the FP64 instantiation) is the reference's FP32 storage. */
static int F(g_mg16) = 0;
static R F(g_lscale) = (R)1;
static inline R F(qst)(R v) {
    if (F(g_mg16) == 1) return (R)or_round_fp16((float)v);
    if (F(g_mg16) == 2) return (R)or_round_bf16((float)v);
    return v;
}
static inline R F(hsign)(int c, int dz, int dy, int dx) {
    int e = (c & 1) * dx + ((c >> 1) & 1) * dy + ((c >> 2) & 1) * dz;
    return (e & 1) ? -(R)HS : (R)HS;
}

static inline R F(tv)(int c, int BK, int BJ, int BI, int z, int y, int x) {
    if ((z >> 1) == BK && (y >> 1) == BJ && (x >> 1) == BI) return F(hsign)(c, z & 1, y & 1, x & 1);
    return 0;
}

static inline int F(mask_at)(const int* fm, int N, int z, int y, int x) {
    if (x < 0 || y < 0 || z < 0 || z >= N || y >= N || x >= N) return 0;
    return fm[IDX(z, y, x)];
}

/* component ax of g(t_{c,B}) at fine cell (z,y,x) */
static inline R F(gt)(const int* fm, int N, int c, int BK, int BJ, int BI, int z, int y, int x, int ax) {
    int m = F(mask_at)(fm, N, z, y, x);
    int ex = ax == 0, ey = ax == 1, ez = ax == 2;
    R t0 = F(tv)(c, BK, BJ, BI, z, y, x);
    R g = 0;
    if ((m >> (2 * ax)) & 1) g += F(tv)(c, BK, BJ, BI, z + ez, y + ey, x + ex) - t0;
    if ((m >> (2 * ax + 1)) & 1) g += t0 - F(tv)(c, BK, BJ, BI, z - ez, y - ey, x - ex);
    return g;
}

static inline void F(axcell)(int ax, int oz, int oy, int ox, int a, int b, int c, int* z, int* y, int* x) {
    if (ax == 0) { *z = oz + c; *y = oy + b; *x = ox + a; }
    else if (ax == 1) { *z = oz + c; *y = oy + a; *x = ox + b; }
    else { *z = oz + a; *y = oy + c; *x = ox + b; }
}

void F(or_probe1)(int N, const int* fm, R* diag, R* off) {
    int m = N / 2;
    size_t m3 = (size_t)m * m * m;
#pragma omp parallel for schedule(static)
    for (int K = 0; K < m; ++K)
        for (int J = 0; J < m; ++J)
            for (int I = 0; I < m; ++I)
                for (int c = 0; c < 8; ++c) {
                    int oz = 2 * K, oy = 2 * J, ox = 2 * I;
                    R dsum = 0;
                    for (int ax = 0; ax < 3; ++ax)
                        for (int b = 0; b < 2; ++b)
                            for (int cc = 0; cc < 2; ++cc)
                                for (int a = 0; a < 4; ++a) {
                                    int z, y, x;
                                    F(axcell)(ax, oz, oy, ox, a - 1, b, cc, &z, &y, &x);
                                    R v = F(gt)(fm, N, c, K, J, I, z, y, x, ax);
                                    dsum += v * v;
                                }
                    size_t li = ((size_t)K * m + J) * m + I;
                    diag[c * m3 + li] = F(qst)(dsum);
                    for (int ax = 0; ax < 3; ++ax) {
                        int nK = K + (ax == 2), nJ = J + (ax == 1), nI = I + (ax == 0);
                        R s = 0;
                        if (nK < m && nJ < m && nI < m) {
                            for (int b = 0; b < 2; ++b)
                                for (int cc = 0; cc < 2; ++cc)
                                    for (int a = 0; a < 2; ++a) {
                                        int z, y, x;
                                        F(axcell)(ax, oz, oy, ox, 1 + a, b, cc, &z, &y, &x);
                                        s += F(gt)(fm, N, c, K, J, I, z, y, x, ax) * F(gt)(fm, N, c, nK, nJ, nI, z, y, x, ax);
                                    }
                        }
                        off[(ax * 8 + c) * m3 + li] = F(qst)(s);
                    }
                }
}

void F(or_galerkin)(int mf, const R* diag, const R* off, R scale, R* dc, R* oc) {
    int m = mf / 2;
    size_t mf3 = (size_t)mf * mf * mf, m3 = (size_t)m * m * m;
#define FI(z, y, x) (((size_t)(z) * mf + (y)) * mf + (x))
#pragma omp parallel for schedule(static)
    for (int K = 0; K < m; ++K)
        for (int J = 0; J < m; ++J)
            for (int I = 0; I < m; ++I)
                for (int c = 0; c < 8; ++c) {
                    int oz = 2 * K, oy = 2 * J, ox = 2 * I;
                    R d = 0;
                    for (int k = 0; k < 8; ++k) d += diag[c * mf3 + FI(oz + (k >> 2), oy + ((k >> 1) & 1), ox + (k & 1))];
                    size_t li = ((size_t)K * m + J) * m + I;
                    for (int ax = 0; ax < 3; ++ax) {
                        R oi = 0, oo = 0;
                        for (int b = 0; b < 2; ++b)
                            for (int cc = 0; cc < 2; ++cc) {
                                int z, y, x;
                                F(axcell)(ax, oz, oy, ox, 0, b, cc, &z, &y, &x);
                                oi += off[(ax * 8 + c) * mf3 + FI(z, y, x)];
                                F(axcell)(ax, oz, oy, ox, 1, b, cc, &z, &y, &x);
                                oo += off[(ax * 8 + c) * mf3 + FI(z, y, x)];
                            }
                        d += (R)2 * oi;
                        oc[(ax * 8 + c) * m3 + li] = F(qst)(oo * scale);
                    }
                    dc[c * m3 + li] = F(qst)(d * scale);
                }
#undef FI
}

/* level hierarchy of PressureSolver3D.__init__ (pressure.py:799-809) */
typedef struct {
    int nlev;
    int m[16];
    R dmin[16];
    R *x[16], *r[16], *diag[16], *off[16];
} F(levels_t);

static void F(levels_init)(F(levels_t) * L, int N) {
    int m = N / 2, level = 1;
    L->nlev = 0;
    for (;;) {
        int l = L->nlev++;
        L->m[l] = m;
        L->dmin[l] = (R)(float)(1.0e-3 * 6.0 * pow(0.25, level - 1));
        size_t m3 = (size_t)m * m * m;
        L->x[l] = (R*)calloc(8 * m3, sizeof(R));
        L->r[l] = (R*)calloc(8 * m3, sizeof(R));
        L->diag[l] = (R*)calloc(8 * m3, sizeof(R));
        L->off[l] = (R*)calloc(24 * m3, sizeof(R));
        if (m % 2 == 1 || m == 1) break;
        m /= 2;
        level += 1;
    }
}

static void F(levels_free)(F(levels_t) * L) {
    for (int l = 0; l < L->nlev; ++l) { free(L->x[l]); free(L->r[l]); free(L->diag[l]); free(L->off[l]); }
}

static void F(levels_build)(F(levels_t) * L, int N, const int* fm) {
    F(or_probe1)(N, fm, L->diag[0], L->off[0]);
    for (int l = 0; l + 1 < L->nlev; ++l)
        F(or_galerkin)(L->m[l], L->diag[l], L->off[l], (R)GALERKIN_SCALE, L->diag[l + 1], L->off[l + 1]);
}

/* offsum pressure.py:271-287 */
static inline R F(offsum)(const R* xv, const R* off, int m, int c, int K, int J, int I) {
    size_t m3 = (size_t)m * m * m;
#define LI(k, j, i) (((size_t)(k) * m + (j)) * m + (i))
    R a = 0;
    if (I + 1 < m) a += off[(0 + c) * m3 + LI(K, J, I)] * xv[c * m3 + LI(K, J, I + 1)];
    if (I > 0) a += off[(0 + c) * m3 + LI(K, J, I - 1)] * xv[c * m3 + LI(K, J, I - 1)];
    if (J + 1 < m) a += off[(8 + c) * m3 + LI(K, J, I)] * xv[c * m3 + LI(K, J + 1, I)];
    if (J > 0) a += off[(8 + c) * m3 + LI(K, J - 1, I)] * xv[c * m3 + LI(K, J - 1, I)];
    if (K + 1 < m) a += off[(16 + c) * m3 + LI(K, J, I)] * xv[c * m3 + LI(K + 1, J, I)];
    if (K > 0) a += off[(16 + c) * m3 + LI(K - 1, J, I)] * xv[c * m3 + LI(K - 1, J, I)];
    return a;
}

/* one colour of red-black GS (k_smooth pressure.py:290-303); red = colour 0 */
static void F(mg_smooth)(F(levels_t) * L, int l, int color, int zero_init) {
    int m = L->m[l];
    size_t m3 = (size_t)m * m * m;
    R *xv = L->x[l], *r = L->r[l], *dg = L->diag[l], *off = L->off[l];
    R dmin = L->dmin[l];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < 8; ++c)
        for (int K = 0; K < m; ++K)
            for (int J = 0; J < m; ++J)
                for (int I = 0; I < m; ++I) {
                    if (((I + J + K) & 1) != color) continue;
                    size_t li = c * m3 + LI(K, J, I);
                    R d = dg[li];
                    if (d <= dmin) { xv[li] = 0; continue; }
                    R acc = r[li];
                    if (!zero_init) acc -= F(offsum)(xv, off, m, c, K, J, I);
                    xv[li] = F(qst)(acc / d);
                }
}

/* residual restriction, average of 8 children (k_resrestrict pressure.py:306-316) */
static void F(mg_restrict)(F(levels_t) * L, int l) {
    int mf = L->m[l], m = L->m[l + 1];
    size_t mf3 = (size_t)mf * mf * mf, m3 = (size_t)m * m * m;
    R *xv = L->x[l], *r = L->r[l], *dg = L->diag[l], *off = L->off[l], *rc = L->r[l + 1];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < 8; ++c)
        for (int K = 0; K < m; ++K)
            for (int J = 0; J < m; ++J)
                for (int I = 0; I < m; ++I) {
                    R acc = 0;
                    for (int k = 0; k < 8; ++k) {
                        int z = 2 * K + (k >> 2), y = 2 * J + ((k >> 1) & 1), x = 2 * I + (k & 1);
                        size_t fi = c * mf3 + (((size_t)z * mf + y) * mf + x);
                        acc += r[fi] - dg[fi] * xv[fi] - F(offsum)(xv, off, mf, c, z, y, x);
                    }
                    rc[c * m3 + LI(K, J, I)] = F(qst)((R)0.125 * acc);
                }
}

/* constant prolongation (k_prolong pressure.py:319-325) */
static void F(mg_prolong)(F(levels_t) * L, int l) {
    int m = L->m[l], mc = L->m[l + 1];
    size_t m3 = (size_t)m * m * m, mc3 = (size_t)mc * mc * mc;
    R *xv = L->x[l], *dg = L->diag[l], *xc = L->x[l + 1];
    R dmin = L->dmin[l];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < 8; ++c)
        for (int K = 0; K < m; ++K)
            for (int J = 0; J < m; ++J)
                for (int I = 0; I < m; ++I) {
                    size_t li = c * m3 + LI(K, J, I);
                    if (dg[li] <= dmin) continue;
                    xv[li] = F(qst)(xv[li] + xc[c * mc3 + (((size_t)(K >> 1) * mc + (J >> 1)) * mc + (I >> 1))]);
                }
}
#undef LI

/* V(1,1) on the level-1 system x1 = V(r1): the plain V-cycle of _vcycle_base
 * (pressure.py:902-931), which the reference's fast V-cycle reproduces with fewer launches
 * (pressure.py:870-873).  Coarsest level: exact diagonal solve when m == 1, else 2m red-black
 * sweep pairs (k_coarsest :328-336 / _vcycle_base :916-922). */
static void F(mg_vcycle)(F(levels_t) * L) {
    int nl = L->nlev;
    for (int l = 0; l + 1 < nl; ++l) {
        F(mg_smooth)(L, l, 0, 1);
        F(mg_smooth)(L, l, 1, 0);
        F(mg_restrict)(L, l);
    }
    int c = nl - 1;
    if (L->m[c] == 1) {
        for (int k = 0; k < 8; ++k) {
            R d = L->diag[c][k];
            L->x[c][k] = d <= L->dmin[c] ? (R)0 : F(qst)(L->r[c][k] / d);
        }
    } else {
        for (int k = 0; k < 2 * L->m[c]; ++k) {
            F(mg_smooth)(L, c, 0, k == 0);
            F(mg_smooth)(L, c, 1, 0);
        }
    }
    for (int l = nl - 2; l >= 0; --l) {
        F(mg_prolong)(L, l);
        F(mg_smooth)(L, l, 1, 0);
        F(mg_smooth)(L, l, 0, 0);
    }
}

/* Hadamard forward transform of a fine field to level 1 (k_fwd_transform :339-349) */
static void F(fwd_transform)(int N, const R* rr, R* r1) {
    int m = N / 2;
    size_t m3 = (size_t)m * m * m;
#pragma omp parallel for schedule(static)
    for (int K = 0; K < m; ++K)
        for (int J = 0; J < m; ++J)
            for (int I = 0; I < m; ++I) {
                R v[8];
                for (int k = 0; k < 8; ++k) v[k] = rr[IDX(2 * K + (k >> 2), 2 * J + ((k >> 1) & 1), 2 * I + (k & 1))];
                for (int c = 0; c < 8; ++c) {
                    R acc = 0;
                    for (int k = 0; k < 8; ++k) acc += F(hsign)(c, k >> 2, (k >> 1) & 1, k & 1) * v[k];
                    r1[c * m3 + ((size_t)K * m + J) * m + I] = F(qst)(acc * F(g_lscale));
                }
            }
}

/* z_out = base + T^T x1, 0 in solids (k_inv_transform :352-365) */
static void F(inv_transform)(int N, const R* x1, const int* fm, const R* base, R* zout) {
    int m = N / 2;
    const R isc = (R)1 / F(g_lscale);
    size_t m3 = (size_t)m * m * m;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                if (fm[i] == 0) { zout[i] = 0; continue; }
                size_t li = ((size_t)(z >> 1) * m + (y >> 1)) * m + (x >> 1);
                R acc = base[i];
                for (int c = 0; c < 8; ++c) acc += F(hsign)(c, z & 1, y & 1, x & 1) * (x1[c * m3 + li] * isc);
                zout[i] = acc;
            }
}

/* The preconditioner z = M^-1 r: damped Jacobi on exact K, V-cycle in the Hadamard basis,
 * damped Jacobi again (k_pre/k_resid/V-cycle/k_inv_transform/k_post_rz, pressure.py:648-710).
 * Returns the partial results rz = r.z (double) and max|r|. */
static void F(precond)(int N, const int* fm, F(levels_t) * L, const R* r, R* z0, R* rr, R* zt, R* zout,
                      double* rz, double* maxr) {
    const R omega = (R)OMEGA_JAC;
    size_t n3 = (size_t)N * N * N;
#pragma omp parallel for schedule(static)
    /* (fm & 63): a fluid cell sealed in by solids has diag(K) = 0 and a zero row and column of K;
     * its Jacobi terms are 0 (the reference divides 0 / 0 there, pressure.py:671, :702) */
    for (size_t i = 0; i < n3; ++i) z0[i] = (fm[i] & 63) ? omega * r[i] / F(kdiag)(fm[i]) : (R)0;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                rr[i] = fm[i] ? r[i] - F(kmul)(z0, fm, N, z, y, x) : (R)0;
            }
    F(fwd_transform)(N, rr, L->r[0]);
    F(mg_vcycle)(L);
    F(inv_transform)(N, L->x[0], fm, z0, zt);
    double* ps = (double*)calloc((size_t)N, sizeof(double));
    double* pm = (double*)calloc((size_t)N, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z) {
        double s = 0, mx = 0;
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                R rv = r[i], zv = 0;
                if (fm[i] & 63) zv = zt[i] + omega * (rv - F(kmul)(zt, fm, N, z, y, x)) / F(kdiag)(fm[i]);
                zout[i] = zv;
                s += (double)rv * (double)zv;
                double a = fabs((double)rv);
                mx = a > mx ? a : mx;
            }
        ps[z] = s;
        pm[z] = mx;
    }
    double s = 0, mx = 0;
    for (int z = 0; z < N; ++z) { s += ps[z]; mx = pm[z] > mx ? pm[z] : mx; }
    free(ps);
    free(pm);
    *rz = s;
    *maxr = mx;
}

/* Exported: the same with level storage `kind` (0 FP32, 1 binary16, 2 bfloat16; FP32 instantiation)
 * and the level-1 right-hand side scaled by 2^-e, 2^e <= max|r| over fluid cells < 2^(e+1) (1 for
 * FP32 storage): the scale the solver takes from the residual, so 16-bit values sit near 1. */
double F(or_precond_apply)(int N, const int* fm, const R* r, R* zout);
double F(or_precond_apply16)(int N, const int* fm, const R* r, R* zout, int kind) {
    size_t n3 = (size_t)N * N * N;
    double mr = 0.0;
    for (size_t i = 0; i < n3; ++i)
        if (fm[i] && fabs((double)r[i]) > mr) mr = fabs((double)r[i]);
    F(g_mg16) = kind;
    F(g_lscale) = kind != 0 && mr > 0.0 && mr < 1e300 ? (R)ldexp(1.0, -ilogb(mr)) : (R)1;
    const double rz = F(or_precond_apply)(N, fm, r, zout);
    F(g_mg16) = 0;
    F(g_lscale) = (R)1;
    return rz;
}

/* Exported: build the level hierarchy for a mask and apply z = M^-1 r once (for tests of the
 * preconditioner's symmetry and of the GPU V-cycle).  Returns r.z. */
double F(or_precond_apply)(int N, const int* fm, const R* r, R* zout) {
    size_t n3 = (size_t)N * N * N;
    F(levels_t) L;
    F(levels_init)(&L, N);
    F(levels_build)(&L, N, fm);
    R* z0 = (R*)calloc(n3, sizeof(R));
    R* rr = (R*)calloc(n3, sizeof(R));
    R* zt = (R*)calloc(n3, sizeof(R));
    double rz, mr;
    F(precond)(N, fm, &L, r, z0, rr, zt, zout, &rz, &mr);
    free(z0); free(rr); free(zt);
    F(levels_free)(&L);
    return rz;
}

/* Exported: level-1 and coarser operators in the reference layout, concatenated level by level
 * (diag then off per level).  Returns the number of levels. */
int F(or_build_ops)(int N, const int* fm, R* out) {
    F(levels_t) L;
    F(levels_init)(&L, N);
    F(levels_build)(&L, N, fm);
    size_t o = 0;
    for (int l = 0; l < L.nlev; ++l) {
        size_t m3 = (size_t)L.m[l] * L.m[l] * L.m[l];
        memcpy(out + o, L.diag[l], 8 * m3 * sizeof(R)); o += 8 * m3;
        memcpy(out + o, L.off[l], 24 * m3 * sizeof(R)); o += 24 * m3;
    }
    int n = L.nlev;
    F(levels_free)(&L);
    return n;
}

/* ---------------------------------------------------------------- PCG projection
 * PressureSolver3D.record (pressure.py:940-950) with k_rhs :602-627, k_fin_maxu :730-735,
 * k_fin_rz :737-771, k_upd_p :712-728, k_apply :166-175.  Dot products, alpha and beta are
 * double; the vectors are R.  Stop test of the reference (normalised max residual, max
 * iterations, rz <= 0, stagnation 4x above the best), plus, when rel_tol > 0, the relative
 * residual ||r||_2 / ||b||_2 <= rel_tol of the recurrence instead of the normalised test.
 * Returns the iteration count; stats[0] = ||b||_2, stats[1] = recurrence ||r||_2,
 * stats[2] = max|u*|. */
int F(or_project)(int N, R* data, const R* solid, const float* prm, R* phi, double* stats) {
    size_t n3 = (size_t)N * N * N;
    int* fm = (int*)calloc(n3, sizeof(int));
    F(or_mask)(N, solid, fm);
    F(levels_t) L;
    F(levels_init)(&L, N);
    F(levels_build)(&L, N, fm);
    R *r = (R*)calloc(n3, sizeof(R)), *z = (R*)calloc(n3, sizeof(R)), *p = (R*)calloc(n3, sizeof(R));
    R *q = (R*)calloc(n3, sizeof(R)), *z0 = (R*)calloc(n3, sizeof(R)), *rr = (R*)calloc(n3, sizeof(R));
    R* zt = (R*)calloc(n3, sizeof(R));
    R* vel = data + n3;
    const int warm = prm[P_WARM] > 0.5f;
    const float tol = prm[P_TOL] > 1.0e-6f ? prm[P_TOL] : 1.0e-6f;
    const int maxit = (int)prm[P_MAXIT];
    const double rel_tol = (double)prm[P_REL_TOL];

    /* per-z-plane partial sums, combined in plane order: independent of the thread count */
    double* pa = (double*)calloc((size_t)N, sizeof(double));
    double* pb = (double*)calloc((size_t)N, sizeof(double));

    /* k_rhs: r = -d(u*) - K phi0; max|u*| over fluid; ||b||^2 */
    double maxu = 0, bb = 0;
    if (!warm) memset(phi, 0, n3 * sizeof(R));
#pragma omp parallel for schedule(static)
    for (int zz = 0; zz < N; ++zz) {
        double sb = 0, mu = 0;
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(zz, y, x);
                if (fm[i] == 0) { r[i] = 0; continue; }
                R b = -F(div2)(vel, fm, N, zz, y, x);
                R rv = b;
                if (warm) rv -= F(kmul)(phi, fm, N, zz, y, x);
                r[i] = rv;
                sb += (double)b * (double)b;
                R ux = vel[i], uy = vel[n3 + i], uz = vel[2 * n3 + i];
                double sp = (double)SQRTR(ux * ux + uy * uy + uz * uz);
                mu = sp > mu ? sp : mu;
            }
        pa[zz] = sb;
        pb[zz] = mu;
    }
    for (int zz = 0; zz < N; ++zz) { bb += pa[zz]; maxu = pb[zz] > maxu ? pb[zz] : maxu; }
    const double thr = (double)(2.0f * tol * (float)maxu);
    const double bnorm = sqrt(bb);

#define RNORM(out)                                                            \
    do {                                                                      \
        _Pragma("omp parallel for schedule(static)")                         \
        for (int zz = 0; zz < N; ++zz) {                                      \
            double s_ = 0;                                                    \
            const R* rp = r + (size_t)zz * N * N;                             \
            for (size_t j = 0; j < (size_t)N * N; ++j) s_ += (double)rp[j] * (double)rp[j]; \
            pa[zz] = s_;                                                      \
        }                                                                     \
        double t_ = 0;                                                        \
        for (int zz = 0; zz < N; ++zz) t_ += pa[zz];                          \
        out = sqrt(t_);                                                       \
    } while (0)

    double rz, mr, alpha = 0, beta = 0, best;
    F(precond)(N, fm, &L, r, z0, rr, zt, z, &rz, &mr);
    double rn = 0;
    RNORM(rn);
    best = mr;
    int it = 0;
#define GO() (it < maxit && rz > 0.0 && mr <= 4.0 * best && \
              (rel_tol > 0.0 ? rn > rel_tol * bnorm : mr > thr))
    int go = GO();
    memcpy(p, z, n3 * sizeof(R));
    while (go) {
        /* k_matvec_dot */
        double pq = 0;
#pragma omp parallel for schedule(static)
        for (int zz = 0; zz < N; ++zz) {
            double s_ = 0;
            for (int y = 0; y < N; ++y)
                for (int x = 0; x < N; ++x) {
                    size_t i = IDX(zz, y, x);
                    R v = fm[i] ? F(kmul)(p, fm, N, zz, y, x) : (R)0;
                    q[i] = v;
                    s_ += (double)v * (double)p[i];
                }
            pa[zz] = s_;
        }
        for (int zz = 0; zz < N; ++zz) pq += pa[zz];
        alpha = pq > 0.0 ? rz / pq : 0.0;
        R af = (R)alpha;
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < n3; ++i) r[i] = r[i] - af * q[i];
        double rz_new;
        F(precond)(N, fm, &L, r, z0, rr, zt, z, &rz_new, &mr);
        RNORM(rn);
        beta = rz > 0.0 ? rz_new / rz : 0.0;
        rz = rz_new;
        it += 1;
        best = mr < best ? mr : best;
        go = GO();
        R bf = (R)beta;
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < n3; ++i) {
            R po = p[i];
            phi[i] = phi[i] + af * po;
            p[i] = z[i] + bf * po;
        }
    }
#undef GO
#undef RNORM
    F(or_apply)(N, fm, phi, vel);
    if (stats) { stats[0] = bnorm; stats[1] = rn; stats[2] = maxu; }
    free(pa); free(pb);
    free(r); free(z); free(p); free(q); free(z0); free(rr); free(zt); free(fm);
    F(levels_free)(&L);
    return it;
}

/* The fused projection (FusedPressure3D.record, pressure_fused.py:398-489, with its default
 * kernel boundaries FUSE_TOP = False / FUSE_MIN_M = inf, :491-499): no CG.  `smooth` weighted
 * Jacobi sweeps on the exact K (k_sweep :41-57; weights: Chebyshev :348-352 or 0.8), then
 * `cycles` x (V-cycle correction phi += T^T V(b - K phi), k_resid_phi / k_fwd_red / V(1,1) /
 * k_inv_transform, + `smooth` sweeps); with tol > 0 (adaptive) more cycles while
 * 0.5 max|b - K phi| >= tol max|u*| and fewer than P_MAXIT cycles have run (k_test :335-347,
 * the tail :463-487); then u -= g(phi) (k_apply_copy :60-72).  A fluid cell sealed in by solids
 * (kdiag = 0) keeps its phi in a sweep (the reference divides 0 / 0 there).  Returns the cycles
 * run; stats[0] = max|u*|, stats[1] = max|b - K phi| at the last test (-1 without the tail). */
static void F(fused_weights)(int smooth, int jacobi, float* w) {
    const int n = smooth < 1 ? 1 : smooth;
    const double lo = 2.0 / 4.0, hi = 2.0;
    for (int i = 0; i < n; ++i)
        w[i] = jacobi ? 0.8f
                      : (float)(1.0 / ((hi + lo) / 2 + (hi - lo) / 2 * cos(3.14159265358979323846 * (2 * i + 1) / (2.0 * n))));
}

static void F(fused_sweep)(int N, const int* fm, const R* vel, const R* src, R* dst, R w, int cold) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                int m = fm[i];
                if (m == 0) { dst[i] = 0; continue; }
                if ((m & 63) == 0) { dst[i] = cold ? (R)0 : src[i]; continue; }
                R b = -F(div2)(vel, fm, N, z, y, x);
                if (cold) dst[i] = w * b / F(kdiag)(m);
                else dst[i] = src[i] + w * (b - F(kmul)(src, fm, N, z, y, x)) / F(kdiag)(m);
            }
}

/* rr = b - K phi (0 in solids); returns max|rr| (as R) */
static R F(fused_resid)(int N, const int* fm, const R* vel, const R* phi, R* rr) {
    R mx = 0;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                R v = 0;
                if (fm[i] != 0) v = -F(div2)(vel, fm, N, z, y, x) - F(kmul)(phi, fm, N, z, y, x);
                rr[i] = v;
                R a = v < 0 ? -v : v;
                mx = a > mx ? a : mx;
            }
    return mx;
}

int F(or_project_fused)(int N, R* data, const R* solid, const float* prm, R* phi, int cycles, int smooth, int jacobi,
                        double* stats) {
    size_t n3 = (size_t)N * N * N;
    int* fm = (int*)calloc(n3, sizeof(int));
    F(or_mask)(N, solid, fm);
    F(levels_t) L;
    F(levels_init)(&L, N);
    F(levels_build)(&L, N, fm);
    R* vel = data + n3;
    R* alt = (R*)calloc(n3, sizeof(R));
    R* rr = (R*)calloc(n3, sizeof(R));
    const int warm = prm[P_WARM] > 0.5f;
    const int ns = smooth < 1 ? 1 : smooth;
    float w[64];
    F(fused_weights)(ns > 64 ? 64 : ns, jacobi, w);
    R* cur = phi;
    int first = 1;
#define OTHER(p) ((p) == phi ? alt : phi)
#define SWEEPS()                                                                        \
    do {                                                                                \
        for (int k_ = 0; k_ < ns && k_ < 64; ++k_) {                                    \
            R* d_ = OTHER(cur);                                                         \
            F(fused_sweep)(N, fm, vel, cur, d_, (R)w[k_], first && !warm);              \
            cur = d_;                                                                   \
            first = 0;                                                                  \
        }                                                                               \
    } while (0)
    SWEEPS();
    for (int c = 0; c < cycles; ++c) {
        F(fused_resid)(N, fm, vel, cur, rr);
        F(fwd_transform)(N, rr, L.r[0]);
        F(mg_vcycle)(&L);
        R* d = OTHER(cur);
        F(inv_transform)(N, L.x[0], fm, cur, d);
        cur = d;
        SWEEPS();
    }
    int it = cycles;
    double rlast = -1.0, umax = 0.0;
    const float tol = prm[P_TOL];
    const int maxit = (int)prm[P_MAXIT];
    if (tol > 0.0f && maxit > cycles) {
        R um = 0;
        for (size_t i = 0; i < n3; ++i) {
            if (fm[i] == 0) continue;
            R ux = vel[i], uy = vel[n3 + i], uz = vel[2 * n3 + i];
            R sp = SQRTR(ux * ux + uy * uy + uz * uz);
            um = sp > um ? sp : um;
        }
        umax = (double)um;
        R rmax = F(fused_resid)(N, fm, vel, cur, rr);
        rlast = (double)rmax;
        int go = it < maxit && (float)0.5f * (float)rmax >= tol * (float)um && um > 0;
        R* loop = cur;
        while (go) {
            F(fwd_transform)(N, rr, L.r[0]);
            F(mg_vcycle)(&L);
            F(inv_transform)(N, L.x[0], fm, loop, loop);
            cur = loop;
            first = 0;
            SWEEPS();
            if (cur != loop) memcpy(loop, cur, n3 * sizeof(R));
            cur = loop;
            rmax = F(fused_resid)(N, fm, vel, cur, rr);
            rlast = (double)rmax;
            it += 1;
            go = it < maxit && (float)0.5f * (float)rmax >= tol * (float)um && um > 0;
        }
    }
#undef SWEEPS
#undef OTHER
    F(or_apply)(N, fm, cur, vel);
    if (cur != phi) memcpy(phi, cur, n3 * sizeof(R));
    if (stats) { stats[0] = umax; stats[1] = rlast; }
    free(alt); free(rr); free(fm);
    F(levels_free)(&L);
    return it;
}

/* ---------------------------------------------------------------- finer scalar grid
 * vfx/sim3d/hires.py.  Fine fields are (3, S, S, S): density, temperature, fuel.  In these
 * functions N is the size of the grid being indexed (IDX), Nc the coarse (velocity) grid's. */

/* f2c constants, from the reference's Python doubles (solver.py:234-239): p_c = a p_f + b */
static inline void F(hires_map)(int Nc, int S, R* a, R* b) {
    *a = (R)((double)Nc / (double)S);
    *b = (R)(0.5 * (double)Nc / (double)S - 0.5);
}

/* advect_sl_fine + advect_mc_fine (hires.py:29-89): MacCormack of the 3 fine channels, RK2
 * traces through the coarse velocity vel (3, Nc^3) in coarse index space, s = dt * inv_dx_c. */
void F(or_hires_advect)(int N, int Nc, const R* src, const R* vel, R s, R* hat, R* out) {
    size_t n3 = (size_t)N * N * N;
    R a, b;
    F(hires_map)(Nc, N, &a, &b);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                R qc[3];
                F(trace)(vel, Nc, a * (R)x + b, a * (R)y + b, a * (R)z + b, -s, qc);
                F(corner_t) c = F(corner)(N, (qc[0] - b) / a, (qc[1] - b) / a, (qc[2] - b) / a);
                for (int ch = 0; ch < 3; ++ch)
                    hat[ch * n3 + IDX(z, y, x)] = F(tri)(src + ch * n3, N, c.z0, c.y0, c.x0, c.fz, c.fy, c.fx);
            }
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                R qc[3], rc[3];
                R px = a * (R)x + b, py = a * (R)y + b, pz = a * (R)z + b;
                F(trace)(vel, Nc, px, py, pz, -s, qc);
                F(trace)(vel, Nc, px, py, pz, s, rc);
                R q[3] = {(qc[0] - b) / a, (qc[1] - b) / a, (qc[2] - b) / a};
                R r[3] = {(rc[0] - b) / a, (rc[1] - b) / a, (rc[2] - b) / a};
                int qx = F(base_idx)(q[0], N), qy = F(base_idx)(q[1], N), qz = F(base_idx)(q[2], N);
                F(corner_t) cr = F(corner)(N, r[0], r[1], r[2]);
                for (int ch = 0; ch < 3; ++ch) {
                    const R* sc = src + ch * n3;
                    R h = hat[ch * n3 + IDX(z, y, x)];
                    R bar = F(tri)(hat + ch * n3, N, cr.z0, cr.y0, cr.x0, cr.fz, cr.fy, cr.fx);
                    R val = h + (R)0.5 * (sc[IDX(z, y, x)] - bar);
                    R lo = sc[IDX(qz, qy, qx)], hi = lo;
                    for (int k = 1; k < 8; ++k) {
                        R v = sc[IDX(qz + (k >> 2), qy + ((k >> 1) & 1), qx + (k & 1))];
                        lo = v < lo ? v : lo;
                        hi = v > hi ? v : hi;
                    }
                    if (val < lo || val > hi) val = h;
                    out[ch * n3 + IDX(z, y, x)] = val;
                }
            }
}

/* sources_fine (hires.py:92-139): the scalar terms of sources at fine-cell centres (fine cell
 * size dxf), fine mass ratio on density, 0 / ambient / 0 in fine solids.  Same extensions as
 * or_sources (ignition ramp, emitter noise with the fine cell index as counter). */
void F(or_hires_sources)(int N, const R* adv, const R* hsolid, const float* prm, R mass_ratio, R dxf, R* out) {
    size_t n3 = (size_t)N * N * N;
    const R amb = (R)prm[P_AMB], dt = (R)prm[P_DT];
    const int n_em = (int)prm[P_N_EM];
    const R ramp = (R)prm[P_IGN_RAMP];
    const R nlo = (R)prm[P_NOISE_LO], nhi = (R)prm[P_NOISE_HI];
    const int noisy = !(prm[P_NOISE_LO] == 1.0f && prm[P_NOISE_HI] == 1.0f);
    const uint32_t seed = (uint32_t)prm[P_SEED], stepno = (uint32_t)prm[P_STEPNO];
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                if (hsolid[i] > (R)0.5) {
                    out[i] = 0; out[n3 + i] = amb; out[2 * n3 + i] = 0;
                    continue;
                }
                R px = ((R)x + (R)0.5) * dxf, py = ((R)y + (R)0.5) * dxf, pz = ((R)z + (R)0.5) * dxf;
                R d = adv[i] * mass_ratio, T = adv[n3 + i], Fu = adv[2 * n3 + i];
                for (int e = 0; e < n_em; ++e) {
                    const float* bb = prm + P_EM0 + e * EM_STRIDE;
                    R rad = (R)bb[3];
                    R ddx = px - (R)bb[0], ddy = py - (R)bb[1], ddz = pz - (R)bb[2];
                    R r2 = (ddx * ddx + ddy * ddy + ddz * ddz) / (rad * rad);
                    if (r2 < (R)1) {
                        R w = ((R)1 - r2) * ((R)1 - r2);
                        R nu = (R)1;
                        if (noisy) {
                            uint32_t u = or_philox_u32(seed, (uint32_t)e, (uint32_t)i, stepno);
                            nu = nlo + (nhi - nlo) * ((R)(u >> 8) * (R)(1.0 / 16777216.0));
                        }
                        R drate = (R)bb[4], frate = (R)bb[6], te = (R)bb[5];
                        if (noisy) { drate = drate * nu; frate = frate * nu; }
                        d += drate * w * dt;
                        Fu += frate * w * dt;
                        if (te > (R)0) {
                            R tex = noisy ? (te - amb) * nu : (te - amb);
                            R tt = amb + tex * w;
                            T = T > tt ? T : tt;
                        }
                    }
                }
                R ign = F(clampr)((T - (R)prm[P_IGN]) / ramp, (R)0, (R)1);
                R burned = Fu * ((R)1 - EXPR(-(R)prm[P_BURN] * dt)) * ign;
                Fu -= burned;
                R tmax = (R)prm[P_TMAX];
                R room = F(clampr)((tmax - T) / (tmax - amb), (R)0, (R)1);
                T += (R)prm[P_BHEAT] * burned * room;
                d += (R)prm[P_SMOKE] * burned;
                T = amb + (T - amb) * EXPR(-(R)prm[P_COOL] * dt);
                d = d * EXPR(-(R)prm[P_DDISS] * dt);
                out[i] = d > (R)0 ? d : (R)0;
                out[n3 + i] = T < tmax ? T : tmax;
                out[2 * n3 + i] = Fu > (R)0 ? Fu : (R)0;
            }
}

/* diffuse_fine (hires.py:142-170): T and F one explicit substep on the fine grid, a missing or
 * solid neighbour contributing the cell's own value (nbv); density copied. */
void F(or_hires_diffuse)(int N, const R* src, const R* hsolid, const float* prm, R nsub, R dxf, R* dst) {
    size_t n3 = (size_t)N * N * N;
    R lam = (R)prm[P_KAPPA] * (R)prm[P_DT] / (nsub * dxf * dxf);
    lam = lam < (R)LAM_CAP ? lam : (R)LAM_CAP;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                dst[i] = src[i];
                for (int c = 1; c <= 2; ++c) {
                    const R* f = src + c * n3;
                    R v = f[i];
                    if (hsolid[i] > (R)0.5 || lam <= (R)0) {
                        dst[c * n3 + i] = v;
                    } else {
#define NB(zz, yy, xx) (F(solid_at)(hsolid, N, zz, yy, xx) ? v : f[IDX(zz, yy, xx)])
                        R sum = NB(z, y, x - 1) + NB(z, y, x + 1) + NB(z, y - 1, x) + NB(z, y + 1, x)
                                + NB(z - 1, y, x) + NB(z + 1, y, x);
#undef NB
                        dst[c * n3 + i] = v + lam * (sum - (R)6 * v);
                    }
                }
            }
}

/* ovl (hires.py:173-176) */
static inline R F(ovl)(int i, int j, R ratio) {
    R hi_a = (R)(j + 1), hi_b = (R)(i + 1) * ratio, lo_a = (R)j, lo_b = (R)i * ratio;
    R len = (hi_a < hi_b ? hi_a : hi_b) - (lo_a > lo_b ? lo_a : lo_b);
    return len > (R)0 ? len : (R)0;
}

/* restrict_scalars (hires.py:179-219): coarse density / temperature / fuel (channels 0, 4, 5 of
 * the (6, N^3) `data`) = overlap-weighted mean of the fine cells, ratio = S / N, span =
 * ceil(S / N) + 1; 0 / ambient / 0 in coarse solids. */
void F(or_hires_restrict)(int N, int S, const R* fine, const R* csolid, const float* prm, R* data) {
    size_t n3 = (size_t)N * N * N, fn3 = (size_t)S * S * S;
    const R ratio = (R)((double)S / (double)N);
    const int span = (S + N - 1) / N + 1;
    const R amb = (R)prm[P_AMB];
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                size_t i = IDX(z, y, x);
                if (csolid[i] > (R)0.5) {
                    data[i] = 0; data[4 * n3 + i] = amb; data[5 * n3 + i] = 0;
                    continue;
                }
                int z0 = (int)FLOORR((R)z * ratio), y0 = (int)FLOORR((R)y * ratio), x0 = (int)FLOORR((R)x * ratio);
                R sw = 0, sd = 0, st = 0, sf = 0;
                for (int k = 0; k < span; ++k) {
                    int zz = z0 + k;
                    R wz = zz < S ? F(ovl)(z, zz, ratio) : (R)0;
                    for (int j = 0; j < span; ++j) {
                        int yy = y0 + j;
                        R wy = yy < S ? F(ovl)(y, yy, ratio) : (R)0;
                        for (int ii = 0; ii < span; ++ii) {
                            int xx = x0 + ii;
                            R wx = xx < S ? F(ovl)(x, xx, ratio) : (R)0;
                            R w = wz * wy * wx;
                            if (w > (R)0) {
                                size_t o = ((size_t)zz * S + yy) * S + xx;
                                sw += w;
                                sd += w * fine[o];
                                st += w * fine[fn3 + o];
                                sf += w * fine[2 * fn3 + o];
                            }
                        }
                    }
                }
                if (sw > (R)0) {
                    data[i] = sd / sw; data[4 * n3 + i] = st / sw; data[5 * n3 + i] = sf / sw;
                } else {
                    data[i] = 0; data[4 * n3 + i] = amb; data[5 * n3 + i] = 0;
                }
            }
}

/* _record_hires (solver.py:241-259): the whole fine stage in place on hi (3, S^3), by the coarse
 * velocity vel (3, Nc^3) at the start of the step.  Returns the fine mass ratio. */
double F(or_hires_stage)(int Nc, int S, R* hi, const R* hsolid, const R* vel, const float* prm, R inv_dx) {
    size_t f3 = (size_t)S * S * S;
    R* hat = (R*)malloc(3 * f3 * sizeof(R));
    R* adv = (R*)malloc(3 * f3 * sizeof(R));
    R s = (R)prm[P_DT] * inv_dx;
    R dxf = (R)(1.0 / (double)S);
    int nd = S > 256 ? (S + 255) / 256 : 1;
    F(or_hires_advect)(S, Nc, hi, vel, s, hat, adv);
    double ratio = F(or_mass_ratio)(S, hi, adv, hsolid);
    R* bufs[2];
    bufs[0] = nd % 2 == 0 ? hi : hat;
    bufs[1] = nd % 2 == 0 ? hat : hi;
    F(or_hires_sources)(S, adv, hsolid, prm, (R)ratio, dxf, bufs[0]);
    for (int k = 0; k < nd; ++k) F(or_hires_diffuse)(S, bufs[k % 2], hsolid, prm, (R)nd, dxf, bufs[(k + 1) % 2]);
    free(hat);
    free(adv);
    return ratio;
}

/* ---------------------------------------------------------------- stages of Solver3D.step
 * _record_advect solver.py:422-430: MacCormack of all 6 channels by the state's own velocity,
 * then the mass ratio.  data -> adv (6ch); returns the ratio. */
double F(or_stage_advect)(int N, const R* data, const R* solid, const float* prm, R inv_dx, R* adv) {
    size_t n3 = (size_t)N * N * N;
    R* hat = (R*)malloc(6 * n3 * sizeof(R));
    R s = (R)prm[P_DT] * inv_dx;
    F(or_advect)(N, 6, data, data + n3, s, hat, adv);
    free(hat);
    return F(or_mass_ratio)(N, data, adv, solid);
}

/* _record_forces solver.py:432-455: sources, n_diff diffusion substeps, curl, smoothed
 * vorticity, body forces, smoothed force applied.  adv (6ch) -> data (6ch).  With a fine grid
 * (hi = (3, S^3) non-NULL) the fine scalars are restricted into data after the diffusion. */
void F(or_stage_forces_hires)(int N, const R* adv, const R* solid, const float* prm, R mass_ratio, R inv_dx, int S,
                              const R* hi, R* data) {
    size_t n3 = (size_t)N * N * N;
    int nd = N > 256 ? (N + 255) / 256 : 1;
    R* hat = (R*)malloc(6 * n3 * sizeof(R));
    R* bufs[2];
    bufs[0] = nd % 2 == 0 ? data : hat;
    bufs[1] = nd % 2 == 0 ? hat : data;
    F(or_sources)(N, adv, solid, prm, mass_ratio, bufs[0]);
    for (int k = 0; k < nd; ++k) F(or_diffuse)(N, bufs[k % 2], solid, prm, (R)nd, bufs[(k + 1) % 2]);
    /* with scalar_res, the coarse scalars the forces read = the fine ones restricted (solver.py:443-445) */
    if (hi) F(or_hires_restrict)(N, S, hi, solid, prm, data);
    R* omega = (R*)malloc(3 * n3 * sizeof(R));
    R* omega_s = (R*)malloc(3 * n3 * sizeof(R));
    R* force = (R*)malloc(3 * n3 * sizeof(R));
    R* ta = (R*)malloc(4 * n3 * sizeof(R));
    R* tb = (R*)malloc(4 * n3 * sizeof(R));
    R inv2dx = (R)(float)(0.5 * (double)inv_dx);
    F(or_curl)(N, data, solid, inv2dx, omega);
    F(or_smooth3)(N, omega, solid, ta, tb, 0, (R)0, omega_s);
    F(or_body_forces)(N, data, solid, omega_s, prm, force);
    F(or_smooth3)(N, force, solid, ta, tb, 1, (R)prm[P_DT], data);
    free(hat); free(omega); free(omega_s); free(force); free(ta); free(tb);
}

void F(or_stage_forces)(int N, const R* adv, const R* solid, const float* prm, R mass_ratio, R inv_dx, R* data) {
    F(or_stage_forces_hires)(N, adv, solid, prm, mass_ratio, inv_dx, 0, NULL, data);
}

#undef F
#undef CAT
#undef CAT2
