// Sources, heat/fuel diffusion, vorticity confinement, buoyancy, wind and the smoothed force.
//
// Restates vfx/sim3d/kernels.py:181-457 (sources, diffuse_heat_fuel, curl, body_forces, the
// separable fluid-renormalised [1,2,1]^3 smoothing sep_x / sep_y / sep_z_store / sep_z_apply,
// here one kernel for the three passes) in the launch order of solver.py:432-455.  "Solid" is read from the uint8 face mask
// (mask == 0 <=> solid > 0.5, pressure.py:115-143); cells outside the box count as solid.
#include "sim3d_kernels.hpp"
#include "sources_ops.hpp"

namespace vfx {

// sources (kernels.py:182-250).  Channels 0..3 (density, velocity) go to `out`, temperature and fuel
// to `out_tf` (channels 4, 5 of that buffer): the diffusion that follows changes only those two,
// so density and velocity are written to their final place here instead of being copied through.
// VEL_ONLY (with the finer scalar grid): only the velocity channels 1..3 are written; density,
// temperature and fuel come from the fine grid's restriction, which overwrites them before
// anything reads them (solver.py:443-445).
template <bool VEL_ONLY>
__global__ void __launch_bounds__(256) k_sources(int N, int zc, const real* __restrict__ adv,
                                                 const uint8_t* __restrict__ fm, const float* __restrict__ prm,
                                                 const DevState* __restrict__ st, real* __restrict__ out,
                                                 real* __restrict__ out_tf) {
    const size_t n3 = (size_t)N * N * N;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= N || y >= N) return;
    const SrcConst k(prm);
    const real dx = (real)prm[P_DX];
    const real ratio = VEL_ONLY ? RL(1.0) : (real)st->s[S_MASS_RATIO];
    const int z1 = min((int)(blockIdx.z + 1) * zc, N);
    for (int z = blockIdx.z * zc; z < z1; ++z) {
        const size_t i = idx3(N, z, y, x);
        if (fm[i] == 0) {
            out[n3 + i] = RL(0.0); out[2 * n3 + i] = RL(0.0); out[3 * n3 + i] = RL(0.0);
            if (!VEL_ONLY) {
                out[i] = RL(0.0); out_tf[4 * n3 + i] = k.amb; out_tf[5 * n3 + i] = RL(0.0);
            }
            continue;
        }
        const real px = ((real)x + RL(0.5)) * dx, py = ((real)y + RL(0.5)) * dx, pz = ((real)z + RL(0.5)) * dx;
        real d = RL(0.0), T = RL(0.0), Fu = RL(0.0);
        if (!VEL_ONLY) {
            d = adv[i] * ratio;
            T = adv[4 * n3 + i];
            Fu = adv[5 * n3 + i];
        }
        real vx = adv[n3 + i], vy = adv[2 * n3 + i], vz = adv[3 * n3 + i];
        cell_sources<!VEL_ONLY, true>(prm, k, px, py, pz, (uint32_t)i, d, T, Fu, vx, vy, vz);
        out[n3 + i] = vx;
        out[2 * n3 + i] = vy;
        out[3 * n3 + i] = vz;
        if (!VEL_ONLY) {
            out[i] = d;
            out_tf[4 * n3 + i] = T;
            out_tf[5 * n3 + i] = Fu;
        }
    }
}

// diffuse_heat_fuel (kernels.py:268-286): T and F (channels 4, 5) one explicit substep with zero
// flux at walls and solids (a missing or solid neighbour contributes the cell's own value).
// Channels 0..3, which the reference copies through, are already in place (k_sources).
__global__ void __launch_bounds__(256) k_diffuse(int N, int zc, const real* __restrict__ src,
                                                 const uint8_t* __restrict__ fm, const float* __restrict__ prm,
                                                 real nsub, real* __restrict__ dst) {
    const int n3 = N * N * N;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= N || y >= N) return;
    const real dx = (real)prm[P_DX];
    real lam = (real)prm[P_KAPPA] * (real)prm[P_DT] / (nsub * dx * dx);
    lam = lam < LAM_CAP ? lam : LAM_CAP;
    const int z1 = min((int)(blockIdx.z + 1) * zc, N);
    for (int z = blockIdx.z * zc; z < z1; ++z) {
        const int i = (z * N + y) * N + x;
        const Nb6 n = nb6(N, x, y, z, i);
        const int m0 = fm[i], mxm = fm[n.xm], mxp = fm[n.xp], mym = fm[n.ym], myp = fm[n.yp], mzm = fm[n.zm],
                  mzp = fm[n.zp];
        const bool sxm = !n.oxm || mxm == 0, sxp = !n.oxp || mxp == 0, sym = !n.oym || mym == 0;
        const bool syp = !n.oyp || myp == 0, szm = !n.ozm || mzm == 0, szp = !n.ozp || mzp == 0;
        real out[2];
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
            const real* f = src + (4 + cc) * n3;
            const real v = f[i], a0 = f[n.xm], a1 = f[n.xp], a2 = f[n.ym], a3 = f[n.yp], a4 = f[n.zm], a5 = f[n.zp];
            const real s = (sxm ? v : a0) + (sxp ? v : a1) + (sym ? v : a2) + (syp ? v : a3) + (szm ? v : a4) +
                            (szp ? v : a5);
            out[cc] = (m0 == 0 || lam <= RL(0.0)) ? v : v + lam * (s - RL(6.0) * v);
        }
        dst[4 * n3 + i] = out[0];
        dst[5 * n3 + i] = out[1];
    }
}

// curl at cell (x, y, z) (kernels.py:297-310): walls and solids count as zero velocity.
__device__ __forceinline__ void curl_at(const real* __restrict__ data, const uint8_t* __restrict__ fm, int N, int x,
                                        int y, int z, real inv2dx, real (&w)[3]) {
    const int n3 = N * N * N;
    const real* vx = data + n3;
    const real* vy = data + 2 * n3;
    const real* vz = data + 3 * n3;
    const int i = (z * N + y) * N + x;
    const Nb6 n = nb6(N, x, y, z, i);
    const int mxm = fm[n.xm], mxp = fm[n.xp], mym = fm[n.ym], myp = fm[n.yp], mzm = fm[n.zm], mzp = fm[n.zp];
    const bool sxm = !n.oxm || mxm == 0, sxp = !n.oxp || mxp == 0, sym = !n.oym || mym == 0;
    const bool syp = !n.oyp || myp == 0, szm = !n.ozm || mzm == 0, szp = !n.ozp || mzp == 0;
    const real vz_yp = vz[n.yp], vz_ym = vz[n.ym], vy_zp = vy[n.zp], vy_zm = vy[n.zm];
    const real vx_zp = vx[n.zp], vx_zm = vx[n.zm], vz_xp = vz[n.xp], vz_xm = vz[n.xm];
    const real vy_xp = vy[n.xp], vy_xm = vy[n.xm], vx_yp = vx[n.yp], vx_ym = vx[n.ym];
    const real dwdy = (syp ? RL(0.0) : vz_yp) - (sym ? RL(0.0) : vz_ym);
    const real dvdz = (szp ? RL(0.0) : vy_zp) - (szm ? RL(0.0) : vy_zm);
    const real dudz = (szp ? RL(0.0) : vx_zp) - (szm ? RL(0.0) : vx_zm);
    const real dwdx = (sxp ? RL(0.0) : vz_xp) - (sxm ? RL(0.0) : vz_xm);
    const real dvdx = (sxp ? RL(0.0) : vy_xp) - (sxm ? RL(0.0) : vy_xm);
    const real dudy = (syp ? RL(0.0) : vx_yp) - (sym ? RL(0.0) : vx_ym);
    w[0] = (dwdy - dvdz) * inv2dx;
    w[1] = (dudz - dwdx) * inv2dx;
    w[2] = (dvdx - dudy) * inv2dx;
}


// the parameters body_force_at reads
struct BodyPrm {
    real amb, alpha, beta, eps, cfloor, dxp, wdx, wdy, wdz, wmark;
    bool wind;
};

__device__ __forceinline__ BodyPrm body_prm(const float* __restrict__ prm) {
    BodyPrm b;
    b.amb = (real)prm[P_AMB]; b.alpha = (real)prm[P_ALPHA]; b.beta = (real)prm[P_BETA]; b.eps = (real)prm[P_EPS];
    b.cfloor = (real)prm[P_CFLOOR]; b.dxp = (real)prm[P_DX];
    b.wdx = (real)prm[P_WIND_X]; b.wdy = (real)prm[P_WIND_Y]; b.wdz = (real)prm[P_WIND_Z]; b.wmark = (real)prm[P_WMARK];
    b.wind = b.wdx * b.wdx + b.wdy * b.wdy + b.wdz * b.wdz > RL(0.0);
    return b;
}

// body_forces (kernels.py:346-387) at fluid cell (x, y, z): buoyancy + vorticity confinement +
// wind marker force from the state f and the smoothed vorticity om (4 channels; channel 3 holds
// |omega_s| = sqrt(a * a + b * b + c * c) of the values k_smooth3<0> wrote, the expression of
// kernels.py:342-343).  Indices within a channel are 32-bit (N <= 512), channel bases are
// pointers.  The gradient's divisor max(xr - xl, 1) is 1 or 2, and a quotient by 1 or 2 equals
// the product by 1 or 0.5 exactly, so it is computed as that product.
__device__ __forceinline__ void body_force_at(const real* __restrict__ f, const real* __restrict__ om,
                                              const BodyPrm& b, int N, int x, int y, int z, real (&F)[3]) {
    const size_t n3 = (size_t)N * N * N;
    const int sz = N * N;
    const int i = (z * N + y) * N + x;
    const real* __restrict__ wm = om + 3 * n3;
    const real d = f[i], T = (f + 4 * n3)[i];
    real fx = RL(0.0), fy = -b.alpha * d + b.beta * (T - b.amb), fz = RL(0.0);
    if (b.eps != RL(0.0)) {
        const int xr = min(x + 1, N - 1), xl = max(x - 1, 0);
        const int yt = min(y + 1, N - 1), yb = max(y - 1, 0);
        const int zf = min(z + 1, N - 1), zb = max(z - 1, 0);
        const int rowb = z * sz + y * N;
        const real hx = xr - xl == 2 ? RL(0.5) : RL(1.0), hy = yt - yb == 2 ? RL(0.5) : RL(1.0);
        const real hz = zf - zb == 2 ? RL(0.5) : RL(1.0);
        const real gx = (wm[rowb + xr] - wm[rowb + xl]) * hx;
        const real gy = (wm[z * sz + yt * N + x] - wm[z * sz + yb * N + x]) * hy;
        const real gz = (wm[zf * sz + y * N + x] - wm[zb * sz + y * N + x]) * hz;
        const real wx = om[i], wy = (om + n3)[i], wz = (om + 2 * n3)[i];
        const real fl = b.cfloor * sqrt(wx * wx + wy * wy + wz * wz);
        const real mag = sqrt(gx * gx + gy * gy + gz * gz + fl * fl) + RL(1.0e-20);
        const real k = b.eps * b.dxp / mag;
        fx += k * (gy * wz - gz * wy);
        fy += k * (gz * wx - gx * wz);
        fz += k * (gx * wy - gy * wx);
    }
    if (b.wind) {
        const real mk = clampf_(d + (f + 5 * n3)[i] + (T - b.amb) / b.wmark, RL(0.0), RL(1.0));
        fx += b.wdx * mk;
        fy += b.wdy * mk;
        fz += b.wdz * mk;
    }
    F[0] = fx;
    F[1] = fy;
    F[2] = fz;
}

// The fluid-renormalised [1,2,1]^3 smoothing B(w a) / B(w) of three channels in ONE pass: sep_x,
// sep_y and sep_z_store / sep_z_apply (kernels.py:392-457) composed, with the field a computed
// where it is staged.  A 256-thread workgroup owns a 64 x 8 column of cells (two rows per thread)
// and marches up z through zt planes.  Each input plane is staged in LDS with a one-cell x / y
// halo, the x pass runs on the 10 staged rows, the y pass per thread, and the y-pass values of
// three consecutive planes stay in registers for the z pass.  Every pass keeps the reference's
// operation order, so the result is the separate kernels' result bit for bit, while no
// intermediate field is written.
// MODE 0: a = the curl of the state's velocity (curl_at, kernels.py:297-310); out (4 channels) =
// the smoothed vorticity and its magnitude sqrt(a^2 + b^2 + c^2) (body_forces' |omega_s| at every
// cell, evaluated once).  MODE 1: a = the body force (body_force_at, kernels.py:346-387, from the
// state and the smoothed vorticity om); fluid cells of the state get velocity += dt * the
// smoothed force.  Tiles are numbered through xcd_block, so the tiles that share halo rows run on
// one XCD.
// zt = SM_ZT = 8 planes per workgroup: 0.91 ms for the forces stage at 256^3 against 0.95 ms with
// 16 (more workgroups in flight outweigh the extra halo planes; 4 is no faster).  Grids that would
// have fewer than 1024 workgroups use fewer planes (smooth3_planes).
constexpr int SM_TX = 64, SM_TY = 8, SM_ZT = 8, SM_RY = SM_TY / 4;   // blockDim (64, 4)
constexpr int SM_HX = SM_TX + 2, SM_HY = SM_TY + 2;

template <int MODE>
__global__ void __launch_bounds__(256) k_smooth3(int N, const real* __restrict__ f, const uint8_t* __restrict__ fm,
                                                 const float* __restrict__ prm, real* __restrict__ out, int zt,
                                                 real inv2dx, const real* __restrict__ om) {
    __shared__ real sa[3][SM_HY][SM_HX];
    __shared__ real sw[SM_HY][SM_HX];
    __shared__ real sx[4][SM_HY][SM_TX];
    const int n3 = N * N * N;
    const int gx = (N + SM_TX - 1) / SM_TX, gy = (N + SM_TY - 1) / SM_TY;
    const int nb = (int)gridDim.x;
    const int t = xcd_block((int)blockIdx.x, nb);
    const int bx = t % gx, by = (t / gx) % gy, bz = t / (gx * gy);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * SM_TX + tx;
    const int x0 = bx * SM_TX, y0 = by * SM_TY, z0 = bz * zt;
    const int z1 = min(z0 + zt, N);
    const int x = x0 + tx;
    const real dt = MODE == 1 ? (real)prm[P_DT] : RL(0.0);
    BodyPrm bp;
    if (MODE == 1) bp = body_prm(prm);
    real ym[SM_RY][4], yc[SM_RY][4], yp[SM_RY][4];
#pragma unroll
    for (int j = 0; j < SM_RY; ++j)
#pragma unroll
        for (int c = 0; c < 4; ++c) ym[j][c] = yc[j][c] = yp[j][c] = RL(0.0);
    // the staged plane: SM_ST slots per thread, loaded one plane ahead into registers so the loads
    // of plane zi + 1 are in flight while plane zi is computed
    constexpr int SM_ST = (SM_HY * SM_HX + 255) / 256;
    real pv[SM_ST][3];
    real pw[SM_ST];
    auto fetch = [&](int zi) {
        const bool zin = zi >= 0 && zi < N;
#pragma unroll
        for (int j = 0; j < SM_ST; ++j) {
            const int k = tid + 256 * j;
            const int r = k / SM_HX, c = k - r * SM_HX;
            const int py = y0 - 1 + r, px = x0 - 1 + c;
            const bool in = k < SM_HY * SM_HX && zin && px >= 0 && px < N && py >= 0 && py < N;
            // outside the box: cell (0, 0, 0) is computed and weighted 0 (every load is issued)
            const int cx = in ? px : 0, cy = in ? py : 0, cz = in ? zi : 0;
            const int gi = (cz * N + cy) * N + cx;
            const int m = fm[gi];
            real v[3];
            if (MODE == 0) {
                curl_at(f, fm, N, cx, cy, cz, inv2dx, v);
            } else {
                body_force_at(f, om, bp, N, cx, cy, cz, v);   // (a solid cell's force is weighted 0)
            }
            pv[j][0] = v[0];
            pv[j][1] = v[1];
            pv[j][2] = v[2];
            pw[j] = (in && m != 0) ? RL(1.0) : RL(0.0);
        }
    };
    fetch(z0 - 1);
    for (int zi = z0 - 1; zi <= z1; ++zi) {
        // stage plane zi: 3 channels and the fluid weight (0 outside the box)
#pragma unroll
        for (int j = 0; j < SM_ST; ++j) {
            const int k = tid + 256 * j;
            if (k < SM_HY * SM_HX) {
                const int r = k / SM_HX, c = k - r * SM_HX;
                sa[0][r][c] = pv[j][0];
                sa[1][r][c] = pv[j][1];
                sa[2][r][c] = pv[j][2];
                sw[r][c] = pw[j];
            }
        }
        if (zi < z1) fetch(zi + 1);
        __syncthreads();
        // x pass (sep_x): s = sum_k q_k w a over x-1, x, x+1 in that order
        for (int k = tid; k < SM_HY * SM_TX; k += 256) {
            const int r = k / SM_TX, c = k - r * SM_TX;
            real s0 = RL(0.0), s1 = RL(0.0), s2 = RL(0.0), ws = RL(0.0);
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                const bool fl = sw[r][c + kk] != RL(0.0);
                const real q = kk == 1 ? RL(2.0) : RL(1.0);
                s0 = s0 + (fl ? q * sa[0][r][c + kk] : RL(0.0));
                s1 = s1 + (fl ? q * sa[1][r][c + kk] : RL(0.0));
                s2 = s2 + (fl ? q * sa[2][r][c + kk] : RL(0.0));
                ws = ws + (fl ? q : RL(0.0));
            }
            sx[0][r][c] = s0;
            sx[1][r][c] = s1;
            sx[2][r][c] = s2;
            sx[3][r][c] = ws;
        }
        __syncthreads();
        // y pass (sep_y) of this thread's rows at plane zi, into the register ring
#pragma unroll
        for (int j = 0; j < SM_RY; ++j) {
            const int ly = ty + 4 * j, y = y0 + ly;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                real v = RL(2.0) * sx[c][ly + 1][tx];
                v = v + (y > 0 ? sx[c][ly][tx] : RL(0.0));
                v = v + (y + 1 < N ? sx[c][ly + 2][tx] : RL(0.0));
                ym[j][c] = yc[j][c];
                yc[j][c] = yp[j][c];
                yp[j][c] = v;
            }
        }
        // z pass (sep_z_ratio) of plane zo = zi - 1 from planes zo - 1, zo, zo + 1
        const int zo = zi - 1;
        if (zo >= z0 && x < N) {
#pragma unroll
            for (int j = 0; j < SM_RY; ++j) {
                const int y = y0 + ty + 4 * j;
                if (y >= N) continue;
                const bool inb0 = zo > 0, inb2 = zo + 1 < N;
                real sv[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    real u = RL(0.0);
                    u = u + (inb0 ? ym[j][c] : RL(0.0));
                    u = u + RL(2.0) * yc[j][c];
                    u = u + (inb2 ? yp[j][c] : RL(0.0));
                    sv[c] = u;
                }
                real r0 = RL(0.0), r1 = RL(0.0), r2 = RL(0.0);
                if (__all(sv[3] == RL(64.0))) {
                    // every cell of the 3x3x3 neighbourhood is fluid: the weight sum is 4^3, and a
                    // quotient by 64 equals the product by 1/64 exactly
                    r0 = sv[0] * RL(0.015625); r1 = sv[1] * RL(0.015625); r2 = sv[2] * RL(0.015625);
                } else if (sv[3] > RL(0.0)) {
                    r0 = sv[0] / sv[3]; r1 = sv[1] / sv[3]; r2 = sv[2] / sv[3];
                }
                const int i = (zo * N + y) * N + x;
                if (MODE == 0) {
                    out[i] = r0; out[n3 + i] = r1; out[2 * n3 + i] = r2;
                    out[3 * n3 + i] = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
                } else if (fm[i] != 0) {
                    out[n3 + i] = out[n3 + i] + dt * r0;
                    out[2 * n3 + i] = out[2 * n3 + i] + dt * r1;
                    out[3 * n3 + i] = out[3 * n3 + i] + dt * r2;
                }
            }
        }
    }
}

// planes per workgroup: SM_ZT, halved (down to 1) while the launch has fewer than 1024 workgroups
static int smooth3_planes(int N) {
    const int cols = ((N + SM_TX - 1) / SM_TX) * ((N + SM_TY - 1) / SM_TY);
    int zt = SM_ZT;
    while (zt > 1 && cols * ((N + zt - 1) / zt) < 1024) zt /= 2;
    return zt;
}

static int smooth3_blocks(int N) {
    const int zt = smooth3_planes(N);
    return ((N + SM_TX - 1) / SM_TX) * ((N + SM_TY - 1) / SM_TY) * ((N + zt - 1) / zt);
}

void launch_sources_diffuse(const Geom& g, const real* adv, const uint8_t* fm, const float* prm,
                            const DevState* st, real* hat, real* data, hipStream_t s) {
    // n_diff = max(1, ceil(N / 256)) substeps, ping-pong ending in data (solver.py:107-109,436-442)
    const int nd = g.N > 256 ? (g.N + 255) / 256 : 1;
    real* b0 = nd % 2 == 0 ? data : hat;
    real* b1 = nd % 2 == 0 ? hat : data;
    // the sources keep one cell per thread whatever the other cell kernels use (151 against 159 us
    // with two at 256^3; the diffusion gains from two: 80 against 89 us)
    const Geom g1 = g.zc == 1 ? g : make_geom(g.N, 1);
    hipLaunchKernelGGL(k_sources<false>, g1.grid, g1.block, 0, s, g1.N, g1.zc, adv, fm, prm, st, data, b0);
    for (int k = 0; k < nd; ++k) {
        real* a = k % 2 == 0 ? b0 : b1;
        real* b = k % 2 == 0 ? b1 : b0;
        hipLaunchKernelGGL(k_diffuse, g.grid, g.block, 0, s, g.N, g.zc, a, fm, prm, (real)nd, b);
    }
}

void launch_sources_vel(const Geom& g, const real* adv, const uint8_t* fm, const float* prm, real* data,
                        hipStream_t s) {
    const Geom g1 = g.zc == 1 ? g : make_geom(g.N, 1);
    hipLaunchKernelGGL(k_sources<true>, g1.grid, g1.block, 0, s, g1.N, g1.zc, adv, fm, prm, nullptr, data, nullptr);
}

void launch_sources_pass(int pass, const Geom& g, const real* src, const uint8_t* fm, const float* prm,
                         const DevState* st, real* dst, hipStream_t s) {
    const Geom g1 = g.zc == 1 ? g : make_geom(g.N, 1);
    if (pass == 0)
        hipLaunchKernelGGL(k_sources<false>, g1.grid, g1.block, 0, s, g1.N, g1.zc, src, fm, prm, st, dst, dst);
    else
        hipLaunchKernelGGL(k_diffuse, g.grid, g.block, 0, s, g.N, g.zc, src, fm, prm, RL(1.0), dst);
}

// the vorticity-confinement chain (kernels.py:297-457): curl + smoothing of omega into omega_s,
// then body forces + smoothing + apply to the velocity; omega and the force are never stored
void launch_vort_forces(const Geom& g, real* data, const uint8_t* fm, const float* prm, real inv2dx, real* omega_s,
                        hipStream_t s) {
    const dim3 sb(SM_TX, 4), sg(smooth3_blocks(g.N));
    hipLaunchKernelGGL(k_smooth3<0>, sg, sb, 0, s, g.N, data, fm, prm, omega_s, smooth3_planes(g.N), inv2dx,
                       nullptr);
    hipLaunchKernelGGL(k_smooth3<1>, sg, sb, 0, s, g.N, data, fm, prm, data, smooth3_planes(g.N), inv2dx, omega_s);
}

}  // namespace vfx
