// Fine-grid kernels of the MG-preconditioned CG projection (pressure.py:595-775, :156-175) and
// the host-side launch sequence of one CG iteration.
//
// The stencil kernels are LDS-tiled: a 256-thread workgroup stages a 32x8x8 fine tile with a
// 2-cell halo (the reach of K = d d^T) and evaluates K from LDS.  Each fused kernel performs the
// reference's arithmetic, value by value, in the reference's order:
//   k_upd_matvec    = k_upd_p of iteration k-1 + k_matvec_dot of iteration k
//                     (phi += alpha p_old; p = z + beta p_old; q = K p; p.q)
//   k_pre_resid_fwd = k_pre + k_resid + k_fwd_red
//                     (r -= alpha q; z0 = omega D^-1 r; rr = r - K z0; level-1 transform; red sweep)
//   k_inv_post      = k_inv_transform + k_post_rz
//                     (zt = z0 + T^T x1; z = zt + omega D^-1 (r - K zt); r.z, max|r|, r.r)
// r and p live in two buffers each (iteration k reads one and writes the other), so a tile never
// reads a neighbour tile's already-updated values.  Every kernel of CG iteration k returns at
// once when the stop test fired at an earlier iteration (DevState::stop_at < k).
#include "pressure_ops.hpp"

namespace vfx {

#define CELL_LOOP_BEGIN                                                 \
    const int x = blockIdx.x * blockDim.x + threadIdx.x;                \
    const int y = blockIdx.y * blockDim.y + threadIdx.y;                \
    const bool live = x < N && y < N;                                   \
    const int z1 = min((int)(blockIdx.z + 1) * zc, N);                  \
    if (live)                                                           \
        for (int cz = blockIdx.z * zc; cz < z1; ++cz) {                 \
            const int i = (cz * N + y) * N + x;
#define CELL_LOOP_END }

constexpr int HN_IT = (HN + TILE_THREADS - 1) / TILE_THREADS;          // halo cells per thread
constexpr int OWN_IT = TX * TY * TZ / TILE_THREADS;                    // own cells per thread
constexpr int HBN8_IT = (HBN * 8 + TILE_THREADS - 1) / TILE_THREADS;  // level-1 halo values per thread

__device__ __forceinline__ int tile_linear() { return (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)); }

// halo cell h of the current tile -> global cell index, or -1 outside the box
__device__ __forceinline__ int halo_global(int N, int h) {
    const int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
    const int gx = (int)blockIdx.x * TX - 2 + hx, gy = (int)blockIdx.y * TY - 2 + hy, gz = (int)blockIdx.z * TZ - 2 + hz;
    if (gx < 0 || gy < 0 || gz < 0 || gx >= N || gy >= N || gz >= N) return -1;
    return (gz * N + gy) * N + gx;
}

// own cell j (0..TX*TY*TZ-1, x fastest) of the current tile -> (global index or -1, halo index)
__device__ __forceinline__ int own_global(int N, int j, int& h) {
    const int lx = j % TX, ly = (j / TX) % TY, lz = j / (TX * TY);
    h = tile_h(lx, ly, lz);
    const int gx = (int)blockIdx.x * TX + lx, gy = (int)blockIdx.y * TY + ly, gz = (int)blockIdx.z * TZ + lz;
    if (gx >= N || gy >= N || gz >= N) return -1;
    return (gz * N + gy) * N + gx;
}

// k_rhs (:602-627): r = -d(u*) - K phi0 (phi = 0 without warm start); block max |u*| over fluid
// and block sum of b^2 (b = -d(u*)) for the relative-residual test.
__global__ void __launch_bounds__(256) k_rhs(int N, int zc, const float* __restrict__ vel,
                                             const uint8_t* __restrict__ fm, float* __restrict__ r,
                                             float* __restrict__ phi, const float* __restrict__ prm,
                                             double* __restrict__ part, int nparts) {
    __shared__ double sh[8];
    const int n3 = N * N * N;
    const bool warm = prm[P_WARM] > 0.5f;
    double mx = 0.0, bb = 0.0;
    CELL_LOOP_BEGIN
        if (!warm) phi[i] = 0.0f;
        const int m = fm[i];
        if (m == 0) { r[i] = 0.0f; continue; }
        const float b = -div2_xyz(vel, n3, m, N, x, y, cz, i);
        float rv = b;
        if (warm) rv -= kmul_xyz(phi, fm, N, x, y, cz, i, m);
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

// k_upd_p of iteration k-1 (:712-728, when upd) + k_matvec_dot of iteration k (:629-646):
// p = z + beta p_old, phi += alpha p_old; q = K p; block partial p.q (double).
__global__ void __launch_bounds__(TILE_THREADS) k_upd_matvec(int N, const float* __restrict__ z,
                                                             const float* __restrict__ p_in, float* __restrict__ p_out,
                                                             float* __restrict__ phi, const uint8_t* __restrict__ fm,
                                                             float* __restrict__ q, double* __restrict__ part, int upd,
                                                             const DevState* __restrict__ st, int k) {
    __shared__ float sp[HN];
    __shared__ uint8_t sm[HN];
    __shared__ double sh[8];
    if (cg_inactive(st, k)) return;
    const float a = (float)st->s[S_ALPHA], b = (float)st->s[S_BETA];
    _Pragma("unroll") for (int hi_ = 0; hi_ < HN_IT; ++hi_) {
        const int h = (int)threadIdx.x + hi_ * TILE_THREADS;
        if (h >= HN) break;
        const int i = halo_global(N, h);
        const int ic = i >= 0 ? i : 0;   // every load in bounds; outside cells staged as mask 0, value 0
        const int m = fm[ic];
        const float zv = z[ic], pv = p_in[ic];
        const float v = upd ? zv + b * pv : pv;
        sp[h] = i >= 0 ? v : 0.0f;
        sm[h] = (uint8_t)(i >= 0 ? m : 0);
    }
    __syncthreads();
    double s = 0.0;
    _Pragma("unroll") for (int ji_ = 0; ji_ < OWN_IT; ++ji_) {
        const int j = (int)threadIdx.x + ji_ * TILE_THREADS;
        int h;
        const int i = own_global(N, j, h);
        if (i < 0) continue;
        const float pn = sp[h];
        if (upd) {
            const float po = p_in[i];
            phi[i] = phi[i] + a * po;
            p_out[i] = pn;
        }
        const float qv = sm[h] ? kmul_tile(sp, sm, h) : 0.0f;
        q[i] = qv;
        s += (double)qv * (double)pn;
    }
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[tile_linear()] = t;
}

// alpha = rz / p.q (the finalize the reference folds into k_pre, :652-659)
__global__ void __launch_bounds__(256) k_fin_pq(const double* __restrict__ part, int nparts,
                                                DevState* __restrict__ st, int k) {
    __shared__ double sh[4];
    if (cg_inactive(st, k)) return;
    const double pq = reduce_parts_sum(part, nparts, sh);
    if (threadIdx.x == 0) st->s[S_ALPHA] = pq > 0.0 ? st->s[S_RZ] / pq : 0.0;
}

// k_pre (:648-672) + k_resid (:674-685) + k_fwd_red (:426-441).  Thread t owns one 2x2x2 block
// of the tile: r -= alpha q on its 8 cells (k > 0; written to r_out), rr = r - K z0 with z0 =
// omega D^-1 r staged for the tile + halo, then the level-1 components and the zero-guess red
// sweep.
__global__ void __launch_bounds__(TILE_THREADS) k_pre_resid_fwd(int N, const float* __restrict__ r_in,
                                                                float* __restrict__ r_out, const float* __restrict__ q,
                                                                const uint8_t* __restrict__ fm, Lvl L, int do_red,
                                                                const DevState* __restrict__ st, int k) {
    __shared__ float sz[HN];
    __shared__ uint8_t sm[HN];
    if (cg_inactive(st, k)) return;
    const float a = (float)st->s[S_ALPHA];
    _Pragma("unroll") for (int hi_ = 0; hi_ < HN_IT; ++hi_) {
        const int h = (int)threadIdx.x + hi_ * TILE_THREADS;
        if (h >= HN) break;
        const int i = halo_global(N, h);
        const int ic = i >= 0 ? i : 0;   // every load in bounds; outside cells staged as mask 0, value 0
        const int m = i >= 0 ? (int)fm[ic] : 0;
        const float r0 = r_in[ic], qv = q[ic];
        const float rv = k > 0 ? r0 - a * qv : r0;
        sz[h] = m ? OMEGA_JAC * rv / kdiag(m) : 0.0f;
        sm[h] = (uint8_t)m;
    }
    __syncthreads();
    const int t = threadIdx.x;
    const int bi = t % BTX, bj = (t / BTX) % BTY, bk = t / (BTX * BTY);
    const int mb = L.m;
    const int I = (int)blockIdx.x * BTX + bi, J = (int)blockIdx.y * BTY + bj, K = (int)blockIdx.z * BTZ + bk;
    if (I >= mb || J >= mb || K >= mb) return;
    float rr[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int dz = c >> 2, dy = (c >> 1) & 1, dx = c & 1;
        const int h = tile_h(2 * bi + dx, 2 * bj + dy, 2 * bk + dz);
        const int i = ((2 * K + dz) * N + 2 * J + dy) * N + 2 * I + dx;
        float rv = r_in[i];
        if (k > 0) {
            rv = rv - a * q[i];
            r_out[i] = rv;
        }
        rr[c] = sm[h] ? rv - kmul_tile(sz, sm, h) : 0.0f;
    }
    const size_t cell = lcell(mb, K, J, I);
    const bool red = do_red && ((I + J + K) & 1) == 0;
    float out[8], xo[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float acc = 0.0f;
#pragma unroll
        for (int s = 0; s < 8; ++s) acc += hsign(c, s >> 2, (s >> 1) & 1, s & 1) * rr[s];
        out[c] = acc;
        if (red) {
            const float d = L.diag[cell * 8 + c];
            xo[c] = d <= L.dmin ? 0.0f : acc / d;
        }
    }
    float4* rp = reinterpret_cast<float4*>(L.r + cell * 8);
    rp[0] = make_float4(out[0], out[1], out[2], out[3]);
    rp[1] = make_float4(out[4], out[5], out[6], out[7]);
    if (red) {
        float4* xp4 = reinterpret_cast<float4*>(L.x + cell * 8);
        xp4[0] = make_float4(xo[0], xo[1], xo[2], xo[3]);
        xp4[1] = make_float4(xo[4], xo[5], xo[6], xo[7]);
    }
}

// k_inv_transform (:352-365) + k_post_rz (:687-710): zt = z0 + T^T x1 (0 in solids) staged for
// the tile + halo from the level-1 values of the tile's blocks and their neighbours, then
// z = zt + omega D^-1 (r - K zt) on the tile and the block partials r.z, max|r|, r.r.  At k = 0
// z is also written to p_init (the first search direction, k_upd_p init).
__global__ void __launch_bounds__(TILE_THREADS) k_inv_post(int N, Lvl L, const float* __restrict__ r,
                                                           const uint8_t* __restrict__ fm, float* __restrict__ z,
                                                           float* __restrict__ p_init, double* __restrict__ part,
                                                           int nparts, const DevState* __restrict__ st, int k) {
    __shared__ float sx[HBN * 8];
    __shared__ float szt[HN];
    __shared__ uint8_t sm[HN];
    __shared__ double sh[8];
    if (cg_inactive(st, k)) return;
    const int mb = L.m;
    const int bx0 = (int)blockIdx.x * BTX - 1, by0 = (int)blockIdx.y * BTY - 1, bz0 = (int)blockIdx.z * BTZ - 1;
    _Pragma("unroll") for (int bi_ = 0; bi_ < HBN8_IT; ++bi_) {
        const int it = (int)threadIdx.x + bi_ * TILE_THREADS;
        if (it >= HBN * 8) break;
        const int b = it >> 3, c = it & 7;
        const int I = bx0 + b % HBX, J = by0 + (b / HBX) % HBY, K = bz0 + b / (HBX * HBY);
        const bool in = I >= 0 && J >= 0 && K >= 0 && I < mb && J < mb && K < mb;
        const float v = L.x[(in ? lcell(mb, K, J, I) : 0) * 8 + c];
        sx[it] = in ? v : 0.0f;
    }
    _Pragma("unroll") for (int hi_ = 0; hi_ < HN_IT; ++hi_) {
        const int h = (int)threadIdx.x + hi_ * TILE_THREADS;
        if (h >= HN) break;
        const int i = halo_global(N, h);
        const int m = fm[i >= 0 ? i : 0];
        sm[h] = (uint8_t)(i >= 0 ? m : 0);
    }
    __syncthreads();
    _Pragma("unroll") for (int hi_ = 0; hi_ < HN_IT; ++hi_) {
        const int h = (int)threadIdx.x + hi_ * TILE_THREADS;
        if (h >= HN) break;
        const int m = sm[h];
        const int i = halo_global(N, h);
        const float rv = r[i >= 0 ? i : 0];
        const int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
        const int hb = ((hz >> 1) * HBY + (hy >> 1)) * HBX + (hx >> 1);
        float acc = m ? OMEGA_JAC * rv / kdiag(m) : 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) acc += hsign(c, hz & 1, hy & 1, hx & 1) * sx[hb * 8 + c];
        szt[h] = m ? acc : 0.0f;
    }
    __syncthreads();
    double s = 0.0, mx = 0.0, rr = 0.0;
    _Pragma("unroll") for (int ji_ = 0; ji_ < OWN_IT; ++ji_) {
        const int j = (int)threadIdx.x + ji_ * TILE_THREADS;
        int h;
        const int i = own_global(N, j, h);
        if (i < 0) continue;
        const int m = sm[h];
        const float rv = r[i];
        float zv = 0.0f;
        if (m) zv = szt[h] + OMEGA_JAC * (rv - kmul_tile(szt, sm, h)) / kdiag(m);
        z[i] = zv;
        if (k == 0) p_init[i] = zv;
        s += (double)rv * (double)zv;
        mx = fmax(mx, fabs((double)rv));
        rr += (double)rv * (double)rv;
    }
    const double a = block_sum(s, sh);
    const double b = block_max(mx, sh);
    const double c = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        const int bl = tile_linear();
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

// The phi update of the stopping iteration J (k_upd_p :723-725 of the last loop body):
// phi += alpha_J p_{J-1}, where p_{J-1} is in p_buf[(J - 1) & 1].
__global__ void __launch_bounds__(256) k_phi_final(int N, int zc, float* __restrict__ phi,
                                                   const float* __restrict__ p0, const float* __restrict__ p1,
                                                   const DevState* __restrict__ st) {
    const int J = st->stop_at;
    if (J < 1 || J == INT_MAX) return;
    const float a = (float)st->s[S_ALPHA];
    const float* __restrict__ p = ((J - 1) & 1) ? p1 : p0;
    CELL_LOOP_BEGIN
        phi[i] = phi[i] + a * p[i];
    CELL_LOOP_END
}

// k_apply (:166-175): u -= g(phi) on fluid, u = 0 in solids
__global__ void __launch_bounds__(256) k_apply(int N, int zc, float* __restrict__ vel,
                                               const uint8_t* __restrict__ fm, const float* __restrict__ phi) {
    const int n3 = N * N * N;
    CELL_LOOP_BEGIN
        const int m = fm[i];
        if (m == 0) {
            vel[i] = 0.0f; vel[n3 + i] = 0.0f; vel[2 * n3 + i] = 0.0f;
            continue;
        }
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
            vel[ax * n3 + i] = vel[ax * n3 + i] - g_ax_xyz(phi, m, N, i, ax == 0 ? x : (ax == 1 ? y : cz), ax);
    CELL_LOOP_END
}

// True residual ||b - K phi||^2 with b = -d(u*) (fresh matvec, FP64 accumulation)
__global__ void __launch_bounds__(256) k_true_res(int N, int zc, const float* __restrict__ vel,
                                                  const uint8_t* __restrict__ fm, const float* __restrict__ phi,
                                                  double* __restrict__ part) {
    __shared__ double sh[8];
    const int n3 = N * N * N;
    double s = 0.0;
    CELL_LOOP_BEGIN
        const int m = fm[i];
        if (m == 0) continue;
        const double v = -(double)div2_xyz(vel, n3, m, N, x, y, cz, i) - (double)kmul_xyz(phi, fm, N, x, y, cz, i, m);
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
    const int n3 = N * N * N;
    CELL_LOOP_BEGIN
        const int m = mask_from_solid(solid, N, cz, y, x);
        out[i] = m == 0 ? 0.0f : 0.5f * fabsf(div2_xyz(vel, n3, m, N, x, y, cz, i));
    CELL_LOOP_END
}

// K p at fluid cells (parity tests of the operator)
__global__ void __launch_bounds__(256) k_kmul(int N, int zc, const float* __restrict__ p,
                                              const uint8_t* __restrict__ fm, float* __restrict__ out) {
    CELL_LOOP_BEGIN
        const int m = fm[i];
        out[i] = m ? kmul_xyz(p, fm, N, x, y, cz, i, m) : 0.0f;
    CELL_LOOP_END
}

// ================================================================= host side
void Pressure::init_tiles() {
    tgrid = dim3((N + TX - 1) / TX, (N + TY - 1) / TY, (N + TZ - 1) / TZ);
    ntiles = (int)(tgrid.x * tgrid.y * tgrid.z);
}

float* Pressure::r_buf(int k) const { return R[k & 1]; }

void Pressure::precond(int k, hipStream_t s) {
    float* r_in = k > 0 ? R[(k - 1) & 1] : R[0];
    float* r_out = R[k & 1];
    hipLaunchKernelGGL(k_pre_resid_fwd, tgrid, dim3(TILE_THREADS), 0, s, N, r_in, r_out, q, fm, lv[0],
                       S > 0 ? 1 : 0, st, k);
    vcycle_levels(k, s);
    hipLaunchKernelGGL(k_inv_post, tgrid, dim3(TILE_THREADS), 0, s, N, lv[0], r_out, fm, z, P[0], part, ntiles, st,
                       k);
}

void Pressure::rhs(const float* vel, const float* prm, hipStream_t s) {
    const Geom& g = geom;
    hipLaunchKernelGGL(k_rhs, g.grid, g.block, 0, s, N, g.zc, vel, fm, R[0], phi, prm, part, nblocks(g));
    hipLaunchKernelGGL(k_fin_rhs, dim3(1), dim3(256), 0, s, part, nblocks(g), prm, st);
}

void Pressure::iteration(int k, const float* prm, hipStream_t s) {
    if (k > 0) {
        // p_{k-1} = z_{k-1} + beta p_{k-2} (k >= 2); p_0 = z_0 was written at k = 0
        const int upd = k >= 2 ? 1 : 0;
        const float* p_in = upd ? P[(k - 2) & 1] : P[0];
        float* p_out = P[(k - 1) & 1];
        hipLaunchKernelGGL(k_upd_matvec, tgrid, dim3(TILE_THREADS), 0, s, N, z, p_in, p_out, phi, fm, q, part_pq,
                           upd, st, k);
        hipLaunchKernelGGL(k_fin_pq, dim3(1), dim3(256), 0, s, part_pq, ntiles, st, k);
    }
    precond(k, s);
    hipLaunchKernelGGL(k_fin_rz, dim3(1), dim3(256), 0, s, part, ntiles, prm, st, k);
}

void Pressure::finish(hipStream_t s) {
    const Geom& g = geom;
    hipLaunchKernelGGL(k_phi_final, g.grid, g.block, 0, s, N, g.zc, phi, P[0], P[1], st);
}

void Pressure::apply(float* vel, hipStream_t s) {
    const Geom& g = geom;
    hipLaunchKernelGGL(k_apply, g.grid, g.block, 0, s, N, g.zc, vel, fm, phi);
}

void Pressure::true_residual(const float* vel, hipStream_t s) {
    const Geom& g = geom;
    hipLaunchKernelGGL(k_true_res, g.grid, g.block, 0, s, N, g.zc, vel, fm, phi, part);
    hipLaunchKernelGGL(k_fin_true, dim3(1), dim3(256), 0, s, part, nblocks(g), st);
}

// One kernel of the projection on the current buffers, for timing (ids relative to
// VFX_K_MATVEC).  Only kernels that leave phi and the iteration state untouched are offered.
bool Pressure::launch_single(int which, hipStream_t s) {
    switch (which) {
        case 0:
            hipLaunchKernelGGL(k_upd_matvec, tgrid, dim3(TILE_THREADS), 0, s, N, z, P[0], P[1], phi, fm, q, part_pq,
                               0, st, 1);
            return true;
        case 1:
            hipLaunchKernelGGL(k_inv_post, tgrid, dim3(TILE_THREADS), 0, s, N, lv[0], R[0], fm, z, P[0], part, ntiles,
                               st, 1);
            return true;
        case 5:
            hipLaunchKernelGGL(k_pre_resid_fwd, tgrid, dim3(TILE_THREADS), 0, s, N, R[0], R[0], q, fm, lv[0],
                               S > 0 ? 1 : 0, st, 0);
            return true;
        default: return launch_single_mg(which, s);
    }
}

void launch_div_abs(const Geom& g, const float* vel, const float* solid, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_div_abs, g.grid, g.block, 0, s, g.N, g.zc, vel, solid, out);
}

void launch_kmul(const Geom& g, const float* p, const uint8_t* fm, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_kmul, g.grid, g.block, 0, s, g.N, g.zc, p, fm, out);
}

}  // namespace vfx
