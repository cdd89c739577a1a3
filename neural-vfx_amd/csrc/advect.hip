// MacCormack advection (K-ADV) and the density mass correction.
//
// Restates vfx/sim3d/kernels.py:66-178 (trilinear helpers, RK2 midpoint trace, advect_sl,
// advect_mc with the 8-corner clamp of Selle et al. 2008) and solver.py:40-76 (fluid density
// sums before / after advection and their ratio).  All channels share one backtrace.
#include "sim3d_kernels.hpp"

namespace vfx {

__device__ __forceinline__ float clampf_(float g, float lo, float hi) {
    return g < lo ? lo : (g > hi ? hi : g);
}

// base_idx / frac (kernels.py:67-74): clamp to [0, n-1], floor, keep one cell of room.  The
// clamp uses fmaxf/fminf so that a NaN coordinate also lands on cell 0 (equal to the reference's
// clamp for every other input).
__device__ __forceinline__ int base_idx(float g, int n) {
    int i = (int)floorf(fminf(fmaxf(g, 0.0f), (float)(n - 1)));
    return i < n - 2 ? i : n - 2;
}

struct Corner {
    size_t o;      // linear index of the (z0, y0, x0) corner
    float fx, fy, fz;
};

__device__ __forceinline__ Corner corner(int N, float px, float py, float pz) {
    int x0 = base_idx(px, N), y0 = base_idx(py, N), z0 = base_idx(pz, N);
    Corner c;
    c.o = idx3(N, z0, y0, x0);
    c.fx = clampf_(px, 0.0f, (float)(N - 1)) - (float)x0;
    c.fy = clampf_(py, 0.0f, (float)(N - 1)) - (float)y0;
    c.fz = clampf_(pz, 0.0f, (float)(N - 1)) - (float)z0;
    return c;
}

// tri (kernels.py:77-85): x first, then y, then z, in the reference's operation order.
__device__ __forceinline__ float tri(const float* __restrict__ f, int N, const Corner& c) {
    const size_t sy = (size_t)N, sz = (size_t)N * N;
    const float* p = f + c.o;
    float a00 = p[0] * (1.0f - c.fx) + p[1] * c.fx;
    float a01 = p[sy] * (1.0f - c.fx) + p[sy + 1] * c.fx;
    float a10 = p[sz] * (1.0f - c.fx) + p[sz + 1] * c.fx;
    float a11 = p[sz + sy] * (1.0f - c.fx) + p[sz + sy + 1] * c.fx;
    float a0 = a00 * (1.0f - c.fy) + a01 * c.fy;
    float a1 = a10 * (1.0f - c.fy) + a11 * c.fy;
    return a0 * (1.0f - c.fz) + a1 * c.fz;
}

// trace (kernels.py:115-120): RK2 midpoint over signed s (cells per m/s).
__device__ __forceinline__ void trace(const float* __restrict__ vel, int N, size_t n3, float px, float py,
                                      float pz, float s, float* q) {
    Corner c1 = corner(N, px, py, pz);
    float v1x = tri(vel, N, c1), v1y = tri(vel + n3, N, c1), v1z = tri(vel + 2 * n3, N, c1);
    float h = 0.5f * s;
    Corner c2 = corner(N, px + h * v1x, py + h * v1y, pz + h * v1z);
    float v2x = tri(vel, N, c2), v2y = tri(vel + n3, N, c2), v2z = tri(vel + 2 * n3, N, c2);
    q[0] = px + s * v2x;
    q[1] = py + s * v2y;
    q[2] = pz + s * v2z;
}

// Pass 1: semi-Lagrangian backward value of C channels -> hat.
__global__ void __launch_bounds__(256) k_adv_sl(int N, int zc, int C, const float* __restrict__ src,
                                                const float* __restrict__ vel, const float* __restrict__ prm,
                                                float inv_dx, float* __restrict__ hat) {
    const size_t n3 = (size_t)N * N * N;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= N || y >= N) return;
    const float s = prm[P_DT] * inv_dx;
    const int z1 = min((int)(blockIdx.z + 1) * zc, N);
    for (int z = blockIdx.z * zc; z < z1; ++z) {
        float q[3];
        trace(vel, N, n3, (float)x, (float)y, (float)z, -s, q);
        Corner c = corner(N, q[0], q[1], q[2]);
        const size_t i = idx3(N, z, y, x);
        for (int ch = 0; ch < C; ++ch) hat[ch * n3 + i] = tri(src + ch * n3, N, c);
    }
}

// Pass 2: MacCormack correction with the 8-corner clamp; optionally the block partial sums of
// fluid density before (src[0]) and after (out[0]) advection (solver.py:46-65), in double.
__global__ void __launch_bounds__(256) k_adv_mc(int N, int zc, int C, const float* __restrict__ src,
                                                const float* __restrict__ vel, const float* __restrict__ prm,
                                                float inv_dx, const float* __restrict__ hat,
                                                float* __restrict__ out, const uint8_t* __restrict__ fm,
                                                double* __restrict__ part, int nparts) {
    __shared__ double sh[8];
    const size_t n3 = (size_t)N * N * N;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const bool live = x < N && y < N;
    const float s = prm[P_DT] * inv_dx;
    double sa = 0.0, sb = 0.0;
    const int z1 = min((int)(blockIdx.z + 1) * zc, N);
    if (live) {
        for (int z = blockIdx.z * zc; z < z1; ++z) {
            float q[3], r[3];
            trace(vel, N, n3, (float)x, (float)y, (float)z, -s, q);
            trace(vel, N, n3, (float)x, (float)y, (float)z, s, r);
            const size_t qo = idx3(N, base_idx(q[2], N), base_idx(q[1], N), base_idx(q[0], N));
            Corner cr = corner(N, r[0], r[1], r[2]);
            const size_t i = idx3(N, z, y, x);
            const size_t sy = (size_t)N, sz = (size_t)N * N;
            for (int ch = 0; ch < C; ++ch) {
                const float* sc = src + ch * n3;
                const float h = hat[ch * n3 + i];
                const float bar = tri(hat + ch * n3, N, cr);
                const float s0 = sc[i];
                float val = h + 0.5f * (s0 - bar);
                const float* p = sc + qo;
                float lo = p[0], hi = lo, v;
                v = p[1];            lo = v < lo ? v : lo; hi = v > hi ? v : hi;
                v = p[sy];           lo = v < lo ? v : lo; hi = v > hi ? v : hi;
                v = p[sy + 1];       lo = v < lo ? v : lo; hi = v > hi ? v : hi;
                v = p[sz];           lo = v < lo ? v : lo; hi = v > hi ? v : hi;
                v = p[sz + 1];       lo = v < lo ? v : lo; hi = v > hi ? v : hi;
                v = p[sz + sy];      lo = v < lo ? v : lo; hi = v > hi ? v : hi;
                v = p[sz + sy + 1];  lo = v < lo ? v : lo; hi = v > hi ? v : hi;
                if (val < lo || val > hi) val = h;
                out[ch * n3 + i] = val;
                if (ch == 0 && part != nullptr && fm[i] != 0) {
                    sa += (double)s0;
                    sb += (double)val;
                }
            }
        }
    }
    if (part != nullptr) {
        double a = block_sum(sa, sh);
        double b = block_sum(sb, sh);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            const int bl = block_linear();
            part[bl] = a;
            part[nparts + bl] = b;
        }
    }
}

// ratio = a / b if a > 0 and b > 1e-30, else 1 (solver.py:67-74).
__global__ void __launch_bounds__(256) k_mass_fin(const double* __restrict__ part, int nparts,
                                                  DevState* __restrict__ st) {
    __shared__ double sh[4];
    double a = reduce_parts_sum(part, nparts, sh);
    double b = reduce_parts_sum(part + nparts, nparts, sh);
    if (threadIdx.x == 0) st->s[S_MASS_RATIO] = (a > 0.0 && b > 1e-30) ? a / b : 1.0;
}

void launch_advect(const Geom& g, int C, const float* src, const float* vel, const float* prm, float inv_dx,
                   float* hat, float* out, const uint8_t* fm, double* part, hipStream_t s) {
    hipLaunchKernelGGL(k_adv_sl, g.grid, g.block, 0, s, g.N, g.zc, C, src, vel, prm, inv_dx, hat);
    hipLaunchKernelGGL(k_adv_mc, g.grid, g.block, 0, s, g.N, g.zc, C, src, vel, prm, inv_dx, hat, out, fm, part,
                       nblocks(g));
}

void launch_adv_pass(int pass, const Geom& g, int C, const float* src, const float* vel, const float* prm,
                     float inv_dx, float* hat, float* out, const uint8_t* fm, double* part, hipStream_t s) {
    if (pass == 0)
        hipLaunchKernelGGL(k_adv_sl, g.grid, g.block, 0, s, g.N, g.zc, C, src, vel, prm, inv_dx, hat);
    else
        hipLaunchKernelGGL(k_adv_mc, g.grid, g.block, 0, s, g.N, g.zc, C, src, vel, prm, inv_dx, hat, out, fm, part,
                           nblocks(g));
}

void launch_mass_fin(const Geom& g, const double* part, DevState* st, hipStream_t s) {
    hipLaunchKernelGGL(k_mass_fin, dim3(1), dim3(256), 0, s, part, nblocks(g), st);
}

}  // namespace vfx
