// Guidance of a high-resolution simulation toward an independent low-resolution one (the paired
// runs of vfx/datagen3d/sim.py:52-79 with vfx/datagen3d/guide3d.py): every step
//
//     before the HR step  vel += k_vel up(L(vel_lr - pool(vel_hr)))                 (fluid cells)
//     after it            s   *= exp(k_s up(clamp(log((L(s_lr)^+ + eps) / (L(pool(s_hr))^+ + eps)))))
//                         for s = density, temperature - ambient; density >= 0, T >= min(T, ambient);
//                         fuel += k_fuel up(L(fuel_lr - pool(fuel_hr)))   (k_fuel = 0 by default)
//
// L is the reference's spectral low-pass exp(-(k / cutoff)^8), k in cycles per domain, applied to
// the mirror-extended n^3 grid through a real FFT (guide3d.py:52-68).  The mirror (half-sample
// even) extension makes that filter exactly a DCT-II along each axis, the radial mask on the
// coefficients, and the inverse DCT (the extension's spectrum is 2 e^{i pi k / 2n} times the
// DCT-II of the grid, and its Nyquist term vanishes), so here it is three n x n matrix products
// forward, the mask, three back: 6 small passes over the n^3 grid instead of two FFTs of (2n)^3.
// pool is the f^3 block mean (f = hr / lr), up the trilinear resize with half-pixel centres.
//
// Also tracking_error (guide3d.py:133-145): relative L2 at res^3 of the two runs.
//
// FP32 build only (the paired runs are FP32 simulations).
#include <cmath>
#include <new>
#include <vector>

#include "ctx.hpp"

#if !VFX_F64

struct vfx_guide {
    int hr = 0, lr = 0, f = 0;
    float k_vel = 0.f, k_scalar = 0.f, k_fuel = 0.f, cutoff = 16.f, eps_d = 0.01f, eps_t = 5.f, log_clamp = 1.f;
    void* pool = nullptr;
    // transform matrices M (out[k] = sum_j M[k][j] in[j]) stored transposed, [j][k]:
    float* fwd = nullptr;    // M[k][j] = cos(pi k (j + 1/2) / n)                      (DCT-II)
    float* inv = nullptr;    // M[j][k] = w_k / n cos(pi k (j + 1/2) / n), w_0 = 1, w_k = 2 (its inverse)
    float* mask = nullptr;   // n^3: exp(-(|k / 2| / cutoff)^8) at coefficient (k_z, k_y, k_x)
    float* a = nullptr;      // 3 n^3 work arrays: low-pass inputs / outputs and the passes' scratch
    float* b = nullptr;
    float* w = nullptr;
    uint8_t* fluid = nullptr;   // hr^3: hr_solid < 0.5, taken at creation like Guide3D.fluid
};

namespace vfx {

constexpr int GUIDE_LR_MAX = 128;
constexpr float GUIDE_AMB = 293.0f;   // AMBIENT_TEMP (vfx/core/fields.py)
constexpr int KB = 8;                 // outputs per thread of a transform pass

// out[c] = M in[c] along axis ax (0 x, 1 y, 2 z) of a (C, n, n, n) array: out[.., k, ..] =
// sum_j M[k][j] in[.., j, ..], optionally times mask at the output element.  One thread per line
// (x fastest where the axis is not x, so a wave's loads and stores are contiguous for the y and z
// passes) and per block of KB outputs; the output block and the channel come from the workgroup
// index, so the KB matrix values of each j are one scalar load from the transposed matrix mt.
__global__ void __launch_bounds__(256) k_axis_mat(const float* __restrict__ in, float* __restrict__ out,
                                                  const float* __restrict__ mt, const float* __restrict__ mask,
                                                  int n, int ax) {
    const int n2 = n * n, n3 = n2 * n;
    const int line = blockIdx.x * blockDim.x + threadIdx.x;
    const int kb = blockIdx.y, c = blockIdx.z;
    if (line >= n2) return;
    int base, stride;
    if (ax == 0) {
        base = line * n;
        stride = 1;
    } else if (ax == 1) {
        base = (line / n) * n2 + line % n;
        stride = n;
    } else {
        base = line;
        stride = n2;
    }
    const float* src = in + (size_t)c * n3 + base;
    float acc[KB];
#pragma unroll
    for (int kk = 0; kk < KB; ++kk) acc[kk] = 0.f;
    const float* m0 = mt + kb * KB;
    // n is a multiple of 16: the loads of 16 steps are issued together
#pragma unroll 16
    for (int j = 0; j < n; ++j) {
        const float v = src[(size_t)j * stride];
        const float* mj = m0 + (size_t)j * n;
#pragma unroll
        for (int kk = 0; kk < KB; ++kk) acc[kk] += mj[kk] * v;
    }
    float* dst = out + (size_t)c * n3 + base;
#pragma unroll
    for (int kk = 0; kk < KB; ++kk) {
        const size_t o = (size_t)(kb * KB + kk) * stride;
        dst[o] = mask ? acc[kk] * mask[base + o] : acc[kk];
    }
}

// the f^3 block mean of HR channel ch around LR cell i
__device__ __forceinline__ float block_mean(const float* __restrict__ hr, int H, int f, int ch, int z, int y, int x) {
    const size_t h3 = (size_t)H * H * H;
    const float* p = hr + ch * h3;
    float s = 0.f;
    for (int k = 0; k < f; ++k)
        for (int j = 0; j < f; ++j) {
            const float* row = p + ((size_t)(z * f + k) * H + (y * f + j)) * H + (size_t)x * f;
            for (int q = 0; q < f; ++q) s += row[q];
        }
    return s / (float)(f * f * f);
}

// Inputs of the low-pass on the LR grid.  mode 0 (velocity): a[c] = lr[1+c] - pool(hr[1+c]);
// mode 1 (scalars): a[c] = lr[s_c] - off_c, b[c] = pool(hr[s_c]) - off_c for s = density,
// temperature (off = 0, ambient); mode 2 (fuel): a[0] = lr[5] - pool(hr[5]).
__global__ void __launch_bounds__(256) k_guide_pool(const float* __restrict__ hr, const float* __restrict__ lr, int H,
                                                    int n, int f, int mode, float* __restrict__ a,
                                                    float* __restrict__ b) {
    const int n3 = n * n * n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const int x = i % n, y = (i / n) % n, z = i / (n * n);
    if (mode == 0) {
        for (int c = 0; c < 3; ++c) a[c * n3 + i] = lr[(1 + c) * n3 + i] - block_mean(hr, H, f, 1 + c, z, y, x);
    } else if (mode == 1) {
        a[i] = lr[i];
        b[i] = block_mean(hr, H, f, 0, z, y, x);
        a[n3 + i] = lr[4 * n3 + i] - GUIDE_AMB;
        b[n3 + i] = block_mean(hr, H, f, 4, z, y, x) - GUIDE_AMB;
    } else {
        a[i] = lr[5 * n3 + i] - block_mean(hr, H, f, 5, z, y, x);
    }
}

// logr = clamp(log((max(L a, 0) + eps) / (max(L b, 0) + eps)), -lc, lc) for density and T - ambient
__global__ void __launch_bounds__(256) k_guide_logratio(float* __restrict__ la, const float* __restrict__ lb, int n3,
                                                        float eps_d, float eps_t, float lc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n3) return;
    const float eps = i < n3 ? eps_d : eps_t;
    const float av = fmaxf(la[i], 0.f), bv = fmaxf(lb[i], 0.f);
    const float r = logf((av + eps) / (bv + eps));
    la[i] = fminf(fmaxf(r, -lc), lc);
}

// trilinear resize coordinates of HR cell d along one axis: the two LR neighbours and weights
struct Up1 {
    int i0, i1;
    float l0, l1;
};

__device__ __forceinline__ Up1 up1(int d, int f, int n) {
    Up1 u;
    float s = ((float)d + 0.5f) / (float)f - 0.5f;
    s = s < 0.f ? 0.f : s;
    u.i0 = min((int)s, n - 1);
    u.i1 = min(u.i0 + 1, n - 1);
    u.l1 = s - (float)u.i0;
    u.l0 = 1.f - u.l1;
    return u;
}

__device__ __forceinline__ float up_at(const float* __restrict__ g, int n, const Up1& ux, const Up1& uy,
                                       const Up1& uz) {
    auto at = [&](int zz, int yy, int xx) { return g[((size_t)zz * n + yy) * n + xx]; };
    const float r00 = ux.l0 * at(uz.i0, uy.i0, ux.i0) + ux.l1 * at(uz.i0, uy.i0, ux.i1);
    const float r01 = ux.l0 * at(uz.i0, uy.i1, ux.i0) + ux.l1 * at(uz.i0, uy.i1, ux.i1);
    const float r10 = ux.l0 * at(uz.i1, uy.i0, ux.i0) + ux.l1 * at(uz.i1, uy.i0, ux.i1);
    const float r11 = ux.l0 * at(uz.i1, uy.i1, ux.i0) + ux.l1 * at(uz.i1, uy.i1, ux.i1);
    return uz.l0 * (uy.l0 * r00 + uy.l1 * r01) + uz.l1 * (uy.l0 * r10 + uy.l1 * r11);
}

// HR update.  mode 0: vel[c] += k up(g[c]) fluid (c = 0..2); mode 1: density *= exp(k up(g0) fluid),
// T = max((T - amb) exp(k up(g1) fluid) + amb, min(T, amb)); mode 2: fuel = max(fuel + k up(g0) fluid, 0).
__global__ void __launch_bounds__(256) k_guide_apply(float* __restrict__ hr, const uint8_t* __restrict__ fluid,
                                                     const float* __restrict__ g, int H, int n, int f, int mode,
                                                     float k) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= H || y >= H) return;
    const size_t h3 = (size_t)H * H * H, n3 = (size_t)n * n * n;
    const size_t i = ((size_t)z * H + y) * H + x;
    const Up1 ux = up1(x, f, n), uy = up1(y, f, n), uz = up1(z, f, n);
    const float fl = fluid[i] ? 1.f : 0.f;
    if (mode == 0) {
        for (int c = 0; c < 3; ++c) hr[(1 + c) * h3 + i] += k * up_at(g + c * n3, n, ux, uy, uz) * fl;
    } else if (mode == 1) {
        const float g0 = expf(k * up_at(g, n, ux, uy, uz) * fl);
        const float g1 = expf(k * up_at(g + n3, n, ux, uy, uz) * fl);
        const float d = hr[i] * g0;
        hr[i] = d > 0.f ? d : 0.f;
        const float t = hr[4 * h3 + i];
        const float tf = t < GUIDE_AMB ? t : GUIDE_AMB;
        const float tn = (t - GUIDE_AMB) * g1 + GUIDE_AMB;
        hr[4 * h3 + i] = tn > tf ? tn : tf;
    } else {
        const float v = hr[5 * h3 + i] + k * up_at(g, n, ux, uy, uz) * fl;
        hr[5 * h3 + i] = v > 0.f ? v : 0.f;
    }
}

__global__ void __launch_bounds__(256) k_guide_fluid(const float* __restrict__ solid, uint8_t* __restrict__ fluid,
                                                     size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) fluid[i] = solid[i] < 0.5f ? 1 : 0;
}

// tracking_error, first pass: the adaptive-average-pool means (window [floor(i N / r),
// ceil((i + 1) N / r)) per axis) of the 6 channels of both runs at res^3, T minus ambient, one
// workgroup per (run, channel, output cell) summing its window in double.  Output (2, 6, res^3):
// hr means, then lr means.
__global__ void __launch_bounds__(256) k_track_pool(const float* __restrict__ hr, int H, const float* __restrict__ lr,
                                                    int L, int res, float* __restrict__ out) {
    __shared__ double sh[4];
    const int r3 = res * res * res;
    const int t = blockIdx.x;
    const int cell = t % r3, ch = (t / r3) % 6, which = t / (6 * r3);
    const int N = which == 0 ? H : L;
    const float* f = (which == 0 ? hr : lr) + (size_t)ch * N * N * N;
    const int cx = cell % res, cy = (cell / res) % res, cz = cell / (res * res);
    const int x0 = (cx * N) / res, x1 = ((cx + 1) * N + res - 1) / res;
    const int y0 = (cy * N) / res, y1 = ((cy + 1) * N + res - 1) / res;
    const int z0 = (cz * N) / res, z1 = ((cz + 1) * N + res - 1) / res;
    const int wx = x1 - x0, wy = y1 - y0, wz = z1 - z0, cnt = wx * wy * wz;
    double s = 0.0;
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
        const int x = x0 + q % wx, y = y0 + (q / wx) % wy, z = z0 + q / (wx * wy);
        s += f[((size_t)z * N + y) * N + x];
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
        float m = (float)(s / (double)cnt);
        if (ch == 4) m -= GUIDE_AMB;
        out[t] = m;
    }
}

// tracking_error, second pass (one workgroup): out = ||a - b|| / (||b|| + 1e-8) over density,
// T - ambient, velocity.
__global__ void __launch_bounds__(FIN_THREADS) k_track_fin(const float* __restrict__ p, int res,
                                                           float* __restrict__ out) {
    __shared__ double sh[FIN_THREADS / 64];
    const int r3 = res * res * res;
    const float* a = p;
    const float* b = p + 6 * r3;
    double num[3] = {0.0, 0.0, 0.0}, den[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < r3; i += blockDim.x) {
        for (int ch = 0; ch < 6; ++ch) {
            const int g = ch == 0 ? 0 : (ch == 4 ? 1 : (ch == 5 ? -1 : 2));
            if (g < 0) continue;
            const double d = (double)a[ch * r3 + i] - (double)b[ch * r3 + i];
            const double bb = (double)b[ch * r3 + i];
            num[g] += d * d;
            den[g] += bb * bb;
        }
    }
    for (int g = 0; g < 3; ++g) {
        const double sn = block_sum(num[g], sh);
        const double sd = block_sum(den[g], sh);
        if (threadIdx.x == 0) out[g] = (float)(sqrt(sn) / (sqrt(sd) + 1e-8));
    }
}

static int grid1(size_t n) { return (int)((n + 255) / 256); }

// L on C <= 3 channels of the LR grid, in place in x; w is C n^3 floats of scratch
static void lowpass(vfx_guide* g, float* x, float* w, int C, hipStream_t s) {
    const int n = g->lr;
    const dim3 grid(grid1((size_t)n * n), n / KB, C);
    hipLaunchKernelGGL(k_axis_mat, grid, dim3(256), 0, s, x, w, g->fwd, nullptr, n, 0);
    hipLaunchKernelGGL(k_axis_mat, grid, dim3(256), 0, s, w, x, g->fwd, nullptr, n, 1);
    hipLaunchKernelGGL(k_axis_mat, grid, dim3(256), 0, s, x, w, g->fwd, g->mask, n, 2);
    hipLaunchKernelGGL(k_axis_mat, grid, dim3(256), 0, s, w, x, g->inv, nullptr, n, 2);
    hipLaunchKernelGGL(k_axis_mat, grid, dim3(256), 0, s, x, w, g->inv, nullptr, n, 1);
    hipLaunchKernelGGL(k_axis_mat, grid, dim3(256), 0, s, w, x, g->inv, nullptr, n, 0);
}

static void apply(vfx_guide* g, float* hr, const float* src, int mode, float k, hipStream_t s) {
    const Geom gh = make_geom(g->hr, 1);
    hipLaunchKernelGGL(k_guide_apply, dim3(gh.grid.x, gh.grid.y, (unsigned)g->hr), gh.block, 0, s, hr, g->fluid, src,
                       g->hr, g->lr, g->f, mode, k);
}

}  // namespace vfx

using namespace vfx;

extern "C" {

int vfx_guide_create(int hr_res, int lr_res, const float* cfg, const float* hr_solid, void* stream, vfx_guide** out) {
    if (!out || !cfg || !hr_solid) return arg_error("null pointer");
    *out = nullptr;
    if (lr_res < 16 || lr_res % 16 != 0 || lr_res > GUIDE_LR_MAX)
        return arg_error("lr_res must be a multiple of 16 in [16, 128]");
    if (hr_res < lr_res || hr_res % lr_res != 0) return arg_error("hr_res must be an integer multiple of lr_res");
    vfx_guide* g = new (std::nothrow) vfx_guide;
    if (!g) return arg_error("out of host memory");
    g->hr = hr_res;
    g->lr = lr_res;
    g->f = hr_res / lr_res;
    g->k_vel = cfg[0];
    g->k_scalar = cfg[1];
    g->k_fuel = cfg[2];
    g->cutoff = cfg[3];
    g->eps_d = cfg[4];
    g->eps_t = cfg[5];
    g->log_clamp = cfg[6];
    const int n = lr_res;
    const size_t n3 = (size_t)n * n * n, h3 = (size_t)hr_res * hr_res * hr_res;
    const size_t floats = 2 * (size_t)n * n + n3 + 9 * n3;
    if (!hip_ok(hipMalloc(&g->pool, floats * sizeof(float) + h3), "guide buffers")) {
        delete g;
        return VFX_E_HIP;
    }
    float* p = static_cast<float*>(g->pool);
    g->fwd = p;
    g->inv = p + (size_t)n * n;
    g->mask = p + 2 * (size_t)n * n;
    g->a = g->mask + n3;
    g->b = g->a + 3 * n3;
    g->w = g->b + 3 * n3;
    g->fluid = reinterpret_cast<uint8_t*>(g->w + 3 * n3);
    // transform matrices and the mask in double, stored as float
    std::vector<float> fw((size_t)n * n), iv((size_t)n * n), mk(n3);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j) {
            const double c = std::cos(pi * k * (j + 0.5) / n);
            fw[(size_t)j * n + k] = (float)c;                               // M[k][j] at [j][k]
            iv[(size_t)k * n + j] = (float)((k == 0 ? 1.0 : 2.0) / n * c);   // M[j][k] at [k][j]
        }
    for (int kz = 0; kz < n; ++kz)
        for (int ky = 0; ky < n; ++ky)
            for (int kx = 0; kx < n; ++kx) {
                const double kk = std::sqrt(0.25 * ((double)kx * kx + (double)ky * ky + (double)kz * kz));
                mk[((size_t)kz * n + ky) * n + kx] = (float)std::exp(-std::pow(kk / (double)g->cutoff, 8.0));
            }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    bool ok = hip_ok(hipMemcpyAsync(g->fwd, fw.data(), fw.size() * sizeof(float), hipMemcpyHostToDevice, s), "guide") &&
              hip_ok(hipMemcpyAsync(g->inv, iv.data(), iv.size() * sizeof(float), hipMemcpyHostToDevice, s), "guide") &&
              hip_ok(hipMemcpyAsync(g->mask, mk.data(), mk.size() * sizeof(float), hipMemcpyHostToDevice, s), "guide");
    if (ok) {
        hipLaunchKernelGGL(k_guide_fluid, dim3(grid1(h3)), dim3(256), 0, s, hr_solid, g->fluid, h3);
        ok = launch_ok("guide fluid mask") && hip_ok(hipStreamSynchronize(s), "guide setup");
    }
    if (!ok) {
        hipFree(g->pool);
        delete g;
        return VFX_E_HIP;
    }
    *out = g;
    return VFX_OK;
}

int vfx_guide_destroy(vfx_guide* g) {
    if (!g) return VFX_OK;
    hipDeviceSynchronize();
    if (g->pool) hipFree(g->pool);
    delete g;
    return VFX_OK;
}

int vfx_guide_velocity(vfx_guide* g, float* hr_data, const float* lr_data, void* stream) {
    if (!g || !hr_data || !lr_data) return arg_error("null pointer");
    if (g->k_vel == 0.f) return VFX_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int n3 = g->lr * g->lr * g->lr;
    hipLaunchKernelGGL(k_guide_pool, dim3(grid1(n3)), dim3(256), 0, s, hr_data, lr_data, g->hr, g->lr, g->f, 0, g->a,
                       nullptr);
    lowpass(g, g->a, g->w, 3, s);
    apply(g, hr_data, g->a, 0, g->k_vel, s);
    if (!launch_ok("guide velocity")) return VFX_E_HIP;
    return VFX_OK;
}

int vfx_guide_scalars(vfx_guide* g, float* hr_data, const float* lr_data, void* stream) {
    if (!g || !hr_data || !lr_data) return arg_error("null pointer");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int n3 = g->lr * g->lr * g->lr;
    if (g->k_scalar != 0.f) {
        // a = s_lr, b = pool(s_hr) (density, T - ambient), then both low-passed
        hipLaunchKernelGGL(k_guide_pool, dim3(grid1(n3)), dim3(256), 0, s, hr_data, lr_data, g->hr, g->lr, g->f, 1,
                           g->a, g->b);
        lowpass(g, g->a, g->w, 2, s);
        lowpass(g, g->b, g->w, 2, s);
        hipLaunchKernelGGL(k_guide_logratio, dim3(grid1(2 * (size_t)n3)), dim3(256), 0, s, g->a, g->b, n3, g->eps_d,
                           g->eps_t, g->log_clamp);
        apply(g, hr_data, g->a, 1, g->k_scalar, s);
    }
    if (g->k_fuel != 0.f) {
        hipLaunchKernelGGL(k_guide_pool, dim3(grid1(n3)), dim3(256), 0, s, hr_data, lr_data, g->hr, g->lr, g->f, 2,
                           g->a, nullptr);
        lowpass(g, g->a, g->w, 1, s);
        apply(g, hr_data, g->a, 2, g->k_fuel, s);
    }
    if (!launch_ok("guide scalars")) return VFX_E_HIP;
    return VFX_OK;
}

int vfx_guide_lowpass(vfx_guide* g, const float* x, int C, float* out, void* stream) {
    if (!g || !x || !out) return arg_error("null pointer");
    if (C < 1 || C > 3) return arg_error("1 to 3 channels");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t n3 = (size_t)g->lr * g->lr * g->lr;
    if (!hip_ok(hipMemcpyAsync(out, x, C * n3 * sizeof(float), hipMemcpyDeviceToDevice, s), "lowpass copy"))
        return VFX_E_HIP;
    lowpass(g, out, g->w, C, s);
    if (!launch_ok("lowpass")) return VFX_E_HIP;
    return VFX_OK;
}

int vfx_tracking_error(const float* hr_data, int hr_res, const float* lr_data, int lr_res, int res, float* scratch,
                       float* out3, void* stream) {
    if (!hr_data || !lr_data || !scratch || !out3) return arg_error("null pointer");
    if (res < 1 || res > lr_res || res > hr_res) return arg_error("res must be in [1, min(hr_res, lr_res)]");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int r3 = res * res * res;
    hipLaunchKernelGGL(k_track_pool, dim3(2 * 6 * r3), dim3(256), 0, s, hr_data, hr_res, lr_data, lr_res, res,
                       scratch);
    hipLaunchKernelGGL(k_track_fin, dim3(1), dim3(FIN_THREADS), 0, s, scratch, res, out3);
    if (!launch_ok("tracking error")) return VFX_E_HIP;
    return VFX_OK;
}

}  // extern "C"

#endif  // !VFX_F64
