// Smoke and fire volume renderer: density + temperature volume (any resolution) and a camera ->
// (out_h, out_w, 3) uint8 sRGB image (the reference's vfx/render3d/renderer.py VolumeRenderer.draw
// with its kernels vfx/render3d/kernels.py, default path: no aux buffers, no temporal resolve, no
// detail field).
//
// Per draw, on one stream:
//   volume prep (skipped with reuse_volume):
//     k_pack_volume   density / temperature -> half2 volume + 32^3 occupancy bricks (kernels.py:124-147)
//     k_light_volume  per light voxel: transmittance toward the key light, sky occlusion, blackbody
//                     emission as the glow source (kernels.py:197-254)
//     glow pyramid    emission pooled to 32^3, blurred; 8^3 level blurred and upsampled (renderer.py:408-416)
//   view:
//     k_march         one ray per pixel at march resolution: front-to-back compositing with empty-brick
//                     skipping and early termination, opaque obstacles at the solid's 0.5 iso-surface,
//                     the floor plane (kernels.py:288-489)
//     bloom           thresholded HDR pooled to 1/4, blurred, a second level at 1/16 (renderer.py:452-457)
//     k_composite     Catmull-Rom upsample, background, bloom, exposure, contrast, ACES, sRGB, dither
//                     (kernels.py:571-600)
// Textures are sampled in software with the hardware convention of normalized coordinates and
// clamp addressing (texel centres at (i + 1/2) / n), in FP32 with exact weights; the volume, light
// and glow grids are stored in FP16 as the reference stores them.
//
// FP32 build only.
#include <hip/hip_fp16.h>

#include <cmath>
#include <new>
#include <utility>
#include <vector>

#include "ctx.hpp"

#if !VFX_F64

namespace vfx {

constexpr int BRICKS = 32;      // empty-space skipping grid over the unit cube (kernels.py:16)
constexpr int GLOW = 32;        // glow volume resolution
constexpr int LUT_N = 256;      // blackbody table entries
constexpr int MAX_TAPS = 16;

// shading parameters (kernels.py:19-52), unpacked from the VFX_RENDER_PARAMS floats
struct Shade {
    float steps, k_abs, k_shadow, k_solid, k_amb, k_emit, emit_min, emit_scale, soot, fire_light, phase_g, rim;
    float ms_a, ms_b, ms_c, exposure, contrast, bloom, bloom_thresh, t_lo, t_hi, d_eps, glow_coarse, floor;
    float3 light_dir, key_col, amb_col, smoke_albedo, bg_top, bg_bot, floor_col, grid_col, obst_col;
};

struct Cam {
    float v[16];   // pos, forward, right, up, tan(fov/2), aspect, jitter offset, 0 (camera_basis)
};

struct Taps {
    float w[MAX_TAPS];
    int r;
};

// ---------------------------------------------------------------- small vector helpers
__device__ __forceinline__ float3 f3(float a, float b, float c) { return make_float3(a, b, c); }
__device__ __forceinline__ float3 add(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 cw(float3 a, float3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(float3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ float3 norm3(float3 a) {
    const float l = len3(a);
    return l > 0.f ? f3(a.x / l, a.y / l, a.z / l) : f3(0.f, 0.f, 0.f);
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float fracf(float x) { return x - floorf(x); }
__device__ __forceinline__ float smoothstep_(float e0, float e1, float x) {
    const float t = clampf((x - e0) / (e1 - e0), 0.f, 1.f);
    return t * t * (3.f - 2.f * t);
}

// slab test against [lo, hi]^3: (t_near, t_far), a miss when t_near > t_far (kernels.py:70-77)
__device__ __forceinline__ float2 box_hit(float3 o, float3 invd, float3 lo, float3 hi) {
    const float3 t0 = cw(sub(lo, o), invd), t1 = cw(sub(hi, o), invd);
    const float nx = fminf(t0.x, t1.x), ny = fminf(t0.y, t1.y), nz = fminf(t0.z, t1.z);
    const float xx = fmaxf(t0.x, t1.x), xy = fmaxf(t0.y, t1.y), xz = fmaxf(t0.z, t1.z);
    return make_float2(fmaxf(fmaxf(nx, ny), nz), fminf(fminf(xx, xy), xz));
}

__device__ __forceinline__ float3 safe_inv(float3 d) {
    const float e = 1.0e-8f;
    const float x = fabsf(d.x) < e ? (d.x < 0.f ? -e : e) : d.x;
    const float y = fabsf(d.y) < e ? (d.y < 0.f ? -e : e) : d.y;
    const float z = fabsf(d.z) < e ? (d.z < 0.f ? -e : e) : d.z;
    return f3(1.f / x, 1.f / y, 1.f / z);
}

// Henyey-Greenstein without the 1 / (4 pi) (kernels.py:89-92)
__device__ __forceinline__ float hg(float cos_t, float g) {
    const float d = 1.f + g * g - 2.f * g * cos_t;
    return (1.f - g * g) / (d * sqrtf(d));
}

// interleaved gradient noise in [0, 1) (kernels.py:116-120)
__device__ __forceinline__ float ign(float x, float y) {
    const float f = 52.9829189f * fracf(0.06711056f * x + 0.00583715f * y);
    return fracf(f);
}

__device__ __forceinline__ float3 lut_lookup(const float* __restrict__ lut, float T, float t_lo, float t_hi) {
    const float f = clampf((T - t_lo) / (t_hi - t_lo), 0.f, 1.f) * (float)(LUT_N - 1);
    const int i0 = min((int)f, LUT_N - 2);
    const float t = f - (float)i0;
    const float* a = lut + 3 * i0;
    return f3(a[0] * (1.f - t) + a[3] * t, a[1] * (1.f - t) + a[4] * t, a[2] * (1.f - t) + a[5] * t);
}

__device__ __forceinline__ float3 emission(const float* __restrict__ lut, const Shade& p, float d, float T) {
    const float e = p.emit_min + (1.f - p.emit_min) * (1.f - expf(-p.k_emit * d));
    return mul(lut_lookup(lut, T, p.t_lo, p.t_hi), e * p.emit_scale);
}

// ---------------------------------------------------------------- texture sampling
// normalized coordinate u -> the two texel indices (clamp addressing) and the weight of the upper one
struct Tw {
    int i0, i1;
    float t;
};

__device__ __forceinline__ Tw tw(float u, int n) {
    const float f = clampf(u, -1.f, 2.f) * (float)n - 0.5f;
    const float fl = floorf(f);
    Tw w;
    w.t = f - fl;
    const int i = (int)fl;
    w.i0 = min(max(i, 0), n - 1);
    w.i1 = min(max(i + 1, 0), n - 1);
    return w;
}

struct Tex3 {
    int D, H, W;
};

template <class F>
__device__ __forceinline__ float tri_(const Tex3& g, float3 q, F at) {
    const Tw wx = tw(q.x, g.W), wy = tw(q.y, g.H), wz = tw(q.z, g.D);
    const float a00 = at(wz.i0, wy.i0, wx.i0) * (1.f - wx.t) + at(wz.i0, wy.i0, wx.i1) * wx.t;
    const float a01 = at(wz.i0, wy.i1, wx.i0) * (1.f - wx.t) + at(wz.i0, wy.i1, wx.i1) * wx.t;
    const float a10 = at(wz.i1, wy.i0, wx.i0) * (1.f - wx.t) + at(wz.i1, wy.i0, wx.i1) * wx.t;
    const float a11 = at(wz.i1, wy.i1, wx.i0) * (1.f - wx.t) + at(wz.i1, wy.i1, wx.i1) * wx.t;
    const float a0 = a00 * (1.f - wy.t) + a01 * wy.t;
    const float a1 = a10 * (1.f - wy.t) + a11 * wy.t;
    return a0 * (1.f - wz.t) + a1 * wz.t;
}

// two channels of a half2 grid in one pass over the 8 texels
__device__ __forceinline__ float2 tex_h2(const __half2* __restrict__ v, const Tex3& g, float3 q) {
    const Tw wx = tw(q.x, g.W), wy = tw(q.y, g.H), wz = tw(q.z, g.D);
    auto at = [&](int z, int y, int x) { return __half22float2(v[((size_t)z * g.H + y) * g.W + x]); };
    const float2 c000 = at(wz.i0, wy.i0, wx.i0), c001 = at(wz.i0, wy.i0, wx.i1);
    const float2 c010 = at(wz.i0, wy.i1, wx.i0), c011 = at(wz.i0, wy.i1, wx.i1);
    const float2 c100 = at(wz.i1, wy.i0, wx.i0), c101 = at(wz.i1, wy.i0, wx.i1);
    const float2 c110 = at(wz.i1, wy.i1, wx.i0), c111 = at(wz.i1, wy.i1, wx.i1);
    float2 r;
    {
        const float a00 = c000.x * (1.f - wx.t) + c001.x * wx.t, a01 = c010.x * (1.f - wx.t) + c011.x * wx.t;
        const float a10 = c100.x * (1.f - wx.t) + c101.x * wx.t, a11 = c110.x * (1.f - wx.t) + c111.x * wx.t;
        r.x = (a00 * (1.f - wy.t) + a01 * wy.t) * (1.f - wz.t) + (a10 * (1.f - wy.t) + a11 * wy.t) * wz.t;
    }
    {
        const float a00 = c000.y * (1.f - wx.t) + c001.y * wx.t, a01 = c010.y * (1.f - wx.t) + c011.y * wx.t;
        const float a10 = c100.y * (1.f - wx.t) + c101.y * wx.t, a11 = c110.y * (1.f - wx.t) + c111.y * wx.t;
        r.y = (a00 * (1.f - wy.t) + a01 * wy.t) * (1.f - wz.t) + (a10 * (1.f - wy.t) + a11 * wy.t) * wz.t;
    }
    return r;
}

// uint8 grid read as value / 255 (normalized read mode)
__device__ __forceinline__ float tex_u8(const uint8_t* __restrict__ v, const Tex3& g, float3 q) {
    return tri_(g, q, [&](int z, int y, int x) { return (float)v[((size_t)z * g.H + y) * g.W + x] * (1.f / 255.f); });
}

// rgb of a (D, H, W, 4) half grid: one 8-byte load per texel, the three channels interpolated alike
__device__ __forceinline__ float3 tex_h4(const __half* __restrict__ v, const Tex3& g, float3 q) {
    const Tw wx = tw(q.x, g.W), wy = tw(q.y, g.H), wz = tw(q.z, g.D);
    const uint2* p = reinterpret_cast<const uint2*>(v);
    float c[8][3];
    const int zs[2] = {wz.i0, wz.i1}, ys[2] = {wy.i0, wy.i1}, xs[2] = {wx.i0, wx.i1};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint2 u = p[((size_t)zs[k >> 2] * g.H + ys[(k >> 1) & 1]) * g.W + xs[k & 1]];
        const float2 rg = __half22float2(*reinterpret_cast<const __half2*>(&u.x));
        const float2 ba = __half22float2(*reinterpret_cast<const __half2*>(&u.y));
        c[k][0] = rg.x;
        c[k][1] = rg.y;
        c[k][2] = ba.x;
    }
    float r[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const float a00 = c[0][ch] * (1.f - wx.t) + c[1][ch] * wx.t, a01 = c[2][ch] * (1.f - wx.t) + c[3][ch] * wx.t;
        const float a10 = c[4][ch] * (1.f - wx.t) + c[5][ch] * wx.t, a11 = c[6][ch] * (1.f - wx.t) + c[7][ch] * wx.t;
        r[ch] = (a00 * (1.f - wy.t) + a01 * wy.t) * (1.f - wz.t) + (a10 * (1.f - wy.t) + a11 * wy.t) * wz.t;
    }
    return f3(r[0], r[1], r[2]);
}

// ---------------------------------------------------------------- prep kernels
__global__ void k_fill_u8(uint8_t* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__device__ __forceinline__ void mark_bricks(uint8_t* __restrict__ occ, int z, int y, int x, int D, int H, int W) {
    const int bz0 = max(z - 1, 0) * BRICKS / D, bz1 = min(z + 1, D - 1) * BRICKS / D;
    const int by0 = max(y - 1, 0) * BRICKS / H, by1 = min(y + 1, H - 1) * BRICKS / H;
    const int bx0 = max(x - 1, 0) * BRICKS / W, bx1 = min(x + 1, W - 1) * BRICKS / W;
    for (int bz = bz0; bz <= bz1; ++bz)
        for (int by = by0; by <= by1; ++by)
            for (int bx = bx0; bx <= bx1; ++bx) occ[(bz * BRICKS + by) * BRICKS + bx] = 1;
}

// pack_volume (kernels.py:124-147)
__global__ void __launch_bounds__(256) k_pack_volume(const float* __restrict__ dens, const float* __restrict__ temp,
                                                     int D, int H, int W, __half2* __restrict__ packed,
                                                     uint8_t* __restrict__ occ, float d_eps, float t_hot) {
    const size_t n = (size_t)D * H * W;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = dens[i], T = temp[i];
    packed[i] = __halves2half2(__float2half_rn(fmaxf(d, 0.f)), __float2half_rn(T));
    if (d > d_eps || T > t_hot) {
        const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)(i / ((size_t)W * H));
        mark_bricks(occ, z, y, x, D, H, W);
    }
}

// separable blur along one axis of a (C, D, H, W) grid, replicate padding
__global__ void __launch_bounds__(256) k_blur_axis(const float* __restrict__ in, float* __restrict__ out, int C, int D,
                                                   int H, int W, int axis, Taps k) {
    const size_t n = (size_t)C * D * H * W;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)((i / ((size_t)W * H)) % D);
    const int len = axis == 0 ? W : (axis == 1 ? H : D);
    const int pos = axis == 0 ? x : (axis == 1 ? y : z);
    const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)W : (size_t)W * H);
    const size_t base = i - (size_t)pos * stride;
    float s = 0.f;
    for (int t = -k.r; t <= k.r; ++t) {
        const int p = min(max(pos + t, 0), len - 1);
        s += k.w[t + k.r] * in[base + (size_t)p * stride];
    }
    out[i] = s;
}

// pack_solid (kernels.py:175-194): smoothed solid -> uint8 + occupancy bricks
__global__ void __launch_bounds__(256) k_pack_solid(const float* __restrict__ s, int D, int H, int W,
                                                    uint8_t* __restrict__ out, uint8_t* __restrict__ occ) {
    const size_t n = (size_t)D * H * W;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = s[i];
    out[i] = (uint8_t)min(max((int)(v * 255.f + 0.5f), 0), 255);
    if (v > 0.f) {
        const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)(i / ((size_t)W * H));
        mark_bricks(occ, z, y, x, D, H, W);
    }
}

// light_volume (kernels.py:203-254)
__global__ void __launch_bounds__(256) k_light_volume(const __half2* __restrict__ vol, Tex3 gv,
                                                      const uint8_t* __restrict__ solid, Tex3 gs, int has_solid,
                                                      const float* __restrict__ lut, Shade p, int L,
                                                      __half2* __restrict__ light, __half* __restrict__ glow_src) {
    const int n = L * L * L;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = i % L, y = (i / L) % L, z = i / (L * L);
    const float inv = 1.f / (float)L;
    const float3 pos = f3(((float)x + 0.5f) * inv, ((float)y + 0.5f) * inv, ((float)z + 0.5f) * inv);
    const float2 here = tex_h2(vol, gv, pos);
    const float3 ld = p.light_dir;
    const float2 tb = box_hit(pos, safe_inv(ld), f3(0.f, 0.f, 0.f), f3(1.f, 1.f, 1.f));
    const float t_exit = tb.y;
    const float ds = 1.5f * inv;
    float od = 0.f;
    float t = (0.25f + 0.75f * ign((float)x + 7.f * (float)z, (float)y + 13.f * (float)z)) * ds;
    while (t < t_exit) {
        const float3 q = add(pos, mul(ld, t));
        const float2 s = tex_h2(vol, gv, q);
        od += p.k_shadow * s.x * ds;
        if (has_solid != 0) od += p.k_solid * tex_u8(solid, gs, q) * ds;
        if (od > 12.f) break;
        t += ds;
    }
    float oda = 0.f;
    const float da = 3.f * inv;
    for (int k = 0; k < 6; ++k) {
        const float qy = pos.y + ((float)k + 0.5f) * da;
        if (qy < 1.f) {
            const float2 s = tex_h2(vol, gv, f3(pos.x, qy, pos.z));
            oda += p.k_amb * s.x * da;
        }
    }
    light[i] = __halves2half2(__float2half_rn(expf(-od)), __float2half_rn(expf(-oda)));
    float3 e = f3(0.f, 0.f, 0.f);
    if (here.y > p.t_lo) e = emission(lut, p, here.x, here.y);
    glow_src[4 * (size_t)i + 0] = __float2half_rn(e.x);
    glow_src[4 * (size_t)i + 1] = __float2half_rn(e.y);
    glow_src[4 * (size_t)i + 2] = __float2half_rn(e.z);
    glow_src[4 * (size_t)i + 3] = __float2half_rn(0.f);
}

// adaptive average pool of the glow source (L^3, 4 half) to (3, 32^3) float
__global__ void __launch_bounds__(256) k_glow_pool(const __half* __restrict__ src, int L, float* __restrict__ out) {
    const int n = GLOW * GLOW * GLOW;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = i % GLOW, y = (i / GLOW) % GLOW, z = i / (GLOW * GLOW);
    const int x0 = x * L / GLOW, x1 = ((x + 1) * L + GLOW - 1) / GLOW;
    const int y0 = y * L / GLOW, y1 = ((y + 1) * L + GLOW - 1) / GLOW;
    const int z0 = z * L / GLOW, z1 = ((z + 1) * L + GLOW - 1) / GLOW;
    float s[3] = {0.f, 0.f, 0.f};
    for (int zz = z0; zz < z1; ++zz)
        for (int yy = y0; yy < y1; ++yy)
            for (int xx = x0; xx < x1; ++xx) {
                const __half* v = src + 4 * (((size_t)zz * L + yy) * L + xx);
                for (int c = 0; c < 3; ++c) s[c] += __half2float(v[c]);
            }
    const float cnt = (float)((z1 - z0) * (y1 - y0) * (x1 - x0));
    for (int c = 0; c < 3; ++c) out[(size_t)c * n + i] = s[c] / cnt;
}

// avg_pool3d by 4: (3, 32^3) -> (3, 8^3)
__global__ void __launch_bounds__(256) k_glow_pool4(const float* __restrict__ in, float* __restrict__ out) {
    constexpr int G8 = GLOW / 4;
    const int n = 3 * G8 * G8 * G8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = i % G8, y = (i / G8) % G8, z = (i / (G8 * G8)) % G8, c = i / (G8 * G8 * G8);
    const float* p = in + (size_t)c * GLOW * GLOW * GLOW;
    float s = 0.f;
    for (int k = 0; k < 4; ++k)
        for (int j = 0; j < 4; ++j)
            for (int q = 0; q < 4; ++q) s += p[((z * 4 + k) * GLOW + (y * 4 + j)) * GLOW + x * 4 + q];
    out[i] = s / 64.f;
}

// torch's linear resize source index (align_corners=False): max(scale (d + 0.5) - 0.5, 0)
struct Lin {
    int i0, i1;
    float l0, l1;
};

__device__ __forceinline__ Lin lin(int d, int in_n, float scale) {
    float s = scale * ((float)d + 0.5f) - 0.5f;
    s = s < 0.f ? 0.f : s;
    Lin r;
    r.i0 = min((int)s, in_n - 1);
    r.i1 = r.i0 + (r.i0 < in_n - 1 ? 1 : 0);
    r.l1 = s - (float)r.i0;
    r.l0 = 1.f - r.l1;
    return r;
}

// glow = g32b + glow_coarse * trilinear(g8b -> 32^3), stored as (32^3, 4) half
__global__ void __launch_bounds__(256) k_glow_combine(const float* __restrict__ g32b, const float* __restrict__ g8b,
                                                      float coarse, __half* __restrict__ glow) {
    constexpr int G8 = GLOW / 4;
    const int n = GLOW * GLOW * GLOW;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = i % GLOW, y = (i / GLOW) % GLOW, z = i / (GLOW * GLOW);
    const Lin lx = lin(x, G8, 0.25f), ly = lin(y, G8, 0.25f), lz = lin(z, G8, 0.25f);
    for (int c = 0; c < 3; ++c) {
        const float* g = g8b + (size_t)c * G8 * G8 * G8;
        auto at = [&](int zz, int yy, int xx) { return g[(zz * G8 + yy) * G8 + xx]; };
        const float up = lz.l0 * (ly.l0 * (lx.l0 * at(lz.i0, ly.i0, lx.i0) + lx.l1 * at(lz.i0, ly.i0, lx.i1)) +
                                  ly.l1 * (lx.l0 * at(lz.i0, ly.i1, lx.i0) + lx.l1 * at(lz.i0, ly.i1, lx.i1))) +
                         lz.l1 * (ly.l0 * (lx.l0 * at(lz.i1, ly.i0, lx.i0) + lx.l1 * at(lz.i1, ly.i0, lx.i1)) +
                                  ly.l1 * (lx.l0 * at(lz.i1, ly.i1, lx.i0) + lx.l1 * at(lz.i1, ly.i1, lx.i1)));
        glow[4 * (size_t)i + c] = __float2half_rn(g32b[(size_t)c * n + i] + coarse * up);
    }
    glow[4 * (size_t)i + 3] = __float2half_rn(0.f);
}

// ---------------------------------------------------------------- march
// light-volume lookup valid outside the cube too (kernels.py:257-267)
__device__ __forceinline__ float2 light_at(const __half2* __restrict__ light, const Tex3& gl, float3 ld, float3 q) {
    const bool inside = q.x >= 0.f && q.x <= 1.f && q.y >= 0.f && q.y <= 1.f && q.z >= 0.f && q.z <= 1.f;
    if (inside) return tex_h2(light, gl, q);
    const float2 tb = box_hit(q, safe_inv(ld), f3(0.f, 0.f, 0.f), f3(1.f, 1.f, 1.f));
    if (tb.x <= tb.y && tb.y > 0.f) {
        const float3 e = add(q, mul(ld, fmaxf(tb.x, 0.f) + 1.0e-3f));
        return make_float2(tex_h2(light, gl, e).x, 1.f);
    }
    return make_float2(1.f, 1.f);
}

// glow lookup, fading with the squared distance outside the cube (kernels.py:270-275)
__device__ __forceinline__ float3 glow_at(const __half* __restrict__ glow, float3 q) {
    const float3 c = f3(clampf(q.x, 0.f, 1.f), clampf(q.y, 0.f, 1.f), clampf(q.z, 0.f, 1.f));
    const float3 g = tex_h4(glow, Tex3{GLOW, GLOW, GLOW}, c);
    const float dout = len3(sub(q, c));
    return mul(g, 1.f / (1.f + 60.f * dout * dout));
}

struct MarchArgs {
    const __half2* vol;
    Tex3 gv;
    const __half2* light;
    Tex3 gl;
    const __half* glow;
    const uint8_t* solid;
    Tex3 gs;
    const uint8_t* occ;
    const uint8_t* occ_s;
    int has_solid;
    float solid_res;
    const float* lut;
};

// aux inputs of the march (return_aux): velocity grid (D, H, W, 4) half in m/s, the previous
// camera, the sim step, the output size, and the premultiplied aux buffer (mh, mw)
struct AuxArgs {
    const __half* vel;
    Tex3 gvel;
    int has_vel;
    Cam cam_prev;
    float vel_dt, out_w, out_h;
    float4* aux;
};

// world point -> continuous output-pixel coordinates (x right, y down) of a camera (kernels.py:278-285)
__device__ __forceinline__ float2 project_px(float3 q, const Cam& cam, float out_w, float out_h) {
    const float3 rel = sub(q, f3(cam.v[0], cam.v[1], cam.v[2]));
    const float z = fmaxf(dot3(rel, f3(cam.v[3], cam.v[4], cam.v[5])), 1.0e-4f);
    const float sx = dot3(rel, f3(cam.v[6], cam.v[7], cam.v[8])) / (z * cam.v[12] * cam.v[13]);
    const float sy = dot3(rel, f3(cam.v[9], cam.v[10], cam.v[11])) / (z * cam.v[12]);
    return make_float2((sx + 1.f) * 0.5f * out_w, (1.f - sy) * 0.5f * out_h);
}

// march (kernels.py:288-489).  AUX also accumulates, from the same samples, the medium-only
// transmittance and the opacity-weighted (T a) mean distance and velocity, and writes the
// premultiplied (alpha, alpha depth, alpha motion_x, alpha motion_y); the colour code is the same.
template <bool AUX>
__global__ void __launch_bounds__(256) k_march(MarchArgs a, Shade p, Cam cam, int h, int w, float4* __restrict__ out,
                                               AuxArgs x) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y * blockDim.y + threadIdx.y;
    if (c >= w || r >= h) return;
    const float3 o = f3(cam.v[0], cam.v[1], cam.v[2]);
    const float3 fw = f3(cam.v[3], cam.v[4], cam.v[5]);
    const float3 rt = f3(cam.v[6], cam.v[7], cam.v[8]);
    const float3 upv = f3(cam.v[9], cam.v[10], cam.v[11]);
    const float th = cam.v[12], aspect = cam.v[13];
    const float sx = (2.f * ((float)c + 0.5f) / (float)w - 1.f) * th * aspect;
    const float sy = (1.f - 2.f * ((float)r + 0.5f) / (float)h) * th;
    const float3 d = norm3(add(add(fw, mul(rt, sx)), mul(upv, sy)));
    const float3 invd = safe_inv(d);
    const float3 ld = p.light_dir;
    float3 rad = f3(0.f, 0.f, 0.f);
    float T = 1.f;
    const float2 tb = box_hit(o, invd, f3(0.f, 0.f, 0.f), f3(1.f, 1.f, 1.f));
    const float t_near = fmaxf(tb.x, 0.f), t_far = tb.y;
    const float ds = 1.f / p.steps;
    const float cos_t = dot3(d, ld);
    const float ph0 = hg(cos_t, p.phase_g) * p.rim + hg(cos_t, -0.25f) * (1.f - p.rim);
    const float ph1 = hg(cos_t, p.phase_g * p.ms_c) * p.rim + hg(cos_t, -0.25f * p.ms_c) * (1.f - p.rim);
    const float ph2 = hg(cos_t, p.phase_g * p.ms_c * p.ms_c) * p.rim + (1.f - p.rim);
    const float a1 = p.ms_a, a2 = p.ms_a * p.ms_a, b1 = p.ms_b, b2 = p.ms_b * p.ms_b;
    float Tm = 1.f, aw = 0.f, ad = 0.f;   // AUX: medium transmittance, sum of weights, sum w t
    float3 av = f3(0.f, 0.f, 0.f);        // AUX: sum w v
    if (t_near < t_far) {
        const float jit = fracf(ign((float)c, (float)r) + cam.v[14]);
        const float t0 = t_near + jit * ds;
        float t = t0, t_prev = t_near;
        while (t < t_far) {
            const float3 q = add(o, mul(d, t));
            const int bx = min(max((int)(q.x * (float)BRICKS), 0), BRICKS - 1);
            const int by = min(max((int)(q.y * (float)BRICKS), 0), BRICKS - 1);
            const int bz = min(max((int)(q.z * (float)BRICKS), 0), BRICKS - 1);
            const int bi = (bz * BRICKS + by) * BRICKS + bx;
            const int o_d = a.occ[bi];
            const int o_s = a.has_solid != 0 ? a.occ_s[bi] : 0;
            if (o_d == 0 && o_s == 0) {
                // skip to the exit of this empty brick, staying on the jittered step lattice
                const float inv_b = 1.f / (float)BRICKS;
                const float3 lo = mul(f3((float)bx, (float)by, (float)bz), inv_b);
                const float2 bb = box_hit(o, invd, lo, add(lo, f3(inv_b, inv_b, inv_b)));
                const float k = ceilf((bb.y - t0) / ds + 1.0e-4f);
                t_prev = t;
                t = fmaxf(t0 + k * ds, t + ds);
                continue;
            }
            if (o_s != 0 && tex_u8(a.solid, a.gs, q) > 0.5f) {
                // refine the surface crossing between t_prev and t, shade the obstacle, stop
                float lo_t = t_prev, hi_t = t;
                for (int it = 0; it < 5; ++it) {
                    const float m = 0.5f * (lo_t + hi_t);
                    if (tex_u8(a.solid, a.gs, add(o, mul(d, m))) > 0.5f) hi_t = m;
                    else lo_t = m;
                }
                const float3 ph = add(o, mul(d, hi_t));
                const float e = 1.f / a.solid_res;
                const float gx = tex_u8(a.solid, a.gs, add(ph, f3(e, 0.f, 0.f))) - tex_u8(a.solid, a.gs, sub(ph, f3(e, 0.f, 0.f)));
                const float gy = tex_u8(a.solid, a.gs, add(ph, f3(0.f, e, 0.f))) - tex_u8(a.solid, a.gs, sub(ph, f3(0.f, e, 0.f)));
                const float gz = tex_u8(a.solid, a.gs, add(ph, f3(0.f, 0.f, e))) - tex_u8(a.solid, a.gs, sub(ph, f3(0.f, 0.f, e)));
                float3 nrm = f3(-gx, -gy, -gz);
                if (len3(nrm) < 1.0e-6f) nrm = f3(-d.x, -d.y, -d.z);
                nrm = norm3(nrm);
                const float3 lq = add(ph, mul(nrm, 0.03f));
                const float2 lv = tex_h2(a.light, a.gl, lq);
                const float3 gl = glow_at(a.glow, lq);
                const float ndl = fmaxf(dot3(nrm, ld), 0.f);
                const float3 hv = norm3(sub(ld, d));
                const float spec = powf(fmaxf(dot3(nrm, hv), 0.f), 40.f) * 0.25f;
                const float hemi = 0.45f + 0.55f * (0.5f + 0.5f * nrm.y);
                const float fres = powf(1.f - fmaxf(dot3(nrm, f3(-d.x, -d.y, -d.z)), 0.f), 4.f) * 0.08f;
                float3 col = cw(p.obst_col, add(add(mul(p.key_col, ndl * lv.x), mul(p.amb_col, hemi * lv.y)),
                                                mul(gl, 2.f * p.fire_light)));
                col = add(add(col, mul(p.key_col, spec * lv.x)), mul(p.amb_col, fres));
                rad = add(rad, mul(col, T));
                T = 0.f;
                break;
            }
            if (o_d != 0) {
                const float2 s = tex_h2(a.vol, a.gv, q);
                const float dn = s.x, Tk = s.y;
                const bool hot = Tk > p.t_lo;
                if (dn > p.d_eps || hot) {
                    const float sigma = p.k_abs * fmaxf(dn, 0.f);
                    const float al = 1.f - expf(-sigma * ds);
                    float3 col = f3(0.f, 0.f, 0.f);
                    if (al > 1.0e-5f) {
                        const float2 lv = tex_h2(a.light, a.gl, q);
                        const float3 gl = glow_at(a.glow, q);
                        const float shd = smoothstep_(p.t_lo, p.t_lo + 500.f, Tk);
                        const float3 alb = mul(p.smoke_albedo, 1.f - p.soot * shd);
                        const float tl = fmaxf(lv.x, 1.0e-6f);
                        const float key = tl * ph0 + a1 * powf(tl, b1) * ph1 + a2 * powf(tl, b2) * ph2;
                        const float3 lit = add(add(mul(p.key_col, key), mul(p.amb_col, lv.y)), mul(gl, p.fire_light));
                        col = mul(cw(alb, lit), al);
                    }
                    if (hot) col = add(col, mul(emission(a.lut, p, dn, Tk), ds));
                    if (AUX) {
                        const float wgt = T * al;
                        if (wgt > 0.f) {
                            aw += wgt;
                            ad += wgt * t;
                            if (x.has_vel != 0) {
                                const float3 v = tex_h4(x.vel, x.gvel, q);
                                av = add(av, mul(v, wgt));
                            }
                        }
                    }
                    rad = add(rad, mul(col, T));
                    T = T * (1.f - al);
                    if (T < 0.004f) {
                        T = 0.f;
                        if (AUX) Tm = 0.f;
                        break;
                    }
                    if (AUX) Tm = T;
                }
            }
            t_prev = t;
            t = t + ds;
        }
    }
    // floor plane y = 0 (behind everything inside the cube)
    if (T > 0.f && d.y < -1.0e-4f && o.y > 0.f) {
        const float tf = -o.y / d.y;
        const float3 pf = add(o, mul(d, tf));
        const float cx = pf.x - 0.5f, cz = pf.z - 0.5f;
        const float rr = sqrtf(cx * cx + cz * cz);
        const float fade = (1.f - smoothstep_(0.35f, 1.9f, rr)) * p.floor;
        if (fade > 0.f) {
            const float pix = tf * 2.f * th / (float)h;
            const float gsz = 0.125f;
            const float gxl = fabsf(pf.x / gsz - roundf(pf.x / gsz)) * gsz;
            const float gzl = fabsf(pf.z / gsz - roundf(pf.z / gsz)) * gsz;
            const float line = 1.f - smoothstep_(0.f, 1.2f * pix + 0.0015f, fminf(gxl, gzl));
            const bool inside = pf.x >= 0.f && pf.x <= 1.f && pf.z >= 0.f && pf.z <= 1.f;
            const float3 lq = f3(pf.x, 0.004f, pf.z);
            const float2 lv = light_at(a.light, a.gl, ld, lq);
            const float3 gl = glow_at(a.glow, lq);
            const float3 base = add(p.floor_col, mul(p.grid_col, line * (inside ? 1.f : 0.45f)));
            const float3 lit = add(add(mul(p.key_col, ld.y * lv.x), mul(p.amb_col, 0.6f + 0.4f * lv.y)),
                                   mul(gl, 2.5f * p.fire_light));
            const float3 fc = cw(base, lit);
            rad = add(rad, mul(fc, T * fade));
            T = T * (1.f - fade);
        }
    }
    out[(size_t)r * w + c] = make_float4(rad.x, rad.y, rad.z, T);
    if (AUX) {
        const float alpha = 1.f - Tm;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (aw > 1.0e-6f && alpha > 0.f) {
            // the opacity-weighted mean point and velocity; that point was at qm - v dt one sim step
            // ago, seen through the previous camera; the current projection is the pixel centre
            const float dep = ad / aw;
            const float3 qm = add(o, mul(d, dep));
            const float3 qp = sub(qm, mul(av, x.vel_dt / aw));
            const float2 cur = make_float2(((float)c + 0.5f) * x.out_w / (float)w, ((float)r + 0.5f) * x.out_h / (float)h);
            const float2 pp = project_px(qp, x.cam_prev, x.out_w, x.out_h);
            v = make_float4(alpha, dep * alpha, (cur.x - pp.x) * alpha, (cur.y - pp.y) * alpha);
        }
        x.aux[(size_t)r * w + c] = v;
    }
}

// (D, H, W) velocity components (3, D, H, W) float -> (D, H, W, 4) half, 4th channel 0
__global__ void __launch_bounds__(256) k_pack_velocity(const float* __restrict__ vel, size_t n, __half* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[4 * i + 0] = __float2half_rn(vel[i]);
    out[4 * i + 1] = __float2half_rn(vel[n + i]);
    out[4 * i + 2] = __float2half_rn(vel[2 * n + i]);
    out[4 * i + 3] = __float2half_rn(0.f);
}

// aux_resolve (kernels.py:603-626): march-resolution premultiplied aux -> (4, Ho, Wo) half alpha,
// depth, motion_x, motion_y (bilinear upsample, then un-premultiplied; 0 where alpha < a_min)
__global__ void __launch_bounds__(256) k_aux_resolve(const float4* __restrict__ aux, int mh, int mw, int Ho, int Wo,
                                                     float a_min, __half* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y * blockDim.y + threadIdx.y;
    if (c >= Wo || r >= Ho) return;
    float4 v;
    if (mh == Ho && mw == Wo) {
        v = aux[(size_t)r * mw + c];
    } else {
        const float u = ((float)c + 0.5f) / (float)Wo, vv = ((float)r + 0.5f) / (float)Ho;
        const float fx = u * (float)mw - 0.5f, fy = vv * (float)mh - 0.5f;
        const float x0f = floorf(fx), y0f = floorf(fy), tx = fx - x0f, ty = fy - y0f;
        const int x0 = min(max((int)x0f, 0), mw - 1), y0 = min(max((int)y0f, 0), mh - 1);
        const int x1 = min(max((int)x0f + 1, 0), mw - 1), y1 = min(max((int)y0f + 1, 0), mh - 1);
        const float4 p00 = aux[(size_t)y0 * mw + x0], p01 = aux[(size_t)y0 * mw + x1];
        const float4 p10 = aux[(size_t)y1 * mw + x0], p11 = aux[(size_t)y1 * mw + x1];
        auto l = [&](float a0, float a1, float b0, float b1) {
            return (a0 * (1.f - tx) + a1 * tx) * (1.f - ty) + (b0 * (1.f - tx) + b1 * tx) * ty;
        };
        v = make_float4(l(p00.x, p01.x, p10.x, p11.x), l(p00.y, p01.y, p10.y, p11.y), l(p00.z, p01.z, p10.z, p11.z),
                        l(p00.w, p01.w, p10.w, p11.w));
    }
    const float al = v.x;
    float dep = 0.f, mx = 0.f, my = 0.f;
    if (al >= a_min) {
        dep = v.y / al;
        mx = v.z / al;
        my = v.w / al;
    }
    const size_t n = (size_t)Ho * Wo, i = (size_t)r * Wo + c;
    out[i] = __float2half_rn(al);
    out[n + i] = __float2half_rn(dep);
    out[2 * n + i] = __float2half_rn(mx);
    out[3 * n + i] = __float2half_rn(my);
}

// ---------------------------------------------------------------- bloom and composite
// relu(hdr - thresh) adaptive-average-pooled to (3, bh, bw)
__global__ void __launch_bounds__(256) k_bloom_pool(const float4* __restrict__ hdr, int mh, int mw, float thresh,
                                                    int bh, int bw, float* __restrict__ out) {
    const int n = bh * bw;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = i % bw, y = i / bw;
    const int x0 = x * mw / bw, x1 = ((x + 1) * mw + bw - 1) / bw;
    const int y0 = y * mh / bh, y1 = ((y + 1) * mh + bh - 1) / bh;
    float s[3] = {0.f, 0.f, 0.f};
    for (int yy = y0; yy < y1; ++yy)
        for (int xx = x0; xx < x1; ++xx) {
            const float4 v = hdr[(size_t)yy * mw + xx];
            s[0] += fmaxf(v.x - thresh, 0.f);
            s[1] += fmaxf(v.y - thresh, 0.f);
            s[2] += fmaxf(v.z - thresh, 0.f);
        }
    const float cnt = (float)((y1 - y0) * (x1 - x0));
    for (int c = 0; c < 3; ++c) out[(size_t)c * n + i] = s[c] / cnt;
}

// avg_pool2d(4, ceil_mode=True): the last window per axis may be partial and averages what it covers
__global__ void __launch_bounds__(256) k_bloom_pool4(const float* __restrict__ in, int bh, int bw, int ph, int pw,
                                                     float* __restrict__ out) {
    const int n = 3 * ph * pw;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = i % pw, y = (i / pw) % ph, c = i / (ph * pw);
    const int x0 = x * 4, x1 = min(x0 + 4, bw), y0 = y * 4, y1 = min(y0 + 4, bh);
    float s = 0.f;
    for (int yy = y0; yy < y1; ++yy)
        for (int xx = x0; xx < x1; ++xx) s += in[((size_t)c * bh + yy) * bw + xx];
    out[i] = s / (float)((y1 - y0) * (x1 - x0));
}

// bloom = b1 + 1.5 bilinear(b2 -> (bh, bw)), stored (bh, bw, 3)
__global__ void __launch_bounds__(256) k_bloom_combine(const float* __restrict__ b1, const float* __restrict__ b2,
                                                       int bh, int bw, int ph, int pw, float* __restrict__ bloom) {
    const int n = bh * bw;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = i % bw, y = i / bw;
    const Lin lx = lin(x, pw, (float)pw / (float)bw), ly = lin(y, ph, (float)ph / (float)bh);
    for (int c = 0; c < 3; ++c) {
        const float* g = b2 + (size_t)c * ph * pw;
        const float up = ly.l0 * (lx.l0 * g[ly.i0 * pw + lx.i0] + lx.l1 * g[ly.i0 * pw + lx.i1]) +
                         ly.l1 * (lx.l0 * g[ly.i1 * pw + lx.i0] + lx.l1 * g[ly.i1 * pw + lx.i1]);
        bloom[(size_t)i * 3 + c] = b1[(size_t)c * n + i] + 1.5f * up;
    }
}

__device__ __forceinline__ float4 v4(const float4* a, int W, int y, int x) { return a[(size_t)y * W + x]; }

// Catmull-Rom upsample of the march buffer clamped to the 2x2 bilinear footprint (kernels.py:512-549)
__device__ float4 sample_cr(const float4* __restrict__ a, int H, int W, float u, float v) {
    const float fx = u * (float)W - 0.5f, fy = v * (float)H - 0.5f;
    const float x0f = floorf(fx), y0f = floorf(fy);
    auto crw = [](float t, float* w) {
        const float t2 = t * t, t3 = t2 * t;
        w[0] = -0.5f * t3 + t2 - 0.5f * t;
        w[1] = 1.5f * t3 - 2.5f * t2 + 1.f;
        w[2] = -1.5f * t3 + 2.f * t2 + 0.5f * t;
        w[3] = 0.5f * t3 - 0.5f * t2;
    };
    float wx[4], wy[4];
    crw(fx - x0f, wx);
    crw(fy - y0f, wy);
    const int xi = (int)x0f, yi = (int)y0f;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 lo = make_float4(1.0e9f, 1.0e9f, 1.0e9f, 1.0e9f), hi = make_float4(-1.0e9f, -1.0e9f, -1.0e9f, -1.0e9f);
    for (int j = 0; j < 4; ++j) {
        const int yy = min(max(yi - 1 + j, 0), H - 1);
        float4 row = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int i = 0; i < 4; ++i) {
            const int xx = min(max(xi - 1 + i, 0), W - 1);
            const float4 s = v4(a, W, yy, xx);
            row.x = row.x + s.x * wx[i];
            row.y = row.y + s.y * wx[i];
            row.z = row.z + s.z * wx[i];
            row.w = row.w + s.w * wx[i];
            if ((i == 1 || i == 2) && (j == 1 || j == 2)) {
                lo = make_float4(fminf(lo.x, s.x), fminf(lo.y, s.y), fminf(lo.z, s.z), fminf(lo.w, s.w));
                hi = make_float4(fmaxf(hi.x, s.x), fmaxf(hi.y, s.y), fmaxf(hi.z, s.z), fmaxf(hi.w, s.w));
            }
        }
        acc.x = acc.x + row.x * wy[j];
        acc.y = acc.y + row.y * wy[j];
        acc.z = acc.z + row.z * wy[j];
        acc.w = acc.w + row.w * wy[j];
    }
    return make_float4(clampf(acc.x, lo.x, hi.x), clampf(acc.y, lo.y, hi.y), clampf(acc.z, lo.z, hi.z),
                       clampf(acc.w, lo.w, hi.w));
}

__device__ float3 sample_b3(const float* __restrict__ a, int H, int W, float u, float v) {
    const float fx = u * (float)W - 0.5f, fy = v * (float)H - 0.5f;
    const float x0f = floorf(fx), y0f = floorf(fy);
    const float tx = fx - x0f, ty = fy - y0f;
    const int x0 = min(max((int)x0f, 0), W - 1), y0 = min(max((int)y0f, 0), H - 1);
    const int x1 = min(max((int)x0f + 1, 0), W - 1), y1 = min(max((int)y0f + 1, 0), H - 1);
    float r[3];
    for (int c = 0; c < 3; ++c) {
        const float a0 = a[((size_t)y0 * W + x0) * 3 + c] * (1.f - tx) + a[((size_t)y0 * W + x1) * 3 + c] * tx;
        const float a1 = a[((size_t)y1 * W + x0) * 3 + c] * (1.f - tx) + a[((size_t)y1 * W + x1) * 3 + c] * tx;
        r[c] = a0 * (1.f - ty) + a1 * ty;
    }
    return f3(r[0], r[1], r[2]);
}

__device__ __forceinline__ float aces(float x) {
    return clampf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.f, 1.f);
}

__device__ __forceinline__ uint8_t to_srgb8(float x, float dither) {
    const float s = x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.f / 2.4f) - 0.055f;
    return (uint8_t)min(max((int)(s * 255.f + 0.5f + dither), 0), 255);
}

// composite (kernels.py:571-600)
__global__ void __launch_bounds__(256) k_composite(const float4* __restrict__ hdr, int mh, int mw,
                                                   const float* __restrict__ bloom, int bh, int bw, Shade p, int Ho,
                                                   int Wo, uint8_t* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y * blockDim.y + threadIdx.y;
    if (c >= Wo || r >= Ho) return;
    const float u = ((float)c + 0.5f) / (float)Wo, v = ((float)r + 0.5f) / (float)Ho;
    const float4 px = (mh == Ho && mw == Wo) ? hdr[(size_t)r * mw + c] : sample_cr(hdr, mh, mw, u, v);
    const float vy = 1.f - v;
    const float vig = 1.f - 0.6f * ((u - 0.5f) * (u - 0.5f) + (vy - 0.6f) * (vy - 0.6f));
    const float3 bg = mul(add(mul(p.bg_bot, 1.f - vy), mul(p.bg_top, vy)), vig);
    float3 col = add(add(f3(px.x, px.y, px.z), mul(bg, px.w)), mul(sample_b3(bloom, bh, bw, u, v), p.bloom));
    col = mul(col, p.exposure);
    if (p.contrast != 1.f) {
        col = f3(0.18f * powf(fmaxf(col.x, 0.f) / 0.18f, p.contrast), 0.18f * powf(fmaxf(col.y, 0.f) / 0.18f, p.contrast),
                 0.18f * powf(fmaxf(col.z, 0.f) / 0.18f, p.contrast));
    }
    const float hh = sinf((float)r * 12.9898f + (float)c * 78.233f) * 43758.5453f;
    const float dither = (hh - floorf(hh)) - 0.5f;
    uint8_t* o = out + ((size_t)r * Wo + c) * 3;
    o[0] = to_srgb8(aces(col.x), dither);
    o[1] = to_srgb8(aces(col.y), dither);
    o[2] = to_srgb8(aces(col.z), dither);
}

// ---------------------------------------------------------------- temporal resolve
// (temporal.py:26-150) The march buffer blended into a history reprojected by camera motion only,
// in (luma-compressed YCoCg, T) space with a 3x3 variance clip of the current frame.

// (rgb, T) -> (Y, Co, Cg, T) scaled by 1 / (1 + max(Y, 0)) so bright fire does not dominate
__device__ __forceinline__ float4 to_c4(float4 v) {
    const float y = 0.25f * v.x + 0.5f * v.y + 0.25f * v.z;
    const float co = 0.5f * v.x - 0.5f * v.z;
    const float cg = -0.25f * v.x + 0.5f * v.y - 0.25f * v.z;
    const float s = 1.f / (1.f + fmaxf(y, 0.f));
    return make_float4(y * s, co * s, cg * s, v.w);
}

__device__ __forceinline__ float4 from_c4(float4 c) {
    const float s = 1.f / fmaxf(1.f - c.x, 1.0e-3f);
    const float y = c.x * s, co = c.y * s, cg = c.z * s;
    return make_float4(y + co - cg, y + cg, y - co - cg, c.w);
}

__device__ __forceinline__ float c4_at(const float4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

// cur: this frame's march buffer; aux: its premultiplied (alpha, alpha depth, ...); hist: the last
// resolved buffer; cam: this draw's camera; tcam: the previous draw's camera with tcam[14] = the
// minimum current weight of this draw and tcam[15] = history valid; out: the resolved buffer
__global__ void __launch_bounds__(256) k_temporal(const float4* __restrict__ cur, const float4* __restrict__ aux,
                                                  const float4* __restrict__ hist, Cam cam, Cam tcam, int H, int W,
                                                  float alpha_base, float gamma, float k_t, float4* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y * blockDim.y + threadIdx.y;
    if (c >= W || r >= H) return;
    const size_t i = (size_t)r * W + c;
    const float4 now = cur[i];
    if (tcam.v[15] < 0.5f) {
        out[i] = now;
        return;
    }
    const float3 o = f3(cam.v[0], cam.v[1], cam.v[2]);
    const float th = cam.v[12], aspect = cam.v[13];
    const float sx = (2.f * ((float)c + 0.5f) / (float)W - 1.f) * th * aspect;
    const float sy = (1.f - 2.f * ((float)r + 0.5f) / (float)H) * th;
    const float3 d = norm3(add(add(f3(cam.v[3], cam.v[4], cam.v[5]), mul(f3(cam.v[6], cam.v[7], cam.v[8]), sx)),
                               mul(f3(cam.v[9], cam.v[10], cam.v[11]), sy)));
    // the reprojected point: the medium's opacity-weighted depth, else the floor plane, else the
    // distance to the domain centre
    const float4 ax = aux[i];
    float dep;
    if (ax.x > 0.02f) dep = ax.y / ax.x;
    else if (d.y < -1.0e-4f && o.y > 0.f) dep = -o.y / d.y;
    else dep = len3(sub(f3(0.5f, 0.5f, 0.5f), o));
    const float3 rel = sub(add(o, mul(d, dep)), f3(tcam.v[0], tcam.v[1], tcam.v[2]));
    const float z = dot3(rel, f3(tcam.v[3], tcam.v[4], tcam.v[5]));
    bool valid = z > 1.0e-3f;
    float px = 0.f, py = 0.f;
    if (valid) {
        const float psx = dot3(rel, f3(tcam.v[6], tcam.v[7], tcam.v[8])) / (z * tcam.v[12] * tcam.v[13]);
        const float psy = dot3(rel, f3(tcam.v[9], tcam.v[10], tcam.v[11])) / (z * tcam.v[12]);
        px = (psx + 1.f) * 0.5f * (float)W;
        py = (1.f - psy) * 0.5f * (float)H;
        valid = px >= 0.f && px <= (float)W && py >= 0.f && py <= (float)H;
    }
    if (!valid) {
        out[i] = now;   // off screen or behind the previous camera: reset
        return;
    }
    const float4 hp = to_c4(sample_cr(hist, H, W, px / (float)W, py / (float)H));
    float4 m1 = make_float4(0.f, 0.f, 0.f, 0.f), m2 = m1;
    float4 lo = make_float4(1.0e9f, 1.0e9f, 1.0e9f, 1.0e9f), hi = make_float4(-1.0e9f, -1.0e9f, -1.0e9f, -1.0e9f);
    for (int j = -1; j <= 1; ++j)
        for (int k = -1; k <= 1; ++k) {
            const float4 sn = to_c4(cur[(size_t)min(max(r + j, 0), H - 1) * W + min(max(c + k, 0), W - 1)]);
            m1 = make_float4(m1.x + sn.x, m1.y + sn.y, m1.z + sn.z, m1.w + sn.w);
            m2 = make_float4(m2.x + sn.x * sn.x, m2.y + sn.y * sn.y, m2.z + sn.z * sn.z, m2.w + sn.w * sn.w);
            lo = make_float4(fminf(lo.x, sn.x), fminf(lo.y, sn.y), fminf(lo.z, sn.z), fminf(lo.w, sn.w));
            hi = make_float4(fmaxf(hi.x, sn.x), fmaxf(hi.y, sn.y), fmaxf(hi.z, sn.z), fmaxf(hi.w, sn.w));
        }
    m1 = make_float4(m1.x / 9.f, m1.y / 9.f, m1.z / 9.f, m1.w / 9.f);
    m2 = make_float4(m2.x / 9.f, m2.y / 9.f, m2.z / 9.f, m2.w / 9.f);
    float cl[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float mk = c4_at(m1, k);
        const float sig = sqrtf(fmaxf(c4_at(m2, k) - mk * mk, 0.f));
        const float a = fmaxf(c4_at(lo, k), mk - gamma * sig);
        const float b = fminf(c4_at(hi, k), mk + gamma * sig);
        cl[k] = clampf(c4_at(hp, k), a, b);
    }
    const float4 nc = to_c4(now);
    // trust the history less where the medium's T changed or its neighbourhood is high-contrast in T
    const float dT = fabsf(nc.w - hp.w);
    const float a0 = fmaxf(alpha_base, tcam.v[14]);
    const float wgt = clampf(a0 + k_t * dT + 0.5f * k_t * (hi.w - lo.w) * dT, a0, 1.f);
    const float4 res = make_float4(cl[0] * (1.f - wgt) + nc.x * wgt, cl[1] * (1.f - wgt) + nc.y * wgt,
                                   cl[2] * (1.f - wgt) + nc.z * wgt, cl[3] * (1.f - wgt) + nc.w * wgt);
    out[i] = from_c4(res);
}

static int nblk(size_t n) { return (int)((n + 255) / 256); }

static Taps gauss_taps(double sigma) {
    Taps k{};
    k.r = std::max(1, (int)std::ceil(2.5 * sigma));
    double s = 0.0;
    for (int t = -k.r; t <= k.r; ++t) s += std::exp(-0.5 * (t / sigma) * (t / sigma));
    for (int t = -k.r; t <= k.r; ++t) k.w[t + k.r] = (float)(std::exp(-0.5 * (t / sigma) * (t / sigma)) / s);
    return k;
}

}  // namespace vfx

using namespace vfx;

// Buffers are (re)allocated when a size grows: at creation, on a new output size or march scale,
// light resolution, volume or solid shape.
struct vfx_render {
    int out_h = 0, out_w = 0, mh = 0, mw = 0, bh = 0, bw = 0, ph = 0, pw = 0, L = 0;
    Shade sh{};
    float* lut = nullptr;
    // volume
    int vD = 0, vH = 0, vW = 0;
    __half2* packed = nullptr;
    uint8_t* occ = nullptr;      // BRICKS^3 (volume), then BRICKS^3 (solid)
    // solid
    int sD = 0, sH = 0, sW = 0;
    float* sblur = nullptr;      // 2 sD sH sW floats
    uint8_t* solid = nullptr;
    bool have_solid = false;
    // light and glow
    __half2* light = nullptr;
    __half* glow_src = nullptr;
    float* gwork = nullptr;      // 3 (3 * 32^3) + 3 (3 * 8^3) floats
    __half* glow = nullptr;      // 32^3 * 4
    // view
    float4* hdr = nullptr;
    float* bwork = nullptr;      // 2 (3 bh bw) + 2 (3 ph pw) floats
    float* bloom = nullptr;      // bh bw 3
    // aux (return_aux): velocity grid and the march-resolution premultiplied aux buffer
    int uD = 0, uH = 0, uW = 0;
    __half* vel = nullptr;       // uD uH uW 4
    float4* aux = nullptr;       // aux_m float4
    size_t aux_m = 0;
    // temporal resolve: parameters, frame counter, the previous draw's camera, history buffers
    bool t_on = false;
    float t_alpha = 0.1f, t_gamma = 1.f, t_kT = 4.f, t_alpha_tick = 0.5f;
    long long t_frame = 0;
    bool t_prev_valid = false;
    float t_prev[16] = {};
    float4* hist = nullptr;      // hist_m float4 each: the last resolved buffer and the next one
    float4* tres = nullptr;
    size_t hist_m = 0;
    const float4* view = nullptr;   // the buffer bloom and composite read (hdr, or the resolved one)
};

static bool ralloc(void** p, size_t bytes) {
    if (*p) {
        (void)hipDeviceSynchronize();
        (void)hipFree(*p);
        *p = nullptr;
    }
    return hip_ok(hipMalloc(p, bytes ? bytes : 4), "render buffers");
}

static void unpack_params(vfx_render* r, const float* v) {
    Shade& s = r->sh;
    int i = 0;
    s.steps = v[i++];
    r->L = (int)v[i++];
    s.k_abs = v[i++]; s.k_shadow = v[i++]; s.k_solid = v[i++]; s.k_amb = v[i++]; s.k_emit = v[i++];
    s.emit_min = v[i++]; s.emit_scale = v[i++]; s.soot = v[i++]; s.fire_light = v[i++]; s.phase_g = v[i++];
    s.rim = v[i++]; s.ms_a = v[i++]; s.ms_b = v[i++]; s.ms_c = v[i++]; s.exposure = v[i++]; s.contrast = v[i++];
    s.bloom = v[i++]; s.bloom_thresh = v[i++]; s.t_lo = v[i++]; s.t_hi = v[i++];
    i++;   // t_ref: the blackbody table's reference temperature, used on the host
    s.d_eps = v[i++];
    const double lx = v[i], ly = v[i + 1], lz = v[i + 2], ll = std::sqrt(lx * lx + ly * ly + lz * lz);
    s.light_dir = make_float3((float)(lx / ll), (float)(ly / ll), (float)(lz / ll));
    i += 3;
    auto col = [&](float3& c) { c = make_float3(v[i], v[i + 1], v[i + 2]); i += 3; };
    col(s.key_col); col(s.amb_col); col(s.smoke_albedo); col(s.bg_top); col(s.bg_bot); col(s.floor_col);
    col(s.grid_col); col(s.obst_col);
    s.glow_coarse = v[i++];
    s.floor = v[i++];
}

static int alloc_light(vfx_render* r) {
    const size_t L3 = (size_t)r->L * r->L * r->L, g3 = (size_t)GLOW * GLOW * GLOW, g8 = g3 / 64;
    if (!ralloc((void**)&r->light, L3 * sizeof(__half2)) || !ralloc((void**)&r->glow_src, L3 * 4 * sizeof(__half)) ||
        !ralloc((void**)&r->gwork, (9 * g3 + 9 * g8) * sizeof(float)) ||
        !ralloc((void**)&r->glow, g3 * 4 * sizeof(__half)))
        return VFX_E_HIP;
    return VFX_OK;
}

static int alloc_output(vfx_render* r, int out_h, int out_w, float scale) {
    r->out_h = out_h;
    r->out_w = out_w;
    r->mh = std::max(8, (int)std::lround(out_h * (double)scale));
    r->mw = std::max(8, (int)std::lround(out_w * (double)scale));
    r->bh = std::max(2, r->mh / 4);
    r->bw = std::max(2, r->mw / 4);
    r->ph = (r->bh + 3) / 4;
    r->pw = (r->bw + 3) / 4;
    const size_t m = (size_t)r->mh * r->mw, b = (size_t)r->bh * r->bw, q = (size_t)r->ph * r->pw;
    if (!ralloc((void**)&r->hdr, m * sizeof(float4)) || !ralloc((void**)&r->bwork, (6 * b + 6 * q) * sizeof(float)) ||
        !ralloc((void**)&r->bloom, 3 * b * sizeof(float)))
        return VFX_E_HIP;
    r->view = r->hdr;
    r->t_prev_valid = false;   // the history is invalid after a resize
    return VFX_OK;
}

static void free_all(vfx_render* r) {
    (void)hipDeviceSynchronize();
    void* ps[] = {r->lut, r->packed, r->occ, r->sblur, r->solid, r->light, r->glow_src, r->gwork, r->glow, r->hdr,
                  r->bwork, r->bloom, r->vel, r->aux, r->hist, r->tres};
    for (void* p : ps)
        if (p) (void)hipFree(p);
}

extern "C" {

int vfx_render_create(int out_h, int out_w, float march_scale, const float* params, const float* lut,
                      vfx_render** out) {
    if (!out || !params || !lut) return arg_error("null pointer");
    *out = nullptr;
    if (out_h < 1 || out_w < 1 || !(march_scale > 0.f)) return arg_error("bad output size or march scale");
    vfx_render* r = new (std::nothrow) vfx_render;
    if (!r) return arg_error("out of host memory");
    unpack_params(r, params);
    if (r->L < 2 || r->L > 512) {
        delete r;
        return arg_error("light_res must be in [2, 512]");
    }
    int rc = VFX_OK;
    if (!ralloc((void**)&r->lut, LUT_N * 3 * sizeof(float)) ||
        !ralloc((void**)&r->occ, 2 * BRICKS * BRICKS * BRICKS) ||
        !hip_ok(hipMemcpy(r->lut, lut, LUT_N * 3 * sizeof(float), hipMemcpyHostToDevice), "render lut"))
        rc = VFX_E_HIP;
    if (rc == VFX_OK) rc = alloc_light(r);
    if (rc == VFX_OK) rc = alloc_output(r, out_h, out_w, march_scale);
    if (rc != VFX_OK) {
        free_all(r);
        delete r;
        return rc;
    }
    *out = r;
    return VFX_OK;
}

int vfx_render_destroy(vfx_render* r) {
    if (!r) return VFX_OK;
    free_all(r);
    delete r;
    return VFX_OK;
}

int vfx_render_set_output(vfx_render* r, int out_h, int out_w, float march_scale) {
    if (!r) return arg_error("null renderer");
    if (out_h < 1 || out_w < 1 || !(march_scale > 0.f)) return arg_error("bad output size or march scale");
    return alloc_output(r, out_h, out_w, march_scale);
}

int vfx_render_set_params(vfx_render* r, const float* params, const float* lut) {
    if (!r || !params || !lut) return arg_error("null pointer");
    const int old_L = r->L;
    unpack_params(r, params);
    if (r->L < 2 || r->L > 512) return arg_error("light_res must be in [2, 512]");
    if (!hip_ok(hipMemcpy(r->lut, lut, LUT_N * 3 * sizeof(float), hipMemcpyHostToDevice), "render lut"))
        return VFX_E_HIP;
    if (r->L != old_L) return alloc_light(r);
    return VFX_OK;
}

}  // extern "C"

// the aux request of one draw (return_aux): velocity (3, uD, uH, uW) float device or NULL, re-packed
// when upload_vel; the previous camera (16 host floats); the sim step; the (4, out_h, out_w) half output
struct AuxReq {
    const float* vel;
    int uD, uH, uW, upload_vel;
    const float* cam_prev;
    float vel_dt;
    __half* out;
};

static int draw_impl(vfx_render* r, const float* dens, const float* temp, int D, int H, int W, const float* solid,
                     int sD, int sH, int sW, int upload_solid, const float* cam, int reuse_volume, uint8_t* out,
                     hipStream_t s, float* ms_prep, float* ms_view, const AuxReq* aq) {
    if (!r || !cam || !out) return arg_error("null pointer");
    if (aq && (!aq->out || !aq->cam_prev)) return arg_error("null aux pointer");
    // argument checks before any state changes
    if (!reuse_volume) {
        if (!dens || !temp || D < 2 || H < 2 || W < 2) return arg_error("volume must be (D, H, W) with sizes >= 2");
        if (solid && (sD < 2 || sH < 2 || sW < 2)) return arg_error("solid must be (D, H, W) with sizes >= 2");
    } else if (r->vD == 0) {
        return arg_error("reuse_volume without a prepared volume");
    }
    if (aq && aq->vel && (aq->uD < 2 || aq->uH < 2 || aq->uW < 2))
        return arg_error("velocity must be (3, D, H, W) with sizes >= 2");
    const bool temporal = r->t_on;
    const bool march_aux = aq || temporal;   // the temporal resolve needs the march's opacity-weighted depth
    const size_t m = (size_t)r->mh * r->mw;
    if (march_aux && r->aux_m != m) {
        if (!ralloc((void**)&r->aux, m * sizeof(float4))) return VFX_E_HIP;
        r->aux_m = m;
    }
    if (temporal && r->hist_m != m) {
        if (!ralloc((void**)&r->hist, m * sizeof(float4)) || !ralloc((void**)&r->tres, m * sizeof(float4)))
            return VFX_E_HIP;
        r->hist_m = m;
        r->t_prev_valid = false;
    }
    Cam c;
    for (int i = 0; i < 16; ++i) c.v[i] = cam[i];
    Cam tc{};
    if (temporal) {
        // per-frame offset of the ray-start jitter (golden-ratio sequence), and the previous draw's
        // camera with this draw's minimum current weight and the history-valid flag
        ++r->t_frame;
        c.v[14] = (float)std::fmod((double)r->t_frame * 0.6180339887498949, 1.0);
        if (r->t_prev_valid) {
            for (int i = 0; i < 16; ++i) tc.v[i] = r->t_prev[i];
            tc.v[14] = reuse_volume ? 0.f : r->t_alpha_tick;
            tc.v[15] = 1.f;
        }
        for (int i = 0; i < 16; ++i) r->t_prev[i] = c.v[i];
        r->t_prev_valid = true;
    }
    AuxArgs x{};
    if (march_aux) {
        for (int i = 0; i < 16; ++i) x.cam_prev.v[i] = aq ? aq->cam_prev[i] : c.v[i];
        x.vel_dt = aq ? aq->vel_dt : 0.f;
        x.out_w = (float)r->out_w;
        x.out_h = (float)r->out_h;
        x.aux = r->aux;
    }
    if (aq) {
        if (aq->vel) {
            const bool shape_new = aq->uD != r->uD || aq->uH != r->uH || aq->uW != r->uW;
            const size_t un = (size_t)aq->uD * aq->uH * aq->uW;
            if (shape_new) {
                if (!ralloc((void**)&r->vel, un * 4 * sizeof(__half))) return VFX_E_HIP;
                r->uD = aq->uD; r->uH = aq->uH; r->uW = aq->uW;
            }
            if (shape_new || aq->upload_vel)
                hipLaunchKernelGGL(k_pack_velocity, dim3(nblk(un)), dim3(256), 0, s, aq->vel, un, r->vel);
        }
        x.vel = r->vel;
        x.gvel = Tex3{r->uD, r->uH, r->uW};
        x.has_vel = aq->vel ? 1 : 0;
    }
    const bool has_solid = solid != nullptr;
    // timing events, destroyed on every return
    struct Events {
        hipEvent_t e[3] = {};
        ~Events() {
            for (auto& x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } evs;
    hipEvent_t* ev = evs.e;
    const bool timed = ms_prep || ms_view;
    if (timed) {
        for (int i = 0; i < 3; ++i)
            if (!hip_ok(hipEventCreate(&ev[i]), "render events")) return VFX_E_HIP;
        (void)hipEventRecord(ev[0], s);
    }
    uint8_t* occ_v = r->occ;
    uint8_t* occ_s = r->occ + BRICKS * BRICKS * BRICKS;
    const int B3 = BRICKS * BRICKS * BRICKS;
    if (!reuse_volume) {
        const size_t n = (size_t)D * H * W;
        if (D != r->vD || H != r->vH || W != r->vW) {
            if (!ralloc((void**)&r->packed, n * sizeof(__half2))) return VFX_E_HIP;
            r->vD = D; r->vH = H; r->vW = W;
        }
        if (has_solid && (upload_solid || sD != r->sD || sH != r->sH || sW != r->sW || !r->have_solid)) {
            const size_t sn = (size_t)sD * sH * sW;
            if (sD != r->sD || sH != r->sH || sW != r->sW) {
                if (!ralloc((void**)&r->sblur, 2 * sn * sizeof(float)) || !ralloc((void**)&r->solid, sn))
                    return VFX_E_HIP;
                r->sD = sD; r->sH = sH; r->sW = sW;
            }
            // smooth the binary mask (Gaussian sigma 1, replicate borders) so the 0.5 iso-surface and
            // its gradient normals are not stair-stepped (renderer.py:227-229)
            const Taps k = gauss_taps(1.0);
            float* a = r->sblur;
            float* b = r->sblur + sn;
            hipLaunchKernelGGL(k_blur_axis, dim3(nblk(sn)), dim3(256), 0, s, solid, a, 1, sD, sH, sW, 2, k);
            hipLaunchKernelGGL(k_blur_axis, dim3(nblk(sn)), dim3(256), 0, s, a, b, 1, sD, sH, sW, 1, k);
            hipLaunchKernelGGL(k_blur_axis, dim3(nblk(sn)), dim3(256), 0, s, b, a, 1, sD, sH, sW, 0, k);
            hipLaunchKernelGGL(k_fill_u8, dim3(nblk(B3)), dim3(256), 0, s, occ_s, (size_t)B3);
            hipLaunchKernelGGL(k_pack_solid, dim3(nblk(sn)), dim3(256), 0, s, a, sD, sH, sW, r->solid, occ_s);
            r->have_solid = true;
        }
        hipLaunchKernelGGL(k_fill_u8, dim3(nblk(B3)), dim3(256), 0, s, occ_v, (size_t)B3);
        hipLaunchKernelGGL(k_pack_volume, dim3(nblk(n)), dim3(256), 0, s, dens, temp, D, H, W, r->packed, occ_v,
                           r->sh.d_eps, r->sh.t_lo);
        // light volume and glow pyramid (renderer.py:397-417)
        const int L = r->L;
        const size_t L3 = (size_t)L * L * L, g3 = (size_t)GLOW * GLOW * GLOW, g8 = g3 / 64;
        hipLaunchKernelGGL(k_light_volume, dim3(nblk(L3)), dim3(256), 0, s, r->packed, Tex3{D, H, W}, r->solid,
                           Tex3{r->sD, r->sH, r->sW}, has_solid ? 1 : 0, r->lut, r->sh, L, r->light, r->glow_src);
        float* g32 = r->gwork;
        float* g32b = g32 + 3 * g3;
        float* gt = g32b + 3 * g3;
        float* g8a = gt + 3 * g3;
        float* g8b = g8a + 3 * g8;
        float* g8t = g8b + 3 * g8;
        const Taps k3 = gauss_taps(1.2);
        hipLaunchKernelGGL(k_glow_pool, dim3(nblk(g3)), dim3(256), 0, s, r->glow_src, L, g32);
        hipLaunchKernelGGL(k_blur_axis, dim3(nblk(3 * g3)), dim3(256), 0, s, g32, gt, 3, GLOW, GLOW, GLOW, 2, k3);
        hipLaunchKernelGGL(k_blur_axis, dim3(nblk(3 * g3)), dim3(256), 0, s, gt, g32b, 3, GLOW, GLOW, GLOW, 1, k3);
        hipLaunchKernelGGL(k_blur_axis, dim3(nblk(3 * g3)), dim3(256), 0, s, g32b, gt, 3, GLOW, GLOW, GLOW, 0, k3);
        hipLaunchKernelGGL(k_glow_pool4, dim3(nblk(3 * g8)), dim3(256), 0, s, g32, g8a);
        const int G8 = GLOW / 4;
        hipLaunchKernelGGL(k_blur_axis, dim3(nblk(3 * g8)), dim3(256), 0, s, g8a, g8t, 3, G8, G8, G8, 2, k3);
        hipLaunchKernelGGL(k_blur_axis, dim3(nblk(3 * g8)), dim3(256), 0, s, g8t, g8b, 3, G8, G8, G8, 1, k3);
        hipLaunchKernelGGL(k_blur_axis, dim3(nblk(3 * g8)), dim3(256), 0, s, g8b, g8t, 3, G8, G8, G8, 0, k3);
        hipLaunchKernelGGL(k_glow_combine, dim3(nblk(g3)), dim3(256), 0, s, gt, g8t, r->sh.glow_coarse, r->glow);
    }
    if (timed) (void)hipEventRecord(ev[1], s);
    // view
    MarchArgs a;
    a.vol = r->packed;
    a.gv = Tex3{r->vD, r->vH, r->vW};
    a.light = r->light;
    a.gl = Tex3{r->L, r->L, r->L};
    a.glow = r->glow;
    a.solid = r->solid;
    a.gs = Tex3{r->sD, r->sH, r->sW};
    a.occ = occ_v;
    a.occ_s = occ_s;
    a.has_solid = has_solid && r->have_solid ? 1 : 0;
    a.solid_res = (float)std::max(r->sD, std::max(r->sH, r->sW));
    a.lut = r->lut;
    const dim3 blk(64, 4);
    if (march_aux)
        hipLaunchKernelGGL(k_march<true>, dim3((r->mw + 63) / 64, (r->mh + 3) / 4), blk, 0, s, a, r->sh, c, r->mh,
                           r->mw, r->hdr, x);
    else
        hipLaunchKernelGGL(k_march<false>, dim3((r->mw + 63) / 64, (r->mh + 3) / 4), blk, 0, s, a, r->sh, c, r->mh,
                           r->mw, r->hdr, x);
    r->view = r->hdr;
    if (temporal) {
        hipLaunchKernelGGL(k_temporal, dim3((r->mw + 63) / 64, (r->mh + 3) / 4), blk, 0, s, r->hdr, r->aux, r->hist, c,
                           tc, r->mh, r->mw, r->t_alpha, r->t_gamma, r->t_kT, r->tres);
        std::swap(r->hist, r->tres);   // the resolved buffer is the next draw's history
        r->view = r->hist;
    }
    const size_t b = (size_t)r->bh * r->bw, q = (size_t)r->ph * r->pw;
    float* x0 = r->bwork;
    float* x1 = x0 + 3 * b;
    float* y0 = x1 + 3 * b;
    float* y1 = y0 + 3 * q;
    const Taps k2 = gauss_taps(2.0);
    hipLaunchKernelGGL(k_bloom_pool, dim3(nblk(b)), dim3(256), 0, s, r->view, r->mh, r->mw, r->sh.bloom_thresh, r->bh,
                       r->bw, x0);
    hipLaunchKernelGGL(k_blur_axis, dim3(nblk(3 * b)), dim3(256), 0, s, x0, x1, 3, 1, r->bh, r->bw, 0, k2);
    hipLaunchKernelGGL(k_blur_axis, dim3(nblk(3 * b)), dim3(256), 0, s, x1, x0, 3, 1, r->bh, r->bw, 1, k2);
    hipLaunchKernelGGL(k_bloom_pool4, dim3(nblk(3 * q)), dim3(256), 0, s, x0, r->bh, r->bw, r->ph, r->pw, y0);
    hipLaunchKernelGGL(k_blur_axis, dim3(nblk(3 * q)), dim3(256), 0, s, y0, y1, 3, 1, r->ph, r->pw, 0, k2);
    hipLaunchKernelGGL(k_blur_axis, dim3(nblk(3 * q)), dim3(256), 0, s, y1, y0, 3, 1, r->ph, r->pw, 1, k2);
    hipLaunchKernelGGL(k_bloom_combine, dim3(nblk(b)), dim3(256), 0, s, x0, y0, r->bh, r->bw, r->ph, r->pw, r->bloom);
    hipLaunchKernelGGL(k_composite, dim3((r->out_w + 63) / 64, (r->out_h + 3) / 4), blk, 0, s, r->view, r->mh, r->mw,
                       r->bloom, r->bh, r->bw, r->sh, r->out_h, r->out_w, out);
    if (aq)
        hipLaunchKernelGGL(k_aux_resolve, dim3((r->out_w + 63) / 64, (r->out_h + 3) / 4), blk, 0, s, r->aux, r->mh,
                           r->mw, r->out_h, r->out_w, 1.0e-3f, aq->out);
    if (!launch_ok("render")) return VFX_E_HIP;
    if (timed) {
        (void)hipEventRecord(ev[2], s);
        (void)hipEventSynchronize(ev[2]);
        float a0 = 0.f, a1 = 0.f;
        (void)hipEventElapsedTime(&a0, ev[0], ev[1]);
        (void)hipEventElapsedTime(&a1, ev[1], ev[2]);
        if (ms_prep) *ms_prep = a0;
        if (ms_view) *ms_view = a1;
    }
    return VFX_OK;
}

extern "C" {

// Volume.draw (renderer.py:281-386), colour only.  dens / temp: (D, H, W) float32 device; solid:
// (sD, sH, sW) float32 device or NULL; upload_solid: the solid changed since the last draw (the
// caller keys it like the reference, renderer.py:224-226); cam: 16 host floats (camera_basis);
// reuse_volume: skip the volume prep and re-render the last prepared volume; out: (out_h, out_w, 3)
// uint8 device.  ms_prep / ms_view (optional, host): device times of the two parts (synchronises).
int vfx_render_draw(vfx_render* r, const float* dens, const float* temp, int D, int H, int W, const float* solid,
                    int sD, int sH, int sW, int upload_solid, const float* cam, int reuse_volume, uint8_t* out,
                    void* stream, float* ms_prep, float* ms_view) {
    return draw_impl(r, dens, temp, D, H, W, solid, sD, sH, sW, upload_solid, cam, reuse_volume, out,
                     reinterpret_cast<hipStream_t>(stream), ms_prep, ms_view, nullptr);
}

// Volume.draw(return_aux=True) (renderer.py:295-393, kernels.py:288-489, :603-626): the same frame
// as vfx_render_draw, plus aux_out (4, out_h, out_w) half device = alpha, depth, motion_x, motion_y.
// vel: (3, uD, uH, uW) float32 device in m/s or NULL (zero motion from the medium); upload_vel: the
// velocity changed since the last aux draw; cam_prev: 16 host floats, the previous sim step's camera
// (the current one when the caller has none); vel_dt: the sim step in seconds.
int vfx_render_draw_aux(vfx_render* r, const float* dens, const float* temp, int D, int H, int W, const float* solid,
                        int sD, int sH, int sW, int upload_solid, const float* cam, int reuse_volume, uint8_t* out,
                        const float* vel, int uD, int uH, int uW, int upload_vel, const float* cam_prev, float vel_dt,
                        uint16_t* aux_out, void* stream, float* ms_prep, float* ms_view) {
    const AuxReq aq{vel, uD, uH, uW, upload_vel, cam_prev, vel_dt, reinterpret_cast<__half*>(aux_out)};
    return draw_impl(r, dens, temp, D, H, W, solid, sD, sH, sW, upload_solid, cam, reuse_volume, out,
                     reinterpret_cast<hipStream_t>(stream), ms_prep, ms_view, &aq);
}

// Temporal amortization (renderer.py:130-163, temporal.py): on != 0 blends each draw's march into
// a history reprojected by camera motion (alpha: base current weight, gamma: clip width, k_t: gain
// of the T change, alpha_tick: current weight on draws with a new volume).  Does not reset the
// history; vfx_render_reset_history does.
int vfx_render_set_temporal(vfx_render* r, int on, float alpha, float gamma, float k_t, float alpha_tick) {
    if (!r) return arg_error("null renderer");
    r->t_on = on != 0;
    r->t_alpha = alpha;
    r->t_gamma = gamma;
    r->t_kT = k_t;
    r->t_alpha_tick = alpha_tick;
    return VFX_OK;
}

int vfx_render_reset_history(vfx_render* r) {
    if (!r) return arg_error("null renderer");
    r->t_prev_valid = false;
    return VFX_OK;
}

// The march buffer of the last draw, after the temporal resolve when that is on: (mh, mw, 4)
// float premultiplied radiance + transmittance, copied to dst (device); its size in *mh, *mw.
int vfx_render_get_hdr(vfx_render* r, float* dst, int* mh, int* mw, void* stream) {
    if (!r) return arg_error("null renderer");
    if (mh) *mh = r->mh;
    if (mw) *mw = r->mw;
    if (!dst) return VFX_OK;
    if (!hip_ok(hipMemcpyAsync(dst, r->view, (size_t)r->mh * r->mw * sizeof(float4), hipMemcpyDeviceToDevice,
                               reinterpret_cast<hipStream_t>(stream)),
                "hdr copy"))
        return VFX_E_HIP;
    return VFX_OK;
}

}  // extern "C"

#endif  // !VFX_F64
