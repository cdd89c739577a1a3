// BF16 MFMA kernels of the learned pressure preconditioner (a 3-level 3D U-Net).
//
// Activations are channels-last (voxel-major, channels fastest) BF16 with the channel count
// padded to a multiple of 16; weights are FP32 masters with BF16 copies in GEMM layout
// [Cout_pad][K_pad], k = tap * Cin_pad + ci (tap = kz * 9 + ky * 3 + kx, the cross-correlation
// of torch's conv3d with padding 1), K_pad = 27 * Cin_pad rounded up to 32.  Every convolution
// is an implicit GEMM on v_mfma_f32_16x16x32_bf16 (M = voxels, N = output channels, K = taps x
// input channels) with FP32 accumulation.  Rounding points (where a value is stored as BF16) are
// explicit so that a CPU reference can reproduce them: conv outputs after bias + ReLU, pooled
// values, gradients with respect to stored activations.
#pragma once

#include "sim3d_common.hpp"

namespace vfx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16_t;

constexpr int NET_THREADS = 256;   // 4 waves
constexpr int WG_CHUNK = 32;       // voxels per MFMA K-step of the weight gradient

__device__ __forceinline__ float bf(bf16_t v) { return (float)v; }
__device__ __forceinline__ bf16_t tobf(float v) { return (bf16_t)v; }   // round to nearest even

// tap t (0..26) -> offsets (dz, dy, dx) in {-1, 0, 1}
__device__ __forceinline__ void tap_off(int t, int& dz, int& dy, int& dx) {
    dz = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dx = t % 3 - 1;
}

// 16 zero bytes: the A operand of a tap outside the grid (or past the 27 taps) is loaded from here,
// so every operand load is unconditional and the loads of several K steps are in flight together
__device__ __align__(16) const unsigned int g_net_zero[4] = {0u, 0u, 0u, 0u};

// Forward convolution 3^3, zero padding: out[v][co] = b[co] + sum_{t, ci} in[v + off(t)][ci] w[co][t][ci]
// A wave computes MT tiles of 16 consecutive voxels x CT*16 output channels over one K slice: the
// K steps (32 = taps x input channels) are split over KS waves of the workgroup, whose partial
// sums are added in LDS in wave order (small grids: more waves per voxel, shorter chains of
// dependent MFMAs); MT > 1 reuses each weight fragment for MT voxel tiles (large grids).
// Output: BF16 (channel stride ocs, offset ooff, ReLU if relu) or, if out32, FP32 channel 0.
template <int CIN, int CT, int KS, int MT>
__global__ void __launch_bounds__(NET_THREADS)
k_conv3(int N, const bf16_t* __restrict__ in, const bf16_t* __restrict__ w, int kp, const float* __restrict__ bias,
        int cout, int relu, bf16_t* __restrict__ out, int ocs, int ooff, float* __restrict__ out32,
        const DevState* __restrict__ gst, int gk) {
    constexpr int WAVES = NET_THREADS / 64;
    constexpr int TPB = WAVES / KS;   // wave groups (of KS waves) per workgroup
    __shared__ floatx4 red[KS > 1 ? WAVES * CT * MT * 64 : 1];
    if (gated_off(gst, gk)) return;
    const int n3 = N * N * N;
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    const int ks = wave % KS;
    const int v0 = ((int)blockIdx.x * TPB + wave / KS) * 16 * MT;
    const int row = lane & 15, grp = lane >> 4;
    int x[MT], y[MT], z[MT];
    bool vin[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const int v = v0 + m * 16 + row;
        vin[m] = v < n3;
        const int vc = vin[m] ? v : 0;
        x[m] = vc % N;
        y[m] = (vc / N) % N;
        z[m] = vc / (N * N);
    }
    const bf16_t* zero = reinterpret_cast<const bf16_t*>(g_net_zero);
    floatx4 acc[MT][CT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[m][c] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int nk = kp / 32;
    const int kb = ks * nk / KS, ke = (ks + 1) * nk / KS;
#pragma unroll 4
    for (int kk = kb; kk < ke; ++kk) {
        const int k = kk * 32 + grp * 8;
        const int t = k / CIN, ci = k % CIN;
        int dz, dy, dx;
        tap_off(t, dz, dy, dx);
        bf16x8 a[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int xx = x[m] + dx, yy = y[m] + dy, zz = z[m] + dz;
            const bool ok = vin[m] && t < 27 && (unsigned)xx < (unsigned)N && (unsigned)yy < (unsigned)N &&
                            (unsigned)zz < (unsigned)N;
            const bf16_t* src = ok ? in + (size_t)((zz * N + yy) * N + xx) * CIN + ci : zero;
            a[m] = *reinterpret_cast<const bf16x8*>(src);
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(w + (size_t)(c * 16 + row) * kp + k);
#pragma unroll
            for (int m = 0; m < MT; ++m) acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b, acc[m][c], 0, 0, 0);
        }
    }
    if (KS > 1) {
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int c = 0; c < CT; ++c) red[((wave * MT + m) * CT + c) * 64 + lane] = acc[m][c];
        __syncthreads();
        if (ks != 0) return;
        for (int j = 1; j < KS; ++j)
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int c = 0; c < CT; ++c) {
                    const floatx4 p = red[(((wave + j) * MT + m) * CT + c) * 64 + lane];
                    acc[m][c] = floatx4{acc[m][c][0] + p[0], acc[m][c][1] + p[1], acc[m][c][2] + p[2],
                                        acc[m][c][3] + p[3]};
                }
    }
    // lane holds rows (voxels) grp*4 + i, column (channel) row of each 16-wide tile
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const int co = c * 16 + row;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int vo = v0 + m * 16 + grp * 4 + i;
                if (vo >= n3 || co >= cout) continue;
                float val = acc[m][c][i] + (bias ? bias[co] : 0.0f);
                if (relu) val = fmaxf(val, 0.0f);
                if (out32) {
                    if (co == 0) out32[vo] = val;
                } else {
                    out[(size_t)vo * ocs + ooff + co] = tobf(val);
                }
            }
        }
}

// The same convolution with the input tile in LDS (grids of size >= 48): a workgroup computes a
// CX x CY x CZ voxel tile (4 waves; wave w owns the tile rows y = w, each an M tile of 16 x
// voxels, over the CZ z planes).  The input tile with its 1-voxel halo is staged 16 channels at
// a time (zero outside the grid); a K step is two taps x those 16 channels (k = tap * Cin_pad +
// channel, the weight layout's), so the A operand of every MFMA is one 16-byte LDS read per lane
// and each weight fragment serves the tile's 4 z planes.  Sums: per 16-channel chunk, taps in
// order, FP32 accumulation as in k_conv3.
constexpr int CX = 16, CY = 4, CZ = 4;
constexpr int SX = CX + 2, SY = CY + 2, SZ = CZ + 2, SV = SX * SY * SZ;   // staged voxels
template <int CIN, int CT>
__global__ void __launch_bounds__(NET_THREADS)
k_conv3l(int N, const bf16_t* __restrict__ in, const bf16_t* __restrict__ w, int kp, const float* __restrict__ bias,
         int cout, int relu, bf16_t* __restrict__ out, int ocs, int ooff, float* __restrict__ out32,
         const DevState* __restrict__ gst, int gk) {
    static_assert(CIN % 16 == 0, "16-channel chunks");
    __shared__ __align__(16) bf16_t sx[SV * 16];
    if (gated_off(gst, gk)) return;
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    const int row = lane & 15, grp = lane >> 4;
    const int tx = (N + CX - 1) / CX, ty = (N + CY - 1) / CY;
    const int b = (int)blockIdx.x;
    const int X0 = (b % tx) * CX, Y0 = ((b / tx) % ty) * CY, Z0 = (b / (tx * ty)) * CZ;
    const bf16_t* zero = reinterpret_cast<const bf16_t*>(g_net_zero);
    floatx4 acc[CZ][CT];
#pragma unroll
    for (int m = 0; m < CZ; ++m)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[m][c] = floatx4{0.f, 0.f, 0.f, 0.f};
    // staging: SV voxels x two 16-byte halves of the chunk's 16 channels
    constexpr int SITEMS = SV * 2, SROUNDS = (SITEMS + NET_THREADS - 1) / NET_THREADS;
    bf16x8 pre[SROUNDS];
    auto load_chunk = [&](int c16) {
#pragma unroll
        for (int j = 0; j < SROUNDS; ++j) {
            const int it = min((int)threadIdx.x + j * NET_THREADS, SITEMS - 1);
            const int v = it >> 1, half = it & 1;
            const int x = X0 - 1 + v % SX, y = Y0 - 1 + (v / SX) % SY, z = Z0 - 1 + v / (SX * SY);
            const bool ok = (unsigned)x < (unsigned)N && (unsigned)y < (unsigned)N && (unsigned)z < (unsigned)N;
            const bf16_t* src = ok ? in + (size_t)((z * N + y) * N + x) * CIN + c16 * 16 + half * 8 : zero;
            pre[j] = *reinterpret_cast<const bf16x8*>(src);
        }
    };
    auto store_chunk = [&]() {
#pragma unroll
        for (int j = 0; j < SROUNDS; ++j) {
            const int it = (int)threadIdx.x + j * NET_THREADS;
            if (j + 1 < SROUNDS || it < SITEMS) *reinterpret_cast<bf16x8*>(sx + it * 8) = pre[j];
        }
    };
    load_chunk(0);
    for (int c16 = 0; c16 < CIN / 16; ++c16) {
        __syncthreads();   // the previous chunk's reads are done
        store_chunk();
        __syncthreads();
        if (c16 + 1 < CIN / 16) load_chunk(c16 + 1);
#pragma unroll 2
        for (int t2 = 0; t2 < 14; ++t2) {
            const int t = 2 * t2 + (grp >> 1);   // this lane group's tap (27: padding)
            int dz, dy, dx;
            tap_off(t, dz, dy, dx);
            const int k = t * CIN + c16 * 16 + (grp & 1) * 8;
            bf16x8 bw[CT];
#pragma unroll
            for (int c = 0; c < CT; ++c)
                bw[c] = *reinterpret_cast<const bf16x8*>(t < 27 ? w + (size_t)(c * 16 + row) * kp + k : zero);
#pragma unroll
            for (int m = 0; m < CZ; ++m) {
                const int sv = ((m + 1 + dz) * SY + (wave + 1 + dy)) * SX + (row + 1 + dx);
                const bf16x8 a = t < 27 ? *reinterpret_cast<const bf16x8*>(sx + sv * 16 + (grp & 1) * 8) : bf16x8{};
#pragma unroll
                for (int c = 0; c < CT; ++c) acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[c], acc[m][c], 0, 0, 0);
            }
        }
    }
    // lane holds rows (x voxels) grp*4 + i, column (channel) row of each 16-wide tile
    const int y = Y0 + wave;
#pragma unroll
    for (int m = 0; m < CZ; ++m) {
        const int z = Z0 + m;
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const int co = c * 16 + row;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int x = X0 + grp * 4 + i;
                if (x >= N || y >= N || z >= N || co >= cout) continue;
                const size_t vo = (size_t)(z * N + y) * N + x;
                float val = acc[m][c][i] + (bias ? bias[co] : 0.0f);
                if (relu) val = fmaxf(val, 0.0f);
                if (out32) {
                    if (co == 0) out32[vo] = val;
                } else {
                    out[vo * ocs + ooff + co] = tobf(val);
                }
            }
        }
    }
}

// Weight gradient of a 3^3 convolution for one tap t = blockIdx.y:
//   part[g][t][co][ci] = sum over the voxels of group g = blockIdx.x of gpre[v][co] in[v + off(t)][ci]
// The group's voxels go through LDS in chunks of 32, transposed so that both MFMA operands read 8
// consecutive voxels of one channel (K = voxels).  Waves split the (co, ci) 16x16 tiles.
template <int CINP, int COUTP>
__global__ void __launch_bounds__(NET_THREADS)
k_conv3_wgrad(int N, const bf16_t* __restrict__ in, const bf16_t* __restrict__ gpre, int vg, float* __restrict__ part) {
    constexpr int TI = COUTP / 16, TJ = CINP / 16, T = TI * TJ, TW = (T + 3) / 4;
    __shared__ __align__(16) bf16_t sg[COUTP][WG_CHUNK];
    __shared__ __align__(16) bf16_t sx[CINP][WG_CHUNK];
    const int n3 = N * N * N;
    const int g = (int)blockIdx.x, t = (int)blockIdx.y;
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6, row = lane & 15, grp = lane >> 4;
    int dz, dy, dx;
    tap_off(t, dz, dy, dx);
    const int vbeg = g * vg, vend = min(vbeg + vg, n3);
    floatx4 acc[TW];
#pragma unroll
    for (int i = 0; i < TW; ++i) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int v0 = vbeg; v0 < vend; v0 += WG_CHUNK) {
        for (int idx = tid; idx < WG_CHUNK * COUTP; idx += NET_THREADS) {
            const int vx = idx / COUTP, co = idx % COUTP, v = v0 + vx;
            sg[co][vx] = v < vend ? gpre[(size_t)v * COUTP + co] : tobf(0.0f);
        }
        for (int idx = tid; idx < WG_CHUNK * CINP; idx += NET_THREADS) {
            const int vx = idx / CINP, ci = idx % CINP, v = v0 + vx;
            bf16_t val = tobf(0.0f);
            if (v < vend) {
                const int x = v % N + dx, y = (v / N) % N + dy, z = v / (N * N) + dz;
                if (x >= 0 && y >= 0 && z >= 0 && x < N && y < N && z < N)
                    val = in[(size_t)((z * N + y) * N + x) * CINP + ci];
            }
            sx[ci][vx] = val;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < TW; ++q) {
            const int ti = wave + 4 * q;
            if (ti < T) {
                const int i = ti / TJ, j = ti % TJ;
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sg[i * 16 + row][grp * 8]);
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sx[j * 16 + row][grp * 8]);
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[q], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < TW; ++q) {
        const int ti = wave + 4 * q;
        if (ti >= T) continue;
        const int i = ti / TJ, j = ti % TJ;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = i * 16 + grp * 4 + r, ci = j * 16 + row;
            part[(((size_t)g * 27 + t) * COUTP + co) * CINP + ci] = acc[q][r];
        }
    }
}

// gw[co][ci][t] (the FP32 parameter layout) = sum over groups (in order) of the partials
__global__ void __launch_bounds__(NET_THREADS)
k_wgrad_reduce(const float* __restrict__ part, int ng, int cinp, int coutp, int cin, int cout, float* __restrict__ gw) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= cout * cin * 27) return;
    const int t = i % 27, ci = (i / 27) % cin, co = i / (27 * cin);
    float s = 0.0f;
    for (int g = 0; g < ng; ++g) s += part[(((size_t)g * 27 + t) * coutp + co) * cinp + ci];
    gw[i] = s;
}

// bias gradient: block partials of sum_v gpre[v][co], then an ordered sum
__global__ void __launch_bounds__(NET_THREADS)
k_bias_part(int n3, const bf16_t* __restrict__ gpre, int coutp, int cout, int vpb, float* __restrict__ part) {
    __shared__ float sh[NET_THREADS];
    const int b = (int)blockIdx.x, tid = (int)threadIdx.x;
    const int v0 = b * vpb, v1 = min(v0 + vpb, n3);
    for (int co = 0; co < cout; ++co) {
        float s = 0.0f;
        for (int v = v0 + tid; v < v1; v += NET_THREADS) s += bf(gpre[(size_t)v * coutp + co]);
        sh[tid] = s;
        __syncthreads();
        for (int w = NET_THREADS / 2; w > 0; w >>= 1) {
            if (tid < w) sh[tid] += sh[tid + w];
            __syncthreads();
        }
        if (tid == 0) part[(size_t)b * cout + co] = sh[0];
        __syncthreads();
    }
}

__global__ void k_bias_reduce(const float* __restrict__ part, int nb, int cout, float* __restrict__ gb) {
    const int co = (int)threadIdx.x;
    if (co >= cout) return;
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s += part[(size_t)b * cout + co];
    gb[co] = s;
}

// GEMM weights from the FP32 parameters w[co][ci][t]:
//   fwd[co][t * cinp + ci] (Cout_pad rows, K_pad columns) and, for the input gradient (the
//   transposed convolution), bwd[ci][t' * coutp + co] = w[co][ci][26 - t'].  Padding is zero.
__global__ void __launch_bounds__(NET_THREADS)
k_pack_weights(const float* __restrict__ w, int cin, int cout, int cinp, int coutp, int kp, int kpt,
               bf16_t* __restrict__ fwd, bf16_t* __restrict__ bwd) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < coutp * kp) {
        const int co = i / kp, k = i % kp, t = k / cinp, ci = k % cinp;
        const float v = (co < cout && ci < cin && t < 27) ? w[((size_t)co * cin + ci) * 27 + t] : 0.0f;
        fwd[i] = tobf(v);
    }
    if (bwd && i < cinp * kpt) {
        const int ci = i / kpt, k = i % kpt, t = k / coutp, co = k % coutp;
        const float v = (co < cout && ci < cin && t < 27) ? w[((size_t)co * cin + ci) * 27 + (26 - t)] : 0.0f;
        bwd[i] = tobf(v);
    }
}

// 2x2x2 average pool of channels [off, off + c) of src (stride cs) at fine size n -> dst (c
// channels); one thread per 8 channels of a coarse voxel (16-byte loads and stores; c, cs and off
// are multiples of 8)
__global__ void __launch_bounds__(NET_THREADS)
k_pool(int n, const bf16_t* __restrict__ src, int cs, int off, int c, bf16_t* __restrict__ dst,
       const DevState* __restrict__ gst, int gk) {
    if (gated_off(gst, gk)) return;
    const int m = n / 2, c8 = c / 8;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)m * m * m * c8) return;
    const int ch = (int)(i % c8) * 8;
    const size_t p = i / c8;
    const int X = (int)(p % m), Y = (int)((p / m) % m), Z = (int)(p / ((size_t)m * m));
    bf16x8 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int x = 2 * X + (k & 1), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z + (k >> 2);
        v[k] = *reinterpret_cast<const bf16x8*>(src + (size_t)((z * n + y) * n + x) * cs + off + ch);
    }
    bf16x8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) s += bf(v[k][e]);
        o[e] = tobf(0.125f * s);
    }
    *reinterpret_cast<bf16x8*>(dst + p * c + ch) = o;
}

// nearest 2x upsample of src (c channels, coarse size m) into channels [off, off + c) of dst
// (stride cs); one thread per 8 channels of a fine voxel
__global__ void __launch_bounds__(NET_THREADS)
k_upsample(int m, const bf16_t* __restrict__ src, int c, bf16_t* __restrict__ dst, int cs, int off,
           const DevState* __restrict__ gst, int gk) {
    if (gated_off(gst, gk)) return;
    const int n = 2 * m, c8 = c / 8;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n * n * c8) return;
    const int ch = (int)(i % c8) * 8;
    const size_t v = i / c8;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((size_t)n * n));
    *reinterpret_cast<bf16x8*>(dst + v * cs + off + ch) =
        *reinterpret_cast<const bf16x8*>(src + (size_t)(((z / 2) * m + y / 2) * m + x / 2) * c + ch);
}

// gradient of the upsample: out[p][ch] = sum over the 8 children of g[child][ch] (g stride gs)
__global__ void __launch_bounds__(NET_THREADS)
k_upsample_bwd(int m, const bf16_t* __restrict__ g, int gs, int c, bf16_t* __restrict__ out) {
    const int n = 2 * m;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)m * m * m * c) return;
    const int ch = (int)(i % c);
    const size_t p = i / c;
    const int X = (int)(p % m), Y = (int)((p / m) % m), Z = (int)(p / ((size_t)m * m));
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int x = 2 * X + (k & 1), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z + (k >> 2);
        s += bf(g[(size_t)((z * n + y) * n + x) * gs + ch]);
    }
    out[i] = tobf(s);
}

// gradient at the pre-activation of a ReLU layer whose output y (stride ys, offset yo) is read
// by the next layer (gradient g, c channels): gpre = y > 0 ? g : 0
__global__ void __launch_bounds__(NET_THREADS)
k_relu_bwd(size_t nv, const bf16_t* __restrict__ g, const bf16_t* __restrict__ y, int ys, int yo, int c,
           bf16_t* __restrict__ gpre) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= nv * c) return;
    const size_t v = i / c;
    const int ch = (int)(i % c);
    gpre[i] = bf(y[v * ys + yo + ch]) > 0.0f ? g[i] : tobf(0.0f);
}

// gradient of a skip activation y = cat[:, off:off + c] (stride cs) that also feeds the pool:
//   g = gcat[v][off + ch] + gpool[parent][ch] / 8 (rounded to BF16), then the ReLU mask of y
__global__ void __launch_bounds__(NET_THREADS)
k_skip_bwd(int n, const bf16_t* __restrict__ gcat, const bf16_t* __restrict__ ycat, int cs, int off,
           const bf16_t* __restrict__ gpool, int c, bf16_t* __restrict__ gpre) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n * n * c) return;
    const int m = n / 2;
    const int ch = (int)(i % c);
    const size_t v = i / c;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((size_t)n * n));
    const size_t p = (size_t)((z / 2) * m + y / 2) * m + x / 2;
    const float s = bf(gcat[v * cs + off + ch]) + 0.125f * bf(gpool[p * c + ch]);
    gpre[i] = bf(ycat[v * cs + off + ch]) > 0.0f ? tobf(s) : tobf(0.0f);
}

// Adam (Kingma & Ba 2015) on the FP32 parameters, bias-corrected with step t
__global__ void __launch_bounds__(NET_THREADS)
k_adam(int n, float* __restrict__ p, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
       float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace vfx
