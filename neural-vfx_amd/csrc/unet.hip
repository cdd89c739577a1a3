// The learned pressure preconditioner: U-Net forward / backward / Adam on BF16 MFMA kernels, its
// scaling as a preconditioner, and the network part of the C ABI (include/vfxsim3d.h, vfx_net_*).
// FP32 build only (the FP64 build of the solver has no network).
#include "../../include/vfxsim3d.h"
#include "unet.hpp"
#include "unet_kernels.hpp"

#if !VFX_F64

#include <cstring>
#include <new>

namespace vfx {

// ---------------------------------------------------------------- scaling and loss kernels
__global__ void __launch_bounds__(256) k_sq_part(size_t n, const real* __restrict__ r, double* __restrict__ part) {
    __shared__ double sh[4];
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += (double)r[i] * (double)r[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// scal[0] = 1 / rms(r), scal[1] = rms(r) (both 1 for r = 0)
__global__ void __launch_bounds__(FIN_THREADS) k_rms_fin(const double* __restrict__ part, int nb, size_t n,
                                                         double* __restrict__ scal) {
    __shared__ double sh[FIN_THREADS / 64];
    const double s = reduce_parts_sum(part, nb, sh);
    if (threadIdx.x == 0) {
        const double rms = sqrt(s / (double)n);
        scal[0] = rms > 0.0 ? 1.0 / rms : 1.0;
        scal[1] = rms > 0.0 ? rms : 1.0;
    }
}

// network input: channel 0 = r / rms(r) (rounded to BF16), channel 1 = fluid indicator, rest 0
__global__ void __launch_bounds__(256) k_pack_input(size_t n, const real* __restrict__ r, const uint8_t* __restrict__ fm,
                                                    const double* __restrict__ scal, bf16_t* __restrict__ X,
                                                    const DevState* __restrict__ gst, int gk) {
    if (gated_off(gst, gk)) return;
    const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float inv = (float)scal[0];
    bf16x8 a = {}, b = {};
    a[0] = tobf((float)r[v] * inv);
    a[1] = tobf(fm[v] ? 1.0f : 0.0f);
    reinterpret_cast<bf16x8*>(X + v * 16)[0] = a;
    reinterpret_cast<bf16x8*>(X + v * 16)[1] = b;
}

// raw two-channel input (ch 0, ch 1 as given; testing)
__global__ void __launch_bounds__(256) k_pack_raw(size_t n, const float* __restrict__ in, bf16_t* __restrict__ X) {
    const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    bf16x8 a = {}, b = {};
    a[0] = tobf(in[v]);
    a[1] = tobf(in[n + v]);
    reinterpret_cast<bf16x8*>(X + v * 16)[0] = a;
    reinterpret_cast<bf16x8*>(X + v * 16)[1] = b;
}

// z = rms(r) * out in fluid cells, 0 in solids
__global__ void __launch_bounds__(256) k_unpack(size_t n, const float* __restrict__ out, const uint8_t* __restrict__ fm,
                                                const double* __restrict__ scal, real* __restrict__ z) {
    const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    z[v] = fm[v] ? (real)(out[v] * (float)scal[1]) : RL(0.0);
}

// d loss / d out of the raw network from an FP32 field (channel 0 of a 16-channel BF16 buffer)
__global__ void __launch_bounds__(256) k_pack_gout(size_t n, const float* __restrict__ g, bf16_t* __restrict__ G) {
    const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    bf16x8 a = {}, b = {};
    a[0] = tobf(g[v]);
    reinterpret_cast<bf16x8*>(G + v * 16)[0] = a;
    reinterpret_cast<bf16x8*>(G + v * 16)[1] = b;
}

// e = f - phi with f = rms * out in fluid cells
__global__ void __launch_bounds__(256) k_err(size_t n, const float* __restrict__ out, const uint8_t* __restrict__ fm,
                                             const double* __restrict__ scal, const real* __restrict__ phi,
                                             real* __restrict__ e) {
    const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float f = fm[v] ? out[v] * (float)scal[1] : 0.0f;
    e[v] = f - phi[v];
}

__global__ void __launch_bounds__(256) k_dot2_part(size_t n, const real* __restrict__ a, const real* __restrict__ b,
                                                   const real* __restrict__ c, const real* __restrict__ d,
                                                   double* __restrict__ part, int nb) {
    __shared__ double sh[4];
    double s = 0.0, t = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        s += (double)a[i] * (double)b[i];
        t += (double)c[i] * (double)d[i];
    }
    const double s1 = block_sum(s, sh);
    const double t1 = block_sum(t, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s1;
        part[nb + blockIdx.x] = t1;
    }
}

// scal[2] = e.Ke, scal[3] = phi.b (= phi.K phi), scal[4] = loss = e.Ke / phi.b
__global__ void __launch_bounds__(FIN_THREADS) k_loss_fin(const double* __restrict__ part, int nb, double* __restrict__ scal) {
    __shared__ double sh[FIN_THREADS / 64];
    const double eke = reduce_parts_sum(part, nb, sh);
    const double pb = reduce_parts_sum(part + nb, nb, sh);
    if (threadIdx.x == 0) {
        scal[2] = eke;
        scal[3] = pb;
        scal[4] = pb > 0.0 ? eke / pb : 0.0;
    }
}

// d loss / d out = rms * 2 K e / (phi.b) in fluid cells
__global__ void __launch_bounds__(256) k_dout(size_t n, const real* __restrict__ ke, const uint8_t* __restrict__ fm,
                                              const double* __restrict__ scal, float* __restrict__ dout) {
    const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    const double pb = scal[3];
    dout[v] = (fm[v] && pb > 0.0) ? (float)(scal[1] * 2.0 * (double)ke[v] / pb) : 0.0f;
}

// ---------------------------------------------------------------- coarse correction
// The network run on the (N/4)^3 grid of 4^3 blocks of an N^3 field: the fine input is the block
// mean, the coarse output is spread back to the block's fluid cells (constant prolongation).

// c[B] = mean of the 64 fine values of block B (solid cells hold 0)
__global__ void __launch_bounds__(256) k_pool4(int N, const real* __restrict__ f, real* __restrict__ c) {
    const int M = N >> 2;
    const size_t m3 = (size_t)M * M * M;
    const size_t b = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (b >= m3) return;
    const int I = (int)(b % M), J = (int)((b / M) % M), K = (int)(b / ((size_t)M * M));
    float acc = 0.0f;
    for (int dz = 0; dz < 4; ++dz)
        for (int dy = 0; dy < 4; ++dy) {
            const real* row = f + ((size_t)(4 * K + dz) * N + (4 * J + dy)) * N + 4 * I;
            const float4 v = *reinterpret_cast<const float4*>(row);
            acc += (v.x + v.y) + (v.z + v.w);
        }
    c[b] = acc * (1.0f / 64.0f);
}

// coarse fluid flag: 1 where any of the block's 64 cells is fluid
__global__ void __launch_bounds__(256) k_mask4(int N, const uint8_t* __restrict__ fm, uint8_t* __restrict__ cm) {
    const int M = N >> 2;
    const size_t m3 = (size_t)M * M * M;
    const size_t b = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (b >= m3) return;
    const int I = (int)(b % M), J = (int)((b / M) % M), K = (int)(b / ((size_t)M * M));
    int any = 0;
    for (int dz = 0; dz < 4; ++dz)
        for (int dy = 0; dy < 4; ++dy) {
            const uint32_t w = *reinterpret_cast<const uint32_t*>(fm + ((size_t)(4 * K + dz) * N + (4 * J + dy)) * N + 4 * I);
            any |= w != 0u;
        }
    cm[b] = any ? 1 : 0;
}

// fine cell v of block B: ADD: z[v] += rms * out[B] (fluid cells; solids keep z); ERR: e[v] =
// rms * out[B] - t[v] in fluid cells, 0 in solids
template <bool ADD>
__global__ void __launch_bounds__(256) k_up4(int N, const float* __restrict__ out, const double* __restrict__ scal,
                                             const uint8_t* __restrict__ fm, real* __restrict__ z,
                                             const real* __restrict__ t) {
    const size_t n3 = (size_t)N * N * N;
    const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (v >= n3) return;
    const int x = (int)(v % N), y = (int)((v / N) % N), zc = (int)(v / ((size_t)N * N));
    const int M = N >> 2;
    const float c = out[((size_t)(zc >> 2) * M + (y >> 2)) * M + (x >> 2)] * (float)scal[1];
    if (ADD) {
        if (fm[v]) z[v] = z[v] + c;
    } else {
        z[v] = fm[v] ? c - t[v] : RL(0.0);
    }
}

// d loss / d out[B] = rms * sum over the block's fluid cells of 2 (K e) / (e_V . K e_V)
__global__ void __launch_bounds__(256) k_dout4(int N, const real* __restrict__ ke, const uint8_t* __restrict__ fm,
                                               const double* __restrict__ scal, float* __restrict__ dout) {
    const int M = N >> 2;
    const size_t m3 = (size_t)M * M * M;
    const size_t b = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (b >= m3) return;
    const int I = (int)(b % M), J = (int)((b / M) % M), K = (int)(b / ((size_t)M * M));
    double acc = 0.0;
    for (int dz = 0; dz < 4; ++dz)
        for (int dy = 0; dy < 4; ++dy)
            for (int dx = 0; dx < 4; ++dx) {
                const size_t v = ((size_t)(4 * K + dz) * N + (4 * J + dy)) * N + 4 * I + dx;
                acc += fm[v] ? (double)ke[v] : 0.0;
            }
    const double pb = scal[3];
    dout[b] = pb > 0.0 ? (float)(scal[1] * 2.0 * acc / pb) : 0.0f;
}

// ---------------------------------------------------------------- launch helpers
static unsigned blocks_for(size_t items, int threads = 256) { return (unsigned)((items + threads - 1) / threads); }

// launch shape of a convolution on an N^3 grid: the LDS-tiled kernel from N = 48 (up to 64 output
// channels; smaller grids give it too few workgroups); otherwise the K split KS and the voxel tiles per wave MT (KS = 1, MT = 4 while there
// are at least 16384 tiles of 16 voxels; KS = 2, 4 below 4096, 1024)
template <int CIN, int CT>
static void conv_t(int N, const bf16_t* in, const bf16_t* w, int kp, const float* bias, int cout, int relu, bf16_t* out,
                   int ocs, int ooff, float* out32, hipStream_t s, const DevState* gst, int gk) {
    const size_t n3 = (size_t)N * N * N;
    const size_t tiles = (n3 + 15) / 16;
    constexpr int W = NET_THREADS / 64;
#define VFX_CONV_LAUNCH(KS, MT)                                                                                      \
    hipLaunchKernelGGL((k_conv3<CIN, CT, KS, MT>), dim3((unsigned)((tiles + (W / KS) * MT - 1) / ((W / KS) * MT))), \
                       dim3(NET_THREADS), 0, s, N, in, w, kp, bias, cout, relu, out, ocs, ooff, out32, gst, gk)
    if (N >= 48 && CT <= 4) {
        const unsigned nb = (unsigned)(((N + CX - 1) / CX) * ((N + CY - 1) / CY) * ((N + CZ - 1) / CZ));
        hipLaunchKernelGGL((k_conv3l<CIN, CT>), dim3(nb), dim3(NET_THREADS), 0, s, N, in, w, kp, bias, cout, relu, out,
                           ocs, ooff, out32, gst, gk);
    } else if (CT <= 4 && tiles >= 16384) VFX_CONV_LAUNCH(1, (CT <= 4 ? 4 : 1));
    else if (tiles >= 4096) VFX_CONV_LAUNCH(1, 1);
    else if (tiles >= 1024) VFX_CONV_LAUNCH(2, 1);
    else VFX_CONV_LAUNCH(4, 1);
#undef VFX_CONV_LAUNCH
}

static bool conv(int cin, int ct, int N, const bf16_t* in, const bf16_t* w, int kp, const float* bias, int cout, int relu,
                 bf16_t* out, int ocs, int ooff, float* out32, hipStream_t s, const DevState* gst = nullptr, int gk = 0) {
#define VFX_CONV(CI, CT_)                                                                  \
    if (cin == CI && ct == CT_) {                                                          \
        conv_t<CI, CT_>(N, in, w, kp, bias, cout, relu, out, ocs, ooff, out32, s, gst, gk); \
        return true;                                                                       \
    }
    VFX_CONV(16, 1) VFX_CONV(16, 2) VFX_CONV(16, 3) VFX_CONV(32, 1) VFX_CONV(32, 2) VFX_CONV(32, 4) VFX_CONV(32, 6)
    VFX_CONV(48, 1) VFX_CONV(64, 2) VFX_CONV(64, 4) VFX_CONV(96, 2)
#undef VFX_CONV
    return false;
}

static int wgrad_groups(size_t n3) {
    size_t ng = (n3 + 2047) / 2048;
    if (ng > 256) ng = 256;
    return (int)(ng < 1 ? 1 : ng);
}

static bool wgrad(int cinp, int coutp, int N, const bf16_t* in, const bf16_t* gpre, float* part, hipStream_t s, int* ng_out) {
    const size_t n3 = (size_t)N * N * N;
    const int ng = wgrad_groups(n3);
    int vg = (int)((n3 + ng - 1) / ng);
    vg = (vg + WG_CHUNK - 1) / WG_CHUNK * WG_CHUNK;
    *ng_out = ng;
#define VFX_WG(CI, CO)                                                                                     \
    if (cinp == CI && coutp == CO) {                                                                       \
        hipLaunchKernelGGL((k_conv3_wgrad<CI, CO>), dim3(ng, 27), dim3(NET_THREADS), 0, s, N, in, gpre, vg, part); \
        return true;                                                                                       \
    }
    VFX_WG(16, 16) VFX_WG(16, 32) VFX_WG(32, 32) VFX_WG(32, 64) VFX_WG(64, 64) VFX_WG(96, 32) VFX_WG(48, 16)
#undef VFX_WG
    return false;
}

// ---------------------------------------------------------------- Net
static size_t al(size_t b) { return (b + 255) / 256 * 256; }

int Net::create(int N_) {
    N = N_;
    // (cin, cout, cinp, coutp, relu) of the 11 convolutions in execution order
    const int spec[NET_LAYERS][5] = {{2, 16, 16, 16, 1},  {16, 16, 16, 16, 1}, {16, 32, 16, 32, 1}, {32, 32, 32, 32, 1},
                                     {32, 64, 32, 64, 1}, {64, 64, 64, 64, 1}, {96, 32, 96, 32, 1}, {32, 32, 32, 32, 1},
                                     {48, 16, 48, 16, 1}, {16, 16, 16, 16, 1}, {16, 1, 16, 16, 0}};
    nparam = 0;
    for (int i = 0; i < NET_LAYERS; ++i) {
        NetLayer& l = L[i];
        l.cin = spec[i][0]; l.cout = spec[i][1]; l.cinp = spec[i][2]; l.coutp = spec[i][3]; l.relu = spec[i][4];
        l.w_off = nparam;
        nparam += (size_t)l.cout * l.cin * 27;
        l.b_off = nparam;
        nparam += (size_t)l.cout;
        l.kp = (27 * l.cinp + 31) / 32 * 32;
        l.kpt = (27 * l.coutp + 31) / 32 * 32;
    }
    const size_t n0 = (size_t)N * N * N, n1 = n0 / 8, n2 = n0 / 64;
    // byte sizes of every buffer, in allocation order
    std::vector<size_t> sz;
    for (int i = 0; i < 4; ++i) sz.push_back(al(nparam * 4));   // p, m, v, g
    for (int i = 0; i < NET_LAYERS; ++i) {
        sz.push_back(al((size_t)L[i].coutp * L[i].kp * 2));
        sz.push_back(al((size_t)L[i].cinp * L[i].kpt * 2));
    }
    const size_t act[] = {16 * n0, 16 * n0, 48 * n0, 16 * n0, 16 * n0,   // X A0 C0 D0a D0b
                          16 * n1, 32 * n1, 96 * n1, 32 * n1, 32 * n1,   // P0 A1 C1 D1a D1b
                          32 * n2, 64 * n2, 64 * n2,                     // P1 A2 B2
                          48 * n0, 16 * n0,                              // gA0 gB0
                          96 * n1, 32 * n1, 32 * n1,                     // gA1 gB1 gC1
                          64 * n2, 64 * n2, 32 * n2};                    // gA2 gB2 gC2
    for (size_t a : act) sz.push_back(al(a * 2));
    sz.push_back(al(n0 * 4));   // out
    size_t wp = 0;
    for (int i = 0; i < NET_LAYERS; ++i) {
        const int lev = (i <= 1 || i >= 8) ? 0 : ((i <= 3 || i >= 6) ? 1 : 2);
        const size_t nl = lev == 0 ? n0 : (lev == 1 ? n1 : n2);
        const size_t need = (size_t)wgrad_groups(nl) * 27 * L[i].coutp * L[i].cinp;
        if (need > wp) wp = need;
    }
    sz.push_back(al(wp * 4));            // wpart
    sz.push_back(al(256 * 64 * 4));      // bpart
    sz.push_back(al(2 * 4096 * 8));      // dpart
    sz.push_back(al(16 * 8));            // scal
    sz.push_back(al(n0 * sizeof(real))); // work
    sz.push_back(al(n0 * sizeof(real))); // work2
    sz.push_back(al(n0));                // cmask
    for (int i = 0; i < 3; ++i) sz.push_back(al(n0 * sizeof(real)));   // csr csrp ccz
    const size_t Nf = (size_t)4 * N;     // k_resid_pool's workgroups on the fine grid: at most
    const size_t rpb = ((Nf + 63) / 64) * ((Nf + 7) / 8) * ((Nf + 3) / 4);   // columns of 4 planes
    sz.push_back(al(2 * rpb * 8));       // lpart
    sz.push_back(al(2 * 256 * 8));       // fpart
    size_t total = 0;
    for (size_t b : sz) total += b;
    if (!hip_ok(hipMalloc(&pool, total), "network allocation")) return VFX_E_HIP;
    if (!hip_ok(hipMemset(pool, 0, total), "network clear")) return VFX_E_HIP;
    char* q = (char*)pool;
    size_t k = 0;
    auto take = [&](void) -> char* { char* r = q; q += sz[k++]; return r; };
    p = (float*)take(); m = (float*)take(); v = (float*)take(); g = (float*)take();
    for (int i = 0; i < NET_LAYERS; ++i) { L[i].wf = (bf16_t*)take(); L[i].wb = (bf16_t*)take(); }
    bf16_t** bufs[] = {&X, &A0, &C0, &D0a, &D0b, &P0, &A1, &C1, &D1a, &D1b, &P1, &A2, &B2,
                       &gA0, &gB0, &gA1, &gB1, &gC1, &gA2, &gB2, &gC2};
    for (bf16_t** b : bufs) *b = (bf16_t*)take();
    out = (float*)take();
    wpart = (float*)take();
    bpart = (float*)take();
    dpart = (double*)take();
    scal = (double*)take();
    work = (float*)take();
    work2 = (float*)take();
    cmask = (uint8_t*)take();
    csr = (real*)take();
    csrp = (real*)take();
    ccz = (real*)take();
    lpart = (double*)take();
    fpart = (double*)take();
    adam_t = 0;
    return VFX_OK;
}

void Net::destroy() {
    if (pool) (void)hipFree(pool);
    pool = nullptr;
}

void Net::pack_weights(hipStream_t s) {
    for (int i = 0; i < NET_LAYERS; ++i) {
        const NetLayer& l = L[i];
        const size_t items = std::max((size_t)l.coutp * l.kp, (size_t)l.cinp * l.kpt);
        hipLaunchKernelGGL(k_pack_weights, dim3(blocks_for(items)), dim3(256), 0, s, p + l.w_off, l.cin, l.cout, l.cinp,
                           l.coutp, l.kp, l.kpt, l.wf, l.wb);
    }
}

// forward of layer i from `in` (cinp channels) into out (stride ocs, offset ooff) or out32
static void layer_fwd(const Net& n, int i, int N, const bf16_t* in, bf16_t* out, int ocs, int ooff, float* out32,
                      hipStream_t s, const DevState* gst, int gk) {
    const NetLayer& l = n.L[i];
    conv(l.cinp, l.coutp / 16, N, in, l.wf, l.kp, n.p + l.b_off, l.cout, l.relu, out, ocs, ooff, out32, s, gst, gk);
}

void Net::forward(hipStream_t s, const DevState* gst, int gk) {
    const int n1 = N / 2, n2 = N / 4;
    const size_t v0 = (size_t)N * N * N, v1 = v0 / 8;
    layer_fwd(*this, 0, N, X, A0, 16, 0, nullptr, s, gst, gk);
    layer_fwd(*this, 1, N, A0, C0, 48, 32, nullptr, s, gst, gk);                 // e0b -> C0[32:48]
    hipLaunchKernelGGL(k_pool, dim3(blocks_for(v1 * 2)), dim3(256), 0, s, N, C0, 48, 32, 16, P0, gst, gk);
    layer_fwd(*this, 2, n1, P0, A1, 32, 0, nullptr, s, gst, gk);
    layer_fwd(*this, 3, n1, A1, C1, 96, 64, nullptr, s, gst, gk);                // e1b -> C1[64:96]
    hipLaunchKernelGGL(k_pool, dim3(blocks_for(v1 / 8 * 4)), dim3(256), 0, s, n1, C1, 96, 64, 32, P1, gst, gk);
    layer_fwd(*this, 4, n2, P1, A2, 64, 0, nullptr, s, gst, gk);
    layer_fwd(*this, 5, n2, A2, B2, 64, 0, nullptr, s, gst, gk);
    hipLaunchKernelGGL(k_upsample, dim3(blocks_for(v1 * 8)), dim3(256), 0, s, n2, B2, 64, C1, 96, 0, gst, gk);
    layer_fwd(*this, 6, n1, C1, D1a, 32, 0, nullptr, s, gst, gk);
    layer_fwd(*this, 7, n1, D1a, D1b, 32, 0, nullptr, s, gst, gk);
    hipLaunchKernelGGL(k_upsample, dim3(blocks_for(v0 * 4)), dim3(256), 0, s, n1, D1b, 32, C0, 48, 0, gst, gk);
    layer_fwd(*this, 8, N, C0, D0a, 16, 0, nullptr, s, gst, gk);
    layer_fwd(*this, 9, N, D0a, D0b, 16, 0, nullptr, s, gst, gk);
    layer_fwd(*this, 10, N, D0b, nullptr, 16, 0, out, s, gst, gk);
}

// weight + bias gradient of layer i from its input `in` and pre-activation gradient `gpre`
static void layer_grads(Net& n, int i, int N, const bf16_t* in, const bf16_t* gpre, hipStream_t s) {
    const NetLayer& l = n.L[i];
    int ng = 1;
    wgrad(l.cinp, l.coutp, N, in, gpre, n.wpart, s, &ng);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for((size_t)l.cout * l.cin * 27)), dim3(NET_THREADS), 0, s, n.wpart, ng,
                       l.cinp, l.coutp, l.cin, l.cout, n.g + l.w_off);
    const size_t n3 = (size_t)N * N * N;
    int nb = (int)((n3 + 4095) / 4096);
    if (nb > 256) nb = 256;
    const int vpb = (int)((n3 + nb - 1) / nb);
    hipLaunchKernelGGL(k_bias_part, dim3(nb), dim3(NET_THREADS), 0, s, (int)n3, gpre, l.coutp, l.cout, vpb, n.bpart);
    hipLaunchKernelGGL(k_bias_reduce, dim3(1), dim3(64), 0, s, n.bpart, nb, l.cout, n.g + l.b_off);
}

// input gradient of layer i (the transposed convolution): gin (cinp channels) from gpre
static void layer_dgrad(const Net& n, int i, int N, const bf16_t* gpre, bf16_t* gin, hipStream_t s) {
    const NetLayer& l = n.L[i];
    conv(l.coutp, l.cinp / 16, N, gpre, l.wb, l.kpt, nullptr, l.cinp, 0, gin, l.cinp, 0, nullptr, s);
}

void Net::backward(hipStream_t s) {
    const int n1 = N / 2, n2 = N / 4;
    const size_t v0 = (size_t)N * N * N, v1 = v0 / 8, v2 = v1 / 8;
    hipLaunchKernelGGL(k_pack_gout, dim3(blocks_for(v0)), dim3(256), 0, s, v0, work, gB0);
    // level 0, decoder
    layer_grads(*this, 10, N, D0b, gB0, s);
    layer_dgrad(*this, 10, N, gB0, gA0, s);                                     // d D0b
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(v0 * 16)), dim3(256), 0, s, v0, gA0, D0b, 16, 0, 16, gB0);
    layer_grads(*this, 9, N, D0a, gB0, s);
    layer_dgrad(*this, 9, N, gB0, gA0, s);                                      // d D0a
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(v0 * 16)), dim3(256), 0, s, v0, gA0, D0a, 16, 0, 16, gB0);
    layer_grads(*this, 8, N, C0, gB0, s);
    layer_dgrad(*this, 8, N, gB0, gA0, s);                                      // d C0 (48 channels)
    hipLaunchKernelGGL(k_upsample_bwd, dim3(blocks_for(v1 * 32)), dim3(256), 0, s, n1, gA0, 48, 32, gB1);   // d D1b
    // level 1, decoder
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(v1 * 32)), dim3(256), 0, s, v1, gB1, D1b, 32, 0, 32, gC1);
    layer_grads(*this, 7, n1, D1a, gC1, s);
    layer_dgrad(*this, 7, n1, gC1, gB1, s);                                     // d D1a
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(v1 * 32)), dim3(256), 0, s, v1, gB1, D1a, 32, 0, 32, gC1);
    layer_grads(*this, 6, n1, C1, gC1, s);
    layer_dgrad(*this, 6, n1, gC1, gA1, s);                                     // d C1 (96 channels)
    hipLaunchKernelGGL(k_upsample_bwd, dim3(blocks_for(v2 * 64)), dim3(256), 0, s, n2, gA1, 96, 64, gA2);   // d B2
    // level 2
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(v2 * 64)), dim3(256), 0, s, v2, gA2, B2, 64, 0, 64, gB2);
    layer_grads(*this, 5, n2, A2, gB2, s);
    layer_dgrad(*this, 5, n2, gB2, gA2, s);                                     // d A2
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(v2 * 64)), dim3(256), 0, s, v2, gA2, A2, 64, 0, 64, gB2);
    layer_grads(*this, 4, n2, P1, gB2, s);
    layer_dgrad(*this, 4, n2, gB2, gC2, s);                                     // d P1 (32 channels)
    // level 1, encoder: e1b feeds the pool and the concat
    hipLaunchKernelGGL(k_skip_bwd, dim3(blocks_for(v1 * 32)), dim3(256), 0, s, n1, gA1, C1, 96, 64, gC2, 32, gB1);
    layer_grads(*this, 3, n1, A1, gB1, s);
    layer_dgrad(*this, 3, n1, gB1, gC1, s);                                     // d A1
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(v1 * 32)), dim3(256), 0, s, v1, gC1, A1, 32, 0, 32, gB1);
    layer_grads(*this, 2, n1, P0, gB1, s);
    layer_dgrad(*this, 2, n1, gB1, gA1, s);                                     // d P0 (16 channels)
    // level 0, encoder
    hipLaunchKernelGGL(k_skip_bwd, dim3(blocks_for(v0 * 16)), dim3(256), 0, s, N, gA0, C0, 48, 32, gA1, 16, gB0);
    layer_grads(*this, 1, N, A0, gB0, s);
    layer_dgrad(*this, 1, N, gB0, gA0, s);                                      // d A0
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(v0 * 16)), dim3(256), 0, s, v0, gA0, A0, 16, 0, 16, gB0);
    layer_grads(*this, 0, N, X, gB0, s);
}

void Net::adam(float lr, float b1, float b2, float eps, hipStream_t s) {
    adam_t += 1;
    const float c1 = 1.0f - std::pow(b1, (float)adam_t), c2 = 1.0f - std::pow(b2, (float)adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(nparam)), dim3(NET_THREADS), 0, s, (int)nparam, p, m, v, g, lr, b1, b2, eps,
                       c1, c2);
    pack_weights(s);
}

void Net::apply(const real* r, const uint8_t* fm, real* z, hipStream_t s) {
    const size_t n0 = (size_t)N * N * N;
    hipLaunchKernelGGL(k_sq_part, dim3(512), dim3(256), 0, s, n0, r, dpart);
    hipLaunchKernelGGL(k_rms_fin, dim3(1), dim3(FIN_THREADS), 0, s, dpart, 512, n0, scal);
    hipLaunchKernelGGL(k_pack_input, dim3(blocks_for(n0)), dim3(256), 0, s, n0, r, fm, scal, X, nullptr, 0);
    forward(s);
    hipLaunchKernelGGL(k_unpack, dim3(blocks_for(n0)), dim3(256), 0, s, n0, out, fm, scal, z);
}

// Coarse correction of a fine residual (the network's grid is N/4 of the fine grid Nf): the block
// means of r -> scaled network input (with the coarse fluid flags), forward, and z += rms * out
// spread over each block's fluid cells.
void Net::apply_coarse(int Nf, const real* r, const uint8_t* fm, real* z, hipStream_t s) {
    const size_t n0 = (size_t)N * N * N, nf = (size_t)Nf * Nf * Nf;
    hipLaunchKernelGGL(k_pool4, dim3(blocks_for(n0)), dim3(256), 0, s, Nf, r, work2);
    hipLaunchKernelGGL(k_mask4, dim3(blocks_for(n0)), dim3(256), 0, s, Nf, fm, cmask);
    hipLaunchKernelGGL(k_sq_part, dim3(512), dim3(256), 0, s, n0, work2, dpart);
    hipLaunchKernelGGL(k_rms_fin, dim3(1), dim3(FIN_THREADS), 0, s, dpart, 512, n0, scal);
    hipLaunchKernelGGL(k_pack_input, dim3(blocks_for(n0)), dim3(256), 0, s, n0, work2, cmask, scal, X, nullptr, 0);
    forward(s);
    hipLaunchKernelGGL(k_up4<true>, dim3(blocks_for(nf)), dim3(256), 0, s, Nf, out, scal, fm, z, nullptr);
}

void Net::coarse_mask(int Nf, const uint8_t* fm, hipStream_t s) {
    const size_t n0 = (size_t)N * N * N;
    hipLaunchKernelGGL(k_mask4, dim3(blocks_for(n0)), dim3(256), 0, s, Nf, fm, cmask);
}

void Net::coarse_forward(const DevState* st, int k, hipStream_t s) {
    const size_t n0 = (size_t)N * N * N;
    hipLaunchKernelGGL(k_pack_input, dim3(blocks_for(n0)), dim3(256), 0, s, n0, work2, cmask, scal, X, st, k);
    forward(s, st, k);
}

// One training step of the coarse correction on a fine pair (r_V, e_V) with K e_V = r_V (the
// residual and the error left by the V-cycle): c = P net(R r_V), loss = e.Ke / e_V.r_V with
// e = c - e_V in the fine energy norm, backward through P (a block sum) and the network, Adam.
// e and K e go to the caller's fine scratch (2 Nf^3 reals).  The loss lands in scal[4].
void net_train_step_coarse(Net& n, const Geom& gf, const real* rv, const real* ev, const uint8_t* fm, float lr,
                           real* scratch, hipStream_t s) {
    const int Nf = gf.N;
    const size_t n0 = (size_t)n.N * n.N * n.N, nf = gf.n3;
    real* e = scratch;
    real* ke = scratch + nf;
    hipLaunchKernelGGL(k_pool4, dim3(blocks_for(n0)), dim3(256), 0, s, Nf, rv, n.work2);
    hipLaunchKernelGGL(k_mask4, dim3(blocks_for(n0)), dim3(256), 0, s, Nf, fm, n.cmask);
    hipLaunchKernelGGL(k_sq_part, dim3(512), dim3(256), 0, s, n0, n.work2, n.dpart);
    hipLaunchKernelGGL(k_rms_fin, dim3(1), dim3(FIN_THREADS), 0, s, n.dpart, 512, n0, n.scal);
    hipLaunchKernelGGL(k_pack_input, dim3(blocks_for(n0)), dim3(256), 0, s, n0, n.work2, n.cmask, n.scal, n.X, nullptr, 0);
    n.forward(s);
    hipLaunchKernelGGL(k_up4<false>, dim3(blocks_for(nf)), dim3(256), 0, s, Nf, n.out, n.scal, fm, e, ev);
    launch_kmul(gf, e, fm, ke, s);
    hipLaunchKernelGGL(k_dot2_part, dim3(512), dim3(256), 0, s, nf, e, ke, ev, rv, n.dpart, 512);
    hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(FIN_THREADS), 0, s, n.dpart, 512, n.scal);
    hipLaunchKernelGGL(k_dout4, dim3(blocks_for(n0)), dim3(256), 0, s, Nf, ke, fm, n.scal, n.work);
    n.backward(s);
    n.adam(lr, 0.9f, 0.999f, 1e-8f, s);
}

// One training step on a (b, phi) pair in the energy norm of K (fm, geometry of the solver):
// forward on b, e = f(b) - phi, loss = e.Ke / phi.b, backward, Adam.  The loss lands in scal[4].
void net_train_step(Net& n, const Geom& g, const real* b, const real* phi, const uint8_t* fm, float lr, hipStream_t s) {
    const size_t n0 = (size_t)n.N * n.N * n.N;
    hipLaunchKernelGGL(k_sq_part, dim3(512), dim3(256), 0, s, n0, b, n.dpart);
    hipLaunchKernelGGL(k_rms_fin, dim3(1), dim3(FIN_THREADS), 0, s, n.dpart, 512, n0, n.scal);
    hipLaunchKernelGGL(k_pack_input, dim3(blocks_for(n0)), dim3(256), 0, s, n0, b, fm, n.scal, n.X, nullptr, 0);
    n.forward(s);
    hipLaunchKernelGGL(k_err, dim3(blocks_for(n0)), dim3(256), 0, s, n0, n.out, fm, n.scal, phi, n.work2);
    launch_kmul(g, n.work2, fm, n.work, s);   // K e
    hipLaunchKernelGGL(k_dot2_part, dim3(512), dim3(256), 0, s, n0, n.work2, n.work, phi, b, n.dpart, 512);
    hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(FIN_THREADS), 0, s, n.dpart, 512, n.scal);
    hipLaunchKernelGGL(k_dout, dim3(blocks_for(n0)), dim3(256), 0, s, n0, n.work, fm, n.scal, n.work);
    n.backward(s);
    n.adam(lr, 0.9f, 0.999f, 1e-8f, s);
}

}  // namespace vfx

using namespace vfx;

struct vfx_net {
    Net net;
};

extern "C" {

int vfx_net_create(int res, vfx_net** out) {
    if (!out) { set_error("null output pointer"); return VFX_E_ARG; }
    *out = nullptr;
    if (res < 8 || res % 4 != 0) { set_error("network grid size must be a multiple of 4, >= 8"); return VFX_E_ARG; }
    vfx_net* n = new (std::nothrow) vfx_net();
    if (!n) { set_error("host allocation"); return VFX_E_HIP; }
    if (int rc = n->net.create(res)) { n->net.destroy(); delete n; return rc; }
    *out = n;
    return VFX_OK;
}

int vfx_net_destroy(vfx_net* n) {
    if (!n) return VFX_OK;
    (void)hipDeviceSynchronize();
    n->net.destroy();
    delete n;
    return VFX_OK;
}

int vfx_net_param_count(vfx_net* n, int* count) {
    if (!n || !count) { set_error("null pointer"); return VFX_E_ARG; }
    *count = (int)n->net.nparam;
    return VFX_OK;
}

int vfx_net_get_params(vfx_net* n, float* host) {
    if (!n || !host) { set_error("null pointer"); return VFX_E_ARG; }
    if (!hip_ok(hipMemcpy(host, n->net.p, n->net.nparam * 4, hipMemcpyDeviceToHost), "parameter copy")) return VFX_E_HIP;
    return VFX_OK;
}

int vfx_net_set_params(vfx_net* n, const float* host) {
    if (!n || !host) { set_error("null pointer"); return VFX_E_ARG; }
    Net& t = n->net;
    if (!hip_ok(hipMemcpy(t.p, host, t.nparam * 4, hipMemcpyHostToDevice), "parameter copy")) return VFX_E_HIP;
    if (!hip_ok(hipMemset(t.m, 0, t.nparam * 4), "moment clear") || !hip_ok(hipMemset(t.v, 0, t.nparam * 4), "moment clear"))
        return VFX_E_HIP;
    t.adam_t = 0;
    t.pack_weights(0);
    if (!launch_ok("weight packing") || !hip_ok(hipDeviceSynchronize(), "weight packing")) return VFX_E_HIP;
    return VFX_OK;
}

int vfx_net_forward(vfx_net* n, const float* in, float* out, void* stream) {
    if (!n || !in || !out) { set_error("null pointer"); return VFX_E_ARG; }
    Net& t = n->net;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t n0 = (size_t)t.N * t.N * t.N;
    hipLaunchKernelGGL(k_pack_raw, dim3(blocks_for(n0)), dim3(256), 0, s, n0, in, t.X);
    t.forward(s);
    if (!hip_ok(hipMemcpyAsync(out, t.out, n0 * 4, hipMemcpyDeviceToDevice, s), "output copy")) return VFX_E_HIP;
    return launch_ok("network forward") ? VFX_OK : VFX_E_HIP;
}

int vfx_net_backward(vfx_net* n, const float* gout, float* grad, void* stream) {
    if (!n || !gout) { set_error("null pointer"); return VFX_E_ARG; }
    Net& t = n->net;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t n0 = (size_t)t.N * t.N * t.N;
    if (!hip_ok(hipMemcpyAsync(t.work, gout, n0 * 4, hipMemcpyDeviceToDevice, s), "gradient copy")) return VFX_E_HIP;
    t.backward(s);
    if (grad && !hip_ok(hipMemcpyAsync(grad, t.g, t.nparam * 4, hipMemcpyDeviceToDevice, s), "gradient copy"))
        return VFX_E_HIP;
    return launch_ok("network backward") ? VFX_OK : VFX_E_HIP;
}

int vfx_net_adam(vfx_net* n, const float* grad, float lr, float b1, float b2, float eps, void* stream) {
    if (!n) { set_error("null pointer"); return VFX_E_ARG; }
    Net& t = n->net;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (grad && !hip_ok(hipMemcpyAsync(t.g, grad, t.nparam * 4, hipMemcpyDeviceToDevice, s), "gradient copy"))
        return VFX_E_HIP;
    t.adam(lr, b1, b2, eps, s);
    return launch_ok("adam") ? VFX_OK : VFX_E_HIP;
}

}  // extern "C"

#endif  // !VFX_F64
