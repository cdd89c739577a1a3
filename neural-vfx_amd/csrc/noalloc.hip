// MacCormack advection without a context and without any device allocation: the variant of
// vfx_advect for callers that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  The kernels are vfx_advect's; their scratch (the backward
// values, the trace corners and the parameter buffer) is one file-scope device array per GPU.
//
// The array's name is a host-side stand-in, not a device address: its address on a device comes
// from hipGetSymbolAddress, taken once per device (with that device current) at the entry point's
// first call there and kept per device.  Every call on one GPU shares that GPU's array, so calls
// that overlap on one GPU (concurrent streams or threads) race.  Nothing a call reads from the
// array is left over from an earlier call: the first kernel writes the whole parameter buffer,
// the backward pass writes every backward value and trace corner the MacCormack pass reads.
#include <mutex>

#include "ctx.hpp"

#if !VFX_F64

constexpr size_t NOALLOC_MAX_CELLS = (size_t)1 << 20;   // grids up to 2^20 cells (96^3 = 884,736; 100^3)
constexpr size_t NOALLOC_HAT = 6 * NOALLOC_MAX_CELLS;  // backward values of up to 6 channels (pair planes)
constexpr size_t NOALLOC_QIDX = NOALLOC_MAX_CELLS;     // trace corners (int)
constexpr size_t NOALLOC_FLOATS = NOALLOC_HAT + NOALLOC_QIDX + P_SIZE;   // 28 MiB + 2.3 KiB

static __device__ __align__(256) float scratch[NOALLOC_FLOATS];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_mu;
float* g_scratch[NOALLOC_MAX_DEVICES] = {};

// the parameter buffer of the advection: dt in P_DT, every other slot 0 (what vfx_advect uploads)
__global__ void __launch_bounds__(256) k_noalloc_prm(float* __restrict__ prm, float dt) {
    for (int i = threadIdx.x; i < P_SIZE; i += blockDim.x) prm[i] = i == P_DT ? dt : 0.0f;
}

// this device's copy of the scratch array (looked up on the first call on the device)
int device_scratch(float** out) {
    int dev = -1;
    if (!hip_ok(hipGetDevice(&dev), "device")) return VFX_E_HIP;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return arg_error("device index above the no-allocation table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[dev] == nullptr) {
        void* p = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "no-allocation scratch address")) return VFX_E_HIP;
        g_scratch[dev] = static_cast<float*>(p);
    }
    *out = g_scratch[dev];
    return VFX_OK;
}

}  // namespace

extern "C" {

int vfx_advect_noalloc(const float* field, int C, const float* vel, float* out, int res, float dt, void* stream) {
    if (!field || !vel || !out) return arg_error("null pointer");
    if (C < 1) return arg_error("channel count must be >= 1");
    if (res < 2 || res % 2 != 0) return arg_error("even grid size >= 2 required");
    const size_t n3 = (size_t)res * res * res;
    if (n3 > NOALLOC_MAX_CELLS) {
        set_error("grid of " + std::to_string(res) + "^3 cells exceeds the no-allocation scratch (at most " +
                  std::to_string(NOALLOC_MAX_CELLS) + " cells)");
        return VFX_E_VALUE;
    }
    float* base = nullptr;
    if (int rc = device_scratch(&base)) return rc;
    float* hat = base;
    int* qidx = reinterpret_cast<int*>(base + NOALLOC_HAT);
    float* prm = base + NOALLOC_HAT + NOALLOC_QIDX;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // the same launch geometry and 1 / dx as a context of this size (vfx_ctx_create)
    const Geom g = make_geom(res, 1);
    const double dx = 1.0 / (double)res;
    const float inv_dx = (float)(1.0 / dx);
    hipLaunchKernelGGL(k_noalloc_prm, dim3(1), dim3(256), 0, s, prm, dt);
    for (int c0 = 0; c0 < C; c0 += 6) {
        const int cc = C - c0 < 6 ? C - c0 : 6;
        launch_advect(g, cc, field + (size_t)c0 * n3, vel, prm, inv_dx, hat, qidx, out + (size_t)c0 * n3, nullptr,
                      nullptr, s);
    }
    return launch_ok("advect (no allocation)") ? VFX_OK : VFX_E_HIP;
}

}  // extern "C"

#endif  // !VFX_F64
