// Host side of the learned pressure preconditioner: a 3-level 3D U-Net (16 / 32 / 64 channels,
// 3^3 convolutions, ReLU, average-pool down, nearest up, skip concatenation), BF16 weights and
// activations with FP32 accumulation, trained with Adam.  FP32 build only.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "sim3d_kernels.hpp"

namespace vfx {

typedef __bf16 bf16_t;

constexpr int NET_LAYERS = 11;

struct NetLayer {
    int cin, cout, cinp, coutp, relu;
    size_t w_off, b_off;   // offsets of weight [cout][cin][27] and bias [cout] in the flat parameters
    int kp, kpt;           // K of the forward and of the transposed GEMM (multiples of 32)
    bf16_t* wf = nullptr;  // [coutp][kp]
    bf16_t* wb = nullptr;  // [cinp][kpt]
};

struct Net {
    int N = 0;             // fine grid size (multiple of 4)
    NetLayer L[NET_LAYERS];
    size_t nparam = 0;
    void* pool = nullptr;
    // FP32 parameters, Adam moments, gradient
    float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    // activations (channels-last BF16), levels 0 (N^3), 1 ((N/2)^3), 2 ((N/4)^3)
    bf16_t *X, *A0, *C0, *D0a, *D0b;
    bf16_t *P0, *A1, *C1, *D1a, *D1b;
    bf16_t *P1, *A2, *B2;
    float* out = nullptr;  // N^3 network output (FP32)
    // gradient scratch per level
    bf16_t *gA0, *gB0, *gA1, *gB1, *gC1, *gA2, *gB2, *gC2;
    float* wpart = nullptr;   // weight-gradient partials
    float* bpart = nullptr;   // bias-gradient partials
    double* dpart = nullptr;  // FP64 dot partials
    double* scal = nullptr;   // device scalars (scale, dots, loss)
    float* work = nullptr;    // N^3 FP32 scratch (K e, dL/dz)
    float* work2 = nullptr;
    uint8_t* cmask = nullptr; // N^3 fluid flags of the coarse correction's input
    // the device-sequenced flexible CG of a fine grid 4 N (coarse correction): block sums of r and
    // of r_prev over fluid cells, the scaled correction c, partials of k_resid_pool and k_fcg_fin
    real *csr = nullptr, *csrp = nullptr, *ccz = nullptr;
    double *lpart = nullptr, *fpart = nullptr;
    int adam_t = 0;

    int create(int N);
    void destroy();
    void pack_weights(hipStream_t s);
    // raw network: X (prepared by the caller) -> out; with gst, every kernel is a no-op when
    // iteration gk of the CG run that DevState gst describes is inactive
    void forward(hipStream_t s, const DevState* gst = nullptr, int gk = 0);
    // d loss / d out (FP32, N^3, in `work`) -> g (flat parameter gradient)
    void backward(hipStream_t s);
    void adam(float lr, float b1, float b2, float eps, hipStream_t s);
    // preconditioner: z = rms(r) * net(r / rms(r), fluid) in fluid cells, 0 in solids
    void apply(const real* r, const uint8_t* fm, real* z, hipStream_t s);
    // coarse correction of a fine field of Nf = 4 N: z += P net(R r) (R: 4^3 block mean, P:
    // constant over each block's fluid cells)
    void apply_coarse(int Nf, const real* r, const uint8_t* fm, real* z, hipStream_t s);
    // the device-sequenced flexible CG (fine grid Nf = 4 N): the coarse fluid flags of the face
    // mask fm, once per projection; then per iteration k the network on the block means in work2
    // (scaled by scal, written by k_resid_pool and k_lrms_fin) into out, gated on DevState st
    void coarse_mask(int Nf, const uint8_t* fm, hipStream_t s);
    void coarse_forward(const DevState* st, int k, hipStream_t s);
};

// one Adam step of the network on a (b, phi) pair in the energy norm of K (mask fm, geometry g of
// a solver of the same size); the loss e.Ke / phi.b lands in n.scal[4]
void net_train_step(Net& n, const Geom& g, const real* b, const real* phi, const uint8_t* fm, float lr, hipStream_t s);
// one Adam step of the coarse correction (network grid N, fine grid gf.N = 4 N) on the residual r_V
// and error e_V that a V-cycle left (K e_V = r_V); scratch: 2 fine fields
void net_train_step_coarse(Net& n, const Geom& gf, const real* rv, const real* ev, const uint8_t* fm, float lr,
                           real* scratch, hipStream_t s);

}  // namespace vfx
