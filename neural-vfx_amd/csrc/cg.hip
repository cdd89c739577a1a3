// Fine-grid kernels of the MG-preconditioned CG projection (pressure.py:595-775, :156-175) and
// the host-side launch sequence of one CG iteration.
//
// The stencil kernels stage what K = d d^T reaches (a 2-cell halo) in LDS and evaluate K there:
// k_upd_matvec_z walks up a 64 x 8 column with a ring of planes, k_pre_resid_fwd and k_inv_post
// stage 32 x 8 x 16 tiles.  Each fused kernel performs the reference's arithmetic, value by value,
// in the reference's order, and the last of its workgroups runs the finaliser of its partials:
//   k_upd_matvec_z  = k_upd_p of iteration k-1 + k_matvec_dot of iteration k
//                     (phi += alpha p_old; p = z + beta p_old; q = K p; p.q; alpha)
//   k_pre_resid_fwd = k_pre + k_resid + k_fwd_red
//                     (r -= alpha q; z0 = omega D^-1 r; rr = r - K z0; level-1 transform; red sweep;
//                     the stop test on r)
//   k_inv_post      = k_inv_transform + k_post_rz + k_fin_rz
//                     (zt = z0 + T^T x1; z = zt + omega D^-1 (r - K zt); r.z, max|r|, r.r; beta and
//                     the stop test)
// r and p live in two buffers each (iteration k reads one and writes the other), so a tile never
// reads a neighbour tile's already-updated values.  Every kernel of CG iteration k returns at
// once when the stop test fired at an earlier iteration (DevState::stop_at < k).
#include "ctx.hpp"
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

constexpr int HB_IT = (HBN + TILE_THREADS - 1) / TILE_THREADS;   // halo blocks staged per thread
// Thread t stages halo blocks t + it TILE_THREADS.  The last round holds the HB_TAIL blocks left
// over; only the first waves of the workgroup, which hold them, load it (a whole wave skips it, a
// wave-uniform test that waits on no load), so the other waves spend no address arithmetic or
// memory instructions on a round with nothing to stage.
constexpr int HB_TAIL = HBN - (HB_IT - 1) * TILE_THREADS;
__device__ __forceinline__ bool stages_round(int t, int it) {
    return it + 1 < HB_IT || t < ((HB_TAIL + 63) & ~63);
}

// The tile of this workgroup.  Tiles are launched as a 1D grid and renumbered with xcd_block, so
// each XCD sweeps a contiguous run of tiles (x fastest, then y, then z) and neighbouring tiles
// share their halo lines in one L2.
struct Tile {
    int x, y, z;   // tile coordinates
    int lin;       // logical tile index (partial-sum slot)
};

__device__ __forceinline__ Tile this_tile(int N) {
    const int tx = (N + TX - 1) / TX, ty = (N + TY - 1) / TY;
    Tile t;
    t.lin = xcd_block((int)blockIdx.x, (int)gridDim.x);
    t.x = t.lin % tx;
    t.y = (t.lin / tx) % ty;
    t.z = t.lin / (tx * ty);
    return t;
}

// Tiles are staged in 2x2x2 blocks (the level-1 cells): halo block b (0..HBN-1) is the level-1
// block (I, J, K) = first block of the tile - 1 + (b % HBX, b / HBX % HBY, b / (HBX HBY)), read
// as four rows of two cells (real2 values, two mask bytes).  N is even, so a block is wholly
// inside or outside the box; one outside is read at the clamped block (field values, scaled by
// in = 0) with its mask at n3, the zero bytes past the end of the mask array.  Staging loads are
// unconditional, so all of a thread's loads are in flight together; only the LDS stores of the
// last, partial round are predicated.
struct HBlock {
    int h0;     // halo index of the block's (0, 0, 0) cell
    int i0;     // fine index of the (clamped) block's (0, 0, 0) cell
    int lc;     // level-1 cell of the clamped block (natural order: operator coefficients)
    int lx;     // its storage index in the level-1 x and r arrays (lsto)
    int o0;     // index of its (0, 0, 0) cell in the tile's own region (TX x TY x TZ), -1 in the halo
    bool in;    // block inside the box
};

__device__ __forceinline__ HBlock halo_block(int N, const Tile& tl, int b) {
    const int bi = b % HBX, bj = (b / HBX) % HBY, bk = b / (HBX * HBY);
    const int mb = N >> 1;
    const int I = tl.x * BTX - 1 + bi, J = tl.y * BTY - 1 + bj, K = tl.z * BTZ - 1 + bk;
    const int CI = min(max(I, 0), mb - 1), CJ = min(max(J, 0), mb - 1), CK = min(max(K, 0), mb - 1);
    HBlock r;
    r.in = I >= 0 && J >= 0 && K >= 0 && I < mb && J < mb && K < mb;
    r.h0 = ((2 * bk) * HY + 2 * bj) * HX + 2 * bi;
    r.i0 = ((2 * CK) * N + 2 * CJ) * N + 2 * CI;
    r.lc = (CK * mb + CJ) * mb + CI;
    r.lx = (int)lsto(mb, (size_t)r.lc, (CI + CJ + CK) & 1);
    const bool own = bi >= 1 && bi <= BTX && bj >= 1 && bj <= BTY && bk >= 1 && bk <= BTZ;
    r.o0 = own ? ((2 * (bk - 1)) * TY + 2 * (bj - 1)) * TX + 2 * (bi - 1) : -1;
    return r;
}

// the halo block of own block t (BTX x BTY x BTZ own blocks, one per thread)
__device__ __forceinline__ int own_hb(int t) {
    return ((t / (BTX * BTY) + 1) * HBY + (t / BTX) % BTY + 1) * HBX + t % BTX + 1;
}

// row rw = dy + 2 dz of a block: offset of its dx = 0 cell in the fine grid and in the halo
__device__ __forceinline__ int row_off(int N, int rw) { return (rw & 1) * N + (rw >> 1) * N * N; }
__device__ __forceinline__ int row_h(int rw) { return (rw & 1) * HX + (rw >> 1) * HX * HY; }
__device__ __forceinline__ int row_o(int rw) { return (rw & 1) * TX + (rw >> 1) * TX * TY; }
constexpr int ON = TX * TY * TZ;   // cells of a tile's own region

__device__ __forceinline__ real2 ld2(const real* __restrict__ f, int i) {
    return *reinterpret_cast<const real2*>(f + i);
}
__device__ __forceinline__ void st2(real* __restrict__ f, int i, real a, real b) {
    *reinterpret_cast<real2*>(f + i) = make_real2(a, b);
}
__device__ __forceinline__ int ldm2(const uint8_t* __restrict__ fm, int i) {
    return *reinterpret_cast<const uint16_t*>(fm + i);
}

// true when every cell of the thread's own block (halo index h0) is INTERIOR in every lane of the
// wave: K is then evaluated without the mask (kmul_tile_interior, kdiag = 6)
__device__ __forceinline__ bool wave_interior(const uint8_t* sm, int h0) {
    bool in = true;
#pragma unroll
    for (int rw = 0; rw < 4; ++rw) {
        const int m2 = *reinterpret_cast<const uint16_t*>(sm + h0 + row_h(rw));
        in = in && (m2 & (INTERIOR | (INTERIOR << 8))) == (INTERIOR | (INTERIOR << 8));
    }
    return __all(in);
}

// true when the four staged rows (two mask bytes each) of a 2x2x2 block are all INTERIOR cells
__device__ __forceinline__ bool block_interior(const int* m2) {
    bool in = true;
#pragma unroll
    for (int rw = 0; rw < 4; ++rw) in = in && (m2[rw] & (INTERIOR | (INTERIOR << 8))) == (INTERIOR | (INTERIOR << 8));
    return in;
}

// K p at halo cell h (0 at a solid cell)
template <bool INT>
__device__ __forceinline__ real kq(const real* P, const uint8_t* M, int h) {
    if (INT) return kmul_tile_interior(P, h);
    return M[h] ? kmul_tile(P, M, h) : RL(0.0);
}

// z0 + T^T x1 at a cell of parity (dz, dy, dx) from hx[c] = HS x_c: hsign(c) x_c = +-(HS x_c)
// exactly, so the sum equals the reference's c-ordered sum of hsign * x
__device__ __forceinline__ real inv_sum(real acc, const real* hx, int dz, int dy, int dx) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int e = (c & 1) * dx + ((c >> 1) & 1) * dy + ((c >> 2) & 1) * dz;
        acc = (e & 1) ? acc - hx[c] : acc + hx[c];
    }
    return acc;
}

// 2^-e with 2^e <= v < 2^(e+1): brings v into [1, 2) exactly (1 for v = 0 or non-finite)
__device__ __forceinline__ double pow2_scale(double v) { return (v > 0.0 && v < 1e300) ? ldexp(1.0, -ilogb(v)) : 1.0; }

// k_rhs (:602-627): r = -d(u*) - K phi0 (WARM; otherwise phi = 0); block max |u*| over fluid,
// block sum of b^2 (b = -d(u*)) for the relative-residual test and block max |r| (the scale of
// the first V-cycle's right-hand side with 16-bit level storage).  Every load is issued
// unconditionally (clamped neighbours; a solid cell's terms are computed and discarded), so they
// are all in flight together.  gated: the rhs of a device-side restart, a no-op unless the
// restart gate found the solve above tolerance (DevState::live).
template <bool WARM>
__global__ void __launch_bounds__(256) k_rhs(int N, int zc, const real* __restrict__ vel,
                                             const uint8_t* __restrict__ fm, real* __restrict__ r,
                                             real* __restrict__ phi, double* __restrict__ part, int nparts,
                                             const DevState* __restrict__ st, int gated) {
    __shared__ double sh[3 * 8];
    if (gated && !st->live) return;
    const int n3 = N * N * N;
    double mx = 0.0, bb = 0.0, mr = 0.0;
    CELL_LOOP_BEGIN
        const int m = fm[i];
        const real b = -div2_xyz(vel, n3, m, N, x, y, cz, i);
        real rv = b;
        if (WARM) rv -= kmul_xyz(phi, fm, N, x, y, cz, i, m);
        const real ux = vel[i], uy = vel[n3 + i], uz = vel[2 * n3 + i];
        const double sp = (double)sqrt(ux * ux + uy * uy + uz * uz);
        const bool fl = m != 0;
        if (!WARM) phi[i] = RL(0.0);
        r[i] = fl ? rv : RL(0.0);
        bb += fl ? (double)b * (double)b : 0.0;
        mr = fl ? fmax(mr, fabs((double)rv)) : mr;
        mx = fl ? fmax(mx, sp) : mx;
    CELL_LOOP_END
    double v3[3] = {mx, bb, mr};
    block_reduce<3, 0b101u>(v3, sh);
    const double a = v3[0], b2 = v3[1], c = v3[2];
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        part[block_linear()] = a;
        part[nparts + block_linear()] = b2;
        part[2 * nparts + block_linear()] = c;
    }
}

// k_fin_maxu (:730-735) + the CG state reset.  The first CG run of a projection (gated = 0) also
// clears the projection's restart bookkeeping; a gated restart run keeps it.
__global__ void __launch_bounds__(FIN_THREADS) k_fin_rhs(const double* __restrict__ part, int nparts,
                                                 const float* __restrict__ prm, DevState* __restrict__ st,
                                                 int gated) {
    __shared__ double sh[3 * FIN_THREADS / 64];
    if (gated && !st->live) return;
    // the three reductions in one pass (each in reduce_parts' order): a third of the load rounds
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    for_parts<3, 16>(part, nparts, {0, nparts, 2 * nparts}, [&](const double (&x)[3]) {
        v0 = fmax(v0, x[0]);
        v1 += x[1];
        v2 = fmax(v2, x[2]);
    });
    double v3[3] = {v0, v1, v2};
    block_reduce<3, 0b101u>(v3, sh);
    const double mx = v3[0], bb = v3[1], mr = v3[2];
    if (threadIdx.x == 0) {
        const float tol = prm[P_TOL] > 1.0e-6f ? prm[P_TOL] : 1.0e-6f;
        st->s[S_LSCALE] = pow2_scale(mr);
        st->s[S_MAXU] = mx;
        st->s[S_THR] = (double)(2.0f * tol * (float)mx);
        st->s[S_BB] = bb;
        st->it = 0;
        st->stop_at = INT_MAX;
        st->early = 0;
        st->conv = 0;
        if (!gated) {
            st->it_total = 0;
            st->restarts = 0;
            st->hit = 0;
            st->live = 1;
            st->fallback = 0;
        }
    }
}

// ---------------------------------------------------------------- finalisers
// The finalisers of a CG iteration run in the last workgroup of the launch whose partials they
// reduce, not as launches of their own.  The partials are written and read with agent-scope
// atomic stores and loads, which go to the device-coherent level and pass by the per-XCD L2s (an
// agent-scope release / acquire would write back and invalidate the whole L2 in every
// workgroup).  Thread 0 of each workgroup waits until its partials are written, then counts the
// workgroup in DevState::done; the workgroup that completes the count runs the finaliser and
// resets the count for the next launch.  A launch whose workgroups return at cg_inactive counts
// nothing.  Every workgroup reads the solver state it needs before it counts, so the finaliser's
// writes reach the launches after this one only.
__device__ __forceinline__ void put_part(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double get_part(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The finalisers' partial reads: PCHUNK per thread are issued before any is used (each is a
// round trip to the coherent level).  Thread t reads i = t, t + blockDim, t + 2 blockDim, ... in
// that order.
constexpr int PCHUNK = 8;
template <int NA, class F>
__device__ __forceinline__ void each_part(const double* __restrict__ part, int nparts, const int (&off)[NA], F&& use) {
    const int B = (int)blockDim.x;
    for (int i0 = (int)threadIdx.x; i0 < nparts; i0 += PCHUNK * B) {
        double v[PCHUNK][NA];
#pragma unroll
        for (int j = 0; j < PCHUNK; ++j) {
            const int i = i0 + j * B;
#pragma unroll
            for (int a = 0; a < NA; ++a) v[j][a] = i < nparts ? get_part(part + off[a] + i) : 0.0;
        }
#pragma unroll
        for (int j = 0; j < PCHUNK; ++j)
            if (i0 + j * B < nparts) use(v[j]);
    }
}

__device__ __forceinline__ bool last_block(DevState* st) {
    __shared__ int last;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_s_waitcnt(0);   // this thread's partial stores are complete
        const unsigned n = gridDim.x * gridDim.y * gridDim.z;
        const unsigned prev = __hip_atomic_fetch_add(&st->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = prev + 1u == n ? 1 : 0;
        if (last) __hip_atomic_store(&st->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return last != 0;
}

// alpha = rz / p.q (the finalize the reference folds into k_pre, :652-659)
__device__ void fin_pq(const double* __restrict__ part, int nparts, DevState* __restrict__ st, double* sh) {
    double v = 0.0;
    each_part<1>(part, nparts, {0}, [&](const double (&x)[1]) { v += x[0]; });
    const double pq = block_sum(v, sh);
    if (threadIdx.x == 0) st->s[S_ALPHA] = pq > 0.0 ? st->s[S_RZ] / pq : 0.0;
}

// k_fin_rz (:737-771) with FP64 scalars, plus the optional relative-residual criterion.  kmax is
// the iteration budget the host enqueued (iterations after it do not exist): the run stops there
// like at P_MAXIT, and DevState::hit records (bit hitbit) a stop that only the budget caused.
__device__ void fin_rz(const double* __restrict__ part, int nparts, const float* __restrict__ prm,
                       DevState* __restrict__ st, int k, int kmax, int hitbit, double* sh) {
    double va = 0.0, vb = 0.0, vc = 0.0;
    each_part<3>(part, nparts, {0, nparts, 2 * nparts}, [&](const double (&x)[3]) {
        va += x[0];
        vb = fmax(vb, x[1]);
        vc += x[2];
    });
    double v3[3] = {va, vb, vc};
    block_reduce<3, 0b010u>(v3, sh);
    const double rz_new = v3[0], mr = v3[1], rr = v3[2];
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
    st->s[S_LSCALE] = pow2_scale(mr);
    st->s[S_RR] = rr;
    st->s[S_BEST] = best;
    const double rel_tol = (double)prm[P_REL_TOL];
    const bool not_conv = rel_tol > 0.0 ? sqrt(rr) > rel_tol * sqrt(st->s[S_BB]) : mr > st->s[S_THR];
    const bool healthy = rz_new > 0.0 && mr <= 4.0 * best;
    const int maxit = (int)prm[P_MAXIT];
    const bool go = k < maxit && k < kmax && healthy && not_conv;
    st->conv = not_conv ? 0 : 1;
    if (!go) {
        st->stop_at = k;
        if (k >= kmax && k < maxit && healthy && not_conv) st->hit |= hitbit;
    }
}

// The stop test of iteration k on r alone, before its V-cycle (the max|r| and r.r partials of
// k_pre_resid_fwd are k_inv_post's, reduced in fin_rz's order): when it fires, the run stops at
// k as fin_rz would have made it, and the V-cycle and k_inv_post of iteration k are skipped
// (their z would not be used).  The reference computes that last z.
// LAST: iteration k is the last of the budget the host enqueued (kmax = k), launched without its
// V-cycle: the run stops here in any case, and unconverged it records the budget stop (bit
// hitbit of DevState::hit) as fin_rz would (its health test on max|r|; r.z is not formed).
template <bool LAST>
__device__ void fin_r(const double* __restrict__ part, int nparts, const float* __restrict__ prm,
                      DevState* __restrict__ st, int k, int hitbit, double* sh) {
    double vb = 0.0, vc = 0.0;
    each_part<2>(part, nparts, {nparts, 2 * nparts}, [&](const double (&x)[2]) {
        vb = fmax(vb, x[0]);
        vc += x[1];
    });
    double v2[2] = {vb, vc};
    block_reduce<2, 0b01u>(v2, sh);
    const double mr = v2[0], rr = v2[1];
    if (threadIdx.x != 0) return;
    const double rel_tol = (double)prm[P_REL_TOL];
    const bool not_conv = rel_tol > 0.0 ? sqrt(rr) > rel_tol * sqrt(st->s[S_BB]) : mr > st->s[S_THR];
    if (not_conv && !LAST) return;
    st->it = k;
    st->s[S_MAXR] = mr;
    st->s[S_RR] = rr;
    st->conv = not_conv ? 0 : 1;
    st->early = 1;
    st->stop_at = k;
    if (LAST && not_conv) {
        const double best = k == 0 ? mr : fmin(st->s[S_BEST], mr);
        if (k < (int)prm[P_MAXIT] && mr <= 4.0 * best) st->hit |= hitbit;
    }
}

// The same finalisers as launches of their own, for launches of many workgroups: there each
// workgroup's wait for its partials and its count, which keep its CU slot busy, cost more than
// the separate launch.
__global__ void __launch_bounds__(FIN_THREADS) k_fin_pq(const double* __restrict__ part, int nparts,
                                                DevState* __restrict__ st, int k) {
    __shared__ double sh[FIN_THREADS / 64];
    if (cg_inactive(st, k)) return;
    fin_pq(part, nparts, st, sh);
}

template <bool LAST>
__global__ void __launch_bounds__(FIN_THREADS) k_fin_r(const double* __restrict__ part, int nparts,
                                               const float* __restrict__ prm, DevState* __restrict__ st, int k,
                                               int hitbit) {
    __shared__ double sh[2 * FIN_THREADS / 64];
    if (cg_inactive(st, k)) return;
    fin_r<LAST>(part, nparts, prm, st, k, hitbit, sh);
}

__global__ void __launch_bounds__(FIN_THREADS) k_fin_rz(const double* __restrict__ part, int nparts,
                                                const float* __restrict__ prm, DevState* __restrict__ st, int k,
                                                int kmax, int hitbit) {
    __shared__ double sh[3 * FIN_THREADS / 64];
    if (cg_inactive(st, k)) return;
    fin_rz(part, nparts, prm, st, k, kmax, hitbit, sh);
}

// coordinate c (-2 .. N+1) of a halo -> the staged cell: the nearest inside 2x2x2 block, same parity
__device__ __forceinline__ int cpar(int c, int N) { return 2 * min(max(c >> 1, 0), (N >> 1) - 1) + (c & 1); }

// ---------------------------------------------------------------- z-marching matvec
// k_upd_p of iteration k-1 (:712-728, when UPD) + k_matvec_dot of iteration k (:629-646):
// p = z + beta p_old, phi += alpha p_old; q = K p; partial p.q (double) per workgroup.
// A walk up z: a workgroup owns a ZX x ZY column of cells over nz planes and keeps
// the planes K reaches (z - 2 .. z + 2, each with a 2-cell xy halo) in a ring of 6 LDS planes.
// While it computes plane z from the ring it has already issued the loads of plane z + 3, so the
// memory traffic of one plane overlaps the arithmetic of the previous one, and one barrier per
// plane suffices (the plane a lane stages is never one the other lanes still read).  Values, the
// staging of cells outside the box (the nearest inside block's cell, scaled by 0, mask 0) and K are
// those of the LDS tiles of the other CG kernels.
constexpr int ZX = 64, ZY = 8, ZC = 32;          // own cells of a workgroup: ZX x ZY x (at most) ZC
constexpr int ZT = (ZX / 2) * ZY;                // threads: two x cells each
constexpr int PW = ZX + 4, PH = ZY + 4;          // a ring plane with its xy halo
constexpr int PS = PW * PH;                      // cells per ring plane
constexpr int PPAIRS = PS / 2;                   // x pairs per ring plane
constexpr int PR = (PPAIRS + ZT - 1) / ZT;       // staging rounds per plane
constexpr int RING = 6;

// K at in-plane index h of the middle plane: P[j] = ring plane c + j - 2, M[j] = masks of plane
// c + j - 1 (kmul_tile's arithmetic)
template <bool INT>
__device__ __forceinline__ real kmul_ring(const real* const (&P)[5], const uint8_t* const (&M)[3], int h) {
    real d = RL(0.0);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        real m2, m1, c, p1, p2;
        int m0 = 0, mp = 0, mm = 0;
        if (ax < 2) {
            const int st = ax == 0 ? 1 : PW;
            m2 = P[2][h - 2 * st]; m1 = P[2][h - st]; c = P[2][h]; p1 = P[2][h + st]; p2 = P[2][h + 2 * st];
            if (!INT) { m0 = M[1][h]; mp = M[1][h + st]; mm = M[1][h - st]; }
        } else {
            m2 = P[0][h]; m1 = P[1][h]; c = P[2][h]; p1 = P[3][h]; p2 = P[4][h];
            if (!INT) { m0 = M[1][h]; mp = M[2][h]; mm = M[0][h]; }
        }
        if (INT) {
            const real gi = (p1 - c) + (c - m1);
            const real gp = (p2 - p1) + (p1 - c);
            const real gm = (c - m1) + (m1 - m2);
            d = d + (gi + gp);
            d = d - (gi + gm);
        } else {
            const real op = (real)((m0 >> (2 * ax)) & 1), om = (real)((m0 >> (2 * ax + 1)) & 1);
            const real opp = (real)((mp >> (2 * ax)) & 1), omp = (real)((mp >> (2 * ax + 1)) & 1);
            const real opm = (real)((mm >> (2 * ax)) & 1), omm = (real)((mm >> (2 * ax + 1)) & 1);
            const real gi = (RL(0.0) + op * (p1 - c)) + om * (c - m1);
            const real gp = (RL(0.0) + opp * (p2 - p1)) + omp * (p1 - c);
            const real gm = (RL(0.0) + opm * (c - m1)) + omm * (m1 - m2);
            d = d + op * (gi + gp);
            d = d - om * (gi + gm);
        }
    }
    return -d;
}

// the raw values of one ring plane a thread stages: x pairs t, t + ZT, ... of plane zp
struct PlaneRegs {
    real2 p[PR], z[PR];
    real c[PR];
    int m[PR];
};

// LEARN: the search direction is formed from z + P c (the coarse correction c of the learned
// preconditioner, constant over each 4^3 block's fluid cells; cz holds (N/4)^3 values)
template <bool UPD, bool LEARN>
__device__ __forceinline__ void plane_load(PlaneRegs& r, const real* __restrict__ p_in, const real* __restrict__ zf,
                                           const real* __restrict__ cz, const uint8_t* __restrict__ fm, int N, int x0,
                                           int y0, int zp) {
    const int zc = cpar(zp, N);
    const int M = N >> 2;
#pragma unroll
    for (int it = 0; it < PR; ++it) {
        const int pr = min((int)threadIdx.x + it * ZT, PPAIRS - 1);
        const int X = x0 - 2 + 2 * (pr % (PW / 2)), Y = y0 - 2 + pr / (PW / 2);
        const int cx = cpar(X, N), cy = cpar(Y, N);
        const int i = (zc * N + cy) * N + cx;
        if (UPD || !LEARN) r.p[it] = ld2(p_in, i);
        if (UPD || LEARN) r.z[it] = ld2(zf, i);
        if (LEARN) r.c[it] = cz[((zc >> 2) * M + (cy >> 2)) * M + (cx >> 2)];
        r.m[it] = ldm2(fm, i);
    }
}

template <bool UPD, bool LEARN>
__device__ __forceinline__ void plane_stage(const PlaneRegs& r, real* __restrict__ sp, uint8_t* __restrict__ sm, real b,
                                            int N, int x0, int y0, int zp) {
    const bool zin = (unsigned)zp < (unsigned)N;
#pragma unroll
    for (int it = 0; it < PR; ++it) {
        const int pr = (int)threadIdx.x + it * ZT;
        if (it + 1 < PR || pr < PPAIRS) {
            const int X = x0 - 2 + 2 * (pr % (PW / 2)), Y = y0 - 2 + pr / (PW / 2);
            const bool in = zin && (unsigned)X < (unsigned)N && (unsigned)Y < (unsigned)N;
            const real s = in ? RL(1.0) : RL(0.0);
            real v0, v1;
            if (UPD || LEARN) {
                real z0 = r.z[it].x, z1 = r.z[it].y;
                if (LEARN) {
                    z0 = z0 + ((r.m[it] & 255) ? r.c[it] : RL(0.0));
                    z1 = z1 + ((r.m[it] >> 8) ? r.c[it] : RL(0.0));
                }
                v0 = UPD ? z0 + b * r.p[it].x : z0;
                v1 = UPD ? z1 + b * r.p[it].y : z1;
            } else {
                v0 = r.p[it].x;
                v1 = r.p[it].y;
            }
            st2(sp, 2 * pr, v0 * s, v1 * s);
            *reinterpret_cast<uint16_t*>(sm + 2 * pr) = (uint16_t)(in ? r.m[it] : 0);
        }
    }
}

// K z of the ring at the two own cells (ox, ox + 1) of the middle plane (0 at non-fluid cells)
__device__ __forceinline__ void ring_kmul2(const real* const (&P)[5], const uint8_t* const (&M)[3], int h, real& q0,
                                           real& q1) {
    const int m2 = *reinterpret_cast<const uint16_t*>(M[1] + h);
    const bool interior = (m2 & (INTERIOR | (INTERIOR << 8))) == (INTERIOR | (INTERIOR << 8));
    if (__all(interior)) {
        q0 = kmul_ring<true>(P, M, h);
        q1 = kmul_ring<true>(P, M, h + 1);
    } else {
        q0 = (m2 & 255) ? kmul_ring<false>(P, M, h) : RL(0.0);
        q1 = (m2 >> 8) ? kmul_ring<false>(P, M, h + 1) : RL(0.0);
    }
}

// LEARN (the learned flexible CG): p = (z + P c) + beta p_old, and at k = 1 (!UPD) p = z + P c is
// written to p_out, which then holds the first search direction.
template <bool UPD, bool LEARN>
__global__ void __launch_bounds__(ZT) k_upd_matvec_z(int N, const real* __restrict__ z, const real* __restrict__ p_in,
                                                     real* __restrict__ p_out, real* __restrict__ phi,
                                                     const uint8_t* __restrict__ fm, real* __restrict__ q,
                                                     double* __restrict__ part, DevState* __restrict__ st, int k,
                                                     int fin, int nz, const real* __restrict__ cz) {
    __shared__ __align__(16) real sp[RING][PS];
    __shared__ __align__(16) uint8_t sm[RING][PS];
    __shared__ double sh[ZT / 64];
    if (cg_inactive(st, k)) return;
    const real a = (real)st->s[S_ALPHA], b = (real)st->s[S_BETA];
    // column: x fastest, then y, then the z chunk (XCD-renumbered, neighbouring columns in one L2)
    const int cx = (N + ZX - 1) / ZX, cy = (N + ZY - 1) / ZY;
    const int lin = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int x0 = (lin % cx) * ZX, y0 = ((lin / cx) % cy) * ZY, z0 = (lin / (cx * cy)) * nz;
    const int z1 = min(z0 + nz, N);
    const int t = (int)threadIdx.x;
    const int lx = 2 * (t % (ZX / 2)), ly = t / (ZX / 2);
    const int ox = x0 + lx, oy = y0 + ly;
    const bool own = ox < N && oy < N;
    const int h = (ly + 2) * PW + lx + 2;
    PlaneRegs pre;
    // planes z0 - 2 .. z0 + 1 into ring slots 0 .. 3 (plane zp lives in slot (zp - z0 + 2) % RING)
    for (int j = 0; j < 4; ++j) {
        plane_load<UPD, LEARN>(pre, p_in, z, cz, fm, N, x0, y0, z0 - 2 + j);
        plane_stage<UPD, LEARN>(pre, sp[j], sm[j], b, N, x0, y0, z0 - 2 + j);
    }
    plane_load<UPD, LEARN>(pre, p_in, z, cz, fm, N, x0, y0, z0 + 2);
    double s = 0.0;
    for (int zc = z0; zc < z1; ++zc) {
        const int base = zc - z0;
        plane_stage<UPD, LEARN>(pre, sp[(base + 4) % RING], sm[(base + 4) % RING], b, N, x0, y0, zc + 2);
        __syncthreads();
        if (zc + 3 <= z1 + 1) plane_load<UPD, LEARN>(pre, p_in, z, cz, fm, N, x0, y0, zc + 3);
        const int oi = (min(zc, N - 1) * N + min(oy, N - 1)) * N + min(ox, N - 2);
        real2 po, ph;
        if (UPD) {
            po = ld2(p_in, oi);
            ph = ld2(phi, oi);
        }
        const real* P[5] = {sp[base % RING], sp[(base + 1) % RING], sp[(base + 2) % RING], sp[(base + 3) % RING],
                            sp[(base + 4) % RING]};
        const uint8_t* M[3] = {sm[(base + 1) % RING], sm[(base + 2) % RING], sm[(base + 3) % RING]};
        real q0, q1;
        ring_kmul2(P, M, h, q0, q1);
        const real pn0 = P[2][h], pn1 = P[2][h + 1];
        s += (double)q0 * (double)pn0;
        s += (double)q1 * (double)pn1;
        if (own) {
            if (UPD) st2(phi, oi, ph.x + a * po.x, ph.y + a * po.y);
            if (UPD || LEARN) st2(p_out, oi, pn0, pn1);
            st2(q, oi, q0, q1);
        }
    }
    const double t2 = block_sum(s, sh);
    if (threadIdx.x == 0) put_part(part + blockIdx.x, t2);
    if (fin != 0 && last_block(st)) fin_pq(part, (int)gridDim.x, st, sh);
}

// planes per column: ZC, halved (down to 2) while the launch has fewer than 1024 workgroups, so
// that small grids still spread over the GPU (64^3: 256 columns of 2 planes)
static int zmarch_planes(int N) {
    const int cols = ((N + ZX - 1) / ZX) * ((N + ZY - 1) / ZY);
    int nz = ZC;
    while (nz > 2 && cols * ((N + nz - 1) / nz) < 1024) nz /= 2;
    return nz;
}

static unsigned zmarch_blocks(int N) {
    const int nz = zmarch_planes(N);
    return (unsigned)(((N + ZX - 1) / ZX) * ((N + ZY - 1) / ZY) * ((N + nz - 1) / nz));
}

// k_pre (:648-672) + k_resid (:674-685) + k_fwd_red (:426-441).  r -= alpha q (UPDR: k > 0;
// written to r_out) and z0 = omega D^-1 r are staged for the tile + halo; then thread t owns one
// 2x2x2 block of the tile: rr = r - K z0 on its 8 cells, the level-1 components and the
// zero-guess red sweep.
template <bool UPDR, class LT>
__global__ void __launch_bounds__(TILE_THREADS) k_pre_resid_fwd(int N, const real* __restrict__ r_in,
                                                                real* __restrict__ r_out, const real* __restrict__ q,
                                                                const uint8_t* __restrict__ fm, LvlT<LT> L, int do_red,
                                                                double* __restrict__ part, int nparts,
                                                                DevState* __restrict__ st, int k,
                                                                const float* __restrict__ prm, int fin, int hitbit) {
    __shared__ __align__(16) real sz[HN];
    __shared__ __align__(16) real sr[ON];   // r of the own region only (the halo needs z0 and the mask)
    __shared__ __align__(16) uint8_t sm[HN];
    __shared__ double sh[2 * TILE_THREADS / 64];
    if (cg_inactive(st, k)) return;
    const Tile tl = this_tile(N);
    const real a = (real)st->s[S_ALPHA];
    const int t = (int)threadIdx.x, n3 = N * N * N;
    const int bi = t % BTX, bj = (t / BTX) % BTY, bk = t / (BTX * BTY);
    const int mb = L.m;
    const int I = tl.x * BTX + bi, J = tl.y * BTY + bj, K = tl.z * BTZ + bk;
    const HBlock ob = halo_block(N, tl, own_hb(t));
    const size_t cell = (size_t)ob.lc;
    const bool red = ob.in && do_red && ((I + J + K) & 1) == 0;
    real dg[8] = {RL(0.0), RL(0.0), RL(0.0), RL(0.0), RL(0.0), RL(0.0), RL(0.0), RL(0.0)};
    if (red) ld8(L.diag + cell * 8, dg);
    const real sc = level_scale<LT>(st);   // 1 for `real` storage
    HBlock hb[HB_IT];
    real2 hr[HB_IT][4], hq[HB_IT][4];
    int hm[HB_IT][4];
#pragma unroll
    for (int it = 0; it < HB_IT; ++it) {
        if (!stages_round(t, it)) continue;
        hb[it] = halo_block(N, tl, min(t + it * TILE_THREADS, HBN - 1));
#pragma unroll
        for (int rw = 0; rw < 4; ++rw) {
            const int o = hb[it].i0 + row_off(N, rw);
            hr[it][rw] = ld2(r_in, o);
            if (UPDR) hq[it][rw] = ld2(q, o);
            hm[it][rw] = ldm2(fm, hb[it].in ? o : n3);
        }
    }
    __builtin_amdgcn_sched_barrier(0);   // every load above is issued before the first use
#pragma unroll
    for (int it = 0; it < HB_IT; ++it) {
        if (it + 1 < HB_IT || t + it * TILE_THREADS < HBN) {
            const real in = hb[it].in ? RL(1.0) : RL(0.0);
            // blocks of INTERIOR cells in every lane: kdiag = 6 (div6 is the division bit for bit)
            const bool fast = __all(block_interior(hm[it]));
#pragma unroll
            for (int rw = 0; rw < 4; ++rw) {
                real rv[2] = {hr[it][rw].x, hr[it][rw].y};
                if (UPDR) {
                    rv[0] = rv[0] - a * hq[it][rw].x;
                    rv[1] = rv[1] - a * hq[it][rw].y;
                }
                const int h = hb[it].h0 + row_h(rw);
                real zv[2];
                if (fast) {
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        rv[dx] = rv[dx] * in;
                        zv[dx] = div6(OMEGA_JAC * rv[dx]);
                    }
                } else {
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        const int m = (hm[it][rw] >> (8 * dx)) & 255;
                        rv[dx] = rv[dx] * in;
                        zv[dx] = kopen(m) ? OMEGA_JAC * rv[dx] / kdiag(m) : RL(0.0);
                    }
                }
                st2(sz, h, zv[0], zv[1]);
                if (hb[it].o0 >= 0) st2(sr, hb[it].o0 + row_o(rw), rv[0], rv[1]);
                *reinterpret_cast<uint16_t*>(sm + h) = (uint16_t)hm[it][rw];
            }
        }
    }
    __syncthreads();
    // max|r| and r.r of the tile, in k_inv_post's order (the same partials it writes): the stop
    // test of this iteration can run before the V-cycle
    double mx = 0.0, rr2 = 0.0;
#pragma unroll
    for (int rw = 0; rw < 4; ++rw)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const real rv = sr[ob.o0 + row_o(rw) + dx];
            mx = fmax(mx, fabs((double)rv));
            rr2 += (double)rv * (double)rv;
        }
    double v2[2] = {mx, rr2};
    block_reduce<2, 0b01u>(v2, sh);
    const double bmx = v2[0], brr = v2[1];
    if (threadIdx.x == 0) {
        put_part(part + nparts + tl.lin, bmx);
        put_part(part + 2 * nparts + tl.lin, brr);
    }
    // fin 1: the stop test on r (fin_r); 2: the budget's last iteration (fin_r<true>)
    if (fin != 0 && last_block(st)) {
        if (fin == 1) fin_r<false>(part, nparts, prm, st, k, hitbit, sh);
        else fin_r<true>(part, nparts, prm, st, k, hitbit, sh);
    }
    if (!ob.in) return;
    real kz[8];
    if (wave_interior(sm, ob.h0)) {
#pragma unroll
        for (int c = 0; c < 8; ++c) kz[c] = kq<true>(sz, sm, ob.h0 + row_h(c >> 1) + (c & 1));
    } else {
#pragma unroll
        for (int c = 0; c < 8; ++c) kz[c] = kq<false>(sz, sm, ob.h0 + row_h(c >> 1) + (c & 1));
    }
    real rr[8];
#pragma unroll
    for (int rw = 0; rw < 4; ++rw) {
        const int h = ob.h0 + row_h(rw), ho = ob.o0 + row_o(rw);
        const real r0 = sr[ho], r1 = sr[ho + 1];
        if (UPDR) st2(r_out, ob.i0 + row_off(N, rw), r0, r1);
        rr[2 * rw] = sm[h] ? r0 - kz[2 * rw] : RL(0.0);
        rr[2 * rw + 1] = sm[h + 1] ? r1 - kz[2 * rw + 1] : RL(0.0);
    }
    real out[8], xo[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        real acc = RL(0.0);
#pragma unroll
        for (int s = 0; s < 8; ++s) acc += hsign(c, s >> 2, (s >> 1) & 1, s & 1) * rr[s];
        out[c] = acc * sc;
        // the red sweep reads the level-1 right-hand side as stored
        const real rs = as_stored<LT>(out[c]);
        xo[c] = dg[c] <= L.dmin ? RL(0.0) : rs / dg[c];
    }
    st8(L.r + (size_t)ob.lx * 8, out);
    if (red) st8(L.x + (size_t)ob.lx * 8, xo);
}

// k_inv_transform (:352-365) + k_post_rz (:687-710): zt = z0 + T^T x1 (0 in solids) staged for
// the tile + halo, one 2x2x2 block at a time from its 8 level-1 values, then
// z = zt + omega D^-1 (r - K zt) on the tile and the block partials r.z, max|r|, r.r.  At k = 0
// z is also written to p_init (the first search direction, k_upd_p init).
template <class LT>
__global__ void __launch_bounds__(TILE_THREADS) k_inv_post(int N, LvlT<LT> L, const real* __restrict__ r,
                                                           const uint8_t* __restrict__ fm, real* __restrict__ z,
                                                           real* __restrict__ p_init, double* __restrict__ part,
                                                           int nparts, DevState* __restrict__ st, int k,
                                                           const float* __restrict__ prm, int fin, int kmax,
                                                           int hitbit) {
    __shared__ __align__(16) real szt[HN];
    __shared__ __align__(16) real sr[ON];   // r of the own region only
    __shared__ __align__(16) uint8_t sm[HN];
    __shared__ double sh[3 * TILE_THREADS / 64];
    if (cg_inactive(st, k)) return;
    const Tile tl = this_tile(N);
    const int t = (int)threadIdx.x, n3 = N * N * N;
    const real isc = RL(1.0) / level_scale<LT>(st);   // undoes the scale of the level-1 right-hand side
    HBlock hb[HB_IT];
    real xv[HB_IT][8];
    real2 hr[HB_IT][4];
    int hm[HB_IT][4];
#pragma unroll
    for (int it = 0; it < HB_IT; ++it) {
        if (!stages_round(t, it)) continue;
        hb[it] = halo_block(N, tl, min(t + it * TILE_THREADS, HBN - 1));
        ld8(L.x + (size_t)hb[it].lx * 8, xv[it]);
#pragma unroll
        for (int rw = 0; rw < 4; ++rw) {
            const int o = hb[it].i0 + row_off(N, rw);
            hr[it][rw] = ld2(r, o);
            hm[it][rw] = ldm2(fm, hb[it].in ? o : n3);
        }
    }
    __builtin_amdgcn_sched_barrier(0);   // every load above is issued before the first use
#pragma unroll
    for (int it = 0; it < HB_IT; ++it) {
        if (it + 1 < HB_IT || t + it * TILE_THREADS < HBN) {
            const real in = hb[it].in ? RL(1.0) : RL(0.0);
            real hx[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) hx[c] = HS * ((xv[it][c] * isc) * in);
            // blocks of INTERIOR cells in every lane: kdiag = 6 (div6 is the division bit for bit)
            const bool fast_b = __all(block_interior(hm[it]));
#pragma unroll
            for (int rw = 0; rw < 4; ++rw) {
                const real rv[2] = {hr[it][rw].x * in, hr[it][rw].y * in};
                real zt[2];
                if (fast_b) {
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) zt[dx] = inv_sum(div6(OMEGA_JAC * rv[dx]), hx, rw >> 1, rw & 1, dx);
                } else {
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        const int m = (hm[it][rw] >> (8 * dx)) & 255;
                        const real acc =
                            inv_sum(kopen(m) ? OMEGA_JAC * rv[dx] / kdiag(m) : RL(0.0), hx, rw >> 1, rw & 1, dx);
                        zt[dx] = m ? acc : RL(0.0);
                    }
                }
                const int h = hb[it].h0 + row_h(rw);
                st2(szt, h, zt[0], zt[1]);
                if (hb[it].o0 >= 0) st2(sr, hb[it].o0 + row_o(rw), rv[0], rv[1]);
                *reinterpret_cast<uint16_t*>(sm + h) = (uint16_t)hm[it][rw];
            }
        }
    }
    __syncthreads();
    // the thread's own block; a block outside the box is staged as mask 0, value 0 (terms 0)
    const HBlock ob = halo_block(N, tl, own_hb(t));
    real kz[8];
    const bool fast = wave_interior(sm, ob.h0);
    if (fast) {
#pragma unroll
        for (int c = 0; c < 8; ++c) kz[c] = kq<true>(szt, sm, ob.h0 + row_h(c >> 1) + (c & 1));
    } else {
#pragma unroll
        for (int c = 0; c < 8; ++c) kz[c] = kq<false>(szt, sm, ob.h0 + row_h(c >> 1) + (c & 1));
    }
    double s = 0.0, mx = 0.0, rr = 0.0;
#pragma unroll
    for (int rw = 0; rw < 4; ++rw) {
        const int h = ob.h0 + row_h(rw), ho = ob.o0 + row_o(rw);
        real zv[2];
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const real rv = sr[ho + dx];
            if (fast) {
                zv[dx] = szt[h + dx] + div6(OMEGA_JAC * (rv - kz[2 * rw + dx]));
            } else {
                const int m = sm[h + dx];
                zv[dx] = kopen(m) ? szt[h + dx] + OMEGA_JAC * (rv - kz[2 * rw + dx]) / kdiag(m) : RL(0.0);
            }
            s += (double)rv * (double)zv[dx];
            mx = fmax(mx, fabs((double)rv));
            rr += (double)rv * (double)rv;
        }
        if (ob.in) {
            const int i = ob.i0 + row_off(N, rw);
            st2(z, i, zv[0], zv[1]);
            if (k == 0) st2(p_init, i, zv[0], zv[1]);
        }
    }
    double v3[3] = {s, mx, rr};
    block_reduce<3, 0b010u>(v3, sh);
    const double a = v3[0], b = v3[1], c = v3[2];
    if (threadIdx.x == 0) {
        const int bl = tl.lin;
        put_part(part + bl, a);
        put_part(part + nparts + bl, b);
        put_part(part + 2 * nparts + bl, c);
    }
    if (fin != 0 && last_block(st)) fin_rz(part, nparts, prm, st, k, kmax, hitbit, sh);
}

// Device-side restart decision after the true residual of a CG run (the host-polled form reads
// the same test on the host): above the relative tolerance, the next gated run restarts CG from
// the true residual with phi as its warm start; otherwise every gated kernel after it is a
// no-op (live = 0, stop_at = -1, which also makes that run's phi update a no-op).
// learned: the run before was the learned flexible CG, which also restarts (the MGPCG fallback)
// after a stall or breakdown, or when it stopped unconverged at its iteration limit.
__global__ void k_restart_gate(const float* __restrict__ prm, DevState* __restrict__ st, int learned) {
    if (threadIdx.x != 0) return;
    const double rel_tol = (double)prm[P_REL_TOL];
    bool done = !(rel_tol > 0.0) || sqrt(st->s[S_TRUE_RR]) <= rel_tol * sqrt(st->s[S_BB]);
    if (learned) {
        if (!st->conv) st->fallback = 1;   // stopped unconverged at the learned limit or budget
        done = done && !st->fallback;
    }
    if (done) {
        st->live = 0;
        st->stop_at = -1;
    } else {
        st->it_total += st->it;
        st->restarts += 1;
        st->live = 1;
    }
}

// The phi update of the stopping iteration J (k_upd_p :723-725 of the last loop body):
// phi += alpha_J p_{J-1}, where p_{J-1} is in p_buf[(J - 1) & 1].
__global__ void __launch_bounds__(256) k_phi_final(int N, int zc, real* __restrict__ phi,
                                                   const real* __restrict__ p0, const real* __restrict__ p1,
                                                   const DevState* __restrict__ st) {
    const int J = st->stop_at;
    if (J < 1 || J == INT_MAX) return;
    const real a = (real)st->s[S_ALPHA];
    const real* __restrict__ p = ((J - 1) & 1) ? p1 : p0;
    CELL_LOOP_BEGIN
        phi[i] = phi[i] + a * p[i];
    CELL_LOOP_END
}

// k_apply (:166-175): u -= g(phi) on fluid, u = 0 in solids
__global__ void __launch_bounds__(256) k_apply(int N, int zc, real* __restrict__ vel,
                                               const uint8_t* __restrict__ fm, const real* __restrict__ phi) {
    const int n3 = N * N * N;
    CELL_LOOP_BEGIN
        const int m = fm[i];
        if (m == 0) {
            vel[i] = RL(0.0); vel[n3 + i] = RL(0.0); vel[2 * n3 + i] = RL(0.0);
            continue;
        }
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
            vel[ax * n3 + i] = vel[ax * n3 + i] - g_ax_xyz(phi, m, N, i, ax == 0 ? x : (ax == 1 ? y : cz), ax);
    CELL_LOOP_END
}

// True residual ||b - K phi||^2 with b = -d(u*) (fresh matvec, FP64 accumulation)
__global__ void __launch_bounds__(256) k_true_res(int N, int zc, const real* __restrict__ vel,
                                                  const uint8_t* __restrict__ fm, const real* __restrict__ phi,
                                                  double* __restrict__ part, const DevState* __restrict__ st,
                                                  int gated) {
    __shared__ double sh[8];
    if (gated && !st->live) return;
    const int n3 = N * N * N;
    double s = 0.0;
    CELL_LOOP_BEGIN
        const int m = fm[i];   // a solid cell's terms are computed (clamped loads) and discarded
        const double v = -(double)div2_xyz(vel, n3, m, N, x, y, cz, i) - (double)kmul_xyz(phi, fm, N, x, y, cz, i, m);
        s += m != 0 ? v * v : 0.0;
    CELL_LOOP_END
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[block_linear()] = t;
}

__global__ void __launch_bounds__(FIN_THREADS) k_fin_true(const double* __restrict__ part, int nparts,
                                                  DevState* __restrict__ st, int gated) {
    __shared__ double sh[FIN_THREADS / 64];
    if (gated && !st->live) return;
    const double s = reduce_parts_sum(part, nparts, sh);
    if (threadIdx.x == 0) st->s[S_TRUE_RR] = s;
}

// |div u| dx with the mask computed from solid (k_div_abs :156-163 + k_mask)
__global__ void __launch_bounds__(256) k_div_abs(int N, int zc, const real* __restrict__ vel,
                                                 const real* __restrict__ solid, real* __restrict__ out) {
    const int n3 = N * N * N;
    CELL_LOOP_BEGIN
        const int m = mask_from_solid(solid, N, cz, y, x);
        out[i] = m == 0 ? RL(0.0) : RL(0.5) * fabs(div2_xyz(vel, n3, m, N, x, y, cz, i));
    CELL_LOOP_END
}

// K p at fluid cells (parity tests of the operator)
__global__ void __launch_bounds__(256) k_kmul(int N, int zc, const real* __restrict__ p,
                                              const uint8_t* __restrict__ fm, real* __restrict__ out) {
    CELL_LOOP_BEGIN
        const int m = fm[i];
        out[i] = m ? kmul_xyz(p, fm, N, x, y, cz, i, m) : RL(0.0);
    CELL_LOOP_END
}

// ================================================================= host side
// Launches of at most FOLD_MAX_BLOCKS workgroups (about one round on the GPU) run their
// finaliser in their last workgroup; larger ones are followed by a finaliser launch.
constexpr int FOLD_MAX_BLOCKS = 512;

void Pressure::init_tiles() {
    tgrid = dim3((N + TX - 1) / TX, (N + TY - 1) / TY, (N + TZ - 1) / TZ);
    ntiles = (int)(tgrid.x * tgrid.y * tgrid.z);
}

real* Pressure::r_buf(int k) const { return R[k & 1]; }

void Pressure::precond(int k, hipStream_t s, const float* prm) {
    precond_front(k, s, prm, prm ? 1 : 0);
    precond_back(k, s);
}

void Pressure::precond_front(int k, hipStream_t s, const float* prm, int fin, int hitbit) {
    real* r_in = k > 0 ? R[(k - 1) & 1] : R[0];
    real* r_out = R[k & 1];
    const int fold = ntiles <= FOLD_MAX_BLOCKS ? fin : 0;
    with_mg_type(mg_kind, [&](auto tag) {
        using LT = decltype(tag);
        const LvlT<LT> L1 = lvl_as<LT>(lv[0]);
        if (k > 0)
            hipLaunchKernelGGL((k_pre_resid_fwd<true, LT>), dim3(ntiles), dim3(TILE_THREADS), 0, s, N, r_in, r_out, q,
                               fm, L1, S > 0 ? 1 : 0, part, ntiles, st, k, prm, fold, hitbit);
        else
            hipLaunchKernelGGL((k_pre_resid_fwd<false, LT>), dim3(ntiles), dim3(TILE_THREADS), 0, s, N, r_in, r_out, q,
                               fm, L1, S > 0 ? 1 : 0, part, ntiles, st, k, prm, fold, hitbit);
    });
    if (fin == 1 && fold == 0)
        hipLaunchKernelGGL(k_fin_r<false>, dim3(1), dim3(FIN_THREADS), 0, s, part, ntiles, prm, st, k, hitbit);
    if (fin == 2 && fold == 0)
        hipLaunchKernelGGL(k_fin_r<true>, dim3(1), dim3(FIN_THREADS), 0, s, part, ntiles, prm, st, k, hitbit);
}

void Pressure::precond_back(int k, hipStream_t s, const float* prm, int kmax, int hitbit) {
    real* r_out = R[k & 1];
    vcycle_levels(k, s);
    const int fold = prm != nullptr && ntiles <= FOLD_MAX_BLOCKS ? 1 : 0;
    with_mg_type(mg_kind, [&](auto tag) {
        using LT = decltype(tag);
        hipLaunchKernelGGL(k_inv_post<LT>, dim3(ntiles), dim3(TILE_THREADS), 0, s, N, lvl_as<LT>(lv[0]), r_out, fm, z,
                           P[0], part, ntiles, st, k, prm, fold, kmax, hitbit);
    });
    if (prm != nullptr && fold == 0)
        hipLaunchKernelGGL(k_fin_rz, dim3(1), dim3(FIN_THREADS), 0, s, part, ntiles, prm, st, k, kmax, hitbit);
}

void Pressure::rhs(const real* vel, const float* prm, bool warm, hipStream_t s, bool gated, bool restart) {
    const Geom& g = restart ? geom_r : geom;
    const int gt = gated ? 1 : 0;
    if (warm)
        hipLaunchKernelGGL(k_rhs<true>, g.grid, g.block, 0, s, N, g.zc, vel, fm, R[0], phi, part, nblocks(g), st, gt);
    else
        hipLaunchKernelGGL(k_rhs<false>, g.grid, g.block, 0, s, N, g.zc, vel, fm, R[0], phi, part, nblocks(g), st, gt);
    hipLaunchKernelGGL(k_fin_rhs, dim3(1), dim3(FIN_THREADS), 0, s, part, nblocks(g), prm, st, gt);
}

void Pressure::restart_gate(const float* prm, hipStream_t s, bool learned) {
    hipLaunchKernelGGL(k_restart_gate, dim3(1), dim3(64), 0, s, prm, st, learned ? 1 : 0);
}

void Pressure::iteration(int k, const float* prm, hipStream_t s, int kmax, int hitbit) {
    if (k > 0) {
        // p_{k-1} = z_{k-1} + beta p_{k-2} (k >= 2); p_0 = z_0 was written at k = 0
        const unsigned zb = zmarch_blocks(N);
        const int fold = (int)zb <= FOLD_MAX_BLOCKS ? 1 : 0;
        if (k >= 2)
            hipLaunchKernelGGL((k_upd_matvec_z<true, false>), dim3(zb), dim3(ZT), 0, s, N, z, P[(k - 2) & 1], P[(k - 1) & 1],
                               phi, fm, q, part_pq, st, k, fold, zmarch_planes(N), nullptr);
        else
            hipLaunchKernelGGL((k_upd_matvec_z<false, false>), dim3(zb), dim3(ZT), 0, s, N, z, P[0], P[0], phi, fm, q,
                               part_pq, st, k, fold, zmarch_planes(N), nullptr);
        if (!fold) hipLaunchKernelGGL(k_fin_pq, dim3(1), dim3(FIN_THREADS), 0, s, part_pq, (int)zb, st, k);
    }
    if (k >= kmax) {
        // the budget's last iteration: the stop test on r decides; its z would not be used
        precond_front(k, s, prm, 2, hitbit);
        return;
    }
    precond_front(k, s, prm, 1);
    precond_back(k, s, prm, kmax, hitbit);
}

void Pressure::finish(hipStream_t s, bool restart) {
    const Geom& g = restart ? geom_r : geom;
    hipLaunchKernelGGL(k_phi_final, g.grid, g.block, 0, s, N, g.zc, phi, P[0], P[1], st);
}

void Pressure::apply(real* vel, hipStream_t s) {
    const Geom& g = geom;
    hipLaunchKernelGGL(k_apply, g.grid, g.block, 0, s, N, g.zc, vel, fm, phi);
}

void Pressure::true_residual(const real* vel, hipStream_t s, bool gated, bool restart) {
    const Geom& g = restart ? geom_r : geom;
    const int gt = gated ? 1 : 0;
    hipLaunchKernelGGL(k_true_res, g.grid, g.block, 0, s, N, g.zc, vel, fm, phi, part, st, gt);
    hipLaunchKernelGGL(k_fin_true, dim3(1), dim3(FIN_THREADS), 0, s, part, nblocks(g), st, gt);
}

// One kernel of the projection on the current buffers, for timing (ids relative to
// VFX_K_MATVEC), in its steady-state form (iteration k >= 2).  What a kernel would write into
// phi, p, q or r goes to `scratch` (3 N^3 floats) instead, so the solver state is untouched.
bool Pressure::launch_single(int which, real* scratch, hipStream_t s) {
    const size_t n3 = (size_t)N * N * N;
    switch (which) {
        case 0:
            hipLaunchKernelGGL((k_upd_matvec_z<true, false>), dim3(zmarch_blocks(N)), dim3(ZT), 0, s, N, z, P[0], scratch,
                               scratch + n3, fm, scratch + 2 * n3, part_pq, st, 2, 0, zmarch_planes(N), nullptr);
            return true;
        case 1:
            with_mg_type(mg_kind, [&](auto tag) {
                using LT = decltype(tag);
                hipLaunchKernelGGL(k_inv_post<LT>, dim3(ntiles), dim3(TILE_THREADS), 0, s, N, lvl_as<LT>(lv[0]), R[0],
                                   fm, z, P[0], part, ntiles, st, 2, nullptr, 0, 0, 0);
            });
            return true;
        case 5:
            with_mg_type(mg_kind, [&](auto tag) {
                using LT = decltype(tag);
                hipLaunchKernelGGL((k_pre_resid_fwd<true, LT>), dim3(ntiles), dim3(TILE_THREADS), 0, s, N, R[0],
                                   scratch, q, fm, lvl_as<LT>(lv[0]), S > 0 ? 1 : 0, part, ntiles, st, 2, nullptr,
                                   0, 0);
            });
            return true;
        default: return launch_single_mg(which, s);
    }
}

void launch_div_abs(const Geom& g, const real* vel, const real* solid, real* out, hipStream_t s) {
    hipLaunchKernelGGL(k_div_abs, g.grid, g.block, 0, s, g.N, g.zc, vel, solid, out);
}

void launch_kmul(const Geom& g, const real* p, const uint8_t* fm, real* out, hipStream_t s) {
    hipLaunchKernelGGL(k_kmul, g.grid, g.block, 0, s, g.N, g.zc, p, fm, out);
}

}  // namespace vfx

#if !VFX_F64
namespace vfx {

// ---------------------------------------------------------------- learned flexible CG on the device
// The coarse-correction preconditioner z = V(r) + P c, c = net(R (r - K V(r))) (R: 4^3 block mean,
// P: constant over a block's fluid cells), sequenced on the device like MGPCG.  Iteration k runs
// the MGPCG iteration's kernels up to the V-cycle's z_V = V(r_k) (k_inv_post without its
// finaliser), then k_resid_pool, the network (gated on the same stop state) and k_fcg_fin, which
// forms the Polak-Ribiere beta = z_k.(r_k - r_{k-1}) / z_{k-1}.r_{k-1} (Golub & Ye 1999) of the
// full z_k without writing it: z_k.r = z_V.r + sum_B c_B sum_{fluid i in B} r_i, and the same for
// r_{k-1}.  The next iteration's k_upd_matvec_z<.., true> forms p = (z_V + P c) + beta p_old.

// t = r - K z_V at fluid cells (0 elsewhere) as 4^3 block means (the network input), the block sums
// of r and r_prev over fluid cells, the partials z_V.r_prev and sum of squared block means (the
// input's scale).  A z-marching column like k_upd_matvec_z; nz is a multiple of 4, so every
// 4^3 block lies in one column: each thread sums its x pair over 4 planes, and after the next
// plane's barrier 32 threads add the 8 pair sums of each of the column's 16 x 2 blocks.
constexpr int PBX = ZX / 4, PBY = ZY / 4;
__global__ void __launch_bounds__(ZT) k_resid_pool(int N, const real* __restrict__ zv, const real* __restrict__ r,
                                                   const real* __restrict__ rp, const uint8_t* __restrict__ fm,
                                                   real* __restrict__ cin, real* __restrict__ sr, real* __restrict__ srp,
                                                   double* __restrict__ part, DevState* __restrict__ st, int k, int nz) {
    __shared__ __align__(16) real sp[RING][PS];
    __shared__ __align__(16) uint8_t sm[RING][PS];
    __shared__ real sacc[3][ZY][ZX / 2];
    __shared__ double sh[ZT / 64];
    if (cg_inactive(st, k)) return;
    const int M = N >> 2;
    const int cx = (N + ZX - 1) / ZX, cy = (N + ZY - 1) / ZY;
    const int lin = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int x0 = (lin % cx) * ZX, y0 = ((lin / cx) % cy) * ZY, z0 = (lin / (cx * cy)) * nz;
    const int z1 = min(z0 + nz, N);
    const int t = (int)threadIdx.x;
    const int lx = 2 * (t % (ZX / 2)), ly = t / (ZX / 2);
    const int ox = x0 + lx, oy = y0 + ly;
    const bool own = ox < N && oy < N;
    const int h = (ly + 2) * PW + lx + 2;
    PlaneRegs pre;
    for (int j = 0; j < 4; ++j) {
        plane_load<false, false>(pre, zv, nullptr, nullptr, fm, N, x0, y0, z0 - 2 + j);
        plane_stage<false, false>(pre, sp[j], sm[j], RL(0.0), N, x0, y0, z0 - 2 + j);
    }
    plane_load<false, false>(pre, zv, nullptr, nullptr, fm, N, x0, y0, z0 + 2);
    // the blocks of the plane group that starts at zg: written by threads 0..31
    double dot = 0.0, sq = 0.0;
    auto flush = [&](int zg) {
        if (t < PBX * PBY) {
            const int bx = t % PBX, by = t / PBX;
            real a[3] = {RL(0.0), RL(0.0), RL(0.0)};
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) a[q] = a[q] + (sacc[q][4 * by + rr][2 * bx] + sacc[q][4 * by + rr][2 * bx + 1]);
            const int BX = (x0 >> 2) + bx, BY = (y0 >> 2) + by;
            if (BX < M && BY < M) {
                const int B = ((zg >> 2) * M + BY) * M + BX;
                const real mean = a[0] * RL(0.015625);
                cin[B] = mean;
                sr[B] = a[1];
                srp[B] = a[2];
                sq += (double)mean * (double)mean;
            }
        }
    };
    real at = RL(0.0), ar = RL(0.0), arp = RL(0.0);
    for (int zc = z0; zc < z1; ++zc) {
        const int base = zc - z0;
        plane_stage<false, false>(pre, sp[(base + 4) % RING], sm[(base + 4) % RING], RL(0.0), N, x0, y0, zc + 2);
        __syncthreads();
        if (base > 0 && (base & 3) == 0) flush(zc - 4);
        if (zc + 3 <= z1 + 1) plane_load<false, false>(pre, zv, nullptr, nullptr, fm, N, x0, y0, zc + 3);
        const int oi = (min(zc, N - 1) * N + min(oy, N - 1)) * N + min(ox, N - 2);
        const real2 rv = ld2(r, oi), pv = ld2(rp, oi);
        const real* P[5] = {sp[base % RING], sp[(base + 1) % RING], sp[(base + 2) % RING], sp[(base + 3) % RING],
                            sp[(base + 4) % RING]};
        const uint8_t* Mk[3] = {sm[(base + 1) % RING], sm[(base + 2) % RING], sm[(base + 3) % RING]};
        real q0, q1;
        ring_kmul2(P, Mk, h, q0, q1);
        const int m2 = *reinterpret_cast<const uint16_t*>(Mk[1] + h);
        const bool f0 = own && (m2 & 255) != 0, f1 = own && (m2 >> 8) != 0;
        at = at + ((f0 ? rv.x - q0 : RL(0.0)) + (f1 ? rv.y - q1 : RL(0.0)));
        ar = ar + ((f0 ? rv.x : RL(0.0)) + (f1 ? rv.y : RL(0.0)));
        arp = arp + ((f0 ? pv.x : RL(0.0)) + (f1 ? pv.y : RL(0.0)));
        dot += f0 ? (double)P[2][h] * (double)pv.x : 0.0;
        dot += f1 ? (double)P[2][h + 1] * (double)pv.y : 0.0;
        if ((base & 3) == 3) {
            sacc[0][ly][lx >> 1] = at;
            sacc[1][ly][lx >> 1] = ar;
            sacc[2][ly][lx >> 1] = arp;
            at = ar = arp = RL(0.0);
        }
    }
    __syncthreads();
    flush(z1 - 4);
    const double d = block_sum(dot, sh);
    const double q2 = block_sum(sq, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = d;
        part[gridDim.x + blockIdx.x] = q2;
    }
}

// the network input's scale from k_resid_pool's partials: scal[0] = 1 / rms, scal[1] = rms of the
// M^3 block means (both 1 for a zero input)
__global__ void __launch_bounds__(FIN_THREADS) k_lrms_fin(const double* __restrict__ part, int nb, int m3,
                                                  double* __restrict__ scal, const DevState* __restrict__ st, int k) {
    __shared__ double sh[FIN_THREADS / 64];
    if (cg_inactive(st, k)) return;
    const double s = reduce_parts_sum(part + nb, nb, sh);
    if (threadIdx.x == 0) {
        const double rms = sqrt(s / (double)m3);
        scal[0] = rms > 0.0 ? 1.0 / rms : 1.0;
        scal[1] = rms > 0.0 ? rms : 1.0;
    }
}

// c_B = rms * out_B on blocks with a fluid cell (0 elsewhere) -> cz; partials c.sr and c.srp; the last
// workgroup: r.z and z.r_prev of the full z, beta, the scalars of the iteration and the flexible
// CG's own stop tests: a breakdown (r.z not positive or not finite) or a stall (||r|| fell less
// than 2x over 3 iterations) stops the run with DevState::fallback set, and the gated MGPCG restart
// run after the true residual finishes the solve from the current phi.
constexpr int FCGF_BLOCKS = 256;
__global__ void __launch_bounds__(256) k_fcg_fin(int m3, const float* __restrict__ out, const double* __restrict__ scal,
                                                 const uint8_t* __restrict__ cmask, const real* __restrict__ sr,
                                                 const real* __restrict__ srp, real* __restrict__ cz,
                                                 const double* __restrict__ pinv, int nparts,
                                                 const double* __restrict__ prp, int nrp, double* __restrict__ fpart,
                                                 DevState* __restrict__ st, int k) {
    __shared__ double sh[256 / 64];
    if (cg_inactive(st, k)) return;
    const float rms = (float)scal[1];
    double a = 0.0, b = 0.0;
    for (int B = (int)(blockIdx.x * blockDim.x + threadIdx.x); B < m3; B += (int)(gridDim.x * blockDim.x)) {
        const real c = cmask[B] ? (real)(out[B] * rms) : RL(0.0);
        cz[B] = c;
        a += (double)c * (double)sr[B];
        b += (double)c * (double)srp[B];
    }
    const double ba = block_sum(a, sh), bb = block_sum(b, sh);
    const int G = (int)gridDim.x;
    if (threadIdx.x == 0) {
        put_part(fpart + blockIdx.x, ba);
        put_part(fpart + G + blockIdx.x, bb);
    }
    if (!last_block(st)) return;
    double rzv = 0.0, mrv = 0.0, rrv = 0.0, zpv = 0.0, crz = 0.0, czp = 0.0;
    each_part<3>(pinv, nparts, {0, nparts, 2 * nparts}, [&](const double (&x)[3]) {
        rzv += x[0];
        mrv = fmax(mrv, x[1]);
        rrv += x[2];
    });
    each_part<1>(prp, nrp, {0}, [&](const double (&x)[1]) { zpv += x[0]; });
    each_part<2>(fpart, G, {0, G}, [&](const double (&x)[2]) {
        crz += x[0];
        czp += x[1];
    });
    const double rz_v = block_sum(rzv, sh), mr = block_max(mrv, sh), rr = block_sum(rrv, sh);
    const double zrp_v = block_sum(zpv, sh), c_rz = block_sum(crz, sh), c_zrp = block_sum(czp, sh);
    if (threadIdx.x != 0) return;
    const double rz = rz_v + c_rz, zrp = zrp_v + c_zrp, rn = sqrt(rr);
    st->it = k;
    st->s[S_MAXR] = mr;
    st->s[S_RR] = rr;
    st->s[S_LSCALE] = pow2_scale(mr);
    st->s[S_BEST] = k == 0 ? mr : fmin(st->s[S_BEST], mr);
    const bool bad = !(rz > 0.0) || !isfinite(rz) || !isfinite(rn);
    const bool stall = k >= 3 && rn > 0.5 * st->s[S_FCG_H0 + k % 3];
    st->s[S_FCG_H0 + k % 3] = rn;
    st->s[S_BETA] = k == 0 ? 0.0 : (rz - zrp) / st->s[S_RZ];
    st->s[S_RZ] = rz;
    st->conv = 0;
    if (bad || stall) {
        st->stop_at = k;
        st->fallback = 1;
    }
}

// planes per column of k_resid_pool: a multiple of 4
static int rpool_planes(int N) { return std::max(4, zmarch_planes(N)); }

unsigned resid_pool_blocks(int N) {
    const int nz = rpool_planes(N);
    return (unsigned)(((N + ZX - 1) / ZX) * ((N + ZY - 1) / ZY) * ((N + nz - 1) / nz));
}

void launch_upd_matvec_learned(const Pressure& P, int k, const real* cz, hipStream_t s) {
    const int N = P.N;
    const unsigned zb = zmarch_blocks(N);
    const int fold = (int)zb <= FOLD_MAX_BLOCKS ? 1 : 0;
    if (k >= 2)
        hipLaunchKernelGGL((k_upd_matvec_z<true, true>), dim3(zb), dim3(ZT), 0, s, N, P.z, P.P[(k - 2) & 1],
                           P.P[(k - 1) & 1], P.phi, P.fm, P.q, P.part_pq, P.st, k, fold, zmarch_planes(N), cz);
    else
        hipLaunchKernelGGL((k_upd_matvec_z<false, true>), dim3(zb), dim3(ZT), 0, s, N, P.z, P.P[0], P.P[0], P.phi, P.fm,
                           P.q, P.part_pq, P.st, k, fold, zmarch_planes(N), cz);
    if (!fold) hipLaunchKernelGGL(k_fin_pq, dim3(1), dim3(FIN_THREADS), 0, s, P.part_pq, (int)zb, P.st, k);
}

void launch_resid_pool(const Pressure& P, int k, real* cin, real* sr, real* srp, double* lpart, double* scal,
                       hipStream_t s) {
    const int N = P.N, M = N >> 2;
    const unsigned nb = resid_pool_blocks(N);
    const real* r = P.R[k & 1];
    const real* rp = k > 0 ? P.R[(k - 1) & 1] : r;
    hipLaunchKernelGGL(k_resid_pool, dim3(nb), dim3(ZT), 0, s, N, P.z, r, rp, P.fm, cin, sr, srp, lpart, P.st, k,
                       rpool_planes(N));
    hipLaunchKernelGGL(k_lrms_fin, dim3(1), dim3(FIN_THREADS), 0, s, lpart, (int)nb, M * M * M, scal, P.st, k);
}

void launch_fcg_fin(const Pressure& P, int k, const float* out, const double* scal, const uint8_t* cmask,
                    const real* sr, const real* srp, real* cz, const double* lpart, double* fpart, hipStream_t s) {
    const int M = P.N >> 2, m3 = M * M * M;
    const int G = std::max(1, std::min(FCGF_BLOCKS, (m3 + 255) / 256));
    hipLaunchKernelGGL(k_fcg_fin, dim3(G), dim3(256), 0, s, m3, out, scal, cmask, sr, srp, cz, P.part, P.ntiles,
                       lpart, (int)resid_pool_blocks(P.N), fpart, P.st, k);
}

// every single-precision bit pattern x: div6(x) against the IEEE quotient x / 6 (two NaN results
// count as equal); the mismatches are added to *bad
__global__ void __launch_bounds__(256) k_check_div6(unsigned long long* __restrict__ bad) {
    unsigned long long n = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
        const float x = __uint_as_float((uint32_t)i);
        const float a = div6(x), b = x / 6.0f;
        if (!(a != a && b != b) && __float_as_uint(a) != __float_as_uint(b)) ++n;
    }
    if (n) atomicAdd(bad, n);
}

}  // namespace vfx

extern "C" int vfx_check_div6(unsigned long long* mismatches, void* stream) {
    using namespace vfx;
    if (!mismatches) return arg_error("null pointer");
    hipLaunchKernelGGL(k_check_div6, dim3(16384), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), mismatches);
    return launch_ok("div6 check") ? VFX_OK : VFX_E_HIP;
}
#endif  // !VFX_F64
