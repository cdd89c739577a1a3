/* vfxsim3d.h — C ABI of the MI355X (gfx950) sim3d smoke/fire solver library (libvfxsim3d.so).
 *
 * The drop-in boundary for the reference's `vfx.sim3d.Solver3D` (vfx/sim3d/solver.py): the
 * Python host layer (neural-vfx_amd/vfx/sim3d/solver.py) keeps the reference API and calls these
 * entry points through ctypes.  Plain pointers and sizes only:
 *   - field pointers are device pointers borrowed from the caller (torch.Tensor.data_ptr());
 *     `data` is (6, N, N, N) float32 [c][z][y][x] with channels density, vel_x, vel_y, vel_z,
 *     temperature, fuel (vfx/core/fields3d.py:12-16); `solid` is (N, N, N) float32, 1 = obstacle;
 *   - `prm_host` is a host array of vfx_p_size() floats in the reference's parameter layout
 *     (vfx/sim3d/kernels.py:11-42, filled like Solver3D._upload, solver.py:331-359);
 *   - `stream` is a hipStream_t (torch.cuda.current_stream().cuda_stream), 0 = null stream.
 * Every function returns 0 on success or a VFX_E_* code; vfx_last_error() gives the message.
 * The library allocates device memory only in vfx_ctx_create and vfx_hires_attach; calls on one
 * context must not overlap (one context per Solver3D, as the reference keeps one solver per grid).
 * vfx_step and vfx_project do not wait for the device (VFX_OPT_SYNC): the host runs ahead by up to
 * 16 steps (a ring of pinned parameter slots, the reference's 4-slot ring at solver.py:331-335).
 */
#ifndef VFXSIM3D_H
#define VFXSIM3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VFX_OK 0
#define VFX_E_ARG 1     /* shape / pointer / size assertion (the reference's AssertionError) */
#define VFX_E_VALUE 2   /* bad value (the reference's ValueError, e.g. >32 emitters) */
#define VFX_E_HIP 3     /* HIP runtime error */

/* stage selectors of vfx_step_stage */
#define VFX_STAGE_ADVECT 0   /* obstacle setup + MacCormack advection + mass ratio: solver.py:422-430 */
#define VFX_STAGE_FORCES 1   /* sources, diffusion, vorticity confinement, forces: solver.py:432-455 */
#define VFX_STAGE_PROJECT 2  /* pressure projection: pressure.py:940-950 */
#define VFX_STAGE_SCALARS 3  /* the finer scalar grid's stage (solver.py:241-259); a no-op without one */

/* options of vfx_set_option */
#define VFX_OPT_TRUE_RESIDUAL 1  /* 1: compute ||b - K phi|| / ||b|| after each projection */
#define VFX_OPT_RESTARTS 2       /* CG restarts from the true residual while it exceeds the
                                    relative tolerance (needs P_REL_TOL > 0); default 0 */
#define VFX_OPT_COARSE_M 3       /* multigrid levels with m <= this run in one workgroup */
#define VFX_OPT_STD_STENCIL 4    /* 0: always read the level operators (no standard-cell table);
                                    results are identical either way, default 1 */
#define VFX_OPT_MG_STORAGE 5     /* multigrid level storage: 0 FP32 (default), 1 FP16, 2 BF16; FP32
                                    arithmetic either way, CG vectors FP32, dots and alpha/beta FP64 */
#define VFX_OPT_FCG_MG 6         /* without a network: 1 runs the projection as flexible CG with the MG
                                    V-cycle as preconditioner (the learned path's solver, checked against
                                    PCG), 2 as flexible CG without a preconditioner (plain CG); 0 off */
#define VFX_OPT_LEARNED_MAXIT 7  /* iteration budget of flexible CG before the MGPCG fallback (100) */
#define VFX_OPT_LEARNED_SWEEPS 8 /* damped-Jacobi corrections z += 0.8 D^-1 (r - K z) after each network
                                    application (0..8, default 2) */
#define VFX_OPT_CG_BUDGET 9      /* device sequencing: 0 (default) adapts the PCG iteration budget from
                                    the counts of recent projections; B > 0 enqueues exactly
                                    min(B, P_MAXIT) iterations (the ones after the stop test are no-ops
                                    on the device) */
#define VFX_OPT_CELLS_PER_THREAD 10  /* z-cells per thread of the cell kernels: 1, 2, 4, 8 (default 2
                                    from 128^3, 1 below); results equal up to FP64 reduction order */
#define VFX_OPT_PRESSURE_MODE 11 /* 0 MGPCG (default); 1 the fused projection (pressure_fused.py):
                                    sweeps + V-cycles as a stationary iteration, adaptive tail while
                                    P_TOL > 0 (up to P_MAXIT cycles); FP32 / FP64 level storage */
#define VFX_OPT_FUSED_CYCLES 12  /* fused: V-cycles always run (pressure_cycles, default 2) */
#define VFX_OPT_FUSED_SMOOTH 13  /* fused: exact-K sweeps before / between / after (pressure_smooth, 4) */
#define VFX_OPT_FUSED_SMOOTHER 14 /* fused: 0 Chebyshev weights on [0.5, 2] (default), 1 Jacobi 0.8 */
#define VFX_OPT_SYNC 15          /* projection sequencing.  0 (default) device: no host synchronisation
                                    in vfx_step / vfx_project (the reference's on-device loop,
                                    pressure.py:940-950).  The host enqueues an iteration budget (the
                                    kernels after the device's stop test are no-ops), and with restarts
                                    on the true residual, a device-side restart decision and one gated
                                    restart run with its own budget.  Budgets follow the counts of recent
                                    projections, read back lazily; a projection that runs out of its budget
                                    stops there, like the reference at pressure_max_iters, is counted in
                                    vfx_pressure_stats out[11], and the budget grows for the next.
                                    1 polled: the host reads the stop flag every iteration and restarts
                                    while the true residual is above tolerance (up to VFX_OPT_RESTARTS).
                                    The learned coarse correction follows this setting; the full-resolution
                                    network's flexible CG is always host-polled. */
#define VFX_OPT_GRAPH 16         /* 1 (default): in device mode vfx_step replays the whole step as a
                                    captured hipGraph, one per set of host-side launch decisions (buffer
                                    pointers, budgets, warm start, options), up to 4 per context
                                    (solver.py:388-420 keeps an LRU of 4 CUDA graphs); 0: direct launches */

typedef struct vfx_ctx vfx_ctx;

/* Number of floats of the parameter buffer (kernels.py:42 P_SIZE = 576). */
int vfx_p_size(void);

/* Allocate every buffer of a res^3 solver (Solver3D.__init__, solver.py:86-125;
 * PressureSolver3D.__init__, pressure.py:789-818).  res must be even (pressure.py:790) and at
 * most 512 (VFX_E_ARG above: the kernels' channel offsets c res^3 + i are 32-bit). */
int vfx_ctx_create(int res, vfx_ctx** out);
int vfx_ctx_destroy(vfx_ctx* ctx);

/* One whole step in place: advection, forces, projection (Solver3D.step, solver.py:146-155). */
int vfx_step(vfx_ctx* ctx, float* data, const float* solid, const float* prm_host, void* stream);

/* One stage of the step (Solver3D.step_profiled's graphs, solver.py:157-184).  The advection
 * stage uploads prm_host; the later stages reuse the uploaded parameters (prm_host may be NULL). */
int vfx_step_stage(vfx_ctx* ctx, int stage, float* data, const float* solid, const float* prm_host, void* stream);

/* Projection only, with dt = 0 forces (Solver3D.project, solver.py:266-272). */
int vfx_project(vfx_ctx* ctx, float* data, const float* solid, const float* prm_host, void* stream);

/* MacCormack advection of C channels by `vel` (3,N,N,N), no obstacles, no mass correction
 * (Solver3D.advect, solver.py:289-308).  out is (C,N,N,N) and must not alias field. */
int vfx_advect(vfx_ctx* ctx, const float* field, int C, const float* vel, float* out, float dt, void* stream);

/* vfx_advect without a context and without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same kernels and results as vfx_advect on a res^3
 * context; the scratch is a fixed 28 MiB device array per GPU (held until the process ends), so
 * res^3 <= 2^20 cells (res even, e.g. up to 100^3; VFX_E_VALUE above).  Calls that overlap on one
 * GPU (concurrent streams or threads) share that array and race. */
int vfx_advect_noalloc(const float* field, int C, const float* vel, float* out, int res, float dt, void* stream);

/* |div u| dx per cell, 0 in solids (record_div + k_div_abs, pressure.py:952-956, :156-163);
 * the caller normalises by max|u| like Solver3D.divergence (solver.py:274-287). */
int vfx_div_abs(vfx_ctx* ctx, const float* data, const float* solid, float* out, void* stream);

/* PCG iterations of the last projection (Solver3D.pressure_iters, solver.py:383-386); waits for
 * the last projection to finish. */
int vfx_pressure_iters(vfx_ctx* ctx, int* out);

/* Diagnostics of the last projection (waits for it): out[0] iterations (all restarts),
 * out[1] max|u*|, out[2] ||b||_2, out[3] recurrence ||r||_2, out[4] max|r|,
 * out[5] true ||b - K phi||_2 (when VFX_OPT_TRUE_RESIDUAL is on, else -1), out[6] restarts,
 * out[7] flexible-CG iterations (learned path), out[8] 1 if it fell back to MGPCG,
 * out[9] 1 if it converged (the true residual judges when computed), out[10] its first CG run's
 * iteration budget (0 when host-polled).  Over the context's life so far: out[11] projections
 * that stopped at a budget unconverged, out[12] unconverged projections, out[13] step graphs
 * captured, out[14] step graph replays (out[11], out[12] count the projections read back so far). */
#define VFX_STATS 15
int vfx_pressure_stats(vfx_ctx* ctx, double* out, int n);

int vfx_set_option(vfx_ctx* ctx, int key, double value);

/* Copy the warm-start pressure phi (N^3 floats, device) out of / into the solver. */
int vfx_get_phi(vfx_ctx* ctx, float* dst, void* stream);
int vfx_set_phi(vfx_ctx* ctx, const float* src, void* stream);

/* Copy the advected state of the last advection stage (6,N,N,N) and the mass ratio.  With a finer
 * scalar grid attached only the velocity channels 1..3 are advected on this grid. */
int vfx_get_advected(vfx_ctx* ctx, float* dst, double* mass_ratio, void* stream);

/* Operator access for parity tests: rebuild masks and multigrid operators for `solid`, then
 * out = K p (vfx_kmul), z = M^-1 r (vfx_precond, the V(1,1) preconditioner), or the level-l
 * operators in component-fastest layout diag [m^3][8], off [m^3][24] (vfx_get_ops; returns
 * the level count in *nlev and the level size in *m; FP32 level storage only).  With 16-bit level
 * storage vfx_precond scales the level-1 right-hand side by the power of two that brings max|r|
 * over fluid cells into [1, 2) (the solver takes the scale from the CG residual). */
int vfx_kmul(vfx_ctx* ctx, const float* solid, const float* p, float* out, void* stream);

/* Self-check with no counterpart in the reference, for tests: runs the CG kernels' division by
 * diag(K) = 6 at INTERIOR cells (a multiply and one fused-multiply-add correction instead of the
 * division sequence) on every single-precision bit pattern and adds the number of results that
 * differ from the IEEE quotient x / 6 to the device counter *mismatches (zero it first).  FP32
 * build only. */
int vfx_check_div6(unsigned long long* mismatches, void* stream);
int vfx_precond(vfx_ctx* ctx, const float* solid, const float* r, float* z, void* stream);
int vfx_get_ops(vfx_ctx* ctx, const float* solid, int level, float* diag, float* off, int* nlev, int* m,
                void* stream);

/* kernel ids of vfx_time_kernel */
#define VFX_K_ADV_SL 0     /* semi-Lagrangian pass of the advection (advect_sl, kernels.py:124-140) */
#define VFX_K_ADV_MC 1     /* MacCormack pass + clamp + mass partials (advect_mc, kernels.py:143-178) */
#define VFX_K_SOURCES 2    /* sources (kernels.py:182-250) */
#define VFX_K_DIFFUSE 3    /* heat + fuel diffusion (kernels.py:268-286) */
#define VFX_K_MATVEC 4     /* p / phi update + q = K p + p.q (k_upd_p + k_matvec_dot, pressure.py:629-646,712-728) */
#define VFX_K_POST_RZ 5    /* inverse transform + post-Jacobi + r.z, max|r|, r.r (pressure.py:352-365,687-710) */
#define VFX_K_SMOOTH0 6    /* one colour of GS on the finest multigrid level (k_smooth_h, pressure.py:290-303) */
#define VFX_K_RESTRICT0 7  /* residual restriction finest -> next level (k_restrict_red, pressure.py:306-316,444-459) */
#define VFX_K_PROLONG0 8   /* prolongation + black sweep on the finest level (k_prolong_black, pressure.py:396-423) */
#define VFX_K_PRE 9        /* r -= alpha q, pre-Jacobi, residual, Hadamard transform, red sweep
                              (k_pre + k_resid + k_fwd_red, pressure.py:648-685,426-441) */
#define VFX_K_COARSE 14    /* single-workgroup coarse V-cycle (k_coarse_block, pressure.py:550-591) */
#define VFX_K_FADV_SL 15   /* fine scalar grid: semi-Lagrangian pass (advect_sl_fine, hires.py:29-48) */
#define VFX_K_FADV_MC 16   /* fine scalar grid: MacCormack pass + mass partials (advect_mc_fine, hires.py:51-89) */
#define VFX_K_FSOURCES 17  /* fine scalar grid: sources (sources_fine, hires.py:92-139) */
#define VFX_K_FDIFFUSE 18  /* fine scalar grid: one diffusion substep (diffuse_fine, hires.py:151-170) */
#define VFX_K_FRESTRICT 19 /* fine -> coarse box restriction (restrict_scalars, hires.py:179-219) */

/* Average device time (ms) of `reps` back-to-back launches of one kernel on the solver's current
 * buffers, bracketed by HIP events on `stream` (measurement support for the benchmark; call it
 * after a step so the buffers hold live data).  Leaves the solver state as it found it. */
int vfx_time_kernel(vfx_ctx* ctx, int which, int reps, void* stream, float* ms_avg);

/* ---------------------------------------------------------------- finer scalar grid
 * Solver3D(res, ..., scalar_res=S) (solver.py:86-125, :187-264; hires.py).  The caller owns the fine
 * fields: hi (3, S, S, S) float32 density, temperature (K), fuel, and hi_solid (S, S, S) float32,
 * 1 = obstacle, both covering the simulation's world box.  Once attached, vfx_step carries the
 * scalars on the fine grid by the res^3 velocity and box-restricts them into data's density,
 * temperature and fuel channels before the forces read them.  S >= res (VFX_E_VALUE otherwise, the
 * reference's ValueError, solver.py:219-220), S <= 1024.  Allocates the fine scratch
 * (7 S^3 floats + S^3 bytes); S = 0 detaches and frees it.  hi_solid is read every step. */
int vfx_hires_attach(vfx_ctx* ctx, int S, float* hi, const float* hi_solid);

/* Solver3D.init_hires_from_state (solver.py:204-215): hi = the density / temperature / fuel of
 * `data` resized trilinearly to S^3 (half-pixel centres), 0 / ambient / 0 where hi_solid >= 0.5. */
int vfx_hires_init(vfx_ctx* ctx, const float* data, float ambient, void* stream);

/* ---------------------------------------------------------------- paired-run guidance
 * vfx/datagen3d/guide3d.py (Guide3D, tracking_error), the guidance between the 64^3 runtime run and
 * the 256^3 run of vfx/datagen3d/sim.py:52-79.  FP32 only.  Fields are (6, n, n, n) float32 states. */
typedef struct vfx_guide vfx_guide;

/* Guide3D.__init__ (guide3d.py:91-100): hr_res = f lr_res for an integer f, lr_res a multiple of 16
 * in [16, 128]; cfg = 7 floats k_vel, k_scalar, k_fuel, cutoff, eps_density, eps_temp, log_clamp
 * (GuideConfig3D, guide3d.py:33-41); hr_solid (hr^3 float32) is read once, as the reference keeps
 * its fluid mask.  Allocates the guide's buffers; synchronises `stream` once. */
int vfx_guide_create(int hr_res, int lr_res, const float* cfg, const float* hr_solid, void* stream, vfx_guide** out);
int vfx_guide_destroy(vfx_guide* guide);

/* Guide3D.velocity (guide3d.py:105-111): before the HR step, in place on hr_data. */
int vfx_guide_velocity(vfx_guide* guide, float* hr_data, const float* lr_data, void* stream);

/* Guide3D.scalars (guide3d.py:113-130): after the HR step, in place on hr_data. */
int vfx_guide_scalars(vfx_guide* guide, float* hr_data, const float* lr_data, void* stream);

/* lowpass (guide3d.py:60-68) of C <= 3 channels (C, lr, lr, lr) into out (may not alias x). */
int vfx_guide_lowpass(vfx_guide* guide, const float* x, int C, float* out, void* stream);

/* tracking_error (guide3d.py:133-145): out3 (device, 3 floats) = relative L2 at res^3 of hr vs lr
 * (density, temperature - ambient, velocity); scratch holds 12 res^3 floats (device). */
int vfx_tracking_error(const float* hr_data, int hr_res, const float* lr_data, int lr_res, int res, float* scratch,
                       float* out3, void* stream);

/* ---------------------------------------------------------------- volume renderer
 * vfx/render3d/renderer.py VolumeRenderer (default path, return_aux and the temporal resolve; no detail
 * field).  FP32 only.  params: VFX_RENDER_PARAMS floats, RenderParams3D's fields in declaration order
 * (renderer.py:31-66) with the 3-tuples flattened: steps, light_res, k_abs, k_shadow, k_solid, k_amb,
 * k_emit, emit_min, emit_scale, soot, fire_light, phase_g, rim, ms_a, ms_b, ms_c, exposure, contrast,
 * bloom, bloom_thresh, t_lo, t_hi, t_ref, d_eps, light_dir[3], key_col[3], amb_col[3],
 * smoke_albedo[3], bg_top[3], bg_bot[3], floor_col[3], grid_col[3], obst_col[3], glow_coarse,
 * floor.  lut: the 256 x 3 float blackbody table (vfx/render/blackbody.py). */
#define VFX_RENDER_PARAMS 53
typedef struct vfx_render vfx_render;

/* VolumeRenderer.__init__ / set_output (renderer.py:96-156): march resolution
 * max(8, round(out * march_scale)); allocates the renderer's buffers. */
int vfx_render_create(int out_h, int out_w, float march_scale, const float* params, const float* lut,
                      vfx_render** out);
int vfx_render_destroy(vfx_render* r);
int vfx_render_set_output(vfx_render* r, int out_h, int out_w, float march_scale);
/* set_params (renderer.py:168-174): a new light_res reallocates the light buffers. */
int vfx_render_set_params(vfx_render* r, const float* params, const float* lut);

/* VolumeRenderer.draw (renderer.py:281-386): dens / temp (D, H, W) float32 device, solid (sD, sH, sW)
 * float32 device or NULL (upload_solid != 0 when it changed since the last draw), cam = 16 host floats
 * (camera_basis, renderer.py:69-82), reuse_volume != 0 re-renders the last prepared volume, out =
 * (out_h, out_w, 3) uint8 device.  ms_prep / ms_view (host, optional): device time of the volume prep
 * and of the view part (synchronises when given). */
int vfx_render_draw(vfx_render* r, const float* dens, const float* temp, int D, int H, int W, const float* solid,
                    int sD, int sH, int sW, int upload_solid, const float* cam, int reuse_volume, uint8_t* out,
                    void* stream, float* ms_prep, float* ms_view);
/* VolumeRenderer.draw(return_aux=True) (renderer.py:295-393; kernels.py:288-489 march_aux, :603-626
 * aux_resolve): the same frame as vfx_render_draw, bit for bit, plus aux_out = (4, out_h, out_w) half
 * device: alpha (1 - the medium's transmittance), opacity-weighted depth along the ray, and motion
 * (x, y) in output pixels (current minus previous projection of the opacity-weighted mean point moved
 * back by vel * vel_dt); depth and motion are 0 where alpha < 1e-3.  vel: (3, uD, uH, uW) float32
 * device in m/s, or NULL for no medium motion (upload_vel != 0 when it changed since the last aux
 * draw); cam_prev: 16 host floats, the previous sim step's camera (pass cam when there is none). */
int vfx_render_draw_aux(vfx_render* r, const float* dens, const float* temp, int D, int H, int W, const float* solid,
                        int sD, int sH, int sW, int upload_solid, const float* cam, int reuse_volume, uint8_t* out,
                        const float* vel, int uD, int uH, int uW, int upload_vel, const float* cam_prev, float vel_dt,
                        uint16_t* aux_out, void* stream, float* ms_prep, float* ms_view);
/* Temporal amortization (renderer.py:130-163, :331-346, :485-492; temporal.py:70-150): on != 0 makes
 * every draw offset its ray-start jitter by a golden-ratio sequence and blend the march into a history
 * reprojected by camera motion (at the medium's opacity-weighted depth, else the floor, else the
 * domain centre), with a 3x3 variance clip in (luma-compressed YCoCg, T); bloom and composite run on
 * the resolved buffer.  alpha: base current weight (0.1), gamma: clip width (1), k_t: gain of the T
 * change (4), alpha_tick: minimum current weight on draws with a new volume (0.5).  The history is
 * invalid after a resize and after vfx_render_reset_history (VolumeRenderer.set_temporal /
 * reset_history, renderer.py:157-163). */
int vfx_render_set_temporal(vfx_render* r, int on, float alpha, float gamma, float k_t, float alpha_tick);
int vfx_render_reset_history(vfx_render* r);
/* The march buffer of the last draw (after the temporal resolve when that is on), (mh, mw, 4) float:
 * premultiplied radiance, transmittance. */
int vfx_render_get_hdr(vfx_render* r, float* dst, int* mh, int* mw, void* stream);

/* Message of the last failed call on this thread. */
int vfx_last_error(char* buf, int n);

/* ---------------------------------------------------------------- learned preconditioner
 * A 3-level 3D U-Net (16 / 32 / 64 channels, 3^3 convolutions, ReLU, average-pool down,
 * nearest-neighbour up, skip concatenation) used as the preconditioner of flexible CG.  BF16
 * weights and activations, FP32 accumulation on MFMA, FP32 master weights trained with Adam.
 * No counterpart in the reference (its solver is MG-preconditioned CG only).  One network object
 * serves one grid size (res % 4 == 0).  Parameters are flat FP32, per convolution in execution
 * order: weight [cout][cin][3][3][3] (torch conv3d layout), then bias [cout]. */
typedef struct vfx_net vfx_net;
int vfx_net_create(int res, vfx_net** out);
int vfx_net_destroy(vfx_net* net);
int vfx_net_param_count(vfx_net* net, int* count);
int vfx_net_get_params(vfx_net* net, float* host);   /* host array of param_count floats */
int vfx_net_set_params(vfx_net* net, const float* host);   /* also resets the Adam state */
/* raw network on a device input (2, res, res, res): channel 0 (residual), channel 1 (fluid 0/1)
 * -> out (res^3), device pointers */
int vfx_net_forward(vfx_net* net, const float* in, float* out, void* stream);
/* after vfx_net_forward: d loss / d out (res^3, device) -> grad (param_count floats, device; may
 * be NULL to keep it inside the network) */
int vfx_net_backward(vfx_net* net, const float* grad_out, float* grad, void* stream);
/* one Adam step with grad (device, or NULL for the last backward's), bias-corrected */
int vfx_net_adam(vfx_net* net, const float* grad, float lr, float beta1, float beta2, float eps, void* stream);
/* forward + energy-norm loss e.K e / phi.b (e = net(b) - phi, K of `solid` as the solver builds
 * it) + backward + Adam, on one (b, phi) pair of the solver's grid size; *loss is written on the
 * host (synchronises) when not NULL */
int vfx_net_train_step(vfx_net* net, vfx_ctx* ctx, const float* solid, const float* b, const float* phi, float lr,
                       double* loss, void* stream);
/* the same for the coarse correction: a network of a quarter of the solver's grid size trained on
 * a fine pair (r_V, e_V), the residual and the error a V-cycle left (K e_V = r_V): c = P net(R r_V)
 * (R the 4^3 block mean, P constant over each block's fluid cells), loss e.K e / e_V.r_V with
 * e = c - e_V in the fine energy norm.  Uses the context's step scratch. */
int vfx_net_train_step_coarse(vfx_net* net, vfx_ctx* ctx, const float* solid, const float* rv, const float* ev,
                              float lr, double* loss, void* stream);
/* attach a network (or NULL to detach) to a solver context: its projections then run flexible CG
 * (Polak-Ribiere form).  A network of the solver's grid size is the preconditioner; one of a
 * quarter of it is the coarse correction of the MG V-cycle, z = V(r) + P net(R (r - K V(r))).
 * Falls back to plain MGPCG from the current pressure when ||r|| falls less than 2x over 3
 * iterations, turns non-finite or VFX_OPT_LEARNED_MAXIT runs out; the true residual judges as for
 * MGPCG.  With VFX_OPT_SYNC 0 the coarse correction's flexible CG is sequenced on the device like
 * MGPCG (an iteration budget, the network's kernels gated on the stop state, the fallback as the
 * gated restart run after the true residual; vfx_step replays as a graph and batched frames
 * accept it); the full-resolution network's is host-polled.  The context does not own the
 * network, and a network serves one context at a time. */
int vfx_ctx_set_net(vfx_ctx* ctx, vfx_net* net);
/* ||r_k||_2 of every flexible-CG iteration of the last host-polled projection (convergence
 * history; empty after a device-sequenced one): up to n values into out (host), the count into
 * *count */
int vfx_fcg_history(vfx_ctx* ctx, double* out, int n, int* count);
/* the preconditioner z = M(r): rms(r) * net(r / rms(r), fluid), 0 in solids */
int vfx_net_apply(vfx_net* net, vfx_ctx* ctx, const float* solid, const float* r, float* z, void* stream);

/* ---------------------------------------------------------------- FP64 run of the same solver
 * The same entry points for FP64 state: fields, multigrid levels and every constant in double
 * (the parameter buffer stays the float layout above).  This is the "FP64 run of the same
 * solver" the FP32 path is measured against; it has no counterpart in the reference, which is
 * FP32 only.  Each function behaves as its FP32 namesake above; contexts of the two precisions
 * are distinct types. */
/* ---- batched frames (FP32): many independent simulations of any sizes as one frame ----
 * No reference counterpart: the reference steps one Solver3D per call (solver.py:146-155) and
 * graph-captures one solver's step (solver.py:388-459).  Each frame enqueues every simulation's
 * step (vfx_step's stages) on one of `streams` HIP streams forked from and joined back into the
 * caller's stream, directly or as one captured hipGraph replayed while the pointers, budgets and
 * launch shapes stay the same.  Projections are MGPCG (or, for a context with a coarse-correction
 * network, the learned flexible CG with its MGPCG fallback) sequenced on the device as in vfx_step
 * (VFX_OPT_SYNC 0) with a per-simulation iteration budget (the device stops at the budget or at
 * convergence; with restarts on, one gated restart run; no host round trip inside a frame).  A
 * budget starts at min(P_MAXIT, 32); vfx_batch_wait reads the frame's state and doubles the budget
 * of a simulation that stopped at it unconverged (vfx_batch_truncated counts those).  Without
 * vfx_batch_wait the budgets stay as they are.  The contexts stay usable on their own between
 * frames; a context may belong to one batch at a time. */
typedef struct vfx_batch vfx_batch;
#define VFX_TUNE_MAX_SIZES 8
#define VFX_TUNE_REPORT 46   /* doubles of the autotuner report (layout at vfx_batch_tune) */
int vfx_batch_create(vfx_ctx* const* ctxs, int n, int streams, vfx_batch** out);
int vfx_batch_destroy(vfx_batch* batch);
/* PCG iteration budget of simulation `sim` (-1: all); default min(P_MAXIT, 32) at the first frame */
int vfx_batch_set_budget(vfx_batch* batch, int sim, int budget);
int vfx_batch_get_budget(vfx_batch* batch, int* budget);   /* n ints (0: not set yet) */
/* Simulation-frames, among the frames vfx_batch_wait has read, whose projection stopped at its
 * budget unconverged (does not wait for the device). */
int vfx_batch_truncated(vfx_batch* batch, long long* count);
/* One frame: simulation i steps data[i] (6,N_i,N_i,N_i), solid[i], parameters prm_host[i * vfx_p_size()
 * ...]; graph 1 replays (capturing when needed) the frame's hipGraph, 0 enqueues it directly.
 * Returns once enqueued. */
int vfx_batch_step(vfx_batch* batch, float* const* data, const float* const* solid, const float* prm_host,
                   int graph, void* stream);
/* Waits for the last frame.  latency_ms[i]: frame start to the end of simulation i's step (GPU
 * wall clock); iters[i]: PCG iterations; converged[i]: the projection met its tolerance within
 * the budget (the true residual judges when it was computed); rel_res[i]: ||r|| / ||b|| at the
 * stop (true residual when computed, else the recurrence); frame_ms: frame start to the last
 * simulation's end.  Any output may be NULL. */
int vfx_batch_wait(vfx_batch* batch, float* latency_ms, int* iters, int* converged, double* rel_res,
                   float* frame_ms);
/* Autotuner: per grid size, VFX_OPT_CELLS_PER_THREAD in {1,2,4} x VFX_OPT_COARSE_M in {4,8,16} by the
 * median of `reps` timed steps of the size's first simulation; then the stream count in {1,2,4,8}
 * by the median frame time of `reps` graph frames.  Every simulation is restored to its state
 * before the call (data and pressure).  report (VFX_TUNE_REPORT doubles, may be NULL): [0] number
 * of sizes; per size k at 1 + 5k: N, cells per thread, coarse size, best ms, ms of the setting
 * before the call; at 41: chosen streams, then frame ms at 1, 2, 4, 8 streams. */
int vfx_batch_tune(vfx_batch* batch, float* const* data, const float* const* solid, const float* prm_host, int reps,
                   void* stream, double* report, int report_len);

typedef struct vfx_ctx_f64 vfx_ctx_f64;
int vfx_p_size_f64(void);
int vfx_last_error_f64(char* buf, int n);
int vfx_ctx_create_f64(int res, vfx_ctx_f64** out);
int vfx_ctx_destroy_f64(vfx_ctx_f64* ctx);
int vfx_step_f64(vfx_ctx_f64* ctx, double* data, const double* solid, const float* prm_host, void* stream);
int vfx_step_stage_f64(vfx_ctx_f64* ctx, int stage, double* data, const double* solid, const float* prm_host,
                       void* stream);
int vfx_project_f64(vfx_ctx_f64* ctx, double* data, const double* solid, const float* prm_host, void* stream);
int vfx_advect_f64(vfx_ctx_f64* ctx, const double* field, int C, const double* vel, double* out, float dt,
                   void* stream);
int vfx_div_abs_f64(vfx_ctx_f64* ctx, const double* data, const double* solid, double* out, void* stream);
int vfx_pressure_iters_f64(vfx_ctx_f64* ctx, int* out);
int vfx_pressure_stats_f64(vfx_ctx_f64* ctx, double* out, int n);
int vfx_set_option_f64(vfx_ctx_f64* ctx, int key, double value);
int vfx_get_phi_f64(vfx_ctx_f64* ctx, double* dst, void* stream);
int vfx_set_phi_f64(vfx_ctx_f64* ctx, const double* src, void* stream);
int vfx_get_advected_f64(vfx_ctx_f64* ctx, double* dst, double* mass_ratio, void* stream);
int vfx_kmul_f64(vfx_ctx_f64* ctx, const double* solid, const double* p, double* out, void* stream);
int vfx_precond_f64(vfx_ctx_f64* ctx, const double* solid, const double* r, double* z, void* stream);
int vfx_get_ops_f64(vfx_ctx_f64* ctx, const double* solid, int level, double* diag, double* off, int* nlev, int* m,
                    void* stream);
int vfx_time_kernel_f64(vfx_ctx_f64* ctx, int which, int reps, void* stream, float* ms_avg);
int vfx_hires_attach_f64(vfx_ctx_f64* ctx, int S, double* hi, const double* hi_solid);
int vfx_hires_init_f64(vfx_ctx_f64* ctx, const double* data, float ambient, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VFXSIM3D_H */
