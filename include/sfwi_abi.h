/*
 * sfwi_abi.h -- C ABI of the MI355X-native SFWI per-shot elastic modeling and
 * gradient library (libsfwi_amd.so, HIP for gfx950).
 *
 * It replaces the device-side C++ interface that the reference's host driver
 * (sfwi-stage{1,2}/main.cpp) calls: class FD2DGPU_ELASTIC and the per-shot
 * free functions declared in sfwi-stage1/FD2DGPU.cuh and
 * sfwi-stage2/FD2DGPU.cuh.  Every entry point below names the reference
 * interface it stands in for.
 *
 * Conventions
 *   - Plain C types only; all host buffers are caller-owned.
 *   - Models and gradients use the reference's .bin layout: allnx*allnz
 *     floats, x-major, z fastest (index ix*allnz + iz).
 *   - Records are trace-major [ntr][nt] floats (index tr*nt + it).
 *   - Return value 0 = success, negative = error; sfwi_last_error() gives the
 *     message.  A context is used by one host thread at a time.
 *   - One GPU per context (the current HIP device at creation).
 */
#ifndef SFWI_ABI_H
#define SFWI_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

#define SFWI_OK 0
#define SFWI_EINVAL (-1)   /* bad argument                                   */
#define SFWI_EHIP (-2)     /* HIP runtime error                              */
#define SFWI_ENOMEM (-3)   /* device allocation failed or over HBM budget    */
#define SFWI_ESTATE (-4)   /* call order (model/wavelet not set, ...)        */
#define SFWI_ENUMERIC (-5) /* a shot's misfit is not finite (NaN or Inf in    */
                           /* its observed traces)                           */
#define SFWI_ECANCELED (-6) /* the request was cancelled before it started   */
#define SFWI_ETIMEOUT (-7) /* sfwi_server_wait_for: not done in time         */

/* Numerical modes. */
#define SFWI_MODE_FAST 0       /* FP32, unsplit interior, FMA stencils          */
#define SFWI_MODE_EXACT 1      /* reference operation order everywhere (split  */
                               /* PML fields on every cell, FP64 promotions)   */

/* Misfit of the gradient and line-search solves.  L2 is the reference's
 * (sfwi-stage1/FD2DGPU.cu:4583-4590); the OT choices put the optimal-transport
 * misfit of sfwi_ot_misfit where the reference has its commented-out OT calls
 * (sfwi-stage1/FD2DGPU.cu:4887-4896, misfit_ot in sfwi-stage1/main.cpp:357,540). */
#define SFWI_MISFIT_L2 0
#define SFWI_MISFIT_OT_TRACE 1    /* sfwi_ot_misfit variant 0, per trace        */
#define SFWI_MISFIT_OT_POWER 2    /* variant 1, per trace, squared signal        */
#define SFWI_MISFIT_OT_GATHER 3   /* variant 2, whole gather of a component      */

/* Forward-wavefield storage used by the reverse-time reconstruction: the
 * 8-cell boundary strips of vx and vz of every step (the reference's
 * record_edge_wavefield, sfwi-stage1/FD2DGPU.cu:3842-3869).  The 16-bit forms
 * keep one FP32 max-abs scale per step and field and store v / scale rounded
 * to nearest even; they halve the strip memory (the largest per-shot buffer). */
#define SFWI_STORE_STRIPS_F32 0   /* FP32                                        */
#define SFWI_STORE_STRIPS_F16 1   /* FP16 with per-step max-abs scale            */
#define SFWI_STORE_STRIPS_BF16 2  /* BF16 with per-step max-abs scale            */
/* The same codes select the storage of the velocity snapshots of
 * SFWI_RECON_SNAPSHOTS (one scale per snapshot and field). */

/* How a gradient solve obtains the forward wavefield for its imaging.
 * STRIPS is the reference's scheme (IMAGE_SINGLESHOT, sfwi-stage1/
 * FD2DGPU.cu:4321-4695): boundary strips saved in the forward loop, the
 * forward field rebuilt by reverse-time stepping (source not subtracted),
 * overwrite injection of the residual, imaging after both adjoint half steps.
 * SNAPSHOTS has no reference counterpart: the forward velocity after the
 * velocity update of every snap_every-th step is kept on the interior plus a
 * 4-cell halo (FP32, or FP16 / BF16 with a max-abs scale per snapshot and
 * field), and the reverse pass runs the adjoint only: the adjoint source
 * (2 r for the L2 misfit, the OT adjoint source as it is) divided by rho is
 * added at the receivers, then the adjoint velocity half step, on snapshot
 * steps the zero-lag imaging from the kept velocity, then the adjoint stress
 * half step, direction +1 (adjoint_dir is not used).  This is the consistent
 * discrete gradient of the misfit with the imaging sampled every snap_every
 * steps (weight snap_every, times dt); with snap_every = 1 and FP32 storage
 * it is the exact gradient of the discrete scheme.  Stage 2: the gradient of
 * the Born misfit with respect to the perturbation.  FAST mode only, no
 * density image, no cpu_mem. */
#define SFWI_RECON_STRIPS 0
#define SFWI_RECON_SNAPSHOTS 1

typedef struct sfwi_ctx sfwi_ctx;

/* Creation parameters: the FD2DGPU_ELASTIC constructor arguments
 * (sfwi-stage1/FD2DGPU.cuh:152-153: sou, dx, dz, dt, nxpml, nzpml, allnx,
 * allnz, scale, pml, nt, nop, ntr) plus the constants the reference hard-codes. */
typedef struct sfwi_params {
    float dx, dz, dt;      /* grid spacing (m) and time step (s)                   */
    int nx, nz;            /* shot window; nxpml = nx + 2 pml, nzpml = nz + 2 pml  */
    int pml;               /* PML width (cells)                                     */
    int nt;                /* time samples                                          */
    int nop;               /* stencil half width, must be 4                         */
    int ntr;               /* maximum traces per shot                               */
    int allnx, allnz;      /* full model size                                       */
    float rr;              /* PML reflection coefficient (reference: 0.001)         */
    float rho_const;       /* > 0: padded density forced to this value (the stage-2
                              medium kernel forces 2000); 0: use the model density */
    int adjoint_dir;       /* adjoint propagation direction: +1 stage 1, -1 stage 2 */
    int max_batch;         /* shots resident on the device at once (>= 1)           */
    int mode;              /* SFWI_MODE_*                                           */
    int storage;           /* SFWI_STORE_*                                          */
    int use_graph;         /* 1: capture each solve (all time steps) of a batch size */
                           /* once into a hipGraph and replay it (L2 misfit; solves */
                           /* with sfwi_set_timing on run directly)                 */
    int march_cols;        /* FAST kernels: columns per wave segment of the column-  */
                           /* march kernels; < 0 selects per-cell kernels; 0 picks */
                           /* per-cell kernels for batches of fewer than 7168 wave */
                           /* tiles (too few to fill the GPU) and 8 columns above  */
    int march_depth;       /* columns of operands prefetched ahead (1, 2, 4; 0: 4, and */
                           /* 2 for the two Born forward kernels, which hold twice */
                           /* the fields; the fused steps' workgroup marches: 1)   */
    int misfit;            /* SFWI_MISFIT_*                                          */
    int march_order;       /* column-march block order: 0 default (= 3); 1 one shot  */
                           /* after another, split over the XCDs; 2 the shots of a  */
                           /* tile back to back; 3 whole shots per XCD              */
    int grad_rho;          /* 1: gradient solves also image the density (sfwi_flush_rho) */
    int cpu_mem;           /* 1: the boundary strips of a gradient solve live in pinned */
                           /* host memory (the reference's cpu_mem, sfwi-stage1/      */
                           /* FD2DGPU.cu:4393-4421, 4545-4560, 4602-4620): the device */
                           /* keeps two chunks of strip_chunk steps per shot, and     */
                           /* completed chunks move over PCIe on a second stream      */
                           /* while the kernels run; results equal those of HBM      */
                           /* strips bit for bit.  Solves with it run without graphs. */
    int strip_chunk;       /* steps per host chunk (the reference's ntt = iointerval; */
                           /* 0 = 100); nt <= 2 strip_chunk keeps every step in HBM   */
    int rbc;               /* 1: random boundary (the reference's rbc, sfwi-stage1/   */
                           /* FD2DGPU.cu:3245-3305, 4330-4334): random Vp / Vs in the */
                           /* PML bands of every window, no damping, no strips; the   */
                           /* reverse pass reconstructs by time reversal alone        */
                           /* (cpu_mem then has nothing to hold)                      */
    unsigned int rbc_seed; /* seed of the random boundary's generator (the reference  */
                           /* seeds cuRAND from time(0); here a run repeats)          */
    int recon;             /* SFWI_RECON_*                                          */
    int snap_every;        /* SFWI_RECON_SNAPSHOTS: steps between snapshots (0 = 4) */
    int fuse_step;         /* FAST column-march forward solves of the background    */
                           /* field: > 0 runs each time step as one fused launch    */
                           /* (velocity and stress together, state double-buffered, */
                           /* identical results): 1 wave-level march (48 rows per  */
                           /* wave), 2 workgroup march (240 rows per workgroup,    */
                           /* z halo shared through LDS), 3 the same, and a        */
                           /* snapshot gradient's adjoint steps fused the same     */
                           /* way; < 0 as two half-step launches; 0 (default,      */
                           /* which sfwi_autotune may change): 3 where the         */
                           /* 240-row tiles take no more lanes than 56-row waves,  */
                           /* else half steps (every choice gives the same bits);  */
                           /* 2 and 3 need a shot's state block (43-52 model-size  */
                           /* fields) below 4 GiB, else SFWI_EINVAL at create      */
    int step_cols;         /* columns per segment of both fused steps (0 = forward
                              16, adjoint 32)                                        */
} sfwi_params;

/* One shot: the fields the reference driver writes into FD2DGPU_ELASTIC
 * before each IMAGE_SINGLESHOT call (sfwi-stage1/main.cpp:596-625). */
typedef struct sfwi_shot {
    int left;              /* first model column of the window                      */
                           /*   (delta_left + record_left_in_v)                     */
    int isx, isz;          /* source column / row on the padded grid                */
    int igz;               /* receiver row on the padded grid                       */
    int ngx_left;          /* coordinate of the window's left edge                  */
    int ntr;               /* traces of this shot (<= sfwi_params.ntr)              */
    const int *gc;         /* receiver coordinates [ntr]                             */
    const float *obs_z;    /* observed vz traces [ntr][nt] (gradient calls)         */
    const float *obs_x;    /* observed vx traces [ntr][nt]                          */
                           /* Both NULL: reuse the observed traces left in HBM by   */
                           /* the previous call for the same batch position (shot   */
                           /* i of a batch of <= max_batch shots sits in slot i);   */
                           /* the trace count must match.                          */
    int obs_id;            /* > 0: the observed traces stored on the device under   */
                           /* this id by sfwi_obs_put (obs_z, obs_x are ignored and */
                           /* nothing is copied); 0: obs_z / obs_x as above         */
} sfwi_shot;

/* Observed traces resident on the device (no reference counterpart: the
 * reference uploads each shot's records before every IMAGE_SINGLESHOT call,
 * record_copytoGPU, sfwi-stage1/FD2DGPU.cu:1327-1333).  sfwi_obs_put stores a
 * shot's traces ([ntr][nt] per component) under obs_id > 0 (replacing what the
 * id held); a shot with that obs_id then reads them in place, so an inversion
 * or a benchmark keeps a whole survey's records in HBM and a solve copies
 * none.  sfwi_obs_clear frees every stored set. */
int sfwi_obs_put(sfwi_ctx *ctx, int obs_id, int ntr, const float *obs_z, const float *obs_x);
int sfwi_obs_clear(sfwi_ctx *ctx);

/* Replaces the FD2DGPU_ELASTIC constructor (sfwi-stage1/FD2DGPU.cu:1366-1655)
 * minus the source wavelet, which sfwi_set_wavelet uploads. */
int sfwi_ctx_create(const sfwi_params *p, sfwi_ctx **out);
/* Replaces ~FD2DGPU_ELASTIC. */
void sfwi_ctx_destroy(sfwi_ctx *ctx);
/* Text of the last error on this context; with ctx NULL, of the calling
 * thread's last failed call without a context (create, server calls). */
const char *sfwi_last_error(const sfwi_ctx *ctx);

/* Source wavelet (nt floats): the `sou` constructor argument, uploaded in
 * sfwi-stage1/FD2DGPU.cu:46-47. */
int sfwi_set_wavelet(sfwi_ctx *ctx, const float *wavelet);
/* Full models (allnx*allnz each): FD2DGPU_ELASTIC::GPUbufferVPVS
 * (sfwi-stage1/FD2DGPU.cu:1322-1327).  rho may be NULL when rho_const > 0.
 * The model is checked first (sfwi_check_model); a bad one is refused with
 * SFWI_EINVAL and the context keeps its previous model. */
int sfwi_set_model(sfwi_ctx *ctx, const float *vp, const float *vs, const float *rho);
/* Host check of a model (no reference counterpart: the reference reads its
 * model files unchecked, sfwi-stage1/main.cpp:1470-1472): every vp finite
 * and > 0, every vs finite and >= 0, every rho finite and > 0 (not checked
 * when rho_const > 0, which replaces it).  SFWI_EINVAL names the first bad
 * cell (ix, iz) and value in sfwi_last_error(NULL).  Needs no GPU. */
int sfwi_check_model(const sfwi_params *p, const float *vp, const float *vs, const float *rho);
/* Model perturbation for Born modeling: FD2DGPU_ELASTIC::GPUbufferM
 * (sfwi-stage1/FD2DGPU.cu:1330-1334) and its window (buffer_image_extrapolation,
 * :1336-1346).  The two stages' Born sources act on different cells
 * (add_born_source_components): stage 2 on every cell at least nop from the
 * grid edge (sfwi-stage2/born.cu:749-755), stage 1 on the interior only, at
 * least pml from the edge (sfwi-stage1/born.cu:830-836).  sfwi_set_perturbation
 * selects the stage-2 cells, sfwi_set_perturbation_region either. */
#define SFWI_BORN_S2 0
#define SFWI_BORN_S1 1
int sfwi_set_perturbation(sfwi_ctx *ctx, const float *dvp, const float *dvs);
int sfwi_set_perturbation_region(sfwi_ctx *ctx, const float *dvp, const float *dvs, int region);

/* Stage-1 per-shot gradient: IMAGE_SINGLESHOT (sfwi-stage1/FD2DGPU.cu:4321-4695,
 * declared sfwi-stage1/FD2DGPU.cuh:466-467).  Adds the L2 misfit to *misfit and
 * accumulates the Vp/Vs gradients and the illumination on the device until
 * sfwi_flush. */
int sfwi_shot_gradient(sfwi_ctx *ctx, const sfwi_shot *shot, double *misfit);
/* Stage-2 per-shot gradient of the Born misfit: IMAGE_SINGLESHOT of stage 2
 * (sfwi-stage2/FD2DGPU.cu:3978-4570, declared sfwi-stage2/FD2DGPU.cuh:455-456),
 * residual = born - obs + background, perturbation from sfwi_set_perturbation. */
int sfwi_shot_born_gradient(sfwi_ctx *ctx, const sfwi_shot *shot, double *misfit);
/* Many shots at once (the reference farms them over MPI ranks,
 * sfwi-stage1/main.cpp:417-634); misfits[i] receives shot i's misfit (+=).
 * The sfwi_batch_* calls solve the shots in batches of up to max_batch; an
 * empty list (nshots = 0, shots may be NULL) does nothing and returns 0.
 * The gradient solves (per-shot and batched) return SFWI_ENUMERIC, naming the
 * first shot, when a misfit is not finite (NaN or Inf in the observed traces):
 * every shot is still solved and its misfit added, so the accumulated images
 * then hold non-finite values until the next sfwi_flush. */
int sfwi_batch_gradient(sfwi_ctx *ctx, int nshots, const sfwi_shot *shots, double *misfits);
int sfwi_batch_born_gradient(sfwi_ctx *ctx, int nshots, const sfwi_shot *shots, double *misfits);

/* Forward solve (born_modeling mode 0 + record_background_wave, the inner loop
 * of CALCULATE_ALPHA_parabola_mb, sfwi-stage1/FD2DGPU.cu:4864-4870):
 * synthetic vz / vx traces [ntr][nt]. */
int sfwi_shot_forward(sfwi_ctx *ctx, const sfwi_shot *shot, float *syn_z, float *syn_x);
/* Forward solves of many shots at once; traces of shot i at
 * syn_z + i*ntr_max*nt (and syn_x likewise), ntr_max = sfwi_params.ntr. */
int sfwi_batch_forward(sfwi_ctx *ctx, int nshots, const sfwi_shot *shots, float *syn_z, float *syn_x);
/* Born forward solve (born_modeling mode 1 + record_scatter, the inner loop of
 * CALCULATE_ALPHA, sfwi-stage2/FD2DGPU.cu:4607-4620): scattered traces b_*
 * and background traces bg_* (either pair may be NULL). */
int sfwi_shot_born_forward(sfwi_ctx *ctx, const sfwi_shot *shot, float *b_z, float *b_x,
                           float *bg_z, float *bg_x);
/* Born forward solves of many shots at once (the same kernels over batches of
 * max_batch shots): scattered traces of shot i at b_z + i*ntr_max*nt. */
int sfwi_batch_born_forward(sfwi_ctx *ctx, int nshots, const sfwi_shot *shots, float *b_z, float *b_x);

/* Born forward solve with double scattering (born_modeling mode 1 then mode
 * 5 each step, sfwi-stage1/FD2DGPU.cu:1855-1860, 1888-1903; the reference
 * keeps the mode-5 calls commented out, e.g. :4742, :5307): the second-order
 * scattered field is driven by the Born source of the updated first-order
 * scattered velocity.  b_*: first-order scattered traces (record_scatter),
 * b2_*: second-order ones (record_scatter with double_scatter, :3610-3617);
 * the reference's Lm + Lm2 (scatter_add, :3236-3241) is b + b2.  Traces of
 * shot i at b_z + i*ntr_max*nt (ntr_max = sfwi_params.ntr); any pointer may
 * be NULL.  Uses the current perturbation and its region; per-cell kernels. */
int sfwi_batch_born2_forward(sfwi_ctx *ctx, int nshots, const sfwi_shot *shots, float *b_z, float *b_x,
                             float *b2_z, float *b2_x);

/* Gauss-Newton Hessian-vector product, ELSRTM_CALCULATE_Hessian_vector
 * (sfwi-stage1/FD2DGPU.cu:5236-5359, declared sfwi-stage1/FD2DGPU.cuh:477-478),
 * called per shot by the inner CG loop of sfwi-stage1/main.cpp:846-935.  The
 * vector p (p_vp, p_vs: allnx*allnz) becomes the perturbation with the stage-1
 * region (as the reference copies it into image_pp_m / image_ps_m); NULL keeps
 * the current perturbation.  Per shot: stage-1 Born forward of p with the
 * background's reconstruction state saved (strips, or snapshots with
 * SFWI_RECON_SNAPSHOTS), the scattered traces injected as the adjoint source,
 * reverse reconstruction and adjoint, Vp / Vs imaging without illumination
 * (hessian_vector_product, :3469-3473).  H p is accumulated into the Vp / Vs
 * gradient accumulators until sfwi_flush (illumination and density untouched).
 * With SFWI_RECON_SNAPSHOTS, snap_every = 1 and FP32 storage it is
 * 2 J^T J p, J the Jacobian of the traces with respect to (vp, vs), up to
 * rounding: symmetric. */
int sfwi_shot_hessian_vector(sfwi_ctx *ctx, const sfwi_shot *shot, const float *p_vp, const float *p_vs);
int sfwi_batch_hessian_vector(sfwi_ctx *ctx, int nshots, const sfwi_shot *shots, const float *p_vp,
                              const float *p_vs);

/* Stage-1 line search, CALCULATE_ALPHA_parabola_mb (sfwi-stage1/FD2DGPU.cu:4819-5028,
 * declared sfwi-stage1/FD2DGPU.cuh:471-472): adds to m[0..2] the misfits of
 * forward solves at vel, vel - step1*dir and vel - step2*dir. */
int sfwi_linesearch_parabola(sfwi_ctx *ctx, const sfwi_shot *shot, const float *dir_vp,
                             const float *dir_vs, float step1, float step2, double m[3]);
/* The same line search over a list of shots (the workers' per-shot
 * CALCULATE_ALPHA_parabola_mb calls, sfwi-stage1/main.cpp:1180, over the
 * iteration's shots): each test model is built once and its windows once per window
 * position, the forward solves run in batches of max_batch, and the per-shot
 * misfits are added to m[0..2] in shot order, so the result equals nshots
 * calls of sfwi_linesearch_parabola bit for bit. */
int sfwi_batch_linesearch_parabola(sfwi_ctx *ctx, int nshots, const sfwi_shot *shots, const float *dir_vp,
                                   const float *dir_vs, float step1, float step2, double m[3]);
/* Stage-2 exact step, CALCULATE_ALPHA (sfwi-stage2/FD2DGPU.cu:4576-4722,
 * declared sfwi-stage2/FD2DGPU.cuh:457-458): fenzi += sum L(cg).r,
 * fenmu += sum L(cg)^2 with r from the current perturbation. */
int sfwi_linesearch_linear(sfwi_ctx *ctx, const sfwi_shot *shot, const float *cg_vp,
                           const float *cg_vs, double *fenzi, double *fenmu);
/* The stage-2 driver's loop of CALCULATE_ALPHA calls over its shots
 * (sfwi-stage2/main.cpp:766-859) as batched solves: both Born solves run over
 * batches of up to max_batch shots, one window of the direction per window
 * position.  Each shot's per-trace FP64 products are added to fenzi / fenmu in
 * shot order, traces in trace order, so the result equals nshots calls of
 * sfwi_linesearch_linear bit for bit.  The observed traces of shot i of a
 * batch stay resident in its slot, as after a gradient call. */
int sfwi_batch_linesearch_linear(sfwi_ctx *ctx, int nshots, const sfwi_shot *shots, const float *cg_vp,
                                 const float *cg_vs, double *fenzi, double *fenmu);
/* No reference counterpart: the stage-2 step's first Born solve recomputes the
 * residual born(dm) - obs + background that the Born gradient call at the
 * same model and images has just formed.  sfwi_batch_born_gradient_res is
 * sfwi_batch_born_gradient (L2 misfit) that also returns each shot's residual
 * (shot i at res_z + i*ntr_max*nt, res_x likewise, ntr_max = sfwi_params.ntr);
 * sfwi_batch_linesearch_linear_res takes those residuals in place of that
 * solve (observed traces not needed) and returns the same sums as
 * sfwi_batch_linesearch_linear bit for bit, one Born solve per shot fewer.
 * The caller keeps model, images and shots unchanged in between. */
int sfwi_batch_born_gradient_res(sfwi_ctx *ctx, int nshots, const sfwi_shot *shots, double *misfits,
                                 float *res_z, float *res_x);
int sfwi_batch_linesearch_linear_res(sfwi_ctx *ctx, int nshots, const sfwi_shot *shots, const float *cg_vp,
                                     const float *cg_vs, const float *res_z, const float *res_x,
                                     double *fenzi, double *fenmu);

/* Gradient flush, FD2DGPU_ELASTIC::imagebuffer_resettozero
 * (sfwi-stage1/FD2DGPU.cu:3646-3673): copies the accumulated Vp and Vs
 * gradients and the illumination (allnx*allnz each) to the host and zeroes
 * them on the device.  Any pointer may be NULL. */
int sfwi_flush(sfwi_ctx *ctx, float *grad_vp, float *grad_vs, float *illum);

/* Density gradient at fixed Vp, Vs (no reference counterpart: the reference
 * images Vp and Vs only, sfwi-stage1/FD2DGPU.cu:3350-3382).  With
 * sfwi_params.grad_rho = 1 every gradient solve also accumulates, per reverse
 * step and interior cell, -(vx_adj dvx + vz_adj dvz)/dt + (vp^2 - 2 vs^2) lg
 * + vs^2 mg, where dv = v(it+1) - v(it) is the reconstruction's velocity step
 * and lg, mg are the lambda and mu correlations of the Vp / Vs imaging, in
 * the same units.  Copies the accumulated image (allnx*allnz) and zeroes it. */
int sfwi_flush_rho(sfwi_ctx *ctx, float *grad_rho);

/* Optimal-transport trace misfit and adjoint source.  Declared but never
 * defined in the reference (OT, OT_power, OT_alltrace,
 * sfwi-stage1/FD2DGPU.cuh:21-25); arguments follow those declarations:
 * observed and synthetic traces [ntr][nt] per component, adjoint source
 * dW/dsyn written to res_*, W2^2 misfit (time units, summed over traces and
 * both components) added to *misfit.
 * variant 0: per trace, signal shifted to positivity by
 *            c = 1.1 max(max|syn|, max|obs|) (c held fixed in the adjoint);
 *         1: per trace, squared signal ("power");
 *         2: all traces of a component as one distribution (flattened
 *            trace after trace, shift as in variant 0 over the gather).
 * Each signal is a piecewise-constant density on unit sample cells; the
 * cumulative sums and quantile inversion run in FP64.  A trace with zero mass
 * on either side contributes 0 and gets a zero adjoint source. */
int sfwi_ot_misfit(sfwi_ctx *ctx, int variant, int ntr, int nt, float dt, const float *obs_x,
                   const float *obs_z, const float *syn_x, const float *syn_z, float *res_x,
                   float *res_z, double *misfit);

/* sfwi_ot_misfit without any device allocation (no reference counterpart:
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results, bit for bit; the workspace (staging, job table,
 * FP64 merge-walk workspace) is one static device array per GPU, about
 * 21 MB, held until the process ends.  Limits: ntr * nt <= 262144 samples per
 * component and ntr <= 4096, else SFWI_EINVAL.  Calls that overlap on one GPU
 * (from any context) share that array and race: serialize them. */
int sfwi_ot_misfit_static(sfwi_ctx *ctx, int variant, int ntr, int nt, float dt, const float *obs_x,
                          const float *obs_z, const float *syn_x, const float *syn_z, float *res_x,
                          float *res_z, double *misfit);

/* ---------------------------------------------------------------------------
 * Serving loop (no reference counterpart: the reference runs one shot per MPI
 * rank at a time).  Requests of any grid, record length and kind enter a FIFO
 * queue; a pool of worker threads, each with its own context and stream,
 * takes the first queued request whose device memory fits the free part of
 * the HBM budget, so work that arrives while other solves run is admitted as
 * workers and memory free up and overlaps with them on the GPU.  Contexts are
 * cached by (params, kind) with their wavefield slots; idle ones are evicted
 * least recently used first when memory is needed.  Each request's result is
 * the one the same solve gives alone.  All buffers of a request are
 * caller-owned and must stay valid until sfwi_server_wait returns for it.
 * Errors: negative codes; the message is sfwi_last_error(NULL) on the calling
 * thread (each host thread keeps its own; a failed request's message is
 * returned to the thread that waits for it).
 * ------------------------------------------------------------------------- */
#define SFWI_REQ_FORWARD 0          /* sfwi_batch_forward                       */
#define SFWI_REQ_BORN_FORWARD 1     /* sfwi_shot_born_forward per shot          */
#define SFWI_REQ_GRADIENT 2         /* sfwi_batch_gradient + sfwi_flush         */
#define SFWI_REQ_BORN_GRADIENT 3    /* sfwi_batch_born_gradient + sfwi_flush    */

typedef struct sfwi_request {
    int kind;                       /* SFWI_REQ_*                                   */
    sfwi_params params;             /* solve shape; max_batch >= nshots             */
    const float *wavelet;           /* nt                                           */
    const float *vp, *vs, *rho;     /* allnx*allnz (rho may be NULL with rho_const) */
    const float *dvp, *dvs;         /* Born kinds                                   */
    int nshots;
    const sfwi_shot *shots;
    float *syn_z, *syn_x;           /* forward kinds: traces of shot i at i*ntr*nt  */
                                    /* (Born forward: scattered traces)            */
    double *misfits;                /* gradient kinds: [nshots] (may be NULL)       */
    float *grad_vp, *grad_vs, *illum; /* gradient kinds: allnx*allnz, summed over   */
                                    /* the request's shots (any may be NULL)       */
    int autotune;                   /* 1: run sfwi_autotune (100 steps) on a newly  */
                                    /* created context before its first solve      */
} sfwi_request;

typedef struct sfwi_server sfwi_server;
/* Device memory one context of this shape needs for `nshots` shots of `kind`. */
size_t sfwi_solve_bytes(const sfwi_params *p, int kind, int nshots);
/* hbm_budget 0: 15/16 of the free HBM; workers in [1, 16]. */
int sfwi_server_create(size_t hbm_budget, int workers, sfwi_server **out);
int sfwi_server_submit(sfwi_server *s, const sfwi_request *req, long long *ticket);
/* Cross-request batching (no reference counterpart), on by default: requests
 * of the same shape and kind (every params field, the kind and the autotune
 * flag equal) queued when a worker takes one join its batch, up to
 * params.max_batch shots in all; each shot keeps its own request's model,
 * perturbation and wavelet, and each request's images are merged over its
 * own shots, so every result equals the request solved alone.  on = 0 runs
 * one request per batch. */
int sfwi_server_batching(sfwi_server *s, int on);
/* Blocks until the request is done; queue_ms = submit -> start, run_ms = start -> end.
 * Returns the request's status: 0, or its error (SFWI_ENUMERIC for a
 * gradient request whose misfit is not finite: its outputs then hold the
 * non-finite values, and the other requests of its batch are unaffected;
 * SFWI_ECANCELED after sfwi_server_cancel), and forgets the ticket. */
int sfwi_server_wait(sfwi_server *s, long long ticket, double *queue_ms, double *run_ms);
/* sfwi_server_wait with a time limit: SFWI_ETIMEOUT if the request is not done
 * within timeout_ms; the ticket then stays valid (wait again, or cancel). */
int sfwi_server_wait_for(sfwi_server *s, long long ticket, double timeout_ms, double *queue_ms, double *run_ms);
/* Cancels a request that is still queued: it never runs, and waiting for it
 * returns SFWI_ECANCELED.  A request that has started (or ended) cannot be
 * cancelled: SFWI_ESTATE. */
int sfwi_server_cancel(sfwi_server *s, long long ticket);
/* Call after every submitted request has been waited for. */
void sfwi_server_destroy(sfwi_server *s);

/* Reference host-driver helper: Ricker wavelet (sfwi-stage1/main.cpp:251-260). */
void sfwi_ricker(float *w, int nt, float dt, float f0);

/* ---- Survey input of the reference drivers (CPU code, no device use) -------- */

#define SFWI_PATH_MAX 1024

/* The 10-line positional parameter file of both drivers (sfwi-stage1/
 * main.cpp:1346-1355; the shipped para_stage1 and para_stage2 both follow it,
 * while sfwi-stage2/main.cpp:946-954 reads one path line fewer). */
typedef struct sfwi_para {
    float dx, dz;
    int pml, allnx, allnz, minshot, maxshot;
    char vp[SFWI_PATH_MAX], vs[SFWI_PATH_MAX], rho[SFWI_PATH_MAX];
    char rec_z[SFWI_PATH_MAX], rec_x[SFWI_PATH_MAX], outdir[SFWI_PATH_MAX];
    float f0;
    int maxiter, iter_round1;
    int sy, gy;                        /* source / receiver depth below the PML, cells */
    int light, rbc, cpu_mem, iointerval;
} sfwi_para;

/* Parse a parameter file as a stream of 22 whitespace-separated fields (the
 * drivers' fscanf); fewer fields, a non-numeric number or a path of
 * SFWI_PATH_MAX characters or more is an error. */
int sfwi_para_read(const char *path, sfwi_para *out);
/* An x-major FP32 model (allnx columns of allnz floats, sfwi-stage1/
 * main.cpp:1468-1473); a file of another size is an error. */
int sfwi_model_read(const char *path, int allnx, int allnz, float *out);
/* Shot index of an SU file, index_shot with the per-shot trace count of
 * sfwi-stage2/main.cpp:121-216 (the stage-1 copy never counts past 1 trace):
 * consecutive traces of equal sx form a shot; row k of table (9 ints) is
 * {k+1, ntr, sx, 0, min(sx,gx), 0, max(sx,gx), 0, first trace}, coordinates
 * divided by coordinate_scale.  nt and dt (s) come from the first header.
 * *nshots gets the shot count; table may be NULL with max_shots = 0 to count. */
int sfwi_su_index(const char *path, int coordinate_scale, int max_shots, int *nt, float *dt, int *nshots,
                  int *table);
/* Traces pos .. pos+ntr-1 of an SU file into data [ntr][nt] and gc = gx /
 * coordinate_scale (read_shot_gather_su, sfwi-stage1/main.cpp:204-225). */
int sfwi_su_read_gather(const char *path, long long pos, int ntr, int nt, int coordinate_scale, float *data,
                        int *gc);
/* Write (or append) one shot gather as SU traces: tracl / tracr / tracf =
 * 1..ntr, fldr = shot, sx, gx[i], offset = gx - sx, scalco = 1, ns = nt,
 * dt in microseconds (no reference counterpart: synthetic records). */
int sfwi_su_write(const char *path, int append, int shot, int ntr, int nt, float dt, int sx, const int *gx,
                  const float *data);
/* A shot's geometry from its index row, as the worker sets it
 * (sfwi-stage1/main.cpp:580-625): left = (xmin - all_left) / dx, isx = pml +
 * (sx - xmin) / dx, isz = pml + sy, igz = pml + gy, ngx_left = xmin, ntr;
 * *nx_shot (may be NULL) gets the shot's aperture (xmax - xmin) / dx + 1.
 * gc, obs_z and obs_x are left to the caller. */
int sfwi_shot_from_table(const int row[9], float dx, int pml, int sy, int gy, int all_left, sfwi_shot *out,
                         int *nx_shot);

/* ---- Host steps of the inversion loops (CPU code, no device use) -----------
 * Model-size arrays are allnx*allnz FP32, index x*allnz + z, the reference's
 * arithmetic types (FP32 arrays, FP64 sums and step lengths).  The caller
 * keeps the conjugate-gradient state between iterations: the previous
 * direction (pp_cg_old, ps_cg_old, zero before iteration 0) and beta_fenmu. */

/* Stage-1 search direction (sfwi-stage1/main.cpp:695-750, 809-839,
 * 1059-1062): g / illumination, zero above cut_layer (the reference uses 26),
 * beta = sum g (g / illum) / beta_fenmu (0 at iter 0), cg = pc + beta cg_old,
 * then dir = cg normalized by the joint max-abs of both components.  Updates
 * pp_cg_old, ps_cg_old (to cg) and *beta_fenmu; *beta (may be NULL) gets beta. */
int sfwi_s1_direction(int allnx, int allnz, int iter, int cut_layer, const float *pp_grad,
                      const float *ps_grad, const float *illum, float *pp_cg_old, float *ps_cg_old,
                      double *beta_fenmu, float *dir_pp, float *dir_ps, double *beta);
/* Stage-1 Gauss-Newton inner CG (sfwi-stage1/main.cpp:784-1063; it runs
 * when max_inner > 0, and the reference ships max_inner = 0, :408).  The
 * caller owns the state: r, p, delta, grad (pp and ps arrays, allnx*allnz
 * each).  sfwi_s1_hcg_init (:786-839): r = normalize_all(cg), p = -r,
 * grad = normalize_all(cg), delta = normalize_all(-cg), cg the CG direction
 * of the iteration (pp_cg_old / ps_cg_old after sfwi_s1_direction).  Per inner
 * iteration the caller sums H p over the shots (sfwi_batch_hessian_vector
 * with p, then sfwi_flush) and calls sfwi_s1_hcg_step (:957-1024): with
 * cut_grad, Hp rows z < cut_layer zeroed (in hp_*); alpha = sum r.r / sum p.Hp;
 * delta += alpha p; grad = -delta; r += alpha Hp; beta = sum r.r (new) /
 * sum r.r (old); p = -r + beta p.  rr_old and rr_new (the reference's
 * hessian_misfit before and after), alpha and beta may be NULL.
 * sfwi_s1_hcg_finish (:1057-1063) normalizes delta and grad; grad is then the
 * search direction of the line search and the update (:1180, 1252). */
int sfwi_s1_hcg_init(int allnx, int allnz, const float *pp_cg, const float *ps_cg, float *r_pp, float *r_ps,
                     float *p_pp, float *p_ps, float *delta_pp, float *delta_ps, float *grad_pp, float *grad_ps);
int sfwi_s1_hcg_step(int allnx, int allnz, int cut_grad, int cut_layer, float *hp_pp, float *hp_ps, float *r_pp,
                     float *r_ps, float *p_pp, float *p_ps, float *delta_pp, float *delta_ps, float *grad_pp,
                     float *grad_ps, double *rr_old, double *rr_new, double *alpha, double *beta);
int sfwi_s1_hcg_finish(int allnx, int allnz, float *delta_pp, float *delta_ps, float *grad_pp, float *grad_ps);
/* Stage-1 step length from the three line-search misfits at 0, step1, step2
 * (cal_step2, sfwi-stage1/main.cpp:65-72), clamped to [-100, 300] (1229-1238). */
double sfwi_s1_alpha(double m0, double m1, double m2, float step1, float step2);
/* Stage-1 model update (sfwi-stage1/main.cpp:1241-1257): vel -= alpha dir, then
 * every cell with z < cut_layer takes the value of the model's first cell. */
int sfwi_s1_update(int allnx, int allnz, int cut_layer, double alpha, const float *dir_pp,
                   const float *dir_ps, float *velp, float *vels);
/* Stage-2 direction (sfwi-stage2/main.cpp:645-703): rows z >= cut_layer (the
 * reference uses 26) take g / illum, rows above are zeroed in pp_grad, ps_grad
 * themselves (both are modified, as in the reference); beta = sum (g^2 / illum)
 * / beta_fenmu (0 at iter 0); pp_cg, ps_cg = pc + beta cg_old. */
int sfwi_s2_direction(int allnx, int allnz, int iter, int cut_layer, float *pp_grad, float *ps_grad,
                      const float *illum, float *pp_cg_old, float *ps_cg_old, double *beta_fenmu,
                      float *pp_cg, float *ps_cg, double *beta);
/* Stage-2 update (sfwi-stage2/main.cpp:861-873): alpha = fenzi / fenmu from
 * sfwi_linesearch_linear summed over shots, image -= alpha cg; *alpha may be NULL. */
int sfwi_s2_update(int allnx, int allnz, double fenzi, double fenmu, const float *pp_cg,
                   const float *ps_cg, float *image_pp, float *image_ps, double *alpha);

/* Measurement hooks (no reference counterpart).  With sample_every = k > 0
 * every k-th time step of a solve brackets each stencil kernel launch with
 * HIP events on the context's stream; sfwi_kernel_time reports the mean
 * device time per launch of one kernel since the last sfwi_set_timing call.
 * Kernel ids: SFWI_K_* below. */
#define SFWI_K_FWD_VEL 0      /* forward velocity half step (whole batch)      */
#define SFWI_K_FWD_STR 1      /* forward stress half step                      */
#define SFWI_K_REC_STR 2      /* reconstruction stress step + derivatives      */
#define SFWI_K_REC_VEL 3      /* reconstruction velocity step                  */
#define SFWI_K_ADJ_STR 4      /* adjoint stress-like step                      */
#define SFWI_K_ADJ_VEL 5      /* adjoint velocity step fused with imaging      */
/* The column-march kernels without the density image image elsewhere in the
 * reverse step (2, 3 and 5 then do not run): */
#define SFWI_K_REC_STR_IMG 6  /* reconstruction stress step + Vp / Vs imaging  */
#define SFWI_K_REC_VEL_ILL 7  /* reconstruction velocity step + illumination   */
#define SFWI_K_ADJ_VEL_ONLY 8 /* adjoint velocity step                         */
#define SFWI_K_SNAP_IMAGE 9   /* snapshot gradient: imaging of one snapshot    */
#define SFWI_K_SNAP_STORE 10  /* snapshot gradient: max-abs scale + packing    */
#define SFWI_K_FWD_STEP 11    /* fused forward time step (velocity + stress)   */
#define SFWI_K_ADJ_STEP 12    /* fused adjoint time step (snapshot gradient)   */
#define SFWI_K_COUNT 13
int sfwi_set_timing(sfwi_ctx *ctx, int sample_every);
int sfwi_kernel_time(sfwi_ctx *ctx, int kernel_id, double *avg_ms, int *launches);
const char *sfwi_kernel_name(int kernel_id);

/* Kernel-family autotuner of the FAST mode (no reference counterpart): for a
 * batch of nshots shots of this context's grid, times a gradient solve of
 * `steps` time steps for each candidate -- per-cell kernels; column march with
 * 8 or 16 columns per segment, prefetch depth 1, 2 or 4, shot-major,
 * shot-interleaved or whole-shots-per-XCD block order; the forward solve in
 * half-step launches or in fused whole-step launches, wave-level or
 * workgroup march (the time-blocking depth, fuse_step 1, 2 or 3) -- and keeps
 * the fastest in the context (read it back
 * with sfwi_get_params).
 * Needs the model and wavelet set.  Accumulated gradients are preserved; the
 * observed records resident in the batch slots are invalidated.  Outputs may
 * be NULL; best_ms is the winner's time for the `steps`-step solve. */
int sfwi_autotune(sfwi_ctx *ctx, int nshots, int steps, int *march_cols, int *march_depth, int *march_order,
                  double *best_ms);

/* The context's current parameters, including what sfwi_autotune chose. */
int sfwi_get_params(const sfwi_ctx *ctx, sfwi_params *out);

/* Library version string. */
const char *sfwi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SFWI_ABI_H */
