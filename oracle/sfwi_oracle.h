/*
 * CPU oracle for SFWI's per-shot elastic modeling + gradient path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the test oracle: tests/, the smoke
 * check in __graft_entry__.py and bench.py's cpu_baseline leg may load it,
 * nothing in sfwi_amd/ may.  It restates the reference algorithm on the CPU in
 * two precisions:
 *   *_f32 : the reference's FP32 operation order, including the places where a
 *           double literal promotes an expression to FP64 (built with
 *           -ffp-contract=off so no FMA is formed that the C source lacks).
 *           This is the uncontracted reading of the source text: the
 *           reference's own nvcc -O3 build contracts a*b+c into FMAs by
 *           default, so ulp-level parity with its binary is unpinned;
 *   *_f64 : every array and every operation in FP64 (accuracy oracle and timed
 *           CPU baseline).
 *
 * All model/gradient arrays are the reference's .bin layout: x-major,
 * z fastest, index ix*allnz + iz.  Records are trace-major [ntr][nt].
 *
 * Pinned independently (tests/test_oracle_pins.py): coefficients, stencils,
 * window / medium / damping, Ricker, the propagators (energy, reversibility),
 * the adjoint half steps (dot-product), Born modeling (linearization), the
 * misfit and line-search sums, the overwrite injection with the last of
 * coincident traces winning, and the imaging step (Vp, Vs and density terms)
 * through or_gradient_consistent_f64, whose images must equal central
 * differences of the misfit.  The faithful gradient's orchestration
 * (tests/test_oracle_orchestration.py, through or_reconstruct_f64 and the
 * planted mistakes of or_config.plant): the strip save / restore geometry
 * with corners skipped (the reconstruction returns the forward field on the
 * imaged interior at every step, and a window one cell off fails), the source
 * timing and scale (closed form of the first sample), the reconstruction
 * without source subtraction (linearity against the exact inverse, which
 * time reversal pins).  The Hessian-vector product and double scattering
 * (tests/test_hessian_oracle.py): the consistent product's symmetry, its
 * quadratic form against the Born energy, central differences of the
 * consistent gradient, the reference's product as the gradient loop on the
 * scattered traces, the stage-1 Born cells, and double scattering as the
 * second-order term of the traces' expansion.  The stage-2 density override
 * (tests/test_oracle_orchestration.py): a run with it equals, bit for bit,
 * the run whose density file is the constant itself, also on the padding.
 * The imaging position in the faithful reverse loop (same file): with the
 * reconstruction made exact and the residual added (plants 8 and 128), the
 * loop's Vp / Vs images equal, to FP64 round-off, the consistent gradient of
 * the wavelet advanced by one sample (imaging after the adjoint stress-like
 * step pairs the derivatives of v(it + 1) with the consistent gradient's
 * adjoint stresses of one iteration later); imaging before that step (plant
 * 256) equals the consistent gradient of the wavelet itself instead.
 *
 * Random boundary (rbc): the reference draws it from cuRAND states seeded with
 * time(0) (sfwi-stage1/FD2DGPU.cu:155-230, 3245-3305), so no run repeats it;
 * this oracle and the library share a project-owned generator instead (the
 * lowbias32 integer hash of C. Wellons, 2018, on seed, cell and attempt),
 * computed in FP32 on both sides.  Pinned by the band values' range and vs/vp
 * ratio, an untouched interior and zero damping.  The gradient's
 * orchestration with it (no strips saved or restored,
 * sfwi-stage1/FD2DGPU.cu:4543, 4602): the exact inverse returns the zero start
 * on the whole grid by reversal alone, the faithful reconstruction minus it
 * is the source alone run backwards (Q5, linearity), and the faithful loop on
 * the random medium meets the same advanced-wavelet identity as above.
 */
#ifndef SFWI_ORACLE_H
#define SFWI_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_config {
    int allnx, allnz;     /* full model size (x-major, z fastest)                       */
    int nx, nz;           /* shot window (nz == allnz in the reference)                 */
    int pml, nt, ntr;     /* PML width, time samples, traces                             */
    float dx, dz, dt;     /* grid spacing (m), time step (s)                             */
    float rr;             /* PML reflection coefficient, 0.001 in the reference           */
    int left;             /* first model column of the window (delta_left+record_left)   */
    int isx, isz, igz;    /* source column/row and receiver row on the padded grid       */
    int ngx_left;         /* coordinate of the window's left edge (receiver placement)   */
    int adjoint_dir;      /* +1 stage 1, -1 stage 2                                       */
    float rho_const;      /* > 0: padded rho forced to this value (stage-2 kernel)       */
    int pml_off;          /* test hook: damping profiles forced to zero                   */
    int nthreads;         /* OpenMP threads, 0 = runtime default                          */
    int rbc;              /* 1: random boundary (random Vp / Vs in the PML bands, no      */
                          /*    damping, no strips: the reverse pass reconstructs by time */
                          /*    reversal alone); 2: the same medium with the strips kept  */
                          /*    (test hook)                                               */
    unsigned rbc_seed;    /* seed of the random boundary                                  */
    int snap_every;       /* snapshot gradient: image every K-th step (<= 1: every step)  */
    int plant;            /* test hook, 0 in use: planted mistakes that the pins must     */
                          /*   catch -- 1: the up strips one row deeper, 2: the left      */
                          /*   strips one column further in, 4: the forward source after  */
                          /*   the velocity update, 8: the reconstruction undoes the      */
                          /*   source (the exact inverse), 16: the Born source without    */
                          /*   the -4 rho vs dvs term, 32: the second-order scattered     */
                          /*   field driven by the background instead of the first order,*/
                          /*   64: the density override after lambda and mu are formed,  */
                          /*   128: the faithful reverse loop adds the residual to the    */
                          /*   totals and the x split parts instead of overwriting the    */
                          /*   totals, 256: it images before the adjoint step             */
} or_config;

/* FD coefficients c1..c4 of the 8th-order staggered stencil (index 0 unused). */
void or_coeffs_f32(float c[5]);
void or_coeffs_f64(double c[5]);

/* Ricker wavelet exactly as the host driver builds it. */
void or_ricker(float *w, int nt, float dt, float f0);

/* Padded medium of one shot window: vp, vs, rho, lambda, mu, dampx, dampz
 * (each nxpml*nzpml, z fastest).  out7 = 7 consecutive arrays. */
void or_setup_f32(const or_config *c, const float *vp, const float *vs, const float *rho, float *out7);
void or_setup_f64(const or_config *c, const float *vp, const float *vs, const float *rho, double *out7);

/* Forward solve: receiver traces of vz and vx, [ntr][nt]. */
void or_forward_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                    const float *wav, const int *gc, float *syn_z, float *syn_x);
void or_forward_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                    const float *wav, const int *gc, double *syn_z, double *syn_x);

/* Stage-1 per-shot gradient.  Accumulates into grad_vp, grad_vs, illum
 * (allnx*allnz) and misfit; writes the residual syn-obs when res_* non-NULL. */
void or_gradient_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                     const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                     float *grad_vp, float *grad_vs, float *illum, double *misfit,
                     float *res_z, float *res_x);
void or_gradient_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                     const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                     double *grad_vp, double *grad_vs, double *illum, double *misfit,
                     double *res_z, double *res_x);

/* Stage-1 gradient that also images the density (no reference counterpart):
 * grad_rho += -(vx_adj dvx + vz_adj dvz)/dt + (vp^2 - 2 vs^2) lg + vs^2 mg per
 * reverse step, dv the reconstructed velocity step, lg / mg the lambda / mu
 * correlations of the Vp / Vs imaging. */
void or_gradient_rho_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                         float *grad_vp, float *grad_vs, float *illum, float *grad_rho, double *misfit);
void or_gradient_rho_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                         double *grad_vp, double *grad_vs, double *illum, double *grad_rho, double *misfit);

/* Reverse loop (reconstruction, adjoint, imaging) driven by a caller-supplied
 * adjoint source res_* [ntr][nt] instead of the L2 residual: the gradient of
 * a misfit whose adjoint source is res_* (OT misfit).  Stage set by c. */
/* Consistent discrete gradient of the L2 misfit (FP64; a test device with no
 * reference counterpart): kept forward fields, additive 2 r / rho injection,
 * imaging between the adjoint half steps, images scaled by dt.  With
 * pml_off = 1 it is the exact gradient of the same scheme, so central
 * differences of the misfit pin the imaging formula (Vp, Vs and density). */
void or_gradient_consistent_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                                const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                                double *gvp, double *gvs, double *grho, double *misfit);
/* Snapshot gradient (FP64; no reference counterpart; the oracle of the
 * library's snapshot reconstruction mode): the consistent gradient above with
 * the imaging decimated to the steps it % K == 0 (K = snap_every), each
 * weighted by K, from the forward velocity after that step's velocity update
 * kept on the interior plus a 4-cell halo.  Reverse step: adjoint source at
 * the receivers (2 r / rho, or adj_* / rho when adj_z, adj_x are given:
 * the derivative of another misfit with respect to the synthetic traces),
 * adjoint velocity half step, imaging, adjoint stress half step, direction
 * +1.  dvp, dvs non-NULL: stage-2 Born forward, residual born - obs +
 * background, background velocity kept.  Accumulates K dt times the Vp / Vs
 * images into gvp, gvs, K sum (vx^2 + vz^2) into gil (allnx*allnz) and the
 * L2 misfit.  With K = 1 the Vp / Vs images equal or_gradient_consistent_f64's. */
void or_gradient_snapshot_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                              const float *dvp, const float *dvs, const float *wav, const int *gc,
                              const float *obs_z, const float *obs_x, const float *adj_z, const float *adj_x,
                              double *gvp, double *gvs, double *gil, double *misfit);
void or_gradient_res_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *wav, const int *gc, const float *res_z, const float *res_x,
                         float *grad_vp, float *grad_vs, float *illum);
void or_gradient_res_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *wav, const int *gc, const float *res_z, const float *res_x,
                         double *grad_vp, double *grad_vs, double *illum);

/* Stage-2 Born forward: scattered traces b_* and background traces bg_*. */
void or_born_forward_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *dvp, const float *dvs, const float *wav, const int *gc,
                         float *b_z, float *b_x, float *bg_z, float *bg_x);
void or_born_forward_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *dvp, const float *dvs, const float *wav, const int *gc,
                         double *b_z, double *b_x, double *bg_z, double *bg_x);

/* Stage-2 per-shot gradient (residual = born - obs + background). */
void or_born_gradient_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                          const float *dvp, const float *dvs, const float *wav, const int *gc,
                          const float *obs_z, const float *obs_x,
                          float *grad_vp, float *grad_vs, float *illum, double *misfit);
void or_born_gradient_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                          const float *dvp, const float *dvs, const float *wav, const int *gc,
                          const float *obs_z, const float *obs_x,
                          double *grad_vp, double *grad_vs, double *illum, double *misfit);

/* Stage-1 Born forward with double scattering (dead code of the reference,
 * kept as its own commented-out calls place it): born_modeling mode 1 then
 * mode 5 each step (sfwi-stage1/FD2DGPU.cu:1855-1860, 1888-1903), both Born
 * sources on the interior only (sfwi-stage1/born.cu:830-880), the second-order
 * field driven by the updated first-order scattered velocity.  b_*: first-order
 * scattered traces, b2_*: second-order ones (record_scatter with
 * double_scatter, :3610-3617). */
void or_born2_forward_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                          const float *dvp, const float *dvs, const float *wav, const int *gc,
                          float *b_z, float *b_x, float *b2_z, float *b2_x);
void or_born2_forward_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                          const float *dvp, const float *dvs, const float *wav, const int *gc,
                          double *b_z, double *b_x, double *b2_z, double *b2_x);

/* Gauss-Newton Hessian-vector product of one shot, ELSRTM_CALCULATE_Hessian_vector
 * (sfwi-stage1/FD2DGPU.cu:5236-5359): stage-1 Born forward of the vector p
 * (interior source) with background strips saved and scattered traces
 * recorded; the gradient's reverse loop with those traces as the adjoint
 * source; Vp / Vs imaging without illumination.  Accumulates H p into
 * gvp, gvs (allnx*allnz). */
void or_hessian_vector_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                           const float *p_vp, const float *p_vs, const float *wav, const int *gc,
                           float *gvp, float *gvs);
void or_hessian_vector_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                           const float *p_vp, const float *p_vs, const float *wav, const int *gc,
                           double *gvp, double *gvs);
/* The same product through the snapshot gradient (FP64, imaging every
 * snap_every-th step, no reference counterpart): the adjoint source 2 b / rho
 * of the scattered traces b.  With snap_every <= 1 and pml_off = 1 it is
 * 2 J^T J p, J the Jacobian of the traces with respect to (vp, vs) on the
 * interior: symmetric, and equal to central differences of the consistent
 * gradient at a model whose own traces are the observed ones. */
void or_hessian_vector_snapshot_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                                    const float *p_vp, const float *p_vs, const float *wav, const int *gc,
                                    double *gvp, double *gvs);

/* Stage-1 line search: misfits at vel, vel - step1*dir, vel - step2*dir. */
void or_parabola_misfits_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                             const float *dir_vp, const float *dir_vs, float step1, float step2,
                             const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                             double m[3]);
/* Stage-2 exact step: fenzi += sum L(cg).r, fenmu += sum L(cg)^2. */
void or_linear_alpha_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *dvp, const float *dvs, const float *cg_vp, const float *cg_vs,
                         const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                         double *fenzi, double *fenmu);

/* Test device for the faithful gradient's orchestration (FP64): its forward
 * loop with strips saved, from init (15 arrays, NULL: zero), then its reverse
 * reconstruction steps, no recording or adjoint.  kind 0: as the gradient
 * (the source left out of the reverse steps); 1: the exact inverse (each
 * reverse velocity step sees its step's source in the stress totals, as the
 * forward one did); 2: the difference of 0 and 1 by linearity (zero state
 * and zero strips, the negated source).  errs[nt] (NULL: skipped): largest difference of
 * the totals after reverse step it from the forward totals before forward
 * step it (velocities on cells at least pml + 4 from every edge, stresses on
 * the imaged interior, at least pml from every edge), relative to the
 * largest forward total there.  out: the state after reverse step 0. */
void or_reconstruct_f64(const or_config *c, const float *vp, const float *vs, const float *rho, const float *wav,
                        const double *init, int kind, double *errs, double *out);

/* Test hooks (FP64): the four stencils of one field, and free propagation of
 * a given state.  state = 15 arrays of nxpml*nzpml in the order
 * vx, vz, txx, tzz, txz, vx_x, vx_z, vz_x, vz_z, txx_x, txx_z, tzz_x, tzz_z, txz_x, txz_z.
 * kind 0: forward steps (v then sigma) with direction dir;
 * kind 1: reconstruction steps (sigma then v) with direction dir;
 * kind 2: adjoint steps (adjoint sigma then adjoint v) with direction dir;
 * kinds 3, 4, 5, 6: one half step only -- forward v, forward sigma,
 * adjoint sigma, adjoint v. */
void or_stencils_f64(int nxpml, int nzpml, const double *f,
                     double *dxp, double *dxm, double *dzp, double *dzm);
void or_propagate_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                      int kind, int dir, int nsteps, double *state);

#ifdef __cplusplus
}
#endif
#endif
