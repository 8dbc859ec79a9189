// sfwi_fd2dgpu.hpp -- header-only C++ adapter with the reference's per-shot
// interface (class FD2DGPU_ELASTIC and the free functions IMAGE_SINGLESHOT,
// CALCULATE_ALPHA_parabola_mb, CALCULATE_ALPHA) over the C ABI of
// libsfwi_amd.so, so a kept host driver (sfwi-stage{1,2}/main.cpp minus MPI)
// compiles against it with its call sites unchanged.
//
// Reference declarations it mirrors:
//   class FD2DGPU_ELASTIC          sfwi-stage1/FD2DGPU.cuh:146-459, sfwi-stage2/FD2DGPU.cuh:149-448
//   IMAGE_SINGLESHOT (stage 1)     sfwi-stage1/FD2DGPU.cuh:466-467
//   IMAGE_SINGLESHOT (stage 2)     sfwi-stage2/FD2DGPU.cuh:455-456
//   CALCULATE_ALPHA_parabola_mb    sfwi-stage1/FD2DGPU.cuh:471-472
//   CALCULATE_ALPHA                sfwi-stage2/FD2DGPU.cuh:457-458
//   ELSRTM_CALCULATE_Hessian_vector sfwi-stage1/FD2DGPU.cuh:477-478
//
// Differences a maintainer sees:
//   * the constructor takes one more, defaulted argument: the stage (1 or 2),
//     which sets the adjoint direction (+1 / -1) and the stage-2 density
//     (2000 kg/m^3) that the two reference builds hard-code;
//   * errors throw sfwi_ref::Error (the reference does not check errors);
//   * rbc = true (random boundary) draws it from a seeded project-owned
//     generator (FD2DGPU_ELASTIC::rbc_seed, default 1) where the reference
//     seeds cuRAND from the clock, and runs without strips, as the reference;
//     cpu_mem = true keeps the strips in pinned host memory in chunks of ntt
//     steps (sfwi_params.cpu_mem); the first gradient call after a flush may
//     switch the setting (the context is rebuilt), a later one may not;
//   * the OT misfit outputs (misfit_ot, misfit_step*_ot) receive the optimal-
//     transport misfit when the context was created with an OT misfit, else
//     they are left untouched (their reference computations are commented out).
#ifndef SFWI_FD2DGPU_HPP
#define SFWI_FD2DGPU_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "sfwi_abi.h"

namespace sfwi_ref {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const sfwi_ctx *ctx) {
    if (rc != SFWI_OK) throw Error(rc, std::string("sfwi: ") + sfwi_last_error(ctx));
}

class FD2DGPU_ELASTIC {
public:
    // Public per-shot fields the driver writes before each shot
    // (sfwi-stage1/main.cpp:596-625).
    int isx = 0, isz = 0, igz = 0;
    int delta_left = 0, record_left_in_v = 0, ngx_left = 0;
    int ntr = 0;
    float cmin = 0.0f;   // left edge of the calculation area (sfwi-stage1/FD2DGPU.cuh:164)
    const int nxpml, nzpml, allnx, allnz, pml, nt, nop;
    const float dx, dz, dt;

    // sfwi-stage1/FD2DGPU.cuh:152-153 (+ stage, misfit and batch extensions).
    FD2DGPU_ELASTIC(const float *sou, float dx1, float dz1, float dt1, int nxpml1, int nzpml1, int allnx1,
                    int allnz1, int /*scale*/, int pml1, int nt1, int nop1, int ntr1, int stage = 1,
                    int misfit = SFWI_MISFIT_L2, int mode = SFWI_MODE_FAST)
        : ntr(ntr1), nxpml(nxpml1), nzpml(nzpml1), allnx(allnx1), allnz(allnz1), pml(pml1), nt(nt1), nop(nop1),
          dx(dx1), dz(dz1), dt(dt1) {
        sfwi_params p{};
        p.dx = dx1; p.dz = dz1; p.dt = dt1;
        p.nx = nxpml1 - 2 * pml1; p.nz = nzpml1 - 2 * pml1;
        p.pml = pml1; p.nt = nt1; p.nop = nop1; p.ntr = ntr1;
        p.allnx = allnx1; p.allnz = allnz1;
        p.rr = 0.001f;                                   // sfwi-stage1/FD2DGPU.cu:460
        p.rho_const = stage == 2 ? 2000.0f : 0.0f;       // sfwi-stage2/FD2DGPU.cu:1111
        p.adjoint_dir = stage == 2 ? -1 : 1;             // sfwi-stage{1,2}/FD2DGPU.cu:4632 / 4462
        p.max_batch = 1;
        p.mode = mode;
        p.storage = SFWI_STORE_STRIPS_F32;
        p.misfit = misfit;
        check(sfwi_ctx_create(&p, &ctx_), nullptr);
        p_ = p;
        sou_.assign(sou, sou + nt1);
        ntr_max_ = ntr1;
        misfit_ = misfit;
        try {
            check(sfwi_set_wavelet(ctx_, sou), ctx_);
        } catch (...) {
            sfwi_ctx_destroy(ctx_);
            throw;
        }
    }
    ~FD2DGPU_ELASTIC() { sfwi_ctx_destroy(ctx_); }
    FD2DGPU_ELASTIC(const FD2DGPU_ELASTIC &) = delete;
    FD2DGPU_ELASTIC &operator=(const FD2DGPU_ELASTIC &) = delete;

    // Full models, stage 1 (sfwi-stage1/FD2DGPU.cuh:296) and stage 2 (rho fixed,
    // sfwi-stage2/FD2DGPU.cuh:297).
    void GPUbufferVPVS(float *vp_all, float *vs_all, float *rho_all) {
        vp_ = vp_all; vs_ = vs_all; rho_ = rho_all;
        check(sfwi_set_model(ctx_, vp_all, vs_all, rho_all), ctx_);
    }
    void GPUbufferVPVS(float *vp_all, float *vs_all) { GPUbufferVPVS(vp_all, vs_all, nullptr); }
    // Model perturbation (sfwi-stage1/FD2DGPU.cuh:237).
    void GPUbufferM(float *pp, float *ps) {
        dvp_ = pp; dvs_ = ps;
        check(sfwi_set_perturbation(ctx_, pp, ps), ctx_);
    }
    // Strip storage of the next gradient call (the per-call cpu_mem and ntt
    // arguments of the reference): a change rebuilds the context with the
    // model, perturbation and wavelet it holds, only while no gradient has
    // been accumulated since the last flush.
    // The random boundary (the per-call rbc argument) is switched the same way;
    // its generator is seeded with rbc_seed (default 1), where the reference
    // seeds cuRAND from the clock.
    void use_host_strips(bool cpu_mem, int ntt, bool rbc = false) {
        if (cpu_mem && ntt < 1) throw Error(SFWI_EINVAL, "sfwi: cpu_mem needs ntt >= 1 (steps per chunk)");
        const int want = cpu_mem ? 1 : 0, chunk = cpu_mem ? ntt : 0, rb = rbc ? 1 : 0;
        if (want == p_.cpu_mem && chunk == p_.strip_chunk && rb == p_.rbc) return;
        if (dirty_) throw Error(SFWI_ESTATE, "sfwi: cpu_mem / ntt / rbc changed with gradients accumulated");
        sfwi_params q = p_;
        q.cpu_mem = want;
        q.strip_chunk = chunk;
        q.rbc = rb;
        if (!q.rbc_seed) q.rbc_seed = rbc_seed;
        sfwi_ctx *c = nullptr;
        check(sfwi_ctx_create(&q, &c), nullptr);
        try {
            check(sfwi_set_wavelet(c, sou_.data()), c);
            if (vp_) check(sfwi_set_model(c, vp_, vs_, rho_), c);
            if (dvp_) check(sfwi_set_perturbation(c, dvp_, dvs_), c);
        } catch (...) {
            sfwi_ctx_destroy(c);
            throw;
        }
        sfwi_ctx_destroy(ctx_);
        ctx_ = c;
        p_ = q;
    }
    // Random boundary of the next forward-only call, the strip settings kept.
    void use_boundary(bool rbc) { use_host_strips(p_.cpu_mem != 0, p_.cpu_mem ? p_.strip_chunk : 0, rbc); }
    unsigned rbc_seed = 1;
    void mark_accumulated() { dirty_ = true; }
    // Observed records of the next shot (sfwi-stage1/FD2DGPU.cuh:387): kept by
    // pointer, read by the next per-shot call.
    void record_copytoGPU(float *record, float *record_x, int *gc) {
        rec_z_ = record; rec_x_ = record_x; gc_ = gc;
    }
    // Window of the next shot (sfwi-stage1/FD2DGPU.cuh:297).
    void bufferVpVsHtoD(int ngx_min) { left_ = ngx_min; }
    // Gradient flush (sfwi-stage1/FD2DGPU.cuh:248): the four Born-gradient
    // buffers stay zero on the live path.
    void imagebuffer_resettozero(float *cpu_pp_grad, float *cpu_ps_grad, float *illumination, float *b_pp,
                                 float *b_ps, float *b_pp2, float *b_ps2) {
        check(sfwi_flush(ctx_, cpu_pp_grad, cpu_ps_grad, illumination), ctx_);
        dirty_ = false;
        const size_t n = (size_t)allnx * allnz;
        for (float *b : {b_pp, b_ps, b_pp2, b_ps2})
            if (b)
                for (size_t i = 0; i < n; ++i) b[i] = 0.0f;
    }

    sfwi_ctx *ctx() const { return ctx_; }
    int misfit_kind() const { return misfit_; }
    const float *vp_all() const { return vp_; }
    const float *vs_all() const { return vs_; }
    const float *rho_all() const { return rho_; }

    // The shot descriptor of the next per-shot call.
    sfwi_shot shot(const float *record, const float *record_x, const int *gc) const {
        if (ntr < 1 || ntr > ntr_max_) throw Error(SFWI_EINVAL, "sfwi: ntr outside the context's trace count");
        sfwi_shot s{};
        s.left = left_;
        s.isx = isx; s.isz = isz; s.igz = igz;
        s.ngx_left = ngx_left;
        s.ntr = ntr;
        s.gc = gc ? gc : gc_;
        s.obs_z = record ? record : rec_z_;
        s.obs_x = record_x ? record_x : rec_x_;
        return s;
    }

private:
    sfwi_ctx *ctx_ = nullptr;
    sfwi_params p_{};
    std::vector<float> sou_;
    bool dirty_ = false;
    int ntr_max_ = 0, misfit_ = SFWI_MISFIT_L2, left_ = 0;
    const float *rec_z_ = nullptr, *rec_x_ = nullptr;
    const int *gc_ = nullptr;
    const float *vp_ = nullptr, *vs_ = nullptr, *rho_ = nullptr;
    const float *dvp_ = nullptr, *dvs_ = nullptr;
};


// Stage-1 per-shot gradient (sfwi-stage1/FD2DGPU.cuh:466-467).
inline void IMAGE_SINGLESHOT(FD2DGPU_ELASTIC &fd, const int /*sxnum*/, const int /*nDim*/, const int /*myid*/,
                             float *record, float *record_x, int *gc, bool rbc, bool cpu_mem, int ntt,
                             int /*iter*/, int /*iter_round1*/, int /*maxiter*/, char * /*snapdir*/,
                             float * /*image_pp*/, float * /*image_ps*/, float * /*cpu_pp_grad_old*/,
                             float * /*cpu_ps_grad_old*/, double &misfit, float *misfit_ot) {
    fd.use_host_strips(cpu_mem, ntt, rbc);
    const sfwi_shot s = fd.shot(record, record_x, gc);
    double m = 0.0;
    check(sfwi_shot_gradient(fd.ctx(), &s, &m), fd.ctx());
    fd.mark_accumulated();
    if (fd.misfit_kind() != SFWI_MISFIT_L2 && misfit_ot) *misfit_ot += (float)m;
    misfit += m;
}

// Stage-2 per-shot gradient of the Born misfit (sfwi-stage2/FD2DGPU.cuh:455-456).
inline void IMAGE_SINGLESHOT(FD2DGPU_ELASTIC &fd, const int /*sxnum*/, const int /*nDim*/, const int /*myid*/,
                             float *record, float *record_x, int *gc, bool rbc, bool cpu_mem, int ntt,
                             int /*iter*/, char * /*snapdir*/, float * /*image_pp*/, float * /*image_ps*/,
                             float * /*cpu_pp_grad_old*/, float * /*cpu_ps_grad_old*/, double &misfit3) {
    fd.use_host_strips(cpu_mem, ntt, rbc);
    const sfwi_shot s = fd.shot(record, record_x, gc);
    check(sfwi_shot_born_gradient(fd.ctx(), &s, &misfit3), fd.ctx());
    fd.mark_accumulated();
}

// Stage-1 parabolic line search (sfwi-stage1/FD2DGPU.cuh:471-472): misfits of
// forward solves at vel, vel - step_1*grad_old and vel - step_2*grad_old.
inline void CALCULATE_ALPHA_parabola_mb(FD2DGPU_ELASTIC &fd, const int /*sxnum*/, const int /*nDim*/,
                                        const int /*myid*/, float *record, float *record_x, int *gc, bool rbc,
                                        bool /*cpu_mem*/, int /*ntt*/, int /*iter*/, float * /*image_pp*/,
                                        float * /*image_ps*/, float *cpu_pp_grad_old, float *cpu_ps_grad_old,
                                        double &misfit_step0, double &misfit_step1, double &misfit_step2,
                                        float step_1, float step_2, float *misfit_step0_ot,
                                        float *misfit_step1_ot, float *misfit_step2_ot) {
    fd.use_boundary(rbc);   // forward solves only: no strips, cpu_mem has no effect
    const sfwi_shot s = fd.shot(record, record_x, gc);
    double m[3] = {0.0, 0.0, 0.0};
    check(sfwi_linesearch_parabola(fd.ctx(), &s, cpu_pp_grad_old, cpu_ps_grad_old, step_1, step_2, m), fd.ctx());
    misfit_step0 += m[0];
    misfit_step1 += m[1];
    misfit_step2 += m[2];
    if (fd.misfit_kind() != SFWI_MISFIT_L2) {
        float *ot[3] = {misfit_step0_ot, misfit_step1_ot, misfit_step2_ot};
        for (int k = 0; k < 3; ++k)
            if (ot[k]) *ot[k] += (float)m[k];
    }
}

// Stage-2 exact step (sfwi-stage2/FD2DGPU.cuh:457-458): fenzi += sum L(cg).r,
// fenmu += sum L(cg)^2, cg = (cpu_pp_grad_old, cpu_ps_grad_old).
inline void CALCULATE_ALPHA(FD2DGPU_ELASTIC &fd, const int /*sxnum*/, const int /*nDim*/, const int /*myid*/,
                            float *record, float *record_x, int *gc, char * /*snapdir*/, bool rbc, bool /*cpu_mem*/,
                            int /*ntt*/, int /*iter*/, float * /*image_pp*/, float * /*image_ps*/,
                            float *cpu_pp_grad_old, float *cpu_ps_grad_old, double &fenzi, double &fenmu) {
    fd.use_boundary(rbc);   // forward solves only: no strips, cpu_mem has no effect
    const sfwi_shot s = fd.shot(record, record_x, gc);
    check(sfwi_linesearch_linear(fd.ctx(), &s, cpu_pp_grad_old, cpu_ps_grad_old, &fenzi, &fenmu), fd.ctx());
}

// Gauss-Newton Hessian-vector product of one shot (sfwi-stage1/FD2DGPU.cuh:477-478),
// the inner CG's per-shot call (sfwi-stage1/main.cpp:928): H p for
// p = (grad_pp, grad_ps) is accumulated with the gradients and read back by
// the flush (imagebuffer_resettozero), as in the reference; the observed
// records are not used.
inline void ELSRTM_CALCULATE_Hessian_vector(FD2DGPU_ELASTIC &fd, const int /*sxnum*/, const int /*nDim*/,
                                            const int /*myid*/, float *record, float *record_x, int *gc, bool rbc,
                                            bool cpu_mem, int ntt, int /*iter*/, float *grad_pp, float *grad_ps) {
    fd.use_host_strips(cpu_mem, ntt, rbc);
    const sfwi_shot s = fd.shot(record, record_x, gc);
    check(sfwi_shot_hessian_vector(fd.ctx(), &s, grad_pp, grad_ps), fd.ctx());
    fd.mark_accumulated();
}

}  // namespace sfwi_ref

#endif  // SFWI_FD2DGPU_HPP
