// Host steps of the inversion loops around the per-shot solves (host code,
// included by sfwi.hip).  They follow the reference drivers' arithmetic type
// by type: FP32 arrays, FP64 accumulators and step lengths, products of two
// floats formed in FP32 before they reach an FP64 sum, the same loop order
// (model column x outer, depth z inner, index x*allnz + z) and no FMA
// contraction (the translation unit is built with -ffp-contract=off).
//
// Stage 1 (sfwi-stage1/main.cpp, update_mp = true, update_mb = false,
// max_inner = 0, sd = true, the reference's configuration):
//   direction : illumination preconditioning and the cut layer (695-717),
//               conjugate gradient with beta = fenzi / fenmu, 0 at iteration
//               0 (719-750), the direction copied and normalized by the joint
//               max-abs of both components (829-839, again at 1059-1062),
//   step      : the parabola through three misfits, cal_step2 (65-72),
//               clamped to [-100, 300] (1221-1238),
//   update    : vel -= alpha * direction, then every cell above the cut
//               layer set to the model's first cell (1241-1257).
// Stage 2 (sfwi-stage2/main.cpp):
//   direction : rows below 26 preconditioned by the illumination, rows above
//               zeroed in the gradient itself, beta from sum g^2 / illum,
//               conjugate gradient (645-703),
//   update    : alpha = fenzi / fenmu from the linear step, image -= alpha *
//               cg (861-873).
#pragma once

namespace {

// normalize_all (sfwi-stage1/main.cpp:35-49): both arrays divided by the
// largest magnitude over the two.
void inv_normalize_all(float *a, float *b, size_t n) {
    float mx = 0.0f;
    for (size_t k = 0; k < n; ++k) mx = (std::fabs(a[k]) > mx) ? std::fabs(a[k]) : mx;
    for (size_t k = 0; k < n; ++k) mx = (std::fabs(b[k]) > mx) ? std::fabs(b[k]) : mx;
    for (size_t k = 0; k < n; ++k) {
        a[k] = a[k] / mx;
        b[k] = b[k] / mx;
    }
}

bool inv_dims_ok(int allnx, int allnz) { return allnx > 0 && allnz > 0; }

}  // namespace

extern "C" {

int sfwi_s1_direction(int allnx, int allnz, int iter, int cut_layer, const float *pp_grad, const float *ps_grad,
                      const float *illum, float *pp_cg_old, float *ps_cg_old, double *beta_fenmu, float *dir_pp,
                      float *dir_ps, double *beta_out) {
    if (!inv_dims_ok(allnx, allnz) || iter < 0 || cut_layer < 0)
        return fail(nullptr, SFWI_EINVAL, "sfwi_s1_direction: allnx=%d allnz=%d iter=%d cut_layer=%d", allnx,
                    allnz, iter, cut_layer);
    if (!pp_grad || !ps_grad || !illum || !pp_cg_old || !ps_cg_old || !beta_fenmu || !dir_pp || !dir_ps)
        return fail(nullptr, SFWI_EINVAL, "sfwi_s1_direction: NULL array");
    const size_t n = (size_t)allnx * allnz;
    // sum_*_grad = grad / illumination, zero above the cut layer
    std::vector<float> spp(n), sps(n);
    for (int i = 0; i < allnx; ++i)
        for (int j = 0; j < allnz; ++j) {
            const size_t k = (size_t)i * allnz + j;
            spp[k] = pp_grad[k] / illum[k];
            sps[k] = ps_grad[k] / illum[k];
        }
    for (int i = 0; i < allnx; ++i)
        for (int j = 0; j < allnz; ++j)
            if (j < cut_layer) {
                const size_t k = (size_t)i * allnz + j;
                spp[k] = 0.0f;
                sps[k] = 0.0f;
            }
    // cg = sum grad; fenzi = sum g * (g / illum) with the raw gradient
    double fenzi = 0.0;
    for (int i = 0; i < allnx; ++i)
        for (int j = 0; j < allnz; ++j) {
            const size_t k = (size_t)i * allnz + j;
            dir_pp[k] = spp[k];
            dir_ps[k] = sps[k];
            const float t = pp_grad[k] * spp[k] + ps_grad[k] * sps[k];
            fenzi += t;
        }
    const double beta = iter == 0 ? 0.0 : fenzi / *beta_fenmu;
    for (size_t k = 0; k < n; ++k) {
        dir_pp[k] = (float)((double)dir_pp[k] + beta * (double)pp_cg_old[k]);
        dir_ps[k] = (float)((double)dir_ps[k] + beta * (double)ps_cg_old[k]);
    }
    std::memcpy(pp_cg_old, dir_pp, n * sizeof(float));
    std::memcpy(ps_cg_old, dir_ps, n * sizeof(float));
    *beta_fenmu = fenzi;
    // the search direction: the CG direction normalized (twice, as the reference)
    inv_normalize_all(dir_pp, dir_ps, n);
    inv_normalize_all(dir_pp, dir_ps, n);
    if (beta_out) *beta_out = beta;
    return 0;
}

// Gauss-Newton inner CG of stage 1 (sfwi-stage1/main.cpp:784-1063, run when
// max_inner > 0).  The state arrays are allnx*allnz FP32 each, in the order
// r, p, delta, grad (pp then ps of each).
int sfwi_s1_hcg_init(int allnx, int allnz, const float *pp_cg, const float *ps_cg, float *r_pp, float *r_ps,
                     float *p_pp, float *p_ps, float *delta_pp, float *delta_ps, float *grad_pp, float *grad_ps) {
    if (!inv_dims_ok(allnx, allnz)) return fail(nullptr, SFWI_EINVAL, "sfwi_s1_hcg_init: allnx=%d allnz=%d", allnx, allnz);
    if (!pp_cg || !ps_cg || !r_pp || !r_ps || !p_pp || !p_ps || !delta_pp || !delta_ps || !grad_pp || !grad_ps)
        return fail(nullptr, SFWI_EINVAL, "sfwi_s1_hcg_init: NULL array");
    const size_t n = (size_t)allnx * allnz;
    // r = cg normalized, p = -r (:809-822)
    std::memcpy(r_pp, pp_cg, n * sizeof(float));
    std::memcpy(r_ps, ps_cg, n * sizeof(float));
    inv_normalize_all(r_pp, r_ps, n);
    for (size_t k = 0; k < n; ++k) {
        p_pp[k] = -r_pp[k];
        p_ps[k] = -r_ps[k];
    }
    // grad = cg, delta = -grad, both normalized (:826-839)
    std::memcpy(grad_pp, pp_cg, n * sizeof(float));
    std::memcpy(grad_ps, ps_cg, n * sizeof(float));
    for (size_t k = 0; k < n; ++k) {
        delta_pp[k] = -grad_pp[k];
        delta_ps[k] = -grad_ps[k];
    }
    inv_normalize_all(delta_pp, delta_ps, n);
    inv_normalize_all(grad_pp, grad_ps, n);
    return 0;
}

int sfwi_s1_hcg_step(int allnx, int allnz, int cut_grad, int cut_layer, float *hp_pp, float *hp_ps, float *r_pp,
                     float *r_ps, float *p_pp, float *p_ps, float *delta_pp, float *delta_ps, float *grad_pp,
                     float *grad_ps, double *rr_old, double *rr_new, double *alpha_out, double *beta_out) {
    if (!inv_dims_ok(allnx, allnz) || cut_layer < 0)
        return fail(nullptr, SFWI_EINVAL, "sfwi_s1_hcg_step: allnx=%d allnz=%d cut_layer=%d", allnx, allnz, cut_layer);
    if (!hp_pp || !hp_ps || !r_pp || !r_ps || !p_pp || !p_ps || !delta_pp || !delta_ps || !grad_pp || !grad_ps)
        return fail(nullptr, SFWI_EINVAL, "sfwi_s1_hcg_step: NULL array");
    const size_t n = (size_t)allnx * allnz;
    // Hp above the cut layer zeroed; alpha = r.r / p.Hp (:961-985)
    double a_num = 0.0, a_den = 0.0;
    for (int i = 0; i < allnx; ++i)
        for (int j = 0; j < allnz; ++j) {
            const size_t k = (size_t)i * allnz + j;
            if (cut_grad && j < cut_layer) {
                hp_pp[k] = 0.0f;
                hp_ps[k] = 0.0f;
            }
            const float rr = r_pp[k] * r_pp[k] + r_ps[k] * r_ps[k];
            a_num += rr;
            const float ph = p_pp[k] * hp_pp[k] + p_ps[k] * hp_ps[k];
            a_den += ph;
        }
    const double alpha = a_num / a_den;
    // delta += alpha p, grad = -delta, r += alpha Hp, beta = r.r / r.r_old (:999-1017)
    double b_num = 0.0;
    for (int i = 0; i < allnx; ++i)
        for (int j = 0; j < allnz; ++j) {
            const size_t k = (size_t)i * allnz + j;
            delta_pp[k] = (float)((double)delta_pp[k] + alpha * (double)p_pp[k]);
            delta_ps[k] = (float)((double)delta_ps[k] + alpha * (double)p_ps[k]);
            grad_pp[k] = -delta_pp[k];
            grad_ps[k] = -delta_ps[k];
            r_pp[k] = (float)((double)r_pp[k] + alpha * (double)hp_pp[k]);
            r_ps[k] = (float)((double)r_ps[k] + alpha * (double)hp_ps[k]);
            const float rr = r_pp[k] * r_pp[k] + r_ps[k] * r_ps[k];
            b_num += rr;
        }
    const double beta = b_num / a_num;
    // p = -r + beta p (:1019-1024)
    for (size_t k = 0; k < n; ++k) {
        p_pp[k] = (float)(-(double)r_pp[k] + beta * (double)p_pp[k]);
        p_ps[k] = (float)(-(double)r_ps[k] + beta * (double)p_ps[k]);
    }
    if (rr_old) *rr_old = a_num;
    if (rr_new) *rr_new = b_num;
    if (alpha_out) *alpha_out = alpha;
    if (beta_out) *beta_out = beta;
    return 0;
}

int sfwi_s1_hcg_finish(int allnx, int allnz, float *delta_pp, float *delta_ps, float *grad_pp, float *grad_ps) {
    if (!inv_dims_ok(allnx, allnz)) return fail(nullptr, SFWI_EINVAL, "sfwi_s1_hcg_finish: allnx=%d allnz=%d", allnx, allnz);
    if (!delta_pp || !delta_ps || !grad_pp || !grad_ps) return fail(nullptr, SFWI_EINVAL, "sfwi_s1_hcg_finish: NULL array");
    const size_t n = (size_t)allnx * allnz;
    inv_normalize_all(delta_pp, delta_ps, n);
    inv_normalize_all(grad_pp, grad_ps, n);
    return 0;
}

double sfwi_s1_alpha(double m0, double m1, double m2, float step1, float step2) {
    const double fenzi = (step1 * step1 - step2 * step2) * m0 + (step2 * step2) * m1 + (-step1 * step1) * m2;
    const double fenmu = (step1 - step2) * m0 + (step2)*m1 + (-step1) * m2;
    double alpha = fenzi / fenmu / 2.0f;
    if (alpha > 300) alpha = 300;
    if (alpha < -100) alpha = -100;
    return alpha;
}

int sfwi_s1_update(int allnx, int allnz, int cut_layer, double alpha, const float *dir_pp, const float *dir_ps,
                   float *velp, float *vels) {
    if (!inv_dims_ok(allnx, allnz) || cut_layer < 0)
        return fail(nullptr, SFWI_EINVAL, "sfwi_s1_update: allnx=%d allnz=%d cut_layer=%d", allnx, allnz, cut_layer);
    if (!dir_pp || !dir_ps || !velp || !vels) return fail(nullptr, SFWI_EINVAL, "sfwi_s1_update: NULL array");
    for (int i = 0; i < allnx; ++i)
        for (int j = 0; j < allnz; ++j) {
            const size_t k = (size_t)i * allnz + j;
            velp[k] = (float)((double)velp[k] - alpha * (double)dir_pp[k]);
            vels[k] = (float)((double)vels[k] - alpha * (double)dir_ps[k]);
        }
    for (int i = 0; i < allnx; ++i)
        for (int j = 0; j < allnz && j < cut_layer; ++j) {
            velp[(size_t)i * allnz + j] = velp[0];
            vels[(size_t)i * allnz + j] = vels[0];
        }
    return 0;
}

int sfwi_s2_direction(int allnx, int allnz, int iter, int cut_layer, float *pp_grad, float *ps_grad,
                      const float *illum, float *pp_cg_old, float *ps_cg_old, double *beta_fenmu, float *pp_cg,
                      float *ps_cg, double *beta_out) {
    if (!inv_dims_ok(allnx, allnz) || iter < 0 || cut_layer < 0)
        return fail(nullptr, SFWI_EINVAL, "sfwi_s2_direction: allnx=%d allnz=%d iter=%d cut_layer=%d", allnx,
                    allnz, iter, cut_layer);
    if (!pp_grad || !ps_grad || !illum || !pp_cg_old || !ps_cg_old || !beta_fenmu || !pp_cg || !ps_cg)
        return fail(nullptr, SFWI_EINVAL, "sfwi_s2_direction: NULL array");
    const size_t n = (size_t)allnx * allnz;
    double fenzi = 0.0;
    for (int i = 0; i < allnx; ++i)
        for (int j = 0; j < allnz; ++j) {
            const size_t k = (size_t)i * allnz + j;
            if (j >= cut_layer) {
                pp_cg[k] = pp_grad[k] / illum[k];
                ps_cg[k] = ps_grad[k] / illum[k];
            } else {
                pp_grad[k] = 0.0f;
                ps_grad[k] = 0.0f;
                pp_cg[k] = pp_grad[k];
                ps_cg[k] = ps_grad[k];
            }
            const float t = pp_grad[k] * pp_grad[k] / illum[k] + ps_grad[k] * ps_grad[k] / illum[k];
            fenzi += t;
        }
    const double beta = iter == 0 ? 0.0 : fenzi / *beta_fenmu;
    for (size_t k = 0; k < n; ++k) {
        pp_cg[k] = (float)((double)pp_cg[k] + beta * (double)pp_cg_old[k]);
        ps_cg[k] = (float)((double)ps_cg[k] + beta * (double)ps_cg_old[k]);
    }
    std::memcpy(pp_cg_old, pp_cg, n * sizeof(float));
    std::memcpy(ps_cg_old, ps_cg, n * sizeof(float));
    *beta_fenmu = fenzi;
    if (beta_out) *beta_out = beta;
    return 0;
}

int sfwi_s2_update(int allnx, int allnz, double fenzi, double fenmu, const float *pp_cg, const float *ps_cg,
                   float *image_pp, float *image_ps, double *alpha_out) {
    if (!inv_dims_ok(allnx, allnz)) return fail(nullptr, SFWI_EINVAL, "sfwi_s2_update: allnx=%d allnz=%d", allnx, allnz);
    if (!pp_cg || !ps_cg || !image_pp || !image_ps) return fail(nullptr, SFWI_EINVAL, "sfwi_s2_update: NULL array");
    const double alpha = fenzi / fenmu;
    for (int i = 0; i < allnx; ++i)
        for (int j = 0; j < allnz; ++j) {
            const size_t k = (size_t)i * allnz + j;
            image_pp[k] = (float)((double)image_pp[k] - alpha * (double)pp_cg[k]);
            image_ps[k] = (float)((double)image_ps[k] - alpha * (double)ps_cg[k]);
        }
    if (alpha_out) *alpha_out = alpha;
    return 0;
}

}  // extern "C"
