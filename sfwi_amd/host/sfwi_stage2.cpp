// Stage-2 inversion driver over the C ABI: the reference's stage-2 program
// (sfwi-stage2/main.cpp) for one GPU, without MPI.  It reads the same
// parameter file layout, Vp / Vs background models and SU records as the
// stage-1 program and runs the stage-2 iterations on the images, which start
// at zero (main.cpp:978-979): Born gradient of every shot with the current
// images as the perturbation (main.cpp:513-594), preconditioned CG direction
// (645-703), exact linear step of the two Born solves per shot (766-859),
// image update (861-873).
//
//   sfwi_stage2 <para file> [niter [max_batch [cut_layer]]]
//
// niter defaults to the file's maxiter, cut_layer to the reference's 26;
// shots outside [minshot, maxshot] are skipped.  The padded density is 2000
// and the adjoint runs in the stage-2 direction, as the stage-2 kernels hard
// code them.  Every aperture width gets a context of its own; shots of a width
// are solved in batches of max_batch (default 64), the step's sums equal the
// reference's per-shot CALCULATE_ALPHA calls added in shot order; the step
// takes each shot's residual from the gradient call in place of its first
// Born solve.  With
// cpu_mem = 1 in the file the boundary strips live in pinned host memory in
// chunks of iointerval steps; with rbc = 1 the PML bands take random
// velocities (seed 1) and no strips are kept.
// Per iteration: one line "iter <k> misfit <m> beta <b> alpha <a> fenzi <z>
// fenmu <u>" on stdout and, as the reference writes them, <outdir>/misfit.txt
// lines and <outdir>/sum_grad_{pp,ps}, sum_cggrad_{pp,ps}, sum_illum and
// sum_image_{pp,ps} _<k>_<allnx>_<allnz>.dat.  (The reference's split
// gradient dumps sum_bgrad*, all_grad* are not written.)
// Exit status: 0 success, 2 input error, 3 library error.
#include <cstdio>
#include <cstdlib>

#include "survey_in.hpp"

using namespace sfwi_host;

int main(int argc, char **argv) {
    if (argc < 2 || argc > 5) {
        std::fprintf(stderr, "usage: %s para_file [niter [max_batch [cut_layer]]]\n", argv[0]);
        return 2;
    }
    std::vector<sfwi_ctx *> ctxs;
    try {
        const Survey sv = read_survey(argv[1]);
        const sfwi_para &p = sv.p;
        const int niter = argc >= 3 ? std::atoi(argv[2]) : p.maxiter;
        const int max_batch = argc >= 4 ? std::atoi(argv[3]) : 64;
        const int cut_layer = argc >= 5 ? std::atoi(argv[4]) : 26;   // sfwi-stage2/main.cpp:648
        if (niter < 0 || max_batch < 1 || cut_layer < 0)
            throw Fail(2, "input error: niter must be >= 0, max_batch >= 1, cut_layer >= 0");
        const size_t nm = (size_t)p.allnx * p.allnz;
        std::vector<float> wav(sv.nt);
        sfwi_ricker(wav.data(), sv.nt, sv.dt, p.f0);
        for (int w : sv.widths) {
            sfwi_params q = width_params(sv, w, max_batch);
            q.rho_const = 2000.0f;   // the stage-2 medium kernel's density
            q.adjoint_dir = -1;      // the stage-2 adjoint's direction
            sfwi_ctx *c = nullptr;
            ck(sfwi_ctx_create(&q, &c), nullptr);
            ctxs.push_back(c);
            ck(sfwi_set_wavelet(c, wav.data()), c);
        }
        mkdir(p.outdir, 0755);
        const std::string outdir = p.outdir;
        std::vector<float> image_pp(nm, 0.0f), image_ps(nm, 0.0f), g[3], t[3];
        std::vector<float> pp_cg(nm), ps_cg(nm), pp_cg_old(nm, 0.0f), ps_cg_old(nm, 0.0f);
        for (int k = 0; k < 3; ++k) { g[k].resize(nm); t[k].resize(nm); }
        double beta_fenmu = 0.0, max_misfit = 0.0;
        // each width's residuals from its gradient calls, handed to the step
        // in place of its first Born solve (the same sums bit for bit)
        const size_t rec = (size_t)sv.ntr_max * sv.nt;
        std::vector<std::vector<float>> res_z(sv.widths.size()), res_x(sv.widths.size());
        for (int iter = 0; iter < niter; ++iter) {
            double misfit = 0.0;
            for (size_t gi = 0; gi < sv.widths.size(); ++gi) {
                sfwi_ctx *c = ctxs[gi];
                const std::vector<sfwi_shot> grp = group_shots(sv, sv.widths[gi]);
                ck(sfwi_set_model(c, sv.vp.data(), sv.vs.data(), nullptr), c);
                ck(sfwi_set_perturbation(c, image_pp.data(), image_ps.data()), c);
                res_z[gi].assign(grp.size() * rec, 0.0f);
                res_x[gi].assign(grp.size() * rec, 0.0f);
                const int nb = std::min<int>(max_batch, (int)grp.size());
                for (size_t b0 = 0; b0 < grp.size(); b0 += nb) {
                    const int n = (int)std::min<size_t>(nb, grp.size() - b0);
                    std::vector<double> m(n, 0.0);
                    ck(sfwi_batch_born_gradient_res(c, n, grp.data() + b0, m.data(), res_z[gi].data() + b0 * rec,
                                                    res_x[gi].data() + b0 * rec),
                       c);
                    for (double v : m) misfit += v;
                }
                ck(sfwi_flush(c, t[0].data(), t[1].data(), t[2].data()), c);
                for (int k = 0; k < 3; ++k)
                    for (size_t i = 0; i < nm; ++i) g[k][i] = gi == 0 ? t[k][i] : g[k][i] + t[k][i];
            }
            if (iter == 0) max_misfit = misfit;
            if (FILE *f = std::fopen((outdir + "/misfit.txt").c_str(), "a")) {
                std::fprintf(f, "i = %d,misfit = %4.3e,rms = %f\n", iter, misfit, misfit / max_misfit);
                std::fclose(f);
            }
            double beta = 0.0;
            ck(sfwi_s2_direction(p.allnx, p.allnz, iter, cut_layer, g[0].data(), g[1].data(), g[2].data(),
                                 pp_cg_old.data(), ps_cg_old.data(), &beta_fenmu, pp_cg.data(), ps_cg.data(), &beta),
               nullptr);
            write_model(out_name(sv, "sum_grad_pp", iter), g[0]);
            write_model(out_name(sv, "sum_grad_ps", iter), g[1]);
            write_model(out_name(sv, "sum_cggrad_pp", iter), pp_cg);
            write_model(out_name(sv, "sum_cggrad_ps", iter), ps_cg);
            write_model(out_name(sv, "sum_illum", iter), g[2]);
            double fenzi = 0.0, fenmu = 0.0;
            for (size_t gi = 0; gi < sv.widths.size(); ++gi) {   // batched; the sums of per-shot calls in shot order
                const std::vector<sfwi_shot> grp = group_shots(sv, sv.widths[gi]);
                ck(sfwi_batch_linesearch_linear_res(ctxs[gi], (int)grp.size(), grp.data(), pp_cg.data(),
                                                    ps_cg.data(), res_z[gi].data(), res_x[gi].data(), &fenzi,
                                                    &fenmu),
                   ctxs[gi]);
            }
            double alpha = 0.0;
            ck(sfwi_s2_update(p.allnx, p.allnz, fenzi, fenmu, pp_cg.data(), ps_cg.data(), image_pp.data(),
                              image_ps.data(), &alpha),
               nullptr);
            std::printf("iter %d misfit %.17g beta %.17g alpha %.17g fenzi %.17g fenmu %.17g\n", iter, misfit, beta,
                        alpha, fenzi, fenmu);
            std::fflush(stdout);
            write_model(out_name(sv, "sum_image_pp", iter), image_pp);
            write_model(out_name(sv, "sum_image_ps", iter), image_ps);
        }
    } catch (const Fail &e) {
        std::fprintf(stderr, "%s\n", e.what());
        for (sfwi_ctx *c : ctxs) sfwi_ctx_destroy(c);
        return e.status;
    }
    for (sfwi_ctx *c : ctxs) sfwi_ctx_destroy(c);
    return 0;
}
