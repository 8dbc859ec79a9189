// Stage-1 inversion driver over the C ABI: the reference's stage-1 program
// (sfwi-stage1/main.cpp) for one GPU, without MPI.  It reads the reference's
// parameter file, x-major FP32 models and SU records, indexes the shots,
// derives each shot's window the way the worker does, and runs the stage-1
// iterations (batched gradients, preconditioned CG direction, parabolic line
// search along it, clamped step, model update with the top-row reset).
//
//   sfwi_stage1 <para file> [niter [max_batch [cut_layer [max_inner]]]]
//
// niter defaults to the file's maxiter, cut_layer to the reference's 26,
// max_inner to the reference's 0 (main.cpp:408); with max_inner > 0 each
// iteration runs the Gauss-Newton inner CG (main.cpp:784-1063: batched
// Hessian-vector products of every shot, the host step, the normalized
// -delta as the search direction) and logs its hessian_misfit lines to
// misfit.txt as the reference does.  Shots whose number lies outside
// [minshot, maxshot] are skipped.  The density is forced to 2000 as the
// reference does (sfwi-stage1/main.cpp:1476-1479).  Where the reference sizes
// its one context by shot 0's aperture and runs other apertures on it
// (main.cpp:610-617), every aperture width here gets a context of its own;
// shots of a width are solved in batches of max_batch (default 64).  With
// cpu_mem = 1 in the file the boundary strips live in pinned host memory in
// chunks of iointerval steps (the reference's ntt), as the reference does;
// with rbc = 1 the PML bands take random velocities (seed 1) and no strips
// are kept.
// Per iteration: one line "iter <k> misfit <m> beta <b> alpha <a>" on stdout,
// and, as the reference writes them, <outdir>/misfit.txt lines and
// <outdir>/vel_p_<k>_<allnx>_<allnz>.dat, vel_s_<k>_<allnx>_<allnz>.dat.
// Exit status: 0 success, 2 input error, 3 library error.
#include <cstdio>
#include <cstdlib>

#include "survey_in.hpp"

using namespace sfwi_host;

int main(int argc, char **argv) {
    if (argc < 2 || argc > 6) {
        std::fprintf(stderr, "usage: %s para_file [niter [max_batch [cut_layer [max_inner]]]]\n", argv[0]);
        return 2;
    }
    std::vector<sfwi_ctx *> ctxs;
    try {
        const Survey sv = read_survey(argv[1]);
        const sfwi_para &p = sv.p;
        const int niter = argc >= 3 ? std::atoi(argv[2]) : p.maxiter;
        const int max_batch = argc >= 4 ? std::atoi(argv[3]) : 64;
        const int cut_layer = argc >= 5 ? std::atoi(argv[4]) : 26;   // sfwi-stage1/main.cpp:410
        const int max_inner = argc >= 6 ? std::atoi(argv[5]) : 0;    // sfwi-stage1/main.cpp:408
        if (niter < 0 || max_batch < 1 || cut_layer < 0 || max_inner < 0)
            throw Fail(2, "input error: niter must be >= 0, max_batch >= 1, cut_layer >= 0, max_inner >= 0");
        const size_t nm = (size_t)p.allnx * p.allnz;
        std::vector<float> vp = sv.vp, vs = sv.vs, rho(nm, 2000.0f);
        std::vector<float> wav(sv.nt);
        sfwi_ricker(wav.data(), sv.nt, sv.dt, p.f0);
        for (int w : sv.widths) {   // one context per aperture width, in order of first appearance
            const sfwi_params q = width_params(sv, w, max_batch);
            sfwi_ctx *c = nullptr;
            ck(sfwi_ctx_create(&q, &c), nullptr);
            ctxs.push_back(c);
            ck(sfwi_set_wavelet(c, wav.data()), c);
        }
        mkdir(p.outdir, 0755);
        const std::string outdir = p.outdir;
        std::vector<float> g[3], t[3], pp_cg_old(nm, 0.0f), ps_cg_old(nm, 0.0f), dir_pp(nm), dir_ps(nm);
        for (int k = 0; k < 3; ++k) { g[k].resize(nm); t[k].resize(nm); }
        double beta_fenmu = 0.0, max_misfit = 0.0;
        const float step1 = 30.0f, step2 = 150.0f;   // sfwi-stage1/main.cpp:373-374
        for (int iter = 0; iter < niter; ++iter) {
            double misfit = 0.0;
            for (size_t gi = 0; gi < sv.widths.size(); ++gi) {
                sfwi_ctx *c = ctxs[gi];
                const std::vector<sfwi_shot> grp = group_shots(sv, sv.widths[gi]);
                ck(sfwi_set_model(c, vp.data(), vs.data(), rho.data()), c);
                const int nb = std::min<int>(max_batch, (int)grp.size());
                for (size_t b0 = 0; b0 < grp.size(); b0 += nb) {
                    const int n = (int)std::min<size_t>(nb, grp.size() - b0);
                    std::vector<double> m(n, 0.0);
                    ck(sfwi_batch_gradient(c, n, grp.data() + b0, m.data()), c);
                    for (double v : m) misfit += v;
                }
                ck(sfwi_flush(c, t[0].data(), t[1].data(), t[2].data()), c);
                for (int k = 0; k < 3; ++k)
                    for (size_t i = 0; i < nm; ++i) g[k][i] = gi == 0 ? t[k][i] : g[k][i] + t[k][i];
            }
            if (iter == 0) max_misfit = misfit;
            double beta = 0.0;
            ck(sfwi_s1_direction(p.allnx, p.allnz, iter, cut_layer, g[0].data(), g[1].data(), g[2].data(),
                                 pp_cg_old.data(), ps_cg_old.data(), &beta_fenmu, dir_pp.data(), dir_ps.data(),
                                 &beta),
               nullptr);
            if (max_inner > 0) {   // Gauss-Newton inner CG: its direction replaces the CG one
                std::vector<float> h[8], hp[2], ht[3];
                for (auto &a : h) a.resize(nm);
                for (auto &a : hp) a.resize(nm);
                for (auto &a : ht) a.resize(nm);
                float *r0 = h[0].data(), *r1 = h[1].data(), *q0 = h[2].data(), *q1 = h[3].data();
                float *d0 = h[4].data(), *d1 = h[5].data(), *e0 = h[6].data(), *e1 = h[7].data();
                ck(sfwi_s1_hcg_init(p.allnx, p.allnz, pp_cg_old.data(), ps_cg_old.data(), r0, r1, q0, q1, d0, d1, e0,
                                    e1),
                   nullptr);
                for (int inner = 0; inner < max_inner; ++inner) {
                    for (size_t gi = 0; gi < sv.widths.size(); ++gi) {   // H p summed over every shot
                        sfwi_ctx *c = ctxs[gi];
                        const std::vector<sfwi_shot> grp = group_shots(sv, sv.widths[gi]);
                        ck(sfwi_batch_hessian_vector(c, (int)grp.size(), grp.data(), q0, q1), c);
                        ck(sfwi_flush(c, ht[0].data(), ht[1].data(), ht[2].data()), c);
                        for (int k = 0; k < 2; ++k)
                            for (size_t i = 0; i < nm; ++i) hp[k][i] = gi == 0 ? ht[k][i] : hp[k][i] + ht[k][i];
                    }
                    double rr0 = 0.0, rr1 = 0.0, ha = 0.0, hb = 0.0;
                    ck(sfwi_s1_hcg_step(p.allnx, p.allnz, 1, cut_layer, hp[0].data(), hp[1].data(), r0, r1, q0, q1, d0,
                                        d1, e0, e1, &rr0, &rr1, &ha, &hb),
                       nullptr);
                    if (FILE *f = std::fopen((outdir + "/misfit.txt").c_str(), "a")) {   // main.cpp:987-1037
                        if (inner == 0) std::fprintf(f, "i = %d,inner = %d,hessian_misfit = %5.3e\n", iter, inner, rr0);
                        std::fprintf(f, "i = %d,inner = %d,hessian_misfit = %5.3e\n", iter, inner + 1, rr1);
                        std::fclose(f);
                    }
                }
                ck(sfwi_s1_hcg_finish(p.allnx, p.allnz, d0, d1, e0, e1), nullptr);
                dir_pp = h[6];
                dir_ps = h[7];
            }
            double ls[3] = {0.0, 0.0, 0.0};
            for (size_t gi = 0; gi < sv.widths.size(); ++gi) {   // batched; the sums of per-shot calls in shot order
                const std::vector<sfwi_shot> grp = group_shots(sv, sv.widths[gi]);
                ck(sfwi_batch_linesearch_parabola(ctxs[gi], (int)grp.size(), grp.data(), dir_pp.data(), dir_ps.data(),
                                                  step1, step2, ls),
                   ctxs[gi]);
            }
            const double alpha = sfwi_s1_alpha(ls[0], ls[1], ls[2], step1, step2);
            ck(sfwi_s1_update(p.allnx, p.allnz, cut_layer, alpha, dir_pp.data(), dir_ps.data(), vp.data(), vs.data()),
               nullptr);
            std::printf("iter %d misfit %.17g beta %.17g alpha %.17g\n", iter, misfit, beta, alpha);
            std::fflush(stdout);
            if (FILE *f = std::fopen((outdir + "/misfit.txt").c_str(), "a")) {
                std::fprintf(f, "i = %d,misfit = %4.3e,rms = %f\n", iter, misfit, misfit / max_misfit);
                std::fprintf(f, "i = %d,alpha = %5.3e\n", iter, alpha);
                std::fclose(f);
            }
            write_model(out_name(sv, "vel_p", iter), vp);
            write_model(out_name(sv, "vel_s", iter), vs);
        }
    } catch (const Fail &e) {
        std::fprintf(stderr, "%s\n", e.what());
        for (sfwi_ctx *c : ctxs) sfwi_ctx_destroy(c);
        return e.status;
    }
    for (sfwi_ctx *c : ctxs) sfwi_ctx_destroy(c);
    return 0;
}
