// Stage-1 inversion driver over the C ABI: the reference's stage-1 program
// (sfwi-stage1/main.cpp) for one GPU, without MPI.  It reads the reference's
// parameter file, x-major FP32 models and SU records, indexes the shots,
// derives each shot's window the way the worker does, and runs the stage-1
// iterations (batched gradients, preconditioned CG direction, parabolic line
// search along it, clamped step, model update with the top-row reset).
//
//   sfwi_stage1 <para file> [niter [max_batch [cut_layer]]]
//
// niter defaults to the file's maxiter, cut_layer to the reference's 26; shots whose number lies outside
// [minshot, maxshot] are skipped.  The density is forced to 2000 as the
// reference does (sfwi-stage1/main.cpp:1476-1479).  Where the reference sizes
// its one context by shot 0's aperture and runs other apertures on it
// (main.cpp:610-617), every aperture width here gets a context of its own;
// shots of a width are solved in batches of max_batch (default 64).  With
// cpu_mem = 1 in the file the boundary strips live in pinned host memory in
// chunks of iointerval steps (the reference's ntt), as the reference does.
// Per iteration: one line "iter <k> misfit <m> beta <b> alpha <a>" on stdout,
// and, as the reference writes them, <outdir>/misfit.txt lines and
// <outdir>/vel_p_<k>_<allnx>_<allnz>.dat, vel_s_<k>_<allnx>_<allnz>.dat.
// Exit status: 0 success, 2 input error, 3 library error.
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "sfwi_abi.h"

namespace {

struct Fail : std::runtime_error {
    int status;
    Fail(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

void ck(int rc, const sfwi_ctx *ctx) {
    if (rc != SFWI_OK) throw Fail(3, std::string("sfwi: ") + sfwi_last_error(ctx));
}
void ck_input(int rc) {
    if (rc != SFWI_OK) throw Fail(2, std::string("input error: ") + sfwi_last_error(nullptr));
}

struct ShotIn {
    sfwi_shot s;
    std::vector<int> gc;
    std::vector<float> rz, rx;
    int width;
};

void write_model(const std::string &path, const std::vector<float> &a) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(a.data(), sizeof(float), a.size(), f) != a.size()) {
        if (f) std::fclose(f);
        throw Fail(2, "cannot write " + path);
    }
    std::fclose(f);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2 || argc > 5) {
        std::fprintf(stderr, "usage: %s para_file [niter [max_batch [cut_layer]]]\n", argv[0]);
        return 2;
    }
    std::vector<sfwi_ctx *> ctxs;
    try {
        sfwi_para p;
        ck_input(sfwi_para_read(argv[1], &p));
        const int niter = argc >= 3 ? std::atoi(argv[2]) : p.maxiter;
        const int max_batch = argc >= 4 ? std::atoi(argv[3]) : 64;
        const int cut_layer = argc >= 5 ? std::atoi(argv[4]) : 26;   // sfwi-stage1/main.cpp:410
        if (niter < 0 || max_batch < 1 || cut_layer < 0)
            throw Fail(2, "input error: niter must be >= 0, max_batch >= 1, cut_layer >= 0");
        if (p.rbc) throw Fail(2, "input error: rbc (random boundary) is not supported; strips are used");
        if (p.cpu_mem && p.iointerval < 1) throw Fail(2, "input error: cpu_mem needs iointerval >= 1 (steps per chunk)");
        const size_t nm = (size_t)p.allnx * p.allnz;
        std::vector<float> vp(nm), vs(nm), rho(nm, 2000.0f);
        ck_input(sfwi_model_read(p.vp, p.allnx, p.allnz, vp.data()));
        ck_input(sfwi_model_read(p.vs, p.allnx, p.allnz, vs.data()));
        // shot index of the record files
        int nt = 0, ns = 0, nt_x = 0, ns_x = 0;
        float dt = 0, dt_x = 0;
        ck_input(sfwi_su_index(p.rec_z, 1, 0, &nt, &dt, &ns, nullptr));
        std::vector<int> table(9 * (size_t)ns), table_x(9 * (size_t)ns);
        ck_input(sfwi_su_index(p.rec_z, 1, ns, &nt, &dt, &ns, table.data()));
        ck_input(sfwi_su_index(p.rec_x, 1, ns, &nt_x, &dt_x, &ns_x, table_x.data()));
        if (nt_x != nt || ns_x != ns || table_x != table) throw Fail(2, "input error: record files index differently");
        int all_left = 1 << 30;
        for (int k = 0; k < ns; ++k) all_left = std::min(all_left, table[9 * k + 4]);
        std::vector<ShotIn> shots;
        for (int k = 0; k < ns; ++k) {
            const int *row = &table[9 * (size_t)k];
            if (row[0] < p.minshot || row[0] > p.maxshot) continue;
            ShotIn in;
            ck_input(sfwi_shot_from_table(row, p.dx, p.pml, p.sy, p.gy, all_left, &in.s, &in.width));
            in.gc.resize(in.s.ntr);
            in.rz.resize((size_t)in.s.ntr * nt);
            in.rx.resize((size_t)in.s.ntr * nt);
            std::vector<int> gx(in.s.ntr);
            ck_input(sfwi_su_read_gather(p.rec_z, row[8], in.s.ntr, nt, 1, in.rz.data(), in.gc.data()));
            ck_input(sfwi_su_read_gather(p.rec_x, row[8], in.s.ntr, nt, 1, in.rx.data(), gx.data()));
            if (gx != in.gc) throw Fail(2, "input error: receiver coordinates differ between the record files");
            shots.push_back(std::move(in));
        }
        if (shots.empty()) throw Fail(2, "input error: no shot between minshot and maxshot");
        for (ShotIn &in : shots) {
            in.s.gc = in.gc.data();
            in.s.obs_z = in.rz.data();
            in.s.obs_x = in.rx.data();
        }
        // one context per aperture width, in order of first appearance
        std::vector<int> widths;
        std::map<int, std::vector<int>> members;
        int ntr_max = 1;
        for (size_t k = 0; k < shots.size(); ++k) {
            if (!members.count(shots[k].width)) widths.push_back(shots[k].width);
            members[shots[k].width].push_back((int)k);
            ntr_max = std::max(ntr_max, shots[k].s.ntr);
        }
        std::vector<float> wav(nt);
        sfwi_ricker(wav.data(), nt, dt, p.f0);
        for (int w : widths) {
            sfwi_params q = {};
            q.dx = p.dx; q.dz = p.dz; q.dt = dt; q.nx = w; q.nz = p.allnz; q.pml = p.pml; q.nt = nt; q.nop = 4;
            q.ntr = ntr_max; q.allnx = p.allnx; q.allnz = p.allnz; q.rr = 0.001f; q.adjoint_dir = 1;
            q.max_batch = std::min<int>(max_batch, (int)members[w].size());
            q.cpu_mem = p.cpu_mem ? 1 : 0;      // strips in pinned host memory, iointerval steps per chunk
            q.strip_chunk = p.cpu_mem ? p.iointerval : 0;
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
            for (size_t gi = 0; gi < widths.size(); ++gi) {
                sfwi_ctx *c = ctxs[gi];
                const std::vector<int> &mem = members[widths[gi]];
                ck(sfwi_set_model(c, vp.data(), vs.data(), rho.data()), c);
                const int nb = std::min<int>(max_batch, (int)mem.size());
                for (size_t b0 = 0; b0 < mem.size(); b0 += nb) {
                    std::vector<sfwi_shot> batch;
                    for (size_t k = b0; k < std::min(mem.size(), b0 + nb); ++k) batch.push_back(shots[mem[k]].s);
                    std::vector<double> m(batch.size(), 0.0);
                    ck(sfwi_batch_gradient(c, (int)batch.size(), batch.data(), m.data()), c);
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
            double ls[3] = {0.0, 0.0, 0.0};
            for (size_t gi = 0; gi < widths.size(); ++gi) {   // batched; the sums of per-shot calls in shot order
                std::vector<sfwi_shot> grp;
                for (int k : members[widths[gi]]) grp.push_back(shots[k].s);
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
            char name[256];
            std::snprintf(name, sizeof(name), "/vel_p_%d_%d_%d.dat", iter, p.allnx, p.allnz);
            write_model(outdir + name, vp);
            std::snprintf(name, sizeof(name), "/vel_s_%d_%d_%d.dat", iter, p.allnx, p.allnz);
            write_model(outdir + name, vs);
        }
    } catch (const Fail &e) {
        std::fprintf(stderr, "%s\n", e.what());
        for (sfwi_ctx *c : ctxs) sfwi_ctx_destroy(c);
        return e.status;
    }
    for (sfwi_ctx *c : ctxs) sfwi_ctx_destroy(c);
    return 0;
}
