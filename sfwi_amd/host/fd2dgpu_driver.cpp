// Host driver over the FD2DGPU_ELASTIC adapter: the per-shot part of the
// reference's worker loop (sfwi-stage1/main.cpp:534-635) without MPI and SU
// parsing.  It reads one survey from a flat binary file, runs one stage-1
// gradient over all shots (optionally the parabolic line search) through the
// same call sequence the reference driver uses, and writes the flushed
// gradients.  With niter it runs that many complete stage-1 iterations the
// way the reference's master rank does (sfwi-stage1/main.cpp:534-1278:
// gradient of every shot, preconditioned CG direction, line search along it,
// clamped parabolic step, model update with the top-row reset) and writes the
// final model.
//
//   fd2dgpu_driver <survey.bin> <out_prefix> [step1 step2 [niter cut_layer [ntt]]]
//
// ntt > 0 passes cpu_mem = true with chunks of ntt steps to the per-shot
// gradient calls, as the reference driver passes its cpu_mem and iointerval
// (boundary strips in host memory).
//
// survey.bin (little endian): int32 header
//   nx nz allnx allnz pml nt ntr_max nshots mode misfit
// float32 dx dz dt, then float32 arrays vp, vs, rho (allnx*allnz, x-major),
// wavelet (nt); per shot: int32 left isx isz igz ngx_left ntr, int32 gc[ntr],
// float32 rec_z[ntr*nt], rec_x[ntr*nt].
// Outputs <out_prefix>pp_grad.bin, ps_grad.bin, illum.bin and one line
// "misfit <value>" (plus "linesearch <m0> <m1> <m2>" with steps) on stdout;
// with niter, one line "iter <k> misfit <m> beta <b> alpha <a>" per iteration
// and <out_prefix>vel_p.bin, vel_s.bin.
// Exit status: 0 success, 2 usage or input error, 3 library error.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "sfwi_fd2dgpu.hpp"

namespace {

struct Reader {
    std::ifstream f;
    explicit Reader(const char *path) : f(path, std::ios::binary) {}
    template <typename T> void read(T *dst, size_t n) {
        f.read(reinterpret_cast<char *>(dst), (std::streamsize)(n * sizeof(T)));
        if (!f) throw std::runtime_error("survey file truncated");
    }
    template <typename T> T one() {
        T v;
        read(&v, 1);
        return v;
    }
};

struct ShotData {
    int left, isx, isz, igz, ngx_left, ntr;
    std::vector<int> gc;
    std::vector<float> rec_z, rec_x;
};

void write_bin(const std::string &path, const std::vector<float> &a) {
    std::ofstream o(path, std::ios::binary);
    o.write(reinterpret_cast<const char *>(a.data()), (std::streamsize)(a.size() * sizeof(float)));
    if (!o) throw std::runtime_error("cannot write " + path);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3 && argc != 5 && argc != 7 && argc != 8) {
        std::fprintf(stderr, "usage: %s survey.bin out_prefix [step1 step2 [niter cut_layer [ntt]]]\n", argv[0]);
        return 2;
    }
    const int niter = argc >= 7 ? std::atoi(argv[5]) : 0, cut_layer = argc >= 7 ? std::atoi(argv[6]) : 0;
    const int ntt = argc == 8 ? std::atoi(argv[7]) : 0;
    if (argc >= 7 && (niter < 1 || cut_layer < 0 || ntt < 0)) {
        std::fprintf(stderr, "input error: niter must be >= 1, cut_layer >= 0 and ntt >= 0\n");
        return 2;
    }
    int nx, nz, allnx, allnz, pml, nt, ntr_max, nshots, mode, misfit_kind;
    float dx, dz, dt;
    std::vector<float> vp, vs, rho, wav;
    std::vector<ShotData> shots;
    try {
        Reader r(argv[1]);
        if (!r.f) throw std::runtime_error(std::string("cannot open ") + argv[1]);
        int h[10];
        r.read(h, 10);
        nx = h[0]; nz = h[1]; allnx = h[2]; allnz = h[3]; pml = h[4]; nt = h[5];
        ntr_max = h[6]; nshots = h[7]; mode = h[8]; misfit_kind = h[9];
        dx = r.one<float>(); dz = r.one<float>(); dt = r.one<float>();
        if (nx < 1 || nz < 1 || allnx < nx || allnz < nz || nt < 1 || ntr_max < 1 || nshots < 0)
            throw std::runtime_error("bad survey header");
        const size_t nm = (size_t)allnx * allnz;
        vp.resize(nm); vs.resize(nm); rho.resize(nm); wav.resize(nt);
        r.read(vp.data(), nm); r.read(vs.data(), nm); r.read(rho.data(), nm); r.read(wav.data(), nt);
        for (int k = 0; k < nshots; ++k) {
            ShotData s;
            int g[6];
            r.read(g, 6);
            s.left = g[0]; s.isx = g[1]; s.isz = g[2]; s.igz = g[3]; s.ngx_left = g[4]; s.ntr = g[5];
            if (s.ntr < 1 || s.ntr > ntr_max) throw std::runtime_error("bad trace count");
            s.gc.resize(s.ntr);
            s.rec_z.resize((size_t)s.ntr * nt);
            s.rec_x.resize((size_t)s.ntr * nt);
            r.read(s.gc.data(), s.gc.size());
            r.read(s.rec_z.data(), s.rec_z.size());
            r.read(s.rec_x.data(), s.rec_x.size());
            shots.push_back(std::move(s));
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "input error: %s\n", e.what());
        return 2;
    }

    try {
        const int nxpml = nx + 2 * pml, nzpml = nz + 2 * pml;
        sfwi_ref::FD2DGPU_ELASTIC image2delastic(wav.data(), dx, dz, dt, nxpml, nzpml, allnx, allnz, 1, pml, nt, 4,
                                                 ntr_max, 1, misfit_kind, mode);
        const size_t nm = (size_t)allnx * allnz;
        std::vector<float> pp_grad(nm), ps_grad(nm), illum(nm), b1(nm), b2(nm), b3(nm), b4(nm);
        // one gradient of every shot on the model vp, vs (the worker loop)
        auto gradient = [&](std::vector<float> &mp, std::vector<float> &ms) {
            image2delastic.GPUbufferVPVS(mp.data(), ms.data(), rho.data());
            double misfit = 0.0;
            float misfit_ot = 0.0f;
            for (int k = 0; k < nshots; ++k) {
                ShotData &s = shots[k];
                image2delastic.ntr = s.ntr;
                image2delastic.record_copytoGPU(s.rec_z.data(), s.rec_x.data(), s.gc.data());
                image2delastic.bufferVpVsHtoD(s.left);
                image2delastic.delta_left = s.left;
                image2delastic.record_left_in_v = 0;
                image2delastic.isx = s.isx;
                image2delastic.isz = s.isz;
                image2delastic.igz = s.igz;
                image2delastic.cmin = (float)s.left;
                image2delastic.ngx_left = s.ngx_left;
                sfwi_ref::IMAGE_SINGLESHOT(image2delastic, k, nxpml * nzpml, 0, s.rec_z.data(), s.rec_x.data(),
                                           s.gc.data(), false, ntt > 0, ntt, 0, 0, 1, nullptr, nullptr, nullptr,
                                           nullptr, nullptr, misfit, &misfit_ot);
            }
            image2delastic.imagebuffer_resettozero(pp_grad.data(), ps_grad.data(), illum.data(), b1.data(),
                                                   b2.data(), b3.data(), b4.data());
            return misfit;
        };
        // the parabolic line search of every shot along (dir_pp, dir_ps)
        auto linesearch = [&](float *dir_pp, float *dir_ps, float step1, float step2, double m[3]) {
            for (int k = 0; k < nshots; ++k) {
                ShotData &s = shots[k];
                image2delastic.ntr = s.ntr;
                image2delastic.record_copytoGPU(s.rec_z.data(), s.rec_x.data(), s.gc.data());
                image2delastic.bufferVpVsHtoD(s.left);
                image2delastic.isx = s.isx;
                image2delastic.isz = s.isz;
                image2delastic.igz = s.igz;
                image2delastic.ngx_left = s.ngx_left;
                sfwi_ref::CALCULATE_ALPHA_parabola_mb(image2delastic, k, nxpml * nzpml, 0, s.rec_z.data(),
                                                      s.rec_x.data(), s.gc.data(), false, false, 0, 0, nullptr,
                                                      nullptr, dir_pp, dir_ps, m[0], m[1], m[2], step1, step2,
                                                      nullptr, nullptr, nullptr);
            }
        };
        const std::string pre = argv[2];
        if (niter > 0) {   // stage-1 iterations (sfwi-stage1/main.cpp:534-1278, master-rank steps)
            const float step1 = std::strtof(argv[3], nullptr), step2 = std::strtof(argv[4], nullptr);
            std::vector<float> vel_p = vp, vel_s = vs, pp_cg_old(nm, 0.0f), ps_cg_old(nm, 0.0f), dir_pp(nm),
                dir_ps(nm);
            double beta_fenmu = 0.0;
            for (int iter = 0; iter < niter; ++iter) {
                const double misfit = gradient(vel_p, vel_s);
                double beta = 0.0;
                sfwi_ref::check(sfwi_s1_direction(allnx, allnz, iter, cut_layer, pp_grad.data(), ps_grad.data(),
                                                  illum.data(), pp_cg_old.data(), ps_cg_old.data(), &beta_fenmu,
                                                  dir_pp.data(), dir_ps.data(), &beta),
                                nullptr);
                double m[3] = {0.0, 0.0, 0.0};
                linesearch(dir_pp.data(), dir_ps.data(), step1, step2, m);
                const double alpha = sfwi_s1_alpha(m[0], m[1], m[2], step1, step2);
                sfwi_ref::check(sfwi_s1_update(allnx, allnz, cut_layer, alpha, dir_pp.data(), dir_ps.data(),
                                               vel_p.data(), vel_s.data()),
                                nullptr);
                std::printf("iter %d misfit %.17g beta %.17g alpha %.17g\n", iter, misfit, beta, alpha);
            }
            write_bin(pre + "vel_p.bin", vel_p);
            write_bin(pre + "vel_s.bin", vel_s);
            return 0;
        }
        const double misfit = gradient(vp, vs);
        write_bin(pre + "pp_grad.bin", pp_grad);
        write_bin(pre + "ps_grad.bin", ps_grad);
        write_bin(pre + "illum.bin", illum);
        std::printf("misfit %.17g\n", misfit);
        if (argc == 5) {
            const float step1 = std::strtof(argv[3], nullptr), step2 = std::strtof(argv[4], nullptr);
            double m[3] = {0.0, 0.0, 0.0};
            linesearch(pp_grad.data(), ps_grad.data(), step1, step2, m);
            std::printf("linesearch %.17g %.17g %.17g\n", m[0], m[1], m[2]);
        }
    } catch (const sfwi_ref::Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 3;
    }
    return 0;
}
