// Survey input shared by the host programs: the reference drivers' parameter
// file, the x-major FP32 Vp and Vs models and SU records read through the C ABI, the shot
// index (sfwi-stage2/main.cpp:121-216), the shot windows derived from it the
// way the workers do (sfwi-stage1/main.cpp:580-625), shots outside
// [minshot, maxshot] skipped, and the shots grouped by aperture width.
#pragma once

#include <sys/stat.h>

#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "sfwi_abi.h"

namespace sfwi_host {

// Exit status carried with the message: 2 input error, 3 library error.
struct Fail : std::runtime_error {
    int status;
    Fail(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void ck(int rc, const sfwi_ctx *ctx) {
    if (rc != SFWI_OK) throw Fail(3, std::string("sfwi: ") + sfwi_last_error(ctx));
}
inline void ck_input(int rc) {
    if (rc != SFWI_OK) throw Fail(2, std::string("input error: ") + sfwi_last_error(nullptr));
}

struct ShotIn {
    sfwi_shot s;
    std::vector<int> gc;
    std::vector<float> rz, rx;
    int width;
};

struct Survey {
    sfwi_para p;
    int nt = 0;
    float dt = 0;
    std::vector<ShotIn> shots;               // shot-number order of the records, in [minshot, maxshot]
    std::vector<int> widths;                 // aperture widths in order of first appearance
    std::map<int, std::vector<int>> members; // shots of each width
    int ntr_max = 1;
    std::vector<float> vp, vs;               // the file's models, allnx x allnz (x-major)
};

inline Survey read_survey(const char *para) {
    Survey sv;
    sfwi_para &p = sv.p;
    ck_input(sfwi_para_read(para, &p));
    if (p.cpu_mem && p.iointerval < 1) throw Fail(2, "input error: cpu_mem needs iointerval >= 1 (steps per chunk)");
    const size_t nm = (size_t)p.allnx * p.allnz;
    sv.vp.resize(nm);
    sv.vs.resize(nm);
    ck_input(sfwi_model_read(p.vp, p.allnx, p.allnz, sv.vp.data()));
    ck_input(sfwi_model_read(p.vs, p.allnx, p.allnz, sv.vs.data()));
    int ns = 0, nt_x = 0, ns_x = 0;
    float dt_x = 0;
    ck_input(sfwi_su_index(p.rec_z, 1, 0, &sv.nt, &sv.dt, &ns, nullptr));
    std::vector<int> table(9 * (size_t)ns), table_x(9 * (size_t)ns);
    ck_input(sfwi_su_index(p.rec_z, 1, ns, &sv.nt, &sv.dt, &ns, table.data()));
    ck_input(sfwi_su_index(p.rec_x, 1, ns, &nt_x, &dt_x, &ns_x, table_x.data()));
    if (nt_x != sv.nt || ns_x != ns || table_x != table) throw Fail(2, "input error: record files index differently");
    int all_left = 1 << 30;
    for (int k = 0; k < ns; ++k) all_left = std::min(all_left, table[9 * k + 4]);
    for (int k = 0; k < ns; ++k) {
        const int *row = &table[9 * (size_t)k];
        if (row[0] < p.minshot || row[0] > p.maxshot) continue;
        ShotIn in;
        ck_input(sfwi_shot_from_table(row, p.dx, p.pml, p.sy, p.gy, all_left, &in.s, &in.width));
        in.gc.resize(in.s.ntr);
        in.rz.resize((size_t)in.s.ntr * sv.nt);
        in.rx.resize((size_t)in.s.ntr * sv.nt);
        std::vector<int> gx(in.s.ntr);
        ck_input(sfwi_su_read_gather(p.rec_z, row[8], in.s.ntr, sv.nt, 1, in.rz.data(), in.gc.data()));
        ck_input(sfwi_su_read_gather(p.rec_x, row[8], in.s.ntr, sv.nt, 1, in.rx.data(), gx.data()));
        if (gx != in.gc) throw Fail(2, "input error: receiver coordinates differ between the record files");
        sv.shots.push_back(std::move(in));
    }
    if (sv.shots.empty()) throw Fail(2, "input error: no shot between minshot and maxshot");
    for (size_t k = 0; k < sv.shots.size(); ++k) {
        ShotIn &in = sv.shots[k];
        in.s.gc = in.gc.data();
        in.s.obs_z = in.rz.data();
        in.s.obs_x = in.rx.data();
        if (!sv.members.count(in.width)) sv.widths.push_back(in.width);
        sv.members[in.width].push_back((int)k);
        sv.ntr_max = std::max(sv.ntr_max, in.s.ntr);
    }
    return sv;
}

// Context parameters of one aperture width: the FD2DGPU_ELASTIC arguments the
// workers pass (sfwi-stage1/main.cpp:534, sfwi-stage2/main.cpp:497), batches of up to max_batch
// shots, cpu_mem / iointerval from the file.
inline sfwi_params width_params(const Survey &sv, int width, int max_batch) {
    const sfwi_para &p = sv.p;
    sfwi_params q = {};
    q.dx = p.dx; q.dz = p.dz; q.dt = sv.dt; q.nx = width; q.nz = p.allnz; q.pml = p.pml; q.nt = sv.nt; q.nop = 4;
    q.ntr = sv.ntr_max; q.allnx = p.allnx; q.allnz = p.allnz; q.rr = 0.001f; q.adjoint_dir = 1;
    q.max_batch = std::min<int>(max_batch, (int)sv.members.at(width).size());
    q.cpu_mem = p.cpu_mem ? 1 : 0;      // strips in pinned host memory, iointerval steps per chunk
    q.strip_chunk = p.cpu_mem ? p.iointerval : 0;
    q.rbc = p.rbc ? 1 : 0;              // random boundary, no strips (cpu_mem then holds nothing)
    q.rbc_seed = 1;                     // the reference seeds from the clock; a run here repeats
    return q;
}

// The shots of one width, in order.
inline std::vector<sfwi_shot> group_shots(const Survey &sv, int width) {
    std::vector<sfwi_shot> g;
    for (int k : sv.members.at(width)) g.push_back(sv.shots[k].s);
    return g;
}

inline void write_model(const std::string &path, const std::vector<float> &a) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(a.data(), sizeof(float), a.size(), f) != a.size()) {
        if (f) std::fclose(f);
        throw Fail(2, "cannot write " + path);
    }
    std::fclose(f);
}

// <outdir>/<name>_<iter>_<allnx>_<allnz>.dat, the reference's output naming.
inline std::string out_name(const Survey &sv, const char *name, int iter) {
    char buf[256];
    std::snprintf(buf, sizeof(buf), "/%s_%d_%d_%d.dat", name, iter, sv.p.allnx, sv.p.allnz);
    return std::string(sv.p.outdir) + buf;
}

}  // namespace sfwi_host
