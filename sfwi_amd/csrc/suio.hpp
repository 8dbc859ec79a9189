// Survey input of the reference drivers (host code, included by sfwi.hip):
// the 10-line positional parameter file, x-major FP32 model files and SU
// (240-byte SEG-Y trace header + FP32 samples, no reel headers) shot records.
//
//   parameter file (sfwi-stage1/main.cpp:1346-1355, read as a stream of
//   whitespace-separated fields like its fscanf calls; both stages use this
//   10-line layout: sfwi-stage2/main.cpp:946-954 reads one path fewer than its
//   shipped para_stage2 holds, so the shared layout is the one both shipped
//   files follow):
//     dx dz pml allnx allnz minshot maxshot
//     vp path / vs path / rho path / record-z path / record-x path / output dir
//     f0 maxiter iter_round1
//     sy gy
//     light rbc cpu_mem iointerval
//   models: allnx columns of allnz floats (sfwi-stage1/main.cpp:1468-1473).
//   SU shot index (sfwi-stage2/main.cpp:121-216, the stage-2 form that counts
//   the traces of each shot; the stage-1 copy never increments the count):
//   traces grouped by consecutive equal sx into rows
//   {shot number (1-based), ntr, sx, 0, min(sx, gx), 0, max(sx, gx), 0, first trace}.
//   gathers: trace pos .. pos+ntr-1, gc = gx / coordinate_scale
//   (sfwi-stage1/main.cpp:204-225).
// Header fields used, by byte offset: tracl 0, tracr 4, fldr 8, tracf 12,
// offset 36, scalco 70, sx 72, gx 80, ns 114, dt 116 (microseconds).
#pragma once

#include <fstream>
#include <sstream>

namespace {

constexpr int SU_HDR = 240;

template <typename T> T su_get(const unsigned char *h, int off) {
    T v;
    std::memcpy(&v, h + off, sizeof(T));
    return v;
}
template <typename T> void su_put(unsigned char *h, int off, T v) { std::memcpy(h + off, &v, sizeof(T)); }

int su_copy_path(char *dst, const std::string &src, const char *what) {
    if (src.size() >= SFWI_PATH_MAX)
        return fail(nullptr, SFWI_EINVAL, "parameter file: %s longer than %d characters", what, SFWI_PATH_MAX - 1);
    std::memcpy(dst, src.c_str(), src.size() + 1);
    return 0;
}

}  // namespace

extern "C" {

int sfwi_para_read(const char *path, sfwi_para *out) {
    if (!path || !out) return fail(nullptr, SFWI_EINVAL, "sfwi_para_read: NULL argument");
    std::ifstream f(path);
    if (!f) return fail(nullptr, SFWI_EINVAL, "cannot open parameter file %s", path);
    // a stream of whitespace-separated fields, as the drivers' fscanf reads it
    std::vector<std::string> tok;
    for (std::string t; f >> t;) tok.push_back(t);
    static const char *names[22] = {"dx", "dz", "pml", "allnx", "allnz", "minshot", "maxshot", "vp path", "vs path",
                                    "rho path", "record-z path", "record-x path", "output path", "f0", "maxiter",
                                    "iter_round1", "sy", "gy", "light", "rbc", "cpu_mem", "iointerval"};
    if (tok.size() < 22)
        return fail(nullptr, SFWI_EINVAL, "parameter file %s: %zu fields, the layout has 22 (missing: %s)", path,
                    tok.size(), names[tok.size()]);
    double v[22];
    for (int k = 0; k < 22; ++k) {
        if (k >= 7 && k < 13) continue;
        char *end = nullptr;
        v[k] = std::strtod(tok[k].c_str(), &end);
        if (!end || *end)
            return fail(nullptr, SFWI_EINVAL, "parameter file %s: %s is '%s', not a number", path, names[k],
                        tok[k].c_str());
    }
    sfwi_para p;
    std::memset(&p, 0, sizeof(p));
    p.dx = (float)v[0]; p.dz = (float)v[1]; p.pml = (int)v[2]; p.allnx = (int)v[3]; p.allnz = (int)v[4];
    p.minshot = (int)v[5]; p.maxshot = (int)v[6];
    char *paths[6] = {p.vp, p.vs, p.rho, p.rec_z, p.rec_x, p.outdir};
    int r;
    for (int k = 0; k < 6; ++k)
        if ((r = su_copy_path(paths[k], tok[7 + k], names[7 + k]))) return r;
    p.f0 = (float)v[13]; p.maxiter = (int)v[14]; p.iter_round1 = (int)v[15];
    p.sy = (int)v[16]; p.gy = (int)v[17];
    p.light = (int)v[18]; p.rbc = (int)v[19]; p.cpu_mem = (int)v[20]; p.iointerval = (int)v[21];
    if (!(p.dx > 0) || !(p.dz > 0) || p.pml < 0 || p.allnx < 1 || p.allnz < 1)
        return fail(nullptr, SFWI_EINVAL, "parameter file %s: bad geometry dx=%g dz=%g pml=%d allnx=%d allnz=%d", path,
                    p.dx, p.dz, p.pml, p.allnx, p.allnz);
    *out = p;
    return 0;
}

int sfwi_model_read(const char *path, int allnx, int allnz, float *out) {
    if (!path || !out || allnx < 1 || allnz < 1) return fail(nullptr, SFWI_EINVAL, "sfwi_model_read: bad argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(nullptr, SFWI_EINVAL, "cannot open model file %s", path);
    const size_t n = (size_t)allnx * allnz;
    f.read(reinterpret_cast<char *>(out), (std::streamsize)(n * sizeof(float)));
    if ((size_t)f.gcount() != n * sizeof(float))
        return fail(nullptr, SFWI_EINVAL, "model file %s holds %lld bytes, %d x %d floats need %zu", path,
                    (long long)f.gcount(), allnx, allnz, n * sizeof(float));
    return 0;
}

int sfwi_su_index(const char *path, int coordinate_scale, int max_shots, int *nt, float *dt, int *nshots,
                  int *table) {
    if (!path || !nt || !dt || !nshots || coordinate_scale == 0 || max_shots < 0 || (max_shots > 0 && !table))
        return fail(nullptr, SFWI_EINVAL, "sfwi_su_index: bad argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(nullptr, SFWI_EINVAL, "cannot open SU file %s", path);
    unsigned char h[SU_HDR];
    f.read(reinterpret_cast<char *>(h), SU_HDR);
    if (f.gcount() != SU_HDR) return fail(nullptr, SFWI_EINVAL, "SU file %s: no trace header", path);
    const int ns = su_get<unsigned short>(h, 114);
    if (ns < 1) return fail(nullptr, SFWI_EINVAL, "SU file %s: ns = 0 in the first trace header", path);
    *nt = ns;
    *dt = (float)(su_get<unsigned short>(h, 116) / 1000000.0);
    const std::streamoff tl = SU_HDR + (std::streamoff)ns * sizeof(float);
    f.seekg(0, std::ios::end);
    const std::streamoff size = f.tellg();
    if (size % tl) return fail(nullptr, SFWI_EINVAL, "SU file %s: %lld bytes is not a whole number of %lld-byte traces",
                               path, (long long)size, (long long)tl);
    const long long ntraces = size / tl;
    int n = 0, ntr = 0, sx0 = 0, xmin = 0, xmax = 0;
    long long first = 0;
    auto close_shot = [&]() {
        if (n > max_shots) return;
        int *row = table + 9 * (size_t)(n - 1);
        row[0] = n; row[1] = ntr; row[2] = sx0 / coordinate_scale; row[3] = 0;
        row[4] = xmin / coordinate_scale; row[5] = 0; row[6] = xmax / coordinate_scale; row[7] = 0;
        row[8] = (int)first;
    };
    for (long long t = 0; t < ntraces; ++t) {
        f.seekg(t * tl);
        f.read(reinterpret_cast<char *>(h), SU_HDR);
        if (f.gcount() != SU_HDR) return fail(nullptr, SFWI_EINVAL, "SU file %s: short read at trace %lld", path, t);
        if (su_get<unsigned short>(h, 114) != ns)
            return fail(nullptr, SFWI_EINVAL, "SU file %s: trace %lld has ns = %d, the first %d", path, t,
                        (int)su_get<unsigned short>(h, 114), ns);
        const int sx = su_get<int>(h, 72), gx = su_get<int>(h, 80);
        if (t == 0 || sx != sx0) {
            if (t > 0) close_shot();
            ++n;
            ntr = 0;
            sx0 = sx;
            first = t;
            xmin = 99999999;
            xmax = -999999;
        }
        ++ntr;
        xmin = std::min(xmin, std::min(sx, gx));
        xmax = std::max(xmax, std::max(sx, gx));
    }
    if (ntraces > 0) close_shot();
    *nshots = n;
    if (n > max_shots && table)
        return fail(nullptr, SFWI_EINVAL, "SU file %s: %d shots, the table holds %d", path, n, max_shots);
    return 0;
}

int sfwi_su_read_gather(const char *path, long long pos, int ntr, int nt, int coordinate_scale, float *data,
                        int *gc) {
    if (!path || !data || !gc || pos < 0 || ntr < 1 || nt < 1 || coordinate_scale == 0)
        return fail(nullptr, SFWI_EINVAL, "sfwi_su_read_gather: bad argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(nullptr, SFWI_EINVAL, "cannot open SU file %s", path);
    const std::streamoff tl = SU_HDR + (std::streamoff)nt * sizeof(float);
    f.seekg(pos * tl);
    unsigned char h[SU_HDR];
    for (int i = 0; i < ntr; ++i) {
        f.read(reinterpret_cast<char *>(h), SU_HDR);
        f.read(reinterpret_cast<char *>(data + (size_t)i * nt), (std::streamsize)(nt * sizeof(float)));
        if (!f) return fail(nullptr, SFWI_EINVAL, "SU file %s: traces %lld..%lld not all present", path, pos,
                            pos + ntr - 1);
        if (su_get<unsigned short>(h, 114) != nt)
            return fail(nullptr, SFWI_EINVAL, "SU file %s: trace %lld has ns = %d, expected %d", path, pos + i,
                        (int)su_get<unsigned short>(h, 114), nt);
        gc[i] = su_get<int>(h, 80) / coordinate_scale;
    }
    return 0;
}

int sfwi_su_write(const char *path, int append, int shot, int ntr, int nt, float dt, int sx, const int *gx,
                  const float *data) {
    if (!path || !gx || !data || ntr < 1 || nt < 1 || nt > 65535 || !(dt > 0))
        return fail(nullptr, SFWI_EINVAL, "sfwi_su_write: bad argument (ntr=%d nt=%d dt=%g)", ntr, nt, dt);
    const long dt_us = std::lround((double)dt * 1e6);
    if (dt_us < 1 || dt_us > 65535) return fail(nullptr, SFWI_EINVAL, "sfwi_su_write: dt %g s does not fit the header", dt);
    std::ofstream f(path, std::ios::binary | (append ? std::ios::app : std::ios::trunc));
    if (!f) return fail(nullptr, SFWI_EINVAL, "cannot open %s for writing", path);
    unsigned char h[SU_HDR];
    for (int i = 0; i < ntr; ++i) {
        std::memset(h, 0, SU_HDR);
        su_put<int>(h, 0, i + 1);        // tracl
        su_put<int>(h, 4, i + 1);        // tracr
        su_put<int>(h, 8, shot);         // fldr
        su_put<int>(h, 12, i + 1);       // tracf
        su_put<int>(h, 36, gx[i] - sx);  // offset
        su_put<short>(h, 70, 1);         // scalco
        su_put<int>(h, 72, sx);
        su_put<int>(h, 80, gx[i]);
        su_put<unsigned short>(h, 114, (unsigned short)nt);
        su_put<unsigned short>(h, 116, (unsigned short)dt_us);
        f.write(reinterpret_cast<const char *>(h), SU_HDR);
        f.write(reinterpret_cast<const char *>(data + (size_t)i * nt), (std::streamsize)(nt * sizeof(float)));
    }
    if (!f) return fail(nullptr, SFWI_EINVAL, "write to %s failed", path);
    return 0;
}

int sfwi_shot_from_table(const int row[9], float dx, int pml, int sy, int gy, int all_left, sfwi_shot *out,
                         int *nx_shot) {
    if (!row || !out || !(dx > 0)) return fail(nullptr, SFWI_EINVAL, "sfwi_shot_from_table: bad argument");
    // sfwi-stage1/main.cpp:580-625 (the worker's per-shot set-up)
    const int ngx_left = row[4], ngx_right = row[6], scx = row[2];
    const int nx_s = (int)((ngx_right - ngx_left) / dx) + 1;
    const int delta_left = (int)((ngx_left - all_left) / dx);
    out->left = delta_left;
    out->isx = pml + (int)((scx - ngx_left) / dx);
    out->isz = pml + sy;
    out->igz = pml + gy;
    out->ngx_left = ngx_left;
    out->ntr = row[1];
    out->obs_id = 0;
    if (nx_shot) *nx_shot = nx_s;
    return 0;
}

}  // extern "C"
