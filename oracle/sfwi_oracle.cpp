// CPU oracle for SFWI's per-shot elastic modeling + gradient path.
//
// TEST INFRASTRUCTURE ONLY: this file is the test oracle of the HIP library in
// sfwi_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it; the product never does.
//
// It restates the reference's live per-shot algorithm, cell by cell, in one
// template instantiated twice:
//   T = float  : the reference's FP32 operation order.  Where the reference's
//                expression contains a double literal (1.0/rho, 5.0*vp, ...),
//                the same promotion to FP64 happens here, because the code is
//                written with the same operand types and built with
//                -ffp-contract=off.
//   T = double : every array and every operation in FP64.
//
// Citations are file:line in the reference repository (285639863/SFWI).
// The layout is the reference's: padded arrays nxpml*nzpml, index
// ix*nzpml+iz (z fastest); model arrays allnx*allnz, index ix*allnz+iz.

#include "sfwi_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int NOP = 4;  // stencil half width and strip half width: sfwi-stage1/main.cpp:1495

// 8th-order staggered coefficients: sfwi-stage1/FD2DGPU.cuh:16 (double
// fractions stored into a float constant array).
template <typename T> struct FD;
template <> struct FD<float> {
    static constexpr float c[5] = {0.0f, (float)(1225.0 / 1024.0), (float)(-245.0 / 3072.0),
                                   (float)(49.0 / 5120.0), (float)(-5.0 / 7168.0)};
};
template <> struct FD<double> {
    static constexpr double c[5] = {0.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0,
                                    -5.0 / 7168.0};
};

// Difference centred half a cell ahead: taps i-3..i+4 (stride s).
//   sum_k c_k (f[i+k] - f[i-k+1])        e.g. sfwi-stage1/born.cu:376 (txx along x)
template <typename T> inline T d_plus(const T *f, long i, long s) {
    T t = 0;
    for (int k = 1; k <= NOP; ++k) t += FD<T>::c[k] * (f[i + k * s] - f[i - (k - 1) * s]);
    return t;
}
// Difference centred half a cell behind: taps i-4..i+3.
//   sum_k c_k (f[i+k-1] - f[i-k])        e.g. sfwi-stage1/born.cu:377 (txz along z)
template <typename T> inline T d_minus(const T *f, long i, long s) {
    T t = 0;
    for (int k = 1; k <= NOP; ++k) t += FD<T>::c[k] * (f[i + (k - 1) * s] - f[i - k * s]);
    return t;
}
// The same stencils applied to m*f evaluated at every tap (the adjoint
// velocity update keeps the moduli inside the derivative,
// sfwi-stage1/born.cu:541-548).
template <typename T> inline T d_plus_w(const T *m, const T *f, long i, long s) {
    T t = 0;
    for (int k = 1; k <= NOP; ++k)
        t += FD<T>::c[k] * ((m[i + k * s]) * f[i + k * s] - (m[i - (k - 1) * s]) * f[i - (k - 1) * s]);
    return t;
}
template <typename T> inline T d_minus_w(const T *m, const T *f, long i, long s) {
    T t = 0;
    for (int k = 1; k <= NOP; ++k)
        t += FD<T>::c[k] * (m[i + (k - 1) * s] * f[i + (k - 1) * s] - m[i - k * s] * f[i - k * s]);
    return t;
}
// (lambda + 2 mu) formed per tap inside the sum, as sfwi-stage1/born.cu:543,546.
template <typename T> inline T d_plus_p(const T *lam, const T *mu, const T *f, long i, long s) {
    T t = 0;
    for (int k = 1; k <= NOP; ++k) {
        long a = i + k * s, b = i - (k - 1) * s;
        t += FD<T>::c[k] * ((lam[a] + 2 * mu[a]) * f[a] - (lam[b] + 2 * mu[b]) * f[b]);
    }
    return t;
}

struct Grid {
    int nx, nz, pml, nxp, nzp, nt, ntr;
    long n;
};

Grid make_grid(const or_config *c) {
    Grid g;
    g.nx = c->nx;
    g.nz = c->nz;
    g.pml = c->pml;
    g.nxp = c->nx + 2 * c->pml;
    g.nzp = c->nz + 2 * c->pml;
    g.nt = c->nt;
    g.ntr = c->ntr;
    g.n = (long)g.nxp * g.nzp;
    return g;
}

// The five wavefield totals and their ten split parts of one propagator.
template <typename T> struct Wave {
    std::vector<T> vx, vz, txx, tzz, txz, vx_x, vx_z, vz_x, vz_z, txx_x, txx_z, tzz_x, tzz_z, txz_x,
        txz_z;
    explicit Wave(long n) {
        for (std::vector<T> *a : all()) a->assign(n, T(0));
    }
    std::vector<std::vector<T> *> all() {
        return {&vx, &vz, &txx, &tzz, &txz, &vx_x, &vx_z, &vz_x, &vz_z,
                &txx_x, &txx_z, &tzz_x, &tzz_z, &txz_x, &txz_z};
    }
};

template <typename T> struct Medium {
    std::vector<T> vp, vs, rho, lam, mu, dampx, dampz, dvp, dvs;
};

// Scalars the kernels take by value (float in the reference).
template <typename T> struct Steps {
    T dx, dz, dt;
};

int nthreads_of(const or_config *c) {
#ifdef _OPENMP
    return c->nthreads > 0 ? c->nthreads : omp_get_max_threads();
#else
    (void)c;
    return 1;
#endif
}

// ---------------------------------------------------------------- set-up ---

// Window of the full model copied into the padded grid, PML filled by
// clamping to the edge: z first, then x (corners come from the x pass).
// sfwi-stage1/FD2DGPU.cu:112-128 (copyvelocityHtoD), :233-240 (copypmlVz),
// :242-249 (copypmlVx), sequence :1348-1364 (bufferVpVsHtoD).
template <typename T>
void window_pad(const Grid &g, int allnz, int left, const float *all, std::vector<T> &out) {
    out.assign(g.n, T(0));
    for (int ix = 0; ix < g.nx; ++ix)
        for (int iz = 0; iz < g.nz; ++iz)
            out[(long)(ix + g.pml) * g.nzp + iz + g.pml] = (T)all[(long)(left + ix) * allnz + iz];
    for (int ix = 0; ix < g.nxp; ++ix)
        for (int iz = 0; iz < g.pml; ++iz) {
            out[(long)ix * g.nzp + iz] = out[(long)ix * g.nzp + g.pml];
            out[(long)ix * g.nzp + g.nzp - g.pml + iz] = out[(long)ix * g.nzp + g.nzp - g.pml - 1];
        }
    for (int ix = 0; ix < g.pml; ++ix)
        for (int iz = 0; iz < g.nzp; ++iz) {
            out[(long)ix * g.nzp + iz] = out[(long)g.pml * g.nzp + iz];
            out[(long)(g.nxp - g.pml + ix) * g.nzp + iz] = out[(long)(g.nxp - g.pml - 1) * g.nzp + iz];
        }
}

// PML profile value, sfwi-stage1/FD2DGPU.cu:460-465 (dampX) and :484-489
// (dampZ):  log10(1/rr) * (5.0*vp/(2.0*pml)) * powf(1.0*d/pml, 4.0).
// FP32 order: log10 of the float 1/rr rounded to float (CUDA log10f
// overload), the 5.0/2.0 factor in double, powf's argument rounded to float
// and x^4 taken as (x^2)^2 in double then rounded to float; product in double,
// stored as float.
inline float pml_profile_f32(float lg, float vp, int dist, int pml) {
    float t = (float)(1.0 * dist / pml);
    double t2 = (double)t * (double)t;
    float p4 = (float)(t2 * t2);
    return (float)((double)lg * (5.0 * (double)vp / (2.0 * pml)) * (double)p4);
}
inline double pml_profile_f64(double lg, double vp, int dist, int pml) {
    double t = 1.0 * dist / pml;
    return lg * (5.0 * vp / (2.0 * pml)) * (t * t * t * t);
}

template <typename T>
void damping(const Grid &g, float rr, bool off, const std::vector<T> &vp, std::vector<T> &dampx,
             std::vector<T> &dampz) {
    // Interior stays 0: d_damp and d_dampz are memset once in the constructor
    // (sfwi-stage1/FD2DGPU.cu:44, FD2DGPU.cu ctor of FD2DGPU_ELASTIC) and the
    // profile kernels only write PML cells.
    dampx.assign(g.n, T(0));
    dampz.assign(g.n, T(0));
    if (off) return;
    const bool f32 = sizeof(T) == sizeof(float);
    const float lg32 = (float)std::log10((double)(1.0f / rr));
    const double lg64 = std::log10(1.0 / (double)rr);
    for (int ix = 0; ix < g.nxp; ++ix)
        for (int iz = 0; iz < g.nzp; ++iz) {
            long i = (long)ix * g.nzp + iz;
            int dxl = -1, dzl = -1;
            if (ix < g.pml) dxl = g.pml - ix;
            if (ix > g.nxp - g.pml) dxl = ix - (g.nxp - g.pml);   // strict >, Q3
            if (iz < g.pml) dzl = g.pml - iz;
            if (iz > g.nzp - g.pml) dzl = iz - (g.nzp - g.pml);
            if (dxl >= 0)
                dampx[i] = f32 ? (T)pml_profile_f32(lg32, (float)vp[i], dxl, g.pml)
                               : (T)pml_profile_f64(lg64, (double)vp[i], dxl, g.pml);
            if (dzl >= 0)
                dampz[i] = f32 ? (T)pml_profile_f32(lg32, (float)vp[i], dzl, g.pml)
                               : (T)pml_profile_f64(lg64, (double)vp[i], dzl, g.pml);
        }
}

// Random boundary, the scheme of random_vel / set_random_boundary
// (sfwi-stage1/FD2DGPU.cu:162-220, 3245-3290) with a project-owned generator:
// k = vmax / pml (vmax the window's largest Vp); a pass over the z bands (the
// top pml rows and their mirror rows at the bottom, depth d = pml - iz) and
// then one over the x bands (columns, d = pml - ix; corners are drawn twice,
// as in the reference) sets vp' = vp + (2u - 1) k d, redrawn until
// 0 <= vp' <= vmax (64 draws, then vp clamped to [0, vmax]), and
// vs' = (vs / vp) vp'.  The mirror cell is nzp-1-iz (nxp-1-ix); the reference
// writes nzpml-iz, one past the band.  u = (h >> 8) 2^-24 of the lowbias32
// hash h; every operation in FP32.
inline unsigned rbc_hash(unsigned x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}
inline float rbc_cell(unsigned seed, unsigned cell, int pass, int d, float kk, float vmax, float v0) {
    const unsigned key = rbc_hash(seed ^ rbc_hash(cell * 2u + (unsigned)pass));
    for (unsigned a = 0; a < 64; ++a) {
        const float u = (float)(rbc_hash(key + a) >> 8) * (1.0f / 16777216.0f);
        const float vn = v0 + ((2.0f * u - 1.0f) * kk) * (float)d;
        if (vn >= 0.0f && vn <= vmax) return vn;
    }
    return std::fmin(std::fmax(v0, 0.0f), vmax);
}
template <typename T>
void random_boundary(const Grid &g, unsigned seed, std::vector<T> &vp, std::vector<T> &vs) {
    float vmax = 0.0f;
    for (long i = 0; i < g.n; ++i) vmax = std::fmax(vmax, (float)vp[i]);
    const float kk = vmax / (float)g.pml;
    auto set = [&](long i, int pass, int d) {
        const float v0 = (float)vp[i], s0 = (float)vs[i];
        const float vn = rbc_cell(seed, (unsigned)i, pass, d, kk, vmax, v0);
        vs[i] = (T)(v0 > 0.0f ? (s0 / v0) * vn : s0);
        vp[i] = (T)vn;
    };
    for (int ix = 0; ix < g.nxp; ++ix)
        for (int iz = 0; iz < g.pml; ++iz) {
            set((long)ix * g.nzp + iz, 0, g.pml - iz);
            set((long)ix * g.nzp + g.nzp - 1 - iz, 0, g.pml - iz);
        }
    for (int ix = 0; ix < g.pml; ++ix)
        for (int iz = 0; iz < g.nzp; ++iz) {
            set((long)ix * g.nzp + iz, 1, g.pml - ix);
            set((long)(g.nxp - 1 - ix) * g.nzp + iz, 1, g.pml - ix);
        }
}

// lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2: sfwi-stage1/FD2DGPU.cu:1116-1124;
// stage 2 forces rho := 2000 on every padded cell inside the same kernel,
// before lambda and mu are formed (sfwi-stage2/FD2DGPU.cu:1106-1113).
// or_config.plant & 64 (test hook) forces it after them instead.
template <typename T>
void build_medium(const or_config *c, const Grid &g, const float *vp, const float *vs,
                  const float *rho, const float *dvp, const float *dvs, Medium<T> &m) {
    window_pad(g, c->allnz, c->left, vp, m.vp);
    window_pad(g, c->allnz, c->left, vs, m.vs);
    window_pad(g, c->allnz, c->left, rho, m.rho);
    const bool late_rho = (c->plant & 64) != 0;
    if (c->rho_const > 0 && !late_rho)
        for (long i = 0; i < g.n; ++i) m.rho[i] = (T)c->rho_const;
    if (c->rbc) random_boundary(g, c->rbc_seed, m.vp, m.vs);
    damping(g, c->rr, c->pml_off != 0 || c->rbc != 0, m.vp, m.dampx, m.dampz);
    m.lam.resize(g.n);
    m.mu.resize(g.n);
    for (long i = 0; i < g.n; ++i) {
        m.lam[i] = m.rho[i] * (m.vp[i] * m.vp[i] - 2 * m.vs[i] * m.vs[i]);
        m.mu[i] = m.rho[i] * m.vs[i] * m.vs[i];
    }
    if (c->rho_const > 0 && late_rho)
        for (long i = 0; i < g.n; ++i) m.rho[i] = (T)c->rho_const;
    if (dvp) window_pad(g, c->allnz, c->left, dvp, m.dvp);   // buffer_image_extrapolation :1336-1345
    if (dvs) window_pad(g, c->allnz, c->left, dvs, m.dvs);
}

// --------------------------------------------------------------- kernels ---

// Velocity update with split PML, sfwi-stage1/born.cu:355-394 (update_vx_vz).
// The 1.0/rho literal promotes the update expression to FP64 (quirk Q1).
template <typename T>
void update_v(const Grid &g, const Steps<T> &st, const Medium<T> &m, Wave<T> &w, int dir, int nth) {
    const long S = g.nzp;
    T *vx = w.vx.data(), *vz = w.vz.data(), *vx_x = w.vx_x.data(), *vx_z = w.vx_z.data();
    T *vz_x = w.vz_x.data(), *vz_z = w.vz_z.data();
    const T *txx = w.txx.data(), *tzz = w.tzz.data(), *txz = w.txz.data();
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const long i = (long)ix * S + iz;
            T d1 = 1 - st.dt * m.dampx[i] / 2, d2 = 1 + st.dt * m.dampx[i] / 2;
            T d3 = 1 - st.dt * m.dampz[i] / 2, d4 = 1 + st.dt * m.dampz[i] / 2;
            T s_txx_x = d_plus(txx, i, S);
            T s_txz_z = d_minus(txz, i, 1);
            T s_tzz_z = d_plus(tzz, i, 1);
            T s_txz_x = d_minus(txz, i, S);
            const double b = 1.0 / (double)m.rho[i];
            vx_x[i] = (T)(((double)(d1 * vx_x[i]) + dir * b * (double)(st.dt / st.dx) * (double)s_txx_x) / (double)d2);
            vx_z[i] = (T)(((double)(d3 * vx_z[i]) + dir * b * (double)(st.dt / st.dz) * (double)s_txz_z) / (double)d4);
            vz_x[i] = (T)(((double)(d1 * vz_x[i]) + dir * b * (double)(st.dt / st.dx) * (double)s_txz_x) / (double)d2);
            vz_z[i] = (T)(((double)(d3 * vz_z[i]) + dir * b * (double)(st.dt / st.dz) * (double)s_tzz_z) / (double)d4);
            vx[i] = vx_x[i] + vx_z[i];
            vz[i] = vz_x[i] + vz_z[i];
        }
}

// Stress update with split PML, sfwi-stage1/born.cu:399-440
// (update_txx_tzz_txz); with `grad` it is update_txx_tzz_txz_grad,
// sfwi-stage1/born.cu:451-499, which also stores the four velocity
// derivatives of the pre-update velocity.  mu is not averaged to the txz
// position (quirk Q2).
template <typename T>
void update_s(const Grid &g, const Steps<T> &st, const Medium<T> &m, Wave<T> &w, int dir, int nth,
              T *vx_gx = nullptr, T *vz_gz = nullptr, T *vx_gz = nullptr, T *vz_gx = nullptr) {
    const long S = g.nzp;
    const T *vx = w.vx.data(), *vz = w.vz.data();
    T *txx = w.txx.data(), *tzz = w.tzz.data(), *txz = w.txz.data();
    T *txx_x = w.txx_x.data(), *txx_z = w.txx_z.data(), *tzz_x = w.tzz_x.data();
    T *tzz_z = w.tzz_z.data(), *txz_x = w.txz_x.data(), *txz_z = w.txz_z.data();
    const T D = (T)dir;
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const long i = (long)ix * S + iz;
            T d1 = 1 - st.dt * m.dampx[i] / 2, d2 = 1 + st.dt * m.dampx[i] / 2;
            T d3 = 1 - st.dt * m.dampz[i] / 2, d4 = 1 + st.dt * m.dampz[i] / 2;
            T s_vx_x = d_minus(vx, i, S);
            T s_vz_z = d_minus(vz, i, 1);
            T s_vx_z = d_plus(vx, i, 1);
            T s_vz_x = d_plus(vz, i, S);
            const T lam = m.lam[i], mu = m.mu[i];
            txx_x[i] = (d1 * txx_x[i] + D * (lam + 2 * mu) * (st.dt / st.dx) * s_vx_x) / d2;
            txx_z[i] = (d3 * txx_z[i] + D * lam * (st.dt / st.dz) * s_vz_z) / d4;
            tzz_x[i] = (d1 * tzz_x[i] + D * lam * (st.dt / st.dx) * s_vx_x) / d2;
            tzz_z[i] = (d3 * tzz_z[i] + D * (lam + 2 * mu) * (st.dt / st.dz) * s_vz_z) / d4;
            txz_x[i] = (d1 * txz_x[i] + D * (st.dt / st.dx) * mu * s_vz_x) / d2;
            txz_z[i] = (d3 * txz_z[i] + D * (st.dt / st.dz) * mu * s_vx_z) / d4;
            txx[i] = txx_x[i] + txx_z[i];
            tzz[i] = tzz_x[i] + tzz_z[i];
            txz[i] = txz_x[i] + txz_z[i];
            if (vx_gx) {
                vx_gx[i] = s_vx_x / st.dx;
                vz_gz[i] = s_vz_z / st.dz;
                vx_gz[i] = s_vx_z / st.dz;
                vz_gx[i] = s_vz_x / st.dx;
            }
        }
}

// Adjoint "stress" update, sfwi-stage1/born.cu:565-603
// (GPUcalculate_elastic_txx_tzz_txz_back): no moduli, txx = txx_x and
// tzz = tzz_z (quirk Q8).
template <typename T>
void adjoint_s(const Grid &g, const Steps<T> &st, const Medium<T> &m, Wave<T> &w, int dir, int nth) {
    const long S = g.nzp;
    const T *vx = w.vx.data(), *vz = w.vz.data();
    T *txx = w.txx.data(), *tzz = w.tzz.data(), *txz = w.txz.data();
    T *txx_x = w.txx_x.data(), *tzz_z = w.tzz_z.data(), *txz_x = w.txz_x.data(), *txz_z = w.txz_z.data();
    const T D = (T)dir;
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const long i = (long)ix * S + iz;
            T d1 = 1 - st.dt * m.dampx[i] / 2, d2 = 1 + st.dt * m.dampx[i] / 2;
            T d3 = 1 - st.dt * m.dampz[i] / 2, d4 = 1 + st.dt * m.dampz[i] / 2;
            T s_vx_x = d_minus(vx, i, S);
            T s_vz_z = d_minus(vz, i, 1);
            T s_vx_z = d_plus(vx, i, 1);
            T s_vz_x = d_plus(vz, i, S);
            txx_x[i] = (d1 * txx_x[i] + D * (st.dt / st.dx) * s_vx_x) / d2;
            tzz_z[i] = (d3 * tzz_z[i] + D * (st.dt / st.dz) * s_vz_z) / d4;
            txz_x[i] = (d1 * txz_x[i] + D * (st.dt / st.dx) * s_vz_x) / d2;
            txz_z[i] = (d3 * txz_z[i] + D * (st.dt / st.dz) * s_vx_z) / d4;
            txx[i] = txx_x[i];
            tzz[i] = tzz_z[i];
            txz[i] = txz_x[i] + txz_z[i];
        }
}

// Adjoint velocity update with the moduli inside the derivatives,
// sfwi-stage1/born.cu:504-562 (GPUcalculate_elastic_vx_vz_back).  FP64
// promotion through 1.0/rho as in update_v.
template <typename T>
void adjoint_v(const Grid &g, const Steps<T> &st, const Medium<T> &m, Wave<T> &w, int dir, int nth) {
    const long S = g.nzp;
    T *vx = w.vx.data(), *vz = w.vz.data(), *vx_x = w.vx_x.data(), *vx_z = w.vx_z.data();
    T *vz_x = w.vz_x.data(), *vz_z = w.vz_z.data();
    const T *txx = w.txx.data(), *tzz = w.tzz.data(), *txz = w.txz.data();
    const T *lam = m.lam.data(), *mu = m.mu.data();
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const long i = (long)ix * S + iz;
            T d1 = 1 - st.dt * m.dampx[i] / 2, d2 = 1 + st.dt * m.dampx[i] / 2;
            T d3 = 1 - st.dt * m.dampz[i] / 2, d4 = 1 + st.dt * m.dampz[i] / 2;
            T s_txx_x = d_plus_p(lam, mu, txx, i, S);
            T s_txz_z = d_minus_w(mu, txz, i, 1);
            T s_tzz_x = d_plus_w(lam, tzz, i, S);
            T s_tzz_z = d_plus_p(lam, mu, tzz, i, 1);
            T s_txz_x = d_minus_w(mu, txz, i, S);
            T s_txx_z = d_plus_w(lam, txx, i, 1);
            const double b = 1.0 / (double)m.rho[i];
            const double ax = (double)(st.dt / st.dx), az = (double)(st.dt / st.dz);
            vx_x[i] = (T)(((double)(d1 * vx_x[i]) + dir * b * ax * (double)s_txx_x + dir * b * ax * (double)s_tzz_x) / (double)d2);
            vx_z[i] = (T)(((double)(d3 * vx_z[i]) + dir * b * az * (double)s_txz_z) / (double)d4);
            vz_x[i] = (T)(((double)(d1 * vz_x[i]) + dir * b * ax * (double)s_txz_x) / (double)d2);
            vz_z[i] = (T)(((double)(d3 * vz_z[i]) + dir * b * az * (double)s_tzz_z + dir * b * az * (double)s_txx_z) / (double)d4);
            vz[i] = vz_x[i] + vz_z[i];
            vx[i] = vx_x[i] + vx_z[i];
        }
}

// Born virtual source into the scattered split stresses from the updated
// velocity of the field it scatters from (the background, or the first-order
// scattered field for the second order).  Stage-2 live kernel
// sfwi-stage2/born.cu:749-801 (add_born_source_components; bounds
// [nop, n - nop), quirk Q9): lo = NOP.  Stage 1 (sfwi-stage1/born.cu:830-880,
// the Born solves of the Hessian-vector product and of double scattering):
// bounds [pml, n - pml), the interior only: lo = pml.  plant & 16 (test hook)
// drops the -4 rho vs dvs term of txx_z and tzz_x.
template <typename T>
void born_source(const Grid &g, const Steps<T> &st, const Medium<T> &m, const Wave<T> &bg, Wave<T> &b,
                 int dir, int nth, int lo = NOP, int plant = 0) {
    const long S = g.nzp;
    const T *vx = bg.vx.data(), *vz = bg.vz.data();
    const T D = (T)dir;
    const T k4 = (plant & 16) ? T(0) : T(4);
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int ix = lo; ix < g.nxp - lo; ++ix)
        for (int iz = lo; iz < g.nzp - lo; ++iz) {
            const long i = (long)ix * S + iz;
            T s_vx_x = d_minus(vx, i, S);
            T s_vz_z = d_minus(vz, i, 1);
            T s_vx_z = d_plus(vx, i, 1);
            T s_vz_x = d_plus(vz, i, S);
            const T rho = m.rho[i], vp = m.vp[i], vs = m.vs[i], dp = m.dvp[i], ds = m.dvs[i];
            b.txx_x[i] = b.txx_x[i] + D * st.dt * (2 * (rho * vp) * (dp) * (s_vx_x) / st.dx);
            b.txx_z[i] = b.txx_z[i] + D * st.dt * ((2 * (rho * vp) * (dp) - k4 * (rho * vs) * (ds)) * (s_vz_z) / st.dz);
            b.tzz_x[i] = b.tzz_x[i] + D * st.dt * ((2 * (rho * vp) * (dp) - k4 * (rho * vs) * (ds)) * (s_vx_x) / st.dx);
            b.tzz_z[i] = b.tzz_z[i] + D * st.dt * (2 * (rho * vp) * (dp) * (s_vz_z) / st.dz);
            b.txz_x[i] = b.txz_x[i] + D * st.dt * ((2 * (rho * vs) * (ds)) * ((s_vz_x) / st.dx));
            b.txz_z[i] = b.txz_z[i] + D * st.dt * ((2 * (rho * vs) * (ds)) * ((s_vx_z) / st.dz));
        }
}

// ------------------------------------------------- sources and receivers ---

// Receiver column on the padded grid: sfwi-stage1/FD2DGPU.cu:3162.
std::vector<int> receiver_columns(const or_config *c, const int *gc) {
    std::vector<int> igx(c->ntr);
    for (int tr = 0; tr < c->ntr; ++tr)
        igx[tr] = c->pml + (int)((float)(gc[tr] - c->ngx_left) / c->dx);
    return igx;
}

// Boundary strips of vx ("omega") and vz ("theta"), N = 4 cells either side
// of the PML interface, corners excluded (quirk Q14):
// sfwi-stage1/FD2DGPU.cu:3842-3869 (save) and :3873-3899 (restore).
template <typename T> struct Strips {
    int nx, nz, pml, nxp, nzp, N;
    int su = 0, sl = 0;                     // test hook (or_config.plant): up / left windows moved by a cell
    std::vector<T> up, down, left, right;   // per field: [field][it][...]
    void init(const Grid &g, int plant = 0) {
        nx = g.nx; nz = g.nz; pml = g.pml; nxp = g.nxp; nzp = g.nzp; N = NOP;
        su = (plant & 1) ? 1 : 0;
        sl = (plant & 2) ? 1 : 0;
        long a = 2L * (long)g.nt * 2 * N * nx, b = 2L * (long)g.nt * 2 * N * nz;
        up.assign(a, T(0)); down.assign(a, T(0)); left.assign(b, T(0)); right.assign(b, T(0));
    }
    template <bool SAVE> void move(T *fvx, T *fvz, int it, int nt) {
        T *f[2] = {fvx, fvz};
        for (int k = 0; k < 2; ++k) {
            long ou = (long)k * nt * 2 * N * nx + (long)it * nx * 2 * N;
            long ol = (long)k * nt * 2 * N * nz + (long)it * 2 * N * nz;
            for (int ix = pml; ix < nxp - pml; ++ix) {
                for (int iz = pml - N + su; iz < pml + N + su; ++iz)
                    sw<SAVE>(f[k][(long)ix * nzp + iz], up[ou + (long)(ix - pml) * 2 * N + (iz - (pml - N + su))]);
                for (int iz = nzp - pml - N; iz < nzp - pml + N; ++iz)
                    sw<SAVE>(f[k][(long)ix * nzp + iz], down[ou + (long)(ix - pml) * 2 * N + (iz - (nzp - pml - N))]);
            }
            for (int ix = pml - N + sl; ix < pml + N + sl; ++ix)
                for (int iz = pml; iz < nzp - pml; ++iz)
                    sw<SAVE>(f[k][(long)ix * nzp + iz], left[ol + (long)(ix - (pml - N + sl)) * nz + (iz - pml)]);
            for (int ix = nxp - pml - N; ix < nxp - pml + N; ++ix)
                for (int iz = pml; iz < nzp - pml; ++iz)
                    sw<SAVE>(f[k][(long)ix * nzp + iz], right[ol + (long)(ix - (nxp - pml - N)) * nz + (iz - pml)]);
        }
    }
    template <bool SAVE> static void sw(T &field, T &buf) {
        if (SAVE) buf = field; else field = buf;
    }
};

// Zero-lag imaging of one reverse step, sfwi-stage1/FD2DGPU.cu:3350-3382
// (doimage_gradient) and :3324-3331 (doimage_illum).
// With grho (no reference counterpart: the reference images Vp and Vs only)
// also the density gradient at fixed Vp, Vs in the same units:
//   -(vx_adj dvx + vz_adj dvz) / dt + (vp^2 - 2 vs^2) lg + vs^2 mg,
// dv = v(it+1) - v(it) of the reconstruction step, the first term the
// d/drho of the velocity update v += dt/rho D sigma through the adjoint
// mapping (adjoint v = a / rho for the exact multiplier a), the others the
// chain rule through lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2.
template <typename T, typename G>
void image_step(const or_config *c, const Grid &g, const Medium<T> &m, const Wave<T> &fw,
                const Wave<T> &bk, const std::vector<T> &vx_gx, const std::vector<T> &vz_gz,
                const std::vector<T> &vx_gz, const std::vector<T> &vz_gx, G *gvp, G *gvs, G *ill,
                int nth, G *grho = nullptr, const std::vector<T> *ovx = nullptr,
                const std::vector<T> *ovz = nullptr, T dt = T(0)) {
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int ix = 0; ix < g.nx; ++ix)
        for (int iz = 0; iz < g.nz; ++iz) {
            const long p = (long)(ix + g.pml) * g.nzp + iz + g.pml;
            const long q = (long)(c->left + ix) * g.nz + iz;
            T lg = -((vx_gx[p] + vz_gz[p]) * (bk.txx[p] + bk.tzz[p]));
            T mg = -2 * (vx_gx[p] * bk.txx[p] + vz_gz[p] * bk.tzz[p]) - (vx_gz[p] + vz_gx[p]) * bk.txz[p];
            T pp = 2 * m.vp[p] * m.rho[p] * lg;
            T ps = -4 * m.vs[p] * m.rho[p] * lg + 2 * m.vs[p] * m.rho[p] * mg;
            T one = 1;
            gvp[q] += one * pp;
            gvs[q] += one * ps;
            T il = fw.vx[p] * fw.vx[p] + fw.vz[p] * fw.vz[p];
            ill[q] += one * il;
            if (grho) {
                T dvx = (*ovx)[p] - fw.vx[p], dvz = (*ovz)[p] - fw.vz[p];
                T rd = -(bk.vx[p] * dvx + bk.vz[p] * dvz) / dt;
                T rg = rd + (m.vp[p] * m.vp[p] - 2 * m.vs[p] * m.vs[p]) * lg + m.vs[p] * m.vs[p] * mg;
                grho[q] += one * rg;
            }
        }
}

// Misfit sum over the residual, i-major then j, each term formed in the
// record precision and accumulated into a double (sfwi-stage1/FD2DGPU.cu:4587-4590).
template <typename T> double l2_sum(const std::vector<T> &rz, const std::vector<T> &rx, int ntr, int nt) {
    double s = 0;
    for (int i = 0; i < ntr; ++i)
        for (int j = 0; j < nt; ++j) {
            long k = (long)i * nt + j;
            s += (rz[k] * rz[k] + rx[k] * rx[k]);
        }
    return s;
}

// ---------------------------------------------------------- per-shot runs ---

template <typename T> struct Run {
    const or_config *c;
    Grid g;
    Steps<T> st;
    Medium<T> m;
    int nth;
    std::vector<int> igx;
    Run(const or_config *cfg, const float *vp, const float *vs, const float *rho, const float *dvp,
        const float *dvs, const int *gc)
        : c(cfg), g(make_grid(cfg)) {
        st.dx = (T)cfg->dx;
        st.dz = (T)cfg->dz;
        st.dt = (T)cfg->dt;
        nth = nthreads_of(cfg);
        build_medium(cfg, g, vp, vs, rho, dvp, dvs, m);
        if (gc) igx = receiver_columns(cfg, gc);
    }
    long at(int ix, int iz) const { return (long)ix * g.nzp + iz; }
    const float *wav = nullptr;   // the forward solve's wavelet
    // Explosive source on txx and tzz before the velocity update,
    // sfwi-stage1/FD2DGPU.cu:2104-2114.
    void add_source(Wave<T> &w, const float *wav, int it) {
        long i = at(c->isx, c->isz);
        w.txx[i] += (T)wav[it];
        w.tzz[i] += (T)wav[it];
    }
    // Receiver sampling, sfwi-stage1/FD2DGPU.cu:3158-3166.
    void record(const Wave<T> &w, int it, std::vector<T> &rz, std::vector<T> &rx) {
        for (int tr = 0; tr < g.ntr; ++tr) {
            long i = at(igx[tr], c->igz);
            rz[(long)tr * g.nt + it] = w.vz[i];
            rx[(long)tr * g.nt + it] = w.vx[i];
        }
    }
    // Adjoint source: the residual overwrites the adjoint velocity at the
    // receivers (stack=false), sfwi-stage1/FD2DGPU.cu:3425-3441 (quirk Q6).
    // Only the totals are written, and the adjoint velocity step rebuilds
    // them from the split parts, so the residual reaches the adjoint field
    // through the stress-like step that follows alone.  or_config.plant & 128
    // (test hook) adds it to the totals and the x split parts instead, as the
    // consistent gradient does.
    void inject(Wave<T> &bk, int it, const std::vector<T> &rz, const std::vector<T> &rx) {
        const bool add = (c->plant & 128) != 0;
        for (int tr = 0; tr < g.ntr; ++tr) {
            long i = at(igx[tr], c->igz);
            if (add) {
                const T az = rz[(long)tr * g.nt + it], ax = rx[(long)tr * g.nt + it];
                bk.vz[i] += az;
                bk.vz_x[i] += az;
                bk.vx[i] += ax;
                bk.vx_x[i] += ax;
            } else {
                bk.vz[i] = rz[(long)tr * g.nt + it];
                bk.vx[i] = rx[(long)tr * g.nt + it];
            }
        }
    }

    // Background forward modeling, born_modeling mode 0
    // (sfwi-stage1/FD2DGPU.cu:1864-1874) inside the loop :4535-4564.
    // init: starting state (zero when NULL); before: called with each step's
    // index and the state before it (test hooks).  or_config.plant & 4 (test
    // hook) adds the source after the velocity update instead.
    void forward(const float *wav_, std::vector<T> &rz, std::vector<T> &rx, Wave<T> *keep = nullptr,
                 Strips<T> *strips = nullptr, const Wave<T> *init = nullptr,
                 const std::function<void(int, const Wave<T> &)> *before = nullptr) {
        wav = wav_;
        Wave<T> w(g.n);
        if (init) w = *init;
        if (c->rbc == 1) strips = nullptr;   // random boundary: nothing saved (FD2DGPU.cu:4543)
        const bool late = (c->plant & 4) != 0;
        for (int it = 0; it < g.nt; ++it) {
            if (before) (*before)(it, w);
            if (!late) add_source(w, wav, it);
            update_v(g, st, m, w, +1, nth);
            if (late) add_source(w, wav, it);
            update_s(g, st, m, w, +1, nth);
            if (!igx.empty()) record(w, it, rz, rx);
            if (strips) strips->template move<true>(w.vx.data(), w.vz.data(), it, g.nt);
        }
        if (keep) *keep = w;
    }

    // One reverse step of the forward field, born_modeling mode 6 inside
    // sfwi-stage1/FD2DGPU.cu:4600-4640: the strips of step it restored (not
    // with the random boundary), the stress step with the velocity
    // derivatives, the velocity step, direction -1; the source is not
    // subtracted (quirk Q5).  The forward step adds its source to the stress
    // totals only, and the stress step rebuilds the totals from the split
    // parts, so the source acts on that step's velocity update alone; the
    // exact inverse therefore lets the reverse velocity step see the source in
    // the totals again.  src_sign (test hooks): +1 does that (the exact
    // inverse), -1 lets it see the negated source (the difference of the
    // faithful and the exact reconstructions, by linearity); or_config.plant
    // & 8 makes the gradient's reverse loop use the exact inverse.  ovx, ovz
    // receive v(it+1) before the velocity step.
    void reconstruct(Wave<T> &fw, Strips<T> *strips, int it, T *gx = nullptr, T *gz = nullptr,
                     T *gxz = nullptr, T *gzx = nullptr, std::vector<T> *ovx = nullptr,
                     std::vector<T> *ovz = nullptr, int src_sign = 0) {
        if (strips && c->rbc != 1) strips->template move<false>(fw.vx.data(), fw.vz.data(), it, g.nt);
        update_s(g, st, m, fw, -1, nth, gx, gz, gxz, gzx);
        if (ovx) {
            *ovx = fw.vx;
            *ovz = fw.vz;
        }
        const long i = at(c->isx, c->isz);
        const T s = (src_sign && wav) ? (T)src_sign * (T)wav[it] : T(0);
        if (s != T(0)) { fw.txx[i] += s; fw.tzz[i] += s; }
        update_v(g, st, m, fw, -1, nth);
        if (s != T(0)) { fw.txx[i] -= s; fw.tzz[i] -= s; }
    }

    // Stage-2 forward, mode 1 (sfwi-stage2/FD2DGPU.cu:1814-1845) inside
    // sfwi-stage2/FD2DGPU.cu:4223-4314: background step, scattered velocity,
    // Born source from the updated background velocity, scattered stress.
    // lo: the Born source's bound (NOP stage 2, pml stage 1).  With b2z / b2x
    // (stage 1, double scattering) each step continues with mode 5
    // (sfwi-stage1/FD2DGPU.cu:1888-1903): the second-order scattered velocity,
    // the Born source from the updated first-order scattered velocity, the
    // second-order scattered stress; its traces are recorded as well
    // (record_scatter with double_scatter, :3610-3617).  plant & 32 (test
    // hook) drives the second order from the background instead.
    void born_forward(const float *wav, std::vector<T> &bz, std::vector<T> &bx, std::vector<T> &gz,
                      std::vector<T> &gx, Wave<T> *keep = nullptr, Strips<T> *strips = nullptr, int lo = NOP,
                      std::vector<T> *b2z = nullptr, std::vector<T> *b2x = nullptr) {
        Wave<T> w(g.n), b(g.n), b2(b2z ? g.n : 1);
        if (c->rbc == 1) strips = nullptr;
        for (int it = 0; it < g.nt; ++it) {
            add_source(w, wav, it);
            update_v(g, st, m, w, +1, nth);
            update_s(g, st, m, w, +1, nth);
            update_v(g, st, m, b, +1, nth);
            born_source(g, st, m, w, b, +1, nth, lo, c->plant);
            update_s(g, st, m, b, +1, nth);
            if (b2z) {
                update_v(g, st, m, b2, +1, nth);
                born_source(g, st, m, (c->plant & 32) ? w : b, b2, +1, nth, lo, c->plant);
                update_s(g, st, m, b2, +1, nth);
                record(b2, it, *b2z, *b2x);
            }
            record(b, it, bz, bx);   // record_scatter: scatter = b, scatter2 = background
            record(w, it, gz, gx);
            if (strips) strips->template move<true>(w.vx.data(), w.vz.data(), it, g.nt);
        }
        if (keep) *keep = w;
    }

    // Reverse loop of IMAGE_SINGLESHOT (sfwi-stage1/FD2DGPU.cu:4600-4640;
    // stage 2 :4379-4510): restore strips, reconstruct (mode 6: stress with
    // derivatives, then velocity, direction -1, source not subtracted, Q5),
    // inject the residual, adjoint step (mode 8) with the stage's direction
    // (Q7), image.  or_config.plant & 256 (test hook) images before the
    // adjoint step instead.
    template <typename G>
    void reverse(Wave<T> &fw, Strips<T> &strips, const std::vector<T> &rz, const std::vector<T> &rx,
                 G *gvp, G *gvs, G *ill, G *grho = nullptr) {
        Wave<T> bk(g.n);
        std::vector<T> vx_gx(g.n, T(0)), vz_gz(g.n, T(0)), vx_gz(g.n, T(0)), vz_gx(g.n, T(0));
        std::vector<T> ovx, ovz;
        for (int it = g.nt - 1; it >= 0; --it) {
            reconstruct(fw, &strips, it, vx_gx.data(), vz_gz.data(), vx_gz.data(), vz_gx.data(),
                        grho ? &ovx : nullptr, grho ? &ovz : nullptr, (c->plant & 8) ? +1 : 0);
            inject(bk, it, rz, rx);
            const bool early = (c->plant & 256) != 0;
            if (early)
                image_step(c, g, m, fw, bk, vx_gx, vz_gz, vx_gz, vz_gx, gvp, gvs, ill, nth, grho, &ovx, &ovz, st.dt);
            adjoint_s(g, st, m, bk, c->adjoint_dir, nth);
            adjoint_v(g, st, m, bk, c->adjoint_dir, nth);
            if (!early)
                image_step(c, g, m, fw, bk, vx_gx, vz_gz, vx_gz, vz_gx, gvp, gvs, ill, nth, grho, &ovx, &ovz, st.dt);
        }
    }
};

template <typename T, typename G>
void gradient_impl(const or_config *c, const float *vp, const float *vs, const float *rho,
                   const float *wav, const int *gc, const float *obs_z, const float *obs_x, G *gvp,
                   G *gvs, G *ill, double *misfit, G *res_z, G *res_x, G *grho = nullptr) {
    Run<T> r(c, vp, vs, rho, nullptr, nullptr, gc);
    const long nr = (long)c->ntr * c->nt;
    std::vector<T> sz(nr), sx(nr);
    Strips<T> strips;
    strips.init(r.g, c->plant);
    Wave<T> fw(1);
    r.forward(wav, sz, sx, &fw, &strips);
    // r = 1.0*syn - obs  (sfwi-stage1/FD2DGPU.cu:3181-3195 with alpha = 1)
    for (long k = 0; k < nr; ++k) {
        sz[k] = (T)1.0f * sz[k] - (T)obs_z[k];
        sx[k] = (T)1.0f * sx[k] - (T)obs_x[k];
    }
    *misfit += l2_sum(sz, sx, c->ntr, c->nt);
    if (res_z)
        for (long k = 0; k < nr; ++k) { res_z[k] = (G)sz[k]; res_x[k] = (G)sx[k]; }
    r.reverse(fw, strips, sz, sx, gvp, gvs, ill, grho);
}

// Forward (strips) then the reverse loop with a caller-supplied adjoint
// source in place of the L2 residual (used with the OT misfit).
template <typename T>
void gradient_res_impl(const or_config *c, const float *vp, const float *vs, const float *rho,
                       const float *wav, const int *gc, const float *res_z, const float *res_x, T *gvp,
                       T *gvs, T *ill) {
    Run<T> r(c, vp, vs, rho, nullptr, nullptr, gc);
    const long nr = (long)c->ntr * c->nt;
    std::vector<T> sz(nr), sx(nr);
    Strips<T> strips;
    strips.init(r.g, c->plant);
    Wave<T> fw(1);
    r.forward(wav, sz, sx, &fw, &strips);
    for (long k = 0; k < nr; ++k) {
        sz[k] = (T)res_z[k];
        sx[k] = (T)res_x[k];
    }
    r.reverse(fw, strips, sz, sx, gvp, gvs, ill);
}

template <typename T, typename G>
void born_gradient_impl(const or_config *c, const float *vp, const float *vs, const float *rho,
                        const float *dvp, const float *dvs, const float *wav, const int *gc,
                        const float *obs_z, const float *obs_x, G *gvp, G *gvs, G *ill, double *misfit) {
    Run<T> r(c, vp, vs, rho, dvp, dvs, gc);
    const long nr = (long)c->ntr * c->nt;
    std::vector<T> bz(nr), bx(nr), gz(nr), gx(nr);
    Strips<T> strips;
    strips.init(r.g, c->plant);
    Wave<T> fw(1);
    r.born_forward(wav, bz, bx, gz, gx, &fw, &strips);
    // r = (b - obs) then r = 1*r + background: sfwi-stage2/FD2DGPU.cu:3119-3130,
    // :3144-3158 (the subtract ignores alpha, quirk Q16).
    for (long k = 0; k < nr; ++k) {
        bz[k] = bz[k] - (T)obs_z[k];
        bx[k] = bx[k] - (T)obs_x[k];
        bz[k] = (T)1.0f * bz[k] + gz[k];
        bx[k] = (T)1.0f * bx[k] + gx[k];
    }
    *misfit += l2_sum(bz, bx, c->ntr, c->nt);
    r.reverse(fw, strips, bz, bx, gvp, gvs, ill);
}

// Gauss-Newton Hessian-vector product of one shot, ELSRTM_CALCULATE_Hessian_vector
// (sfwi-stage1/FD2DGPU.cu:5236-5359): the vector p is the perturbation of a
// stage-1 Born forward solve (mode 1, source on the interior only) that
// saves the background strips and records the scattered traces; the reverse
// loop of IMAGE_SINGLESHOT then runs with those traces as the adjoint source
// (overwrite injection, stack = false, :3425-3447) and images with
// doimage_gradient alone (hessian_vector_product, :3469-3473: no
// illumination).  Accumulates H p into gvp, gvs.
template <typename T, typename G>
void hessian_impl(const or_config *c, const float *vp, const float *vs, const float *rho, const float *pvp,
                  const float *pvs, const float *wav, const int *gc, G *gvp, G *gvs) {
    Run<T> r(c, vp, vs, rho, pvp, pvs, gc);
    const long nr = (long)c->ntr * c->nt;
    std::vector<T> bz(nr), bx(nr), gz(nr), gx(nr);
    Strips<T> strips;
    strips.init(r.g, c->plant);
    Wave<T> fw(1);
    r.born_forward(wav, bz, bx, gz, gx, &fw, &strips, c->pml);
    std::vector<G> ill((size_t)c->allnx * c->allnz, G(0));   // not part of the product
    r.reverse(fw, strips, bz, bx, gvp, gvs, ill.data());
}

}  // namespace

// ------------------------------------------------------------- C entries ---

extern "C" {

void or_coeffs_f32(float c[5]) { for (int k = 0; k < 5; ++k) c[k] = FD<float>::c[k]; }
void or_coeffs_f64(double c[5]) { for (int k = 0; k < 5; ++k) c[k] = FD<double>::c[k]; }

// sfwi-stage1/main.cpp:251-260 (PI = 3.14159265358979, main.cpp:8).
void or_ricker(float *w, int nt, float dt, float f0) {
    const double pi = 3.14159265358979;
    float nk = pi * pi * f0 * f0 * dt * dt;
    int t0 = (int)std::ceil(1.0 / (f0 * dt));
    for (int t = 0; t < nt; ++t)
        w[t] = (1.0 - 2.0 * nk * (t - t0) * (t - t0)) * std::exp(-nk * (t - t0) * (t - t0));
}

}  // extern "C"

template <typename T>
static void setup_impl(const or_config *c, const float *vp, const float *vs, const float *rho, T *out) {
    Grid g = make_grid(c);
    Medium<T> m;
    build_medium(c, g, vp, vs, rho, nullptr, nullptr, m);
    const std::vector<T> *arr[7] = {&m.vp, &m.vs, &m.rho, &m.lam, &m.mu, &m.dampx, &m.dampz};
    for (int k = 0; k < 7; ++k) std::memcpy(out + (long)k * g.n, arr[k]->data(), g.n * sizeof(T));
}

extern "C" {
void or_setup_f32(const or_config *c, const float *vp, const float *vs, const float *rho, float *o) { setup_impl(c, vp, vs, rho, o); }
void or_setup_f64(const or_config *c, const float *vp, const float *vs, const float *rho, double *o) { setup_impl(c, vp, vs, rho, o); }

}  // extern "C"

template <typename T>
static void forward_impl(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *wav, const int *gc, T *syn_z, T *syn_x) {
    Run<T> r(c, vp, vs, rho, nullptr, nullptr, gc);
    const long nr = (long)c->ntr * c->nt;
    std::vector<T> sz(nr), sx(nr);
    r.forward(wav, sz, sx);
    std::memcpy(syn_z, sz.data(), nr * sizeof(T));
    std::memcpy(syn_x, sx.data(), nr * sizeof(T));
}

// Consistent discrete gradient (test device, no reference counterpart): the
// exact gradient of the L2 misfit sum_(traces, steps) r^2 with respect to the
// window's vp, vs and rho for the same forward scheme, built to pin the imaging
// formula of image_step by finite differences.  It differs from the faithful
// gradient (gradient_impl) only in the orchestration: the forward velocities
// and derivatives are kept instead of reconstructed, the residual is added
// (2 r / rho, the transpose of receiver sampling under a = rho v_adj) instead
// of overwriting the adjoint velocity, and each step images between the
// adjoint velocity-like and stress-like half steps (the adjoint of the stress
// update pairs with the stresses it produced); the result is scaled by dt.
// With the PML off the adjoint half steps are exact transposes (dot-product
// pins), so these images must match central differences of the misfit.
template <typename T>
static void gradient_consistent_impl(const or_config *c, const float *vp, const float *vs, const float *rho,
                                     const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                                     T *gvp, T *gvs, T *grho, double *misfit) {
    Run<T> r(c, vp, vs, rho, nullptr, nullptr, gc);
    const Grid &g = r.g;
    const long n = g.n, nr = (long)c->ntr * c->nt;
    std::vector<T> sz(nr), sx(nr);
    std::vector<std::vector<T>> keep(c->nt);   // per step: vx, vz before and after the velocity update, 4 derivatives
    Wave<T> w(n);
    for (int it = 0; it < c->nt; ++it) {
        r.add_source(w, wav, it);
        std::vector<T> &k = keep[it];
        k.resize(8 * n);
        std::copy(w.vx.begin(), w.vx.end(), k.begin());
        std::copy(w.vz.begin(), w.vz.end(), k.begin() + n);
        update_v(g, r.st, r.m, w, +1, r.nth);
        std::copy(w.vx.begin(), w.vx.end(), k.begin() + 2 * n);
        std::copy(w.vz.begin(), w.vz.end(), k.begin() + 3 * n);
        update_s(g, r.st, r.m, w, +1, r.nth, k.data() + 4 * n, k.data() + 5 * n, k.data() + 6 * n, k.data() + 7 * n);
        r.record(w, it, sz, sx);
    }
    for (long q = 0; q < nr; ++q) {
        sz[q] = sz[q] - (T)obs_z[q];
        sx[q] = sx[q] - (T)obs_x[q];
    }
    *misfit += l2_sum(sz, sx, c->ntr, c->nt);
    Wave<T> bk(n), fw(n);
    std::vector<T> ovx(n), ovz(n), d[4], ill((size_t)c->allnx * c->allnz);
    for (int f = 0; f < 4; ++f) d[f].resize(n);
    const size_t nm = (size_t)c->allnx * c->allnz;
    std::vector<T> ivp(nm, T(0)), ivs(nm, T(0)), irho(nm, T(0));
    for (int it = c->nt - 1; it >= 0; --it) {
        for (int tr = 0; tr < g.ntr; ++tr) {   // the split parts keep summing to the totals
            const long i = r.at(r.igx[tr], c->igz);
            const T az = 2 * sz[(long)tr * c->nt + it] / r.m.rho[i];
            const T ax = 2 * sx[(long)tr * c->nt + it] / r.m.rho[i];
            bk.vz[i] += az;
            bk.vz_x[i] += az;
            bk.vx[i] += ax;
            bk.vx_x[i] += ax;
        }
        adjoint_v(g, r.st, r.m, bk, +1, r.nth);
        const std::vector<T> &k = keep[it];
        std::copy(k.begin(), k.begin() + n, fw.vx.begin());
        std::copy(k.begin() + n, k.begin() + 2 * n, fw.vz.begin());
        std::copy(k.begin() + 2 * n, k.begin() + 3 * n, ovx.begin());
        std::copy(k.begin() + 3 * n, k.begin() + 4 * n, ovz.begin());
        for (int f = 0; f < 4; ++f) std::copy(k.begin() + (4 + f) * n, k.begin() + (5 + f) * n, d[f].begin());
        image_step(c, g, r.m, fw, bk, d[0], d[1], d[2], d[3], ivp.data(), ivs.data(), ill.data(), r.nth,
                   irho.data(), &ovx, &ovz, r.st.dt);
        adjoint_s(g, r.st, r.m, bk, +1, r.nth);
    }
    for (size_t q = 0; q < nm; ++q) {
        gvp[q] = r.st.dt * ivp[q];
        gvs[q] = r.st.dt * ivs[q];
        grho[q] = r.st.dt * irho[q];
    }
}

// Test device (no reference counterpart) for the faithful gradient's
// orchestration: its forward loop (Run::forward, strips saved) from `init`
// (zero when NULL), then its reverse reconstruction steps (Run::reconstruct)
// with no recording and no adjoint.  kind 0: as the gradient (the source
// left out of the reverse velocity steps, quirk Q5); 1: the exact inverse
// (each reverse velocity step sees its step's source, as the forward one did);
// 2: their difference by linearity -- from a zero state with zero strips,
// each reverse velocity step seeing the negated source.  errs[it] (may be NULL): the largest difference
// between the reconstructed totals after reverse step it and the forward
// totals before forward step it -- velocities on the cells at least pml + 4
// from every edge, stresses on the imaged interior (at least pml from every
// edge) -- divided by the largest forward total there.  out: the state after
// reverse step 0 (15 arrays of nxpml*nzpml).
template <typename T>
static void reconstruct_impl(const or_config *c, const float *vp, const float *vs, const float *rho,
                             const float *wav, const double *init, int kind, double *errs, double *out) {
    Run<T> r(c, vp, vs, rho, nullptr, nullptr, nullptr);
    const Grid &g = r.g;
    const long n = g.n;
    Strips<T> strips;
    strips.init(g, c->plant);
    Wave<T> w(n);
    // checked cells: velocities at least pml + 4 from every edge, stresses on
    // the imaged interior (at least pml from every edge)
    auto inside = [&](long i, int f) {
        const int lo = f < 2 ? g.pml + NOP : g.pml;
        const int ix = (int)(i / g.nzp), iz = (int)(i % g.nzp);
        return ix >= lo && ix < g.nxp - lo && iz >= lo && iz < g.nzp - lo;
    };
    std::vector<std::vector<T>> keep;
    if (kind != 2) {
        Wave<T> w0(n);
        if (init) {
            std::vector<std::vector<T> *> a = w0.all();
            for (int k = 0; k < 15; ++k)
                for (long i = 0; i < n; ++i) (*a[k])[i] = (T)init[k * n + i];
        }
        if (errs) keep.resize(g.nt);
        std::function<void(int, const Wave<T> &)> hook = [&](int it, const Wave<T> &x) {
            std::vector<T> &k = keep[it];
            k.resize(5 * n);
            const std::vector<T> *t[5] = {&x.vx, &x.vz, &x.txx, &x.tzz, &x.txz};
            for (int f = 0; f < 5; ++f) std::copy(t[f]->begin(), t[f]->end(), k.begin() + f * n);
        };
        std::vector<T> rz, rx;
        r.forward(wav, rz, rx, &w, &strips, init ? &w0 : nullptr, errs ? &hook : nullptr);
    } else {
        r.wav = wav;
    }
    T scale = 0;
    if (errs)
        for (const std::vector<T> &k : keep)
            for (long q = 0; q < 5 * n; ++q)
                if (inside(q % n, (int)(q / n))) scale = std::max(scale, std::abs(k[q]));
    const int sign = kind == 1 ? +1 : kind == 2 ? -1 : ((c->plant & 8) ? +1 : 0);
    for (int it = g.nt - 1; it >= 0; --it) {
        r.reconstruct(w, &strips, it, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, sign);
        if (errs) {
            const std::vector<T> *t[5] = {&w.vx, &w.vz, &w.txx, &w.tzz, &w.txz};
            T e = 0;
            for (int f = 0; f < 5; ++f)
                for (long i = 0; i < n; ++i)
                    if (inside(i, f)) e = std::max(e, std::abs((*t[f])[i] - keep[it][f * n + i]));
            errs[it] = scale > 0 ? (double)(e / scale) : (double)e;
        }
    }
    std::vector<std::vector<T> *> a = w.all();
    for (int k = 0; k < 15; ++k)
        for (long i = 0; i < n; ++i) out[k * n + i] = (double)(*a[k])[i];
}

extern "C" {
void or_reconstruct_f64(const or_config *c, const float *vp, const float *vs, const float *rho, const float *wav,
                        const double *init, int kind, double *errs, double *out) {
    reconstruct_impl<double>(c, vp, vs, rho, wav, init, kind, errs, out);
}
}  // extern "C"

// Snapshot gradient (no reference counterpart; the oracle of the library's
// recon = snapshots mode).  The consistent gradient above with the imaging
// decimated in time: the forward velocity after the velocity update of every
// step it with it % K == 0 is kept on the imaged interior plus a 4-cell halo
// (what its derivatives need), and the reverse pass images only at those
// steps, each with weight K.  Per reverse step: add the adjoint source at the
// receivers (2 r / rho for the L2 misfit, or a caller-supplied source adj / rho,
// the derivative of another misfit with respect to the synthetic traces),
// adjoint velocity half step, imaging from the kept velocity and the adjoint
// stresses, adjoint stress half step; direction +1.  With K = 1 the Vp / Vs
// images equal gradient_consistent_impl's (a CPU test checks it), which
// finite differences of the misfit pin.  Born (dvp, dvs given): the forward
// solve is the stage-2 one and the residual born - obs + background; the
// background velocity is kept.  Images scaled by K dt; illumination
// K sum (vx^2 + vz^2) of the kept velocity.
template <typename T>
static void gradient_snapshot_impl(const or_config *c, const float *vp, const float *vs, const float *rho,
                                   const float *dvp, const float *dvs, const float *wav, const int *gc,
                                   const float *obs_z, const float *obs_x, const float *adj_z,
                                   const float *adj_x, T *gvp, T *gvs, T *gil, double *misfit,
                                   bool hess = false) {
    const bool born = dvp && dvs;
    Run<T> r(c, vp, vs, rho, born ? dvp : nullptr, born ? dvs : nullptr, gc);
    const Grid &g = r.g;
    const long n = g.n, nr = (long)c->ntr * c->nt;
    const int K = c->snap_every > 1 ? c->snap_every : 1;
    const int nsnap = (c->nt + K - 1) / K;
    // kept region: columns [pml-4, pml+nx+4), rows [pml-4, pml+nz+4)
    const int x0 = g.pml - NOP, z0 = g.pml - NOP, RX = g.nx + 2 * NOP, RZ = g.nz + 2 * NOP;
    const long RN = (long)RX * RZ;
    std::vector<T> keep((size_t)nsnap * 2 * RN);
    auto save = [&](const Wave<T> &w, int k) {
        T *o = keep.data() + (size_t)k * 2 * RN;
        for (int x = 0; x < RX; ++x)
            for (int z = 0; z < RZ; ++z) {
                const long i = r.at(x0 + x, z0 + z), q = (long)x * RZ + z;
                o[q] = w.vx[i];
                o[RN + q] = w.vz[i];
            }
    };
    std::vector<T> sz(nr), sx(nr), bz, bx;
    Wave<T> w(n), b(born ? n : 1);
    if (born) { bz.resize(nr); bx.resize(nr); }
    for (int it = 0; it < c->nt; ++it) {
        r.add_source(w, wav, it);
        update_v(g, r.st, r.m, w, +1, r.nth);
        if (it % K == 0) save(w, it / K);
        update_s(g, r.st, r.m, w, +1, r.nth);
        if (born) {
            update_v(g, r.st, r.m, b, +1, r.nth);
            born_source(g, r.st, r.m, w, b, +1, r.nth, hess ? g.pml : NOP, c->plant);
            update_s(g, r.st, r.m, b, +1, r.nth);
            r.record(b, it, bz, bx);
        }
        r.record(w, it, sz, sx);
    }
    for (long q = 0; q < nr; ++q) {   // residual (stage 2: born - obs + background)
        if (hess) {                   // Hessian-vector product: the scattered traces
            sz[q] = bz[q];
            sx[q] = bx[q];
        } else if (born) {
            sz[q] = (bz[q] - (T)obs_z[q]) + sz[q];
            sx[q] = (bx[q] - (T)obs_x[q]) + sx[q];
        } else {
            sz[q] = sz[q] - (T)obs_z[q];
            sx[q] = sx[q] - (T)obs_x[q];
        }
    }
    if (!hess) *misfit += l2_sum(sz, sx, c->ntr, c->nt);
    Wave<T> bk(n);
    std::vector<T> fvx(n, T(0)), fvz(n, T(0));
    const size_t nm = (size_t)c->allnx * c->allnz;
    std::vector<T> ivp(nm, T(0)), ivs(nm, T(0)), ill(nm, T(0));
    const T dx = r.st.dx, dz = r.st.dz;
    for (int it = c->nt - 1; it >= 0; --it) {
        for (int tr = 0; tr < g.ntr; ++tr) {   // the split parts keep summing to the totals
            const long i = r.at(r.igx[tr], c->igz), o = (long)tr * c->nt + it;
            const T az = (adj_z ? (T)adj_z[o] : 2 * sz[o]) / r.m.rho[i];
            const T ax = (adj_x ? (T)adj_x[o] : 2 * sx[o]) / r.m.rho[i];
            bk.vz[i] += az;
            bk.vz_x[i] += az;
            bk.vx[i] += ax;
            bk.vx_x[i] += ax;
        }
        adjoint_v(g, r.st, r.m, bk, +1, r.nth);
        if (it % K == 0) {
            const T *o = keep.data() + (size_t)(it / K) * 2 * RN;
            for (int x = 0; x < RX; ++x)
                for (int z = 0; z < RZ; ++z) {
                    const long i = r.at(x0 + x, z0 + z), q = (long)x * RZ + z;
                    fvx[i] = o[q];
                    fvz[i] = o[RN + q];
                }
#pragma omp parallel for schedule(static) num_threads(r.nth)
            for (int ix = 0; ix < g.nx; ++ix)
                for (int iz = 0; iz < g.nz; ++iz) {
                    const long p = r.at(ix + g.pml, iz + g.pml);
                    const long q = (long)(c->left + ix) * g.nz + iz;
                    const T gx = d_minus(fvx.data(), p, (long)g.nzp) / dx;   // the stress update's derivatives
                    const T gz = d_minus(fvz.data(), p, 1L) / dz;
                    const T gxz = d_plus(fvx.data(), p, 1L) / dz;
                    const T gzx = d_plus(fvz.data(), p, (long)g.nzp) / dx;
                    const T lg = -((gx + gz) * (bk.txx[p] + bk.tzz[p]));
                    const T mg = -2 * (gx * bk.txx[p] + gz * bk.tzz[p]) - (gxz + gzx) * bk.txz[p];
                    ivp[q] += 2 * r.m.vp[p] * r.m.rho[p] * lg;
                    ivs[q] += -4 * r.m.vs[p] * r.m.rho[p] * lg + 2 * r.m.vs[p] * r.m.rho[p] * mg;
                    ill[q] += fvx[p] * fvx[p] + fvz[p] * fvz[p];
                }
        }
        adjoint_s(g, r.st, r.m, bk, +1, r.nth);
    }
    const T wk = (T)K * r.st.dt;
    for (size_t q = 0; q < nm; ++q) {
        gvp[q] += wk * ivp[q];
        gvs[q] += wk * ivs[q];
        if (gil) gil[q] += (T)K * ill[q];
    }
}

extern "C" {
void or_gradient_consistent_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                                const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                                double *gvp, double *gvs, double *grho, double *misfit) {
    gradient_consistent_impl<double>(c, vp, vs, rho, wav, gc, obs_z, obs_x, gvp, gvs, grho, misfit);
}
void or_gradient_snapshot_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                              const float *dvp, const float *dvs, const float *wav, const int *gc,
                              const float *obs_z, const float *obs_x, const float *adj_z, const float *adj_x,
                              double *gvp, double *gvs, double *gil, double *misfit) {
    gradient_snapshot_impl<double>(c, vp, vs, rho, dvp, dvs, wav, gc, obs_z, obs_x, adj_z, adj_x, gvp, gvs, gil,
                                   misfit);
}
void or_hessian_vector_snapshot_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                                    const float *p_vp, const float *p_vs, const float *wav, const int *gc,
                                    double *gvp, double *gvs) {
    double unused = 0.0;
    gradient_snapshot_impl<double>(c, vp, vs, rho, p_vp, p_vs, wav, gc, nullptr, nullptr, nullptr, nullptr, gvp,
                                   gvs, nullptr, &unused, true);
}
}  // extern "C"

extern "C" {
void or_forward_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                    const float *wav, const int *gc, float *syn_z, float *syn_x) {
    forward_impl<float>(c, vp, vs, rho, wav, gc, syn_z, syn_x);
}
void or_forward_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                    const float *wav, const int *gc, double *syn_z, double *syn_x) {
    forward_impl<double>(c, vp, vs, rho, wav, gc, syn_z, syn_x);
}

void or_gradient_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                     const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                     float *gvp, float *gvs, float *ill, double *misfit, float *res_z, float *res_x) {
    gradient_impl<float, float>(c, vp, vs, rho, wav, gc, obs_z, obs_x, gvp, gvs, ill, misfit, res_z, res_x);
}
void or_gradient_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                     const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                     double *gvp, double *gvs, double *ill, double *misfit, double *res_z, double *res_x) {
    gradient_impl<double, double>(c, vp, vs, rho, wav, gc, obs_z, obs_x, gvp, gvs, ill, misfit, res_z, res_x);
}
void or_gradient_rho_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                         float *gvp, float *gvs, float *ill, float *grho, double *misfit) {
    gradient_impl<float, float>(c, vp, vs, rho, wav, gc, obs_z, obs_x, gvp, gvs, ill, misfit, nullptr, nullptr,
                                grho);
}
void or_gradient_rho_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                         double *gvp, double *gvs, double *ill, double *grho, double *misfit) {
    gradient_impl<double, double>(c, vp, vs, rho, wav, gc, obs_z, obs_x, gvp, gvs, ill, misfit, nullptr,
                                  nullptr, grho);
}
void or_gradient_res_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *wav, const int *gc, const float *res_z, const float *res_x,
                         float *gvp, float *gvs, float *ill) {
    gradient_res_impl<float>(c, vp, vs, rho, wav, gc, res_z, res_x, gvp, gvs, ill);
}
void or_gradient_res_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *wav, const int *gc, const float *res_z, const float *res_x,
                         double *gvp, double *gvs, double *ill) {
    gradient_res_impl<double>(c, vp, vs, rho, wav, gc, res_z, res_x, gvp, gvs, ill);
}

}  // extern "C"

template <typename T>
static void born_forward_impl(const or_config *c, const float *vp, const float *vs, const float *rho,
                              const float *dvp, const float *dvs, const float *wav, const int *gc,
                              T *b_z, T *b_x, T *bg_z, T *bg_x) {
    Run<T> r(c, vp, vs, rho, dvp, dvs, gc);
    const long nr = (long)c->ntr * c->nt;
    std::vector<T> bz(nr), bx(nr), gz(nr), gx(nr);
    r.born_forward(wav, bz, bx, gz, gx);
    std::memcpy(b_z, bz.data(), nr * sizeof(T));
    std::memcpy(b_x, bx.data(), nr * sizeof(T));
    std::memcpy(bg_z, gz.data(), nr * sizeof(T));
    std::memcpy(bg_x, gx.data(), nr * sizeof(T));
}

template <typename T>
static void born2_forward_impl(const or_config *c, const float *vp, const float *vs, const float *rho,
                               const float *dvp, const float *dvs, const float *wav, const int *gc,
                               T *b_z, T *b_x, T *b2_z, T *b2_x) {
    Run<T> r(c, vp, vs, rho, dvp, dvs, gc);
    const long nr = (long)c->ntr * c->nt;
    std::vector<T> bz(nr), bx(nr), gz(nr), gx(nr), cz(nr), cx(nr);
    r.born_forward(wav, bz, bx, gz, gx, nullptr, nullptr, c->pml, &cz, &cx);
    std::memcpy(b_z, bz.data(), nr * sizeof(T));
    std::memcpy(b_x, bx.data(), nr * sizeof(T));
    std::memcpy(b2_z, cz.data(), nr * sizeof(T));
    std::memcpy(b2_x, cx.data(), nr * sizeof(T));
}

extern "C" {
void or_born2_forward_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                          const float *dvp, const float *dvs, const float *wav, const int *gc,
                          float *b_z, float *b_x, float *b2_z, float *b2_x) {
    born2_forward_impl<float>(c, vp, vs, rho, dvp, dvs, wav, gc, b_z, b_x, b2_z, b2_x);
}
void or_born2_forward_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                          const float *dvp, const float *dvs, const float *wav, const int *gc,
                          double *b_z, double *b_x, double *b2_z, double *b2_x) {
    born2_forward_impl<double>(c, vp, vs, rho, dvp, dvs, wav, gc, b_z, b_x, b2_z, b2_x);
}
void or_hessian_vector_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                           const float *p_vp, const float *p_vs, const float *wav, const int *gc,
                           float *gvp, float *gvs) {
    hessian_impl<float, float>(c, vp, vs, rho, p_vp, p_vs, wav, gc, gvp, gvs);
}
void or_hessian_vector_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                           const float *p_vp, const float *p_vs, const float *wav, const int *gc,
                           double *gvp, double *gvs) {
    hessian_impl<double, double>(c, vp, vs, rho, p_vp, p_vs, wav, gc, gvp, gvs);
}
void or_born_forward_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *dvp, const float *dvs, const float *wav, const int *gc,
                         float *b_z, float *b_x, float *bg_z, float *bg_x) {
    born_forward_impl<float>(c, vp, vs, rho, dvp, dvs, wav, gc, b_z, b_x, bg_z, bg_x);
}
void or_born_forward_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *dvp, const float *dvs, const float *wav, const int *gc,
                         double *b_z, double *b_x, double *bg_z, double *bg_x) {
    born_forward_impl<double>(c, vp, vs, rho, dvp, dvs, wav, gc, b_z, b_x, bg_z, bg_x);
}

void or_born_gradient_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                          const float *dvp, const float *dvs, const float *wav, const int *gc,
                          const float *obs_z, const float *obs_x, float *gvp, float *gvs, float *ill,
                          double *misfit) {
    born_gradient_impl<float, float>(c, vp, vs, rho, dvp, dvs, wav, gc, obs_z, obs_x, gvp, gvs, ill, misfit);
}
void or_born_gradient_f64(const or_config *c, const float *vp, const float *vs, const float *rho,
                          const float *dvp, const float *dvs, const float *wav, const int *gc,
                          const float *obs_z, const float *obs_x, double *gvp, double *gvs, double *ill,
                          double *misfit) {
    born_gradient_impl<double, double>(c, vp, vs, rho, dvp, dvs, wav, gc, obs_z, obs_x, gvp, gvs, ill, misfit);
}

// Stage-1 line search: three forward solves at vel, vel-step1*d and
// vel-step2*d (test model formed in float on the host),
// sfwi-stage1/FD2DGPU.cu:4819-5028 (CALCULATE_ALPHA_parabola_mb).
void or_parabola_misfits_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                             const float *dir_vp, const float *dir_vs, float step1, float step2,
                             const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                             double m[3]) {
    const long nm = (long)c->allnx * c->allnz, nr = (long)c->ntr * c->nt;
    std::vector<float> tp(nm), ts(nm), sz(nr), sx(nr);
    const float steps[3] = {0.0f, step1, step2};
    for (int k = 0; k < 3; ++k) {
        const float *mp = vp, *ms = vs;
        if (k > 0) {
            for (long i = 0; i < nm; ++i) {
                tp[i] = vp[i] - steps[k] * dir_vp[i];
                ts[i] = vs[i] - steps[k] * dir_vs[i];
            }
            mp = tp.data();
            ms = ts.data();
        }
        forward_impl<float>(c, mp, ms, rho, wav, gc, sz.data(), sx.data());
        std::vector<float> rz(nr), rx(nr);
        for (long i = 0; i < nr; ++i) {
            rz[i] = 1.0f * sz[i] - obs_z[i];
            rx[i] = 1.0f * sx[i] - obs_x[i];
        }
        m[k] += l2_sum(rz, rx, c->ntr, c->nt);
    }
}

// Stage-2 exact step: sfwi-stage2/FD2DGPU.cu:4576-4722 (CALCULATE_ALPHA).
void or_linear_alpha_f32(const or_config *c, const float *vp, const float *vs, const float *rho,
                         const float *dvp, const float *dvs, const float *cg_vp, const float *cg_vs,
                         const float *wav, const int *gc, const float *obs_z, const float *obs_x,
                         double *fenzi, double *fenmu) {
    const long nr = (long)c->ntr * c->nt;
    std::vector<float> bz(nr), bx(nr), gz(nr), gx(nr), lz(nr), lx(nr);
    born_forward_impl<float>(c, vp, vs, rho, dvp, dvs, wav, gc, bz.data(), bx.data(), gz.data(), gx.data());
    for (long k = 0; k < nr; ++k) {
        bz[k] = bz[k] - obs_z[k];
        bx[k] = bx[k] - obs_x[k];
        bz[k] = 1.0f * bz[k] + gz[k];
        bx[k] = 1.0f * bx[k] + gx[k];
    }
    born_forward_impl<float>(c, vp, vs, rho, cg_vp, cg_vs, wav, gc, lz.data(), lx.data(), gz.data(), gx.data());
    for (int i = 0; i < c->ntr; ++i)
        for (int j = 0; j < c->nt; ++j) {
            long k = (long)i * c->nt + j;
            *fenmu += (lz[k] * lz[k] + lx[k] * lx[k]);
            *fenzi += (lz[k] * bz[k] + lx[k] * bx[k]);
        }
}

void or_stencils_f64(int nxpml, int nzpml, const double *f, double *dxp, double *dxm, double *dzp,
                     double *dzm) {
    const long S = nzpml;
    for (int ix = NOP; ix < nxpml - NOP; ++ix)
        for (int iz = NOP; iz < nzpml - NOP; ++iz) {
            long i = (long)ix * S + iz;
            dxp[i] = d_plus(f, i, S);
            dxm[i] = d_minus(f, i, S);
            dzp[i] = d_plus(f, i, 1);
            dzm[i] = d_minus(f, i, 1);
        }
}

void or_propagate_f64(const or_config *c, const float *vp, const float *vs, const float *rho, int kind,
                      int dir, int nsteps, double *state) {
    Run<double> r(c, vp, vs, rho, nullptr, nullptr, nullptr);
    Wave<double> w(r.g.n);
    auto arrs = w.all();
    for (size_t k = 0; k < arrs.size(); ++k)
        std::memcpy(arrs[k]->data(), state + (long)k * r.g.n, r.g.n * sizeof(double));
    for (int s = 0; s < nsteps; ++s) {
        if (kind == 0) {
            update_v(r.g, r.st, r.m, w, dir, r.nth);
            update_s(r.g, r.st, r.m, w, dir, r.nth);
        } else if (kind == 1) {
            update_s(r.g, r.st, r.m, w, dir, r.nth);
            update_v(r.g, r.st, r.m, w, dir, r.nth);
        } else if (kind == 2) {
            adjoint_s(r.g, r.st, r.m, w, dir, r.nth);
            adjoint_v(r.g, r.st, r.m, w, dir, r.nth);
        } else if (kind == 3) {
            update_v(r.g, r.st, r.m, w, dir, r.nth);
        } else if (kind == 4) {
            update_s(r.g, r.st, r.m, w, dir, r.nth);
        } else if (kind == 5) {
            adjoint_s(r.g, r.st, r.m, w, dir, r.nth);
        } else {
            adjoint_v(r.g, r.st, r.m, w, dir, r.nth);
        }
    }
    for (size_t k = 0; k < arrs.size(); ++k)
        std::memcpy(state + (long)k * r.g.n, arrs[k]->data(), r.g.n * sizeof(double));
}

}  // extern "C"
