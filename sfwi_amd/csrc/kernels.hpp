// Device code of the SFWI per-shot elastic path for gfx950 (MI355X).
//
// Data layout (HBM): every padded field is nxpml*nzpml floats, index
// ix*nzpml + iz (z fastest, the reference's layout), one contiguous block per
// shot slot.  A batch of shots is one launch: blockIdx.y selects the shot,
// blockIdx.x a run of 256 consecutive cells of the update domain
// [4, nxpml-4) x [4, nzpml-4) in z-fastest order, so every load that is not a
// stencil tap is a fully coalesced 1 KiB wave access and the x taps are the
// same runs shifted by whole columns.  Workgroups are remapped so that each of
// the 8 XCDs owns a contiguous eighth of every shot's cells: the medium arrays
// shared by all shots of a batch then stay in that XCD's L2.
//
// Two numerical modes share the kernels:
//   EXACT = true : every cell keeps the reference's split PML fields and the
//                  reference's operation order, including the FP64 promotion
//                  of the velocity updates (1.0/rho) and no FMA contraction
//                  (the translation unit is built with -ffp-contract=off).
//   EXACT = false: split fields only where the reference's split state can
//                  differ from the total (the damped PML, the 4-cell strips
//                  that the reconstruction overwrites, the source cell for the
//                  stresses and the receiver row for the adjoint velocity);
//                  everywhere else a cell carries only the five totals, which
//                  is the same update without the split.  FP32 with explicit
//                  FMAs in the stencil sums.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace sfwi {

constexpr int NOP = 4;        // stencil half width (reference nop)
constexpr int NSTRIP = 4;     // strip half width (reference N = nop)
constexpr int BLK = 256;      // threads per workgroup

// 8th-order staggered coefficients, float as in sfwi-stage1/FD2DGPU.cuh:16.
__constant__ float kC[5] = {0.0f, (float)(1225.0 / 1024.0), (float)(-245.0 / 3072.0),
                            (float)(49.0 / 5120.0), (float)(-5.0 / 7168.0)};

// Device pointer into global memory (HBM).  The field pointers reach the
// kernels inside descriptors read from memory, where the compiler cannot tell
// their address space and would emit flat instructions (which also count
// against the scalar-memory wait counter, so every scalar-load wait would
// drain the vector loads in flight).  Indexing through this wrapper emits
// global loads and stores with the uniform base in SGPRs and a 32-bit
// unsigned byte offset: indices are non-negative and below 2^30.
template <class T> struct gptr {
    T *p;
    gptr() = default;
    __host__ __device__ gptr(T *q) : p(q) {}
    template <class U, class = typename std::enable_if<std::is_convertible<U *, T *>::value>::type>
    __host__ __device__ gptr(const gptr<U> &o) : p(o.p) {}
    __host__ __device__ operator T *() const { return p; }
    __host__ __device__ gptr operator+(size_t n) const { return gptr(p + n); }
    using G = T __attribute__((address_space(1)));
    using GB = typename std::conditional<std::is_const<T>::value, const char, char>::type
        __attribute__((address_space(1)));
    __device__ G &operator[](unsigned i) const { return *(G *)((GB *)p + i * (unsigned)sizeof(T)); }
};
using fptr = gptr<float>;
using cfptr = gptr<const float>;

// Pointers of one propagator's fields.
struct Fields {
    fptr vx, vz, txx, tzz, txz;
    fptr vx_x, vx_z, vz_x, vz_z;
    fptr txx_x, txx_z, tzz_x, tzz_z, txz_x, txz_z;
};

// Padded medium of one shot window.
struct Medium {
    cfptr vp, vs, rho, lam, mu, dampx, dampz;
    cfptr bu;          // 1/rho (FP32 fast mode)
    cfptr icp, ics;    // imaging coefficients 2 vp rho, 2 vs rho (fast mode)
    cfptr dvp, dvs;    // padded perturbation (Born)
};

// Per-shot device descriptor, read with scalar loads (uniform per block).
struct DevShot {
    Medium m;
    Fields fw, bk, bn;                     // forward/reconstructed, adjoint, Born scattered
    Fields fw2;                            // second buffer of the forward state for the fused step march:
                                           // its own totals and velocity split parts, fw's stress split parts
    Fields bn2;                            // second-order scattered field (double scattering)
    Fields bk2;                            // second buffer of the adjoint state for the fused adjoint step of
                                           // the snapshot gradient (totals and velocity split parts; bk's
                                           // stress split parts)
    cfptr wav;                             // the shot's own wavelet [nt] (NULL: the context's)
    fptr dgx, dgz, dgxz, dgzx;             // velocity derivatives for imaging
    fptr rz, rx;                           // synthetic (then residual) traces [ntr][nt]
    fptr r2z, r2x;                         // background traces of a Born solve
    cfptr oz, ox;                          // observed traces
    gptr<const int> igx;                   // receiver columns [ntr]
    gptr<const int> inj;                   // traces grouped by receiver cell, trace order within a cell [ntr]
    gptr<const int> inj_off;               // group k = inj[inj_off[k] .. inj_off[k+1]) [ninj + 1]
    gptr<const int> rtr;                   // trace of each column of the receiver row, -1 none, -2 several [nxp]
    fptr ainj;                             // adjoint source by step and column, [nt][2 (z, x)][nxp] (snapshot
                                           // gradients whose receivers have a cell each)
    fptr strip;                            // boundary strips, [nt][2 fields][SF] (FP32 storage)
    fptr stage;                            // one step's strips [2][SF] before packing (16-bit storage)
    gptr<unsigned short> strip_h;          // packed strips [nt][2][SF] (FP16 / BF16 storage)
    fptr strip_scale;                      // max-abs scale per step and field [nt][2]
    fptr snap32;                           // velocity snapshots, FP32 [nsnap][2][snap_n]
    gptr<unsigned short> snap16;           // velocity snapshots, FP16 / BF16 of v / scale [nsnap][2][snap_n]
    gptr<unsigned> snap_max;               // max |v| per snapshot and field as FP32 bits [nsnap][2]
    fptr snap_part;                        // partial maxima of the snapshot being stored [2][SNAP_PART]
    fptr gvp, gvs, gil;                    // per-shot gradient accumulators [nx][nz]
    fptr grho;                             // per-shot density gradient [nx][nz] (grad_rho)
    fptr vold_x, vold_z;                   // v(it+1) before the reconstruction velocity step (grad_rho)
    gptr<double> trace_sum;                // per-trace partial misfit [ntr]
    int isx, isz, igz, ntr, ninj, left;
};

// Geometry shared by every shot of a batch.
struct Geo {
    int nxp, nzp, pml, nx, nz, nt, allnz;
    int nzd, ndom;          // update domain height and cell count
    int per_xcd;            // blocks per XCD in the remapped grid
    int cm_xs;              // column-march segment width (FAST kernels)
    int cm_nbands, cm_waves;// row bands per column strip, waves per shot
    int cm_per_xcd;         // column-march workgroups per XCD
    int cm_order;           // 0: grid.y = shot; 1: 1-D grid, shots of a tile adjacent; 2: 1-D, whole shots per XCD
    int cm_nshots;          // shots of the launch (cm_order 1, 2)
    int cm_wg;              // column-march workgroups per shot
    int sf;                 // strip floats per field and step
    int store;              // strip storage: 0 FP32, 1 FP16, 2 BF16 (per-step max-abs scale)
    int grad_rho;           // 1: also image the density
    int snap_k;             // snapshot reconstruction: steps between snapshots (0: strips)
    int snap_rz, snap_n;    // snapshot region height nz + 8 and cells (nx + 8)(nz + 8), origin (pml-4, pml-4)
    int im_nbands, im_waves, im_wg, im_per_xcd;   // snapshot imaging march: row bands, waves and workgroups
                                                  // per shot over the interior, workgroups per XCD (launch)
    int fs_xs, fs_nbands, fs_waves, fs_wg, fs_per_xcd;   // fused step march: segment width, 48-row bands,
                                                         // waves and workgroups per shot, per XCD (launch)
    int ws_xs, ws_ntiles, ws_wg, ws_per_xcd;   // workgroup fused step: segment width, 240-row tiles,
                                               // workgroups per shot, per XCD (launch)
    int snap_np;            // partial maxima per field and shot of a snapshot being stored (stride)
    float inv_nzd;
    float dt, dx, dz, dtdx, dtdz, idx, idz;
};

// ------------------------------------------------------------- helpers ---

// Cell of this thread in the update domain, XCD-aware block order.
__device__ inline bool cell_of(const Geo &g, int &ix, int &iz) {
    const int b = blockIdx.x;
    const int L = (b & 7) * g.per_xcd + (b >> 3);
    const int l = L * BLK + (int)threadIdx.x;
    if (l >= g.ndom) return false;
    int x = (int)(((float)l + 0.5f) * g.inv_nzd);
    int z = l - x * g.nzd;
    if (z < 0) { --x; z += g.nzd; } else if (z >= g.nzd) { ++x; z -= g.nzd; }
    ix = x + NOP;
    iz = z + NOP;
    return true;
}

// Cells whose split state must be kept in fast mode: within pml+4 of an edge
// (damped PML plus the strip band of the reconstruction).
__device__ inline bool in_band(const Geo &g, int ix, int iz) {
    const int w = g.pml + NSTRIP;
    return ix < w || ix >= g.nxp - w || iz < w || iz >= g.nzp - w;
}

// sum_k c_k (f[i+k s] - f[i-(k-1) s])   taps i-3..i+4
template <bool EXACT, class P> __device__ inline float d_plus(const P &f, int i, int s) {
    if (EXACT) {
        float t = 0.0f;
        t += kC[1] * (f[i + s] - f[i]);
        t += kC[2] * (f[i + 2 * s] - f[i - s]);
        t += kC[3] * (f[i + 3 * s] - f[i - 2 * s]);
        t += kC[4] * (f[i + 4 * s] - f[i - 3 * s]);
        return t;
    } else {
        float t = kC[1] * (f[i + s] - f[i]);
        t = __builtin_fmaf(kC[2], f[i + 2 * s] - f[i - s], t);
        t = __builtin_fmaf(kC[3], f[i + 3 * s] - f[i - 2 * s], t);
        t = __builtin_fmaf(kC[4], f[i + 4 * s] - f[i - 3 * s], t);
        return t;
    }
}
// sum_k c_k (f[i+(k-1) s] - f[i-k s])   taps i-4..i+3
template <bool EXACT, class P> __device__ inline float d_minus(const P &f, int i, int s) {
    if (EXACT) {
        float t = 0.0f;
        t += kC[1] * (f[i] - f[i - s]);
        t += kC[2] * (f[i + s] - f[i - 2 * s]);
        t += kC[3] * (f[i + 2 * s] - f[i - 3 * s]);
        t += kC[4] * (f[i + 3 * s] - f[i - 4 * s]);
        return t;
    } else {
        float t = kC[1] * (f[i] - f[i - s]);
        t = __builtin_fmaf(kC[2], f[i + s] - f[i - 2 * s], t);
        t = __builtin_fmaf(kC[3], f[i + 2 * s] - f[i - 3 * s], t);
        t = __builtin_fmaf(kC[4], f[i + 3 * s] - f[i - 4 * s], t);
        return t;
    }
}

// Damping factors of one cell: d1 = 1 - dt*a/2, d2 = 1 + dt*a/2 (reference order).
struct Damp {
    float d1, d2, d3, d4;
};
__device__ inline Damp damp_of(const Geo &g, const Medium &m, int i) {
    const float ax = m.dampx[i], az = m.dampz[i];
    Damp d;
    d.d1 = 1.0f - g.dt * ax / 2.0f;
    d.d2 = 1.0f + g.dt * ax / 2.0f;
    d.d3 = 1.0f - g.dt * az / 2.0f;
    d.d4 = 1.0f + g.dt * az / 2.0f;
    return d;
}

// --------------------------------------------------- one-cell updates ---

__device__ inline void vel_update_fast(const Geo &g, const Medium &m, const Fields &F, int i, bool split,
                                       int dir, float a, float b, float c, float e);

// Velocity update (sfwi-stage1/born.cu:355-394, update_vx_vz).
template <bool EXACT>
__device__ inline void vel_cell(const Geo &g, const Medium &m, const Fields &F, int i, bool split,
                                int dir) {
    const int S = g.nzp;
    const float a = d_plus<EXACT>(F.txx, i, S);   // Dx+ txx
    const float b = d_minus<EXACT>(F.txz, i, 1);  // Dz- txz
    const float c = d_plus<EXACT>(F.tzz, i, 1);   // Dz+ tzz
    const float e = d_minus<EXACT>(F.txz, i, S);  // Dx- txz
    if (EXACT) {
        const Damp d = damp_of(g, m, i);
        const double r = 1.0 / (double)m.rho[i];
        const float vxx = (float)(((double)(d.d1 * F.vx_x[i]) + dir * r * (double)g.dtdx * (double)a) / (double)d.d2);
        const float vxz = (float)(((double)(d.d3 * F.vx_z[i]) + dir * r * (double)g.dtdz * (double)b) / (double)d.d4);
        const float vzx = (float)(((double)(d.d1 * F.vz_x[i]) + dir * r * (double)g.dtdx * (double)e) / (double)d.d2);
        const float vzz = (float)(((double)(d.d3 * F.vz_z[i]) + dir * r * (double)g.dtdz * (double)c) / (double)d.d4);
        F.vx_x[i] = vxx;
        F.vx_z[i] = vxz;
        F.vz_x[i] = vzx;
        F.vz_z[i] = vzz;
        F.vx[i] = vxx + vxz;
        F.vz[i] = vzx + vzz;
    } else {
        vel_update_fast(g, m, F, i, split, dir, a, b, c, e);
    }
}

// Velocity derivatives used by the stress update.
struct VGrad {
    float xx, zz, xz, zx;   // Dx- vx, Dz- vz, Dz+ vx, Dx+ vz
};
template <bool EXACT> __device__ inline VGrad vgrad(const Geo &g, const Fields &F, int i) {
    VGrad v;
    v.xx = d_minus<EXACT>(F.vx, i, g.nzp);
    v.zz = d_minus<EXACT>(F.vz, i, 1);
    v.xz = d_plus<EXACT>(F.vx, i, 1);
    v.zx = d_plus<EXACT>(F.vz, i, g.nzp);
    return v;
}

// Born source terms of one cell from the background velocity derivatives
// (sfwi-stage2/born.cu:787-797, add_born_source_components).
struct BornSrc {
    float xx_x, xx_z, zz_x, zz_z, xz_x, xz_z;
};
template <bool EXACT>
__device__ inline BornSrc born_src(const Geo &g, const Medium &m, const VGrad &v, int i, int dir) {
    const float rho = m.rho[i], vp = m.vp[i], vs = m.vs[i], dp = m.dvp[i], ds = m.dvs[i];
    const float D = (float)dir;
    BornSrc s;
    if (EXACT) {
        s.xx_x = D * g.dt * (2.0f * (rho * vp) * (dp) * (v.xx) / g.dx);
        s.xx_z = D * g.dt * ((2.0f * (rho * vp) * (dp) - 4.0f * (rho * vs) * (ds)) * (v.zz) / g.dz);
        s.zz_x = D * g.dt * ((2.0f * (rho * vp) * (dp) - 4.0f * (rho * vs) * (ds)) * (v.xx) / g.dx);
        s.zz_z = D * g.dt * (2.0f * (rho * vp) * (dp) * (v.zz) / g.dz);
        s.xz_x = D * g.dt * ((2.0f * (rho * vs) * (ds)) * ((v.zx) / g.dx));
        s.xz_z = D * g.dt * ((2.0f * (rho * vs) * (ds)) * ((v.xz) / g.dz));
    } else {
        const float P = 2.0f * (rho * vp) * dp;
        const float Sv = 2.0f * (rho * vs) * ds;
        const float Q = __builtin_fmaf(-2.0f, Sv, P);
        const float kx = D * g.dt * g.idx, kz = D * g.dt * g.idz;
        s.xx_x = kx * P * v.xx;
        s.xx_z = kz * Q * v.zz;
        s.zz_x = kx * Q * v.xx;
        s.zz_z = kz * P * v.zz;
        s.xz_x = kx * Sv * v.zx;
        s.xz_z = kz * Sv * v.xz;
    }
    return s;
}

// ------------------------------------------------- FAST-mode updates ---
// The update of one cell given its stencil values, split into the loads of
// the cell's own state (a marching kernel issues them one column ahead), the
// arithmetic and the stores.  Split cells follow the reference's split-PML
// form; unsplit cells advance the totals directly.

__device__ inline Damp damp_v(const Geo &g, float ax, float az) {
    Damp d;
    d.d1 = 1.0f - g.dt * ax / 2.0f;
    d.d2 = 1.0f + g.dt * ax / 2.0f;
    d.d3 = 1.0f - g.dt * az / 2.0f;
    d.d4 = 1.0f + g.dt * az / 2.0f;
    return d;
}

// Own state of a velocity update (forward, reconstruction or adjoint).
struct VelCell {
    float vx, vz, vx_x, vx_z, vz_x, vz_z, bu, ax, az;
};
__device__ inline VelCell vel_load(const Medium &m, const Fields &F, int i, bool split) {
    VelCell s;
    s.bu = m.bu[i];
    if (split) {
        s.vx_x = F.vx_x[i]; s.vx_z = F.vx_z[i]; s.vz_x = F.vz_x[i]; s.vz_z = F.vz_z[i];
        s.ax = m.dampx[i]; s.az = m.dampz[i];
        s.vx = 0.f; s.vz = 0.f;
    } else {
        s.vx = F.vx[i]; s.vz = F.vz[i];
        s.vx_x = s.vx_z = s.vz_x = s.vz_z = s.ax = s.az = 0.f;
    }
    return s;
}
// a, b, c, e drive vx_x, vx_z, vz_z, vz_x: forward a = Dx+ txx, b = Dz- txz,
// c = Dz+ tzz, e = Dx- txz; adjoint a = Dx+ p1, b = Dz- p2, c = Dz+ p3, e = Dx- p2.
__device__ inline void vel_math(const Geo &g, VelCell &s, bool split, int dir, float a, float b, float c,
                                float e) {
    const float bu = (float)dir * s.bu;
    if (split) {
        const Damp d = damp_v(g, s.ax, s.az);
        const float bx = bu * g.dtdx, bz = bu * g.dtdz;
        s.vx_x = __builtin_fmaf(d.d1, s.vx_x, bx * a) / d.d2;
        s.vx_z = __builtin_fmaf(d.d3, s.vx_z, bz * b) / d.d4;
        s.vz_x = __builtin_fmaf(d.d1, s.vz_x, bx * e) / d.d2;
        s.vz_z = __builtin_fmaf(d.d3, s.vz_z, bz * c) / d.d4;
        s.vx = s.vx_x + s.vx_z;
        s.vz = s.vz_x + s.vz_z;
    } else {
        s.vx = __builtin_fmaf(bu, __builtin_fmaf(g.dtdx, a, g.dtdz * b), s.vx);
        s.vz = __builtin_fmaf(bu, __builtin_fmaf(g.dtdx, e, g.dtdz * c), s.vz);
    }
}
__device__ inline void vel_store(const Fields &F, int i, const VelCell &s, bool split) {
    if (split) {
        F.vx_x[i] = s.vx_x; F.vx_z[i] = s.vx_z; F.vz_x[i] = s.vz_x; F.vz_z[i] = s.vz_z;
    }
    F.vx[i] = s.vx;
    F.vz[i] = s.vz;
}
__device__ inline void vel_update_fast(const Geo &g, const Medium &m, const Fields &F, int i, bool split,
                                       int dir, float a, float b, float c, float e) {
    VelCell s = vel_load(m, F, i, split);
    vel_math(g, s, split, dir, a, b, c, e);
    vel_store(F, i, s, split);
}

// Own state of a stress update.
struct StrCell {
    float txx, tzz, txz, xx_x, xx_z, zz_x, zz_z, xz_x, xz_z, lam, mu, ax, az;
};
__device__ inline StrCell str_load(const Medium &m, const Fields &F, int i, bool split) {
    StrCell s;
    s.lam = m.lam[i];
    s.mu = m.mu[i];
    if (split) {
        s.xx_x = F.txx_x[i]; s.xx_z = F.txx_z[i]; s.zz_x = F.tzz_x[i];
        s.zz_z = F.tzz_z[i]; s.xz_x = F.txz_x[i]; s.xz_z = F.txz_z[i];
        s.ax = m.dampx[i]; s.az = m.dampz[i];
        s.txx = s.tzz = s.txz = 0.f;
    } else {
        s.txx = F.txx[i]; s.tzz = F.tzz[i]; s.txz = F.txz[i];
        s.xx_x = s.xx_z = s.zz_x = s.zz_z = s.xz_x = s.xz_z = s.ax = s.az = 0.f;
    }
    return s;
}
template <bool HAS_SRC>
__device__ inline void str_math(const Geo &g, StrCell &s, const VGrad &v, bool split, int dir,
                                const BornSrc &src) {
    const float lam = s.lam, mu = s.mu;
    const float D = (float)dir;
    const float l2m = __builtin_fmaf(2.0f, mu, lam);
    const float ax = D * g.dtdx * v.xx, az = D * g.dtdz * v.zz;
    const float gx = D * g.dtdx * v.zx, gz = D * g.dtdz * v.xz;
    if (split) {
        const Damp d = damp_v(g, s.ax, s.az);
        if (HAS_SRC) {
            s.xx_x += src.xx_x; s.xx_z += src.xx_z; s.zz_x += src.zz_x;
            s.zz_z += src.zz_z; s.xz_x += src.xz_x; s.xz_z += src.xz_z;
        }
        s.xx_x = __builtin_fmaf(d.d1, s.xx_x, l2m * ax) / d.d2;
        s.xx_z = __builtin_fmaf(d.d3, s.xx_z, lam * az) / d.d4;
        s.zz_x = __builtin_fmaf(d.d1, s.zz_x, lam * ax) / d.d2;
        s.zz_z = __builtin_fmaf(d.d3, s.zz_z, l2m * az) / d.d4;
        s.xz_x = __builtin_fmaf(d.d1, s.xz_x, mu * gx) / d.d2;
        s.xz_z = __builtin_fmaf(d.d3, s.xz_z, mu * gz) / d.d4;
        s.txx = s.xx_x + s.xx_z;
        s.tzz = s.zz_x + s.zz_z;
        s.txz = s.xz_x + s.xz_z;
    } else {
        float sxx = __builtin_fmaf(l2m, ax, lam * az);
        float szz = __builtin_fmaf(lam, ax, l2m * az);
        float sxz = mu * (gx + gz);
        if (HAS_SRC) {
            sxx += src.xx_x + src.xx_z;
            szz += src.zz_x + src.zz_z;
            sxz += src.xz_x + src.xz_z;
        }
        s.txx += sxx;
        s.tzz += szz;
        s.txz += sxz;
    }
}
__device__ inline void str_store(const Fields &F, int i, const StrCell &s, bool split) {
    if (split) {
        F.txx_x[i] = s.xx_x; F.txx_z[i] = s.xx_z; F.tzz_x[i] = s.zz_x;
        F.tzz_z[i] = s.zz_z; F.txz_x[i] = s.xz_x; F.txz_z[i] = s.xz_z;
    }
    F.txx[i] = s.txx;
    F.tzz[i] = s.tzz;
    F.txz[i] = s.txz;
}
template <bool HAS_SRC>
__device__ inline void str_update_fast(const Geo &g, const Medium &m, const Fields &F, const VGrad &v, int i,
                                       bool split, int dir, const BornSrc &s) {
    StrCell c = str_load(m, F, i, split);
    str_math<HAS_SRC>(g, c, v, split, dir, s);
    str_store(F, i, c, split);
}

// Own state of the adjoint stress-like update (txx = txx_x, tzz = tzz_z).
struct AdjStrCell {
    float txx, tzz, txz, xx_x, zz_z, xz_x, xz_z, ax, az;
};
__device__ inline AdjStrCell adj_str_load(const Medium &m, const Fields &F, int i, bool split) {
    AdjStrCell s;
    if (split) {
        s.xx_x = F.txx_x[i]; s.zz_z = F.tzz_z[i]; s.xz_x = F.txz_x[i]; s.xz_z = F.txz_z[i];
        s.ax = m.dampx[i]; s.az = m.dampz[i];
        s.txx = s.tzz = s.txz = 0.f;
    } else {
        s.txx = F.txx[i]; s.tzz = F.tzz[i]; s.txz = F.txz[i];
        s.xx_x = s.zz_z = s.xz_x = s.xz_z = s.ax = s.az = 0.f;
    }
    return s;
}
__device__ inline void adj_str_math(const Geo &g, AdjStrCell &s, const VGrad &v, bool split, int dir) {
    const float D = (float)dir;
    if (split) {
        const Damp d = damp_v(g, s.ax, s.az);
        s.xx_x = __builtin_fmaf(d.d1, s.xx_x, D * g.dtdx * v.xx) / d.d2;
        s.zz_z = __builtin_fmaf(d.d3, s.zz_z, D * g.dtdz * v.zz) / d.d4;
        s.xz_x = __builtin_fmaf(d.d1, s.xz_x, D * g.dtdx * v.zx) / d.d2;
        s.xz_z = __builtin_fmaf(d.d3, s.xz_z, D * g.dtdz * v.xz) / d.d4;
        s.txx = s.xx_x;
        s.tzz = s.zz_z;
        s.txz = s.xz_x + s.xz_z;
    } else {
        s.txx = __builtin_fmaf(D * g.dtdx, v.xx, s.txx);
        s.tzz = __builtin_fmaf(D * g.dtdz, v.zz, s.tzz);
        s.txz = __builtin_fmaf(D, __builtin_fmaf(g.dtdx, v.zx, g.dtdz * v.xz), s.txz);
    }
}
__device__ inline void adj_str_store(const Fields &F, int i, const AdjStrCell &s, bool split) {
    if (split) {
        F.txx_x[i] = s.xx_x; F.tzz_z[i] = s.zz_z; F.txz_x[i] = s.xz_x; F.txz_z[i] = s.xz_z;
    }
    F.txx[i] = s.txx;
    F.tzz[i] = s.tzz;
    F.txz[i] = s.txz;
}
__device__ inline void adj_str_update_fast(const Geo &g, const Medium &m, const Fields &F, const VGrad &v,
                                           int i, bool split, int dir) {
    AdjStrCell s = adj_str_load(m, F, i, split);
    adj_str_math(g, s, v, split, dir);
    adj_str_store(F, i, s, split);
}

// sx = Dx+[(lam+2mu) txx + lam tzz], sz = Dz+[(lam+2mu) tzz + lam txx],
// sxz_z = Dz-[mu txz], sxz_x = Dx-[mu txz]: the velocity update's form.
__device__ inline void adj_vel_update_fast(const Geo &g, const Medium &m, const Fields &F, int i, bool split,
                                           int dir, float sx, float sz, float sxz_z, float sxz_x) {
    vel_update_fast(g, m, F, i, split, dir, sx, sxz_z, sz, sxz_x);
}

// Products whose derivatives the adjoint velocity update takes.
__device__ inline float adj_p1(float lam, float mu, float txx, float tzz) {
    return __builtin_fmaf(__builtin_fmaf(2.0f, mu, lam), txx, lam * tzz);
}

// Zero-lag imaging of one interior cell (q = its index in the per-shot image),
// FAST mode: the per-shot accumulators sum the time-independent factors'
// cofactors, lg = -(Dvx + Dvz)(sxx + szz) for Vp and -2 lg + mg for Vs;
// k_merge multiplies them by 2 vp rho and 2 vs rho once per shot (the
// reference multiplies every step, sfwi-stage1/FD2DGPU.cu:3350-3382: the same
// sum up to rounding).
__device__ inline void image_fast(const DevShot &d, int i, int q) {
    const float gx = d.dgx[i], gz = d.dgz[i], gxz = d.dgxz[i], gzx = d.dgzx[i];
    const float sxx = d.bk.txx[i], szz = d.bk.tzz[i], sxz = d.bk.txz[i];
    const float vx = d.fw.vx[i], vz = d.fw.vz[i];
    const float lg = -((gx + gz) * (sxx + szz));
    const float mg = __builtin_fmaf(-(gxz + gzx), sxz, -2.0f * __builtin_fmaf(gx, sxx, gz * szz));
    d.gvp[q] = d.gvp[q] + lg;
    d.gvs[q] = d.gvs[q] + __builtin_fmaf(-2.0f, lg, mg);
    d.gil[q] = __builtin_fmaf(vx, vx, __builtin_fmaf(vz, vz, d.gil[q]));
}

// Vp / Vs image cofactors of one cell (the arithmetic of image_fast; k_merge
// applies the coefficients 2 vp rho, 2 vs rho): velocity derivatives
// gx = Dx- vx / dx, gz = Dz- vz / dz, gxz = Dz+ vx / dz, gzx = Dx+ vz / dx,
// adjoint stresses and the sums so far; returns the updated sums in op, os.
__device__ inline void image_vpvs(float gx, float gz, float gxz, float gzx, float sxx, float szz, float sxz,
                                  float ip, float is, float &op, float &os) {
    const float lg = -((gx + gz) * (sxx + szz));
    const float mg = __builtin_fmaf(-(gxz + gzx), sxz, -2.0f * __builtin_fmaf(gx, sxx, gz * szz));
    op = ip + lg;
    os = is + __builtin_fmaf(-2.0f, lg, mg);
}

// Density term of one interior cell (FAST; both FAST kernel families call
// it with the same operands):  -(vx_adj dvx + vz_adj dvz) / dt
// + (vp^2 - 2 vs^2) lg + vs^2 mg, dv = v(it+1) - v(it) of the reconstruction.
__device__ inline float rho_term_fast(const Geo &g, float lg, float mg, float bvx, float bvz, float dvx,
                                      float dvz, float vp, float vs) {
    const float rd = -__builtin_fmaf(bvx, dvx, bvz * dvz) / g.dt;
    const float vs2 = vs * vs;
    return __builtin_fmaf(__builtin_fmaf(vp, vp, -2.0f * vs2), lg, __builtin_fmaf(vs2, mg, rd));
}

// Stress update (sfwi-stage1/born.cu:399-440, update_txx_tzz_txz); with
// a Born source the terms are added to the split parts first
// (sfwi-stage2/FD2DGPU.cu:1839-1840 order: source, then update).
template <bool EXACT, bool HAS_SRC>
__device__ inline void str_cell(const Geo &g, const Medium &m, const Fields &F, const VGrad &v, int i,
                                bool split, int dir, const BornSrc &s) {
    const float lam = m.lam[i], mu = m.mu[i];
    if (EXACT) {
        const Damp d = damp_of(g, m, i);
        const float D = (float)dir;
        float xx_x = F.txx_x[i], xx_z = F.txx_z[i], zz_x = F.tzz_x[i], zz_z = F.tzz_z[i];
        float xz_x = F.txz_x[i], xz_z = F.txz_z[i];
        if (HAS_SRC) {
            xx_x = xx_x + s.xx_x; xx_z = xx_z + s.xx_z; zz_x = zz_x + s.zz_x;
            zz_z = zz_z + s.zz_z; xz_x = xz_x + s.xz_x; xz_z = xz_z + s.xz_z;
        }
        xx_x = (d.d1 * xx_x + D * (lam + 2.0f * mu) * g.dtdx * v.xx) / d.d2;
        xx_z = (d.d3 * xx_z + D * lam * g.dtdz * v.zz) / d.d4;
        zz_x = (d.d1 * zz_x + D * lam * g.dtdx * v.xx) / d.d2;
        zz_z = (d.d3 * zz_z + D * (lam + 2.0f * mu) * g.dtdz * v.zz) / d.d4;
        xz_x = (d.d1 * xz_x + D * g.dtdx * mu * v.zx) / d.d2;
        xz_z = (d.d3 * xz_z + D * g.dtdz * mu * v.xz) / d.d4;
        F.txx_x[i] = xx_x; F.txx_z[i] = xx_z; F.tzz_x[i] = zz_x;
        F.tzz_z[i] = zz_z; F.txz_x[i] = xz_x; F.txz_z[i] = xz_z;
        F.txx[i] = xx_x + xx_z;
        F.tzz[i] = zz_x + zz_z;
        F.txz[i] = xz_x + xz_z;
    } else {
        str_update_fast<HAS_SRC>(g, m, F, v, i, split, dir, s);
    }
}

// Adjoint stress-like update (sfwi-stage1/born.cu:565-603):
// txx = txx_x, tzz = tzz_z, txz = txz_x + txz_z.
template <bool EXACT>
__device__ inline void adj_str_cell(const Geo &g, const Medium &m, const Fields &F, int i, bool split,
                                    int dir) {
    const VGrad v = vgrad<EXACT>(g, F, i);
    if (EXACT) {
        const float D = (float)dir;
        const Damp d = damp_of(g, m, i);
        const float xx = (d.d1 * F.txx_x[i] + D * g.dtdx * v.xx) / d.d2;
        const float zz = (d.d3 * F.tzz_z[i] + D * g.dtdz * v.zz) / d.d4;
        const float xzx = (d.d1 * F.txz_x[i] + D * g.dtdx * v.zx) / d.d2;
        const float xzz = (d.d3 * F.txz_z[i] + D * g.dtdz * v.xz) / d.d4;
        F.txx_x[i] = xx;
        F.tzz_z[i] = zz;
        F.txz_x[i] = xzx;
        F.txz_z[i] = xzz;
        F.txx[i] = xx;
        F.tzz[i] = zz;
        F.txz[i] = xzx + xzz;
    } else {
        adj_str_update_fast(g, m, F, v, i, split, dir);
    }
}

// Adjoint velocity update with the moduli inside the derivatives
// (sfwi-stage1/born.cu:504-562).
template <bool EXACT>
__device__ inline void adj_vel_cell(const Geo &g, const Medium &m, const Fields &F, int i, bool split,
                                    int dir) {
    const int S = g.nzp;
    const cfptr lam = m.lam;
    const cfptr mu = m.mu;
    const cfptr txx = F.txx;
    const cfptr tzz = F.tzz;
    const cfptr txz = F.txz;
    if (EXACT) {
        float s_txx_x = 0.f, s_txz_z = 0.f, s_tzz_x = 0.f, s_tzz_z = 0.f, s_txz_x = 0.f, s_txx_z = 0.f;
#pragma unroll
        for (int k = 1; k <= NOP; ++k) {
            const int pa = i + k * S, pb = i - (k - 1) * S;          // x taps of Dx+
            const int qa = i + (k - 1) * S, qb = i - k * S;          // x taps of Dx-
            const int za = i + k, zb = i - (k - 1);                  // z taps of Dz+
            const int wa = i + (k - 1), wb = i - k;                  // z taps of Dz-
            s_txx_x += kC[k] * ((lam[pa] + 2.0f * mu[pa]) * txx[pa] - (lam[pb] + 2.0f * mu[pb]) * txx[pb]);
            s_txz_z += kC[k] * (mu[wa] * txz[wa] - mu[wb] * txz[wb]);
            s_tzz_x += kC[k] * ((lam[pa]) * tzz[pa] - (lam[pb]) * tzz[pb]);
            s_tzz_z += kC[k] * ((lam[za] + 2.0f * mu[za]) * tzz[za] - (lam[zb] + 2.0f * mu[zb]) * tzz[zb]);
            s_txz_x += kC[k] * (mu[qa] * txz[qa] - mu[qb] * txz[qb]);
            s_txx_z += kC[k] * ((lam[za]) * txx[za] - (lam[zb]) * txx[zb]);
        }
        const Damp d = damp_of(g, m, i);
        const double r = 1.0 / (double)m.rho[i];
        const double ax = (double)g.dtdx, az = (double)g.dtdz;
        const float vxx = (float)(((double)(d.d1 * F.vx_x[i]) + dir * r * ax * (double)s_txx_x + dir * r * ax * (double)s_tzz_x) / (double)d.d2);
        const float vxz = (float)(((double)(d.d3 * F.vx_z[i]) + dir * r * az * (double)s_txz_z) / (double)d.d4);
        const float vzx = (float)(((double)(d.d1 * F.vz_x[i]) + dir * r * ax * (double)s_txz_x) / (double)d.d2);
        const float vzz = (float)(((double)(d.d3 * F.vz_z[i]) + dir * r * az * (double)s_tzz_z + dir * r * az * (double)s_txx_z) / (double)d.d4);
        F.vx_x[i] = vxx;
        F.vx_z[i] = vxz;
        F.vz_x[i] = vzx;
        F.vz_z[i] = vzz;
        F.vz[i] = vzx + vzz;
        F.vx[i] = vxx + vxz;
    } else {
        // Fast form: Dx+ of (lam+2mu) txx + lam tzz, Dz+ of (lam+2mu) tzz + lam txx.
        float sx = 0.f, sz = 0.f, sxz_z = 0.f, sxz_x = 0.f;
#pragma unroll
        for (int k = 1; k <= NOP; ++k) {
            const int pa = i + k * S, pb = i - (k - 1) * S;
            const int qa = i + (k - 1) * S, qb = i - k * S;
            const int za = i + k, zb = i - (k - 1);
            const int wa = i + (k - 1), wb = i - k;
            const float la = lam[pa], lb = lam[pb];
            const float fa = adj_p1(la, mu[pa], txx[pa], tzz[pa]);
            const float fb = adj_p1(lb, mu[pb], txx[pb], tzz[pb]);
            sx = __builtin_fmaf(kC[k], fa - fb, sx);
            const float lza = lam[za], lzb = lam[zb];
            const float ga = adj_p1(lza, mu[za], tzz[za], txx[za]);
            const float gb = adj_p1(lzb, mu[zb], tzz[zb], txx[zb]);
            sz = __builtin_fmaf(kC[k], ga - gb, sz);
            sxz_z = __builtin_fmaf(kC[k], mu[wa] * txz[wa] - mu[wb] * txz[wb], sxz_z);
            sxz_x = __builtin_fmaf(kC[k], mu[qa] * txz[qa] - mu[qb] * txz[qb], sxz_x);
        }
        adj_vel_update_fast(g, m, F, i, split, dir, sx, sz, sxz_z, sxz_x);
    }
}

// ------------------------------------------------------------ kernels ---

// 16-bit encodings of a value in [-1, 1]: FP16 (store 1) or BF16 (store 2),
// both rounded to nearest even.
__device__ inline unsigned short pack16(int store, float x) {
    if (store == 1) return __builtin_bit_cast(unsigned short, (_Float16)x);
    const unsigned u = __builtin_bit_cast(unsigned, x);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float unpack16(int store, unsigned short h) {
    if (store == 1) return (float)__builtin_bit_cast(_Float16, h);
    return __builtin_bit_cast(float, (unsigned)h << 16);
}

// Strip slots of a padded cell: up, down, left, right (-1 when outside).
// Layout per field and step: up [nx][2N] | down [nx][2N] | left [2N][nz] | right [2N][nz]
// (sfwi-stage1/FD2DGPU.cu:3842-3869 index formulas).
__device__ inline void strip_slots(const Geo &g, int ix, int iz, int s[4]) {
    const int N = NSTRIP, p = g.pml;
    s[0] = s[1] = s[2] = s[3] = -1;
    if (ix >= p && ix < g.nxp - p) {
        if (iz >= p - N && iz < p + N) s[0] = (ix - p) * 2 * N + (iz - (p - N));
        if (iz >= g.nzp - p - N && iz < g.nzp - p + N)
            s[1] = 2 * N * g.nx + (ix - p) * 2 * N + (iz - (g.nzp - p - N));
    }
    if (iz >= p && iz < g.nzp - p) {
        if (ix >= p - N && ix < p + N) s[2] = 4 * N * g.nx + (ix - (p - N)) * g.nz + (iz - p);
        if (ix >= g.nxp - p - N && ix < g.nxp - p + N)
            s[3] = 4 * N * g.nx + 2 * N * g.nz + (ix - (g.nxp - p - N)) * g.nz + (iz - p);
    }
}

// Explosive source on txx and tzz of every shot (sfwi-stage1/FD2DGPU.cu:2104-2114).
// buf: the forward-state buffer the step reads (1: DevShot::fw2, fused step march).
__global__ void k_source(Geo g, const DevShot *__restrict__ shots, const float *__restrict__ wav,
                         int it, int nshots, int buf) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    const DevShot &d = shots[s];
    const Fields &F = buf ? d.fw2 : d.fw;
    const int i = d.isx * g.nzp + d.isz;
    const float w = d.wav ? d.wav[it] : wav[it];
    F.txx[i] += w;
    F.tzz[i] += w;
}

// Velocity half step of every shot.  FWD also stores the boundary strips of
// the new velocity; BORN (1 or 2) also advances the scattered field, and
// BORN == 2 the second-order scattered field.
template <bool EXACT, bool SAVE, int BORN>
__global__ void __launch_bounds__(BLK) k_vel(Geo g, const DevShot *__restrict__ shots, int it, int dir) {
    int ix, iz;
    if (!cell_of(g, ix, iz)) return;
    const DevShot &d = shots[blockIdx.y];
    const int i = ix * g.nzp + iz;
    const bool split = in_band(g, ix, iz);
    if (g.grad_rho && dir < 0) {   // reconstruction: keep v(it+1) for the density image
        d.vold_x[i] = d.fw.vx[i];
        d.vold_z[i] = d.fw.vz[i];
    }
    vel_cell<EXACT>(g, d.m, d.fw, i, split, dir);
    if (BORN) vel_cell<EXACT>(g, d.m, d.bn, i, split, dir);
    if (BORN == 2) vel_cell<EXACT>(g, d.m, d.bn2, i, split, dir);
    if (SAVE && split) {
        int sl[4];
        strip_slots(g, ix, iz, sl);
        if (sl[0] >= 0 || sl[1] >= 0 || sl[2] >= 0 || sl[3] >= 0) {
            const float vx = d.fw.vx[i], vz = d.fw.vz[i];
            const fptr base = g.store ? d.stage : d.strip + (size_t)it * 2 * g.sf;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (sl[k] >= 0) {
                    base[sl[k]] = vx;
                    base[g.sf + sl[k]] = vz;
                }
        }
    }
}

// Stress half step of every shot.  GRAD stores the four velocity derivatives
// (update_txx_tzz_txz_grad, sfwi-stage1/born.cu:451-499); BORN also advances
// the scattered stresses with the Born source of the background velocity;
// BORN == 2 then the second-order scattered stresses with the Born source of
// the updated first-order scattered velocity (born_modeling mode 5,
// sfwi-stage1/FD2DGPU.cu:1888-1903: its velocity step, the source from b_vx,
// b_vz, its stress step).
template <bool EXACT, bool GRAD, int BORN>
__global__ void __launch_bounds__(BLK) k_str(Geo g, const DevShot *__restrict__ shots, int dir) {
    int ix, iz;
    if (!cell_of(g, ix, iz)) return;
    const DevShot &d = shots[blockIdx.y];
    const int i = ix * g.nzp + iz;
    const bool split = in_band(g, ix, iz) || (ix == d.isx && iz == d.isz);
    const VGrad v = vgrad<EXACT>(g, d.fw, i);
    BornSrc none{};
    str_cell<EXACT, false>(g, d.m, d.fw, v, i, split, dir, none);
    if (GRAD) {
        if (EXACT) {
            d.dgx[i] = v.xx / g.dx;
            d.dgz[i] = v.zz / g.dz;
            d.dgxz[i] = v.xz / g.dz;
            d.dgzx[i] = v.zx / g.dx;
        } else {
            d.dgx[i] = v.xx * g.idx;
            d.dgz[i] = v.zz * g.idz;
            d.dgxz[i] = v.xz * g.idz;
            d.dgzx[i] = v.zx * g.idx;
        }
    }
    if (BORN) {
        const BornSrc s = born_src<EXACT>(g, d.m, v, i, dir);
        const VGrad vb = vgrad<EXACT>(g, d.bn, i);
        str_cell<EXACT, true>(g, d.m, d.bn, vb, i, in_band(g, ix, iz), dir, s);
        if (BORN == 2) {   // the first-order velocity is not changed by its stress step
            const BornSrc s2 = born_src<EXACT>(g, d.m, vb, i, dir);
            const VGrad vb2 = vgrad<EXACT>(g, d.bn2, i);
            str_cell<EXACT, true>(g, d.m, d.bn2, vb2, i, in_band(g, ix, iz), dir, s2);
        }
    }
}

// Restore the boundary strips of step `it` into vx, vz
// (sfwi-stage1/FD2DGPU.cu:3873-3899).
__global__ void k_strip_load(Geo g, const DevShot *__restrict__ shots, int it) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.sf) return;
    const DevShot &d = shots[blockIdx.y];
    const int N = NSTRIP, p = g.pml;
    int ix, iz;
    if (t < 2 * N * g.nx) {
        ix = p + t / (2 * N);
        iz = p - N + t % (2 * N);
    } else if (t < 4 * N * g.nx) {
        const int u = t - 2 * N * g.nx;
        ix = p + u / (2 * N);
        iz = g.nzp - p - N + u % (2 * N);
    } else if (t < 4 * N * g.nx + 2 * N * g.nz) {
        const int u = t - 4 * N * g.nx;
        ix = p - N + u / g.nz;
        iz = p + u % g.nz;
    } else {
        const int u = t - 4 * N * g.nx - 2 * N * g.nz;
        ix = g.nxp - p - N + u / g.nz;
        iz = p + u % g.nz;
    }
    const int i = ix * g.nzp + iz;
    if (g.store) {
        const gptr<unsigned short> h = d.strip_h + (size_t)it * 2 * g.sf;
        const fptr sc = d.strip_scale + 2 * (size_t)it;
        d.fw.vx[i] = unpack16(g.store, h[t]) * sc[0];
        d.fw.vz[i] = unpack16(g.store, h[g.sf + t]) * sc[1];
    } else {
        const fptr base = d.strip + (size_t)it * 2 * g.sf;
        d.fw.vx[i] = base[t];
        d.fw.vz[i] = base[g.sf + t];
    }
}

// 16-bit strip storage of step `it`: per field, scale = max |v| over the
// step's strip, stored value = v / scale rounded to nearest even in FP16
// (store 1) or BF16 (store 2).  One workgroup per (field, shot).
__global__ void __launch_bounds__(256) k_strip_pack(Geo g, const DevShot *__restrict__ shots, int it) {
    __shared__ float red[256];
    const DevShot &d = shots[blockIdx.y];
    const int f = blockIdx.x;
    const fptr v = d.stage + (size_t)f * g.sf;
    float m = 0.0f;
    for (int k = threadIdx.x; k < g.sf; k += 256) m = fmaxf(m, fabsf(v[k]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    const float scale = red[0];
    const float inv = scale > 0.0f ? 1.0f / scale : 0.0f;
    const gptr<unsigned short> h = d.strip_h + ((size_t)it * 2 + f) * g.sf;
    for (int k = threadIdx.x; k < g.sf; k += 256) h[k] = pack16(g.store, v[k] * inv);
    if (threadIdx.x == 0) d.strip_scale[2 * (size_t)it + f] = scale;
}

// Receiver sampling (sfwi-stage1/FD2DGPU.cu:3158-3166): vz, vx of the
// forward field into rz, rx; with BORN the scattered field goes to rz, rx and
// the background to r2z, r2x (record_scatter, sfwi-stage2/FD2DGPU.cu:3511-3519);
// BORN == 2 puts the second-order scattered field in r2z, r2x instead
// (record_scatter with double_scatter, sfwi-stage1/FD2DGPU.cu:3610-3617).
// buf: the forward-state buffer holding the step's result (1: DevShot::fw2).
template <int BORN>
__global__ void k_record(Geo g, const DevShot *__restrict__ shots, int it, int buf) {
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    const DevShot &d = shots[blockIdx.y];
    if (tr >= d.ntr) return;
    const Fields &F = buf ? d.fw2 : d.fw;
    const int i = d.igx[tr] * g.nzp + d.igz;
    const size_t o = (size_t)tr * g.nt + it;
    if (BORN) {
        const Fields &S = BORN == 2 ? d.bn2 : F;
        d.rz[o] = d.bn.vz[i];
        d.rx[o] = d.bn.vx[i];
        d.r2z[o] = S.vz[i];
        d.r2x[o] = S.vx[i];
    } else {
        d.rz[o] = F.vz[i];
        d.rx[o] = F.vx[i];
    }
}

// Residual in place of the synthetic traces: stage 1 r = 1*syn - obs
// (sfwi-stage1/FD2DGPU.cu:3181-3195); stage 2 r = (b - obs) then 1*r + bg
// (sfwi-stage2/FD2DGPU.cu:3119-3158).
template <bool BORN>
__global__ void k_residual(Geo g, const DevShot *__restrict__ shots) {
    const DevShot &d = shots[blockIdx.y];
    const size_t n = (size_t)d.ntr * g.nt;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        if (BORN) {
            float z = d.rz[k] - d.oz[k], x = d.rx[k] - d.ox[k];
            d.rz[k] = 1.0f * z + d.r2z[k];
            d.rx[k] = 1.0f * x + d.r2x[k];
        } else {
            d.rz[k] = 1.0f * d.rz[k] - d.oz[k];
            d.rx[k] = 1.0f * d.rx[k] - d.ox[k];
        }
    }
}

// Per-trace misfit partial sums: each term formed in FP32 and accumulated in
// FP64 along the trace (sfwi-stage1/FD2DGPU.cu:4587-4590).
__global__ void k_trace_l2(Geo g, const DevShot *__restrict__ shots) {
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    const DevShot &d = shots[blockIdx.y];
    if (tr >= d.ntr) return;
    const fptr z = d.rz + (size_t)tr * g.nt, x = d.rx + (size_t)tr * g.nt;
    double s = 0.0;
    for (int j = 0; j < g.nt; ++j) s += (double)(z[j] * z[j] + x[j] * x[j]);
    d.trace_sum[tr] = s;
}
// Shot misfit = ordered sum of the trace sums.
__global__ void k_shot_l2(const DevShot *__restrict__ shots, double *out, int nshots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    const DevShot &d = shots[s];
    double m = 0.0;
    for (int tr = 0; tr < d.ntr; ++tr) m += d.trace_sum[tr];
    out[s] = m;
}

// Adjoint source: the residual overwrites the adjoint velocity at the
// receivers (sfwi-stage1/FD2DGPU.cu:3425-3441, stack = false).  `inj` lists
// the traces to inject, one per receiver cell (the last trace of coincident
// receivers, the sequential order's winner), so threads never collide.
__global__ void k_inject(Geo g, const DevShot *__restrict__ shots, int it) {
    const DevShot &d = shots[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.ninj) return;
    const int tr = d.inj[d.inj_off[k + 1] - 1];
    const int i = d.igx[tr] * g.nzp + d.igz;
    d.bk.vz[i] = d.rz[(size_t)tr * g.nt + it];
    d.bk.vx[i] = d.rx[(size_t)tr * g.nt + it];
}

// Adjoint source of the snapshot (consistent) gradient: fac r / rho added at
// the receivers, to the totals and to the x parts of the split adjoint
// velocity (the receiver row is split in the adjoint velocity step), so the
// parts keep summing to the totals.  fac = 2 for the L2 misfit sum r^2 (r the
// residual), 1 when r is already the derivative of the misfit (OT).  The
// traces of one receiver cell are added in trace order by one thread.
__global__ void k_inject_add(Geo g, const DevShot *__restrict__ shots, int it, float fac, int par) {
    const DevShot &d = shots[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.ninj) return;
    const Fields &B = par ? d.bk2 : d.bk;   // the buffer holding the current adjoint state
    const int i = d.igx[d.inj[d.inj_off[k]]] * g.nzp + d.igz;
    const float rho = d.m.rho[i];
    float vz = B.vz[i], vzx = B.vz_x[i], vx = B.vx[i], vxx = B.vx_x[i];
    for (int j = d.inj_off[k]; j < d.inj_off[k + 1]; ++j) {
        const size_t o = (size_t)d.inj[j] * g.nt + it;
        const float az = fac * d.rz[o] / rho, ax = fac * d.rx[o] / rho;
        vz += az; vzx += az; vx += ax; vxx += ax;
    }
    B.vz[i] = vz; B.vz_x[i] = vzx; B.vx[i] = vx; B.vx_x[i] = vxx;
}

// The adjoint source of every reverse step as a column table, fac r / rho at
// the receiver row (the value k_inject_add adds), -0 elsewhere (adding it
// changes nothing, not even a zero's sign): the fused adjoint step reads one
// value per column and component with no index load and adds it on the whole
// receiver row.  Needs every receiver in a cell of its own.
__global__ void k_inject_table(Geo g, const DevShot *__restrict__ shots, float fac) {
    const DevShot &d = shots[blockIdx.y];
    const size_t n = (size_t)g.nt * 2 * g.nxp;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(q % g.nxp), f = (int)((q / g.nxp) & 1);
        const int it = (int)(q / (2 * (size_t)g.nxp));
        const int tr = d.rtr[x];
        float v = -0.0f;   // x + (-0) == x for every x, zeros of either sign included
        if (tr >= 0) {
            const float rho = d.m.rho[x * g.nzp + d.igz];
            const size_t o = (size_t)tr * g.nt + it;
            v = f ? fac * d.rx[o] / rho : fac * d.rz[o] / rho;
        }
        d.ainj[q] = v;
    }
}

template <bool EXACT>
__global__ void __launch_bounds__(BLK) k_adj_str(Geo g, const DevShot *__restrict__ shots, int dir) {
    int ix, iz;
    if (!cell_of(g, ix, iz)) return;
    const DevShot &d = shots[blockIdx.y];
    const int i = ix * g.nzp + iz;
    adj_str_cell<EXACT>(g, d.m, d.bk, i, in_band(g, ix, iz), dir);
}

// Adjoint velocity step fused with the zero-lag imaging of this reverse step
// (sfwi-stage1/FD2DGPU.cu:3350-3382 doimage_gradient, :3324-3331 doimage_illum);
// IMAGE = false: the adjoint velocity step alone (snapshot gradient, which
// images in k_snap_image).
template <bool EXACT, bool IMAGE = true>
__global__ void __launch_bounds__(BLK) k_adj_vel_image(Geo g, const DevShot *__restrict__ shots, int dir) {
    int ix, iz;
    if (!cell_of(g, ix, iz)) return;
    const DevShot &d = shots[blockIdx.y];
    const int i = ix * g.nzp + iz;
    const bool split = in_band(g, ix, iz) || iz == d.igz;
    adj_vel_cell<EXACT>(g, d.m, d.bk, i, split, dir);
    if (!IMAGE) return;
    const int jx = ix - g.pml, jz = iz - g.pml;
    if (jx < 0 || jx >= g.nx || jz < 0 || jz >= g.nz) return;
    const int q = jx * g.nz + jz;
    const float gx = d.dgx[i], gz = d.dgz[i], gxz = d.dgxz[i], gzx = d.dgzx[i];
    const float sxx = d.bk.txx[i], szz = d.bk.tzz[i], sxz = d.bk.txz[i];
    const float vx = d.fw.vx[i], vz = d.fw.vz[i];
    if (EXACT) {
        const float lg = -((gx + gz) * (sxx + szz));
        const float mg = -2.0f * (gx * sxx + gz * szz) - (gxz + gzx) * sxz;
        const float vp = d.m.vp[i], vs = d.m.vs[i], rho = d.m.rho[i];
        const float pp = 2.0f * vp * rho * lg;
        const float ps = -4.0f * vs * rho * lg + 2.0f * vs * rho * mg;
        d.gvp[q] += 1.0f * pp;
        d.gvs[q] += 1.0f * ps;
        d.gil[q] += 1.0f * (vx * vx + vz * vz);
        if (g.grad_rho) {
            const float dvx = d.vold_x[i] - vx, dvz = d.vold_z[i] - vz;
            const float rd = -(d.bk.vx[i] * dvx + d.bk.vz[i] * dvz) / g.dt;
            const float rg = rd + (vp * vp - 2.0f * vs * vs) * lg + vs * vs * mg;
            d.grho[q] += 1.0f * rg;
        }
    } else {
        image_fast(d, i, q);
        if (g.grad_rho) {
            const float lg = -((gx + gz) * (sxx + szz));
            const float mg = __builtin_fmaf(-(gxz + gzx), sxz, -2.0f * __builtin_fmaf(gx, sxx, gz * szz));
            d.grho[q] += rho_term_fast(g, lg, mg, d.bk.vx[i], d.bk.vz[i], d.vold_x[i] - vx, d.vold_z[i] - vz,
                                       d.m.vp[i], d.m.vs[i]);
        }
    }
}

// Per-shot accumulators added into the full-model images.  One thread per
// model cell (allnx columns of nz rows) adds the shots whose window covers
// its column, in batch order: windows at different offsets overlap, so a
// thread per window cell would race with the thread of another shot that
// maps to the same model cell, and the batch order makes the result equal to
// the shots merged one at a time, bitwise.  SCALE (FAST mode): the Vp and Vs
// accumulators hold cofactors, multiplied here by the imaging coefficients
// 2 vp rho and 2 vs rho of the shot's window (a rounded product, then the sum).
// wv, wi: weights of the Vp / Vs and illumination sums (1 for the strip
// gradient, which leaves every product unchanged; K dt and K for the
// snapshot gradient, which images every K-th step and includes dt).
// ALL = false (the Hessian-vector product, doimage_gradient alone): the
// Vp / Vs images only.
template <bool SCALE, bool ALL = true>
__global__ void k_merge(Geo g, const DevShot *__restrict__ shots, int nshots, int allnx, float *gvp, float *gvs,
                        float *gil, float *grho, float wv, float wi) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= allnx * g.nz) return;
    const int ox = o / g.nz, jz = o - ox * g.nz;
    const bool rho = ALL && g.grad_rho;
    float a = gvp[o], b = gvs[o], c = ALL ? gil[o] : 0.0f, r = rho ? grho[o] : 0.0f;
    for (int s = 0; s < nshots; ++s) {
        const DevShot &d = shots[s];
        const int jx = ox - d.left;
        if (jx < 0 || jx >= g.nx) continue;
        const int q = jx * g.nz + jz;
        if (SCALE) {
            const int i = (jx + g.pml) * g.nzp + jz + g.pml;
            a += (d.m.icp[i] * wv) * d.gvp[q];
            b += (d.m.ics[i] * wv) * d.gvs[q];
        } else {
            a += d.gvp[q];
            b += d.gvs[q];
        }
        if (ALL) c += wi * d.gil[q];
        if (rho) r += d.grho[q];
    }
    gvp[o] = a;
    gvs[o] = b;
    if (ALL) gil[o] = c;
    if (rho) grho[o] = r;
}

// ----------------------------------------------------- velocity snapshots ---
// The snapshot gradient keeps the forward velocity after the velocity update
// of every step it with it % snap_k == 0 on the region the imaging's
// derivatives read: the nx x nz interior plus a 4-cell halo, columns
// [pml-4, pml+nx+4) x rows [pml-4, pml+nz+4), index x * snap_rz + z in region
// coordinates.  FP32, or FP16 / BF16 of v / scale with one FP32 max-abs scale
// per snapshot and field (rounded to nearest even).

__device__ inline int snap_cell(const Geo &g, int t) {   // padded index of region cell t
    const int x = t / g.snap_rz, z = t - x * g.snap_rz;
    return (x + g.pml - NOP) * g.nzp + (z + g.pml - NOP);
}

// Max |v| of snapshot k (16-bit storage) in two passes without atomics:
// k_snap_max gives block b of a shot the region cells b BLK + j + m SNAP_PART
// BLK and writes that block's maximum of each field to the shot's partials;
// k_snap_scale reduces a shot's SNAP_PART partials to the snapshot's scales.
constexpr int SNAP_PART = 256;   // blocks (partial maxima per field) of k_snap_max per shot

__device__ inline float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__global__ void __launch_bounds__(BLK) k_snap_max(Geo g, const DevShot *__restrict__ shots, int buf) {
    __shared__ float red[2][BLK / 64];
    const DevShot &d = shots[blockIdx.y];
    const Fields &F = buf ? d.fw2 : d.fw;
    float mx = 0.0f, mz = 0.0f;
    for (int t = blockIdx.x * BLK + threadIdx.x; t < g.snap_n; t += SNAP_PART * BLK) {
        const int i = snap_cell(g, t);
        mx = fmaxf(mx, fabsf(F.vx[i]));
        mz = fmaxf(mz, fabsf(F.vz[i]));
    }
    mx = wave_max(mx);
    mz = wave_max(mz);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wv] = mx;
        red[1][wv] = mz;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        float m = red[threadIdx.x][0];
#pragma unroll
        for (int k = 1; k < BLK / 64; ++k) m = fmaxf(m, red[threadIdx.x][k]);
        d.snap_part[threadIdx.x * g.snap_np + blockIdx.x] = m;
    }
}

// One wave per shot: the scales of snapshot k from the np partial maxima of
// each field (written by k_snap_max or by the workgroup fused forward step),
// stored as FP32 bits (non-negative floats order as integers).
__global__ void __launch_bounds__(64) k_snap_scale(Geo g, const DevShot *__restrict__ shots, int k, int np) {
    const DevShot &d = shots[blockIdx.x];
    float mx = 0.0f, mz = 0.0f;
    for (int j = threadIdx.x; j < np; j += 64) {
        mx = fmaxf(mx, d.snap_part[j]);
        mz = fmaxf(mz, d.snap_part[g.snap_np + j]);
    }
    mx = wave_max(mx);
    mz = wave_max(mz);
    if (threadIdx.x == 0) {
        d.snap_max[2 * (size_t)k] = __float_as_uint(mx);
        d.snap_max[2 * (size_t)k + 1] = __float_as_uint(mz);
    }
}

// Store snapshot k: FP32, or v / scale in 16 bits.  buf: the forward-state
// buffer holding the new velocity (1: DevShot::fw2).  Each thread packs
// SNAP_PER cells BLK apart (every load and store instruction of a wave stays
// contiguous) and issues all their loads before its first store: one cell per
// thread left each wave two loads in flight and launched four times the
// waves, for 3.2 TB/s.
constexpr int SNAP_PER = 4;
__global__ void __launch_bounds__(BLK) k_snap_pack(Geo g, const DevShot *__restrict__ shots, int k, int buf) {
    const DevShot &d = shots[blockIdx.y];
    const Fields &F = buf ? d.fw2 : d.fw;
    const int t0 = blockIdx.x * (BLK * SNAP_PER) + threadIdx.x;
    float vx[SNAP_PER], vz[SNAP_PER];
#pragma unroll
    for (int j = 0; j < SNAP_PER; ++j) {
        const int t = t0 + j * BLK;
        if (t < g.snap_n) {
            const int i = snap_cell(g, t);
            vx[j] = F.vx[i];
            vz[j] = F.vz[i];
        }
    }
    // snapshot k starts at a 64-bit offset; indices within it stay below 2^31
    const size_t o = (size_t)k * 2 * g.snap_n;
    if (g.store == 0) {
        const fptr s = d.snap32 + o;
#pragma unroll
        for (int j = 0; j < SNAP_PER; ++j) {
            const int t = t0 + j * BLK;
            if (t < g.snap_n) {
                s[t] = vx[j];
                s[g.snap_n + t] = vz[j];
            }
        }
    } else {
        const gptr<unsigned short> s = d.snap16 + o;
        const float sx = __uint_as_float(d.snap_max[2 * (size_t)k]), sz = __uint_as_float(d.snap_max[2 * (size_t)k + 1]);
        const float ix = sx > 0.0f ? 1.0f / sx : 0.0f, iz = sz > 0.0f ? 1.0f / sz : 0.0f;
#pragma unroll
        for (int j = 0; j < SNAP_PER; ++j) {
            const int t = t0 + j * BLK;
            if (t < g.snap_n) {
                s[t] = pack16(g.store, vx[j] * ix);
                s[g.snap_n + t] = pack16(g.store, vz[j] * iz);
            }
        }
    }
}

// Value of snapshot k's field f (0 vx, 1 vz) at region cell t, unscaled.
__device__ inline float snap_val(const Geo &g, const DevShot &d, int k, int f, int t) {
    const size_t o = ((size_t)k * 2 + f) * g.snap_n;
    return g.store == 0 ? (d.snap32 + o)[t] : unpack16(g.store, (d.snap16 + o)[t]);
}

// Zero-lag imaging of the snapshot gradient at snapshot k, one thread per
// interior cell: the stress update's velocity derivatives of the kept
// velocity (Dx- vx, Dz- vz, Dz+ vx, Dx+ vz), the adjoint stresses of the
// current reverse step (between its adjoint velocity and stress half steps)
// and the Vp / Vs cofactor sums (k_merge applies 2 vp rho, 2 vs rho and the
// weight); illumination vx^2 + vz^2 of the kept velocity.
__global__ void __launch_bounds__(BLK) k_snap_image(Geo g, const DevShot *__restrict__ shots, int k, int par) {
    const DevShot &d = shots[blockIdx.y];
    const Fields &B = par ? d.bk2 : d.bk;   // the buffer holding the current adjoint state
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nx * g.nz) return;
    const int jx = q / g.nz, jz = q - jx * g.nz;
    const int R = g.snap_rz;
    const int t = (jx + NOP) * R + (jz + NOP);   // region cell of this interior cell
    const int bx = 0, bz = 1;   // fields of snapshot k
    float sx = 1.0f, sz = 1.0f;
    if (g.store != 0) {
        sx = __uint_as_float(d.snap_max[2 * (size_t)k]);
        sz = __uint_as_float(d.snap_max[2 * (size_t)k + 1]);
    }
    // sum_k c_k (f[t+(k-1)s] - f[t-ks]) and sum_k c_k (f[t+ks] - f[t-(k-1)s])
    auto dm = [&](int b, int s) {
        float a = kC[1] * (snap_val(g, d, k, b, t) - snap_val(g, d, k, b, t - s));
        a = __builtin_fmaf(kC[2], snap_val(g, d, k, b, t + s) - snap_val(g, d, k, b, t - 2 * s), a);
        a = __builtin_fmaf(kC[3], snap_val(g, d, k, b, t + 2 * s) - snap_val(g, d, k, b, t - 3 * s), a);
        a = __builtin_fmaf(kC[4], snap_val(g, d, k, b, t + 3 * s) - snap_val(g, d, k, b, t - 4 * s), a);
        return a;
    };
    auto dp = [&](int b, int s) {
        float a = kC[1] * (snap_val(g, d, k, b, t + s) - snap_val(g, d, k, b, t));
        a = __builtin_fmaf(kC[2], snap_val(g, d, k, b, t + 2 * s) - snap_val(g, d, k, b, t - s), a);
        a = __builtin_fmaf(kC[3], snap_val(g, d, k, b, t + 3 * s) - snap_val(g, d, k, b, t - 2 * s), a);
        a = __builtin_fmaf(kC[4], snap_val(g, d, k, b, t + 4 * s) - snap_val(g, d, k, b, t - 3 * s), a);
        return a;
    };
    const float gx = dm(bx, R) * (sx * g.idx), gz = dm(bz, 1) * (sz * g.idz);
    const float gxz = dp(bx, 1) * (sx * g.idz), gzx = dp(bz, R) * (sz * g.idx);
    const float vx = snap_val(g, d, k, bx, t) * sx, vz = snap_val(g, d, k, bz, t) * sz;
    const int i = (jx + g.pml) * g.nzp + (jz + g.pml);
    float op, os;
    image_vpvs(gx, gz, gxz, gzx, B.txx[i], B.tzz[i], B.txz[i], d.gvp[q], d.gvs[q], op, os);
    d.gvp[q] = op;
    d.gvs[q] = os;
    d.gil[q] = __builtin_fmaf(vx, vx, __builtin_fmaf(vz, vz, d.gil[q]));
}

// ------------------------------------------------------------- set-up ---

// Window of a full-model array into the padded grid with edge clamping; one
// pass equals the reference's copy, z-pad, x-pad sequence
// (sfwi-stage1/FD2DGPU.cu:112-128, 233-249).
__global__ void k_window(Geo g, const float *__restrict__ all, int left, float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nxp * g.nzp) return;
    const int ix = i / g.nzp, iz = i - ix * g.nzp;
    const int sx = min(max(ix, g.pml), g.nxp - g.pml - 1) - g.pml;
    const int sz = min(max(iz, g.pml), g.nzp - g.pml - 1) - g.pml;
    out[i] = all[(size_t)(left + sx) * g.allnz + sz];
}

// Window of the perturbation for the stage-1 Born source, which acts on the
// interior cells only (sfwi-stage1/born.cu:830-836, bounds [pml, n - pml)):
// the interior as k_window, zero on the PML cells, so the source there is 0.
__global__ void k_window_interior(Geo g, const float *__restrict__ all, int left, float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nxp * g.nzp) return;
    const int ix = i / g.nzp, iz = i - ix * g.nzp;
    const bool in = ix >= g.pml && ix < g.nxp - g.pml && iz >= g.pml && iz < g.nzp - g.pml;
    out[i] = in ? all[(size_t)(left + ix - g.pml) * g.allnz + (iz - g.pml)] : 0.0f;
}

// PML profile (sfwi-stage1/FD2DGPU.cu:443-493) and Lame parameters
// (sfwi-stage1/FD2DGPU.cu:1116-1124), plus the fast-mode coefficients.
// `lg` is log10(1/rr) rounded to float.
__global__ void k_medium(Geo g, float lg, float rho_const, float *vp, float *vs, float *rho, float *lam,
                         float *mu, float *dampx, float *dampz, float *bu, float *icp, float *ics) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nxp * g.nzp) return;
    const int ix = i / g.nzp, iz = i - ix * g.nzp;
    if (rho_const > 0.0f) rho[i] = rho_const;
    const float p = vp[i], s = vs[i], r = rho[i];
    int dxl = -1, dzl = -1;
    if (ix < g.pml) dxl = g.pml - ix;
    if (ix > g.nxp - g.pml) dxl = ix - (g.nxp - g.pml);
    if (iz < g.pml) dzl = g.pml - iz;
    if (iz > g.nzp - g.pml) dzl = iz - (g.nzp - g.pml);
    float ax = 0.0f, az = 0.0f;
    if (dxl >= 0) {
        const float t = (float)(1.0 * dxl / g.pml);
        const double t2 = (double)t * (double)t;
        ax = (float)((double)lg * (5.0 * (double)p / (2.0 * g.pml)) * (double)(float)(t2 * t2));
    }
    if (dzl >= 0) {
        const float t = (float)(1.0 * dzl / g.pml);
        const double t2 = (double)t * (double)t;
        az = (float)((double)lg * (5.0 * (double)p / (2.0 * g.pml)) * (double)(float)(t2 * t2));
    }
    dampx[i] = ax;
    dampz[i] = az;
    lam[i] = r * (p * p - 2.0f * s * s);
    mu[i] = r * s * s;
    bu[i] = 1.0f / r;
    icp[i] = 2.0f * p * r;
    ics[i] = 2.0f * s * r;
}

// Random boundary (rbc): the scheme of random_vel / set_random_boundary
// (sfwi-stage1/FD2DGPU.cu:162-220, 3245-3290) with a project-owned generator,
// the lowbias32 integer hash (C. Wellons, 2018) of seed, cell and attempt,
// where the reference seeds cuRAND with time(0).  k = vmax / pml; PASS 0 sets
// the z bands (top rows iz < pml and their mirrors nzp-1-iz, depth
// d = pml - iz), PASS 1 then the x bands (columns, d = pml - ix; corners drawn
// twice, as in the reference): vp' = vp + (2u - 1) k d redrawn until
// 0 <= vp' <= vmax (64 draws, then vp clamped), vs' = (vs / vp) vp'.  FP32
// throughout; the mirror is one cell inside the reference's nzpml-iz.
__device__ inline unsigned rbc_hash(unsigned x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}
// Largest (non-negative) Vp of the window: float bits order like unsigned ints.
__global__ void k_vmax(int n, const float *__restrict__ vp, unsigned *vmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMax(vmax, __float_as_uint(fmaxf(vp[i], 0.0f)));
}
__device__ inline void rbc_set(unsigned seed, int i, int pass, int d, float kk, float vmax, float *vp, float *vs) {
    const float v0 = vp[i], s0 = vs[i];
    const unsigned key = rbc_hash(seed ^ rbc_hash((unsigned)i * 2u + (unsigned)pass));
    float vn = fminf(fmaxf(v0, 0.0f), vmax);
    for (unsigned a = 0; a < 64; ++a) {
        const float u = __uint2float_rz(rbc_hash(key + a) >> 8) * (1.0f / 16777216.0f);
        const float t = __fadd_rn(v0, __fmul_rn(__fmul_rn(__fsub_rn(__fmul_rn(2.0f, u), 1.0f), kk), (float)d));
        if (t >= 0.0f && t <= vmax) { vn = t; break; }
    }
    vs[i] = v0 > 0.0f ? __fmul_rn(__fdiv_rn(s0, v0), vn) : s0;
    vp[i] = vn;
}
template <int PASS>
__global__ void k_rbc_band(Geo g, unsigned seed, const unsigned *vmax_bits, float *vp, float *vs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const float vmax = __uint_as_float(*vmax_bits), kk = __fdiv_rn(vmax, (float)g.pml);
    if (PASS == 0) {
        if (t >= g.nxp * g.pml) return;
        const int ix = t / g.pml, iz = t - ix * g.pml;
        rbc_set(seed, ix * g.nzp + iz, 0, g.pml - iz, kk, vmax, vp, vs);
        rbc_set(seed, ix * g.nzp + g.nzp - 1 - iz, 0, g.pml - iz, kk, vmax, vp, vs);
    } else {
        if (t >= g.pml * g.nzp) return;
        const int ix = t / g.nzp, iz = t - ix * g.nzp;
        rbc_set(seed, ix * g.nzp + iz, 1, g.pml - ix, kk, vmax, vp, vs);
        rbc_set(seed, (g.nxp - 1 - ix) * g.nzp + iz, 1, g.pml - ix, kk, vmax, vp, vs);
    }
}

// Test model of the stage-1 line search: v - step*d (host order,
// sfwi-stage1/FD2DGPU.cu:4902-4907).
__global__ void k_axpy_model(int n, const float *__restrict__ v, const float *__restrict__ d, float step,
                             float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v[i] - step * d[i];
}

// Stage-2 step sums: fenzi += sum L.r, fenmu += sum L^2, per trace in FP64
// along the trace, then in trace order (sfwi-stage2/FD2DGPU.cu:4707-4712).
__global__ void k_trace_dot(Geo g, int ntr, const float *lz, const float *lx, const float *rz,
                            const float *rx, double *num, double *den) {
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr >= ntr) return;
    double a = 0.0, b = 0.0;
    const size_t o = (size_t)tr * g.nt;
    for (int j = 0; j < g.nt; ++j) {
        const float z = lz[o + j], x = lx[o + j];
        b += (double)(z * z + x * x);
        a += (double)(z * rz[o + j] + x * rx[o + j]);
    }
    num[tr] = a;
    den[tr] = b;
}

}  // namespace sfwi
