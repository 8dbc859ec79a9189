// Column-march stencil kernels of the FAST mode.
//
// Each wave64 owns a band of 64 consecutive rows (z) of one shot -- 56 output
// rows plus a 4-row halo on each side -- and marches along x over a segment
// of columns.  Per field with x taps it keeps an eight-column sliding window
// in VGPRs and loads one new column per step; the z taps are formed from the
// lane's own value with DPP whole-wave shifts (v_mov_dpp wave_shl/wave_shr,
// plain VALU moves, no LDS).  The march is software-pipelined: every operand
// of column c+1 (new window columns and the cell's own state) is loaded while
// column c is computed, so memory latency overlaps the arithmetic.  Each
// output cell then costs about one coalesced wave load per operand, against
// sixteen per stencil field in the per-cell kernels whose limit is the
// texture-address unit.  Waves are independent (no LDS, no barriers);
// workgroups are spread so that each XCD owns a contiguous range of column
// segments, keeping the medium shared by a batch's shots in that XCD's L2.
//
// The segment width is a runtime choice (Geo::cm_xs); the per-cell update
// math is the FAST-mode code of kernels.hpp (load / math / store), so every
// FAST kernel family produces identical results.
#pragma once
#include "kernels.hpp"

namespace sfwi {

constexpr int CM_ROWS = 56;   // output rows per wave (64 lanes minus 2 x 4 halo)
// Prefetch depth of the imaging operands (at most the march depth): one
// column ahead measured the same as deeper rings, with fewer registers.
constexpr int IMG_DEPTH = 1;
constexpr int img_depth(int pd) { return pd < IMG_DEPTH ? pd : IMG_DEPTH; }

struct Seg {
    int x0, x1, zb;   // columns [x0, x1), first output row zb
    int shot;
};

// Block -> (shot, wave tile).  Block b runs on XCD b & 7; each XCD owns the
// contiguous range L = (b & 7) * per_xcd + (b >> 3) of the launch's work.
// Order 0: grid.y is the shot, L the workgroup tile of that shot.  Order 1:
// one 1-D grid over (tile, shot) pairs with the shot fastest, so the shots of
// one tile run back to back on one XCD and read the medium they share from
// that XCD's L2.  Order 2: one 1-D grid over (shot, tile) pairs with the tile
// fastest, so each XCD owns whole shots: the split PML band, which costs
// about twice the interior per cell, is then spread evenly over the XCDs
// instead of landing on the two that own the grid's left and right edges.
__device__ inline bool seg_of(const Geo &g, Seg &s) {
    const int b = blockIdx.x;
    const int L = (b & 7) * g.cm_per_xcd + (b >> 3);
    int tile = L;
    s.shot = blockIdx.y;
    if (g.cm_order == 1) {
        tile = L / g.cm_nshots;
        s.shot = L - tile * g.cm_nshots;
    } else if (g.cm_order == 2) {
        s.shot = L / g.cm_wg;
        tile = L - s.shot * g.cm_wg;
        if (s.shot >= g.cm_nshots) return false;
    }
    // the wave index is uniform across the wave: say so, so segment bounds,
    // loop trip counts and prefetch guards are scalar (no exec-mask branches)
    const int W = tile * (BLK / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (W >= g.cm_waves) return false;
    const int seg = W / g.cm_nbands;
    const int band = W - seg * g.cm_nbands;
    s.x0 = NOP + seg * g.cm_xs;
    s.x1 = min(s.x0 + g.cm_xs, g.nxp - NOP);
    s.zb = NOP + band * CM_ROWS;
    return true;
}

// Value of lane l+1 / l-1 (whole-wave DPP shifts; the end lanes get 0 and
// belong to the halo, whose results are discarded).
__device__ inline float nb_up(float v) { return __builtin_amdgcn_update_dpp(0.0f, v, 0x130, 0xf, 0xf, false); }
__device__ inline float nb_dn(float v) { return __builtin_amdgcn_update_dpp(0.0f, v, 0x138, 0xf, 0xf, false); }

//   z_plus : sum_k c_k (f[z+k]   - f[z-k+1])
//   z_minus: sum_k c_k (f[z+k-1] - f[z-k])
__device__ inline float z_plus(float f) {
    const float p1 = nb_up(f), p2 = nb_up(p1), p3 = nb_up(p2), p4 = nb_up(p3);
    const float m1 = nb_dn(f), m2 = nb_dn(m1), m3 = nb_dn(m2);
    float t = kC[1] * (p1 - f);
    t = __builtin_fmaf(kC[2], p2 - m1, t);
    t = __builtin_fmaf(kC[3], p3 - m2, t);
    t = __builtin_fmaf(kC[4], p4 - m3, t);
    return t;
}
__device__ inline float z_minus(float f) {
    const float p1 = nb_up(f), p2 = nb_up(p1), p3 = nb_up(p2);
    const float m1 = nb_dn(f), m2 = nb_dn(m1), m3 = nb_dn(m2), m4 = nb_dn(m3);
    float t = kC[1] * (f - m1);
    t = __builtin_fmaf(kC[2], p1 - m2, t);
    t = __builtin_fmaf(kC[3], p2 - m3, t);
    t = __builtin_fmaf(kC[4], p3 - m4, t);
    return t;
}

// x stencils from an 8-column window; both plus (w[j] = f(c-3+j)) and minus
// (w[j] = f(c-4+j)) take the form sum_k c_k (w[3+k] - w[4-k]).
__device__ inline float x_win(const float *w) {
    float t = kC[1] * (w[4] - w[3]);
    t = __builtin_fmaf(kC[2], w[5] - w[2], t);
    t = __builtin_fmaf(kC[3], w[6] - w[1], t);
    t = __builtin_fmaf(kC[4], w[7] - w[0], t);
    return t;
}
__device__ inline void shift(float *w) {
#pragma unroll
    for (int j = 0; j < 7; ++j) w[j] = w[j + 1];
}

// Every load comes from a valid address (the row clamped into the grid), and
// the zero of an out-of-grid lane is a select applied when the value is
// consumed a column later, not when it is loaded: an instruction that reads a
// load's result, or writes a register that a load in flight targets, makes
// the wave wait for that load, which would serialise the prefetch.
struct Lane {
    int z, zc;          // row, row clamped into the grid (load address)
    bool zin, act;
};
__device__ inline Lane lane_of(const Geo &g, const Seg &s) {
    const int lane = threadIdx.x & 63;
    Lane l;
    l.z = s.zb - 4 + lane;
    l.zc = min(l.z, g.nzp - 1);
    l.zin = l.z < g.nzp;
    l.act = lane >= 4 && lane < 60 && l.z < g.nzp - NOP;
    return l;
}
template <class P> __device__ inline float ld(const P &f, const Geo &g, const Lane &l, int x) {
    return f[x * g.nzp + l.zc];
}
__device__ inline float zs(const Lane &l, float v) { return l.zin ? v : 0.0f; }
// True on every lane when some lane of the wave holds p (wave-uniform value).
__device__ inline bool any_lane(bool p) { return __ballot(p) != 0; }

// Own-cell loaders of the marches, into ring slots that persist across
// columns: an active lane loads the totals when it is unsplit and the split
// parts and damping when it is split.  Both are lane-masked branches with no
// else, so a slot keeps its registers (nothing writes a register that a load
// in flight targets, which would make the wave wait for it), a wave that
// straddles the band edge fetches only the cache lines each variant needs, and
// a variant no lane uses is skipped.  The update math reads only the variant
// its lane uses.
__device__ inline void vel_load_w(const Medium &m, const Fields &F, int i, bool act, bool sp, VelCell &s) {
    if (act) s.bu = m.bu[i];
    if (act && !sp) { s.vx = F.vx[i]; s.vz = F.vz[i]; }
    if (act && sp) {
        s.vx_x = F.vx_x[i]; s.vx_z = F.vx_z[i]; s.vz_x = F.vz_x[i]; s.vz_z = F.vz_z[i];
        s.ax = m.dampx[i]; s.az = m.dampz[i];
    }
}
__device__ inline void str_load_w(const Medium &m, const Fields &F, int i, bool act, bool sp, StrCell &s) {
    if (act) { s.lam = m.lam[i]; s.mu = m.mu[i]; }
    if (act && !sp) { s.txx = F.txx[i]; s.tzz = F.tzz[i]; s.txz = F.txz[i]; }
    if (act && sp) {
        s.xx_x = F.txx_x[i]; s.xx_z = F.txx_z[i]; s.zz_x = F.tzz_x[i];
        s.zz_z = F.tzz_z[i]; s.xz_x = F.txz_x[i]; s.xz_z = F.txz_z[i];
        s.ax = m.dampx[i]; s.az = m.dampz[i];
    }
}
__device__ inline void adj_str_load_w(const Medium &m, const Fields &F, int i, bool act, bool sp, AdjStrCell &s) {
    if (act && !sp) { s.txx = F.txx[i]; s.tzz = F.tzz[i]; s.txz = F.txz[i]; }
    if (act && sp) {
        s.xx_x = F.txx_x[i]; s.zz_z = F.tzz_z[i]; s.xz_x = F.txz_x[i]; s.xz_z = F.txz_z[i];
        s.ax = m.dampx[i]; s.az = m.dampz[i];
    }
}

// Image index of the lane's cell in column c (its offset in the per-shot
// image of the interior), or -1 outside the interior or on a halo lane.
__device__ inline int img_index(const Geo &g, const Lane &l, int c) {
    const int jx = c - g.pml, jz = l.z - g.pml;
    return (l.act && jx >= 0 && jx < g.nx && jz >= 0 && jz < g.nz) ? jx * g.nz + jz : -1;
}

// Operands of the zero-lag imaging of one cell (loaded a column ahead, only
// when some lane of the wave images).  sxx and szz are the adjoint stresses
// the velocity update loads anyway; the caller fills them from its ring.
struct ImgCell {
    float gx, gz, gxz, gzx, sxx, szz, sxz, vx, vz, ip, is, il;
    float ox, oz, vp, vs, ir;   // density image (grad_rho): v(it+1), vp, vs, accumulator
};
__device__ inline void img_load(const DevShot &d, int i, int q, bool any, ImgCell &c) {
    if (any) {
        c.gx = d.dgx[i]; c.gz = d.dgz[i]; c.gxz = d.dgxz[i]; c.gzx = d.dgzx[i];
        c.sxz = d.bk.txz[i];
        c.vx = d.fw.vx[i]; c.vz = d.fw.vz[i];
        c.ip = d.gvp[q]; c.is = d.gvs[q]; c.il = d.gil[q];
    }
}
__device__ inline void img_load_rho(const DevShot &d, int i, int q, bool any, ImgCell &c) {
    if (any) {
        c.ox = d.vold_x[i]; c.oz = d.vold_z[i];
        c.vp = d.m.vp[i]; c.vs = d.m.vs[i];
        c.ir = d.grho[q];
    }
}
// The arithmetic of image_fast on prefetched operands.
__device__ inline void img_store(const DevShot &d, int q, const ImgCell &c) {
    float op, os;
    image_vpvs(c.gx, c.gz, c.gxz, c.gzx, c.sxx, c.szz, c.sxz, c.ip, c.is, op, os);
    d.gvp[q] = op;
    d.gvs[q] = os;
    d.gil[q] = __builtin_fmaf(c.vx, c.vx, __builtin_fmaf(c.vz, c.vz, c.il));
}

// Velocity half step (SAVE: store strips of the new velocity; BORN: also the
// scattered field; ILL: reconstruction, also accumulate the illumination
// vx^2 + vz^2 of the new velocity).  PD columns of operands are in flight
// ahead of the one being computed (a ring indexed at compile time: the march
// is unrolled by PD).
template <bool SAVE, bool BORN, int PD, bool ILL = false>
__global__ void __launch_bounds__(BLK) km_vel(Geo g, const DevShot *__restrict__ shots, int it, int dir) {
    Seg s;
    if (!seg_of(g, s)) return;
    const DevShot d = shots[s.shot];   // by value: the field pointers stay in registers (stores cannot alias them)
    const Lane l = lane_of(g, s);
    constexpr int NF = BORN ? 2 : 1;
    float wxx[NF][8], wxz[NF][8];
    float qxx[NF][PD], qxz[NF][PD], qzz[NF][PD];
    VelCell qst[NF][PD] = {};
    auto fetch = [&](int f, int j, int c) {   // operands of column c into ring slot j
        const Fields &F = f ? d.bn : d.fw;
        qxx[f][j] = ld(F.txx, g, l, c + 4);
        qxz[f][j] = ld(F.txz, g, l, c + 3);
        qzz[f][j] = ld(F.tzz, g, l, c);
        const bool sp = in_band(g, c, l.z);
        vel_load_w(d.m, F, c * g.nzp + l.zc, l.act, sp, qst[f][j]);
    };
    constexpr int IPD = img_depth(PD);
    float qil[IPD] = {};
    auto ifetch = [&](int j, int c) {   // illumination so far, IPD columns ahead
        const int q = img_index(g, l, c);
        if (any_lane(q >= 0)) qil[j] = d.gil[q < 0 ? 0 : q];
    };
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const Fields &F = f ? d.bn : d.fw;
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            wxx[f][j] = zs(l, ld(F.txx, g, l, s.x0 - 3 + j));
            wxz[f][j] = zs(l, ld(F.txz, g, l, s.x0 - 4 + j));
        }
#pragma unroll
        for (int j = 0; j < PD; ++j)
            if (s.x0 + j < s.x1) fetch(f, j, s.x0 + j);
    }
    if (ILL) {
#pragma unroll
        for (int j = 0; j < IPD; ++j)
            if (s.x0 + j < s.x1) ifetch(j, s.x0 + j);
    }
    for (int c0 = s.x0; c0 < s.x1; c0 += PD) {
#pragma unroll
        for (int j = 0; j < PD; ++j) {
            const int c = c0 + j;
            if (c >= s.x1) break;
            const bool split = in_band(g, c, l.z);
            const int i = c * g.nzp + l.z;
            float il = 0.0f;
            if (ILL) {
                il = qil[j % IPD];
                if (c + IPD < s.x1) ifetch(j % IPD, c + IPD);
            }
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const Fields &F = f ? d.bn : d.fw;
                wxx[f][7] = zs(l, qxx[f][j]);
                wxz[f][7] = zs(l, qxz[f][j]);
                const float tzz = zs(l, qzz[f][j]);
                VelCell st = qst[f][j];
                if (c + PD < s.x1) fetch(f, j, c + PD);
                const float a = x_win(wxx[f]);        // Dx+ txx
                const float e = x_win(wxz[f]);        // Dx- txz
                const float cz = z_plus(tzz);         // Dz+ tzz
                const float bz = z_minus(wxz[f][4]);  // Dz- txz (txz at column c)
                shift(wxx[f]);
                shift(wxz[f]);
                if (l.act) {
                    if (g.grad_rho && dir < 0 && f == 0) {   // reconstruction: keep v(it+1)
                        d.vold_x[i] = split ? F.vx[i] : st.vx;
                        d.vold_z[i] = split ? F.vz[i] : st.vz;
                    }
                    vel_math(g, st, split, dir, a, bz, cz, e);
                    vel_store(F, i, st, split);
                    if (ILL && f == 0) {
                        const int q = img_index(g, l, c);
                        if (q >= 0) d.gil[q] = __builtin_fmaf(st.vx, st.vx, __builtin_fmaf(st.vz, st.vz, il));
                    }
                    if (SAVE && f == 0 && split) {
                        int sl[4];
                        strip_slots(g, c, l.z, sl);
                        const fptr base = g.store ? d.stage : d.strip + (size_t)it * 2 * g.sf;
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            if (sl[k] >= 0) {
                                base[sl[k]] = st.vx;
                                base[g.sf + sl[k]] = st.vz;
                            }
                    }
                }
            }
        }
    }
}

// Stress half step (GRAD: store the velocity derivatives; BORN: also the
// scattered stresses with the Born source of the background velocity; IMG:
// reconstruction, also form the Vp / Vs images of the step from the velocity
// derivatives, which stay in registers, and the adjoint stresses that the
// adjoint stress-like step of the same time step has produced).
template <bool GRAD, bool BORN, int PD, bool IMG = false>
__global__ void __launch_bounds__(BLK) km_str(Geo g, const DevShot *__restrict__ shots, int dir) {
    Seg s;
    if (!seg_of(g, s)) return;
    const DevShot d = shots[s.shot];   // by value: the field pointers stay in registers (stores cannot alias them)
    const Lane l = lane_of(g, s);
    constexpr int NF = BORN ? 2 : 1;
    float wvx[NF][8], wvz[NF][8];   // vx: columns c-4.., vz: columns c-3..
    float qvx[NF][PD], qvz[NF][PD];
    StrCell qst[NF][PD] = {};
    auto split_of = [&](int f, int c) {
        const bool band = in_band(g, c, l.z);
        return f ? band : (band || (c == d.isx && l.z == d.isz));
    };
    auto fetch = [&](int f, int j, int c) {
        const Fields &F = f ? d.bn : d.fw;
        qvx[f][j] = ld(F.vx, g, l, c + 3);
        qvz[f][j] = ld(F.vz, g, l, c + 4);
        const bool sp = split_of(f, c);
        str_load_w(d.m, F, c * g.nzp + l.zc, l.act, sp, qst[f][j]);
    };
    struct SImg {
        float sxx, szz, sxz, ip, is;
    };
    constexpr int IPD = img_depth(PD);
    SImg qim[IPD] = {};
    auto ifetch = [&](int j, int c) {   // imaging operands, IPD columns ahead
        const int q = img_index(g, l, c);
        if (any_lane(q >= 0)) {
            const int i = c * g.nzp + l.zc, r = q < 0 ? 0 : q;
            qim[j].sxx = d.bk.txx[i];
            qim[j].szz = d.bk.tzz[i];
            qim[j].sxz = d.bk.txz[i];
            qim[j].ip = d.gvp[r];
            qim[j].is = d.gvs[r];
        }
    };
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const Fields &F = f ? d.bn : d.fw;
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            wvx[f][j] = zs(l, ld(F.vx, g, l, s.x0 - 4 + j));
            wvz[f][j] = zs(l, ld(F.vz, g, l, s.x0 - 3 + j));
        }
#pragma unroll
        for (int j = 0; j < PD; ++j)
            if (s.x0 + j < s.x1) fetch(f, j, s.x0 + j);
    }
    if (IMG) {
#pragma unroll
        for (int j = 0; j < IPD; ++j)
            if (s.x0 + j < s.x1) ifetch(j, s.x0 + j);
    }
    const BornSrc none{};
    for (int c0 = s.x0; c0 < s.x1; c0 += PD) {
#pragma unroll
        for (int j = 0; j < PD; ++j) {
            const int c = c0 + j;
            if (c >= s.x1) break;
            const int i = c * g.nzp + l.z;
            SImg im{};
            if (IMG) {
                im = qim[j % IPD];
                if (c + IPD < s.x1) ifetch(j % IPD, c + IPD);
            }
            VGrad v[NF];
            StrCell st[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                wvx[f][7] = zs(l, qvx[f][j]);
                wvz[f][7] = zs(l, qvz[f][j]);
                st[f] = qst[f][j];
                if (c + PD < s.x1) fetch(f, j, c + PD);
                v[f].xx = x_win(wvx[f]);            // Dx- vx
                v[f].zx = x_win(wvz[f]);            // Dx+ vz
                v[f].zz = z_minus(wvz[f][3]);       // Dz- vz (vz at column c)
                v[f].xz = z_plus(wvx[f][4]);        // Dz+ vx (vx at column c)
                shift(wvx[f]);
                shift(wvz[f]);
            }
            if (l.act) {
                const bool sp = split_of(0, c);
                str_math<false>(g, st[0], v[0], sp, dir, none);
                str_store(d.fw, i, st[0], sp);
                if (GRAD) {
                    d.dgx[i] = v[0].xx * g.idx;
                    d.dgz[i] = v[0].zz * g.idz;
                    d.dgxz[i] = v[0].xz * g.idz;
                    d.dgzx[i] = v[0].zx * g.idx;
                }
                if (IMG) {
                    const int q = img_index(g, l, c);
                    if (q >= 0) {
                        float op, os;
                        image_vpvs(v[0].xx * g.idx, v[0].zz * g.idz, v[0].xz * g.idz, v[0].zx * g.idx, im.sxx,
                                   im.szz, im.sxz, im.ip, im.is, op, os);
                        d.gvp[q] = op;
                        d.gvs[q] = os;
                    }
                }
                if (BORN) {
                    const bool band = split_of(1, c);
                    const BornSrc src = born_src<false>(g, d.m, v[0], i, dir);
                    str_math<true>(g, st[NF - 1], v[NF - 1], band, dir, src);
                    str_store(d.bn, i, st[NF - 1], band);
                }
            }
        }
    }
}

// Adjoint stress-like step.
template <int PD>
__global__ void __launch_bounds__(BLK) km_adj_str(Geo g, const DevShot *__restrict__ shots, int dir) {
    Seg s;
    if (!seg_of(g, s)) return;
    const DevShot d = shots[s.shot];   // by value: the field pointers stay in registers (stores cannot alias them)
    const Lane l = lane_of(g, s);
    const Fields &F = d.bk;
    float wvx[8], wvz[8], qvx[PD], qvz[PD];
    AdjStrCell qst[PD] = {};
    auto fetch = [&](int j, int c) {
        qvx[j] = ld(F.vx, g, l, c + 3);
        qvz[j] = ld(F.vz, g, l, c + 4);
        const bool sp = in_band(g, c, l.z);
        adj_str_load_w(d.m, F, c * g.nzp + l.zc, l.act, sp, qst[j]);
    };
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        wvx[j] = zs(l, ld(F.vx, g, l, s.x0 - 4 + j));
        wvz[j] = zs(l, ld(F.vz, g, l, s.x0 - 3 + j));
    }
#pragma unroll
    for (int j = 0; j < PD; ++j)
        if (s.x0 + j < s.x1) fetch(j, s.x0 + j);
    for (int c0 = s.x0; c0 < s.x1; c0 += PD) {
#pragma unroll
        for (int j = 0; j < PD; ++j) {
            const int c = c0 + j;
            if (c >= s.x1) break;
            const bool split = in_band(g, c, l.z);
            const int i = c * g.nzp + l.z;
            wvx[7] = zs(l, qvx[j]);
            wvz[7] = zs(l, qvz[j]);
            AdjStrCell st = qst[j];
            if (c + PD < s.x1) fetch(j, c + PD);
            VGrad v;
            v.xx = x_win(wvx);
            v.zx = x_win(wvz);
            v.zz = z_minus(wvz[3]);
            v.xz = z_plus(wvx[4]);
            shift(wvx);
            shift(wvz);
            if (l.act) {
                adj_str_math(g, st, v, split, dir);
                adj_str_store(F, i, st, split);
            }
        }
    }
}

// Adjoint velocity step with the moduli inside the derivatives; IMAGE: fused
// with the zero-lag imaging from the stored velocity derivatives (RHO: also
// the density image).  The imaging operands have a prefetch ring of their
// own, IPD columns deep (IPD divides PD).
template <int PD, bool IMAGE, bool RHO, int IPD>
__global__ void __launch_bounds__(BLK) km_adj_vel_image(Geo g, const DevShot *__restrict__ shots, int dir) {
    static_assert(IPD >= 1 && PD % IPD == 0, "the imaging ring depth divides the operand ring depth");
    Seg s;
    if (!seg_of(g, s)) return;
    const DevShot d = shots[s.shot];   // by value: the field pointers stay in registers (stores cannot alias them)
    const Lane l = lane_of(g, s);
    const Fields &F = d.bk;
    const Medium &m = d.m;
    // p1 = (lam+2mu) txx + lam tzz on columns c-3..c+4; p2 = mu txz on c-4..c+3;
    // p3 = (lam+2mu) tzz + lam txx at column c.
    // Each column's lam, mu, txx, tzz are loaded once, four columns ahead for
    // p1, and kept: one column later their mu forms p2 (with txz, the only
    // other stencil load), four columns later they form p3.  The raw operands
    // travel in the prefetch ring; products are formed when the column is
    // consumed (a product at load time would wait for the load).
    struct Raw {
        float lam, mu, xx, zz;
    };
    auto RL = [&](int x) {
        const int i = x * g.nzp + l.zc;
        Raw r;
        r.lam = m.lam[i];
        r.mu = m.mu[i];
        r.xx = F.txx[i];
        r.zz = F.tzz[i];
        return r;
    };
    auto P1v = [&](const Raw &r) { return l.zin ? adj_p1(r.lam, r.mu, r.xx, r.zz) : 0.0f; };
    auto P3v = [&](const Raw &r) { return l.zin ? adj_p1(r.lam, r.mu, r.zz, r.xx) : 0.0f; };
    auto P2v = [&](float mu, float xz) { return l.zin ? mu * xz : 0.0f; };
    auto split_of = [&](int c) { return in_band(g, c, l.z) || l.z == d.igz; };
    float w1[8], w2[8];
    Raw qr[PD], hist[4];      // hist[k]: the operands of column c + k
    float qxz[PD];
    VelCell qst[PD] = {};
    ImgCell qim[IPD] = {};
    auto fetch = [&](int j, int c) {
        qr[j] = RL(c + 4);
        qxz[j] = F.txz[(c + 3) * g.nzp + l.zc];
        const bool sp = split_of(c);
        vel_load_w(m, F, c * g.nzp + l.zc, l.act, sp, qst[j]);
    };
    auto ifetch = [&](int j, int c) {
        const int q = img_index(g, l, c);
        img_load(d, c * g.nzp + l.zc, q < 0 ? 0 : q, any_lane(q >= 0), qim[j]);
        if (RHO) img_load_rho(d, c * g.nzp + l.zc, q < 0 ? 0 : q, any_lane(q >= 0), qim[j]);
    };
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const Raw r = RL(s.x0 - 3 + j);
        w1[j] = P1v(r);
        if (j >= 3) hist[j - 3] = r;
        const int i = (s.x0 - 4 + j) * g.nzp + l.zc;
        w2[j] = P2v(m.mu[i], F.txz[i]);
    }
#pragma unroll
    for (int j = 0; j < PD; ++j)
        if (s.x0 + j < s.x1) fetch(j, s.x0 + j);
    if (IMAGE) {
#pragma unroll
        for (int j = 0; j < IPD; ++j)
            if (s.x0 + j < s.x1) ifetch(j, s.x0 + j);
    }
    for (int c0 = s.x0; c0 < s.x1; c0 += PD) {
#pragma unroll
        for (int j = 0; j < PD; ++j) {
            const int c = c0 + j;
            if (c >= s.x1) break;
            const bool split = split_of(c);
            const int i = c * g.nzp + l.z;
            const Raw r4 = qr[j];                // column c + 4
            const Raw r0 = hist[0];              // column c
            w1[7] = P1v(r4);
            w2[7] = P2v(hist[3].mu, qxz[j]);   // mu, txz at column c + 3
            const float p3 = P3v(r0);
#pragma unroll
            for (int k = 0; k < 3; ++k) hist[k] = hist[k + 1];
            hist[3] = r4;
            VelCell st = qst[j];
            ImgCell im{};
            if (IMAGE) {
                im = qim[j % IPD];
                im.sxx = r0.xx;                  // txx, tzz at column c
                im.szz = r0.zz;
            }
            if (c + PD < s.x1) fetch(j, c + PD);
            if (IMAGE && c + IPD < s.x1) ifetch(j % IPD, c + IPD);
            const float sx = x_win(w1);          // Dx+ p1
            const float sxz_x = x_win(w2);       // Dx- p2
            const float sz = z_plus(p3);         // Dz+ p3
            const float sxz_z = z_minus(w2[4]);  // Dz- p2 (p2 at column c)
            shift(w1);
            shift(w2);
            if (l.act) {
                vel_math(g, st, split, dir, sx, sxz_z, sz, sxz_x);
                vel_store(F, i, st, split);
                const int q = IMAGE ? img_index(g, l, c) : -1;
                if (q >= 0) {
                    img_store(d, q, im);
                    if (RHO) {
                        const float lg = -((im.gx + im.gz) * (im.sxx + im.szz));
                        const float mg = __builtin_fmaf(-(im.gxz + im.gzx), im.sxz,
                                                        -2.0f * __builtin_fmaf(im.gx, im.sxx, im.gz * im.szz));
                        d.grho[q] = im.ir + rho_term_fast(g, lg, mg, st.vx, st.vz, im.ox - im.vx, im.oz - im.vz,
                                                          im.vp, im.vs);
                    }
                }
            }
        }
    }
}


// Zero-lag imaging of the snapshot gradient at snapshot k as a column march
// over the interior (the arithmetic of k_snap_image, so both kernel families
// give identical images).  A wave owns 56 interior rows of one shot (plus 4
// halo rows each side, which the region holds) and 8 interior columns; the
// x derivatives come from 8-column register windows of the kept vx (columns
// c-4..c+3, Dx-) and vz (c-3..c+4, Dx+), the z derivatives from DPP shifts of
// the lane's own column value.  ST: snapshot storage (0 FP32, 1 FP16, 2 BF16).
// Block -> (shot, tile): each XCD owns a contiguous range of whole shots.
template <int ST>
__global__ void __launch_bounds__(BLK) km_snap_image(Geo g, const DevShot *__restrict__ shots, int nshots, int k,
                                                     int par) {
    const int b = blockIdx.x;
    const int L = (b & 7) * g.im_per_xcd + (b >> 3);
    const int shot = L / g.im_wg;
    if (shot >= nshots) return;
    const int W = (L - shot * g.im_wg) * (BLK / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (W >= g.im_waves) return;
    const int seg = W / g.im_nbands, band = W - seg * g.im_nbands;
    const int x0 = seg * 8, x1 = min(x0 + 8, g.nx);   // interior columns
    const DevShot d = shots[shot];
    const Fields &B = par ? d.bk2 : d.bk;   // the buffer holding the current adjoint state
    const int lane = threadIdx.x & 63;
    const int jz = band * CM_ROWS - 4 + lane;           // interior row of this lane (halo lanes outside)
    const int R = g.snap_rz;
    const int rz = min(jz + NOP, R - 1);                 // region row (clamped load address)
    const bool act = lane >= 4 && lane < 60 && jz < g.nz;
    const size_t bx = (size_t)k * 2 * g.snap_n, bz = bx + g.snap_n;   // 64-bit offsets of the two fields
    float sx = 1.0f, sz = 1.0f;
    if (ST != 0) {
        sx = __uint_as_float(d.snap_max[2 * (size_t)k]);
        sz = __uint_as_float(d.snap_max[2 * (size_t)k + 1]);
    }
    auto val = [&](size_t base, int rx) -> float {   // unscaled snapshot value at region column rx
        const int o = rx * R + rz;   // within one field of one snapshot
        if (ST == 0) return (d.snap32 + base)[o];
        return unpack16(ST, (d.snap16 + base)[o]);
    };
    // region column of interior column c is c + 4: vx window of column c covers
    // region columns c .. c+7, vz window c+1 .. c+8
    float wx[8], wz[8];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        wx[j] = val(bx, x0 + j);
        wz[j] = val(bz, x0 + 1 + j);
    }
    float nx_ = val(bx, x0 + 7), nz_ = val(bz, x0 + 8);
    for (int c = x0; c < x1; ++c) {
        wx[7] = nx_;
        wz[7] = nz_;
        if (c + 1 < x1) {   // next column's window entries
            nx_ = val(bx, c + 8);
            nz_ = val(bz, c + 9);
        }
        const int i = (c + g.pml) * g.nzp + (jz + g.pml);
        const int q = c * g.nz + jz;
        float sxx = 0.0f, szz = 0.0f, sxz = 0.0f, ip = 0.0f, is = 0.0f, il = 0.0f;
        if (act) {
            sxx = B.txx[i]; szz = B.tzz[i]; sxz = B.txz[i];
            ip = d.gvp[q]; is = d.gvs[q]; il = d.gil[q];
        }
        const float vxc = wx[4], vzc = wz[3];             // kept velocity at column c
        const float gx = x_win(wx) * (sx * g.idx), gz = z_minus(vzc) * (sz * g.idz);
        const float gxz = z_plus(vxc) * (sx * g.idz), gzx = x_win(wz) * (sz * g.idx);
        shift(wx);
        shift(wz);
        if (act) {
            const float vx = vxc * sx, vz = vzc * sz;
            float op, os;
            image_vpvs(gx, gz, gxz, gzx, sxx, szz, sxz, ip, is, op, os);
            d.gvp[q] = op;
            d.gvs[q] = os;
            d.gil[q] = __builtin_fmaf(vx, vx, __builtin_fmaf(vz, vz, il));
        }
    }
}

}  // namespace sfwi
