// Fused forward time step of the FAST mode: velocity and stress updates of
// one time step in a single column march (time blocking of depth one step).
//
// The two half steps of a time step are split into two launches elsewhere
// (km_vel, km_str), and each re-reads what the other wrote: the velocity half
// reads the stresses as stencil inputs, the stress half the velocities, so an
// unsplit cell moves 72 B per step instead of the 52 B of one update.  Here a
// wave owns 48 rows (lanes 8..55 of 64) and a segment of columns [x0, x1),
// and marches x from x0 - 4 to x1 + 4: at column c it forms the new velocity
// on lanes 4..59 from the old stresses (8-column register windows of txx and
// txz, DPP z taps), keeps the new velocity in 9-column windows, and forms the
// new stresses of column c - 4, whose stencil needs the new velocity of
// columns c - 8 .. c.  The velocity of the 4 halo columns on each side and
// of the 4 halo rows is computed again by the neighbouring waves, so every
// wave reads only old values: the velocity totals, the velocity split parts
// and the stress totals live in two buffers (DevShot::fw and fw2), read from
// the buffer of the step's parity and written to the other.  The stress
// split parts are read and written by their owner wave only and stay in
// place (fw2's stress split pointers are fw's).  The old stresses a cell's
// stress update needs are the stencil windows' own values, so an unsplit
// cell's old state is fetched once.  The per-cell arithmetic is vel_math /
// str_math on the same stencil sums as km_vel / km_str, so the results equal
// those kernels' bit for bit.
#pragma once
#include "kernels_march.hpp"

namespace sfwi {

constexpr int FS_ROWS = 48;   // output rows per wave of the fused step (64 lanes minus 2 x 8 halo)

// The buffer holding a shot's forward state after `steps` fused steps.
__device__ inline const Fields &fw_buf(const DevShot &d, int b) { return b ? d.fw2 : d.fw; }

template <bool SAVE, int PD>
__global__ void __launch_bounds__(BLK) km_fwd_step(Geo g, const DevShot *__restrict__ shots, int nshots, int it,
                                                   int par) {
    const int b = blockIdx.x;
    const int L = (b & 7) * g.fs_per_xcd + (b >> 3);
    const int shot = L / g.fs_wg;
    if (shot >= nshots) return;
    const int W = (L - shot * g.fs_wg) * (BLK / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (W >= g.fs_waves) return;
    const int seg = W / g.fs_nbands, band = W - seg * g.fs_nbands;
    const int x0 = NOP + seg * g.fs_xs, x1 = min(x0 + g.fs_xs, g.nxp - NOP);
    const int zb = NOP + band * FS_ROWS;
    const DevShot d = shots[shot];   // by value: the field pointers stay in registers
    const Fields &R = par ? d.fw2 : d.fw;    // the state before the step
    const Fields &Wt = par ? d.fw : d.fw2;   // the state after it
    const int lane = threadIdx.x & 63;
    const int z = zb - 8 + lane;
    const int zc = min(max(z, 0), g.nzp - 1);
    const bool zin = z >= 0 && z < g.nzp;
    const bool vrow = z >= NOP && z < g.nzp - NOP;              // rows the update domain holds
    const bool vlane = lane >= 4 && lane < 60 && vrow;          // velocity formed here (halo rows included)
    const bool own = lane >= 8 && lane < 56 && z < g.nzp - NOP;  // rows this wave writes
    auto colc = [&](int c) { return min(max(c, 0), g.nxp - 1); };
    auto colin = [&](int c) { return c >= 0 && c < g.nxp; };
    auto in_dom = [&](int c) { return c >= NOP && c < g.nxp - NOP; };
    auto ld = [&](cfptr f, int c) { return f[colc(c) * g.nzp + zc]; };
    // per-column operands, PD columns ahead: old stresses entering the windows
    // (txx at c + 4, txz at c + 3, tzz at c), the old velocity state of
    // column c, and the stress split state and moduli of column c - 4
    float qxx[PD], qxz[PD], qzz[PD];
    VelCell qv[PD] = {};
    StrCell qs[PD] = {};
    auto fetch = [&](int j, int c) {
        qxx[j] = ld(R.txx, c + 4);
        qxz[j] = ld(R.txz, c + 3);
        qzz[j] = ld(R.tzz, c);
        const int i = colc(c) * g.nzp + zc;
        vel_load_w(d.m, R, i, vlane && in_dom(c), in_band(g, c, z), qv[j]);
        const int cp = c - 4;
        const bool sown = own && cp >= x0 && cp < x1;
        const bool sp = in_band(g, cp, z) || (cp == d.isx && z == d.isz);
        const int ip = colc(cp) * g.nzp + zc;
        if (sown) { qs[j].lam = d.m.lam[ip]; qs[j].mu = d.m.mu[ip]; }
        if (sown && sp) {
            qs[j].xx_x = R.txx_x[ip]; qs[j].xx_z = R.txx_z[ip]; qs[j].zz_x = R.tzz_x[ip];
            qs[j].zz_z = R.tzz_z[ip]; qs[j].xz_x = R.txz_x[ip]; qs[j].xz_z = R.txz_z[ip];
            qs[j].ax = d.m.dampx[ip]; qs[j].az = d.m.dampz[ip];
        }
    };
    auto zs = [&](int c, float v) { return (zin && colin(c)) ? v : 0.0f; };
    const int cs = x0 - 4, ce = x1 + 4;
    float wxx[8], wxz[8], tzh[5] = {}, vxn[9] = {}, vzn[9] = {};
    float txx_m4 = 0.0f;   // txx of column c - 4 (left the txx window a column ago)
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        wxx[j] = zs(cs - 3 + j, ld(R.txx, cs - 3 + j));
        wxz[j] = zs(cs - 4 + j, ld(R.txz, cs - 4 + j));
    }
#pragma unroll
    for (int j = 0; j < PD; ++j)
        if (cs + j < ce) fetch(j, cs + j);
    const BornSrc none{};
    for (int c0 = cs; c0 < ce; c0 += PD) {
#pragma unroll
        for (int j = 0; j < PD; ++j) {
            const int c = c0 + j;
            if (c >= ce) break;
            wxx[7] = zs(c + 4, qxx[j]);
            wxz[7] = zs(c + 3, qxz[j]);
            tzh[4] = zs(c, qzz[j]);
            VelCell vs = qv[j];
            StrCell ss = qs[j];
            if (c + PD < ce) fetch(j, c + PD);
            // velocity of column c
            const float a = x_win(wxx);        // Dx+ txx
            const float e = x_win(wxz);        // Dx- txz
            const float cz = z_plus(tzh[4]);   // Dz+ tzz
            const float bz = z_minus(wxz[4]);  // Dz- txz (txz at column c)
            float vx_c = 0.0f, vz_c = 0.0f;
            if (vlane && in_dom(c)) {
                const bool sp = in_band(g, c, z);
                vel_math(g, vs, sp, 1, a, bz, cz, e);
                vx_c = vs.vx;
                vz_c = vs.vz;
                if (own && c >= x0 && c < x1) {
                    const int i = c * g.nzp + z;
                    vel_store(Wt, i, vs, sp);
                    if (SAVE && sp) {
                        int sl[4];
                        strip_slots(g, c, z, sl);
                        const fptr base = g.store ? d.stage : d.strip + (size_t)it * 2 * g.sf;
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            if (sl[k] >= 0) {
                                base[sl[k]] = vs.vx;
                                base[g.sf + sl[k]] = vs.vz;
                            }
                    }
                }
            }
            // old stresses of column c - 4, then the windows move on
            const float sxx4 = txx_m4, sxz4 = wxz[0], szz4 = tzh[0];
            txx_m4 = wxx[0];
            shift(wxx);
            shift(wxz);
#pragma unroll
            for (int k = 0; k < 4; ++k) tzh[k] = tzh[k + 1];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                vxn[k] = vxn[k + 1];
                vzn[k] = vzn[k + 1];
            }
            vxn[8] = vx_c;
            vzn[8] = vz_c;
            // stresses of column c - 4 from the new velocities of columns c - 8 .. c
            const int cp = c - 4;
            if (cp >= x0) {
                VGrad v;
                v.xx = x_win(&vxn[0]);       // Dx- vx
                v.zx = x_win(&vzn[1]);       // Dx+ vz
                v.zz = z_minus(vzn[4]);      // Dz- vz (vz at column c - 4)
                v.xz = z_plus(vxn[4]);       // Dz+ vx (vx at column c - 4)
                if (own && cp < x1) {
                    const bool sp = in_band(g, cp, z) || (cp == d.isx && z == d.isz);
                    if (!sp) {
                        ss.txx = sxx4;
                        ss.tzz = szz4;
                        ss.txz = sxz4;
                    }
                    str_math<false>(g, ss, v, sp, 1, none);
                    str_store(Wt, cp * g.nzp + z, ss, sp);
                }
            }
        }
    }
}

// z taps of a column held in LDS (col[t + 4] is lane t's row; 4 pad entries
// each side): the sums of z_plus / z_minus in the same order, so a column
// formed from LDS taps equals one formed from DPP taps bit for bit.
__device__ inline float lz_plus(const float *col, int t, float f) {
    const float p1 = col[t + 5], p2 = col[t + 6], p3 = col[t + 7], p4 = col[t + 8];
    const float m1 = col[t + 3], m2 = col[t + 2], m3 = col[t + 1];
    float r = kC[1] * (p1 - f);
    r = __builtin_fmaf(kC[2], p2 - m1, r);
    r = __builtin_fmaf(kC[3], p3 - m2, r);
    r = __builtin_fmaf(kC[4], p4 - m3, r);
    return r;
}
__device__ inline float lz_minus(const float *col, int t, float f) {
    const float p1 = col[t + 5], p2 = col[t + 6], p3 = col[t + 7];
    const float m1 = col[t + 3], m2 = col[t + 2], m3 = col[t + 1], m4 = col[t];
    float r = kC[1] * (f - m1);
    r = __builtin_fmaf(kC[2], p1 - m2, r);
    r = __builtin_fmaf(kC[3], p2 - m3, r);
    r = __builtin_fmaf(kC[4], p3 - m4, r);
    return r;
}

constexpr int WS_ROWS = BLK - 16;   // output rows of the workgroup fused step (256 lanes minus 2 x 8 halo)
constexpr int WS_PAD = BLK + 8;     // one LDS column: 4 pad entries each side

// The fields of one allocation addressed through a buffer resource: a field
// is the byte offset of its start from the allocation's (the scalar offset
// operand of each access), a cell its byte offset inside the field (one
// vector register that every field at that cell shares).  The workgroup
// marches address a shot's state block and its medium window this way: with a
// 64-bit pointer per field they held more uniform values than a wave has
// scalar registers, the compiler parked the rest in vector-register lanes and
// read them back every column, and every load in a lane-masked branch formed
// its own 64-bit address.  Word 3 (0x00020000) is the gfx9-family raw-buffer
// setting (32-bit data format); the record count is the maximum, so no access
// is range-checked away.  Every offset stays below 4 GiB: the host allows the
// workgroup marches only where a shot's state block does.  Stores carry the
// non-temporal hint (cache policy 2): the new state is read again only by the
// next step, after far more traffic than the caches hold, and written through
// it no longer leaves dirty lines for the following kernels to write back
// (on the 1200 x 600 wave: forward 367 -> 360 us, adjoint 358 -> 353, and the
// imaging and snapshot-store launches between them 162 -> 136 and 61 -> 50).
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ explicit Buf(const void *base)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, -1, 0x00020000)) {}
    __device__ float ld(unsigned f, int i) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (unsigned)i * 4u, f, 0));
    }
    __device__ void st(unsigned f, int i, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (unsigned)i * 4u, f, 2);
    }
};
__device__ inline unsigned boff(const void *p, const void *base) {
    return (unsigned)((const char *)p - (const char *)base);
}
// Byte offsets of one state buffer's totals and velocity split parts, and of
// the stress split parts (which have one buffer only), from the state block.
struct FOff {
    unsigned vx, vz, txx, tzz, txz, vx_x, vx_z, vz_x, vz_z;
};
struct SOff {
    unsigned txx_x, txx_z, tzz_x, tzz_z, txz_x, txz_z;
};
__device__ inline FOff foff(const Fields &F, const void *b) {
    return FOff{boff(F.vx, b),   boff(F.vz, b),   boff(F.txx, b),  boff(F.tzz, b), boff(F.txz, b),
                boff(F.vx_x, b), boff(F.vx_z, b), boff(F.vz_x, b), boff(F.vz_z, b)};
}
__device__ inline SOff soff(const Fields &F, const void *b) {
    return SOff{boff(F.txx_x, b), boff(F.txx_z, b), boff(F.tzz_x, b),
                boff(F.tzz_z, b), boff(F.txz_x, b), boff(F.txz_z, b)};
}
// Byte offsets of the medium fields the marches read, from the window's start.
struct MOff {
    unsigned lam, mu, bu, dampx, dampz;
};
__device__ inline MOff moff(const Medium &m) {
    return MOff{boff(m.lam, m.vp), boff(m.mu, m.vp), boff(m.bu, m.vp), boff(m.dampx, m.vp), boff(m.dampz, m.vp)};
}
// The value a ring slot holds on the lanes that skip its conditional load: no
// such lane reads the slot, and leaving it unspecified lets each load write
// its ring register directly (keeping the slot's previous value there costs a
// register copy per operand and column).
__device__ inline float unset() {
    float f;
    asm volatile("" : "=v"(f));   // a register as it stands: no instruction
    return f;
}
template <class T> __device__ inline void unset_all(T &v) {
    static_assert(sizeof(T) % sizeof(float) == 0, "a cell of floats");
    float *f = reinterpret_cast<float *>(&v);
#pragma unroll
    for (unsigned k = 0; k < sizeof(T) / sizeof(float); ++k) f[k] = unset();
}
// vel_load_w, vel_store, str_store and adj_str_store through the two bases.
__device__ inline void vel_load_b(const Buf &S, const Buf &M, const MOff &mo, const FOff &F, int i, bool act,
                                  bool sp, VelCell &s) {
    if (act) s.bu = M.ld(mo.bu, i);
    if (act && !sp) { s.vx = S.ld(F.vx, i); s.vz = S.ld(F.vz, i); }
    if (act && sp) {
        s.vx_x = S.ld(F.vx_x, i); s.vx_z = S.ld(F.vx_z, i); s.vz_x = S.ld(F.vz_x, i); s.vz_z = S.ld(F.vz_z, i);
        s.ax = M.ld(mo.dampx, i); s.az = M.ld(mo.dampz, i);
    }
}
__device__ inline void vel_store_b(const Buf &S, const FOff &F, int i, const VelCell &s, bool split) {
    if (split) {
        S.st(F.vx_x, i, s.vx_x); S.st(F.vx_z, i, s.vx_z); S.st(F.vz_x, i, s.vz_x); S.st(F.vz_z, i, s.vz_z);
    }
    S.st(F.vx, i, s.vx);
    S.st(F.vz, i, s.vz);
}
__device__ inline void str_store_b(const Buf &S, const FOff &F, const SOff &P, int i, const StrCell &s,
                                   bool split) {
    if (split) {
        S.st(P.txx_x, i, s.xx_x); S.st(P.txx_z, i, s.xx_z); S.st(P.tzz_x, i, s.zz_x);
        S.st(P.tzz_z, i, s.zz_z); S.st(P.txz_x, i, s.xz_x); S.st(P.txz_z, i, s.xz_z);
    }
    S.st(F.txx, i, s.txx);
    S.st(F.tzz, i, s.tzz);
    S.st(F.txz, i, s.txz);
}
__device__ inline void adj_str_store_b(const Buf &S, const FOff &F, const SOff &P, int i, const AdjStrCell &s,
                                       bool split) {
    if (split) {
        S.st(P.txx_x, i, s.xx_x); S.st(P.tzz_z, i, s.zz_z); S.st(P.txz_x, i, s.xz_x); S.st(P.txz_z, i, s.xz_z);
    }
    S.st(F.txx, i, s.txx);
    S.st(F.tzz, i, s.tzz);
    S.st(F.txz, i, s.txz);
}

// Fused forward time step with the z halo shared inside the workgroup.
// The four waves of a workgroup stack along z: thread t owns row
// Z0 - 8 + t of one column strip [x0, x1), so a workgroup covers 256
// consecutive rows and writes the 240 rows Z0 .. Z0 + 239.  The z taps come
// from LDS instead of DPP shifts: at column c every thread puts the old tzz
// and txz of column c and the new vx, vz of column c - 1 into a
// double-buffered LDS column, one barrier, then reads the eight neighbours of
// its row, so the rows where two waves meet are formed once, and only the
// workgroup's top and bottom 8 rows are halo (the wave-level fused march
// spends 16 of every 64 lanes on it).  Velocity z derivatives are formed a
// column after their velocity and wait three columns in registers for the
// stress update of their column.  The x direction, the buffers and the
// per-cell arithmetic are those of km_fwd_step, so the results equal the
// half-step pair's bit for bit.  With wav set the step also adds the
// explosive source (k_source) and records the receivers' new velocity
// (k_record), so a time step of the forward solve is one launch.
template <bool SAVE, int PD>
__global__ void __launch_bounds__(BLK) km_fwd_step_wg(Geo g, const DevShot *__restrict__ shots, int nshots,
                                                      int it, int par, int smax, const float *__restrict__ wav,
                                                      int ts) {
    __shared__ float zt[2][4][WS_PAD];
    __shared__ float red[2][BLK / 64];
    const int b = blockIdx.x;
    const int L = (b & 7) * g.ws_per_xcd + (b >> 3);
    const int shot = L / g.ws_wg;
    if (shot >= nshots) return;   // uniform over the workgroup: no barrier is left waiting
    const int w = L - shot * g.ws_wg;
    const int seg = w / g.ws_ntiles, tile = w - seg * g.ws_ntiles;
    const int x0 = NOP + seg * g.ws_xs, x1 = min(x0 + g.ws_xs, g.nxp - NOP);
    const int Z0 = NOP + tile * WS_ROWS;
    const DevShot d = shots[shot];   // by value: the field pointers stay in registers
    const Buf S(d.fw.vx), M(d.m.vp);   // the shot's state block (fw.vx leads it) and medium window
    const FOff R = foff(par ? d.fw2 : d.fw, d.fw.vx);    // the state before the step
    const FOff Wt = foff(par ? d.fw : d.fw2, d.fw.vx);   // the state after it
    const SOff P = soff(d.fw, d.fw.vx);                  // the stress split parts (one buffer)
    const MOff mo = moff(d.m);
    const int t = threadIdx.x;
    const int z = Z0 - 8 + t;
    const int zc = min(max(z, 0), g.nzp - 1);
    const bool zin = z >= 0 && z < g.nzp;
    const bool vrow = z >= NOP && z < g.nzp - NOP;            // rows the update domain holds
    const bool vlane = t >= 4 && t < BLK - 4 && vrow;         // velocity formed here (halo rows included)
    const bool own = t >= 8 && t < BLK - 8 && z < g.nzp - NOP;  // rows this workgroup writes
    auto colc = [&](int c) { return min(max(c, 0), g.nxp - 1); };
    auto colin = [&](int c) { return c >= 0 && c < g.nxp; };
    auto in_dom = [&](int c) { return c >= NOP && c < g.nxp - NOP; };
    auto ld = [&](unsigned f, int c) { return S.ld(f, colc(c) * g.nzp + zc); };
    float qxx[PD], qxz[PD], qzz[PD];
    int qtr[PD];   // trace recorded at column c (-1: none), through the ring like every operand
    VelCell qv[PD] = {};
    StrCell qs[PD] = {};
    auto fetch = [&](int j, int c) {
        // the column's trace index is the same on every lane; kept a per-lane
        // value (the empty asm hides that), or the compiler moves it to a
        // scalar register right at the load, which waits for the load
        int cj = colc(c);
        asm volatile("" : "+v"(cj));
        qtr[j] = d.rtr[cj];
        qxx[j] = ld(R.txx, c + 4);
        qxz[j] = ld(R.txz, c + 3);
        qzz[j] = ld(R.tzz, c);
        const int i = colc(c) * g.nzp + zc;
        unset_all(qv[j]);   // lanes that load no variant never read the slot
        vel_load_b(S, M, mo, R, i, vlane && in_dom(c), in_band(g, c, z), qv[j]);
        const int cp = c - 4;
        const bool sown = own && cp >= x0 && cp < x1;
        const bool sp = in_band(g, cp, z) || (cp == d.isx && z == d.isz);
        const int ip = colc(cp) * g.nzp + zc;
        unset_all(qs[j]);
        if (sown) { qs[j].lam = M.ld(mo.lam, ip); qs[j].mu = M.ld(mo.mu, ip); }
        if (sown && sp) {
            qs[j].xx_x = S.ld(P.txx_x, ip); qs[j].xx_z = S.ld(P.txx_z, ip); qs[j].zz_x = S.ld(P.tzz_x, ip);
            qs[j].zz_z = S.ld(P.tzz_z, ip); qs[j].xz_x = S.ld(P.txz_x, ip); qs[j].xz_z = S.ld(P.txz_z, ip);
            qs[j].ax = M.ld(mo.dampx, ip); qs[j].az = M.ld(mo.dampz, ip);
        }
    };
    auto zs = [&](int c, float v) { return (zin && colin(c)) ? v : 0.0f; };
    // the explosive source of step ts, added where the old txx, tzz of the
    // source cell enter (k_source's addition, made by every reader instead of
    // stored; a select, so no other value changes, not even a zero's sign)
    const float wsrc = !wav ? 0.0f : d.wav ? d.wav[ts] : wav[ts];
    const bool srow = wav && z == d.isz;   // wav NULL: k_source has added it
    auto src = [&](int c, float v) { return (srow && c == d.isx) ? v + wsrc : v; };
    if (t < 8) {   // pad entries: read only by halo rows, whose results are dropped
        const int e = t < 4 ? t : BLK + t;
#pragma unroll
        for (int k = 0; k < 8; ++k) zt[k >> 2][k & 3][e] = 0.0f;
    }
    const int cs = x0 - 4, ce = x1 + 4;
    float wxx[8], wxz[8], tzh[5] = {}, vxn[9] = {}, vzn[9] = {};
    float hxz[4] = {}, hzz[4] = {};   // Dz+ vx, Dz- vz of columns c - 4 .. c - 1
    float txx_m4 = 0.0f;   // txx of column c - 4 (left the txx window a column ago)
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        wxx[j] = src(cs - 3 + j, zs(cs - 3 + j, ld(R.txx, cs - 3 + j)));
        wxz[j] = zs(cs - 4 + j, ld(R.txz, cs - 4 + j));
    }
#pragma unroll
    for (int j = 0; j < PD; ++j)
        if (cs + j < ce) fetch(j, cs + j);
    const BornSrc none{};
    // smax: a snapshot step; keep the max |v| of the new velocity on the
    // snapshot region (interior plus a 4-cell halo) over the cells this
    // workgroup stores, for the snapshot's max-abs scale
    const bool zreg = z >= g.pml - NOP && z < g.pml + g.nz + NOP;
    float mvx = 0.0f, mvz = 0.0f;
    for (int c0 = cs; c0 < ce; c0 += PD) {
#pragma unroll
        for (int j = 0; j < PD; ++j) {
            const int c = c0 + j;
            if (c >= ce) break;   // uniform over the workgroup
            wxx[7] = src(c + 4, zs(c + 4, qxx[j]));
            wxz[7] = zs(c + 3, qxz[j]);
            tzh[4] = src(c, zs(c, qzz[j]));
            VelCell vs = qv[j];
            StrCell ss = qs[j];
            const int trc = qtr[j];
            if (c + PD < ce) fetch(j, c + PD);
            float(*col)[WS_PAD] = zt[c & 1];
            col[0][t + 4] = tzh[4];   // old tzz, txz of column c
            col[1][t + 4] = wxz[4];
            col[2][t + 4] = vxn[8];   // new vx, vz of column c - 1
            col[3][t + 4] = vzn[8];
            __syncthreads();
            // velocity of column c
            const float a = x_win(wxx);                    // Dx+ txx
            const float e = x_win(wxz);                    // Dx- txz
            const float cz = lz_plus(col[0], t, tzh[4]);   // Dz+ tzz
            const float bz = lz_minus(col[1], t, wxz[4]);  // Dz- txz (txz at column c)
            hxz[3] = lz_plus(col[2], t, vxn[8]);           // Dz+ vx of column c - 1
            hzz[3] = lz_minus(col[3], t, vzn[8]);          // Dz- vz of column c - 1
            float vx_c = 0.0f, vz_c = 0.0f;
            if (vlane && in_dom(c)) {
                const bool sp = in_band(g, c, z);
                vel_math(g, vs, sp, 1, a, bz, cz, e);
                vx_c = vs.vx;
                vz_c = vs.vz;
                if (own && c >= x0 && c < x1) {
                    const int i = c * g.nzp + z;
                    vel_store_b(S, Wt, i, vs, sp);
                    if (wav && z == d.igz && trc >= 0) {   // the receiver here records (k_record)
                        const size_t o = (size_t)trc * g.nt + ts;
                        d.rz[o] = vs.vz;
                        d.rx[o] = vs.vx;
                    }
                    if (smax && zreg && c >= g.pml - NOP && c < g.pml + g.nx + NOP) {
                        mvx = fmaxf(mvx, fabsf(vs.vx));
                        mvz = fmaxf(mvz, fabsf(vs.vz));
                    }
                    if (SAVE && sp) {
                        int sl[4];
                        strip_slots(g, c, z, sl);
                        const fptr base = g.store ? d.stage : d.strip + (size_t)it * 2 * g.sf;
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            if (sl[k] >= 0) {
                                base[sl[k]] = vs.vx;
                                base[g.sf + sl[k]] = vs.vz;
                            }
                    }
                }
            }
            // old stresses of column c - 4, then the windows move on
            const float sxx4 = txx_m4, sxz4 = wxz[0], szz4 = tzh[0];
            const float dxz4 = hxz[0], dzz4 = hzz[0];
            txx_m4 = wxx[0];
            shift(wxx);
            shift(wxz);
#pragma unroll
            for (int k = 0; k < 4; ++k) tzh[k] = tzh[k + 1];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                hxz[k] = hxz[k + 1];
                hzz[k] = hzz[k + 1];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                vxn[k] = vxn[k + 1];
                vzn[k] = vzn[k + 1];
            }
            vxn[8] = vx_c;
            vzn[8] = vz_c;
            // stresses of column c - 4 from the new velocities of columns c - 8 .. c
            const int cp = c - 4;
            if (cp >= x0 && own && cp < x1) {
                VGrad v;
                v.xx = x_win(&vxn[0]);   // Dx- vx
                v.zx = x_win(&vzn[1]);   // Dx+ vz
                v.zz = dzz4;             // Dz- vz (vz at column c - 4)
                v.xz = dxz4;             // Dz+ vx (vx at column c - 4)
                const bool sp = in_band(g, cp, z) || (cp == d.isx && z == d.isz);
                if (!sp) {
                    ss.txx = sxx4;
                    ss.tzz = szz4;
                    ss.txz = sxz4;
                }
                str_math<false>(g, ss, v, sp, 1, none);
                str_store_b(S, Wt, P, cp * g.nzp + z, ss, sp);
            }
        }
    }
    if (smax) {   // this workgroup's partial maxima (k_snap_scale reduces them)
        mvx = wave_max(mvx);
        mvz = wave_max(mvz);
        if ((t & 63) == 0) {
            red[0][t >> 6] = mvx;
            red[1][t >> 6] = mvz;
        }
        __syncthreads();
        if (t < 2) {
            float mm = red[t][0];
#pragma unroll
            for (int k = 1; k < BLK / 64; ++k) mm = fmaxf(mm, red[t][k]);
            d.snap_part[t * g.snap_np + w] = mm;
        }
    }
}

// Fused adjoint time step of the snapshot gradient, the workgroup march of
// km_fwd_step_wg applied to the adjoint pair: the adjoint velocity of column
// c (km_adj_vel_image's arithmetic: x taps of p1 = (lam + 2 mu) txx + lam tzz
// and p2 = mu txz from 8-column windows, z taps of p3 = (lam + 2 mu) tzz +
// lam txx and of p2 from LDS), then the adjoint stress-like update of column
// c - 4 (km_adj_str's arithmetic) from the new adjoint velocities of columns
// c - 8 .. c.  Each column's lam, mu, txx, tzz are loaded once, four columns
// ahead, and kept for p2 (mu, one column later) and p3 (four columns later).
// The adjoint totals and velocity split parts live in two buffers
// (DevShot::bk and bk2, the step's parity picks the one read), the stress
// split parts stay in place; results equal km_adj_vel_image + km_adj_str bit
// for bit.  Used by the snapshot gradient, whose imaging reads the adjoint
// stresses before the step (neither half step of the pair had changed them
// at that point either).  With inj set it also adds the step's adjoint
// source (k_inject_add) from the column table k_inject_table built, to the
// split parts of the receiver row (its totals are recomputed from them).
template <int PD>
__global__ void __launch_bounds__(BLK) km_adj_step_wg(Geo g, const DevShot *__restrict__ shots, int nshots,
                                                      int par, int dir, int it, int inj) {
    __shared__ float zt[2][4][WS_PAD];
    // old txx, tzz, txz of each row over the last 9 columns, kept from their
    // load as stencil operands (txx, tzz of column c + 4, txz of c + 3) for the
    // stress update of their column eight and seven columns later; a thread
    // reads only its own row, so no barrier guards the ring
    __shared__ float ring[3][9][BLK];
    const int b = blockIdx.x;
    const int L = (b & 7) * g.ws_per_xcd + (b >> 3);
    const int shot = L / g.ws_wg;
    if (shot >= nshots) return;   // uniform over the workgroup: no barrier is left waiting
    const int w = L - shot * g.ws_wg;
    const int seg = w / g.ws_ntiles, tile = w - seg * g.ws_ntiles;
    const int x0 = NOP + seg * g.ws_xs, x1 = min(x0 + g.ws_xs, g.nxp - NOP);
    const int Z0 = NOP + tile * WS_ROWS;
    const DevShot d = shots[shot];   // by value: the field pointers stay in registers
    const Buf S(d.fw.vx), M(d.m.vp);   // the shot's state block (fw.vx leads it) and medium window
    const FOff R = foff(par ? d.bk2 : d.bk, d.fw.vx);    // the adjoint state before the step
    const FOff Wt = foff(par ? d.bk : d.bk2, d.fw.vx);   // the state after it
    const SOff P = soff(d.bk, d.fw.vx);                  // the stress split parts (one buffer)
    const MOff mo = moff(d.m);
    const int t = threadIdx.x;
    const int z = Z0 - 8 + t;
    const int zc = min(max(z, 0), g.nzp - 1);
    const bool zin = z >= 0 && z < g.nzp;
    const bool vrow = z >= NOP && z < g.nzp - NOP;            // rows the update domain holds
    const bool vlane = t >= 4 && t < BLK - 4 && vrow;         // velocity formed here (halo rows included)
    const bool own = t >= 8 && t < BLK - 8 && z < g.nzp - NOP;  // rows this workgroup writes
    auto colc = [&](int c) { return min(max(c, 0), g.nxp - 1); };
    auto colin = [&](int c) { return c >= 0 && c < g.nxp; };
    auto in_dom = [&](int c) { return c >= NOP && c < g.nxp - NOP; };
    auto ok = [&](int c) { return zin && colin(c); };
    struct Raw {
        float lam, mu, xx, zz;
    };
    // per-column operands, PD columns ahead: lam, mu, txx, tzz of column c + 4,
    // txz of column c + 3, the adjoint velocity state of column c, the split
    // state of the stress column c - 4
    Raw qr[PD];
    float qxz[PD], qaz[PD], qax[PD];   // qaz, qax: the adjoint source at column c (inj)
    // Own-cell state in shared slots: a lane uses either its totals or its
    // split parts, never both, and a VGPR is per lane, so one set of registers
    // holds whichever the lane needs (adjoint velocity: bu, then vx, vz or
    // vx_x, vx_z, vz_x, vz_z, ax, az; stress column: txx, tzz, txz or xx_x,
    // zz_z, xz_x, xz_z, ax, az).  Five registers fewer per ring slot.
    float qv[PD][7] = {}, qs[PD][6] = {};
    const bool rrow = inj && z == d.igz;   // the receiver row takes the adjoint source (k_inject_add)
    const size_t ia = (size_t)it * 2 * g.nxp;
    auto fetch = [&](int j, int c) {
        if (inj) {   // one address per wave and column
            qaz[j] = d.ainj[ia + colc(c)];
            qax[j] = d.ainj[ia + g.nxp + colc(c)];
        }
        const int i4 = colc(c + 4) * g.nzp + zc;
        qr[j].lam = M.ld(mo.lam, i4);
        qr[j].mu = M.ld(mo.mu, i4);
        qr[j].xx = S.ld(R.txx, i4);
        qr[j].zz = S.ld(R.tzz, i4);
        qxz[j] = S.ld(R.txz, colc(c + 3) * g.nzp + zc);
        const int i = colc(c) * g.nzp + zc;
        const bool va = vlane && in_dom(c), vsp = in_band(g, c, z) || z == d.igz;
        float *u = qv[j];
#pragma unroll
        for (int k = 0; k < 7; ++k) u[k] = unset();   // lanes that load no variant never read the slot
        if (va) u[6] = M.ld(mo.bu, i);
        if (va && !vsp) { u[0] = S.ld(R.vx, i); u[1] = S.ld(R.vz, i); }
        if (va && vsp) {
            u[0] = S.ld(R.vx_x, i); u[1] = S.ld(R.vx_z, i); u[2] = S.ld(R.vz_x, i); u[3] = S.ld(R.vz_z, i);
            u[4] = M.ld(mo.dampx, i); u[5] = M.ld(mo.dampz, i);
        }
        const int cp = c - 4;
        const bool sown = own && cp >= x0 && cp < x1, ssp = in_band(g, cp, z);
        const int ip = colc(cp) * g.nzp + zc;
        float *q = qs[j];
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k] = unset();
        if (sown && ssp) {   // (the totals of an unsplit cell come from the ring)
            q[0] = S.ld(P.txx_x, ip); q[1] = S.ld(P.tzz_z, ip); q[2] = S.ld(P.txz_x, ip); q[3] = S.ld(P.txz_z, ip);
            q[4] = M.ld(mo.dampx, ip); q[5] = M.ld(mo.dampz, ip);
        }
    };
    auto P1 = [&](const Raw &r, int c) { return ok(c) ? adj_p1(r.lam, r.mu, r.xx, r.zz) : 0.0f; };
    auto P3 = [&](const Raw &r, int c) { return ok(c) ? adj_p1(r.lam, r.mu, r.zz, r.xx) : 0.0f; };
    auto P2 = [&](float mu, float xz, int c) { return ok(c) ? mu * xz : 0.0f; };
    if (t < 8) {   // pad entries: read only by halo rows, whose results are dropped
        const int e = t < 4 ? t : BLK + t;
#pragma unroll
        for (int k = 0; k < 8; ++k) zt[k >> 2][k & 3][e] = 0.0f;
    }
    const int cs = x0 - 4, ce = x1 + 4;
    // w1: p1 of c-3 .. c+4; w2: p2 of c-4 .. c+3; hist[k]: lam, mu, txx, tzz of
    // column c+k (k = 0..3); vxn / vzn: new adjoint velocity of c-8 .. c.  The
    // stress update of column c - 4 takes that column's old stresses from the
    // LDS ring (reloading them cost 12 B per cell; keeping them eight columns
    // in registers costs a wave per SIMD).
    float w1[8], w2[8], vxn[9] = {}, vzn[9] = {};
    float hxz[4] = {}, hzz[4] = {};   // Dz+ vx, Dz- vz of columns c - 4 .. c - 1
    Raw hist[4];
#pragma unroll
    for (int k = 0; k < 7; ++k) {   // p1 of cs-3 .. cs+3, p2 of cs-4 .. cs+2
        const int x1c = cs - 3 + k, x2c = cs - 4 + k;
        const int i1 = colc(x1c) * g.nzp + zc, i2 = colc(x2c) * g.nzp + zc;
        Raw r{M.ld(mo.lam, i1), M.ld(mo.mu, i1), S.ld(R.txx, i1), S.ld(R.tzz, i1)};
        if (!ok(x1c)) r = Raw{0.0f, 0.0f, 0.0f, 0.0f};
        w1[k] = P1(r, x1c);
        if (k >= 3) hist[k - 3] = r;   // columns cs .. cs+3
        w2[k] = P2(M.ld(mo.mu, i2), S.ld(R.txz, i2), x2c);
    }
#pragma unroll
    for (int j = 0; j < PD; ++j)
        if (cs + j < ce) fetch(j, cs + j);
    for (int c0 = cs; c0 < ce; c0 += PD) {
#pragma unroll
        for (int j = 0; j < PD; ++j) {
            const int c = c0 + j;
            if (c >= ce) break;   // uniform over the workgroup
            Raw r4 = qr[j];                       // column c + 4
            {   // the loaded values, before the masking below
                const int sw = (c + 4) % 9, sz3 = (c + 3) % 9;
                ring[0][sw][t] = r4.xx;
                ring[1][sw][t] = r4.zz;
                ring[2][sz3][t] = qxz[j];
            }
            if (!ok(c + 4)) r4 = Raw{0.0f, 0.0f, 0.0f, 0.0f};
            w1[7] = P1(r4, c + 4);
            w2[7] = P2(hist[3].mu, qxz[j], c + 3);   // mu, txz of column c + 3
            const float p3 = P3(hist[0], c);         // column c
            // each lane's update reads only the variant its cell uses
            VelCell vs;
            vs.vx = vs.vx_x = qv[j][0];
            vs.vz = vs.vx_z = qv[j][1];
            vs.vz_x = qv[j][2]; vs.vz_z = qv[j][3]; vs.ax = qv[j][4]; vs.az = qv[j][5]; vs.bu = qv[j][6];
            AdjStrCell ss;
            ss.txx = ss.xx_x = qs[j][0];
            ss.tzz = ss.zz_z = qs[j][1];
            ss.txz = ss.xz_x = qs[j][2];
            ss.xz_z = qs[j][3]; ss.ax = qs[j][4]; ss.az = qs[j][5];
            if (rrow) {   // receiver row cells are split: their parts take the source (-0 where none)
                vs.vz_x += qaz[j];
                vs.vx_x += qax[j];
            }
            if (c + PD < ce) fetch(j, c + PD);
            float(*col)[WS_PAD] = zt[c & 1];
            col[0][t + 4] = p3;       // p3, p2 of column c
            col[1][t + 4] = w2[4];
            col[2][t + 4] = vxn[8];   // new vx, vz of column c - 1
            col[3][t + 4] = vzn[8];
            __syncthreads();
            // adjoint velocity of column c
            const float sx = x_win(w1);                       // Dx+ p1
            const float sxz_x = x_win(w2);                    // Dx- p2
            const float sz = lz_plus(col[0], t, p3);          // Dz+ p3
            const float sxz_z = lz_minus(col[1], t, w2[4]);   // Dz- p2 (p2 at column c)
            hxz[3] = lz_plus(col[2], t, vxn[8]);              // Dz+ vx of column c - 1
            hzz[3] = lz_minus(col[3], t, vzn[8]);             // Dz- vz of column c - 1
            float vx_c = 0.0f, vz_c = 0.0f;
            if (vlane && in_dom(c)) {
                const bool sp = in_band(g, c, z) || z == d.igz;
                vel_math(g, vs, sp, dir, sx, sxz_z, sz, sxz_x);
                vx_c = vs.vx;
                vz_c = vs.vz;
                if (own && c >= x0 && c < x1) vel_store_b(S, Wt, c * g.nzp + z, vs, sp);
            }
            // the windows move on
            const float dxz4 = hxz[0], dzz4 = hzz[0];
            shift(w1);
            shift(w2);
#pragma unroll
            for (int k = 0; k < 3; ++k) hist[k] = hist[k + 1];
            hist[3] = r4;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                hxz[k] = hxz[k + 1];
                hzz[k] = hzz[k + 1];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                vxn[k] = vxn[k + 1];
                vzn[k] = vzn[k + 1];
            }
            vxn[8] = vx_c;
            vzn[8] = vz_c;
            // adjoint stresses of column c - 4 from the new velocities of columns c - 8 .. c
            const int cp = c - 4;
            if (cp >= x0 && own && cp < x1) {
                VGrad v;
                v.xx = x_win(&vxn[0]);   // Dx- vx
                v.zx = x_win(&vzn[1]);   // Dx+ vz
                v.zz = dzz4;             // Dz- vz (vz at column c - 4)
                v.xz = dxz4;             // Dz+ vx (vx at column c - 4)
                const bool sp = in_band(g, cp, z);
                if (!sp) {
                    const int sr = (c + 5) % 9;   // column c - 4
                    ss.txx = ring[0][sr][t];
                    ss.tzz = ring[1][sr][t];
                    ss.txz = ring[2][sr][t];
                }
                adj_str_math(g, ss, v, sp, dir);
                adj_str_store_b(S, Wt, P, cp * g.nzp + z, ss, sp);
            }
        }
    }
}

// Copy a shot's forward state (velocity and stress totals, velocity split
// parts) from the second buffer to the first, after an odd number of fused
// steps, so the kernels that follow find it in DevShot::fw.
__global__ void k_fw_from_fw2(Geo g, const DevShot *__restrict__ shots) {
    const DevShot &d = shots[blockIdx.y];
    const int n = g.nxp * g.nzp;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        d.fw.vx[i] = d.fw2.vx[i]; d.fw.vz[i] = d.fw2.vz[i];
        d.fw.txx[i] = d.fw2.txx[i]; d.fw.tzz[i] = d.fw2.tzz[i]; d.fw.txz[i] = d.fw2.txz[i];
        d.fw.vx_x[i] = d.fw2.vx_x[i]; d.fw.vx_z[i] = d.fw2.vx_z[i];
        d.fw.vz_x[i] = d.fw2.vz_x[i]; d.fw.vz_z[i] = d.fw2.vz_z[i];
    }
}

}  // namespace sfwi
