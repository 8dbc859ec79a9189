// Host side of libsfwi_amd: context, device memory, batch orchestration and
// the C ABI declared in include/sfwi_abi.h.
//
// One context = one GPU, one HIP stream.  A solve runs a batch of up to
// max_batch shots: each shot owns a slot (its wavefields, records, strips and
// gradient accumulators in HBM) and every stencil launch advances all shots of
// the batch by one half step.  Padded medium windows are built on the device
// once per window position and shared by every shot that uses it.
#include "sfwi_abi.h"
#include "kernels.hpp"
#include "kernels_march.hpp"
#include "kernels_step.hpp"
#include "ot.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace sfwi;

// --------------------------------------------- static OT workspace ---
// The workspace of sfwi_ot_misfit_static, which allocates nothing: one
// file-scope device array per GPU, held until the process ends.  Layout:
// [flag int, total double] | job table | per-job misfits | staging of the four
// gathers | FP64 merge-walk workspace.  Limits: at most NA_MAX_SAMPLES samples
// per component (ntr * nt) and NA_MAX_TRACES traces.
namespace {
constexpr size_t NA_MAX_SAMPLES = size_t(1) << 18;
constexpr int NA_MAX_TRACES = 4096;
constexpr size_t na_align(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t NA_JOBS = 256;
constexpr size_t NA_MIS = NA_JOBS + na_align(2 * NA_MAX_TRACES * sizeof(OtJob));
constexpr size_t NA_IO = NA_MIS + na_align(2 * NA_MAX_TRACES * sizeof(double));
constexpr size_t NA_WS = NA_IO + na_align(4 * NA_MAX_SAMPLES * sizeof(float));
constexpr size_t NA_WS_BYTES = na_align(2 * 4 * (NA_MAX_SAMPLES + 1) * sizeof(double));
constexpr size_t NA_BYTES = NA_WS + NA_WS_BYTES;
}  // namespace

static __device__ __align__(256) char scratch[NA_BYTES];

namespace {

// Device address of `scratch` per GPU, from hipGetSymbolAddress with that GPU
// current (the array's name is no device address in host code).
std::mutex g_scratch_m;
constexpr int NA_MAX_DEVICES = 64;
char *g_scratch[NA_MAX_DEVICES] = {};

// Message of the last failed call without a context (create, server calls):
// one per host thread, so concurrent callers never share or clobber it.
thread_local std::string g_create_error;

struct Slot {
    char *block = nullptr;      // fw(15) + fw2(9) + bk(15) + derivatives(4) [+ v(it+1)(2)] [+ bk2(9)] fields
    char *born = nullptr;       // bn(15) fields
    char *born2 = nullptr;      // bn2(15) fields (double scattering)
    char *rec = nullptr;        // rz rx r2z r2x oz ox, ints, accumulators, trace sums
    float *strip = nullptr;     // FP32 [ring][2][sf], or 16-bit: scales [ring][2], stage [2][sf], [ring][2][sf]
    char *hstrip = nullptr;     // host strips (cpu_mem, pinned): FP32 [nt][2][sf], or 16-bit [nt][2][sf] + scales [nt][2]
    char *snap = nullptr;       // velocity snapshots: max bits [nsnap][2] (256-B aligned), then FP32 or 16-bit values
    float *ainj = nullptr;      // adjoint source by step and column [nt][2][nxp] (fused adjoint of snapshot gradients)
    Fields fw{}, fw2{}, bk{}, bn{}, bn2{}, bk2{};
    float *dg[4]{};
    float *rz, *rx, *r2z, *r2x, *oz, *ox;
    int *igx, *inj, *inj_off, *rtr;
    float *gvp, *gvs, *gil, *grho;
    float *vold[2]{};
    double *trace_sum;
    int obs_ntr = 0;                 // traces of the observed records resident in oz/ox (0: none)
    std::vector<int> h_igx, h_inj, h_off, h_rtr;   // host copies kept alive for async uploads
};

struct ObsSet {        // observed traces resident under an id (sfwi_obs_put)
    float *base = nullptr;   // [ntr][nt] vz, then [ntr][nt] vx
    int ntr = 0;
};

struct Window {
    float *base = nullptr;
    Medium m{};
    bool has_pert = false;
};

// Solves: forward, forward + misfit, gradient, Born forward, Born gradient,
// Born forward with double scattering, Gauss-Newton Hessian-vector product.
enum Kind { K_FWD, K_FWD_MISFIT, K_GRAD, K_BORN_FWD, K_BORN_GRAD, K_BORN2_FWD, K_HESS };
inline bool kind_born(Kind k) { return k == K_BORN_FWD || k == K_BORN_GRAD || k == K_BORN2_FWD || k == K_HESS; }
inline bool kind_grad(Kind k) { return k == K_GRAD || k == K_BORN_GRAD || k == K_HESS; }
inline bool kind_forward_only(Kind k) { return k == K_FWD || k == K_BORN_FWD || k == K_BORN2_FWD; }

}  // namespace

struct sfwi_ctx {
    sfwi_params p{};
    Geo g{};
    bool fold = true;       // the prepared batch's sources and receivers lie in the update domain and no two
                            // receivers share a cell: the workgroup marches add the source, record the traces
                            // and inject the adjoint source themselves
    bool wg_ok = true;      // a shot's state block lies below 4 GiB, which the workgroup marches address with
                            // 32-bit byte offsets
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    size_t ncell = 0, nmodel = 0, nrec = 0, nint = 0;
    float *wav = nullptr;
    float *mvp = nullptr, *mvs = nullptr, *mrho = nullptr, *mdvp = nullptr, *mdvs = nullptr;
    float *tmp_model = nullptr;   // 4 full-model scratch arrays (line-search models)
    float *Gvp = nullptr, *Gvs = nullptr, *Gil = nullptr, *Grho = nullptr;
    float *res_keep = nullptr;    // 2 record arrays: residual kept across two Born solves
    float *res_batch = nullptr;   // the same for each shot of a batched linear line search
    double *dot_batch = nullptr;  // its per-trace FP64 sums
    size_t res_batch_cap = 0, dot_batch_cap = 0;
    std::vector<Slot> slots;
    DevShot *d_shots = nullptr;
    std::vector<DevShot> h_shots;
    std::map<int, Window> windows;
    std::map<int, ObsSet> obs;    // resident observed traces by id
    double *d_out = nullptr;      // per-shot misfits [max_batch] + 2 scalars
    unsigned *d_vmax = nullptr;   // largest Vp of a window (random boundary)
    double *d_num = nullptr, *d_den = nullptr;
    bool have_model = false, have_wav = false, have_pert = false;
    int pert_region = SFWI_BORN_S2;   // where the Born source acts (sfwi_set_perturbation_region)
    bool defer_merge = false;          // run_batch leaves the per-shot images unmerged (solve_merged)
    float *part_buf = nullptr;         // models and wavelets of the parts of a merged solve
    size_t part_buf_cap = 0;
    // optimal-transport misfit: job table, FP64 workspace, per-job misfits, I/O staging
    OtJob *ot_jobs = nullptr;
    double *ot_ws = nullptr, *ot_mis = nullptr;
    int *ot_first = nullptr;
    float *ot_io = nullptr;
    int *ot_bad = nullptr;        // set by k_ot when a merge walk stops at its bound
    bool ot_armed = false;        // an OT launch since the flag was last read
    size_t ot_jobs_cap = 0, ot_ws_cap = 0, ot_mis_cap = 0, ot_first_cap = 0, ot_io_cap = 0;
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;   // captured solves by (kind, batch size)
    // boundary strips in host memory (cpu_mem): the device keeps `ring` steps
    // per slot (nt, or two chunks of `chunk` steps); completed chunks are
    // copied on `cstream`, ordered against the kernels by the events
    int ring = 0, chunk = 0;
    hipStream_t cstream = nullptr;
    hipEvent_t ev_ring[2]{}, ev_copy[2]{};
    std::vector<OtJob> ot_h;          // host copies kept alive for the async uploads
    std::vector<int> ot_first_h;
    // timing
    int timing_every = 0;
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_kid;
    size_t ev_used = 0;
    double kt_ms[SFWI_K_COUNT] = {};
    int kt_n[SFWI_K_COUNT] = {};
};

namespace {

int fail(sfwi_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf; else g_create_error = buf;
    return code;
}

// Index of the first value that is not finite, or -1.
long bad_value(const float *v, long n) {
    for (long i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return i;
    return -1;
}

// A model the propagator cannot use: some vp not finite or <= 0, vs not
// finite or < 0, or (when the file's density is used) rho not finite or <= 0.
// Writes the first such cell to msg.
bool bad_model(const sfwi_params &p, const float *vp, const float *vs, const float *rho, char *msg, size_t cap) {
    const long n = (long)p.allnx * p.allnz;
    const bool use_rho = rho && !(p.rho_const > 0);
    for (long q = 0; q < n; ++q) {
        const char *what = nullptr;
        float v = 0.0f;
        if (!(std::isfinite(vp[q]) && vp[q] > 0.0f)) { what = "vp"; v = vp[q]; }
        else if (!(std::isfinite(vs[q]) && vs[q] >= 0.0f)) { what = "vs"; v = vs[q]; }
        else if (use_rho && !(std::isfinite(rho[q]) && rho[q] > 0.0f)) { what = "rho"; v = rho[q]; }
        if (what) {
            snprintf(msg, cap, "%s = %g at model cell (ix %ld, iz %ld): vp must be finite and > 0, vs finite "
                     "and >= 0, rho finite and > 0", what, v, q / p.allnz, q % p.allnz);
            return true;
        }
    }
    return false;
}

#define HIPCK(c, expr)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return fail(c, SFWI_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Resolve the static workspace's address on `dev` (once per device; the
// caller has made `dev` current).
int scratch_init(sfwi_ctx *c, int dev) {
    if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(c, SFWI_EINVAL, "device %d beyond the static workspace table", dev);
    std::lock_guard<std::mutex> lk(g_scratch_m);
    if (g_scratch[dev]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(c, SFWI_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[dev] = (char *)p;
    return 0;
}

int dalloc(sfwi_ctx *c, void **p, size_t bytes) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess && bytes + (64u << 20) > fr)
        return fail(c, SFWI_ENOMEM, "device allocation of %zu MiB exceeds free HBM (%zu MiB free)",
                    bytes >> 20, fr >> 20);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(c, SFWI_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return 0;
}

// ------------------------------------------------------------- timing ---

int tbeg(sfwi_ctx *c, int kid, int it) {
    if (c->timing_every <= 0 || (it % c->timing_every) != 0) return -1;
    if (c->ev_used + 2 > c->ev.size()) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return -1;
            c->ev.push_back(e);
            c->ev_kid.push_back(0);
        }
    }
    const int idx = (int)c->ev_used;
    c->ev_used += 2;
    c->ev_kid[idx] = kid;
    hipEventRecord(c->ev[idx], c->stream);
    return idx;
}
void tend(sfwi_ctx *c, int idx) {
    if (idx >= 0) hipEventRecord(c->ev[idx + 1], c->stream);
}
void tcollect(sfwi_ctx *c) {
    for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]) == hipSuccess) {
            c->kt_ms[c->ev_kid[k]] += ms;
            c->kt_n[c->ev_kid[k]] += 1;
        }
    }
    c->ev_used = 0;
}

// ------------------------------------------------------------ windows ---

int build_window(sfwi_ctx *c, Window &w, int left, const float *vp, const float *vs, const float *rho,
                 const float *dvp, const float *dvs, bool interior = false) {
    const size_t n = c->ncell;
    if (!w.base) {
        int r = dalloc(c, (void **)&w.base, 12 * n * sizeof(float));
        if (r) return r;
    }
    float *a[12];
    for (int k = 0; k < 12; ++k) a[k] = w.base + k * n;
    const dim3 blk(256), grd((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_window, grd, blk, 0, c->stream, c->g, vp, left, a[0]);
    hipLaunchKernelGGL(k_window, grd, blk, 0, c->stream, c->g, vs, left, a[1]);
    hipLaunchKernelGGL(k_window, grd, blk, 0, c->stream, c->g, rho, left, a[2]);
    float lg = (float)std::log10((double)(1.0f / c->p.rr));
    if (c->p.rbc) {   // random Vp / Vs in the PML bands, no damping
        HIPCK(c, hipMemsetAsync(c->d_vmax, 0, sizeof(unsigned), c->stream));
        hipLaunchKernelGGL(k_vmax, grd, blk, 0, c->stream, (int)n, a[0], c->d_vmax);
        const dim3 gz((unsigned)((c->g.nxp * c->g.pml + 255) / 256)), gx((unsigned)((c->g.pml * c->g.nzp + 255) / 256));
        hipLaunchKernelGGL(k_rbc_band<0>, gz, blk, 0, c->stream, c->g, c->p.rbc_seed, c->d_vmax, a[0], a[1]);
        hipLaunchKernelGGL(k_rbc_band<1>, gx, blk, 0, c->stream, c->g, c->p.rbc_seed, c->d_vmax, a[0], a[1]);
        lg = 0.0f;
    }
    hipLaunchKernelGGL(k_medium, grd, blk, 0, c->stream, c->g, lg, c->p.rho_const, a[0], a[1], a[2], a[3],
                       a[4], a[5], a[6], a[7], a[8], a[9]);
    w.has_pert = dvp && dvs;
    if (w.has_pert && interior) {   // stage-1 Born source: interior cells only
        hipLaunchKernelGGL(k_window_interior, grd, blk, 0, c->stream, c->g, dvp, left, a[10]);
        hipLaunchKernelGGL(k_window_interior, grd, blk, 0, c->stream, c->g, dvs, left, a[11]);
    } else if (w.has_pert) {
        hipLaunchKernelGGL(k_window, grd, blk, 0, c->stream, c->g, dvp, left, a[10]);
        hipLaunchKernelGGL(k_window, grd, blk, 0, c->stream, c->g, dvs, left, a[11]);
    }
    HIPCK(c, hipGetLastError());
    w.m.vp = a[0]; w.m.vs = a[1]; w.m.rho = a[2]; w.m.lam = a[3]; w.m.mu = a[4];
    w.m.dampx = a[5]; w.m.dampz = a[6]; w.m.bu = a[7]; w.m.icp = a[8]; w.m.ics = a[9];
    w.m.dvp = a[10]; w.m.dvs = a[11];
    return 0;
}

void drop_windows(sfwi_ctx *c) {
    for (auto &kv : c->windows) hipFree(kv.second.base);
    c->windows.clear();
}

int model_window(sfwi_ctx *c, int left, bool born, const Medium **out) {
    auto it = c->windows.find(left);
    if (it != c->windows.end() && (!born || it->second.has_pert)) {
        *out = &it->second.m;
        return 0;
    }
    if (c->windows.size() >= 256 && it == c->windows.end()) drop_windows(c);
    Window &w = c->windows[left];
    int r = build_window(c, w, left, c->mvp, c->mvs, c->mrho, c->have_pert ? c->mdvp : nullptr,
                         c->have_pert ? c->mdvs : nullptr, c->pert_region == SFWI_BORN_S1);
    if (r) return r;
    *out = &w.m;
    return 0;
}

// -------------------------------------------------------------- slots ---

// Bytes before the snapshot values of a slot: the max-abs words (256-B
// aligned), then the partial maxima of the snapshot being stored.
size_t snap_words(const sfwi_ctx *c) {
    const size_t nsnap = (c->p.nt + c->g.snap_k - 1) / c->g.snap_k;
    return (nsnap * 2 * sizeof(unsigned) + 255) & ~(size_t)255;
}
size_t snap_head(const sfwi_ctx *c) { return snap_words(c) + 2 * (size_t)c->g.snap_np * sizeof(float); }

// Model-size fields of one batch slot's block: fw (15), fw2 (9), bk (15), the
// velocity derivatives (4), v(it+1) with the density image (2), bk2 (9) with
// snapshot gradients.
int slot_fields(const sfwi_ctx *c) { return 43 + 2 * c->g.grad_rho + (c->g.snap_k > 0 ? 9 : 0); }

int ensure_slots(sfwi_ctx *c, int n, bool born, bool strips, bool snaps, bool born2 = false) {
    const size_t cell = c->ncell * sizeof(float);
    while ((int)c->slots.size() < n) c->slots.emplace_back();
    for (int s = 0; s < n; ++s) {
        Slot &S = c->slots[s];
        if (!S.block) {
            int r = dalloc(c, (void **)&S.block, slot_fields(c) * cell);
            if (r) return r;
            float *f = (float *)S.block;
            fptr *fw[15] = {&S.fw.vx, &S.fw.vz, &S.fw.txx, &S.fw.tzz, &S.fw.txz, &S.fw.vx_x, &S.fw.vx_z,
                              &S.fw.vz_x, &S.fw.vz_z, &S.fw.txx_x, &S.fw.txx_z, &S.fw.tzz_x, &S.fw.tzz_z,
                              &S.fw.txz_x, &S.fw.txz_z};
            fptr *bk[15] = {&S.bk.vx, &S.bk.vz, &S.bk.txx, &S.bk.tzz, &S.bk.txz, &S.bk.vx_x, &S.bk.vx_z,
                              &S.bk.vz_x, &S.bk.vz_z, &S.bk.txx_x, &S.bk.txx_z, &S.bk.tzz_x, &S.bk.tzz_z,
                              &S.bk.txz_x, &S.bk.txz_z};
            // fw (15 fields) | fw2 (9: totals and velocity split parts) | bk (15) | derivatives (4) | v(it+1) (2)
            for (int k = 0; k < 15; ++k) *fw[k] = f + k * c->ncell;
            S.fw2 = S.fw;   // the stress split parts are shared
            fptr *f2[9] = {&S.fw2.vx, &S.fw2.vz, &S.fw2.txx, &S.fw2.tzz, &S.fw2.txz, &S.fw2.vx_x, &S.fw2.vx_z,
                           &S.fw2.vz_x, &S.fw2.vz_z};
            for (int k = 0; k < 9; ++k) *f2[k] = f + (15 + k) * c->ncell;
            for (int k = 0; k < 15; ++k) *bk[k] = f + (24 + k) * c->ncell;
            for (int k = 0; k < 4; ++k) S.dg[k] = f + (39 + k) * c->ncell;
            if (c->g.grad_rho)
                for (int k = 0; k < 2; ++k) S.vold[k] = f + (43 + k) * c->ncell;
            if (c->g.snap_k > 0) {   // the fused adjoint step's second buffer: totals and velocity split parts
                S.bk2 = S.bk;
                fptr *b2[9] = {&S.bk2.vx, &S.bk2.vz, &S.bk2.txx, &S.bk2.tzz, &S.bk2.txz, &S.bk2.vx_x, &S.bk2.vx_z,
                               &S.bk2.vz_x, &S.bk2.vz_z};
                for (int k = 0; k < 9; ++k) *b2[k] = f + (43 + 2 * c->g.grad_rho + k) * c->ncell;
            }
            // records, receiver tables, accumulators, per-trace sums
            const size_t rbytes = 6 * c->nrec * sizeof(float) + (3 * (size_t)c->p.ntr + 1) * sizeof(int) +
                                  4 * c->nint * sizeof(float) + (size_t)c->p.ntr * sizeof(double) +
                                  (size_t)c->g.nxp * sizeof(int) + 64;
            r = dalloc(c, (void **)&S.rec, rbytes);
            if (r) return r;
            char *q = S.rec;
            S.trace_sum = (double *)q; q += (size_t)c->p.ntr * sizeof(double);
            float *rr = (float *)q;
            S.rz = rr; S.rx = rr + c->nrec; S.r2z = rr + 2 * c->nrec; S.r2x = rr + 3 * c->nrec;
            S.oz = rr + 4 * c->nrec; S.ox = rr + 5 * c->nrec;
            float *acc = rr + 6 * c->nrec;
            S.gvp = acc; S.gvs = acc + c->nint; S.gil = acc + 2 * c->nint; S.grho = acc + 3 * c->nint;
            S.igx = (int *)(acc + 4 * c->nint);
            S.inj = S.igx + c->p.ntr;
            S.inj_off = S.inj + c->p.ntr;
            S.rtr = S.inj_off + c->p.ntr + 1;
        }
        if (born && !S.born) {
            int r = dalloc(c, (void **)&S.born, 15 * cell);
            if (r) return r;
            float *f = (float *)S.born;
            fptr *bn[15] = {&S.bn.vx, &S.bn.vz, &S.bn.txx, &S.bn.tzz, &S.bn.txz, &S.bn.vx_x, &S.bn.vx_z,
                              &S.bn.vz_x, &S.bn.vz_z, &S.bn.txx_x, &S.bn.txx_z, &S.bn.tzz_x, &S.bn.tzz_z,
                              &S.bn.txz_x, &S.bn.txz_z};
            for (int k = 0; k < 15; ++k) *bn[k] = f + k * c->ncell;
        }
        if (born2 && !S.born2) {
            int r = dalloc(c, (void **)&S.born2, 15 * cell);
            if (r) return r;
            float *f = (float *)S.born2;
            fptr *bn[15] = {&S.bn2.vx, &S.bn2.vz, &S.bn2.txx, &S.bn2.tzz, &S.bn2.txz, &S.bn2.vx_x, &S.bn2.vx_z,
                            &S.bn2.vz_x, &S.bn2.vz_z, &S.bn2.txx_x, &S.bn2.txx_z, &S.bn2.tzz_x, &S.bn2.tzz_z,
                            &S.bn2.txz_x, &S.bn2.txz_z};
            for (int k = 0; k < 15; ++k) *bn[k] = f + k * c->ncell;
        }
        if (snaps && !S.snap) {
            const size_t nsnap = (c->p.nt + c->g.snap_k - 1) / c->g.snap_k;
            const size_t vals = nsnap * 2 * (size_t)c->g.snap_n * (c->g.store ? 2 : 4);
            int r = dalloc(c, (void **)&S.snap, snap_head(c) + vals);
            if (r) return r;
            r = dalloc(c, (void **)&S.ainj, 2 * (size_t)c->p.nt * c->g.nxp * sizeof(float));
            if (r) return r;
        }
        if (strips && !S.strip) {
            const size_t nv = (size_t)c->ring * 2 * c->g.sf;
            const size_t bytes = c->g.store ? nv * 2 + (size_t)c->ring * 2 * sizeof(float) + 2 * (size_t)c->g.sf * sizeof(float) + 256
                                            : nv * sizeof(float);
            int r = dalloc(c, (void **)&S.strip, bytes);
            if (r) return r;
            if (c->ring < c->p.nt) {
                const size_t hv = (size_t)c->p.nt * 2 * c->g.sf;
                const size_t hb = c->g.store ? hv * 2 + (size_t)c->p.nt * 2 * sizeof(float) : hv * sizeof(float);
                if (hipHostMalloc((void **)&S.hstrip, hb, hipHostMallocDefault) != hipSuccess) {
                    S.hstrip = nullptr;
                    return fail(c, SFWI_ENOMEM, "pinned host allocation of %zu MiB for the strips failed", hb >> 20);
                }
            }
        }
    }
    return 0;
}

int check_shot(sfwi_ctx *c, const sfwi_shot &s, int k) {
    const Geo &g = c->g;
    if (s.ntr < 1 || s.ntr > c->p.ntr) return fail(c, SFWI_EINVAL, "shot %d: ntr=%d outside [1,%d]", k, s.ntr, c->p.ntr);
    if (!s.gc) return fail(c, SFWI_EINVAL, "shot %d: gc is NULL", k);
    if (s.left < 0 || s.left + g.nx > c->p.allnx)
        return fail(c, SFWI_EINVAL, "shot %d: window [%d,%d) outside the model (allnx=%d)", k, s.left,
                    s.left + g.nx, c->p.allnx);
    if (s.isx < 0 || s.isx >= g.nxp || s.isz < 0 || s.isz >= g.nzp)
        return fail(c, SFWI_EINVAL, "shot %d: source (%d,%d) outside the padded grid", k, s.isx, s.isz);
    if (s.igz < 0 || s.igz >= g.nzp) return fail(c, SFWI_EINVAL, "shot %d: receiver row %d outside", k, s.igz);
    return 0;
}

// Receiver columns (sfwi-stage1/FD2DGPU.cu:3162) and the de-duplicated list
// of injected traces (last trace per cell, the sequential order's winner).
int receiver_tables(sfwi_ctx *c, const sfwi_shot &s, int k, Slot &S) {
    S.h_igx.resize(s.ntr);
    for (int tr = 0; tr < s.ntr; ++tr) {
        const int x = c->g.pml + (int)((float)(s.gc[tr] - s.ngx_left) / c->p.dx);
        if (x < 0 || x >= c->g.nxp)
            return fail(c, SFWI_EINVAL, "shot %d trace %d: receiver column %d outside the padded grid", k, tr, x);
        S.h_igx[tr] = x;
    }
    // traces grouped by receiver cell (cells in order of their first trace,
    // traces in order within a cell): the overwrite injection takes the last
    // of each group, the additive one sums the group in trace order
    std::vector<int> group(c->g.nxp, -1), count;
    for (int tr = 0; tr < s.ntr; ++tr) {
        int &gi = group[S.h_igx[tr]];
        if (gi < 0) { gi = (int)count.size(); count.push_back(0); }
        ++count[gi];
    }
    S.h_off.assign(count.size() + 1, 0);
    for (size_t k = 0; k < count.size(); ++k) S.h_off[k + 1] = S.h_off[k] + count[k];
    std::vector<int> fill(S.h_off.begin(), S.h_off.end() - 1);
    S.h_inj.assign(s.ntr, 0);
    for (int tr = 0; tr < s.ntr; ++tr) S.h_inj[fill[group[S.h_igx[tr]]]++] = tr;
    // the trace of each column of the receiver row: -1 none, -2 several
    S.h_rtr.assign(c->g.nxp, -1);
    for (int tr = 0; tr < s.ntr; ++tr) {
        int &v = S.h_rtr[S.h_igx[tr]];
        v = v == -1 ? tr : -2;
    }
    return 0;
}

// Prepare n shots in slots 0..n-1: medium windows, tables, records, zeroed state.
// shot_m / shot_wav (n each, a merged solve): every shot's own medium window
// and wavelet.
int prepare(sfwi_ctx *c, Kind kind, int n, const sfwi_shot *shots, const Medium *override_m,
            const std::map<int, Window> *override_w = nullptr, const Medium *const *shot_m = nullptr,
            const float *const *shot_wav = nullptr) {
    const bool born = kind_born(kind), born2 = kind == K_BORN2_FWD;
    const bool grad = kind_grad(kind);
    const bool obs = kind == K_GRAD || kind == K_BORN_GRAD || kind == K_FWD_MISFIT;
    if (!c->have_model && !shot_m) return fail(c, SFWI_ESTATE, "model not set (sfwi_set_model)");
    if (!c->have_wav && !shot_wav) return fail(c, SFWI_ESTATE, "wavelet not set (sfwi_set_wavelet)");
    if (born && !override_m && !shot_m && !c->have_pert)
        return fail(c, SFWI_ESTATE, "perturbation not set (sfwi_set_perturbation)");
    if (n < 1 || n > c->p.max_batch) return fail(c, SFWI_EINVAL, "batch of %d shots outside [1,%d]", n, c->p.max_batch);
    const bool snaps = grad && c->g.snap_k > 0;
    int r = ensure_slots(c, n, born, grad && !c->p.rbc && !snaps, snaps, born2);
    if (r) return r;
    c->h_shots.resize(n);
    c->fold = true;
    auto dom = [&](int x, int z) { return x >= NOP && x < c->g.nxp - NOP && z >= NOP && z < c->g.nzp - NOP; };
    for (int k = 0; k < n; ++k) {
        const sfwi_shot &s = shots[k];
        if ((r = check_shot(c, s, k))) return r;
        const ObsSet *os = nullptr;
        if (obs && s.obs_id > 0) {
            const auto f = c->obs.find(s.obs_id);
            if (f == c->obs.end()) return fail(c, SFWI_EINVAL, "shot %d: no observed traces stored under id %d", k, s.obs_id);
            if (f->second.ntr != s.ntr)
                return fail(c, SFWI_EINVAL, "shot %d: id %d holds %d traces, the shot has %d", k, s.obs_id,
                            f->second.ntr, s.ntr);
            os = &f->second;
        } else if (obs && (!s.obs_z || !s.obs_x) && (s.obs_z || s.obs_x || c->slots[k].obs_ntr != s.ntr))
            return fail(c, SFWI_EINVAL, "shot %d: observed traces missing (and no resident records of %d traces "
                        "in batch slot %d)", k, s.ntr, k);
        Slot &S = c->slots[k];
        if ((r = receiver_tables(c, s, k, S))) return r;
        const Medium *m = override_m ? override_m : shot_m ? shot_m[k] : nullptr;
        if (!m && override_w) {   // a medium per window position (batched line search)
            const auto ow = override_w->find(s.left);
            if (ow == override_w->end()) return fail(c, SFWI_EINVAL, "shot %d: no test-model window", k);
            m = &ow->second.m;
        }
        if (!m && (r = model_window(c, s.left, born, &m))) return r;
        DevShot d{};
        d.m = *m;
        d.fw = S.fw; d.fw2 = S.fw2; d.bk = S.bk; d.bn = S.bn; d.bn2 = S.bn2; d.bk2 = S.bk2;
        d.dgx = S.dg[0]; d.dgz = S.dg[1]; d.dgxz = S.dg[2]; d.dgzx = S.dg[3];
        d.rz = S.rz; d.rx = S.rx; d.r2z = S.r2z; d.r2x = S.r2x; d.oz = S.oz; d.ox = S.ox;
        if (os) {   // read in place from the resident set
            d.oz = os->base;
            d.ox = os->base + (size_t)os->ntr * c->p.nt;
        }
        d.igx = S.igx; d.inj = S.inj; d.inj_off = S.inj_off; d.rtr = S.rtr; d.strip = S.strip; d.ainj = S.ainj;
        if (snaps) {
            const size_t head = snap_head(c);
            d.snap_max = (unsigned *)S.snap;
            d.snap_part = (float *)(S.snap + snap_words(c));
            d.snap32 = (float *)(S.snap + head);
            d.snap16 = (unsigned short *)(S.snap + head);
        }
        if (c->g.store && S.strip) {
            const size_t nv = (size_t)c->p.nt * 2 * c->g.sf;
            char *q = (char *)S.strip;
            d.strip_scale = (float *)q;
            d.stage = d.strip_scale + (size_t)c->ring * 2;
            d.strip_h = (unsigned short *)(float *)(d.stage + 2 * (size_t)c->g.sf);
            (void)nv;
        }
        d.gvp = S.gvp; d.gvs = S.gvs; d.gil = S.gil; d.grho = S.grho; d.trace_sum = S.trace_sum;
        d.vold_x = S.vold[0]; d.vold_z = S.vold[1];
        d.isx = s.isx; d.isz = s.isz; d.igz = s.igz; d.ntr = s.ntr; d.ninj = (int)S.h_off.size() - 1;
        d.left = s.left;
        d.wav = shot_wav ? shot_wav[k] : nullptr;
        c->h_shots[k] = d;
        // the workgroup marches add the source, record the receivers and inject
        // the adjoint source themselves where these lie in their update domain
        // and every receiver has a cell of its own
        if (!dom(s.isx, s.isz)) c->fold = false;
        for (int x : S.h_igx)
            if (!dom(x, s.igz) || S.h_rtr[x] < 0) c->fold = false;
        const size_t cell = c->ncell * sizeof(float);
        HIPCK(c, hipMemsetAsync(S.block, 0, (grad ? 43 : 24) * cell, c->stream));
        if (grad && snaps) HIPCK(c, hipMemsetAsync((float *)S.bk2.vx, 0, 9 * cell, c->stream));
        if (born) HIPCK(c, hipMemsetAsync(S.born, 0, 15 * cell, c->stream));
        if (born2) HIPCK(c, hipMemsetAsync(S.born2, 0, 15 * cell, c->stream));
        HIPCK(c, hipMemcpyAsync(S.igx, S.h_igx.data(), s.ntr * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCK(c, hipMemcpyAsync(S.inj, S.h_inj.data(), S.h_inj.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCK(c, hipMemcpyAsync(S.inj_off, S.h_off.data(), S.h_off.size() * sizeof(int), hipMemcpyHostToDevice,
                                c->stream));
        HIPCK(c, hipMemcpyAsync(S.rtr, S.h_rtr.data(), S.h_rtr.size() * sizeof(int), hipMemcpyHostToDevice,
                                c->stream));
        if (snaps) HIPCK(c, hipMemsetAsync(S.snap, 0, snap_words(c), c->stream));   // scales of FP32 snapshots: 0
        if (obs && !os && s.obs_z) {   // NULL: the records resident in this slot are reused
            const size_t rb = (size_t)s.ntr * c->p.nt * sizeof(float);
            HIPCK(c, hipMemcpyAsync(S.oz, s.obs_z, rb, hipMemcpyHostToDevice, c->stream));
            HIPCK(c, hipMemcpyAsync(S.ox, s.obs_x, rb, hipMemcpyHostToDevice, c->stream));
            S.obs_ntr = s.ntr;
        }
        if (grad) HIPCK(c, hipMemsetAsync(S.gvp, 0, 4 * c->nint * sizeof(float), c->stream));
    }
    HIPCK(c, hipMemcpyAsync(c->d_shots, c->h_shots.data(), n * sizeof(DevShot), hipMemcpyHostToDevice, c->stream));
    return 0;
}

// ------------------------------------------------------------- solves ---

// Column-march geometry of a segment width (0: 8 columns).
void set_march_geo(Geo &g, int cols) {
    g.cm_xs = cols > 0 ? cols : 8;
    g.cm_nbands = (g.nzd + CM_ROWS - 1) / CM_ROWS;
    g.cm_waves = ((g.nxp - 2 * NOP + g.cm_xs - 1) / g.cm_xs) * g.cm_nbands;
    g.cm_wg = (g.cm_waves + BLK / 64 - 1) / (BLK / 64);
    g.cm_per_xcd = (g.cm_wg + 7) / 8;
}

void drop_graphs(sfwi_ctx *c) {
    for (auto &kv : c->graphs) hipGraphExecDestroy(kv.second);
    c->graphs.clear();
}

// ----------------------------------------------------- optimal transport ---

constexpr int OT_MAX_BLOCKS = 1024;   // resident OT workgroups (4 per CU); each owns a workspace

// Grow-only device buffer; waits for the stream before freeing the old one.
int grow(sfwi_ctx *c, void **p, size_t *cap, size_t need) {
    if (need <= *cap) return 0;
    if (*p) {
        HIPCK(c, hipStreamSynchronize(c->stream));
        HIPCK(c, hipFree(*p));
        *p = nullptr;
        *cap = 0;
    }
    int r = dalloc(c, p, need);
    if (r) return r;
    *cap = need;
    return 0;
}

// Enqueue k_ot over the jobs in c->ot_h (longest job n_max samples).
int ot_launch(sfwi_ctx *c, size_t n_max, double dt2) {
    const size_t nj = c->ot_h.size();
    const int blocks = (int)std::min<size_t>(nj, OT_MAX_BLOCKS);
    const size_t stride = 4 * (n_max + 1);
    int r;
    if ((r = grow(c, (void **)&c->ot_jobs, &c->ot_jobs_cap, nj * sizeof(OtJob)))) return r;
    if ((r = grow(c, (void **)&c->ot_ws, &c->ot_ws_cap, blocks * stride * sizeof(double)))) return r;
    HIPCK(c, hipMemcpyAsync(c->ot_jobs, c->ot_h.data(), nj * sizeof(OtJob), hipMemcpyHostToDevice, c->stream));
    if (!c->ot_armed) HIPCK(c, hipMemsetAsync(c->ot_bad, 0, sizeof(int), c->stream));
    c->ot_armed = true;
    hipLaunchKernelGGL(k_ot, dim3(blocks), dim3(OT_THREADS), 0, c->stream, c->ot_jobs, (int)nj, c->ot_ws, stride,
                       dt2, c->ot_bad);
    HIPCK(c, hipGetLastError());
    return 0;
}

// OT misfit of a prepared batch in place of the L2 residual: the synthetic
// traces in rz / rx (stage 2: born + background) become the adjoint source,
// per-shot misfits go to d_out.
int ot_batch(sfwi_ctx *c, int n, bool born) {
    const Geo &g = c->g;
    const int var = c->p.misfit;
    const bool gather = var == SFWI_MISFIT_OT_GATHER;
    if (born) {
        const dim3 grid(256, (unsigned)n);
        hipLaunchKernelGGL(k_born_total, grid, dim3(BLK), 0, c->stream, g, c->d_shots);
    }
    c->ot_h.clear();
    c->ot_first_h.assign(1, 0);
    size_t nj = 0;
    for (int s = 0; s < n; ++s) nj += gather ? 2 : 2 * (size_t)c->h_shots[s].ntr;
    int r;
    if ((r = grow(c, (void **)&c->ot_mis, &c->ot_mis_cap, nj * sizeof(double)))) return r;
    if ((r = grow(c, (void **)&c->ot_first, &c->ot_first_cap, (n + 1) * sizeof(int)))) return r;
    size_t n_max = 0;
    for (int s = 0; s < n; ++s) {
        const DevShot &d = c->h_shots[s];
        float *syn[2] = {d.rx, d.rz};
        const float *obs[2] = {d.ox, d.oz};
        for (int comp = 0; comp < 2; ++comp) {
            if (gather) {
                const size_t len = (size_t)d.ntr * g.nt;
                c->ot_h.push_back({syn[comp], obs[comp], syn[comp], c->ot_mis + c->ot_h.size(), (int)len, 0});
                n_max = std::max(n_max, len);
            } else {
                for (int tr = 0; tr < d.ntr; ++tr) {
                    const size_t o = (size_t)tr * g.nt;
                    c->ot_h.push_back({syn[comp] + o, obs[comp] + o, syn[comp] + o, c->ot_mis + c->ot_h.size(), g.nt,
                                       var == SFWI_MISFIT_OT_POWER});
                }
                n_max = std::max(n_max, (size_t)g.nt);
            }
        }
        c->ot_first_h.push_back((int)c->ot_h.size());
    }
    HIPCK(c, hipMemcpyAsync(c->ot_first, c->ot_first_h.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice,
                            c->stream));
    if ((r = ot_launch(c, n_max, (double)c->p.dt * (double)c->p.dt))) return r;
    hipLaunchKernelGGL(k_ot_shot_sum, dim3((n + 63) / 64), dim3(64), 0, c->stream, c->ot_mis, c->ot_first,
                       c->d_out, n);
    HIPCK(c, hipGetLastError());
    return 0;
}

// Copy steps [k chunk, min((k+1) chunk, nt)) of every slot's boundary strips
// between ring half k % 2 on the device and the slot's pinned host buffer, on
// the copy stream (cpu_mem).
int strip_chunk_copy(sfwi_ctx *c, int n, int k, bool to_host) {
    const size_t per = 2 * (size_t)c->g.sf;   // values of one step
    const size_t t0 = (size_t)k * c->chunk;
    const size_t steps = std::min<size_t>(c->chunk, (size_t)c->g.nt - t0);   // g.nt: the solve's record length
    const size_t r0 = (size_t)(k & 1) * c->chunk;
    const hipMemcpyKind kd = to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
    auto cp = [&](void *dev, void *host, size_t bytes) {
        return to_host ? hipMemcpyAsync(host, dev, bytes, kd, c->cstream) : hipMemcpyAsync(dev, host, bytes, kd, c->cstream);
    };
    for (int s = 0; s < n; ++s) {
        Slot &S = c->slots[s];
        if (!c->g.store) {
            HIPCK(c, cp(S.strip + r0 * per, (float *)S.hstrip + t0 * per, steps * per * sizeof(float)));
        } else {
            float *dscale = S.strip;
            unsigned short *dh = (unsigned short *)(dscale + (size_t)c->ring * 2 + per);
            unsigned short *hh = (unsigned short *)S.hstrip;
            float *hscale = (float *)(hh + (size_t)c->p.nt * per);
            HIPCK(c, cp(dh + r0 * per, hh + t0 * per, steps * per * sizeof(unsigned short)));
            HIPCK(c, cp(dscale + r0 * 2, hscale + t0 * 2, steps * 2 * sizeof(float)));
        }
    }
    return 0;
}

// Merge the per-shot images of batch slots [off, off + n) into the context's
// accumulators, as the end of a gradient solve of `kind` does: FAST mode
// applies the imaging coefficients, the snapshot gradient weights its sums
// by snap_k dt and snap_k, the Hessian product merges the Vp / Vs images only.
void merge_range(sfwi_ctx *c, Kind kind, int off, int n) {
    const Geo &g = c->g;
    const DevShot *D = c->d_shots + off;
    const dim3 mg((unsigned)((c->nmodel + 255) / 256)), blk(BLK);
    const bool fast = c->p.mode != SFWI_MODE_EXACT, snap = g.snap_k > 0, hess = kind == K_HESS;
    const float K = (float)g.snap_k;
    const float wv = snap ? K * g.dt : 1.0f, wi = snap ? K : 1.0f;
    hipStream_t st = c->stream;
    float *a = c->Gvp, *b = c->Gvs, *il = c->Gil, *r = c->Grho;
    const int ax = c->p.allnx;
    if (fast && hess) hipLaunchKernelGGL((k_merge<true, false>), mg, blk, 0, st, g, D, n, ax, a, b, il, r, wv, wi);
    else if (fast) hipLaunchKernelGGL((k_merge<true>), mg, blk, 0, st, g, D, n, ax, a, b, il, r, wv, wi);
    else if (hess) hipLaunchKernelGGL((k_merge<false, false>), mg, blk, 0, st, g, D, n, ax, a, b, il, r, wv, wi);
    else hipLaunchKernelGGL((k_merge<false>), mg, blk, 0, st, g, D, n, ax, a, b, il, r, wv, wi);
}

// PD: operand prefetch depth of the column-march kernels; HPD: the same for
// the two Born forward kernels, which carry two fields and twice the registers
// (at depth 4 they fit 2 waves per SIMD and run 6-11 % slower than at depth 2;
// the imaging reconstruction kernels gain 1 % at depth 2, which the next
// kernel loses again, and the others are 1-4 % faster at depth 4).  FPD: the
// wave-level fused step's; WPD: the workgroup forward step's, APD: the
// workgroup adjoint step's.  At depth 1 the two keep 4 waves per SIMD (123-125
// VGPRs); at depth 2 the forward needs 151 and runs at 3 waves, 3-5 % slower
// (377-386 us against 365 on the 1200 x 600 wave), and the adjoint, which
// keeps the old stresses of its stress column in an LDS ring, 169 (2 waves,
// 398 us against 357; without the ring it needed 150 and gained 1 %).
template <bool E, int PD, int HPD, int FPD = 2, int WPD = FPD, int APD = WPD>
int run_batch(sfwi_ctx *c, Kind kind, int n) {
    const Geo &g = c->g;
    const int nt = g.nt;
    const bool born = kind_born(kind), born2 = kind == K_BORN2_FWD, hess = kind == K_HESS;
    const bool grad = kind_grad(kind);
    const dim3 blk(BLK);
    const dim3 cells((unsigned)(8 * g.per_xcd), (unsigned)n);
    const dim3 src1((unsigned)((n + 63) / 64));
    const dim3 trg((unsigned)((c->p.ntr + 63) / 64), (unsigned)n);
    hipStream_t st = c->stream;
    const DevShot *D = c->d_shots;

    // column-march launches: block order per sfwi_params.march_order
    Geo gm = g;
    dim3 tiles((unsigned)(8 * g.cm_per_xcd), (unsigned)n);
    const int order = c->p.march_order ? c->p.march_order : 3;
    if (order >= 2) {
        gm.cm_order = order - 1;
        gm.cm_nshots = n;
        gm.cm_per_xcd = (int)(((long)g.cm_wg * n + 7) / 8);
        tiles = dim3((unsigned)(8 * gm.cm_per_xcd), 1);
    }
    // per-cell kernels on request (march_cols < 0), or automatically (march_cols 0)
    // when the batch has too few wave tiles to fill the GPU: measured crossover
    // near 7 march waves per SIMD (1024 SIMDs)
    // (double scattering has per-cell kernels only)
    const bool percell = E || born2 || c->p.march_cols < 0 || (c->p.march_cols == 0 && (long)n * g.cm_waves < 7168);
    // Strips in host memory (cpu_mem): step it lives at ring position it % ring
    // (two chunks).  Forward: a chunk is copied to the host once written (the
    // last two stay on the device), and a ring half is rewritten only after its
    // previous chunk has left.  Reverse: the last two chunks are resident;
    // once a chunk is consumed its half takes the chunk two below it, copied
    // while the chunk in between is computed.
    // With the random boundary (rbc) nothing is saved: the reverse pass
    // reconstructs by time reversal alone (sfwi-stage1/FD2DGPU.cu:4543, 4602).
    const bool snap = grad && g.snap_k > 0;   // velocity snapshots instead of strips
    const bool save = grad && !c->p.rbc && !snap;
    const bool hs = save && c->ring < nt;
    const dim3 snapg((unsigned)((g.snap_n + BLK * SNAP_PER - 1) / (BLK * SNAP_PER)), (unsigned)n);   // k_snap_pack
    const int ch = c->chunk, klast = (nt - 1) / ch;
    // Fused step march (velocity and stress of a step in one launch, the state
    // double-buffered) for FAST column-march forward solves of the background
    // field when fuse_step > 0 (sfwi_autotune sets it where it is faster).
    // fuse_step 0 (automatic): the workgroup marches (3: forward and, for a
    // snapshot gradient, adjoint) where their 240-row tiles take no more lanes
    // per column than the 56-row waves of the half-step kernels (1200 x 600
    // grid: 3 x 256 = 12 x 64, 5.5 % faster; 401 x 301: 2 x 256 > 7 x 64, 9 %
    // slower), the half steps elsewhere.  Every choice gives the same bits.
    const int fstep = c->p.fuse_step != 0 ? c->p.fuse_step
                      : (!percell && c->wg_ok && g.ws_ntiles * BLK <= g.cm_nbands * 64 ? 3 : 0);
    const bool fused = !percell && !born && fstep > 0;
    // fuse_step 2, 3: the workgroup form (z halo shared through LDS, 240-row tiles)
    const bool fwg = fused && fstep >= 2;
    Geo gs = g;
    gs.fs_per_xcd = (int)(((long)g.fs_wg * n + 7) / 8);
    gs.ws_per_xcd = (int)(((long)g.ws_wg * n + 7) / 8);
    const dim3 steps_grid((unsigned)(8 * (fwg ? gs.ws_per_xcd : gs.fs_per_xcd)));
    // the adjoint march's own segment width (step_cols, else 32: on the 1200 x
    // 600 wave 367 us against 377 at 16; the forward march is fastest at 16)
    Geo ga = gs;
    ga.ws_xs = c->p.step_cols > 0 ? c->p.step_cols : 32;
    ga.ws_wg = ((g.nxp - 2 * NOP + ga.ws_xs - 1) / ga.ws_xs) * g.ws_ntiles;
    ga.ws_per_xcd = (int)(((long)ga.ws_wg * n + 7) / 8);
    const dim3 wa_grid((unsigned)(8 * ga.ws_per_xcd));
    for (int it = 0; it < nt && fused; ++it) {
        const int sp = hs ? it % c->ring : it;
        const int par = it & 1;   // buffer read by this step; the step writes the other
        if (hs && it % ch == 0 && it / ch >= 2) HIPCK(c, hipStreamWaitEvent(st, c->ev_copy[(it / ch) & 1], 0));
        // the workgroup march adds the source and records the receivers itself
        // (when they lie in its update domain: else k_source / k_record do)
        const bool fold = fwg && c->fold;
        if (!fold) hipLaunchKernelGGL(k_source, src1, dim3(64), 0, st, g, D, c->wav, it, n, par);
        int t = tbeg(c, SFWI_K_FWD_STEP, it);
        // a 16-bit snapshot step: the workgroup march also forms the partial
        // maxima of the snapshot's max-abs scales (k_snap_max's pass)
        const int smax = fwg && snap && g.store && it % g.snap_k == 0;
        if (fwg) {
            const float *wsrc = fold ? c->wav : nullptr;
            if (save)
                hipLaunchKernelGGL((km_fwd_step_wg<true, WPD>), steps_grid, blk, 0, st, gs, D, n, sp, par, smax,
                                   wsrc, it);
            else
                hipLaunchKernelGGL((km_fwd_step_wg<false, WPD>), steps_grid, blk, 0, st, gs, D, n, sp, par, smax,
                                   wsrc, it);
        } else if (save) {
            hipLaunchKernelGGL((km_fwd_step<true, FPD>), steps_grid, blk, 0, st, gs, D, n, sp, par);
        } else {
            hipLaunchKernelGGL((km_fwd_step<false, FPD>), steps_grid, blk, 0, st, gs, D, n, sp, par);
        }
        tend(c, t);
        if (save && g.store) hipLaunchKernelGGL(k_strip_pack, dim3(2, (unsigned)n), dim3(256), 0, st, g, D, sp);
        if (snap && it % g.snap_k == 0) {   // keep the new velocity (snapshot it / snap_k)
            t = tbeg(c, SFWI_K_SNAP_STORE, it);
            if (g.store) {
                if (!smax) hipLaunchKernelGGL(k_snap_max, dim3(SNAP_PART, (unsigned)n), blk, 0, st, g, D, 1 - par);
                hipLaunchKernelGGL(k_snap_scale, dim3((unsigned)n), dim3(64), 0, st, g, D, it / g.snap_k,
                                   smax ? g.ws_wg : SNAP_PART);
            }
            hipLaunchKernelGGL(k_snap_pack, snapg, blk, 0, st, g, D, it / g.snap_k, 1 - par);
            tend(c, t);
        }
        if (!fold) hipLaunchKernelGGL(k_record<false>, trg, dim3(64), 0, st, g, D, it, 1 - par);
        if (hs && it % ch == ch - 1 && it / ch <= klast - 2) {   // chunk written: to the host
            const int k = it / ch;
            HIPCK(c, hipEventRecord(c->ev_ring[k & 1], st));
            HIPCK(c, hipStreamWaitEvent(c->cstream, c->ev_ring[k & 1], 0));
            if (int r = strip_chunk_copy(c, n, k, true)) return r;
            HIPCK(c, hipEventRecord(c->ev_copy[k & 1], c->cstream));
        }
    }
    if (fused && (nt & 1) && grad && !snap)   // the final state is in the second buffer: the reverse reads fw
        hipLaunchKernelGGL(k_fw_from_fw2, dim3(1024, (unsigned)n), blk, 0, st, g, D);
    for (int it = 0; it < nt && !fused; ++it) {
        const int sp = hs ? it % c->ring : it;
        if (hs && it % ch == 0 && it / ch >= 2) HIPCK(c, hipStreamWaitEvent(st, c->ev_copy[(it / ch) & 1], 0));
        hipLaunchKernelGGL(k_source, src1, dim3(64), 0, st, g, D, c->wav, it, n, 0);
        int t = tbeg(c, SFWI_K_FWD_VEL, it);
        if (percell) {
            if (born2) {
                hipLaunchKernelGGL((k_vel<E, false, 2>), cells, blk, 0, st, g, D, sp, 1);
            } else if (born) {
                if (save) hipLaunchKernelGGL((k_vel<E, true, true>), cells, blk, 0, st, g, D, sp, 1);
                else hipLaunchKernelGGL((k_vel<E, false, true>), cells, blk, 0, st, g, D, sp, 1);
            } else {
                if (save) hipLaunchKernelGGL((k_vel<E, true, false>), cells, blk, 0, st, g, D, sp, 1);
                else hipLaunchKernelGGL((k_vel<E, false, false>), cells, blk, 0, st, g, D, sp, 1);
            }
        } else {
            if (born) {
                if (save) hipLaunchKernelGGL((km_vel<true, true, HPD>), tiles, blk, 0, st, gm, D, sp, 1);
                else hipLaunchKernelGGL((km_vel<false, true, HPD>), tiles, blk, 0, st, gm, D, sp, 1);
            } else {
                if (save) hipLaunchKernelGGL((km_vel<true, false, PD>), tiles, blk, 0, st, gm, D, sp, 1);
                else hipLaunchKernelGGL((km_vel<false, false, PD>), tiles, blk, 0, st, gm, D, sp, 1);
            }
        }
        tend(c, t);
        if (save && g.store) hipLaunchKernelGGL(k_strip_pack, dim3(2, (unsigned)n), dim3(256), 0, st, g, D, sp);
        if (snap && it % g.snap_k == 0) {   // keep the new velocity (snapshot it / snap_k)
            t = tbeg(c, SFWI_K_SNAP_STORE, it);
            if (g.store) {   // max-abs scales: per-block partials, then one wave per shot
                hipLaunchKernelGGL(k_snap_max, dim3(SNAP_PART, (unsigned)n), blk, 0, st, g, D, 0);
                hipLaunchKernelGGL(k_snap_scale, dim3((unsigned)n), dim3(64), 0, st, g, D, it / g.snap_k, SNAP_PART);
            }
            hipLaunchKernelGGL(k_snap_pack, snapg, blk, 0, st, g, D, it / g.snap_k, 0);
            tend(c, t);
        }
        t = tbeg(c, SFWI_K_FWD_STR, it);
        if (percell) {
            if (born2) hipLaunchKernelGGL((k_str<E, false, 2>), cells, blk, 0, st, g, D, 1);
            else if (born) hipLaunchKernelGGL((k_str<E, false, true>), cells, blk, 0, st, g, D, 1);
            else hipLaunchKernelGGL((k_str<E, false, false>), cells, blk, 0, st, g, D, 1);
        } else {
            if (born) hipLaunchKernelGGL((km_str<false, true, HPD>), tiles, blk, 0, st, gm, D, 1);
            else hipLaunchKernelGGL((km_str<false, false, PD>), tiles, blk, 0, st, gm, D, 1);
        }
        tend(c, t);
        if (born2) hipLaunchKernelGGL(k_record<2>, trg, dim3(64), 0, st, g, D, it, 0);
        else if (born) hipLaunchKernelGGL(k_record<1>, trg, dim3(64), 0, st, g, D, it, 0);
        else hipLaunchKernelGGL(k_record<0>, trg, dim3(64), 0, st, g, D, it, 0);
        if (hs && it % ch == ch - 1 && it / ch <= klast - 2) {   // chunk written: to the host
            const int k = it / ch;
            HIPCK(c, hipEventRecord(c->ev_ring[k & 1], st));
            HIPCK(c, hipStreamWaitEvent(c->cstream, c->ev_ring[k & 1], 0));
            if (int r = strip_chunk_copy(c, n, k, true)) return r;
            HIPCK(c, hipEventRecord(c->ev_copy[k & 1], c->cstream));
        }
    }
    HIPCK(c, hipGetLastError());
    if (kind_forward_only(kind)) return 0;

    // The Hessian-vector product injects the scattered traces as they are
    // recorded (rz, rx): no residual, no misfit.
    if (hess) {
    } else if (c->p.misfit == SFWI_MISFIT_L2) {
        const dim3 resg(256, (unsigned)n);
        if (born) hipLaunchKernelGGL(k_residual<true>, resg, blk, 0, st, g, D);
        else hipLaunchKernelGGL(k_residual<false>, resg, blk, 0, st, g, D);
        hipLaunchKernelGGL(k_trace_l2, trg, dim3(64), 0, st, g, D);
        hipLaunchKernelGGL(k_shot_l2, src1, dim3(64), 0, st, D, c->d_out, n);
        HIPCK(c, hipGetLastError());
    } else {
        int r = ot_batch(c, n, born);
        if (r) return r;
    }
    if (!grad) return 0;

    const int adir = c->p.adjoint_dir;
    if (snap) {
        // Snapshot (consistent) gradient: per reverse step the adjoint source
        // fac r / rho is added at the receivers, then the adjoint velocity half
        // step, on snapshot steps the imaging from the kept velocity and the
        // adjoint stresses, then the adjoint stress half step, direction +1;
        // no reconstruction.  Images weighted by snap_k dt, illumination by
        // snap_k, at the merge.
        const float fac = (hess || c->p.misfit == SFWI_MISFIT_L2) ? 2.0f : 1.0f;
        const dim3 ig((unsigned)((c->nint + BLK - 1) / BLK), (unsigned)n);
        Geo gi = g;   // imaging march: whole shots per XCD
        gi.im_per_xcd = (int)(((long)g.im_wg * n + 7) / 8);
        const dim3 igm((unsigned)(8 * gi.im_per_xcd));
        // fuse_step 3: each reverse step is also one workgroup march (the
        // adjoint state double-buffered, bk / bk2 by the step's parity); the
        // imaging reads the adjoint stresses before it, which neither half step
        // of the pair changes before the imaging either
        const bool afused = !percell && fstep == 3;
        // the adjoint source as a column table, added by the march itself
        // (receivers in the update domain, one per cell); else k_inject_add
        const int ainj = afused && c->fold;
        if (ainj)
            hipLaunchKernelGGL(k_inject_table, dim3((unsigned)std::min<size_t>(4096, ((size_t)nt * 2 * g.nxp + 255) / 256),
                                                    (unsigned)n), dim3(256), 0, st, g, D, fac);
        for (int it = nt - 1; it >= 0 && afused; --it) {
            const int apar = (nt - 1 - it) & 1;   // buffer holding the adjoint state before this step
            if (!ainj) hipLaunchKernelGGL(k_inject_add, trg, dim3(64), 0, st, g, D, it, fac, apar);
            if (it % g.snap_k == 0) {
                int t = tbeg(c, SFWI_K_SNAP_IMAGE, it);
                const int k = it / g.snap_k;
                if (g.store == 0) hipLaunchKernelGGL(km_snap_image<0>, igm, blk, 0, st, gi, D, n, k, apar);
                else if (g.store == 1) hipLaunchKernelGGL(km_snap_image<1>, igm, blk, 0, st, gi, D, n, k, apar);
                else hipLaunchKernelGGL(km_snap_image<2>, igm, blk, 0, st, gi, D, n, k, apar);
                tend(c, t);
            }
            int t = tbeg(c, SFWI_K_ADJ_STEP, it);
            hipLaunchKernelGGL((km_adj_step_wg<APD>), wa_grid, blk, 0, st, ga, D, n, apar, 1, it, ainj);
            tend(c, t);
        }
        for (int it = nt - 1; it >= 0 && !afused; --it) {
            hipLaunchKernelGGL(k_inject_add, trg, dim3(64), 0, st, g, D, it, fac, 0);
            int t = tbeg(c, SFWI_K_ADJ_VEL_ONLY, it);
            if (percell) hipLaunchKernelGGL((k_adj_vel_image<false, false>), cells, blk, 0, st, g, D, 1);
            else hipLaunchKernelGGL((km_adj_vel_image<PD, false, false, 1>), tiles, blk, 0, st, gm, D, 1);
            tend(c, t);
            if (it % g.snap_k == 0) {
                t = tbeg(c, SFWI_K_SNAP_IMAGE, it);
                const int k = it / g.snap_k;
                if (percell) hipLaunchKernelGGL(k_snap_image, ig, blk, 0, st, g, D, k, 0);
                else if (g.store == 0) hipLaunchKernelGGL(km_snap_image<0>, igm, blk, 0, st, gi, D, n, k, 0);
                else if (g.store == 1) hipLaunchKernelGGL(km_snap_image<1>, igm, blk, 0, st, gi, D, n, k, 0);
                else hipLaunchKernelGGL(km_snap_image<2>, igm, blk, 0, st, gi, D, n, k, 0);
                tend(c, t);
            }
            t = tbeg(c, SFWI_K_ADJ_STR, it);
            if (percell) hipLaunchKernelGGL(k_adj_str<false>, cells, blk, 0, st, g, D, 1);
            else hipLaunchKernelGGL(km_adj_str<PD>, tiles, blk, 0, st, gm, D, 1);
            tend(c, t);
        }
        if (!c->defer_merge) merge_range(c, kind, 0, n);
        HIPCK(c, hipGetLastError());
        return 0;
    }
    const dim3 stripg((unsigned)((g.sf + 255) / 256), (unsigned)n);
    auto rev_begin = [&](int it) -> int {   // the step's strips restored (after their chunk has arrived)
        if (hs && it % ch == ch - 1 && it / ch <= klast - 2)
            HIPCK(c, hipStreamWaitEvent(st, c->ev_copy[(it / ch) & 1], 0));
        if (save) hipLaunchKernelGGL(k_strip_load, stripg, blk, 0, st, g, D, hs ? it % c->ring : it);
        return 0;
    };
    auto rev_end = [&](int it) -> int {     // chunk consumed: its ring half takes the chunk two below
        if (hs && it % ch == 0 && it / ch >= 2) {
            const int k = it / ch;
            HIPCK(c, hipEventRecord(c->ev_ring[k & 1], st));
            HIPCK(c, hipStreamWaitEvent(c->cstream, c->ev_ring[k & 1], 0));
            if (int r = strip_chunk_copy(c, n, k - 2, false)) return r;
            HIPCK(c, hipEventRecord(c->ev_copy[k & 1], c->cstream));
        }
        return 0;
    };
    // Column-march reverse step without the density image: the adjoint
    // stress-like step runs first, so the reconstruction's stress step can
    // image from the velocity derivatives in its registers and the reverse
    // velocity step can add the illumination; the adjoint velocity step is then
    // imaging-free.  The reconstruction and the adjoint fields are independent
    // within a step, so the order changes no value, and the images are the
    // same arithmetic on the same operands (identical results).
    const bool march_img = !percell && !g.grad_rho;
    for (int it = nt - 1; it >= 0 && march_img; --it) {
        if (int r = rev_begin(it)) return r;
        hipLaunchKernelGGL(k_inject, trg, dim3(64), 0, st, g, D, it);
        int t = tbeg(c, SFWI_K_ADJ_STR, it);
        hipLaunchKernelGGL(km_adj_str<PD>, tiles, blk, 0, st, gm, D, adir);
        tend(c, t);
        t = tbeg(c, SFWI_K_REC_STR_IMG, it);
        hipLaunchKernelGGL((km_str<false, false, PD, true>), tiles, blk, 0, st, gm, D, -1);
        tend(c, t);
        t = tbeg(c, SFWI_K_REC_VEL_ILL, it);
        hipLaunchKernelGGL((km_vel<false, false, PD, true>), tiles, blk, 0, st, gm, D, it, -1);
        tend(c, t);
        t = tbeg(c, SFWI_K_ADJ_VEL_ONLY, it);
        hipLaunchKernelGGL((km_adj_vel_image<PD, false, false, 1>), tiles, blk, 0, st, gm, D, adir);
        tend(c, t);
        if (int r = rev_end(it)) return r;
    }
    for (int it = nt - 1; it >= 0 && !march_img; --it) {
        if (int r = rev_begin(it)) return r;
        int t = tbeg(c, SFWI_K_REC_STR, it);
        if (percell) hipLaunchKernelGGL((k_str<E, true, false>), cells, blk, 0, st, g, D, -1);
        else hipLaunchKernelGGL((km_str<true, false, PD>), tiles, blk, 0, st, gm, D, -1);
        tend(c, t);
        t = tbeg(c, SFWI_K_REC_VEL, it);
        if (percell) hipLaunchKernelGGL((k_vel<E, false, false>), cells, blk, 0, st, g, D, it, -1);
        else hipLaunchKernelGGL((km_vel<false, false, PD>), tiles, blk, 0, st, gm, D, it, -1);
        tend(c, t);
        hipLaunchKernelGGL(k_inject, trg, dim3(64), 0, st, g, D, it);
        t = tbeg(c, SFWI_K_ADJ_STR, it);
        if (percell) hipLaunchKernelGGL(k_adj_str<E>, cells, blk, 0, st, g, D, adir);
        else hipLaunchKernelGGL(km_adj_str<PD>, tiles, blk, 0, st, gm, D, adir);
        tend(c, t);
        t = tbeg(c, SFWI_K_ADJ_VEL, it);
        if (percell) hipLaunchKernelGGL(k_adj_vel_image<E>, cells, blk, 0, st, g, D, adir);
        else if (g.grad_rho)
            hipLaunchKernelGGL((km_adj_vel_image<PD, true, true, img_depth(PD)>), tiles, blk, 0, st, gm, D, adir);
        else
            hipLaunchKernelGGL((km_adj_vel_image<PD, true, false, img_depth(PD)>), tiles, blk, 0, st, gm, D, adir);
        tend(c, t);
        if (int r = rev_end(it)) return r;
    }
    if (!c->defer_merge) merge_range(c, kind, 0, n);
    HIPCK(c, hipGetLastError());
    return 0;
}

int run_direct(sfwi_ctx *c, Kind kind, int n) {
    if (c->p.mode == SFWI_MODE_EXACT) return run_batch<true, 1, 1>(c, kind, n);
    switch (c->p.march_depth) {
    case 1: return run_batch<false, 1, 1, 1>(c, kind, n);
    case 2: return run_batch<false, 2, 2, 2>(c, kind, n);
    case 4: return run_batch<false, 4, 4, 4>(c, kind, n);
    default: return run_batch<false, 4, 2, 2, 1, 1>(c, kind, n);   // automatic
    }
}

// With use_graph the whole solve of a (kind, batch size) pair -- every time
// step of the forward and reverse loops -- is captured once from the
// context's stream into a hipGraph and replayed afterwards.  Every kernel
// argument is a per-context constant (the geometry, the device shot table,
// the wavelet and output buffers) or the step index; the per-shot state the
// kernels read (slot and medium pointers, geometry) lives in the device shot
// table that prepare() uploads before each replay, so one graph serves any
// shots of that batch size.  The capture holds kernels only (no memset or
// host copy).  Solves that need host work between kernels (the OT misfit),
// solves whose strips live in host memory (copies on a second stream) and
// timed solves run directly.
int run(sfwi_ctx *c, Kind kind, int n) {
    const bool host_work = c->p.misfit != SFWI_MISFIT_L2 && !kind_forward_only(kind) && kind != K_HESS;
    if (!c->p.use_graph || host_work || c->timing_every > 0 || c->ring < c->p.nt) return run_direct(c, kind, n);
    const std::pair<int, int> key((int)kind, n);
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
        HIPCK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int r = run_direct(c, kind, n);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (r) {
            if (graph) hipGraphDestroy(graph);
            return r;
        }
        if (e != hipSuccess) return fail(c, SFWI_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(c, SFWI_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
        it = c->graphs.emplace(key, exec).first;
    }
    HIPCK(c, hipGraphLaunch(it->second, c->stream));
    return 0;
}

// After the stream has drained: an OT merge walk that stopped at its bound
// (levels above the last cumulative level, which the clamping rules out)
// makes the solve fail instead of returning a truncated misfit.
int check_ot(sfwi_ctx *c) {
    if (!c->ot_armed) return 0;
    int bad = 0;
    c->ot_armed = false;
    HIPCK(c, hipMemcpy(&bad, c->ot_bad, sizeof(int), hipMemcpyDeviceToHost));
    if (bad) return fail(c, SFWI_EHIP, "OT misfit: a merge walk did not end within its bound (invalid CDF levels)");
    return 0;
}

int finish(sfwi_ctx *c) {
    HIPCK(c, hipStreamSynchronize(c->stream));
    tcollect(c);
    return check_ot(c);
}

// A shot's observed traces into its slot: from the resident set of its
// obs_id, else from the host arrays.
int load_obs(sfwi_ctx *c, const sfwi_shot &s, Slot &S) {
    const size_t rb = (size_t)s.ntr * c->p.nt * sizeof(float);
    if (s.obs_id > 0) {
        const auto f = c->obs.find(s.obs_id);
        if (f == c->obs.end() || f->second.ntr != s.ntr)
            return fail(c, SFWI_EINVAL, "no observed traces of %d traces stored under id %d", s.ntr, s.obs_id);
        const float *b = f->second.base;
        HIPCK(c, hipMemcpyAsync(S.oz, b, rb, hipMemcpyDeviceToDevice, c->stream));
        HIPCK(c, hipMemcpyAsync(S.ox, b + (size_t)s.ntr * c->p.nt, rb, hipMemcpyDeviceToDevice, c->stream));
    } else {
        if (!s.obs_z || !s.obs_x) return fail(c, SFWI_EINVAL, "observed traces missing");
        HIPCK(c, hipMemcpyAsync(S.oz, s.obs_z, rb, hipMemcpyHostToDevice, c->stream));
        HIPCK(c, hipMemcpyAsync(S.ox, s.obs_x, rb, hipMemcpyHostToDevice, c->stream));
    }
    S.obs_ntr = s.ntr;
    return 0;
}

// Batched solve with misfit output; shots processed in groups of max_batch.
int copy_records(sfwi_ctx *c, const float *src, float *dst, int ntr);

int solve_misfits(sfwi_ctx *c, Kind kind, int nshots, const sfwi_shot *shots, double *misfits,
                  float *res_z = nullptr, float *res_x = nullptr) {
    if (nshots < 0 || (nshots > 0 && !shots)) return fail(c, SFWI_EINVAL, "bad shot list");
    int bad = -1;   // first shot whose misfit is not finite
    double bad_m = 0.0;
    for (int b0 = 0; b0 < nshots; b0 += c->p.max_batch) {
        const int n = std::min(c->p.max_batch, nshots - b0);
        int r = prepare(c, kind, n, shots + b0, nullptr);
        if (r) return r;
        if ((r = run(c, kind, n))) return r;
        std::vector<double> m(n);
        HIPCK(c, hipMemcpyAsync(m.data(), c->d_out, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        for (int k = 0; k < n && res_z; ++k) {   // each shot's residual (the adjoint source)
            const size_t o = (size_t)(b0 + k) * c->nrec;
            if ((r = copy_records(c, c->slots[k].rz, res_z + o, shots[b0 + k].ntr)) ||
                (r = copy_records(c, c->slots[k].rx, res_x + o, shots[b0 + k].ntr)))
                return r;
        }
        if ((r = finish(c))) return r;
        if (misfits)
            for (int k = 0; k < n; ++k) misfits[b0 + k] += m[k];
        for (int k = 0; k < n && bad < 0 && kind != K_HESS; ++k)
            if (!std::isfinite(m[k])) { bad = b0 + k; bad_m = m[k]; }
    }
    if (bad >= 0)
        return fail(c, SFWI_ENUMERIC, "shot %d: misfit %g is not finite (NaN or Inf in its observed traces); "
                                      "the accumulated images hold it until the next flush", bad, bad_m);
    return 0;
}

int copy_records(sfwi_ctx *c, const float *src, float *dst, int ntr) {
    if (dst)
        HIPCK(c, hipMemcpyAsync(dst, src, (size_t)ntr * c->p.nt * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    return 0;
}

// One part of a merged solve: a request's shots with its own model,
// perturbation (Born kinds) and wavelet, and its outputs.
struct Part {
    const float *vp, *vs, *rho, *dvp, *dvs, *wav;
    int nshots;
    const sfwi_shot *shots;
    float *syn_z, *syn_x;                  // forward kinds: traces of shot i at i*ntr*nt
    double *misfits;                       // gradient kinds: [nshots]
    float *gvp, *gvs, *gil;                // gradient kinds: allnx*allnz, the part's shots summed
};

// Several requests of one shape and kind solved as one batch: every shot
// gets the medium window of its own request's model and its request's
// wavelet; each part's images are merged over its own shots only, into
// accumulators that start at zero, so a part's results equal the same
// request solved alone bit for bit.  The context's accumulators are zero on
// return (its model, wavelet and perturbation are not used or changed).
// bad (may be NULL) receives, per part, its first shot whose misfit is not
// finite, or -1.
int solve_merged(sfwi_ctx *c, Kind kind, const std::vector<Part> &parts, std::vector<int> *bad) {
    int n = 0;
    for (const Part &q : parts) n += q.nshots;
    if (n < 1 || n > c->p.max_batch) return fail(c, SFWI_EINVAL, "merged batch of %d shots outside [1,%d]", n, c->p.max_batch);
    const bool born = kind_born(kind), grad = kind_grad(kind);
    const size_t nm = c->nmodel, per = 5 * nm + (size_t)c->p.nt;
    int r = grow(c, (void **)&c->part_buf, &c->part_buf_cap, parts.size() * per * sizeof(float));
    if (r) return r;
    std::vector<std::map<int, Window>> tw(parts.size());
    auto done = [&](int rc) {
        hipStreamSynchronize(c->stream);
        for (auto &mw : tw)
            for (auto &kv : mw) hipFree(kv.second.base);
        c->defer_merge = false;
        return rc;
    };
    std::vector<sfwi_shot> shots;
    std::vector<const Medium *> sm;
    std::vector<const float *> sw;
    for (size_t p = 0; p < parts.size(); ++p) {
        const Part &q = parts[p];
        if (!q.vp || !q.vs || !q.wav || (!q.rho && !(c->p.rho_const > 0)) || (born && (!q.dvp || !q.dvs)))
            return done(fail(c, SFWI_EINVAL, "part %zu: model, wavelet or perturbation missing", p));
        float *b = c->part_buf + p * per;
        const size_t mb = nm * sizeof(float);
        HIPCK(c, hipMemcpyAsync(b, q.vp, mb, hipMemcpyHostToDevice, c->stream));
        HIPCK(c, hipMemcpyAsync(b + nm, q.vs, mb, hipMemcpyHostToDevice, c->stream));
        if (q.rho) HIPCK(c, hipMemcpyAsync(b + 2 * nm, q.rho, mb, hipMemcpyHostToDevice, c->stream));
        else HIPCK(c, hipMemsetAsync(b + 2 * nm, 0, mb, c->stream));
        if (born) {
            HIPCK(c, hipMemcpyAsync(b + 3 * nm, q.dvp, mb, hipMemcpyHostToDevice, c->stream));
            HIPCK(c, hipMemcpyAsync(b + 4 * nm, q.dvs, mb, hipMemcpyHostToDevice, c->stream));
        }
        HIPCK(c, hipMemcpyAsync(b + 5 * nm, q.wav, c->p.nt * sizeof(float), hipMemcpyHostToDevice, c->stream));
        for (int k = 0; k < q.nshots; ++k) {
            const int left = q.shots[k].left;
            if (left < 0 || left + c->g.nx > c->p.allnx)
                return done(fail(c, SFWI_EINVAL, "part %zu shot %d: window outside the model", p, k));
            Window &w = tw[p][left];
            if (!w.base && (r = build_window(c, w, left, b, b + nm, b + 2 * nm, born ? b + 3 * nm : nullptr,
                                             born ? b + 4 * nm : nullptr)))
                return done(r);
            shots.push_back(q.shots[k]);
            sm.push_back(&w.m);
            sw.push_back(b + 5 * nm);
        }
    }
    if ((r = prepare(c, kind, n, shots.data(), nullptr, nullptr, sm.data(), sw.data()))) return done(r);
    c->defer_merge = true;
    if ((r = run_direct(c, kind, n))) return done(r);
    c->defer_merge = false;
    std::vector<double> mis(n, 0.0);
    if (grad) HIPCK(c, hipMemcpyAsync(mis.data(), c->d_out, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    const size_t ib = c->nmodel * sizeof(float);
    int off = 0;
    for (const Part &q : parts) {
        if (grad) {   // this part's images alone
            HIPCK(c, hipMemsetAsync(c->Gvp, 0, 4 * ib, c->stream));
            merge_range(c, kind, off, q.nshots);
            HIPCK(c, hipGetLastError());
            if (q.gvp) HIPCK(c, hipMemcpyAsync(q.gvp, c->Gvp, ib, hipMemcpyDeviceToHost, c->stream));
            if (q.gvs) HIPCK(c, hipMemcpyAsync(q.gvs, c->Gvs, ib, hipMemcpyDeviceToHost, c->stream));
            if (q.gil) HIPCK(c, hipMemcpyAsync(q.gil, c->Gil, ib, hipMemcpyDeviceToHost, c->stream));
        } else {
            for (int k = 0; k < q.nshots; ++k) {
                const size_t o = (size_t)k * c->nrec;
                const Slot &S = c->slots[off + k];
                if ((r = copy_records(c, S.rz, q.syn_z ? q.syn_z + o : nullptr, q.shots[k].ntr)) ||
                    (r = copy_records(c, S.rx, q.syn_x ? q.syn_x + o : nullptr, q.shots[k].ntr)))
                    return done(r);
            }
        }
        off += q.nshots;
    }
    if (grad) HIPCK(c, hipMemsetAsync(c->Gvp, 0, 4 * ib, c->stream));
    if ((r = finish(c))) return done(r);
    off = 0;
    if (bad) bad->assign(parts.size(), -1);
    for (size_t p = 0; p < parts.size(); ++p) {
        const Part &q = parts[p];
        if (grad && q.misfits)
            for (int k = 0; k < q.nshots; ++k) q.misfits[k] = mis[off + k];
        for (int k = q.nshots - 1; k >= 0 && grad && bad && kind != K_HESS; --k)
            if (!std::isfinite(mis[off + k])) (*bad)[p] = k;
        off += q.nshots;
    }
    return done(0);
}

}  // namespace

// ================================================================ C ABI ===

extern "C" {

const char *sfwi_version(void) { return "sfwi_amd 0.1 (gfx950)"; }

const char *sfwi_last_error(const sfwi_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int sfwi_check_model(const sfwi_params *p, const float *vp, const float *vs, const float *rho) {
    if (!p || !vp || !vs) return fail(nullptr, SFWI_EINVAL, "NULL argument");
    if (p->allnx < 1 || p->allnz < 1) return fail(nullptr, SFWI_EINVAL, "model of %d x %d cells", p->allnx, p->allnz);
    if (!rho && !(p->rho_const > 0)) return fail(nullptr, SFWI_EINVAL, "rho is NULL and rho_const is not set");
    char msg[160];
    if (bad_model(*p, vp, vs, rho, msg, sizeof msg)) return fail(nullptr, SFWI_EINVAL, "%s", msg);
    return 0;
}

int sfwi_ctx_create(const sfwi_params *p, sfwi_ctx **out) {
    if (!p || !out) return fail(nullptr, SFWI_EINVAL, "NULL argument");
    *out = nullptr;
    if (p->nop != 4) return fail(nullptr, SFWI_EINVAL, "nop=%d: only the 8th-order stencil (nop=4) exists", p->nop);
    if (p->pml < 8) return fail(nullptr, SFWI_EINVAL, "pml=%d: must be >= 8 so the strips lie inside the grid", p->pml);
    if (p->nx < 8 || p->nz < 8 || p->nt < 1 || p->ntr < 1 || p->max_batch < 1)
        return fail(nullptr, SFWI_EINVAL, "bad sizes nx=%d nz=%d nt=%d ntr=%d max_batch=%d", p->nx, p->nz, p->nt,
                    p->ntr, p->max_batch);
    if (p->nz != p->allnz) return fail(nullptr, SFWI_EINVAL, "nz (%d) must equal allnz (%d)", p->nz, p->allnz);
    if (p->nx > p->allnx) return fail(nullptr, SFWI_EINVAL, "nx (%d) exceeds allnx (%d)", p->nx, p->allnx);
    if (!(p->dx > 0 && p->dz > 0 && p->dt > 0 && p->rr > 0))
        return fail(nullptr, SFWI_EINVAL, "dx, dz, dt and rr must be positive");
    if (p->adjoint_dir != 1 && p->adjoint_dir != -1) return fail(nullptr, SFWI_EINVAL, "adjoint_dir must be +1 or -1");
    if (p->mode != SFWI_MODE_FAST && p->mode != SFWI_MODE_EXACT) return fail(nullptr, SFWI_EINVAL, "unknown mode %d", p->mode);
    if (p->storage < SFWI_STORE_STRIPS_F32 || p->storage > SFWI_STORE_STRIPS_BF16)
        return fail(nullptr, SFWI_EINVAL, "unknown strip storage %d", p->storage);
    if (p->march_order < 0 || p->march_order > 3) return fail(nullptr, SFWI_EINVAL, "unknown march_order %d", p->march_order);
    if (p->misfit < SFWI_MISFIT_L2 || p->misfit > SFWI_MISFIT_OT_GATHER)
        return fail(nullptr, SFWI_EINVAL, "unknown misfit %d", p->misfit);
    if (p->cpu_mem != 0 && p->cpu_mem != 1) return fail(nullptr, SFWI_EINVAL, "cpu_mem must be 0 or 1");
    if (p->rbc != 0 && p->rbc != 1) return fail(nullptr, SFWI_EINVAL, "rbc must be 0 or 1");
    if (p->rbc && p->pml < 1) return fail(nullptr, SFWI_EINVAL, "rbc needs pml >= 1 (the random bands)");
    if (p->strip_chunk < 0) return fail(nullptr, SFWI_EINVAL, "strip_chunk must be >= 0");
    if (p->recon != SFWI_RECON_STRIPS && p->recon != SFWI_RECON_SNAPSHOTS)
        return fail(nullptr, SFWI_EINVAL, "unknown recon %d", p->recon);
    if (p->snap_every < 0) return fail(nullptr, SFWI_EINVAL, "snap_every must be >= 0");
    if (p->step_cols < 0) return fail(nullptr, SFWI_EINVAL, "step_cols must be >= 0");
    if (p->recon == SFWI_RECON_SNAPSHOTS) {
        if (p->mode != SFWI_MODE_FAST)
            return fail(nullptr, SFWI_EINVAL, "recon=snapshots runs in FAST mode only (EXACT restates the strips)");
        if (p->grad_rho) return fail(nullptr, SFWI_EINVAL, "recon=snapshots does not image the density");
        if (p->cpu_mem) return fail(nullptr, SFWI_EINVAL, "recon=snapshots keeps its snapshots in HBM (no cpu_mem)");
    }
    const long nxp = p->nx + 2L * p->pml, nzp = p->nz + 2L * p->pml;
    if (nxp * nzp >= (1L << 31) / 16) return fail(nullptr, SFWI_EINVAL, "grid too large for 32-bit cell indices");
    // a shot's state block (every model-size field of a batch slot)
    const size_t sblock = (size_t)(43 + 2 * p->grad_rho + (p->recon == SFWI_RECON_SNAPSHOTS ? 9 : 0)) *
                          (size_t)(nxp * nzp) * sizeof(float);
    if (p->fuse_step >= 2 && sblock > 0xFFFFFFFFull)
        return fail(nullptr, SFWI_EINVAL, "fuse_step %d: the workgroup marches need a shot's state block below "
                    "4 GiB (%zu MiB on this grid)", p->fuse_step, sblock >> 20);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(nullptr, SFWI_EHIP, "no HIP device");

    sfwi_ctx *c = new sfwi_ctx();
    c->p = *p;
    c->device = dev;
    Geo &g = c->g;
    g.nxp = (int)nxp; g.nzp = (int)nzp; g.pml = p->pml; g.nx = p->nx; g.nz = p->nz; g.nt = p->nt;
    g.allnz = p->allnz;
    g.nzd = g.nzp - 2 * NOP;
    g.ndom = (g.nxp - 2 * NOP) * g.nzd;
    const int nb = (g.ndom + BLK - 1) / BLK;
    g.per_xcd = (nb + 7) / 8;
    set_march_geo(g, p->march_cols);
    g.sf = 4 * NSTRIP * (p->nx + p->nz);
    g.store = p->storage;
    g.grad_rho = p->grad_rho ? 1 : 0;
    if (p->recon == SFWI_RECON_SNAPSHOTS) {
        g.snap_k = p->snap_every > 0 ? p->snap_every : 4;
        g.snap_rz = p->nz + 2 * NOP;
        g.snap_n = (p->nx + 2 * NOP) * g.snap_rz;
        g.im_nbands = (p->nz + CM_ROWS - 1) / CM_ROWS;
        g.im_waves = ((p->nx + 7) / 8) * g.im_nbands;
        g.im_wg = (g.im_waves + BLK / 64 - 1) / (BLK / 64);
    }
    g.fs_xs = p->step_cols > 0 ? p->step_cols : 16;
    g.fs_nbands = (g.nzd + FS_ROWS - 1) / FS_ROWS;
    g.fs_waves = ((g.nxp - 2 * NOP + g.fs_xs - 1) / g.fs_xs) * g.fs_nbands;
    g.fs_wg = (g.fs_waves + BLK / 64 - 1) / (BLK / 64);
    g.ws_xs = p->step_cols > 0 ? p->step_cols : 16;
    g.ws_ntiles = (g.nzd + WS_ROWS - 1) / WS_ROWS;
    g.ws_wg = ((g.nxp - 2 * NOP + g.ws_xs - 1) / g.ws_xs) * g.ws_ntiles;
    g.snap_np = std::max(SNAP_PART, g.ws_wg);   // k_snap_max's blocks or the fused step's workgroups
    g.inv_nzd = 1.0f / (float)g.nzd;
    g.dt = p->dt; g.dx = p->dx; g.dz = p->dz;
    g.dtdx = p->dt / p->dx; g.dtdz = p->dt / p->dz;
    g.idx = 1.0f / p->dx; g.idz = 1.0f / p->dz;
    c->ncell = (size_t)nxp * nzp;
    c->wg_ok = sblock <= 0xFFFFFFFFull;
    c->nmodel = (size_t)p->allnx * p->allnz;
    c->nrec = (size_t)p->ntr * p->nt;
    c->nint = (size_t)p->nx * p->nz;

    c->chunk = p->strip_chunk > 0 ? p->strip_chunk : 100;
    c->ring = (p->cpu_mem && 2L * c->chunk < p->nt) ? 2 * c->chunk : p->nt;

    auto bail = [&](int code) { g_create_error = c->err; sfwi_ctx_destroy(c); return code; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(c, SFWI_EHIP, "hipStreamCreate failed"));
    if (int rs = scratch_init(c, dev)) return bail(rs);   // this device is current
    if (c->ring < p->nt) {
        if (hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking) != hipSuccess)
            return bail(fail(c, SFWI_EHIP, "hipStreamCreate failed (strip copies)"));
        for (int k = 0; k < 2; ++k)
            if (hipEventCreateWithFlags(&c->ev_ring[k], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&c->ev_copy[k], hipEventDisableTiming) != hipSuccess)
                return bail(fail(c, SFWI_EHIP, "hipEventCreate failed (strip copies)"));
    }
    int r;
    if ((r = dalloc(c, (void **)&c->wav, p->nt * sizeof(float)))) return bail(r);
    if ((r = dalloc(c, (void **)&c->mvp, 5 * c->nmodel * sizeof(float)))) return bail(r);
    c->mvs = c->mvp + c->nmodel; c->mrho = c->mvp + 2 * c->nmodel;
    c->mdvp = c->mvp + 3 * c->nmodel; c->mdvs = c->mvp + 4 * c->nmodel;
    if ((r = dalloc(c, (void **)&c->tmp_model, 4 * c->nmodel * sizeof(float)))) return bail(r);
    if ((r = dalloc(c, (void **)&c->Gvp, 4 * c->nmodel * sizeof(float)))) return bail(r);
    c->Gvs = c->Gvp + c->nmodel; c->Gil = c->Gvp + 2 * c->nmodel; c->Grho = c->Gvp + 3 * c->nmodel;
    if ((r = dalloc(c, (void **)&c->res_keep, 2 * c->nrec * sizeof(float)))) return bail(r);
    if ((r = dalloc(c, (void **)&c->d_shots, p->max_batch * sizeof(DevShot)))) return bail(r);
    if ((r = dalloc(c, (void **)&c->d_out, (p->max_batch + 4) * sizeof(double)))) return bail(r);
    if ((r = dalloc(c, (void **)&c->d_num, 2 * (size_t)p->ntr * sizeof(double)))) return bail(r);
    if ((r = dalloc(c, (void **)&c->d_vmax, 256))) return bail(r);   // vmax word + OT flag
    c->ot_bad = (int *)((char *)c->d_vmax + 128);
    c->d_den = c->d_num + p->ntr;
    if (hipMemsetAsync(c->Gvp, 0, 4 * c->nmodel * sizeof(float), c->stream) != hipSuccess ||
        hipMemsetAsync(c->mvp, 0, 5 * c->nmodel * sizeof(float), c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return bail(fail(c, SFWI_EHIP, "initial clears failed"));
    *out = c;
    return 0;
}

void sfwi_ctx_destroy(sfwi_ctx *c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->cstream) hipStreamSynchronize(c->cstream);
    for (auto &kv : c->graphs) hipGraphExecDestroy(kv.second);
    for (Slot &S : c->slots) {
        hipFree(S.block);
        hipFree(S.born);
        hipFree(S.born2);
        hipFree(S.rec);
        hipFree(S.strip);
        hipFree(S.snap);
        hipFree(S.ainj);
        if (S.hstrip) hipHostFree(S.hstrip);
    }
    for (int k = 0; k < 2; ++k) {
        if (c->ev_ring[k]) hipEventDestroy(c->ev_ring[k]);
        if (c->ev_copy[k]) hipEventDestroy(c->ev_copy[k]);
    }
    if (c->cstream) hipStreamDestroy(c->cstream);
    drop_windows(c);
    for (auto &kv : c->obs) hipFree(kv.second.base);
    for (hipEvent_t e : c->ev) hipEventDestroy(e);
    hipFree(c->wav); hipFree(c->mvp); hipFree(c->tmp_model); hipFree(c->Gvp); hipFree(c->res_keep);
    hipFree(c->d_shots); hipFree(c->d_out); hipFree(c->d_num); hipFree(c->res_batch); hipFree(c->dot_batch); hipFree(c->d_vmax);
    hipFree(c->ot_jobs); hipFree(c->ot_ws); hipFree(c->ot_mis); hipFree(c->ot_first); hipFree(c->ot_io);
    hipFree(c->part_buf);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int sfwi_set_wavelet(sfwi_ctx *c, const float *w) {
    if (!c || !w) return fail(c, SFWI_EINVAL, "NULL argument");
    const long bw = bad_value(w, c->p.nt);
    if (bw >= 0) return fail(c, SFWI_EINVAL, "wavelet sample %ld is %g, not finite", bw, w[bw]);
    HIPCK(c, hipMemcpyAsync(c->wav, w, c->p.nt * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    c->have_wav = true;
    return 0;
}

int sfwi_set_model(sfwi_ctx *c, const float *vp, const float *vs, const float *rho) {
    if (!c || !vp || !vs) return fail(c, SFWI_EINVAL, "NULL model");
    if (!rho && !(c->p.rho_const > 0)) return fail(c, SFWI_EINVAL, "rho is NULL and rho_const is not set");
    char msg[160];
    if (bad_model(c->p, vp, vs, rho, msg, sizeof msg)) return fail(c, SFWI_EINVAL, "%s", msg);
    const size_t b = c->nmodel * sizeof(float);
    HIPCK(c, hipStreamSynchronize(c->stream));
    drop_windows(c);
    HIPCK(c, hipMemcpyAsync(c->mvp, vp, b, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(c->mvs, vs, b, hipMemcpyHostToDevice, c->stream));
    if (rho) HIPCK(c, hipMemcpyAsync(c->mrho, rho, b, hipMemcpyHostToDevice, c->stream));
    else HIPCK(c, hipMemsetAsync(c->mrho, 0, b, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    c->have_model = true;
    return 0;
}

int sfwi_set_perturbation_region(sfwi_ctx *c, const float *dvp, const float *dvs, int region) {
    if (!c || !dvp || !dvs) return fail(c, SFWI_EINVAL, "NULL perturbation");
    if (region != SFWI_BORN_S1 && region != SFWI_BORN_S2) return fail(c, SFWI_EINVAL, "unknown Born region %d", region);
    const long bp = bad_value(dvp, (long)c->nmodel), bs = bad_value(dvs, (long)c->nmodel);
    if (bp >= 0 || bs >= 0)
        return fail(c, SFWI_EINVAL, "perturbation %s is not finite at cell %ld", bp >= 0 ? "dvp" : "dvs",
                    bp >= 0 ? bp : bs);
    const size_t b = c->nmodel * sizeof(float);
    HIPCK(c, hipStreamSynchronize(c->stream));
    drop_windows(c);
    HIPCK(c, hipMemcpyAsync(c->mdvp, dvp, b, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(c->mdvs, dvs, b, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    c->have_pert = true;
    c->pert_region = region;
    return 0;
}

int sfwi_set_perturbation(sfwi_ctx *c, const float *dvp, const float *dvs) {
    return sfwi_set_perturbation_region(c, dvp, dvs, SFWI_BORN_S2);
}

int sfwi_batch_born_forward(sfwi_ctx *c, int nshots, const sfwi_shot *shots, float *b_z, float *b_x) {
    if (!c || nshots < 0 || (nshots > 0 && !shots)) return fail(c, SFWI_EINVAL, "bad shot list");
    for (int b0 = 0; b0 < nshots; b0 += c->p.max_batch) {
        const int n = std::min(c->p.max_batch, nshots - b0);
        int r = prepare(c, K_BORN_FWD, n, shots + b0, nullptr);
        if (r || (r = run(c, K_BORN_FWD, n))) return r;
        for (int k = 0; k < n; ++k) {
            const size_t o = (size_t)(b0 + k) * c->nrec;
            if ((r = copy_records(c, c->slots[k].rz, b_z ? b_z + o : nullptr, shots[b0 + k].ntr)) ||
                (r = copy_records(c, c->slots[k].rx, b_x ? b_x + o : nullptr, shots[b0 + k].ntr)))
                return r;
        }
        if ((r = finish(c))) return r;
    }
    return 0;
}

int sfwi_batch_born2_forward(sfwi_ctx *c, int nshots, const sfwi_shot *shots, float *b_z, float *b_x, float *b2_z,
                             float *b2_x) {
    if (!c || nshots < 0 || (nshots > 0 && !shots)) return fail(c, SFWI_EINVAL, "bad shot list");
    for (int b0 = 0; b0 < nshots; b0 += c->p.max_batch) {
        const int n = std::min(c->p.max_batch, nshots - b0);
        int r = prepare(c, K_BORN2_FWD, n, shots + b0, nullptr);
        if (r || (r = run(c, K_BORN2_FWD, n))) return r;
        for (int k = 0; k < n; ++k) {
            const size_t o = (size_t)(b0 + k) * c->nrec;
            const Slot &S = c->slots[k];
            const int ntr = shots[b0 + k].ntr;
            if ((r = copy_records(c, S.rz, b_z ? b_z + o : nullptr, ntr)) ||
                (r = copy_records(c, S.rx, b_x ? b_x + o : nullptr, ntr)) ||
                (r = copy_records(c, S.r2z, b2_z ? b2_z + o : nullptr, ntr)) ||
                (r = copy_records(c, S.r2x, b2_x ? b2_x + o : nullptr, ntr)))
                return r;
        }
        if ((r = finish(c))) return r;
    }
    return 0;
}

int sfwi_batch_hessian_vector(sfwi_ctx *c, int nshots, const sfwi_shot *shots, const float *p_vp,
                              const float *p_vs) {
    if (!c) return SFWI_EINVAL;
    if (!p_vp != !p_vs) return fail(c, SFWI_EINVAL, "p_vp and p_vs must both be given or both be NULL");
    if (p_vp) {
        int r = sfwi_set_perturbation_region(c, p_vp, p_vs, SFWI_BORN_S1);
        if (r) return r;
    } else if (!c->have_pert) {
        return fail(c, SFWI_ESTATE, "no vector: give p_vp, p_vs or set the perturbation");
    }
    return solve_misfits(c, K_HESS, nshots, shots, nullptr);
}

int sfwi_shot_hessian_vector(sfwi_ctx *c, const sfwi_shot *shot, const float *p_vp, const float *p_vs) {
    if (!c || !shot) return fail(c, SFWI_EINVAL, "NULL argument");
    return sfwi_batch_hessian_vector(c, 1, shot, p_vp, p_vs);
}

int sfwi_obs_put(sfwi_ctx *c, int id, int ntr, const float *obs_z, const float *obs_x) {
    if (!c || !obs_z || !obs_x) return fail(c, SFWI_EINVAL, "NULL argument");
    if (id <= 0) return fail(c, SFWI_EINVAL, "obs_id must be positive (got %d)", id);
    if (ntr < 1 || ntr > c->p.ntr) return fail(c, SFWI_EINVAL, "ntr=%d outside [1,%d]", ntr, c->p.ntr);
    ObsSet &o = c->obs[id];
    if (o.base && o.ntr != ntr) {   // a different size: reallocate
        HIPCK(c, hipStreamSynchronize(c->stream));
        hipFree(o.base);
        o.base = nullptr;
    }
    const size_t rb = (size_t)ntr * c->p.nt * sizeof(float);
    if (!o.base) {
        int r = dalloc(c, (void **)&o.base, 2 * rb);
        if (r) {
            c->obs.erase(id);
            return r;
        }
    }
    o.ntr = ntr;
    HIPCK(c, hipMemcpyAsync(o.base, obs_z, rb, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(o.base + (size_t)ntr * c->p.nt, obs_x, rb, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int sfwi_obs_clear(sfwi_ctx *c) {
    if (!c) return fail(c, SFWI_EINVAL, "NULL context");
    HIPCK(c, hipStreamSynchronize(c->stream));
    for (auto &kv : c->obs) hipFree(kv.second.base);
    c->obs.clear();
    return 0;
}

int sfwi_shot_gradient(sfwi_ctx *c, const sfwi_shot *shot, double *misfit) {
    if (!c) return SFWI_EINVAL;
    return solve_misfits(c, K_GRAD, 1, shot, misfit);
}
int sfwi_shot_born_gradient(sfwi_ctx *c, const sfwi_shot *shot, double *misfit) {
    if (!c) return SFWI_EINVAL;
    return solve_misfits(c, K_BORN_GRAD, 1, shot, misfit);
}
int sfwi_batch_gradient(sfwi_ctx *c, int nshots, const sfwi_shot *shots, double *misfits) {
    if (!c) return SFWI_EINVAL;
    return solve_misfits(c, K_GRAD, nshots, shots, misfits);
}
int sfwi_batch_born_gradient(sfwi_ctx *c, int nshots, const sfwi_shot *shots, double *misfits) {
    if (!c) return SFWI_EINVAL;
    return solve_misfits(c, K_BORN_GRAD, nshots, shots, misfits);
}
int sfwi_batch_born_gradient_res(sfwi_ctx *c, int nshots, const sfwi_shot *shots, double *misfits, float *res_z,
                                 float *res_x) {
    if (!c) return SFWI_EINVAL;
    if (!res_z || !res_x) return fail(c, SFWI_EINVAL, "NULL residuals");
    if (c->p.misfit != SFWI_MISFIT_L2) return fail(c, SFWI_EINVAL, "residuals are those of the L2 misfit");
    return solve_misfits(c, K_BORN_GRAD, nshots, shots, misfits, res_z, res_x);
}

int sfwi_shot_forward(sfwi_ctx *c, const sfwi_shot *shot, float *syn_z, float *syn_x) {
    if (!c || !shot) return fail(c, SFWI_EINVAL, "NULL argument");
    int r = prepare(c, K_FWD, 1, shot, nullptr);
    if (r || (r = run(c, K_FWD, 1))) return r;
    if ((r = copy_records(c, c->slots[0].rz, syn_z, shot->ntr)) || (r = copy_records(c, c->slots[0].rx, syn_x, shot->ntr)))
        return r;
    return finish(c);
}

int sfwi_batch_forward(sfwi_ctx *c, int nshots, const sfwi_shot *shots, float *syn_z, float *syn_x) {
    if (!c || nshots < 0 || (nshots > 0 && !shots)) return fail(c, SFWI_EINVAL, "bad shot list");
    for (int b0 = 0; b0 < nshots; b0 += c->p.max_batch) {
        const int n = std::min(c->p.max_batch, nshots - b0);
        int r = prepare(c, K_FWD, n, shots + b0, nullptr);
        if (r || (r = run(c, K_FWD, n))) return r;
        for (int k = 0; k < n; ++k) {
            const size_t o = (size_t)(b0 + k) * c->nrec;
            if ((r = copy_records(c, c->slots[k].rz, syn_z ? syn_z + o : nullptr, shots[b0 + k].ntr)) ||
                (r = copy_records(c, c->slots[k].rx, syn_x ? syn_x + o : nullptr, shots[b0 + k].ntr)))
                return r;
        }
        if ((r = finish(c))) return r;
    }
    return 0;
}

int sfwi_shot_born_forward(sfwi_ctx *c, const sfwi_shot *shot, float *b_z, float *b_x, float *bg_z, float *bg_x) {
    if (!c || !shot) return fail(c, SFWI_EINVAL, "NULL argument");
    int r = prepare(c, K_BORN_FWD, 1, shot, nullptr);
    if (r || (r = run(c, K_BORN_FWD, 1))) return r;
    const Slot &S = c->slots[0];
    if ((r = copy_records(c, S.rz, b_z, shot->ntr)) || (r = copy_records(c, S.rx, b_x, shot->ntr)) ||
        (r = copy_records(c, S.r2z, bg_z, shot->ntr)) || (r = copy_records(c, S.r2x, bg_x, shot->ntr)))
        return r;
    return finish(c);
}

int sfwi_linesearch_parabola(sfwi_ctx *c, const sfwi_shot *shot, const float *dir_vp, const float *dir_vs,
                             float step1, float step2, double m[3]) {
    if (!c || !shot || !dir_vp || !dir_vs || !m) return fail(c, SFWI_EINVAL, "NULL argument");
    if (!c->have_model) return fail(c, SFWI_ESTATE, "model not set");
    const size_t b = c->nmodel * sizeof(float);
    float *tvp = c->tmp_model, *tvs = tvp + c->nmodel, *dvp = tvp + 2 * c->nmodel, *dvs = tvp + 3 * c->nmodel;
    HIPCK(c, hipMemcpyAsync(dvp, dir_vp, b, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(dvs, dir_vs, b, hipMemcpyHostToDevice, c->stream));
    const float steps[3] = {0.0f, step1, step2};
    Window w;
    for (int k = 0; k < 3; ++k) {
        const float *mp = c->mvp, *ms = c->mvs;
        if (k > 0) {
            const dim3 grd((unsigned)((c->nmodel + 255) / 256));
            hipLaunchKernelGGL(k_axpy_model, grd, dim3(256), 0, c->stream, (int)c->nmodel, c->mvp, dvp, steps[k], tvp);
            hipLaunchKernelGGL(k_axpy_model, grd, dim3(256), 0, c->stream, (int)c->nmodel, c->mvs, dvs, steps[k], tvs);
            mp = tvp;
            ms = tvs;
        }
        int r = build_window(c, w, shot->left, mp, ms, c->mrho, nullptr, nullptr);
        if (r) { hipFree(w.base); return r; }
        if ((r = prepare(c, K_FWD_MISFIT, 1, shot, &w.m)) || (r = run(c, K_FWD_MISFIT, 1))) { hipFree(w.base); return r; }
        double v = 0;
        if (hipMemcpyAsync(&v, c->d_out, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            (r = finish(c))) { hipFree(w.base); return r ? r : fail(c, SFWI_EHIP, "copy failed"); }
        m[k] += v;
    }
    hipFree(w.base);
    return 0;
}

int sfwi_batch_linesearch_parabola(sfwi_ctx *c, int nshots, const sfwi_shot *shots, const float *dir_vp,
                                   const float *dir_vs, float step1, float step2, double m[3]) {
    if (!c || (nshots > 0 && !shots) || nshots < 0 || !dir_vp || !dir_vs || !m)
        return fail(c, SFWI_EINVAL, "NULL argument or negative shot count");
    if (!c->have_model) return fail(c, SFWI_ESTATE, "model not set");
    if (nshots == 0) return 0;
    const size_t b = c->nmodel * sizeof(float);
    float *tvp = c->tmp_model, *tvs = tvp + c->nmodel, *dvp = tvp + 2 * c->nmodel, *dvs = tvp + 3 * c->nmodel;
    HIPCK(c, hipMemcpyAsync(dvp, dir_vp, b, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(dvs, dir_vs, b, hipMemcpyHostToDevice, c->stream));
    const float steps[3] = {0.0f, step1, step2};
    std::map<int, Window> tw;   // one window of each test model per window position in the shot list
    auto done = [&](int r) {
        hipStreamSynchronize(c->stream);
        for (auto &kv : tw) hipFree(kv.second.base);
        return r;
    };
    std::vector<double> v(c->p.max_batch);
    for (int k = 0; k < 3; ++k) {
        const float *mp = c->mvp, *ms = c->mvs;
        if (k > 0) {
            const dim3 grd((unsigned)((c->nmodel + 255) / 256));
            hipLaunchKernelGGL(k_axpy_model, grd, dim3(256), 0, c->stream, (int)c->nmodel, c->mvp, dvp, steps[k], tvp);
            hipLaunchKernelGGL(k_axpy_model, grd, dim3(256), 0, c->stream, (int)c->nmodel, c->mvs, dvs, steps[k], tvs);
            mp = tvp;
            ms = tvs;
        }
        std::map<int, bool> built;
        for (int s = 0; s < nshots; ++s) {
            if (built.count(shots[s].left)) continue;
            built[shots[s].left] = true;
            if (int r = build_window(c, tw[shots[s].left], shots[s].left, mp, ms, c->mrho, nullptr, nullptr))
                return done(r);
        }
        for (int b0 = 0; b0 < nshots; b0 += c->p.max_batch) {
            const int n = std::min(c->p.max_batch, nshots - b0);
            int r = prepare(c, K_FWD_MISFIT, n, shots + b0, nullptr, &tw);
            if (r || (r = run(c, K_FWD_MISFIT, n))) return done(r);
            if (hipMemcpyAsync(v.data(), c->d_out, n * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
                return done(fail(c, SFWI_EHIP, "copy failed"));
            if ((r = finish(c))) return done(r);
            for (int j = 0; j < n; ++j) m[k] += v[j];   // shot order, as the per-shot calls add
        }
    }
    return done(0);
}

int sfwi_linesearch_linear(sfwi_ctx *c, const sfwi_shot *shot, const float *cg_vp, const float *cg_vs,
                           double *fenzi, double *fenmu) {
    if (!c || !shot || !cg_vp || !cg_vs || !fenzi || !fenmu) return fail(c, SFWI_EINVAL, "NULL argument");
    if ((!shot->obs_z || !shot->obs_x) && shot->obs_id <= 0) return fail(c, SFWI_EINVAL, "observed traces missing");
    // Solve 1: residual r = born(dm) - obs + background.
    int r = prepare(c, K_BORN_FWD, 1, shot, nullptr);
    if (r || (r = run(c, K_BORN_FWD, 1))) return r;
    Slot &S = c->slots[0];
    if ((r = load_obs(c, *shot, S))) return r;   // slot 0 now holds this shot's records
    const size_t rb = (size_t)shot->ntr * c->p.nt * sizeof(float);
    hipLaunchKernelGGL(k_residual<true>, dim3(256, 1), dim3(BLK), 0, c->stream, c->g, c->d_shots);
    HIPCK(c, hipMemcpyAsync(c->res_keep, S.rz, rb, hipMemcpyDeviceToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(c->res_keep + c->nrec, S.rx, rb, hipMemcpyDeviceToDevice, c->stream));
    // Solve 2: L(cg) = scattered traces with the perturbation cg.
    const size_t b = c->nmodel * sizeof(float);
    float *cvp = c->tmp_model + 2 * c->nmodel, *cvs = c->tmp_model + 3 * c->nmodel;
    HIPCK(c, hipMemcpyAsync(cvp, cg_vp, b, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(cvs, cg_vs, b, hipMemcpyHostToDevice, c->stream));
    Window w;
    if ((r = build_window(c, w, shot->left, c->mvp, c->mvs, c->mrho, cvp, cvs))) { hipFree(w.base); return r; }
    if ((r = prepare(c, K_BORN_FWD, 1, shot, &w.m)) || (r = run(c, K_BORN_FWD, 1))) { hipFree(w.base); return r; }
    hipLaunchKernelGGL(k_trace_dot, dim3((shot->ntr + 63) / 64), dim3(64), 0, c->stream, c->g, shot->ntr, S.rz, S.rx,
                       c->res_keep, c->res_keep + c->nrec, c->d_num, c->d_den);
    std::vector<double> num(shot->ntr), den(shot->ntr);
    if (hipMemcpyAsync(num.data(), c->d_num, shot->ntr * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(den.data(), c->d_den, shot->ntr * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess) {
        hipFree(w.base);
        return fail(c, SFWI_EHIP, "copy failed");
    }
    r = finish(c);
    hipFree(w.base);
    if (r) return r;
    // The per-trace FP64 partial sums are added in trace order.
    for (int tr = 0; tr < shot->ntr; ++tr) { *fenzi += num[tr]; *fenmu += den[tr]; }
    return 0;
}

}  // extern "C"

namespace {

// The batched linear step; with res_z / res_x (the residuals of a Born
// gradient call at the same model and images, [nshots][ntr][nt]) the first
// Born solve is replaced by their upload.
int linesearch_linear_batch(sfwi_ctx *c, int nshots, const sfwi_shot *shots, const float *cg_vp,
                            const float *cg_vs, const float *res_z, const float *res_x, double *fenzi,
                            double *fenmu) {
    if (!c || (nshots > 0 && !shots) || nshots < 0 || !cg_vp || !cg_vs || !fenzi || !fenmu)
        return fail(c, SFWI_EINVAL, "NULL argument or negative shot count");
    if (!c->have_model) return fail(c, SFWI_ESTATE, "model not set");
    if (!c->have_pert) return fail(c, SFWI_ESTATE, "perturbation not set (sfwi_set_perturbation)");
    if (!res_z != !res_x) return fail(c, SFWI_EINVAL, "res_z and res_x must both be given or both be NULL");
    for (int s = 0; s < nshots && !res_z; ++s)
        if ((!shots[s].obs_z || !shots[s].obs_x) && shots[s].obs_id <= 0)
            return fail(c, SFWI_EINVAL, "shot %d: observed traces missing", s);
    if (nshots == 0) return 0;
    const int nb = std::min(c->p.max_batch, nshots);
    // residuals of a batch ([2 nb][nrec]: z, x per slot) and per-trace sums ([2 nb][ntr])
    int r = grow(c, (void **)&c->res_batch, &c->res_batch_cap, 2 * (size_t)nb * c->nrec * sizeof(float));
    if (r || (r = grow(c, (void **)&c->dot_batch, &c->dot_batch_cap, 2 * (size_t)nb * c->p.ntr * sizeof(double))))
        return r;
    const size_t b = c->nmodel * sizeof(float);
    float *cvp = c->tmp_model + 2 * c->nmodel, *cvs = c->tmp_model + 3 * c->nmodel;
    HIPCK(c, hipMemcpyAsync(cvp, cg_vp, b, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(cvs, cg_vs, b, hipMemcpyHostToDevice, c->stream));
    std::map<int, Window> tw;   // one window of the direction per window position in the shot list
    auto done = [&](int rc) {
        hipStreamSynchronize(c->stream);
        for (auto &kv : tw) hipFree(kv.second.base);
        return rc;
    };
    for (int s = 0; s < nshots; ++s) {
        if (tw.count(shots[s].left)) continue;
        if ((r = build_window(c, tw[shots[s].left], shots[s].left, c->mvp, c->mvs, c->mrho, cvp, cvs)))
            return done(r);
    }
    std::vector<double> sums(2 * (size_t)nb * c->p.ntr);
    for (int b0 = 0; b0 < nshots; b0 += nb) {
        const int n = std::min(nb, nshots - b0);
        if (res_z) {   // residuals handed over by the Born gradient call
            for (int k = 0; k < n; ++k) {
                const size_t rb = (size_t)shots[b0 + k].ntr * c->p.nt * sizeof(float);
                const size_t o = (size_t)(b0 + k) * c->nrec;
                float *keep = c->res_batch + 2 * (size_t)k * c->nrec;
                if (hipMemcpyAsync(keep, res_z + o, rb, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
                    hipMemcpyAsync(keep + c->nrec, res_x + o, rb, hipMemcpyHostToDevice, c->stream) != hipSuccess)
                    return done(fail(c, SFWI_EHIP, "copy failed"));
            }
        }
        // Solve 1 over the batch: residual r = born(dm) - obs + background, kept per slot.
        if (!res_z && ((r = prepare(c, K_BORN_FWD, n, shots + b0, nullptr)) || (r = run(c, K_BORN_FWD, n))))
            return done(r);
        for (int k = 0; k < n && !res_z; ++k)
            if ((r = load_obs(c, shots[b0 + k], c->slots[k]))) return done(r);
        if (!res_z) hipLaunchKernelGGL(k_residual<true>, dim3(256, n), dim3(BLK), 0, c->stream, c->g, c->d_shots);
        for (int k = 0; k < n && !res_z; ++k) {
            Slot &S = c->slots[k];
            const size_t rb = (size_t)shots[b0 + k].ntr * c->p.nt * sizeof(float);
            float *keep = c->res_batch + 2 * (size_t)k * c->nrec;
            if (hipMemcpyAsync(keep, S.rz, rb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
                hipMemcpyAsync(keep + c->nrec, S.rx, rb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
                return done(fail(c, SFWI_EHIP, "copy failed"));
        }
        // Solve 2 over the batch: L(cg), then the per-trace products of each shot.
        if ((r = prepare(c, K_BORN_FWD, n, shots + b0, nullptr, &tw)) || (r = run(c, K_BORN_FWD, n)))
            return done(r);
        for (int k = 0; k < n; ++k) {
            Slot &S = c->slots[k];
            const int ntr = shots[b0 + k].ntr;
            const float *keep = c->res_batch + 2 * (size_t)k * c->nrec;
            double *num = c->dot_batch + 2 * (size_t)k * c->p.ntr;
            hipLaunchKernelGGL(k_trace_dot, dim3((ntr + 63) / 64), dim3(64), 0, c->stream, c->g, ntr, S.rz, S.rx,
                               keep, keep + c->nrec, num, num + c->p.ntr);
        }
        if (hipMemcpyAsync(sums.data(), c->dot_batch, 2 * (size_t)n * c->p.ntr * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            return done(fail(c, SFWI_EHIP, "copy failed"));
        if ((r = finish(c))) return done(r);
        // shot order, each shot's traces in trace order: the per-shot calls' sums
        for (int k = 0; k < n; ++k) {
            const double *num = sums.data() + 2 * (size_t)k * c->p.ntr, *den = num + c->p.ntr;
            for (int tr = 0; tr < shots[b0 + k].ntr; ++tr) { *fenzi += num[tr]; *fenmu += den[tr]; }
        }
    }
    return done(0);
}

}  // namespace

extern "C" {

int sfwi_batch_linesearch_linear(sfwi_ctx *c, int nshots, const sfwi_shot *shots, const float *cg_vp,
                                 const float *cg_vs, double *fenzi, double *fenmu) {
    return linesearch_linear_batch(c, nshots, shots, cg_vp, cg_vs, nullptr, nullptr, fenzi, fenmu);
}

int sfwi_batch_linesearch_linear_res(sfwi_ctx *c, int nshots, const sfwi_shot *shots, const float *cg_vp,
                                     const float *cg_vs, const float *res_z, const float *res_x, double *fenzi,
                                     double *fenmu) {
    if (!res_z || !res_x) return fail(c, SFWI_EINVAL, "NULL residuals");
    return linesearch_linear_batch(c, nshots, shots, cg_vp, cg_vs, res_z, res_x, fenzi, fenmu);
}

int sfwi_flush_rho(sfwi_ctx *c, float *grho) {
    if (!c) return SFWI_EINVAL;
    if (!c->p.grad_rho) return fail(c, SFWI_ESTATE, "the context was created without grad_rho");
    const size_t b = c->nmodel * sizeof(float);
    if (grho) HIPCK(c, hipMemcpyAsync(grho, c->Grho, b, hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemsetAsync(c->Grho, 0, b, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int sfwi_flush(sfwi_ctx *c, float *gvp, float *gvs, float *gil) {
    if (!c) return SFWI_EINVAL;
    const size_t b = c->nmodel * sizeof(float);
    if (gvp) HIPCK(c, hipMemcpyAsync(gvp, c->Gvp, b, hipMemcpyDeviceToHost, c->stream));
    if (gvs) HIPCK(c, hipMemcpyAsync(gvs, c->Gvs, b, hipMemcpyDeviceToHost, c->stream));
    if (gil) HIPCK(c, hipMemcpyAsync(gil, c->Gil, b, hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemsetAsync(c->Gvp, 0, 3 * b, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int sfwi_set_timing(sfwi_ctx *c, int every) {
    if (!c) return SFWI_EINVAL;
    c->timing_every = every > 0 ? every : 0;
    for (int k = 0; k < SFWI_K_COUNT; ++k) { c->kt_ms[k] = 0; c->kt_n[k] = 0; }
    return 0;
}

// Kernel-family autotuner of the FAST mode: for a batch of nshots shots of
// this context's grid, times a gradient solve of `steps` time steps (zero
// observed traces, one warm-up run per candidate) for each candidate and keeps
// the fastest in the context.  Candidates: per-cell kernels; column march with
// 8 or 16 columns per segment, prefetch depth 1, 2 or 4, shot-major,
// shot-interleaved or whole-shots-per-XCD block order; the forward solve in
// half steps or in fused whole steps (wave-level or workgroup march).  Accumulated gradients are preserved; the
// observed records resident in the batch slots are invalidated.
int sfwi_autotune(sfwi_ctx *c, int nshots, int steps, int *march_cols, int *march_depth, int *march_order,
                  double *best_ms) {
    if (!c) return SFWI_EINVAL;
    if (c->p.mode != SFWI_MODE_FAST) return fail(c, SFWI_ESTATE, "the autotuner chooses among FAST kernels");
    if (!c->have_model || !c->have_wav) return fail(c, SFWI_ESTATE, "set the model and the wavelet first");
    const int n = std::min(std::max(nshots, 1), c->p.max_batch);
    const int ns = std::min(std::max(steps, 2), c->p.nt);
    const int ntr = std::min(c->p.ntr, c->p.nx);
    std::vector<int> gc(ntr);
    for (int k = 0; k < ntr; ++k) gc[k] = (int)((float)k * c->p.dx);
    std::vector<float> zero((size_t)ntr * c->p.nt, 0.0f);
    std::vector<sfwi_shot> shots(n);
    for (int k = 0; k < n; ++k) {
        sfwi_shot &s = shots[k];
        s.left = 0;
        s.isx = c->g.pml + (int)((k + 0.5) * c->p.nx / n);
        s.isz = c->g.pml + 1;
        s.igz = c->g.pml;
        s.ngx_left = 0;
        s.ntr = ntr;
        s.gc = gc.data();
        s.obs_z = s.obs_x = zero.data();
    }
    const size_t acc = 4 * c->nmodel * sizeof(float);
    HIPCK(c, hipMemcpyAsync(c->tmp_model, c->Gvp, acc, hipMemcpyDeviceToDevice, c->stream));
    // kernel family, segment, prefetch depth, block order, and the time-blocking
    // depth of the forward solve: half steps (fuse 0) or whole fused steps
    // (1: wave-level march, 2: workgroup march with the z halo in LDS, 3: the
    // same for a snapshot gradient's adjoint steps too)
    struct Cand { int cols, depth, order, fuse; };
    const Cand cands[] = {{-1, 1, 1, 0}, {8, 4, 3, 0}, {8, 2, 3, 0}, {8, 1, 3, 0}, {16, 2, 3, 0}, {8, 4, 1, 0},
                          {8, 2, 2, 0}, {8, 2, 3, 1}, {8, 1, 3, 1}, {8, 2, 3, 2}, {8, 1, 3, 2}, {8, 1, 3, 3}};
    const sfwi_params keep = c->p;
    const int nt_keep = c->g.nt;
    c->g.nt = ns;                              // a shorter record inside the same buffers
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int r = 0;
    double best = 1e300;
    Cand win = cands[1];
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) r = fail(c, SFWI_EHIP, "hipEventCreate");
    for (const Cand &cd : cands) {
        if (r) break;
        if (cd.fuse >= 2 && !c->wg_ok) continue;   // the workgroup marches' 4 GiB state-block bound
        c->p.march_cols = cd.cols;
        c->p.march_depth = cd.depth;
        c->p.march_order = cd.order;
        c->p.fuse_step = cd.fuse;
        set_march_geo(c->g, cd.cols);
        float ms = 0.0f;
        for (int rep = 0; rep < 2 && !r; ++rep) {
            if ((r = prepare(c, K_GRAD, n, shots.data(), nullptr))) break;
            if (rep && hipEventRecord(e0, c->stream) != hipSuccess) { r = fail(c, SFWI_EHIP, "hipEventRecord"); break; }
            if ((r = run_direct(c, K_GRAD, n))) break;
            if (rep && hipEventRecord(e1, c->stream) != hipSuccess) { r = fail(c, SFWI_EHIP, "hipEventRecord"); break; }
            if (hipStreamSynchronize(c->stream) != hipSuccess) { r = fail(c, SFWI_EHIP, "autotune solve failed"); break; }
        }
        if (!r && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best) {
            best = ms;
            win = cd;
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    c->g.nt = nt_keep;
    c->p = keep;
    c->p.march_cols = win.cols;
    c->p.march_depth = win.depth;
    c->p.march_order = win.order;
    c->p.fuse_step = win.fuse;
    set_march_geo(c->g, win.cols);
    drop_graphs(c);                            // captured with another kernel family
    for (Slot &S : c->slots) S.obs_ntr = 0;
    HIPCK(c, hipMemcpyAsync(c->Gvp, c->tmp_model, acc, hipMemcpyDeviceToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (r) return r;
    if (march_cols) *march_cols = win.cols;
    if (march_depth) *march_depth = win.depth;
    if (march_order) *march_order = win.order;
    if (best_ms) *best_ms = best;
    return 0;
}

int sfwi_get_params(const sfwi_ctx *c, sfwi_params *out) {
    if (!c || !out) return SFWI_EINVAL;
    *out = c->p;
    return 0;
}

int sfwi_kernel_time(sfwi_ctx *c, int kid, double *avg_ms, int *launches) {
    if (!c || kid < 0 || kid >= SFWI_K_COUNT) return fail(c, SFWI_EINVAL, "bad kernel id %d", kid);
    if (avg_ms) *avg_ms = c->kt_n[kid] ? c->kt_ms[kid] / c->kt_n[kid] : 0.0;
    if (launches) *launches = c->kt_n[kid];
    return 0;
}

const char *sfwi_kernel_name(int kid) {
    static const char *names[SFWI_K_COUNT] = {
        "k_vel(forward)", "k_str(forward)", "k_str(reconstruct)", "k_vel(reconstruct)", "k_adj_str",
        "k_adj_vel_image", "k_str(reconstruct+image)", "k_vel(reconstruct+illum)", "k_adj_vel",
        "k_snap_image", "k_snap_store", "k_step(forward)", "k_step(adjoint)"};
    return (kid >= 0 && kid < SFWI_K_COUNT) ? names[kid] : "?";
}

// Host-driver helper, same arithmetic as sfwi-stage1/main.cpp:251-260:
// Nk in double stored as float; the exponent's argument is float (Nk is a
// float and t-t0 an int), so the float overload of exp is taken.
void sfwi_ricker(float *w, int nt, float dt, float f0) {
    const double pi = 3.14159265358979;
    const float nk = pi * pi * f0 * f0 * dt * dt;
    const int t0 = (int)std::ceil(1.0 / (f0 * dt));
    for (int t = 0; t < nt; ++t)
        w[t] = (1.0 - 2.0 * nk * (t - t0) * (t - t0)) * std::exp(-nk * (t - t0) * (t - t0));
}

int sfwi_ot_misfit(sfwi_ctx *c, int variant, int ntr, int nt, float dt, const float *obs_x, const float *obs_z,
                   const float *syn_x, const float *syn_z, float *res_x, float *res_z, double *misfit) {
    if (!c) return fail(c, SFWI_EINVAL, "NULL context");
    if (!obs_x || !obs_z || !syn_x || !syn_z || !res_x || !res_z || !misfit)
        return fail(c, SFWI_EINVAL, "NULL argument");
    if (variant < 0 || variant > 2) return fail(c, SFWI_EINVAL, "unknown OT variant %d", variant);
    if (ntr < 1 || nt < 1) return fail(c, SFWI_EINVAL, "ntr=%d nt=%d must be positive", ntr, nt);
    if (!(dt > 0)) return fail(c, SFWI_EINVAL, "dt must be positive");
    const size_t m = (size_t)ntr * nt;
    if (m >= (1u << 30)) return fail(c, SFWI_EINVAL, "gather of %zu samples too large", m);
    int r;
    if ((r = grow(c, (void **)&c->ot_io, &c->ot_io_cap, 4 * m * sizeof(float)))) return r;
    float *ox = c->ot_io, *oz = ox + m, *rx = oz + m, *rz = rx + m;
    const size_t nj = variant == 2 ? 2 : 2 * (size_t)ntr;
    if ((r = grow(c, (void **)&c->ot_mis, &c->ot_mis_cap, nj * sizeof(double)))) return r;
    HIPCK(c, hipMemcpyAsync(ox, obs_x, m * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(oz, obs_z, m * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(rx, syn_x, m * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(rz, syn_z, m * sizeof(float), hipMemcpyHostToDevice, c->stream));
    c->ot_h.clear();
    float *syn[2] = {rx, rz};
    const float *obs[2] = {ox, oz};
    for (int comp = 0; comp < 2; ++comp) {
        if (variant == 2) {
            c->ot_h.push_back({syn[comp], obs[comp], syn[comp], c->ot_mis + c->ot_h.size(), (int)m, 0});
        } else {
            for (int tr = 0; tr < ntr; ++tr) {
                const size_t o = (size_t)tr * nt;
                c->ot_h.push_back({syn[comp] + o, obs[comp] + o, syn[comp] + o, c->ot_mis + c->ot_h.size(), nt,
                                   variant == 1});
            }
        }
    }
    if ((r = ot_launch(c, variant == 2 ? m : (size_t)nt, (double)dt * (double)dt))) return r;
    std::vector<double> mis(nj);
    HIPCK(c, hipMemcpyAsync(mis.data(), c->ot_mis, nj * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemcpyAsync(res_x, rx, m * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemcpyAsync(res_z, rz, m * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    if ((r = check_ot(c))) return r;
    double total = 0.0;
    for (double v : mis) total += v;
    *misfit += total;
    return 0;
}

int sfwi_ot_misfit_static(sfwi_ctx *c, int variant, int ntr, int nt, float dt, const float *obs_x,
                          const float *obs_z, const float *syn_x, const float *syn_z, float *res_x, float *res_z,
                          double *misfit) {
    if (!c) return fail(c, SFWI_EINVAL, "NULL context");
    if (!obs_x || !obs_z || !syn_x || !syn_z || !res_x || !res_z || !misfit)
        return fail(c, SFWI_EINVAL, "NULL argument");
    if (variant < 0 || variant > 2) return fail(c, SFWI_EINVAL, "unknown OT variant %d", variant);
    if (ntr < 1 || nt < 1) return fail(c, SFWI_EINVAL, "ntr=%d nt=%d must be positive", ntr, nt);
    if (!(dt > 0)) return fail(c, SFWI_EINVAL, "dt must be positive");
    const size_t m = (size_t)ntr * nt;
    if (m > NA_MAX_SAMPLES || ntr > NA_MAX_TRACES)
        return fail(c, SFWI_EINVAL, "gather of %d x %d exceeds the static workspace (at most %zu samples per "
                    "component and %d traces)", ntr, nt, NA_MAX_SAMPLES, NA_MAX_TRACES);
    char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_m);
        base = g_scratch[c->device];
    }
    if (!base) return fail(c, SFWI_ESTATE, "static workspace not resolved on device %d", c->device);
    int *bad = (int *)base;
    double *total_d = (double *)(base + 8);
    OtJob *jobs = (OtJob *)(base + NA_JOBS);
    double *mis = (double *)(base + NA_MIS);
    float *ox = (float *)(base + NA_IO), *oz = ox + m, *rx = oz + m, *rz = rx + m;
    double *ws = (double *)(base + NA_WS);
    const size_t n_max = variant == 2 ? m : (size_t)nt;
    const size_t nj = variant == 2 ? 2 : 2 * (size_t)ntr;
    const size_t stride = 4 * (n_max + 1);
    const size_t fit = NA_WS_BYTES / (stride * sizeof(double));   // >= 2: n_max <= NA_MAX_SAMPLES
    const int blocks = (int)std::min(std::min(nj, fit), (size_t)OT_MAX_BLOCKS);
    HIPCK(c, hipMemcpyAsync(ox, obs_x, m * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(oz, obs_z, m * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(rx, syn_x, m * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(rz, syn_z, m * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemsetAsync(bad, 0, sizeof(int), c->stream));   // the walk flag starts clear on every call
    hipLaunchKernelGGL(k_ot_jobs, dim3((unsigned)((nj + 63) / 64)), dim3(64), 0, c->stream, jobs, variant, ntr, nt,
                       rx, rz, ox, oz, mis);
    hipLaunchKernelGGL(k_ot, dim3(blocks), dim3(OT_THREADS), 0, c->stream, jobs, (int)nj, ws, stride,
                       (double)dt * (double)dt, bad);
    hipLaunchKernelGGL(k_sum_in_order, dim3(1), dim3(64), 0, c->stream, mis, (int)nj, total_d);
    HIPCK(c, hipGetLastError());
    double total = 0.0;
    int flag = 0;
    HIPCK(c, hipMemcpyAsync(&total, total_d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemcpyAsync(&flag, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemcpyAsync(res_x, rx, m * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemcpyAsync(res_z, rz, m * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (flag) return fail(c, SFWI_EHIP, "OT misfit: a merge walk did not end within its bound (invalid CDF levels)");
    *misfit += total;
    return 0;
}

}  // extern "C"

#include "server.hpp"
#include "inversion.hpp"
#include "suio.hpp"
