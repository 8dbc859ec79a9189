// Optimal-transport (quadratic Wasserstein) trace misfit and adjoint source.
//
// The reference declares OT, OT_power and OT_alltrace (sfwi-stage1/FD2DGPU.cuh:21-25)
// with the interface (obs_x, obs_z, cal_x, cal_z, residual_x, residual_z,
// length, tracenum, double& misfit) and never defines them.  This is the 1-D
// W2^2 between piecewise-constant densities on unit sample cells:
//
//   transform   u = s + c, c = 1.1 max(max|s|, max|o|)   (shift variants)
//               u = s^2                                   (power variant)
//   CDFs        P_j = sum_{i<j} u_i / M,  Q likewise, cumulative sums in FP64
//   misfit      W = int_0^1 (F^-1(y) - G^-1(y))^2 dy, exact on each segment
//               between merged breakpoints of P and Q
//   adjoint     g_j = -(B_j + sum_{c>j} A_c) with A_c = int_c h, B_c = int_c h (x-c),
//               h(x) = 2 (x - G^-1(F(x))), then dW/ds_j = u'(s_j) (g_j - sum p g) / M
//   units       misfit * dt^2, adjoint source * dt^2.
//
// One workgroup of 256 threads owns one job (one trace of one component, or a
// whole flattened gather) at a time; thread t owns the contiguous cell chunk
// [t L, (t+1) L), L = ceil(n / 256).  P, Q, A, B live in a per-workgroup FP64
// workspace of 4 (n + 1) doubles; the scans over chunk totals run on one
// thread (256 terms) so the result does not depend on timing.
#pragma once
#include "ot_walk.hpp"

namespace sfwi {

constexpr int OT_THREADS = 256;

struct OtJob {
    const float *syn;   // n samples; may alias res
    const float *obs;   // n samples
    float *res;         // adjoint source, n samples
    double *misfit;     // W2^2 * dt^2 of this job
    int n;
    int power;          // 1: u = s^2, 0: u = s + c
};

__device__ inline double ot_block_sum(double v, double *red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
        for (int k = 0; k < OT_THREADS; ++k) s += red[k];
        red[OT_THREADS] = s;
    }
    __syncthreads();
    const double s = red[OT_THREADS];
    __syncthreads();
    return s;
}

__device__ inline float ot_block_max(float v, float *redf) {
    const int t = threadIdx.x;
    redf[t] = v;
    __syncthreads();
    if (t == 0) {
        float m = 0.0f;
        for (int k = 0; k < OT_THREADS; ++k) m = fmaxf(m, redf[k]);
        redf[OT_THREADS] = m;
    }
    __syncthreads();
    const float m = redf[OT_THREADS];
    __syncthreads();
    return m;
}

// Exclusive prefix (from the left) of the per-thread values, and the total.
__device__ inline double ot_block_prefix(double v, double *red, double *total) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
        for (int k = 0; k < OT_THREADS; ++k) {
            const double x = red[k];
            red[k] = s;
            s += x;
        }
        red[OT_THREADS] = s;
    }
    __syncthreads();
    const double p = red[t];
    *total = red[OT_THREADS];
    __syncthreads();
    return p;
}

// Exclusive suffix (from the right) of the per-thread values.
__device__ inline double ot_block_suffix(double v, double *red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
        for (int k = OT_THREADS - 1; k >= 0; --k) {
            const double x = red[k];
            red[k] = s;
            s += x;
        }
    }
    __syncthreads();
    const double p = red[t];
    __syncthreads();
    return p;
}

__global__ void __launch_bounds__(OT_THREADS)
k_ot(const OtJob *__restrict__ jobs, int njobs, double *__restrict__ ws, size_t ws_stride, double dt2, int *bad) {
    __shared__ double red[OT_THREADS + 1];
    __shared__ float redf[OT_THREADS + 1];
    const int t = threadIdx.x;
    double *P = ws + (size_t)blockIdx.x * ws_stride;
    for (int jb = blockIdx.x; jb < njobs; jb += gridDim.x) {
        const OtJob J = jobs[jb];
        const int n = J.n;
        double *Q = P + (n + 1), *A = Q + (n + 1), *B = A + (n + 1);
        const int L = (n + OT_THREADS - 1) / OT_THREADS;
        const int j0 = min(n, t * L), j1 = min(n, j0 + L);

        double c = 0.0;
        if (!J.power) {
            float m = 0.0f;
            for (int j = j0; j < j1; ++j) m = fmaxf(m, fmaxf(fabsf(J.syn[j]), fabsf(J.obs[j])));
            c = 1.1 * (double)ot_block_max(m, redf);
        }
        // chunk sums, then the FP64 cumulative sums P and Q
        double su = 0.0, sv = 0.0;
        for (int j = j0; j < j1; ++j) {
            const double s = J.syn[j], o = J.obs[j];
            su += J.power ? s * s : s + c;
            sv += J.power ? o * o : o + c;
        }
        double Mu, Mv;
        double cu = ot_block_prefix(su, red, &Mu);
        double cv = ot_block_prefix(sv, red, &Mv);
        if (!(Mu > 0.0 && Mv > 0.0)) {
            for (int j = j0; j < j1; ++j) J.res[j] = 0.0f;
            if (t == 0) *J.misfit = 0.0;
            __syncthreads();
            continue;
        }
        if (t == 0) { P[0] = 0.0; Q[0] = 0.0; }
        // The running sums and the totals add the same terms in different
        // orders, so a fraction can round above 1 where the mass still to come
        // is below the rounding of the total (a trace whose energy has ended):
        // clamp to 1, which the last cell holds exactly.  The merge walk below
        // relies on every P[i] <= Q[n] = 1.
        for (int j = j0; j < j1; ++j) {
            const double s = J.syn[j], o = J.obs[j];
            cu += J.power ? s * s : s + c;
            cv += J.power ? o * o : o + c;
            P[j + 1] = fmin(cu / Mu, 1.0);
            Q[j + 1] = fmin(cv / Mv, 1.0);
        }
        if (j0 < j1 && j1 == n) { P[n] = 1.0; Q[n] = 1.0; }   // owner of the last cell
        __syncthreads();

        // per cell of F: exact segment integrals
        double w = 0.0;
        for (int i = j0; i < j1; ++i) {
            const double Pi = P[i], Pe = P[i + 1], pp = Pe - Pi;
            double a = 0.0, b = 0.0;
            if (pp > 0.0) {
                // a walk that does not end the cell means levels above Q[n]: flag it
                if (!ot_cell_walk(Q, n, i, Pi, Pe, w, a, b)) *bad = 1;
            } else {
                // empty cell: T = G^-1 just above the level Pi
                const int k = ot_floor_index(Q, n, Pi);
                const double tau = (Q[k + 1] > Pi) ? k + (Pi - Q[k]) / (Q[k + 1] - Q[k]) : (double)n;
                a = 2.0 * ((i + 0.5) - tau);
                b = 2.0 * ((i - tau) / 2.0 + 1.0 / 3.0);
            }
            A[i] = a;
            B[i] = b;
        }
        const double W = ot_block_sum(w, red) / 12.0;

        // g_j = -(B_j + sum_{c>j} A_c), overwriting A; sum_j p_j g_j
        double chunk = 0.0;
        for (int i = j0; i < j1; ++i) chunk += A[i];
        double sfx = ot_block_suffix(chunk, red);
        double pg = 0.0;
        for (int i = j1 - 1; i >= j0; --i) {
            const double ai = A[i];
            const double g = -(B[i] + sfx);
            sfx += ai;
            A[i] = g;
            pg += (P[i + 1] - P[i]) * g;
        }
        const double gbar = ot_block_sum(pg, red);
        for (int j = j0; j < j1; ++j) {
            const double du = J.power ? 2.0 * (double)J.syn[j] : 1.0;
            J.res[j] = (float)(du * (A[j] - gbar) / Mu * dt2);
        }
        if (t == 0) *J.misfit = W * dt2;
        __syncthreads();   // the workspace is reused by the next job of this block
    }
}

// Job table of one gather on the device (the layout sfwi_ot_misfit builds on
// the host): components x then z; per trace (variants 0, 1) or one job of the
// whole flattened gather (variant 2); the adjoint source overwrites the
// synthetic traces; job j's misfit goes to mis[j].
__global__ void k_ot_jobs(OtJob *jobs, int variant, int ntr, int nt, float *syn_x, float *syn_z,
                          const float *obs_x, const float *obs_z, double *mis) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = variant == 2 ? 1 : ntr;
    if (j >= 2 * per) return;
    const int comp = j / per, tr = j - comp * per;
    float *syn = comp ? syn_z : syn_x;
    const float *obs = comp ? obs_z : obs_x;
    OtJob J;
    if (variant == 2) {
        J.syn = syn; J.obs = obs; J.n = ntr * nt; J.power = 0;
    } else {
        const size_t o = (size_t)tr * nt;
        J.syn = syn + o; J.obs = obs + o; J.n = nt; J.power = variant == 1;
    }
    J.res = const_cast<float *>(J.syn);
    J.misfit = mis + j;
    jobs[j] = J;
}

// Sum of n values in index order by one thread (the order of a host loop).
__global__ void k_sum_in_order(const double *v, int n, double *out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += v[k];
    *out = s;
}

// Stage-2 synthetic traces for the OT misfit: born + background, in place
// (the L2 residual is born - obs + background, sfwi-stage2/FD2DGPU.cu:3119-3158).
__global__ void k_born_total(Geo g, const DevShot *__restrict__ shots) {
    const DevShot &d = shots[blockIdx.y];
    const size_t n = (size_t)d.ntr * g.nt;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        d.rz[k] += d.r2z[k];
        d.rx[k] += d.r2x[k];
    }
}

// Per-shot OT misfit: ordered sum over the shot's jobs (x traces, then z).
__global__ void k_ot_shot_sum(const double *__restrict__ mis, const int *__restrict__ first, double *out,
                              int nshots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    double m = 0.0;
    for (int j = first[s]; j < first[s + 1]; ++j) m += mis[j];
    out[s] = m;
}

}  // namespace sfwi
