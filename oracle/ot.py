"""Optimal-transport trace misfit and adjoint source -- CPU oracle (numpy, FP64).

TEST INFRASTRUCTURE ONLY: the oracle of the HIP implementation
(sfwi_ot_misfit and the OT misfit option of the gradient path).  Only tests/,
the smoke check and bench.py's cpu_baseline leg may import it.

The reference declares the interface but never defines it (OT, OT_power,
OT_alltrace, sign: sfwi-stage1/FD2DGPU.cuh:21-25; OT_W2_CAL :404; commented
call sites sfwi-stage1/FD2DGPU.cu:4887-4896), so this restates the textbook
1-D quadratic Wasserstein distance (Villani, "Topics in Optimal
Transportation", thm 2.18: W2^2 = int_0^1 (F^-1(y) - G^-1(y))^2 dy; its
application to FWI traces as in Yang, Engquist, Sun & Hamfeldt, Geophysics 83
(2018) R43-R62) for one fixed discretisation:

* A trace is turned into a non-negative signal by a transform (variant 0:
  u = s + c with c = 1.1 max(max|s|, max|o|) over the trace; variant 1:
  u = s^2; variant 2: as variant 0 over the whole gather of a component,
  flattened trace after trace) and read as a piecewise-constant density on
  unit cells [j, j+1) of the sample axis, normalised to unit mass.
* F, G (cumulative sums in FP64) are piecewise linear; F^-1, G^-1 are
  piecewise linear between the merged breakpoints, so W2^2 is integrated
  exactly segment by segment.
* The first variation of W2^2 with respect to the mass of cell j is
  g_j = -int_j^{j+1} int_s^n 2 (x - T(x)) dx ds with T = G^-1 o F, again
  exact per segment.  For an empty cell of F, T is G^-1 just above the cell's
  level.  The adjoint source is the chain rule through the normalisation and
  the transform: dW/ds_j = u'(s_j) (g_j - sum_i p_i g_i) / M.  The shift c of
  variants 0 and 2 is held constant in the adjoint.
* Time units: misfit = W2^2 * dt^2, adjoint source = dW/ds * dt^2.
* A trace whose transformed syn or obs has zero mass contributes nothing
  (misfit 0, adjoint source 0).
"""
import numpy as np

SHIFT_FACTOR = 1.1


def w2_cells(a, b):
    """W2^2 between the piecewise-constant densities with cell masses a and b
    (non-negative, any positive total) in sample units, and the first variation
    g (per cell of a) and the normalised masses p of a."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = a.size
    ca, cb = np.cumsum(a), np.cumsum(b)
    P = np.concatenate(([0.0], ca / ca[-1]))
    Q = np.concatenate(([0.0], cb / cb[-1]))
    pp, qq = np.diff(P), np.diff(Q)
    y = np.union1d(P, Q)
    ya, yb = y[:-1], y[1:]
    # the cells holding segment [ya, yb]: last breakpoint <= ya (no breakpoint
    # lies strictly inside a merged segment)
    i = np.clip(np.searchsorted(P, ya, "right") - 1, 0, n - 1)
    k = np.clip(np.searchsorted(Q, ya, "right") - 1, 0, n - 1)
    xa = i + (ya - P[i]) / pp[i]
    xb = i + (yb - P[i]) / pp[i]
    ha = 2.0 * (xa - (k + (ya - Q[k]) / qq[k]))
    hb = 2.0 * (xb - (k + (yb - Q[k]) / qq[k]))
    W = float(np.sum((yb - ya) * (ha * ha + ha * hb + hb * hb)) / 12.0)
    A = np.zeros(n)
    B = np.zeros(n)
    dx = xb - xa
    ua, ub = xa - i, xb - i
    np.add.at(A, i, dx * (ha + hb) / 2.0)
    np.add.at(B, i, dx * (ha * (2 * ua + ub) + hb * (ua + 2 * ub)) / 6.0)
    z = np.nonzero(pp == 0)[0]
    if z.size:
        kz = np.minimum(np.searchsorted(Q, P[z], "right") - 1, n - 1)
        inside = Q[kz + 1] > P[z]
        tau = np.where(inside, kz + (P[z] - Q[kz]) / np.where(inside, qq[kz], 1.0), float(n))
        A[z] = 2.0 * ((z + 0.5) - tau)
        B[z] = 2.0 * ((z - tau) / 2.0 + 1.0 / 3.0)
    suffix = np.concatenate((np.cumsum(A[::-1])[::-1][1:], [0.0]))   # sum_{c > j} A_c
    g = -(B + suffix)
    return W, g, pp


def ot_signal(s, o, variant, dt, shift=None):
    """Misfit (time units) and adjoint source of one signal pair."""
    s = np.asarray(s, np.float64)
    o = np.asarray(o, np.float64)
    if variant == 1:
        u, v, du = s * s, o * o, 2.0 * s
    else:
        c = SHIFT_FACTOR * max(np.abs(s).max(), np.abs(o).max()) if shift is None else shift
        u, v, du = s + c, o + c, np.ones_like(s)
    mu, mv = u.sum(), v.sum()
    if not (mu > 0 and mv > 0):
        return 0.0, np.zeros_like(s)
    W, g, p = w2_cells(u, v)
    gbar = float(np.dot(p, g))
    res = du * (g - gbar) / mu
    return W * dt * dt, res * dt * dt


def ot_misfit(obs_x, obs_z, syn_x, syn_z, dt, variant=0):
    """Gather misfit (sum over traces and both components) and adjoint sources
    res_x, res_z ([ntr][nt] float32, as the library writes them)."""
    total = 0.0
    out = []
    for obs, syn in ((obs_x, syn_x), (obs_z, syn_z)):
        obs = np.asarray(obs, np.float32)
        syn = np.asarray(syn, np.float32)
        res = np.zeros(syn.shape, np.float64)
        if variant == 2:
            m, r = ot_signal(syn.ravel(), obs.ravel(), 2, dt)
            res = r.reshape(syn.shape)
            total += m
        else:
            for tr in range(syn.shape[0]):
                m, r = ot_signal(syn[tr], obs[tr], variant, dt)
                res[tr] = r
                total += m
        out.append(res)
    return out[0], out[1], total
