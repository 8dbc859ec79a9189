"""Closed-form pin of the propagator: the homogeneous acoustic limit (Vs = 0)
against the analytic 2-D Green's function.

With mu = 0 the stresses stay isotropic (txx = tzz = -p) and the scheme
solves the acoustic system rho v_t = -grad p.  The explosive source adds the
wavelet to txx and tzz before each velocity update, and the stress update
rebuilds the source cell's totals from split parts that never hold it
(sfwi-stage1/FD2DGPU.cu:2104-2114 with the split source cell), so the
wavelet acts as an instantaneous stress w(t) on one cell of area dx dz: a
stress-rate source S(t) = dx dz w'(t) delta(x - xs).  Then

    p_tt - c^2 lap p = -S'(t) delta,   rho v = grad (G *_t S),
    G(r, t) = H(ct - r) / (2 pi c sqrt(c^2 t^2 - r^2)),
    (G *_t S)(r, t) = 1 / (2 pi c^2) int_0^acosh(ct/r) S(t - (r/c) cosh u) du.

vx lives half a cell to the right of the stress cell, vz half a cell below,
and the record of step it holds v at (it + 1/2) dt (leapfrog).  The FP64
oracle must match this to the scheme's second-order time error: the error
falls 4x when dt halves, while the 8th-order space error is already below it
(dx 10 m and 5 m give the same error).  A source added after the velocity
update (plant 4) never reaches the field."""
import numpy as np

from oracle import oracle as O

VP, RHO, F0 = 2000.0, 1000.0, 12.0
T0 = 1.6 / F0   # the wavelet starts at exp(-(1.6 pi)^2) ~ 1e-11 of its peak: no jump at t = 0


def _ricker(t):
    a = (np.pi * F0 * (t - T0)) ** 2
    return (1 - 2 * a) * np.exp(-a)


def _dricker(t):
    a = (np.pi * F0 * (t - T0)) ** 2
    return -2 * (np.pi * F0) ** 2 * (t - T0) * np.exp(-a) * (3 - 2 * a)


def _analytic(dx, xs, zs, xr, zr, times, nu=2000):
    """vx, vz of the Green's function solution at the staggered positions of
    receiver cell (xr, zr) (metres of the stress cell)."""
    def conv(r):   # (G *_t S)(r, t) by the cosh substitution (smooth integrand)
        u = np.linspace(0.0, 1.0, nu)[None, :] * np.arccosh(np.maximum(VP * times / r, 1.0))[:, None]
        tt = times[:, None] - (r / VP) * np.cosh(u)
        s = np.where(tt >= 0, _dricker(tt), 0.0) * dx * dx
        return np.trapezoid(s, u, axis=1) / (2 * np.pi * VP ** 2)
    out = []
    for j, (x, z) in enumerate(((xr + 0.5 * dx, zr), (xr, zr + 0.5 * dx))):
        r = np.hypot(x - xs, z - zs)
        h = 1e-3 * dx
        dr = (conv(r + h) - conv(r - h)) / (2 * h)
        out.append(dr * ((x - xs) if j == 0 else (z - zs)) / r / RHO)
    return out


def _error(dx, dt, shift=0.5, plant=0):
    L = 800.0
    n, pml = int(L / dx), int(100 / dx)
    nt = int(round(0.4 / dt))
    vp = np.full((n, n), VP, np.float32)
    isx = isz = pml + n // 2
    igz = isz - int(150 / dx)                                  # receivers 150 m above the source
    gc = (np.arange(-200, 201, 50) + L / 2).astype(np.int32)   # offsets -200 .. 200 m
    c = O.config(allnx=n, allnz=n, nx=n, nz=n, pml=pml, nt=nt, ntr=len(gc), dx=dx, dz=dx, dt=dt, left=0,
                 isx=isx, isz=isz, igz=igz, ngx_left=0, plant=plant)
    dt32 = float(np.float32(dt))
    w = _ricker(np.arange(nt) * dt32).astype(np.float32)
    rz, rx = O.forward(c, vp, np.zeros_like(vp), np.full_like(vp, RHO), w, gc, "f64")
    times = (np.arange(nt) + shift) * dt32
    num, ana = [], []
    for k, g in enumerate(gc):
        ax, az = _analytic(dx, isx * dx, isz * dx, (pml + int(g / dx)) * dx, igz * dx, times)
        num += [rx[k], rz[k]]
        ana += [ax, az]
    num, ana = np.array(num), np.array(ana)
    return float(np.linalg.norm(num - ana) / np.linalg.norm(ana))


def test_acoustic_limit_matches_the_green_function():
    e1 = _error(10.0, 0.0005)
    e2 = _error(10.0, 0.00025)
    assert e1 <= 3e-3, e1
    assert 3.5 <= e1 / e2 <= 4.5, (e1, e2)               # second order in time
    assert abs(_error(5.0, 0.0005) / e1 - 1) <= 0.1      # space error already below it
    for shift in (0.0, 1.0):                             # the record is v at (it + 1/2) dt
        assert _error(10.0, 0.0005, shift) >= 5 * e1


def test_late_source_misses_the_green_function():
    assert _error(10.0, 0.0005, plant=4) >= 0.99
