"""Pins of the CPU oracle against mathematics, independent of its own formulas.

The reference ships no tests or golden vectors and cannot run here (CUDA +
MPI), so each part of the oracle is checked against a property that a
plausible restatement mistake (a dropped term, a wrong sign or tap, swapped
operands) would break:

  * FD coefficients: Taylor moments of the 8th-order staggered stencil.
  * Stencil staggering: D+ / D- of a sinusoid equal its derivative half a
    cell ahead / behind.
  * Medium set-up: edge-clamped window, PML profile and Lame parameters in
    closed form.
  * Propagator: discrete energy of the leapfrog is conserved with the PML
    off (needs every coupling sign and the lambda / mu placement right), and
    reconstruction steps invert forward steps exactly.
  * Adjoint kernels: each adjoint half step is the transpose of the matching
    forward half step under a = rho v_adj, b = -sigma_adj (dot-product test).
  * Born modeling: the exact discrete linearization of forward modeling
    (central differences converge at second order to the Born traces).
  * Composite solves: misfit, line-search misfits and the linear step sums
    recomputed from independent forward solves.
  * Adjoint injection: the residual overwrites the adjoint velocity and the
    last of two coincident traces wins (neither additive nor first-wins).
  * Random boundary (rbc): band velocities within [0, vmax] with the vs / vp
    ratio kept, the interior and density untouched, no damping, a repeatable
    seed.  Its reconstruction without strips is the propagator's inverse
    (pinned above with the PML off); that the reference then skips the strips
    and, as everywhere, leaves the source in the reconstruction is only
    restated (parity unpinned).
  * Imaging: with kept forward fields, additive injection and imaging between
    the adjoint half steps, the imaging step gives the exact gradient of the
    L2 misfit (PML off): central differences converge to it at second order
    for Vp, Vs and density.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sfwi_amd import workloads as W
from tests.helpers import oracle_cfg, rel_l2


def test_fd_coefficients_taylor_moments():
    for prec in ("f64", "f32"):
        c = O.coeffs(prec).astype(np.float64)[1:]
        odd = 2 * np.arange(1, 5) - 1
        tol = 1e-15 if prec == "f64" else 2e-7
        assert abs((c * odd).sum() - 1.0) < tol
        for p in (3, 5, 7):
            assert abs((c * odd.astype(float) ** p).sum()) < tol * 10 ** (p - 1)


def test_stencils_are_staggered_derivatives():
    nxp, nzp = 72, 64
    kx, kz = 2 * np.pi / 37.0, 2 * np.pi / 29.0
    ix = np.arange(nxp)[:, None]
    iz = np.arange(nzp)[None, :]
    f = np.sin(kx * ix + 0.3) * np.ones((1, nzp))
    dxp, dxm, dzp, dzm = O.stencils(f)
    sl = (slice(4, nxp - 4), slice(4, nzp - 4))
    assert np.abs(dxp[sl] - (kx * np.cos(kx * (ix + 0.5) + 0.3) * np.ones((1, nzp)))[sl]).max() < 1e-9
    assert np.abs(dxm[sl] - (kx * np.cos(kx * (ix - 0.5) + 0.3) * np.ones((1, nzp)))[sl]).max() < 1e-9
    assert np.abs(dzp[sl]).max() < 1e-12 and np.abs(dzm[sl]).max() < 1e-12
    g = np.ones((nxp, 1)) * np.sin(kz * iz - 0.7)
    dxp, dxm, dzp, dzm = O.stencils(g)
    assert np.abs(dzp[sl] - (np.ones((nxp, 1)) * kz * np.cos(kz * (iz + 0.5) - 0.7))[sl]).max() < 1e-9
    assert np.abs(dzm[sl] - (np.ones((nxp, 1)) * kz * np.cos(kz * (iz - 0.5) - 0.7))[sl]).max() < 1e-9


def test_ricker_closed_form():
    nt, dt, f0 = 400, 0.001, 15.0
    w = O.ricker(nt, dt, f0)
    t0 = int(np.ceil(1.0 / np.float32(f0 * np.float32(dt))))
    assert w[t0] == 1.0
    k = np.arange(1, t0)
    assert np.array_equal(w[t0 + k], w[t0 - k])
    ref = (1 - 2 * (np.pi * f0 * dt * (np.arange(nt) - t0)) ** 2) * np.exp(-(np.pi * f0 * dt * (np.arange(nt) - t0)) ** 2)
    assert np.abs(w - ref).max() < 1e-6


def _cfg(nx=40, nz=36, pml=8, **kw):
    return O.config(allnx=nx + 5, allnz=nz, nx=nx, nz=nz, pml=pml, nt=10, ntr=1, dx=10.0, dz=10.0,
                    dt=0.001, left=3, isx=pml + 4, isz=pml + 2, igz=pml, ngx_left=30, **kw)


def test_window_padding_and_medium_closed_form():
    rng = np.random.default_rng(0)
    c = _cfg()
    vp = rng.uniform(1800, 3000, (c.allnx, c.allnz)).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    rho = rng.uniform(1900, 2400, (c.allnx, c.allnz)).astype(np.float32)
    for prec in ("f64", "f32"):
        m = O.setup(c, vp, vs, rho, prec)
        nxp, nzp = c.nx + 2 * c.pml, c.nz + 2 * c.pml
        sx = np.clip(np.arange(nxp) - c.pml, 0, c.nx - 1) + c.left
        sz = np.clip(np.arange(nzp) - c.pml, 0, c.nz - 1)
        for k, a in enumerate((vp, vs, rho)):
            assert np.array_equal(m[k], a[np.ix_(sx, sz)].astype(m.dtype))
        P, S, R = (a[np.ix_(sx, sz)].astype(np.float64) for a in (vp, vs, rho))
        assert np.allclose(m[3], R * (P * P - 2 * S * S), rtol=1e-6)   # lambda
        assert np.allclose(m[4], R * S * S, rtol=1e-6)                 # mu
        # PML profile: 3 * 5 vp / (2 pml) * (d / pml)^4, left edge d = pml - ix,
        # right edge d = ix - (nxpml - pml) (strict >), zero in between.
        dx_ = np.zeros((nxp, nzp))
        for ix in range(nxp):
            d = c.pml - ix if ix < c.pml else (ix - (nxp - c.pml) if ix > nxp - c.pml else None)
            if d is not None:
                dx_[ix] = 3.0 * 5.0 * P[ix] / (2.0 * c.pml) * (d / c.pml) ** 4
        assert np.allclose(m[5], dx_, rtol=2e-6, atol=0)
        assert not m[5][c.pml:nxp - c.pml + 1].any()
        assert m[5][0].max() > 0 and m[5][nxp - 1].max() > 0 and not m[5][nxp - c.pml].any()


def _medium_f64(c, sv):
    return O.setup(c, sv.vp, sv.vs, sv.rho, "f64")


def _state(nxp, nzp, rng, adjoint=False, margin=10, vscale=1.0, sscale=1.0):
    """Random smooth fields vanishing near the edges; split parts consistent
    with the totals (x part = total)."""
    X, Z = np.meshgrid(np.arange(nxp), np.arange(nzp), indexing="ij")
    win = ((X >= margin) & (X < nxp - margin) & (Z >= margin) & (Z < nzp - margin)).astype(float)
    f = []
    for _ in range(5):
        a = W.gaussian_smooth(rng.standard_normal((nxp, nzp)).astype(np.float32), 1.5).astype(np.float64)
        f.append(a * win)
    vx, vz, txx, tzz, txz = f
    vx, vz = vx * vscale, vz * vscale
    txx, tzz, txz = txx * sscale, tzz * sscale, txz * sscale
    z = np.zeros_like(vx)
    if adjoint:   # adjoint stresses: txx = txx_x, tzz = tzz_z
        parts = [vx, z, vz, z, txx, z, z, tzz, txz, z]
    else:
        parts = [vx, z, vz, z, txx, z, tzz, z, txz, z]
    return np.stack([vx, vz, txx, tzz, txz] + parts).astype(np.float64)


def _tot(s):
    return s[:5]


def test_energy_is_conserved_without_pml():
    sv = W.tiny(nx=48, nz=40, pml=8, seed=17)
    c = oracle_cfg(sv, sv.shots[0], pml_off=1)
    m = _medium_f64(c, sv)
    rho, lam, mu = m[2], m[3], m[4]
    A = lam + 2 * mu
    det = A * A - lam * lam
    # velocities small against stresses (impedance ~ 5e6)
    s0 = _state(m.shape[1], m.shape[2], np.random.default_rng(3), vscale=1e-6)

    def energy(s, s_next):
        txx, tzz, txz = s[2], s[3], s[4]
        e_s = ((A * txx * txx - 2 * lam * txx * tzz + A * tzz * tzz) / det + txz * txz / mu).sum()
        e_v = (rho * (s[0] * s_next[0] + s[1] * s_next[1])).sum()
        return 0.5 * (e_s + e_v)

    es = []
    s = s0
    for k in range(6):
        s1 = O.propagate(c, sv.vp, sv.vs, sv.rho, s, 0, +1, 1)
        es.append(energy(s, s1))
        s = O.propagate(c, sv.vp, sv.vs, sv.rho, s1, 0, +1, 39)
    es = np.array(es)
    assert np.abs(es - es[0]).max() <= 1e-9 * abs(es[0]), es


def test_reconstruction_inverts_forward_without_pml():
    sv = W.tiny(nx=48, nz=40, pml=8, seed=19)
    c = oracle_cfg(sv, sv.shots[0], pml_off=1)
    nxp, nzp = sv.nx + 2 * sv.pml, sv.allnz + 2 * sv.pml
    s0 = _state(nxp, nzp, np.random.default_rng(5), vscale=1e-6)
    s1 = O.propagate(c, sv.vp, sv.vs, sv.rho, s0, 0, +1, 120)
    assert rel_l2(_tot(s1), _tot(s0)) > 0.1
    back = O.propagate(c, sv.vp, sv.vs, sv.rho, s1, 1, -1, 120)
    assert rel_l2(_tot(back), _tot(s0)) < 1e-10


@pytest.mark.parametrize("fwd_kind,adj_kind", [(3, 5), (4, 6)])
def test_adjoint_half_steps_are_transposes(fwd_kind, adj_kind):
    """A forward half step is F = I + K, K mapping stresses to velocities
    (kind 3) or velocities to stresses (kind 4).  The reference's adjoint
    stress (kind 5) / velocity (kind 6) half step applied to
    (v_adj, s_adj) = (a / rho, -b) must realise I + K^T:
    <K x, y> = <x, K^T y>.  x is zero where K writes and y is zero where K^T
    writes, so both increments are formed without cancellation."""
    sv = W.tiny(nx=48, nz=40, pml=8, seed=23)
    c = oracle_cfg(sv, sv.shots[0], pml_off=1)
    rho = _medium_f64(c, sv)[2]
    nxp, nzp = rho.shape
    rng = np.random.default_rng(7)
    v_out = fwd_kind == 3                     # K writes velocities
    x = _state(nxp, nzp, rng, vscale=0.0 if v_out else 1.0, sscale=1.0 if v_out else 0.0)
    y = _state(nxp, nzp, rng, vscale=1.0 if v_out else 0.0, sscale=0.0 if v_out else 1.0)
    kx = _tot(O.propagate(c, sv.vp, sv.vs, sv.rho, x, fwd_kind, +1, 1)) - _tot(x)
    lhs = float((kx * _tot(y)).sum())
    a_vx, a_vz, b_xx, b_zz, b_xz = _tot(y)
    adj = np.zeros_like(x)
    adj[0], adj[1] = a_vx / rho, a_vz / rho
    adj[2], adj[3], adj[4] = -b_xx, -b_zz, -b_xz
    adj[5], adj[7] = adj[0], adj[1]                         # vx_x, vz_x
    adj[9], adj[12], adj[13] = adj[2], adj[3], adj[4]       # txx_x, tzz_z, txz_x
    fa = O.propagate(c, sv.vp, sv.vs, sv.rho, adj, adj_kind, +1, 1)
    kty = np.stack([rho * fa[0], rho * fa[1], -fa[2], -fa[3], -fa[4]]) - _tot(y)
    rhs = float((_tot(x) * kty).sum())
    scale = float(np.abs(kx * _tot(y)).sum())
    assert abs(lhs) > 1e-3 * scale
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)


def _exact_grid(a, s=64.0):
    return (np.round(np.asarray(a, np.float64) * s) / s).astype(np.float32)


def test_born_is_the_linearization_of_forward_modeling():
    sv = W.tiny(nt=300, seed=13)
    sh = sv.shots[0]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    c = oracle_cfg(sv, sh, rho_const=2000.0)
    vp0, vs0 = _exact_grid(sv.vp0), _exact_grid(sv.vs0)   # so vp0 +- eps*dvp is exact in float32
    X, Z = np.meshgrid(np.arange(sv.allnx), np.arange(sv.allnz), indexing="ij")
    bump = np.exp(-((X - 24) ** 2 + (Z - 22) ** 2) / (2 * 4.0 ** 2))
    dvp, dvs = _exact_grid(64.0 * bump), _exact_grid(-32.0 * bump)
    bz, bx, _, _ = O.born_forward(c, vp0, vs0, sv.rho0, dvp, dvs, w, sh.gc, "f64")
    errs = []
    for eps in (0.125, 0.0625, 0.03125):
        e = np.float32(eps)
        pz, px = O.forward(c, vp0 + e * dvp, vs0 + e * dvs, sv.rho0, w, sh.gc, "f64")
        mz, mx = O.forward(c, vp0 - e * dvp, vs0 - e * dvs, sv.rho0, w, sh.gc, "f64")
        errs.append(max(rel_l2(bz, (pz - mz) / (2 * eps)), rel_l2(bx, (px - mx) / (2 * eps))))
    assert errs[-1] < 5e-5
    for a, b in zip(errs, errs[1:]):
        assert 3.5 < a / b < 4.5, errs      # second-order convergence


def test_misfit_and_line_search_compose_from_forward_solves():
    sv = W.tiny(nt=200, seed=29)
    sh = sv.shots[1]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    c = oracle_cfg(sv, sh)
    oz, ox = O.forward(c, sv.vp, sv.vs, sv.rho, w, sh.gc, "f32")
    sz, sx = O.forward(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, "f32")
    mis = O.gradient(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox, "f32")[3]
    direct = float((((sz - oz).astype(np.float64)) ** 2 + ((sx - ox).astype(np.float64)) ** 2).sum())
    assert abs(mis - direct) <= 1e-6 * direct
    d = np.full(sv.vp.shape, 0.5, np.float32)
    m = O.parabola_misfits(c, sv.vp0, sv.vs0, sv.rho0, d, d, 30.0, 150.0, w, sh.gc, oz, ox)
    for k, step in enumerate((0.0, 30.0, 150.0)):
        tz, tx = O.forward(c, sv.vp0 - np.float32(step) * d, sv.vs0 - np.float32(step) * d, sv.rho0, w,
                           sh.gc, "f32")
        ref = float((((tz - oz).astype(np.float64)) ** 2 + ((tx - ox).astype(np.float64)) ** 2).sum())
        assert abs(m[k] - ref) <= 1e-6 * ref
    assert m[0] == pytest.approx(mis, rel=1e-12)


def test_linear_step_sums_compose_from_born_solves():
    sv = W.tiny(nt=200, seed=31)
    sh = sv.shots[0]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    c = oracle_cfg(sv, sh, rho_const=2000.0, adjoint_dir=-1)
    oz, ox = O.forward(c, sv.vp, sv.vs, sv.rho, w, sh.gc, "f32")
    dvp, dvs = (sv.vp - sv.vp0) * 0.5, (sv.vs - sv.vs0) * 0.5
    cgp, cgs = dvp * 0.3, dvs * -0.2
    fz, fm = O.linear_alpha(c, sv.vp0, sv.vs0, sv.rho0, dvp, dvs, cgp, cgs, w, sh.gc, oz, ox)
    bz, bx, gz, gx = O.born_forward(c, sv.vp0, sv.vs0, sv.rho0, dvp, dvs, w, sh.gc, "f32")
    rz, rx = (bz - oz) + gz, (bx - ox) + gx
    lz, lx, _, _ = O.born_forward(c, sv.vp0, sv.vs0, sv.rho0, cgp, cgs, w, sh.gc, "f32")
    d = lambda a, b: float((a.astype(np.float64) * b.astype(np.float64)).sum())
    assert fz == pytest.approx(d(lz, rz) + d(lx, rx), rel=1e-6)
    assert fm == pytest.approx(d(lz, lz) + d(lx, lx), rel=1e-6)


def test_adjoint_injection_overwrites_and_the_last_coincident_trace_wins():
    """The residual overwrites the adjoint velocity at each receiver cell
    (sfwi-stage1/FD2DGPU.cu:3425-3441): two traces on one cell leave only the
    later one's residual, which an additive injection (r1 + r2) or a
    first-wins order would not reproduce."""
    sv = W.tiny(nt=160, seed=23)
    base = sv.shots[0]
    g0, g1 = int(base.gc[10]), int(base.gc[30])
    r = np.random.default_rng(4)
    res = [W.gaussian_smooth(r.standard_normal((1, sv.nt)).astype(np.float32), 3.0)[0] for _ in range(3)]
    w = W.ricker(sv.nt, sv.dt, sv.f0)

    def grad(gc, rows):
        sh = W.Shot(sx=base.sx, gc=np.array(gc, np.int32), ngx_left=base.ngx_left, left=base.left)
        c = oracle_cfg(sv, sh)
        rz = np.ascontiguousarray(np.stack(rows), np.float32)
        return O.gradient_res(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, rz, rz * np.float32(0.5))

    dup = grad([g0, g1, g1], [res[0], res[1], res[2]])
    last = grad([g0, g1], [res[0], res[2]])
    first = grad([g0, g1], [res[0], res[1]])
    summed = grad([g0, g1], [res[0], res[1] + res[2]])
    for a, b in zip(dup, last):
        assert np.array_equal(a, b)
    assert rel_l2(dup[0], first[0]) > 1e-2 and rel_l2(dup[0], summed[0]) > 1e-2


@pytest.mark.parametrize("param", ["vp", "vs", "rho"])
def test_imaging_formula_is_the_discrete_gradient(param):
    """The imaging step (zero-lag products of the velocity derivatives with the
    adjoint stresses, the chain rule to Vp / Vs through lambda, mu and the
    density term), used with kept forward fields, additive injection and
    imaging between the adjoint half steps (or_gradient_consistent_f64), is
    the exact gradient of the L2 misfit with the PML off: central differences
    of the misfit converge to it at second order.  The faithful gradient uses
    the same imaging step with the reference's orchestration."""
    sv = W.tiny(nx=40, nz=32, pml=8, nt=160, seed=29)
    sh = sv.shots[0]
    c = oracle_cfg(sv, sh, pml_off=1)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    oz, ox = O.forward(c, sv.vp, sv.vs, sv.rho, w, sh.gc, "f32")
    base = [sv.vp0, sv.vs0, sv.rho0]
    k = ["vp", "vs", "rho"].index(param)
    g = O.gradient_consistent(c, *base, w, sh.gc, oz, ox)[k]
    d = W.gaussian_smooth(np.random.default_rng(3).standard_normal(sv.vp0.shape).astype(np.float32), 2.0)
    d = d.astype(np.float64)
    d[0, :] = d[-1, :] = 0                    # edge cells also fill the padding, which is not imaged
    d[:, 0] = d[:, -1] = 0
    errs = []
    for eps in (4e-3, 2e-3, 1e-3):
        m = base[k].astype(np.float64)
        s = eps * np.abs(m).max() / np.abs(d).max()
        mp, mm = (m + s * d).astype(np.float32), (m - s * d).astype(np.float32)
        step = (mp.astype(np.float64) - mm.astype(np.float64)) / 2    # the perturbation actually applied
        jp = O.gradient_consistent(c, *[mp if j == k else base[j] for j in range(3)], w, sh.gc, oz, ox)[3]
        jm = O.gradient_consistent(c, *[mm if j == k else base[j] for j in range(3)], w, sh.gc, oz, ox)[3]
        an = float((g * step).sum())
        errs.append(abs((jp - jm) / 2 - an) / abs(an))
    assert errs[-1] < 5e-4, errs
    assert 3.0 < errs[0] / errs[1] < 5.0 and 3.0 < errs[1] / errs[2] < 5.0, errs     # second order


def test_random_boundary_medium():
    """rbc = 1: every PML-band cell's Vp is redrawn within [0, vmax] (vmax the
    window's largest Vp) with vs / vp kept; the interior, the density and the
    seed's repeatability hold; damping is zero; another seed gives another
    boundary."""
    sv = W.tiny(nt=10, seed=4)
    c0 = _cfg(nx=40, nz=36, pml=8)
    vp, vs, rho = sv.vp[:45, :36].copy(), sv.vs[:45, :36].copy(), sv.rho[:45, :36].copy()
    base = O.setup(c0, vp, vs, rho)
    a = O.setup(_cfg(nx=40, nz=36, pml=8, rbc=1, rbc_seed=11), vp, vs, rho)
    b = O.setup(_cfg(nx=40, nz=36, pml=8, rbc=1, rbc_seed=11), vp, vs, rho)
    d = O.setup(_cfg(nx=40, nz=36, pml=8, rbc=1, rbc_seed=12), vp, vs, rho)
    assert np.array_equal(a, b) and not np.array_equal(a[0], d[0])
    p = 8
    band = np.ones(a[0].shape, bool)
    band[p:-p, p:-p] = False
    inner = ~band
    assert np.array_equal(a[0][inner], base[0][inner]) and np.array_equal(a[1][inner], base[1][inner])
    assert np.array_equal(a[2], base[2])
    assert not a[5].any() and not a[6].any()
    vmax = base[0].max()
    assert a[0][band].min() >= 0 and a[0][band].max() <= vmax
    assert np.mean(a[0][band] != base[0][band]) > 0.95          # redrawn, not copied
    ratio0 = base[1][band] / base[0][band]
    assert np.allclose(a[1][band] / a[0][band], ratio0, rtol=1e-5)
    # the Lame parameters follow the new velocities
    assert np.allclose(a[4], a[2] * a[1] ** 2, rtol=1e-6)
    # the perturbation grows with the depth into the band: outer rows spread more
    dev = np.abs(a[0] - base[0])
    assert dev[p:-p, 0].mean() > dev[p:-p, p - 1].mean()

