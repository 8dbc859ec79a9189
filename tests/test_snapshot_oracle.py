"""Pins of the oracle's snapshot gradient (or_gradient_snapshot_f64), the
checker of the library's snapshot reconstruction mode.

  * With the imaging at every step (K = 1) its Vp / Vs images equal those of
    or_gradient_consistent_f64, which central differences of the misfit pin
    (tests/test_oracle_pins.py): same injection, half-step order, velocity
    derivatives and chain rule.
  * Decimation (K > 1) samples the zero-lag correlation every K steps with
    weight K; for a band-limited wavefield sampled well above the correlation's
    Nyquist rate the sum converges to the full-rate one as K falls.
  * Stage 2: the Born misfit |born(dm) + background - obs|^2 is exactly
    quadratic in dm, so a central difference of it is exact (up to rounding)
    for any step: with the PML off, the snapshot Born gradient must equal
    central differences with respect to dvp and dvs.
  * A caller-supplied adjoint source equal to 2 r reproduces the L2 gradient;
    with observed traces equal to the synthetic ones the images vanish.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sfwi_amd import workloads as W
from tests.helpers import oracle_cfg, rel_l2


def _case(f0=25.0, nt=160, pml_off=0, seed=29, **kw):
    sv = W.tiny(nx=40, nz=32, pml=8, nt=nt, seed=seed, f0=f0)
    sh = sv.shots[0]
    c = oracle_cfg(sv, sh, pml_off=pml_off, **kw)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    oz, ox = O.forward(c, sv.vp, sv.vs, sv.rho, w, sh.gc, "f32")
    return sv, sh, c, w, oz, ox


def test_every_step_snapshot_gradient_is_the_consistent_gradient():
    sv, sh, c, w, oz, ox = _case()
    gvp, gvs, _, mis = O.gradient_snapshot(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox)
    cvp, cvs, _, cmis = O.gradient_consistent(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox)
    assert mis == cmis
    assert rel_l2(gvp, cvp) < 1e-13 and rel_l2(gvs, cvs) < 1e-13
    assert np.abs(cvp).max() > 0 and np.abs(cvs).max() > 0


def test_decimated_imaging_converges_to_the_full_rate_sum():
    # f0 = 10 Hz: the correlation's band ends near 50 Hz; K dt <= 4 ms samples it above Nyquist
    sv, sh, c, w, oz, ox = _case(f0=10.0, nt=400)
    full = O.gradient_snapshot(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox)
    errs = []
    for K in (8, 4, 2):
        c.snap_every = K
        g = O.gradient_snapshot(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox)
        errs.append(max(rel_l2(g[0], full[0]), rel_l2(g[1], full[1]), rel_l2(g[2], full[2])))
        assert g[3] == full[3]   # the forward solve and the misfit do not depend on K
    assert errs[1] < 1e-2 and errs[2] < 1e-3, errs
    assert errs[0] > errs[1] > errs[2], errs


@pytest.mark.parametrize("param", ["dvp", "dvs"])
def test_born_snapshot_gradient_is_exact_for_the_quadratic_born_misfit(param):
    sv, sh, c, w, _, _ = _case(pml_off=1, nt=160, rho_const=2000.0)
    rng = np.random.default_rng(11)
    dvp = (0.05 * (sv.vp - sv.vp0)).astype(np.float32)
    dvs = (0.05 * (sv.vs - sv.vs0)).astype(np.float32)
    # observed = background + born of the true perturbation (a stage-2 inversion's data)
    bz, bx, gz, gx = O.born_forward(c, sv.vp0, sv.vs0, sv.rho0, (sv.vp - sv.vp0), (sv.vs - sv.vs0), w, sh.gc, "f64")
    oz, ox = (bz + gz).astype(np.float32), (bx + gx).astype(np.float32)
    gvp, gvs, _, _ = O.gradient_snapshot(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox, dvp=dvp, dvs=dvs)
    g = gvp if param == "dvp" else gvs
    d = W.gaussian_smooth(rng.standard_normal(sv.vp0.shape).astype(np.float32), 2.0).astype(np.float64)
    d[0, :] = d[-1, :] = 0
    d[:, 0] = d[:, -1] = 0
    base = dvp if param == "dvp" else dvs
    for scale in (1.0, 0.25):
        s = scale * np.abs(base).max() / np.abs(d).max()
        mp, mm = (base + s * d).astype(np.float32), (base - s * d).astype(np.float32)
        step = (mp.astype(np.float64) - mm.astype(np.float64)) / 2

        def J(p):
            a = (p, dvs) if param == "dvp" else (dvp, p)
            return O.gradient_snapshot(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox, dvp=a[0], dvs=a[1])[3]
        fd = (J(mp) - J(mm)) / 2
        an = float((g * step).sum())
        assert abs(fd - an) <= 1e-7 * abs(an), (param, scale, fd, an)


def test_supplied_adjoint_source_and_zero_residual():
    sv, sh, c, w, oz, ox = _case()
    c.snap_every = 2
    ref = O.gradient_snapshot(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox)
    sz, sx = O.forward(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, "f64")
    az, ax = (2.0 * (sz - oz)).astype(np.float32), (2.0 * (sx - ox)).astype(np.float32)
    g = O.gradient_snapshot(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox, adj_z=az, adj_x=ax)
    assert rel_l2(g[0], ref[0]) < 1e-6 and rel_l2(g[1], ref[1]) < 1e-6
    assert np.array_equal(g[2], ref[2])   # illumination: the forward field only
    z = O.gradient_snapshot(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, sz.astype(np.float32), sx.astype(np.float32))
    assert np.abs(z[0]).max() < 1e-5 * np.abs(ref[0]).max() and np.abs(z[1]).max() < 1e-5 * np.abs(ref[1]).max()
