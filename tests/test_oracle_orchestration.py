"""Pins of the faithful gradient's orchestration in the oracle
(sfwi-stage1/FD2DGPU.cu:4321-4695), each by a property the mathematics fixes
and each with a planted mistake (or_config.plant, a test hook of the oracle)
that must fail it.  The GPU's EXACT mode reproduces the oracle bit for bit
(tests/test_gpu_parity.py, tests/test_gpu_configs.py), so these pins carry
over to it.

  * Strips (A13, quirk Q14; FD2DGPU.cu:3842-3899): from a source-free state
    with the PML on, in FP64, the reverse run that restores the 8-cell strips
    around the PML interface every step returns the forward field to
    round-off at every step -- the stresses on the whole imaged interior, the
    velocities from pml + 4 inwards -- although the damped PML itself is not
    reversible and the strip corners are not saved.  With velocities restored
    on rows [a, a + 8) the stresses are exact from row a + 4 on, and the
    update rows must be undamped, so a = pml - 4 is the only window that
    works: a window one cell off, inwards or outwards, fails.
  * Source (A5; FD2DGPU.cu:2104-2114): the explosive source enters txx and tzz
    before the velocity update of its step, unscaled, so the first recorded
    sample is the closed form dt / (rho dx) c_k w[0] on the columns around the
    source (and -dt / (rho dz) c_1 w[0] for vz at the source).  A source added
    after the velocity update fails.
  * No source subtraction (quirk Q5; FD2DGPU.cu:4600-4640): the source enters
    the stress totals before a forward velocity update and the stress update
    rebuilds the totals from split parts that never held it, so it acts on
    that velocity update alone; the exact inverse lets each reverse velocity
    update see it again.  With the PML off that exact inverse returns the
    zero starting state (time reversal), while the faithful reconstruction
    minus the exact one equals the source alone run through the same reverse
    steps (zero state, zero strips, negated source) -- linearity -- and that
    difference is not zero.  A reconstruction that undoes the source fails.
    With the random boundary (no strips, no damping) the same holds on the
    whole grid.
  * Imaging position (FD2DGPU.cu:4600-4640): with the reconstruction exact
    and the residual added (plants 8, 128), the faithful loop's images equal
    the FD-pinned consistent gradient of the wavelet advanced by one sample;
    imaging before the adjoint stress-like step (plant 256) matches the
    unshifted wavelet instead.  Also on the random-boundary medium.
  * Stage-2 density override (sfwi-stage2/FD2DGPU.cu:1106-1113): equal to a
    constant density file; an override after lambda and mu (plant 64) fails.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sfwi_amd import workloads as W
from tests.helpers import oracle_cfg, rel_l2
from tests.test_oracle_pins import _state


def _strip_case(plant=0):
    sv = W.tiny(nx=48, nz=40, pml=8, nt=300, seed=31)
    c = oracle_cfg(sv, sv.shots[0], plant=plant)
    nxp, nzp = sv.nx + 2 * sv.pml, sv.allnz + 2 * sv.pml
    s0 = _state(nxp, nzp, np.random.default_rng(8), vscale=1e-6)
    w0 = np.zeros(sv.nt, np.float32)
    return sv, c, s0, w0


def test_strip_restore_returns_the_forward_interior_every_step():
    sv, c, s0, w0 = _strip_case()
    out, e = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w0, 0, init=s0, errs=True)
    assert e.max() <= 1e-10, e.max()
    # the final state is the initial one on the checked cells
    assert rel_l2(out[:2, 12:-12, 12:-12], s0[:2, 12:-12, 12:-12]) <= 1e-10
    assert rel_l2(out[2:5, 8:-8, 8:-8], s0[2:5, 8:-8, 8:-8]) <= 1e-10


def test_strip_window_one_cell_off_fails():
    for plant in (1, 2):   # up strips one row deeper; left strips one column further in
        sv, c, s0, w0 = _strip_case(plant)
        _, e = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w0, 0, init=s0, errs=True)
        assert e.max() > 1e-6, (plant, e.max())


def _source_case(plant=0):
    n, pml = 40, 8
    vp = np.full((n, n), 2500.0, np.float32)
    vs = np.full((n, n), 1400.0, np.float32)
    rho = np.full((n, n), 2100.0, np.float32)
    isx = isz = pml + 16
    gc = (np.arange(10, 23) * 10).astype(np.int32)            # columns pml + 10 .. pml + 22 on the source row
    c = O.config(allnx=n, allnz=n, nx=n, nz=n, pml=pml, nt=3, ntr=len(gc), dx=10.0, dz=10.0, dt=0.001, left=0,
                 isx=isx, isz=isz, igz=isz, ngx_left=0, pml_off=1, plant=plant)
    wav = np.array([1.0, 0.5, -0.25], np.float32)
    return c, vp, vs, rho, wav, gc


def test_source_enters_before_the_velocity_update_unscaled():
    c, vp, vs, rho, wav, gc = _source_case()
    rz, rx = O.forward(c, vp, vs, rho, wav, gc, "f64")
    ck = O.coeffs("f64")
    cols = c.pml + gc // 10
    ex = np.zeros(len(gc))
    ez = np.zeros(len(gc))
    a = float(np.float32(0.001)) / 10.0 / 2100.0 * float(wav[0])   # dt arrives as the FP32 value
    for k in range(1, 5):
        ex[cols == c.isx - k] = a * ck[k]          # Dx+ txx at the source's left
        ex[cols == c.isx + k - 1] = -a * ck[k]     # and right
    ez[cols == c.isx] = -a * ck[1]                 # Dz+ tzz at the source row
    assert np.abs(ex).max() > 0
    np.testing.assert_allclose(rx[:, 0], ex, rtol=1e-12, atol=1e-12 * a)
    np.testing.assert_allclose(rz[:, 0], ez, rtol=1e-12, atol=1e-12 * a)


def test_source_after_the_velocity_update_fails():
    c, vp, vs, rho, wav, gc = _source_case(plant=4)
    rz, rx = O.forward(c, vp, vs, rho, wav, gc, "f64")
    assert not np.abs(rx[:, 0]).any() and not np.abs(rz[:, 0]).any()   # the closed form above is missed


def _q5_case(pml_off=0, plant=0, **kw):
    sv = W.tiny(nx=48, nz=40, pml=8, nt=200, seed=33)
    sh = sv.shots[0]
    c = oracle_cfg(sv, sh, pml_off=pml_off, plant=plant, **kw)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    return sv, c, w


def test_reconstruction_keeps_the_source():
    sv, c, w = _q5_case()
    F, _ = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w, 0)
    S, _ = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w, 1)
    A, _ = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w, 2)
    assert np.abs(A).max() > 0
    assert rel_l2(F - S, A) <= 1e-10, rel_l2(F - S, A)
    # with the PML off the subtracting reconstruction undoes the forward solve
    sv, c, w = _q5_case(pml_off=1)
    S0, _ = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w, 1)
    A0, _ = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w, 2)
    assert np.abs(S0).max() <= 1e-9 * np.abs(A0).max()


def test_random_boundary_reconstruction_by_reversal_alone():
    """With the random boundary the reverse run restores no strips
    (FD2DGPU.cu:4543, 4602) and the band has no damping: the exact inverse
    returns the zero start on the whole grid, and the faithful reconstruction
    minus it is the source alone run through the same reverse steps."""
    sv, c, w = _q5_case(rbc=1, rbc_seed=5)
    F, _ = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w, 0)
    S, _ = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w, 1)
    A, _ = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w, 2)
    assert np.abs(A).max() > 0
    assert np.abs(S[:5]).max() <= 1e-9 * np.abs(A[:5]).max()   # the five totals, every cell
    assert rel_l2(F - S, A) <= 1e-10, rel_l2(F - S, A)


def test_reconstruction_that_subtracts_the_source_fails():
    sv, c, w = _q5_case(plant=8)
    F, _ = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w, 0)
    S, _ = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w, 1)
    A, _ = O.reconstruct(c, sv.vp, sv.vs, sv.rho, w, 2)
    assert rel_l2(F - S, A) > 0.5


def _rho_case(rho_const, plant=0, const_file=False):
    sv = W.tiny(nx=48, nz=40, nt=240, seed=33)
    sh = sv.shots[0]
    c = oracle_cfg(sv, sh, rho_const=rho_const, adjoint_dir=-1, plant=plant)
    rho = np.full_like(sv.rho, 2000.0) if const_file else sv.rho
    dvp, dvs = (sv.vp - sv.vp0) * 0.5, (sv.vs - sv.vs0) * 0.5
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    return sv, sh, c, rho, dvp, dvs, w


def test_density_override_equals_a_constant_density_file():
    """Stage 2's rho := 2000 (sfwi-stage2/FD2DGPU.cu:1106-1113) replaces the
    density on every padded cell before lambda and mu are formed, so a run
    with it equals, bit for bit, the run whose density file holds 2000
    everywhere -- background traces and Born traces, FP32 and FP64.  An
    override formed after lambda and mu (plant 64) fails."""
    assert not np.all(_rho_case(0.0)[3] == 2000.0)   # the file's own density differs
    for prec in ("f32", "f64"):
        outs = []
        for rho_const, const_file, plant in ((2000.0, False, 0), (0.0, True, 0), (2000.0, False, 64)):
            sv, sh, c, rho, dvp, dvs, w = _rho_case(rho_const, plant, const_file)
            outs.append(O.born_forward(c, sv.vp, sv.vs, rho, dvp, dvs, w, sh.gc, prec))
        for a, b in zip(outs[0], outs[1]):
            assert np.array_equal(a, b), prec
        assert not all(np.array_equal(a, b) for a, b in zip(outs[2], outs[1])), prec


def _imaging_case(seed, **kw):
    sv = W.tiny(nx=48, nz=40, pml=8, nt=200, seed=seed)
    sh = sv.shots[0]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    w[0] = 0.0                                   # so v(1) = 0
    w_adv = np.append(w[1:], np.float32(0.0))    # the wavelet one sample earlier
    rng = np.random.default_rng(seed)
    rz = rng.standard_normal((len(sh.gc), sv.nt)).astype(np.float32)
    rx = rng.standard_normal((len(sh.gc), sv.nt)).astype(np.float32)
    zero = np.zeros_like(rz)
    base = dict(rho_const=2048.0, adjoint_dir=1, **kw)   # 2048: r * rho and its quotient are exact

    def faithful(plant):
        c = oracle_cfg(sv, sh, plant=plant, **base)
        return O.gradient_res(c, sv.vp, sv.vs, sv.rho, w, sh.gc, rz, rx, "f64")[:2]

    def consistent(wav):
        c = oracle_cfg(sv, sh, snap_every=1, **base)
        return O.gradient_snapshot(c, sv.vp, sv.vs, sv.rho, wav, sh.gc, zero, zero,
                                   adj_z=rz * 2048.0, adj_x=rx * 2048.0)[:2]

    return float(np.float32(sv.dt)), faithful, consistent, w, w_adv


@pytest.mark.parametrize("kw", [dict(pml_off=1), dict(rbc=1, rbc_seed=3)])
def test_faithful_imaging_position(kw):
    """The faithful reverse loop (FD2DGPU.cu:4600-4640) images after the
    adjoint stress-like step, with the derivatives of v(it + 1) from the
    reconstruction's stress step.  With the reconstruction made exact (plant
    8, no damping: the PML off, or the random boundary, whose reverse run uses
    no strips) and the residual added instead of overwritten (plant 128), the
    adjoint states it images are those of the consistent gradient (FD-pinned,
    K = 1) one iteration later -- the adjoint velocity step is the identity on
    a stress-free field -- paired with forward derivatives one step later:
    so, with w[0] = 0, dt times its Vp / Vs images equal, to FP64 round-off,
    the consistent gradient of the wavelet advanced by one sample (adjoint
    source r rho for an injection of r).  Imaging before the adjoint step
    (plant 256) instead equals the consistent gradient of the wavelet itself."""
    dt, faithful, consistent, w, w_adv = _imaging_case(37, **kw)
    f = faithful(8 | 128)
    e = faithful(8 | 128 | 256)
    adv, same = consistent(w_adv), consistent(w)
    for k in range(2):
        assert np.abs(same[k]).max() > 0
        assert rel_l2(dt * f[k], adv[k]) <= 1e-9, (k, rel_l2(dt * f[k], adv[k]))
        assert rel_l2(dt * e[k], same[k]) <= 1e-9, (k, rel_l2(dt * e[k], same[k]))
        assert rel_l2(dt * f[k], same[k]) > 1e-3 and rel_l2(dt * e[k], adv[k]) > 1e-3
