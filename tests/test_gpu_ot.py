"""GPU parity of the optimal-transport misfit (sfwi_ot_misfit and the OT misfit
option of the gradient path) against the numpy FP64 oracle (oracle/ot.py).

Both sides compute in FP64 with different summation orders, so the misfit is
compared at 1e-9 relative and the FP32 adjoint source at 1e-6 of its maximum.
One limit of FP64 applies to the power transform: where the mass still to come
in a trace (1 - F) falls below ~1e-10 of the total, the CDF sits within a few
ulp of 1 and which tail cells read as empty depends on the summation order;
there the adjoint source is compared at 1e-4 of its maximum instead.
The gradient path is compared with the oracle composed as forward -> OT
adjoint source -> reverse loop: in EXACT mode the synthetic traces are
bit-identical, so the gradients agree to 1e-6; in FAST mode to the usual 1e-4.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import ot as OT
from sfwi_amd import workloads as W
from tests.helpers import context, lib_shot, oracle_cfg, rel_l2

pytestmark = pytest.mark.gpu


def _gather(ntr, nt, seed, dead=True):
    rng = np.random.default_rng(seed)
    t = np.arange(nt)
    ox = np.zeros((ntr, nt), np.float32)
    oz = np.zeros((ntr, nt), np.float32)
    for tr in range(ntr):
        for arr, f in ((ox, 0.03), (oz, 0.05)):
            t0 = rng.uniform(0.2, 0.6) * nt
            arr[tr] = np.sin(f * (t - t0)) * np.exp(-((t - t0) / (0.08 * nt + 1)) ** 2)
    sx = np.roll(ox, 3, axis=1) * 0.9 + 0.01 * rng.standard_normal((ntr, nt)).astype(np.float32)
    sz = np.roll(oz, -5, axis=1) * 1.1
    if dead and ntr > 2:
        oz[1] = 0
        sz[1] = 0
    return ox, oz, sx.astype(np.float32), sz.astype(np.float32)


def _resolved(syn, variant):
    """Samples whose remaining (right-tail) mass is resolvable in FP64."""
    if variant != 1:
        return np.ones(syn.shape, bool)
    u = syn.astype(np.float64) ** 2
    tail = np.cumsum(u[:, ::-1], axis=1)[:, ::-1]
    tot = np.maximum(tail[:, :1], 1e-300)
    return tail / tot > 1e-10


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("ntr,nt", [(7, 300), (1, 1), (3, 257), (2, 6000), (5, 40)])
def test_ot_misfit_matches_oracle(variant, ntr, nt):
    dt = float(np.float32(0.002))
    ox, oz, sx, sz = _gather(ntr, nt, seed=ntr * 1000 + nt)
    rx, rz, m = OT.ot_misfit(ox, oz, sx, sz, dt, variant)
    sv = W.tiny()
    with context(sv) as ctx:
        gm, grx, grz = ctx.ot_misfit(ox, oz, sx, sz, dt, variant)
        # twice in a row on the same context: the workspace is reused
        gm2, grx2, grz2 = ctx.ot_misfit(ox, oz, sx, sz, dt, variant)
    assert gm == pytest.approx(m, rel=1e-9, abs=1e-300)
    assert gm2 == gm and np.array_equal(grx2, grx) and np.array_equal(grz2, grz)
    for a, b, syn in ((grx, rx, sx), (grz, rz, sz)):
        scale = max(np.abs(b).max(), 1e-300)
        ok = _resolved(syn, variant)
        assert np.abs(a - b)[ok].max(initial=0.0) <= 1e-6 * scale
        assert np.abs(a - b).max() <= 1e-4 * scale


def test_ot_identical_traces_zero():
    """Zero misfit; an adjoint source negligible against that of a one-sample
    shift (cells whose squared signal underflows keep a one-sided derivative)."""
    ox, oz, _, _ = _gather(4, 500, seed=3)
    with context(W.tiny()) as ctx:
        for variant in (0, 1, 2):
            m, rx, rz = ctx.ot_misfit(ox, oz, ox, oz, 0.001, variant)
            _, sx, sz = ctx.ot_misfit(ox, oz, np.roll(ox, 1, axis=1), np.roll(oz, 1, axis=1), 0.001, variant)
            assert abs(m) <= 1e-20
            assert np.abs(rx).max() <= 1e-6 * np.abs(sx).max()
            assert np.abs(rz).max() <= 1e-6 * np.abs(sz).max()


def test_ot_rejects_bad_arguments():
    ox, oz, sx, sz = _gather(2, 50, seed=1)
    with context(W.tiny()) as ctx:
        with pytest.raises(Exception, match="variant"):
            ctx.ot_misfit(ox, oz, sx, sz, 0.001, 7)
        with pytest.raises(Exception, match="dt"):
            ctx.ot_misfit(ox, oz, sx, sz, 0.0, 0)


@pytest.mark.parametrize("misfit,variant,mode,tol", [
    ("ot_power", 1, "exact", 1e-6), ("ot_trace", 0, "exact", 1e-6), ("ot_gather", 2, "fast", 1e-4)])
def test_gradient_with_ot_misfit(misfit, variant, mode, tol):
    sv = W.tiny(nt=320)
    sh = sv.shots[1]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    c = oracle_cfg(sv, sh)
    oz, ox = O.forward(c, sv.vp, sv.vs, sv.rho, w, sh.gc, "f32")
    sz, sx = O.forward(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, "f32")
    rx, rz, m = OT.ot_misfit(ox, oz, sx, sz, float(np.float32(sv.dt)), variant)
    gvp, gvs, gil = O.gradient_res(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, rz.astype(np.float32),
                                   rx.astype(np.float32))
    with context(sv, mode, misfit=misfit) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        gm = ctx.gradient(lib_shot(sv, sh, oz, ox))
        hvp, hvs, hil = ctx.flush()
    assert gm == pytest.approx(m, rel=1e-9 if mode == "exact" else 1e-4)
    assert rel_l2(hvp, gvp) <= tol and rel_l2(hvs, gvs) <= tol
    assert rel_l2(hil, gil) <= tol


def test_born_gradient_with_ot_misfit():
    sv = W.tiny(nt=320)
    sh = sv.shots[0]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    c = oracle_cfg(sv, sh, rho_const=2000.0, adjoint_dir=-1)
    oz, ox = O.forward(oracle_cfg(sv, sh), sv.vp, sv.vs, sv.rho, w, sh.gc, "f32")
    dvp, dvs = (sv.vp - sv.vp0) * 0.5, (sv.vs - sv.vs0) * 0.5
    bz, bx, gz, gx = O.born_forward(c, sv.vp0, sv.vs0, sv.rho0, dvp, dvs, w, sh.gc, "f32")
    sz, sx = bz + gz, bx + gx                          # FP32 sums, as the library forms them
    rx, rz, m = OT.ot_misfit(ox, oz, sx, sz, float(np.float32(sv.dt)), 1)
    gvp, gvs, gil = O.gradient_res(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, rz.astype(np.float32),
                                   rx.astype(np.float32))
    with context(sv, "exact", misfit="ot_power", rho_const=2000.0, adjoint_dir=-1) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        ctx.set_perturbation(dvp, dvs)
        gm = ctx.born_gradient(lib_shot(sv, sh, oz, ox))
        hvp, hvs, hil = ctx.flush()
    assert gm == pytest.approx(m, rel=1e-9)
    assert rel_l2(hvp, gvp) <= 1e-6 and rel_l2(hvs, gvs) <= 1e-6


def _ended_gather(ntr, nt, seed):
    """Traces whose signal ends before the record does (an exactly zero tail)."""
    rng = np.random.default_rng(seed)
    t = np.arange(nt)
    ox = np.zeros((ntr, nt), np.float32)
    for tr in range(ntr):
        t0 = rng.uniform(0.03, 0.5) * nt
        te = t0 + rng.uniform(0.07, 0.4) * nt
        ox[tr] = np.where((t > t0) & (t < te), np.sin(0.08 * (t - t0)) * np.exp(-(t - t0) / rng.uniform(50, 800)),
                          0.0) * 1e-9
    sx = np.roll(ox, 5, axis=1) * 1.05
    oz = np.roll(ox, -3, axis=1)
    sz = np.roll(oz, 2, axis=1) * 0.95
    return ox, oz, sx.astype(np.float32), sz.astype(np.float32)


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_ot_traces_with_ended_signal(variant):
    """Traces whose energy has died out before the end of the record: with the
    power transform the running cumulative fraction of such a trace can round
    above 1 before its last cell (the mass still to come is below the rounding
    of the total; 60 of these 256 traces do so in the kernel's summation order),
    where the merge walk of the kernel must still end; the fractions are
    clamped to 1.  (The shift transforms add a constant per sample, so their
    fractions stay below 1 - 1/n.)  Misfit and adjoint source against the FP64
    oracle as in test_ot_misfit_matches_oracle; dt is the FP32 value the
    library receives."""
    dt = float(np.float32(0.001))
    ox, oz, sx, sz = _ended_gather(64, 3000, seed=41)
    rx, rz, m = OT.ot_misfit(ox, oz, sx, sz, dt, variant)
    with context(W.tiny()) as ctx:
        gm, grx, grz = ctx.ot_misfit(ox, oz, sx, sz, dt, variant)
    assert gm == pytest.approx(m, rel=1e-9, abs=1e-300)
    for a, b, syn in ((grx, rx, sx), (grz, rz, sz)):
        scale = max(np.abs(b).max(), 1e-300)
        ok = _resolved(syn, variant)
        assert np.abs(a - b)[ok].max(initial=0.0) <= 1e-6 * scale
        assert np.abs(a - b).max() <= 1e-4 * scale
