"""Full-size workloads, checked through properties that do not need the
oracle at that size:

  * the 401 x 301 shot batch (nt 3000, 30-cell PML), 3 of its shots: the FAST
    batched path with the column-march kernels (forced: 3 shots are below the
    automatic family switch, which would pick the per-cell kernels) against
    the EXACT path (the reference's operation order, bit-identical to the
    oracle at small sizes) within the FAST bar (gradients and misfit 1e-4),
    the per-cell FAST kernels equal to the column march bit for bit, and a
    batch equal to its shots run one at a time, bit for bit;
  * the 1200 x 600 mixed-precision grid (nt 6000, 40-cell PML), 2 shots, with
    the column-march kernels: the 16-bit boundary-strip storage within the
    accuracy bar (gradient 1e-3) of the FP32 strips, which are within 2e-6 of
    FP64 on this grid;
  * the 401 x 301 batch with each optimal-transport misfit: the gradient
    completes with finite images and its per-shot misfits equal the
    standalone OT misfit of the same traces;
  * the 401 x 301 batch's stage-2 Born gradient and density image: FAST
    within its bar of EXACT, batch = alone bitwise.
"""
import numpy as np
import pytest

from sfwi_amd import workloads as W
from sfwi_amd.elastic import Context, Shot
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _ctx(sv, n, **kw):
    return Context(dx=sv.dx, dz=sv.dz, dt=sv.dt, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt, ntr=sv.ntr,
                   allnx=sv.allnx, allnz=sv.allnz, max_batch=n, **kw)


def _shots(sv, recs=None):
    out = []
    for k, s in enumerate(sv.shots):
        isx, isz, igz = W.shot_geometry(sv, s)
        oz, ox = (None, None) if recs is None else (recs[0][k][:len(s.gc)], recs[1][k][:len(s.gc)])
        out.append(Shot(s.left, isx, isz, igz, s.ngx_left, s.gc, oz, ox))
    return out


def _records(sv, w):
    with _ctx(sv, len(sv.shots)) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp, sv.vs, sv.rho)
        return ctx.batch_forward(_shots(sv))


def test_shot_batch_full_size():
    sv = W.shot_batch(nshots=3)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    shots = _shots(sv, _records(sv, w))
    res = {}
    for mode, cols in (("exact", 0), ("fast", 8), ("percell", -1)):
        with _ctx(sv, 3, mode="exact" if mode == "exact" else "fast", march_cols=cols) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            m = ctx.batch_gradient(shots)
            g = ctx.flush()
            if mode == "fast":
                alone_m, acc = [], None
                for s in shots:
                    alone_m.append(ctx.batch_gradient([s])[0])
                    a = ctx.flush()
                    acc = [x.copy() for x in a] if acc is None else [x + y for x, y in zip(acc, a)]
                assert np.array_equal(m, np.array(alone_m))
                for x, y in zip(g, acc):
                    assert np.array_equal(x, y)
        res[mode] = (m, g)
    (me, ge), (mf, gf) = res["exact"], res["fast"]
    assert np.abs(ge[0]).max() > 0
    for a, b in zip(gf, ge):
        assert rel_l2(a, b) <= 1e-4
    assert np.all(np.abs(mf - me) <= 1e-4 * np.abs(me))
    assert np.array_equal(res["percell"][0], mf)
    for a, b in zip(res["percell"][1], gf):
        assert np.array_equal(a, b)


def test_mixed_precision_full_size_strip_storage():
    sv = W.mixed_precision(nshots=2)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    shots = _shots(sv, _records(sv, w))
    out = {}
    for st in ("f32", "f16", "bf16"):
        with _ctx(sv, 2, storage=st, march_cols=8) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            m = ctx.batch_gradient(shots)
            out[st] = (m, ctx.flush())
    ref = out["f32"][1]
    for st, bar in (("f16", 1e-3), ("bf16", 1e-3)):
        g = out[st][1]
        assert rel_l2(g[0], ref[0]) <= bar and rel_l2(g[1], ref[1]) <= bar, st
        assert np.array_equal(out[st][0], out["f32"][0])     # the forward pass and its misfit do not use strips


@pytest.mark.parametrize("misfit,variant", [("ot_trace", 0), ("ot_power", 1), ("ot_gather", 2)])
def test_shot_batch_full_size_ot_misfits(misfit, variant):
    """The 401 x 301 shot batch (nt 3000), 3 shots, with each optimal-transport
    misfit: the batched gradient runs to the end on these full-length traces
    (whose signal has ended before the record in some traces), gives finite,
    nonzero images, and its per-shot misfits equal the standalone OT misfit of
    the same synthetic and observed traces."""
    sv = W.shot_batch(nshots=3)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    obs = _records(sv, w)
    shots = _shots(sv, obs)
    with _ctx(sv, 3, misfit=misfit) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        m = ctx.batch_gradient(shots)
        g = ctx.flush()
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        syn = ctx.batch_forward(_shots(sv))
        for k, s in enumerate(sv.shots):
            n = len(s.gc)
            sm, _, _ = ctx.ot_misfit(obs[1][k][:n], obs[0][k][:n], syn[1][k][:n], syn[0][k][:n],
                                     float(np.float32(sv.dt)), variant)
            assert sm == pytest.approx(m[k], rel=1e-12)
    for a in g:
        assert np.all(np.isfinite(a))
    assert np.abs(g[0]).max() > 0 and np.abs(g[1]).max() > 0


def test_shot_batch_full_size_born_gradient():
    """The stage-2 Born gradient on the 401 x 301 batch (nt 3000), 3 shots,
    rho := 2000 and adjoint direction -1 as in stage 2, with a smooth model
    perturbation: FAST within the FAST bar (1e-4) of EXACT, and the batch equal
    to its shots run one at a time, bit for bit."""
    sv = W.shot_batch(nshots=3)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    kw = dict(rho_const=2000.0, adjoint_dir=-1)
    with _ctx(sv, 3, **kw) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp, sv.vs, sv.rho)
        obs = ctx.batch_forward(_shots(sv))
    shots = _shots(sv, obs)
    dvp = (0.5 * (sv.vp - sv.vp0)).astype(np.float32)
    dvs = (0.5 * (sv.vs - sv.vs0)).astype(np.float32)
    res = {}
    for mode in ("exact", "fast"):
        with _ctx(sv, 3, mode=mode, **kw) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            ctx.set_perturbation(dvp, dvs)
            m = ctx.batch_gradient(shots, born=True)
            g = ctx.flush()
            if mode == "fast":
                alone = [ctx.batch_gradient([s], born=True)[0] for s in shots]
                g1 = ctx.flush()
                assert np.array_equal(m, np.array(alone))
                for a, b in zip(g, g1):
                    assert np.array_equal(a, b)
        res[mode] = (m, g)
    (me, ge), (mf, gf) = res["exact"], res["fast"]
    assert np.abs(ge[0]).max() > 0
    for a, b in zip(gf, ge):
        assert rel_l2(a, b) <= 1e-4
    assert np.all(np.abs(mf - me) <= 1e-4 * np.abs(me))


def test_shot_batch_full_size_density_gradient():
    """The density image on the 401 x 301 batch (nt 3000), 3 shots: FAST (the
    column-march path with stored derivatives) within 1e-3 of EXACT for the
    density and 1e-4 for Vp / Vs, and the batch equal to its shots run one at
    a time, bit for bit."""
    sv = W.shot_batch(nshots=3)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    shots = _shots(sv, _records(sv, w))
    res = {}
    for mode in ("exact", "fast"):
        with _ctx(sv, 3, mode=mode, grad_rho=True) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            ctx.batch_gradient(shots)
            g = list(ctx.flush()) + [ctx.flush_rho()]
            if mode == "fast":
                for s in shots:
                    ctx.batch_gradient([s])
                g1 = list(ctx.flush()) + [ctx.flush_rho()]
                for a, b in zip(g, g1):
                    assert np.array_equal(a, b)
        res[mode] = g
    ge, gf = res["exact"], res["fast"]
    assert np.abs(ge[3]).max() > 0
    assert rel_l2(gf[0], ge[0]) <= 1e-4 and rel_l2(gf[1], ge[1]) <= 1e-4
    assert rel_l2(gf[3], ge[3]) <= 1e-3
