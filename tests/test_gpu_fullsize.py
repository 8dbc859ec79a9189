"""Full-size workloads, checked through properties that do not need the
oracle at that size:

  * the 401 x 301 shot batch (nt 3000, 30-cell PML), 3 of its shots: the FAST
    batched column-march path against the EXACT path (the reference's
    operation order, bit-identical to the oracle at small sizes) within the
    FAST bar (gradients and misfit 1e-4), and a batch equal to its shots run
    one at a time, bit for bit;
  * the 1200 x 600 mixed-precision grid (nt 6000, 40-cell PML), 2 shots: the
    16-bit boundary-strip storage within the accuracy bar (gradient 1e-3) of
    the FP32 strips, which are within 2e-6 of FP64 on this grid.
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
    for mode in ("exact", "fast"):
        with _ctx(sv, 3, mode=mode) as ctx:
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


def test_mixed_precision_full_size_strip_storage():
    sv = W.mixed_precision(nshots=2)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    shots = _shots(sv, _records(sv, w))
    out = {}
    for st in ("f32", "f16", "bf16"):
        with _ctx(sv, 2, storage=st) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            m = ctx.batch_gradient(shots)
            out[st] = (m, ctx.flush())
    ref = out["f32"][1]
    for st, bar in (("f16", 1e-3), ("bf16", 1e-3)):
        g = out[st][1]
        assert rel_l2(g[0], ref[0]) <= bar and rel_l2(g[1], ref[1]) <= bar, st
        assert np.array_equal(out[st][0], out["f32"][0])     # the forward pass and its misfit do not use strips
