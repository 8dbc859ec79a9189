"""Failure handling, forced from the host (every kernel runs to its normal
end): a model or wavelet the propagator cannot use is refused before it
reaches the device and the context keeps its previous one; a shot whose
observed traces hold NaN gives a non-finite misfit, reported as
SFWI_ENUMERIC naming the shot, while every other shot's misfit is the clean
one and the context is clean again after a flush; in the serving loop such a
request fails alone -- the requests batched with it equal their results
alone -- a bad request is refused at submit, a queued request can be
cancelled, and a wait can be bounded in time."""
import numpy as np
import pytest

from sfwi_amd import _lib
from sfwi_amd import workloads as W
from sfwi_amd.elastic import Server, SfwiError
from tests.helpers import context, lib_shot
from tests.test_gpu_server import _alone, _same, _same_shape_requests, _submit

pytestmark = pytest.mark.gpu


def _case():
    sv = W.tiny(nx=48, nz=40, nt=200, seed=71, nshots=3)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    with context(sv, max_batch=3) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp, sv.vs, sv.rho)
        oz, ox = ctx.batch_forward([lib_shot(sv, s) for s in sv.shots])
    shots = [lib_shot(sv, s, oz[k][:len(s.gc)].copy(), ox[k][:len(s.gc)].copy()) for k, s in enumerate(sv.shots)]
    return sv, w, shots


def _gradient(ctx, shots):
    m = ctx.batch_gradient(shots)
    return m, ctx.flush()


def _eq(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_bad_model_and_wavelet_are_refused(mode):
    sv, w, shots = _case()
    with context(sv, mode, max_batch=3) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        ref = _gradient(ctx, shots)
        bad = sv.vp0.copy()
        bad[5, 7] = np.nan
        with pytest.raises(SfwiError) as e:
            ctx.set_model(bad, sv.vs0, sv.rho0)
        assert e.value.code == _lib.SFWI_EINVAL and "(ix 5, iz 7)" in str(e.value)
        neg = sv.vs0.copy()
        neg[0, 3] = -1.0
        with pytest.raises(SfwiError, match="vs = -1"):
            ctx.set_model(sv.vp0, neg, sv.rho0)
        wb = w.copy()
        wb[3] = np.inf
        with pytest.raises(SfwiError, match="wavelet sample 3"):
            ctx.set_wavelet(wb)
        pb = np.zeros_like(sv.vp0)
        pb[1, 1] = np.nan
        with pytest.raises(SfwiError, match="not finite"):
            ctx.set_perturbation(pb, pb)
        again = _gradient(ctx, shots)   # the previous model and wavelet are still in place
    assert np.array_equal(again[0], ref[0])
    _eq(again[1], ref[1])


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_nan_observed_traces_are_reported(mode):
    sv, w, shots = _case()
    with context(sv, mode, max_batch=3) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        ref = _gradient(ctx, shots)
        bad = _case()[2]
        bad[1].obs_z[4, 50] = np.nan
        with pytest.raises(SfwiError) as e:
            ctx.batch_gradient(bad)
        assert e.value.code == _lib.SFWI_ENUMERIC and "shot 1" in str(e.value)
        m = e.value.misfits
        assert np.isnan(m[1]) and m[0] == ref[0][0] and m[2] == ref[0][2]
        assert np.isnan(ctx.flush()[0]).any()   # the images held it until the flush
        again = _gradient(ctx, shots)           # clean again afterwards
    assert np.array_equal(again[0], ref[0])
    _eq(again[1], ref[1])


@pytest.mark.parametrize("batching", [True, False])
def test_a_failing_request_does_not_poison_its_batch(batching):
    reqs = _same_shape_requests("gradient", 4)
    alone = [_alone(r) for r in reqs]
    reqs[2]["shots"][0].obs_x[3, 20] = np.nan
    with Server(workers=1, batching=batching) as srv:
        tickets = [_submit(srv, r) for r in reqs]
        outs = []
        for k, t in enumerate(tickets):
            if k == 2:
                with pytest.raises(SfwiError) as e:
                    srv.wait(t)
                assert e.value.code == _lib.SFWI_ENUMERIC and "shot 0" in str(e.value)
                assert np.isnan(e.value.outputs["misfits"][0])
            else:
                outs.append((k, srv.wait(t)))
        # the worker's cached context is clean for the next request
        after = srv.wait(_submit(srv, reqs[0]))
    for k, o in outs:
        _same(o, alone[k])
    _same(after, alone[0])


def test_bad_request_is_refused_at_submit():
    reqs = _same_shape_requests("gradient", 1)
    r = dict(reqs[0])
    r["vs"] = r["vs"].copy()
    r["vs"][2, 2] = np.inf
    with Server(workers=1) as srv:
        with pytest.raises(SfwiError) as e:
            _submit(srv, r)
        assert e.value.code == _lib.SFWI_EINVAL and "(ix 2, iz 2)" in str(e.value)
        _same(srv.wait(_submit(srv, reqs[0])), _alone(reqs[0]))


def test_cancel_and_bounded_wait():
    reqs = _same_shape_requests("gradient", 4)
    alone = [_alone(r) for r in reqs]
    with Server(workers=1, batching=False) as srv:
        tickets = [_submit(srv, r) for r in reqs]
        srv.cancel(tickets[3])                   # still queued behind the first three
        with pytest.raises(SfwiError) as e:
            srv.cancel(tickets[3])
        assert e.value.code == _lib.SFWI_ESTATE
        with pytest.raises(SfwiError) as e:
            srv.wait(tickets[2], timeout=0.0)   # queued or running: not done at once
        assert e.value.code == _lib.SFWI_ETIMEOUT
        outs = [srv.wait(t, timeout=120.0) for t in tickets[:3]]
        with pytest.raises(SfwiError) as e:
            srv.wait(tickets[3])
        assert e.value.code == _lib.SFWI_ECANCELED
    for o, a in zip(outs, alone):
        _same(o, a)
