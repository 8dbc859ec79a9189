"""Observed traces resident on the device (sfwi_obs_put, sfwi_shot.obs_id):
a solve that reads a shot's stored traces in place gives the results of the
same solve with the traces passed from host memory, bit for bit (gradients,
misfits, the stage-2 linear step sums), and unknown ids or a trace-count
mismatch are errors."""
import numpy as np
import pytest

from oracle import oracle as O
from sfwi_amd import workloads as W
from sfwi_amd._lib import SfwiError
from sfwi_amd.elastic import Shot
from tests.helpers import context, lib_shot, oracle_cfg

pytestmark = pytest.mark.gpu


def _setup(nt=200, nshots=3):
    sv = W.tiny(nt=nt, nshots=nshots, seed=4)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    obs = []
    for sh in sv.shots:
        obs.append(O.forward(oracle_cfg(sv, sh), sv.vp, sv.vs, sv.rho, w, sh.gc, "f32"))
    return sv, w, obs


def _id_shot(sv, sh, k):
    isx, isz, igz = W.shot_geometry(sv, sh)
    return Shot(sh.left, isx, isz, igz, sh.ngx_left, sh.gc, obs_id=k + 1)


@pytest.mark.parametrize("recon", ["strips", "snapshots"])
def test_resident_observed_traces_equal_host_traces(recon):
    sv, w, obs = _setup()
    with context(sv, "fast", max_batch=3, recon=recon, march_cols=8) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        for k in range(3):
            ctx.obs_put(k + 1, *obs[k])
        # the stored sets are found by id in any batch position; the images
        # are merged in batch order, so each order is compared with itself
        for order in ([0, 1, 2], [2, 0, 1]):
            m_host = ctx.batch_gradient([lib_shot(sv, sv.shots[k], *obs[k]) for k in order])
            g_host = ctx.flush()
            m_id = ctx.batch_gradient([_id_shot(sv, sv.shots[k], k) for k in order])
            g_id = ctx.flush()
            assert np.array_equal(m_id, m_host)
            for a, b in zip(g_id, g_host):
                assert np.array_equal(a, b)
        with pytest.raises(SfwiError):
            ctx.batch_gradient([_id_shot(sv, sv.shots[0], 6)])      # nothing stored under id 7
        bad = _id_shot(sv, sv.shots[0], 0)
        bad.c.ntr = bad.c.ntr - 1
        with pytest.raises(SfwiError):
            ctx.batch_gradient([bad])                               # trace count differs
        ctx.obs_clear()
        with pytest.raises(SfwiError):
            ctx.batch_gradient([_id_shot(sv, sv.shots[0], 0)])


def test_resident_observed_traces_in_the_linear_step():
    sv, w, obs = _setup(nshots=2)
    dvp = (0.5 * (sv.vp - sv.vp0)).astype(np.float32)
    dvs = (0.5 * (sv.vs - sv.vs0)).astype(np.float32)
    cg_vp = (sv.vp - sv.vp0).astype(np.float32)
    cg_vs = (sv.vs - sv.vs0).astype(np.float32)
    with context(sv, "fast", max_batch=2, rho_const=2000.0, adjoint_dir=-1) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        ctx.set_perturbation(dvp, dvs)
        host = ctx.batch_linesearch_linear([lib_shot(sv, sh, *obs[k]) for k, sh in enumerate(sv.shots)],
                                           cg_vp, cg_vs)
        for k in range(2):
            ctx.obs_put(k + 1, *obs[k])
        ids = ctx.batch_linesearch_linear([_id_shot(sv, sh, k) for k, sh in enumerate(sv.shots)], cg_vp, cg_vs)
        one = ctx.linesearch_linear(_id_shot(sv, sv.shots[0], 0), cg_vp, cg_vs)
        one_host = ctx.linesearch_linear(lib_shot(sv, sv.shots[0], *obs[0]), cg_vp, cg_vs)
    assert tuple(host) == tuple(ids)
    assert tuple(one) == tuple(one_host)
