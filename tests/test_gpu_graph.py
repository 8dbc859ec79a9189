"""hipGraph replay of whole solves (use_graph): every replay equals direct
execution bitwise -- for gradients, forward and Born solves, two batch sizes,
with other solves and host synchronisations between the replays."""
import numpy as np
import pytest

from sfwi_amd import workloads as W
from tests.helpers import context, lib_shot

pytestmark = pytest.mark.gpu


def _setup(ctx, sv, w):
    ctx.set_wavelet(w)
    ctx.set_model(sv.vp0, sv.vs0, sv.rho0)


def test_graph_replays_equal_direct_execution():
    sv = W.tiny(nt=200)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    dvp, dvs = (sv.vp - sv.vp0) * 0.5, (sv.vs - sv.vs0) * 0.5
    with context(sv, max_batch=2) as ref:
        ref.set_wavelet(w)
        ref.set_model(sv.vp, sv.vs, sv.rho)
        oz, ox = ref.batch_forward([lib_shot(sv, s) for s in sv.shots])
        shots = [lib_shot(sv, s, oz[k], ox[k]) for k, s in enumerate(sv.shots)]
        _setup(ref, sv, w)
        m_ref = ref.batch_gradient(shots)
        g_ref = ref.flush()
        m1_ref = ref.batch_gradient(shots[:1])
        g1_ref = ref.flush()
        f_ref = ref.batch_forward(shots)
        ref.set_perturbation(dvp, dvs)
        mb_ref = ref.batch_gradient(shots, born=True)
        gb_ref = ref.flush()
    with context(sv, max_batch=2, use_graph=True) as ctx:
        _setup(ctx, sv, w)
        ctx.set_perturbation(dvp, dvs)
        for rep in range(3):
            m = ctx.batch_gradient(shots)
            g = ctx.flush()
            assert np.array_equal(m, m_ref), rep
            for a, b in zip(g, g_ref):
                assert np.array_equal(a, b), rep
            f = ctx.batch_forward(shots)
            assert np.array_equal(f[0], f_ref[0]) and np.array_equal(f[1], f_ref[1])
            m1 = ctx.batch_gradient(shots[:1])
            g1 = ctx.flush()
            assert np.array_equal(m1, m1_ref)
            for a, b in zip(g1, g1_ref):
                assert np.array_equal(a, b)
            mb = ctx.batch_gradient(shots, born=True)
            gb = ctx.flush()
            assert np.array_equal(mb, mb_ref)
            for a, b in zip(gb, gb_ref):
                assert np.array_equal(a, b)
        # timed solves run directly and still agree
        ctx.set_timing(5)
        m = ctx.batch_gradient(shots)
        g = ctx.flush()
        ctx.set_timing(0)
        assert np.array_equal(m, m_ref)
        for a, b in zip(g, g_ref):
            assert np.array_equal(a, b)
