"""Kernel-family autotuner: it keeps accumulated gradients, picks one of its
candidates, and the context then computes exactly what a context created
with that choice computes."""
import numpy as np
import pytest

from sfwi_amd import workloads as W
from tests.helpers import context, lib_shot

pytestmark = pytest.mark.gpu


def test_autotune_choice_and_preserved_gradients():
    sv = W.tiny(nt=200)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    with context(sv, max_batch=2) as ref:
        ref.set_wavelet(w)
        ref.set_model(sv.vp, sv.vs, sv.rho)
        oz, ox = ref.batch_forward([lib_shot(sv, s) for s in sv.shots])
    shots = [lib_shot(sv, s, oz[k], ox[k]) for k, s in enumerate(sv.shots)]
    with context(sv, max_batch=2) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        m0 = ctx.batch_gradient(shots[:1])            # accumulated before tuning
        cols, depth, order, ms = ctx.autotune(2, steps=60)
        fuse = ctx.params().fuse_step
        assert (cols, depth, order, fuse) in {(-1, 1, 1, 0), (8, 4, 3, 0), (8, 2, 3, 0), (8, 1, 3, 0), (16, 2, 3, 0),
                                              (8, 4, 1, 0), (8, 2, 2, 0), (8, 2, 3, 1), (8, 1, 3, 1)}
        assert ms > 0
        m1 = ctx.batch_gradient(shots)
        g = ctx.flush()
    with context(sv, max_batch=2, march_cols=cols, march_depth=depth, march_order=order, fuse_step=fuse) as chk:
        chk.set_wavelet(w)
        chk.set_model(sv.vp0, sv.vs0, sv.rho0)
        n0 = chk.batch_gradient(shots[:1])
        n1 = chk.batch_gradient(shots)
        h = chk.flush()
    assert np.array_equal(m0, n0) and np.array_equal(m1, n1)
    for a, b in zip(g, h):
        assert np.array_equal(a, b)
