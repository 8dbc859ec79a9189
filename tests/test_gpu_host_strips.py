"""Boundary strips in pinned host memory (sfwi_params.cpu_mem, the reference's
cpu_mem, sfwi-stage1/FD2DGPU.cu:4393-4421, 4545-4560, 4602-4620): the device
keeps two chunks of strip_chunk steps per shot and the other chunks move over
PCIe on a second stream while the kernels run.  Gradients, illumination and
misfits equal those of strips held in HBM bit for bit, for a chunk length that
leaves a partial last chunk (50 of 240 steps), one step (a ring of two steps),
an odd length (7) and half the record (120: every step stays on the device),
for FP32 and 16-bit strips, both FAST kernel families, EXACT mode, a batch of
three shots with two solves in a row, and graph capture requested (host strips
run without a graph).  The reference's own chunk copy skips chunk 0 when
ntt = 1 (its copy condition excludes it = 0); here every chunk is restored.
"""
import numpy as np
import pytest

from sfwi_amd import workloads as W
from tests.helpers import context, lib_shot

pytestmark = pytest.mark.gpu

CASES = [("fast", 8, "f32", 50), ("fast", 8, "f32", 1), ("fast", 8, "f32", 7), ("fast", 8, "f32", 120),
         ("fast", 8, "bf16", 50), ("fast", 8, "f16", 7), ("fast", -1, "f32", 7), ("exact", 0, "f32", 50)]


def _survey():
    sv = W.tiny(nt=240, nshots=3, seed=12)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    with context(sv, max_batch=3) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp, sv.vs, sv.rho)
        oz, ox = ctx.batch_forward([lib_shot(sv, s) for s in sv.shots])
    shots = [lib_shot(sv, s, oz[k][:len(s.gc)], ox[k][:len(s.gc)]) for k, s in enumerate(sv.shots)]
    return sv, w, shots


def _solve(sv, w, shots, **kw):
    out = []
    with context(sv, max_batch=3, **kw) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        for _ in range(2):
            m = ctx.batch_gradient(shots)
            out.append((m, ctx.flush()))
    return out


@pytest.fixture(scope="module")
def survey():
    return _survey()


@pytest.mark.parametrize("mode,cols,storage,chunk", CASES)
def test_host_strips_equal_hbm_strips(survey, mode, cols, storage, chunk):
    sv, w, shots = survey
    ref = _solve(sv, w, shots, mode=mode, march_cols=cols, storage=storage)
    got = _solve(sv, w, shots, mode=mode, march_cols=cols, storage=storage, cpu_mem=True, strip_chunk=chunk)
    assert np.abs(ref[0][1][0]).max() > 0
    for (m0, g0), (m1, g1) in zip(ref, got):
        assert np.array_equal(m0, m1)
        for a, b in zip(g0, g1):
            assert np.array_equal(a, b), "max |d| = %g" % np.abs(a - b).max()
    for a, b in zip(ref[0][1], ref[1][1]):      # the two solves in a row agree too
        assert np.array_equal(a, b)


def test_host_strips_with_graph_requested(survey):
    sv, w, shots = survey
    ref = _solve(sv, w, shots)
    got = _solve(sv, w, shots, use_graph=True, cpu_mem=True, strip_chunk=50)
    for (m0, g0), (m1, g1) in zip(ref, got):
        assert np.array_equal(m0, m1)
        for a, b in zip(g0, g1):
            assert np.array_equal(a, b)
