"""The random boundary (rbc, the reference's alternative to the strips,
sfwi-stage1/FD2DGPU.cu:3245-3305, 4330-4334, 4543, 4602): random Vp / Vs in the
PML bands of every window from the shared project-owned generator, no
damping, nothing saved for the reverse pass.  Against the oracle with the
same seed: EXACT bit-identical traces and gradients (both stages), FAST
within the usual bars; the column-march and per-cell FAST kernels, batched and
one-at-a-time solves agree bit for bit; cpu_mem has no strips to hold and
changes nothing.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sfwi_amd import workloads as W
from tests.helpers import context, lib_shot, oracle_cfg, rel_l2

pytestmark = pytest.mark.gpu

SEED = 77


def _case(nt=240, seed=31):
    sv = W.tiny(nt=nt, seed=seed)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    recs = [O.forward(oracle_cfg(sv, sh), sv.vp, sv.vs, sv.rho, w, sh.gc, "f32") for sh in sv.shots]
    return sv, w, recs


def test_forward_and_gradient_exact_bitwise():
    sv, w, recs = _case()
    sh, (oz, ox) = sv.shots[1], recs[1]
    c = oracle_cfg(sv, sh, rbc=1, rbc_seed=SEED)
    fz, fx = O.forward(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, "f32")
    ref = O.gradient(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox, "f32")
    with context(sv, "exact", rbc=True, rbc_seed=SEED) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        gz, gx = ctx.forward(lib_shot(sv, sh))
        m = ctx.gradient(lib_shot(sv, sh, oz, ox))
        g = ctx.flush()
    assert np.array_equal(gz, fz) and np.array_equal(gx, fx)
    for a, b in zip(g, ref[:3]):
        assert np.abs(b).max() > 0
        assert np.array_equal(a, b), "max |d| = %g" % np.abs(a - b).max()
    assert abs(m - ref[3]) <= 1e-12 * abs(ref[3])
    # the boundary changed the solve: the PML gradient differs
    pml = O.gradient(oracle_cfg(sv, sh), sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox, "f32")
    assert rel_l2(g[0], pml[0]) > 1e-3


def test_gradient_fast_within_tolerance():
    sv, w, recs = _case()
    sh, (oz, ox) = sv.shots[1], recs[1]
    c = oracle_cfg(sv, sh, rbc=1, rbc_seed=SEED)
    ref = O.gradient(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox, "f32")
    ref64 = O.gradient(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox, "f64")
    for a, b in zip(ref[:3], ref64[:3]):
        assert rel_l2(a, b) <= 0.2e-4, "ill-conditioned case"
    with context(sv, "fast", rbc=True, rbc_seed=SEED) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        m = ctx.gradient(lib_shot(sv, sh, oz, ox))
        g = ctx.flush()
    for a, b in zip(g, ref[:3]):
        assert rel_l2(a, b) <= 1e-4
    assert abs(m - ref[3]) <= 1e-4 * abs(ref[3])


def test_born_gradient_exact_bitwise():
    sv, w, recs = _case(nt=200, seed=33)
    sh, (oz, ox) = sv.shots[0], recs[0]
    kw = dict(rho_const=2000.0, adjoint_dir=-1)
    dvp = (0.5 * (sv.vp - sv.vp0)).astype(np.float32)
    dvs = (0.5 * (sv.vs - sv.vs0)).astype(np.float32)
    c = oracle_cfg(sv, sh, rbc=1, rbc_seed=SEED, **kw)
    ref = O.born_gradient(c, sv.vp0, sv.vs0, sv.rho0, dvp, dvs, w, sh.gc, oz, ox, "f32")
    with context(sv, "exact", rbc=True, rbc_seed=SEED, **kw) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, None)
        ctx.set_perturbation(dvp, dvs)
        m = ctx.batch_gradient([lib_shot(sv, sh, oz, ox)], born=True)
        g = ctx.flush()
    for a, b in zip(g, ref[:3]):
        assert np.abs(b).max() > 0
        assert np.array_equal(a, b), "max |d| = %g" % np.abs(a - b).max()
    assert abs(m[0] - ref[3]) <= 1e-12 * abs(ref[3])


def test_fast_families_batches_and_host_strips_agree():
    sv, w, recs = _case(nt=200, seed=35)
    shots = [lib_shot(sv, sh, oz, ox) for sh, (oz, ox) in zip(sv.shots, recs)]
    out = []
    for mc, mb, cpu_mem in ((-1, 1, False), (8, len(shots), False), (8, len(shots), True)):
        with context(sv, "fast", max_batch=mb, march_cols=mc, rbc=True, rbc_seed=SEED, cpu_mem=cpu_mem,
                     strip_chunk=16) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            m = np.concatenate([ctx.batch_gradient(shots[i:i + mb]) for i in range(0, len(shots), mb)])
            out.append((m, ctx.flush()))
    (m0, g0) = out[0]
    assert np.abs(g0[0]).max() > 0
    for m, g in out[1:]:
        assert np.array_equal(m, m0)
        for a, b in zip(g, g0):
            assert np.array_equal(a, b), "max |d| = %g" % np.abs(a - b).max()


def test_line_search_exact():
    sv, w, recs = _case(nt=200, seed=37)
    sh, (oz, ox) = sv.shots[1], recs[1]
    d = np.full(sv.vp.shape, 0.5, np.float32)
    c = oracle_cfg(sv, sh, rbc=1, rbc_seed=SEED)
    ref = O.parabola_misfits(c, sv.vp0, sv.vs0, sv.rho0, d, d, 30.0, 150.0, w, sh.gc, oz, ox)
    with context(sv, "exact", rbc=True, rbc_seed=SEED) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        m = ctx.linesearch_parabola(lib_shot(sv, sh, oz, ox), d, d, 30.0, 150.0)
    assert np.all(np.abs(m - ref) <= 1e-12 * np.abs(ref))
