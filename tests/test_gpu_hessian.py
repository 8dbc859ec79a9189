"""GPU parity of the Gauss-Newton Hessian-vector product
(sfwi_batch_hessian_vector, ELSRTM_CALCULATE_Hessian_vector of
sfwi-stage1/FD2DGPU.cu:5236-5359), of the double-scattering Born solve
(sfwi_batch_born2_forward, born_modeling mode 5, :1888-1903) and of the
stage-1 inner CG loop (sfwi-stage1/main.cpp:784-1063) built on them, against
the oracle, whose products are pinned in tests/test_hessian_oracle.py.

  * EXACT mode equals the oracle's FP32 restatement bit for bit (products
    summed over the shots in batch order), for vectors inside the interior
    and for vectors that reach the window edge (the stage-1 Born source
    ignores the PML cells);
  * FAST mode (per-cell and column-march kernels) within 1e-4 relative L2;
  * the snapshot (consistent) product with every step imaged, FP32 and no PML
    damping: within 1e-4 of the oracle's FP64 one and symmetric to 1e-4;
  * the inner CG: two Gauss-Newton iterations give the oracle's search
    direction bit for bit in EXACT mode.
"""
import numpy as np
import pytest

from oracle import inversion as I
from oracle import oracle as O
from sfwi_amd import workloads as W
from sfwi_amd.inversion import Stage1
from tests.helpers import context, lib_shot, oracle_cfg, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _case(nt=200, seed=11, nshots=2):
    sv = W.tiny(nx=48, nz=40, pml=8, nt=nt, seed=seed, nshots=nshots)
    return sv, W.ricker(sv.nt, sv.dt, sv.f0)


def _vec(sv, seed, edge=False):
    r = np.random.default_rng(seed)
    X, Z = np.meshgrid(np.arange(sv.nx), np.arange(sv.allnz), indexing="ij")
    a = 50.0 * np.exp(-((X - r.uniform(15, 33)) ** 2 + (Z - r.uniform(14, 26)) ** 2) / 18.0)
    b = 30.0 * np.exp(-((X - r.uniform(15, 33)) ** 2 + (Z - r.uniform(14, 26)) ** 2) / 18.0)
    if edge:   # reaches the left edge of the window
        a = a + 40.0 * np.exp(-((X - 1) ** 2 + (Z - 20) ** 2) / 8.0)
    return a.astype(np.float32), b.astype(np.float32)


def _oracle_hp(sv, w, p, prec="f32", **kw):
    tot = None
    for sh in sv.shots:
        h = O.hessian_vector(oracle_cfg(sv, sh, **kw), sv.vp0, sv.vs0, sv.rho0, *p, w, sh.gc, prec)
        tot = h if tot is None else tuple((a + b).astype(a.dtype) for a, b in zip(tot, h))
    return tot


def _gpu_hp(sv, w, p, mode, **kw):
    with context(sv, mode, max_batch=len(sv.shots), **kw) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        ctx.hessian_vector([lib_shot(sv, sh) for sh in sv.shots], *p)
        return ctx.flush()


@pytest.mark.parametrize("edge", [False, True])
def test_hessian_vector_exact_bitwise(edge):
    sv, w = _case()
    p = _vec(sv, 1, edge)
    want = _oracle_hp(sv, w, p)
    gvp, gvs, gil = _gpu_hp(sv, w, p, "exact")
    assert np.abs(want[0]).max() > 0 and np.abs(want[1]).max() > 0
    assert np.array_equal(gvp, want[0]), "max |d| = %g" % np.abs(gvp - want[0]).max()
    assert np.array_equal(gvs, want[1]), "max |d| = %g" % np.abs(gvs - want[1]).max()
    assert not gil.any()   # no illumination in the product


@pytest.mark.parametrize("cols", [-1, 8])
def test_hessian_vector_fast(cols):
    sv, w = _case()
    p = _vec(sv, 2, edge=True)
    want = _oracle_hp(sv, w, p)
    gvp, gvs, _ = _gpu_hp(sv, w, p, "fast", march_cols=cols)
    assert rel_l2(gvp, want[0]) <= TOL and rel_l2(gvs, want[1]) <= TOL


def test_hessian_vector_keeps_the_perturbation_and_accumulates():
    """NULL vector: the current (stage-1) perturbation; two calls add up."""
    sv, w = _case()
    p = _vec(sv, 3)
    shots = [lib_shot(sv, sh) for sh in sv.shots]
    with context(sv, "exact", max_batch=2) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        ctx.hessian_vector(shots[:1], *p)
        a0 = ctx.flush()
        ctx.hessian_vector(shots[1:])
        a1 = ctx.flush()
        ctx.hessian_vector(shots)
        one = ctx.flush()
        ctx.hessian_vector(shots)
        ctx.hessian_vector(shots[:1])
        ctx.hessian_vector(shots[1:])
        two = ctx.flush()
    for k in range(2):   # the accumulator adds each shot's product in call order
        assert np.array_equal(one[k], (a0[k] + a1[k]).astype(np.float32))
        assert np.array_equal(two[k], ((one[k] + a0[k]).astype(np.float32) + a1[k]).astype(np.float32))


def test_snapshot_hessian_is_symmetric():
    sv, w = _case(nt=240)
    x, y = _vec(sv, 4), _vec(sv, 5)
    kw = dict(recon="snapshots", snap_every=1, rr=1.0, march_cols=8)   # rr = 1: no damping
    hx = _gpu_hp(sv, w, x, "fast", **kw)
    hy = _gpu_hp(sv, w, y, "fast", **kw)
    a = float(np.sum(hx[0].astype(np.float64) * y[0]) + np.sum(hx[1].astype(np.float64) * y[1]))
    b = float(np.sum(hy[0].astype(np.float64) * x[0]) + np.sum(hy[1].astype(np.float64) * x[1]))
    assert abs(a - b) <= TOL * abs(a), (a, b)
    want = None
    for sh in sv.shots:
        h = O.hessian_vector_snapshot(oracle_cfg(sv, sh, pml_off=1, snap_every=1), sv.vp0, sv.vs0, sv.rho0, *x, w,
                                      sh.gc)
        want = h if want is None else (want[0] + h[0], want[1] + h[1])
    assert rel_l2(hx[0], want[0]) <= TOL and rel_l2(hx[1], want[1]) <= TOL


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_born2_forward(mode):
    sv, w = _case()
    q, r = _vec(sv, 6, edge=True)
    with context(sv, mode, max_batch=2) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        ctx.set_perturbation(q, r, region="s1")
        out = ctx.batch_born2_forward([lib_shot(sv, sh) for sh in sv.shots])
    for k, sh in enumerate(sv.shots):
        want = O.born2_forward(oracle_cfg(sv, sh), sv.vp0, sv.vs0, sv.rho0, q, r, w, sh.gc)
        ntr = len(sh.gc)
        assert np.abs(want[2]).max() > 0
        for a, b in zip(out, want):
            if mode == "exact":
                assert np.array_equal(a[k, :ntr], b), "max |d| = %g" % np.abs(a[k, :ntr] - b).max()
            else:
                assert rel_l2(a[k, :ntr], b) <= TOL


def test_gauss_newton_inner_loop_exact():
    sv, w = _case(nt=160)
    shots = [lib_shot(sv, sh) for sh in sv.shots]
    g1, g2 = _vec(sv, 7)
    il = np.ones_like(g1)
    st = Stage1(sv.allnx, sv.allnz, cut_layer=4, max_inner=2)
    st.direction(g1 * np.float32(1e-6), g2 * np.float32(1e-6), il)
    with context(sv, "exact", max_batch=2) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        d1, d2, hist = st.gauss_newton([(ctx, shots)])
    orc = I.s1_hcg_init(st.pp_cg_old, st.ps_cg_old)
    ohist = []
    for _ in range(2):
        hp = _oracle_hp(sv, w, orc["p"])
        ohist.append(I.s1_hcg_step(orc, hp[0], hp[1], True, 4))
    I.s1_hcg_finish(orc)
    assert hist == ohist
    assert np.array_equal(d1, orc["grad"][0]) and np.array_equal(d2, orc["grad"][1])
    assert all(h[2] > 0 for h in hist)   # p.Hp > 0: the Gauss-Newton Hessian is positive along p


@pytest.mark.parametrize("extra,okw", [(dict(rbc=True, rbc_seed=11), dict(rbc=1, rbc_seed=11)),
                                       (dict(cpu_mem=True, strip_chunk=30), {}),
                                       (dict(use_graph=True), {})])
def test_hessian_vector_storage_variants_exact(extra, okw):
    """The random boundary (no strips: time reversal, the oracle's rbc), strips
    in pinned host memory and graph replays: EXACT = the oracle bit for bit."""
    sv, w = _case()
    p = _vec(sv, 8, edge=True)
    want = _oracle_hp(sv, w, p, **okw)
    with context(sv, "exact", max_batch=len(sv.shots), **extra) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        for _ in range(3):   # with use_graph the later calls replay the captured solve
            ctx.hessian_vector([lib_shot(sv, sh) for sh in sv.shots], *p)
            gvp, gvs, _ = ctx.flush()
            assert np.array_equal(gvp, want[0]) and np.array_equal(gvs, want[1])
