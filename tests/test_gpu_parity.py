"""GPU parity of the HIP library against the CPU oracle (C ABI via ctypes).

Tolerances (relative L2 against the oracle's FP32 restatement of the
reference, on identical seeded inputs):
  * EXACT mode: bit-identical traces, gradients and illumination (the same
    operation order, FP64 promotions and no FMA contraction); misfit sums to
    1e-12 (the FP64 partial sums are grouped per trace).
  * FAST mode: traces <= 1e-5, gradients / illumination / misfit <= 1e-4.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sfwi_amd import workloads as W
from tests.helpers import context, lib_shot, oracle_cfg, rel_l2

pytestmark = pytest.mark.gpu

TRACE_TOL = 1e-5
GRAD_TOL = 1e-4


def _survey(kind, nt=200):
    if kind == "elastic":
        return W.tiny(nt=nt)
    if kind == "water":          # Vs = 0 in a top water layer (plain_small's feature)
        sv = W.tiny(nt=nt, seed=11)
        for a in (sv.vs, sv.vs0):
            a[:, :6] = 0.0
        return sv
    if kind == "acoustic":
        return W.tiny(nt=nt, elastic=False, seed=5)
    raise ValueError(kind)


def _conditioned(c, sv, sh, w, oz, ox, ref32, frac=0.2):
    """The FP32-vs-FP64 spread of the oracle itself must sit well below the
    parity bar, or a comparison at that bar measures rounding noise (a
    residual that is a near-cancellation of the direct wave is such a case)."""
    ref64 = O.gradient(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox, "f64")
    for a, b in zip(ref32[:3], ref64[:3]):
        assert rel_l2(a, b) <= frac * GRAD_TOL, "ill-conditioned case: %g" % rel_l2(a, b)


def _observed(sv, sh, w):
    c = oracle_cfg(sv, sh)
    return O.forward(c, sv.vp, sv.vs, sv.rho, w, sh.gc, "f32")


@pytest.mark.parametrize("kind", ["elastic", "water", "acoustic"])
def test_forward_exact_bitwise(kind):
    sv = _survey(kind)
    sh = sv.shots[0]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    oz, ox = _observed(sv, sh, w)
    with context(sv, "exact") as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp, sv.vs, sv.rho)
        gz, gx = ctx.forward(lib_shot(sv, sh))
    assert np.abs(oz).max() > 0
    assert np.array_equal(gz, oz), "max |d| = %g" % np.abs(gz - oz).max()
    assert np.array_equal(gx, ox), "max |d| = %g" % np.abs(gx - ox).max()


@pytest.mark.parametrize("kind", ["elastic", "water"])
def test_forward_fast_within_tolerance(kind):
    sv = _survey(kind)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    with context(sv, "fast") as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp, sv.vs, sv.rho)
        for sh in sv.shots:
            oz, ox = _observed(sv, sh, w)
            gz, gx = ctx.forward(lib_shot(sv, sh))
            assert rel_l2(gz, oz) <= TRACE_TOL
            assert rel_l2(gx, ox) <= TRACE_TOL


def _gradient_case(kind):
    sv = _survey(kind, nt=320)
    sh = sv.shots[1]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    oz, ox = _observed(sv, sh, w)
    c = oracle_cfg(sv, sh)
    ref = O.gradient(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox, "f32")
    return sv, sh, w, oz, ox, ref


@pytest.mark.parametrize("kind", ["elastic", "water"])
def test_gradient_exact_bitwise(kind):
    sv, sh, w, oz, ox, (g_vp, g_vs, ill, mis, _, _) = _gradient_case(kind)
    with context(sv, "exact") as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        m = ctx.gradient(lib_shot(sv, sh, oz, ox))
        gvp, gvs, gil = ctx.flush()
    assert np.abs(g_vp).max() > 0 and np.abs(g_vs).max() > 0
    assert np.array_equal(gvp, g_vp), "max |d| = %g" % np.abs(gvp - g_vp).max()
    assert np.array_equal(gvs, g_vs), "max |d| = %g" % np.abs(gvs - g_vs).max()
    assert np.array_equal(gil, ill), "max |d| = %g" % np.abs(gil - ill).max()
    assert abs(m - mis) <= 1e-12 * abs(mis)


@pytest.mark.parametrize("kind", ["elastic", "water"])
def test_gradient_fast_within_tolerance(kind):
    sv, sh, w, oz, ox, ref = _gradient_case(kind)
    g_vp, g_vs, ill, mis = ref[:4]
    _conditioned(oracle_cfg(sv, sh), sv, sh, w, oz, ox, ref)
    with context(sv, "fast") as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        m = ctx.gradient(lib_shot(sv, sh, oz, ox))
        gvp, gvs, gil = ctx.flush()
    assert rel_l2(gvp, g_vp) <= GRAD_TOL
    assert rel_l2(gvs, g_vs) <= GRAD_TOL
    assert rel_l2(gil, ill) <= GRAD_TOL
    assert abs(m - mis) <= GRAD_TOL * abs(mis)


def test_batch_equals_alone():
    """Shots advanced together in one batch give the same per-shot results as
    the same shots run one at a time (no cross-shot coupling, fixed order)."""
    sv = W.tiny(nt=160, nshots=4, seed=3)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    shots = []
    for sh in sv.shots:
        oz, ox = _observed(sv, sh, w)
        shots.append(lib_shot(sv, sh, oz, ox))
    with context(sv, "fast", max_batch=4) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        mb = ctx.batch_gradient(shots)
        gb = ctx.flush()
        alone_m, acc = [], None
        for s in shots:
            alone_m.append(ctx.gradient(s))
            g = ctx.flush()
            acc = [a.copy() for a in g] if acc is None else [a + b for a, b in zip(acc, g)]
    assert np.array_equal(mb, np.array(alone_m))
    for a, b in zip(gb, acc):
        assert np.array_equal(a, b)


def test_window_and_edge_geometry():
    """Partial aperture: the window starts inside the model, receivers cover
    part of it, one coincident receiver pair, source near the PML."""
    base = W.tiny(nx=64, nt=320, seed=9)
    nxw = 40
    left = 12
    gc = (np.arange(left + 3, left + 30) * 10).astype(np.int32)
    gc = np.concatenate([gc, gc[5:6]])                 # coincident receiver
    shot = W.Shot(sx=int((left + 2) * 10), gc=gc, ngx_left=int(left * 10), left=left)
    sv = W.Survey(base.name, base.vp, base.vs, base.rho, base.vp0, base.vs0, base.rho0, base.dx,
                  base.dz, base.dt, base.nt, base.pml, base.f0, 1, 0, nxw, [shot])
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    c = oracle_cfg(sv, shot)
    oz, ox = O.forward(c, sv.vp, sv.vs, sv.rho, w, gc, "f32")
    ref = O.gradient(c, sv.vp0, sv.vs0, sv.rho0, w, gc, oz, ox, "f32")
    _conditioned(c, sv, shot, w, oz, ox, ref)
    for mode in ("exact", "fast"):
        with context(sv, mode) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp, sv.vs, sv.rho)
            gz, gx = ctx.forward(lib_shot(sv, shot))
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            m = ctx.gradient(lib_shot(sv, shot, oz, ox))
            gvp, gvs, gil = ctx.flush()
        if mode == "exact":
            assert np.array_equal(gz, oz) and np.array_equal(gx, ox)
            assert np.array_equal(gvp, ref[0]) and np.array_equal(gvs, ref[1])
        else:
            assert rel_l2(gz, oz) <= TRACE_TOL and rel_l2(gx, ox) <= TRACE_TOL
            assert rel_l2(gvp, ref[0]) <= GRAD_TOL and rel_l2(gvs, ref[1]) <= GRAD_TOL
        assert abs(m - ref[3]) <= GRAD_TOL * abs(ref[3])
        # nothing lands outside the window's columns
        assert not gvp[:left].any() and not gvp[left + nxw:].any()


def _born_case():
    sv = W.tiny(nt=320, seed=21)
    dvp = (sv.vp - sv.vp0).astype(np.float32)
    dvs = (sv.vs - sv.vs0).astype(np.float32)
    return sv, dvp, dvs


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_born_forward(mode):
    sv, dvp, dvs = _born_case()
    sh = sv.shots[0]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    c = oracle_cfg(sv, sh, rho_const=2000.0)
    ref = O.born_forward(c, sv.vp0, sv.vs0, sv.rho0, dvp, dvs, w, sh.gc, "f32")
    with context(sv, mode, rho_const=2000.0, adjoint_dir=-1) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, None)
        ctx.set_perturbation(dvp, dvs)
        out = ctx.born_forward(lib_shot(sv, sh))
    assert np.abs(ref[0]).max() > 0
    for a, b in zip(out, ref):
        if mode == "exact":
            assert np.array_equal(a, b), "max |d| = %g" % np.abs(a - b).max()
        else:
            assert rel_l2(a, b) <= TRACE_TOL


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_born_gradient(mode):
    sv, dvp, dvs = _born_case()
    sh = sv.shots[1]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    c = oracle_cfg(sv, sh, rho_const=2000.0, adjoint_dir=-1)
    oz, ox = O.forward(c, sv.vp, sv.vs, sv.rho, w, sh.gc, "f32")
    g_vp, g_vs, ill, mis = O.born_gradient(c, sv.vp0, sv.vs0, sv.rho0, dvp * 0.5, dvs * 0.5, w, sh.gc,
                                          oz, ox, "f32")
    with context(sv, mode, rho_const=2000.0, adjoint_dir=-1) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, None)
        ctx.set_perturbation(dvp * 0.5, dvs * 0.5)
        m = ctx.born_gradient(lib_shot(sv, sh, oz, ox))
        gvp, gvs, gil = ctx.flush()
    if mode == "exact":
        assert np.array_equal(gvp, g_vp) and np.array_equal(gvs, g_vs) and np.array_equal(gil, ill)
        assert abs(m - mis) <= 1e-12 * abs(mis)
    else:
        assert rel_l2(gvp, g_vp) <= GRAD_TOL and rel_l2(gvs, g_vs) <= GRAD_TOL
        assert rel_l2(gil, ill) <= GRAD_TOL
        assert abs(m - mis) <= GRAD_TOL * abs(mis)


def test_linesearch_parabola():
    sv = W.tiny(nt=160, seed=4)
    sh = sv.shots[0]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    oz, ox = _observed(sv, sh, w)
    rng = np.random.default_rng(0)
    d_vp = W.gaussian_smooth(rng.standard_normal(sv.vp.shape).astype(np.float32), 3.0)
    d_vs = W.gaussian_smooth(rng.standard_normal(sv.vp.shape).astype(np.float32), 3.0)
    d_vp /= np.abs(d_vp).max()
    d_vs /= np.abs(d_vs).max()
    c = oracle_cfg(sv, sh)
    ref = O.parabola_misfits(c, sv.vp0, sv.vs0, sv.rho0, d_vp, d_vs, 30.0, 150.0, w, sh.gc, oz, ox)
    for mode in ("exact", "fast"):
        with context(sv, mode) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            m = ctx.linesearch_parabola(lib_shot(sv, sh, oz, ox), d_vp, d_vs, 30.0, 150.0)
        tol = 1e-12 if mode == "exact" else GRAD_TOL
        assert np.all(np.abs(m - ref) <= tol * np.abs(ref)), (mode, m, ref)


def test_linesearch_linear():
    sv, dvp, dvs = _born_case()
    sh = sv.shots[0]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    c = oracle_cfg(sv, sh, rho_const=2000.0, adjoint_dir=-1)
    oz, ox = O.forward(c, sv.vp, sv.vs, sv.rho, w, sh.gc, "f32")
    cg_vp, cg_vs = dvp * 0.3, dvs * -0.2
    ref = O.linear_alpha(c, sv.vp0, sv.vs0, sv.rho0, dvp * 0.5, dvs * 0.5, cg_vp, cg_vs, w, sh.gc, oz, ox)
    for mode in ("exact", "fast"):
        with context(sv, mode, rho_const=2000.0, adjoint_dir=-1) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, None)
            ctx.set_perturbation(dvp * 0.5, dvs * 0.5)
            fz, fm = ctx.linesearch_linear(lib_shot(sv, sh, oz, ox), cg_vp, cg_vs)
        tol = 1e-12 if mode == "exact" else GRAD_TOL
        assert abs(fz - ref[0]) <= tol * abs(ref[0]) and abs(fm - ref[1]) <= tol * abs(ref[1]), (mode, fz, fm, ref)


def test_single_trace_and_errors():
    sv = W.tiny(nt=120, seed=2)
    sh = sv.shots[0]
    one = W.Shot(sx=sh.sx, gc=sh.gc[17:18].copy(), ngx_left=0, left=0)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    c = oracle_cfg(sv, one)
    oz, ox = O.forward(c, sv.vp, sv.vs, sv.rho, w, one.gc, "f32")
    with context(sv, "exact") as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp, sv.vs, sv.rho)
        gz, gx = ctx.forward(lib_shot(sv, one))
        assert np.array_equal(gz, oz) and np.array_equal(gx, ox)
        # a receiver outside the padded grid is rejected, not written
        bad = W.Shot(sx=sh.sx, gc=np.array([10 * 10000], np.int32), ngx_left=0, left=0)
        with pytest.raises(Exception, match="receiver column"):
            ctx.forward(lib_shot(sv, bad))
        # the context stays usable after an error
        gz2, _ = ctx.forward(lib_shot(sv, one))
        assert np.array_equal(gz2, oz)


@pytest.mark.parametrize("cols,depth,order", [(8, 1, 2), (8, 1, 1), (16, 2, 2), (37, 4, 1), (8, 0, 0), (8, 4, 3),
                                               (16, 1, 3), (37, 2, 3)])
def test_march_kernels_match_per_cell(cols, depth, order):
    """The column-march FAST kernels use the per-cell FAST update math in the
    same operation order, so gradients, misfits and Born traces agree bitwise
    for any segment width (37: segments that do not divide the grid), any
    prefetch depth (0: the default) and every block order (0: the default)."""
    sv, dvp, dvs = _born_case()
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    shots = []
    for sh in sv.shots:
        oz, ox = _observed(sv, sh, w)
        shots.append(lib_shot(sv, sh, oz, ox))
    out = {}
    for mc in (-1, cols):
        with context(sv, "fast", max_batch=2, march_cols=mc, march_depth=depth, march_order=order) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            m = ctx.batch_gradient(shots)
            g = ctx.flush()
            ctx.set_perturbation(dvp, dvs)
            b = ctx.born_forward(shots[0])
            mb = ctx.batch_gradient(shots, born=True)
            gb = ctx.flush()
        out[mc] = (m, g, b, mb, gb)
    a, b = out[-1], out[cols]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[3], b[3])
    for x, y in zip(a[1] + a[2] + a[4], b[1] + b[2] + b[4]):
        assert np.array_equal(x, y), "max |d| = %g" % np.abs(x - y).max()


@pytest.mark.parametrize("cols,depth,order", [(8, 4, 3), (8, 1, 2), (16, 2, 1), (37, 4, 3), (8, 0, 0)])
def test_march_kernels_match_per_cell_multiband(cols, depth, order):
    """As above on a grid tall enough for several row bands per column strip:
    nz = 134 with an 8-cell PML gives a 142-row update domain, three 56-row
    march bands of which the last holds 30 rows (and three interior bands of
    56, 56 and 22 rows for the snapshot imaging march), so the seams between
    bands -- halo lanes, band offsets, a partial last band -- are compared.
    Strip gradients, Born traces and gradients, and snapshot gradients (BF16)."""
    sv = W.tiny(nx=44, nz=134, pml=8, nt=200, seed=23)
    dvp = (sv.vp - sv.vp0).astype(np.float32)
    dvs = (sv.vs - sv.vs0).astype(np.float32)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    shots = []
    for sh in sv.shots:
        oz, ox = _observed(sv, sh, w)
        shots.append(lib_shot(sv, sh, oz, ox))
    out = {}
    for mc in (-1, cols):
        with context(sv, "fast", max_batch=2, march_cols=mc, march_depth=depth, march_order=order) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            m = ctx.batch_gradient(shots)
            g = ctx.flush()
            ctx.set_perturbation(dvp, dvs)
            b = ctx.born_forward(shots[0])
            mb = ctx.batch_gradient(shots, born=True)
            gb = ctx.flush()
        with context(sv, "fast", max_batch=2, march_cols=mc, march_depth=depth, march_order=order,
                     recon="snapshots", snap_every=3, storage="bf16") as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            ms = ctx.batch_gradient(shots)
            gs = ctx.flush()
        out[mc] = (m, g, b, mb, gb, ms, gs)
    a, b = out[-1], out[cols]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[3], b[3]) and np.array_equal(a[5], b[5])
    assert np.abs(a[1][0]).max() > 0 and np.abs(a[6][0]).max() > 0
    for x, y in zip(a[1] + a[2] + a[4] + a[6], b[1] + b[2] + b[4] + b[6]):
        assert np.array_equal(x, y), "max |d| = %g" % np.abs(x - y).max()


def test_resident_observed_records():
    """Shots with NULL observed traces reuse the records left in their batch
    slot by the previous call: same misfits and gradients, bitwise."""
    sv = W.tiny(nt=200)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    recs = [_observed(sv, s, w) for s in sv.shots]
    with context(sv, max_batch=len(sv.shots)) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        with pytest.raises(Exception, match="resident"):
            ctx.batch_gradient([lib_shot(sv, s) for s in sv.shots])
        m1 = ctx.batch_gradient([lib_shot(sv, s, *r) for s, r in zip(sv.shots, recs)])
        g1 = ctx.flush()
        m2 = ctx.batch_gradient([lib_shot(sv, s) for s in sv.shots])
        g2 = ctx.flush()
    assert np.array_equal(m1, m2)
    for a, b in zip(g1, g2):
        assert np.array_equal(a, b)


def _overlap_case():
    """Three shots whose 57-column windows start at columns 0, 11 and 23 of an
    80-column model (overlapping, so their images land on shared model cells),
    49 rows: the padded grid's last x segment is one column wide (65 update
    columns: 8 segments of 8 and one of 1) and its last row tile holds one
    output row (57 update rows: 56 + 1).  Shot 1 has a coincident receiver
    pair, shot 2 its source next to the PML."""
    base = W.tiny(nx=80, nz=49, pml=8, nt=240, seed=13)
    nxw = 57
    shots = []
    for k, left in enumerate((0, 11, 23)):
        gc = (np.arange(left + 2, left + nxw - 2, 2) * 10).astype(np.int32)
        if k == 1:
            gc = np.concatenate([gc, gc[4:5]])
        sx = (left + (1 if k == 2 else nxw // 2)) * 10
        shots.append(W.Shot(sx=int(sx), gc=gc, ngx_left=int(left * 10), left=left))
    sv = W.Survey(base.name, base.vp, base.vs, base.rho, base.vp0, base.vs0, base.rho0, base.dx,
                  base.dz, base.dt, base.nt, base.pml, base.f0, 2, 1, nxw, shots)
    return sv, max(len(s.gc) for s in shots)


def test_overlapping_windows_batch_exact():
    """A batch of shots with overlapping windows at different offsets, EXACT
    mode: the per-shot images are merged into the model without lost updates
    and in batch order, so the batch equals the oracle's per-shot FP32 images
    summed in FP32 in that order, bitwise.  (This case is ill-conditioned for
    a FAST-vs-oracle comparison: the oracle's own FP32 and FP64 gradients
    differ by a few per cent, so FAST is checked against EXACT structure
    instead, in the test below.)"""
    sv, ntr = _overlap_case()
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    obs = [_observed(sv, s, w) for s in sv.shots]
    ref = [np.zeros((sv.allnx, sv.allnz), np.float32) for _ in range(3)]
    mis = []
    for s, (oz, ox) in zip(sv.shots, obs):
        r = O.gradient(oracle_cfg(sv, s), sv.vp0, sv.vs0, sv.rho0, w, s.gc, oz, ox, "f32")
        for k in range(3):
            ref[k] = ref[k] + r[k]
        mis.append(r[3])
    with context(sv, "exact", max_batch=3, ntr=ntr) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        m = ctx.batch_gradient([lib_shot(sv, s, *o) for s, o in zip(sv.shots, obs)])
        g = ctx.flush()
    assert np.abs(ref[0]).max() > 0
    for x, y in zip(g, ref):
        assert np.array_equal(x, y), "max |d| = %g" % np.abs(x - y).max()
    assert np.all(np.abs(m - np.array(mis)) <= 1e-12 * np.abs(np.array(mis)))


@pytest.mark.parametrize("depth,order", [(4, 3), (1, 1)])
def test_overlapping_windows_batch(depth, order):
    """The same batch in FAST mode: the batch equals the shots merged one call
    at a time, bitwise; the column-march reverse step (imaging inside the
    reconstruction kernels; with the density image the stored-derivative
    arrangement) equals the per-cell FAST kernels bitwise on this grid's
    one-column segment and one-row tile."""
    sv, ntr = _overlap_case()
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    lib = [lib_shot(sv, s, *_observed(sv, s, w)) for s in sv.shots]
    for grad_rho in (False, True):
        out = {}
        for mc in (-1, 8):
            with context(sv, "fast", max_batch=3, ntr=ntr, march_cols=mc, march_depth=depth, march_order=order,
                         grad_rho=grad_rho) as ctx:
                ctx.set_wavelet(w)
                ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
                m = ctx.batch_gradient(lib)
                g = list(ctx.flush()) + ([ctx.flush_rho()] if grad_rho else [])
                for s in lib:
                    ctx.batch_gradient([s])
                g1 = list(ctx.flush()) + ([ctx.flush_rho()] if grad_rho else [])
            assert np.abs(g[0]).max() > 0
            for x, y in zip(g, g1):
                assert np.array_equal(x, y), "batch vs one at a time: max |d| = %g" % np.abs(x - y).max()
            out[mc] = (m, g)
        (ma, ga), (mb, gb) = out[-1], out[8]
        assert np.array_equal(ma, mb)
        for x, y in zip(ga, gb):
            assert np.array_equal(x, y), "grad_rho %s: max |d| = %g" % (grad_rho, np.abs(x - y).max())


def test_batch_linesearch_linear_equals_per_shot_calls():
    """sfwi_batch_linesearch_linear over shots with overlapping windows at
    three offsets, in batches of 2 + 1 (max_batch 2), equals the per-shot
    sfwi_linesearch_linear calls in shot order, bit for bit, in both modes,
    starting from zero and from non-zero sums; the observed records stay
    resident in their slots for a following gradient call."""
    sv, ntr = _overlap_case()
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    lib = [lib_shot(sv, s, *_observed(sv, s, w)) for s in sv.shots]
    dvp = (0.5 * (sv.vp - sv.vp0)).astype(np.float32)
    dvs = (0.5 * (sv.vs - sv.vs0)).astype(np.float32)
    rng = np.random.default_rng(6)
    c1 = (rng.standard_normal(sv.vp.shape) * 0.1).astype(np.float32)
    c2 = (rng.standard_normal(sv.vp.shape) * 0.05).astype(np.float32)
    for mode in ("fast", "exact"):
        with context(sv, mode, max_batch=2, ntr=ntr, rho_const=2000.0, adjoint_dir=-1) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, None)
            ctx.set_perturbation(dvp, dvs)
            per0 = (0.0, 0.0)
            for s in lib:
                per0 = ctx.linesearch_linear(s, c1, c2, *per0)
            start = (0.25 * per0[0], 0.5 * per0[1])
            per = start
            for s in lib:
                per = ctx.linesearch_linear(s, c1, c2, *per)
            # a per-shot call leaves its shot's records in slot 0
            m_one = ctx.batch_gradient([lib_shot(sv, sv.shots[2])], born=True)
            g_one = ctx.flush()
            bat0 = ctx.batch_linesearch_linear(lib, c1, c2)
            bat = ctx.batch_linesearch_linear(lib, c1, c2, *start)
            # the last batch was shot 2 alone, in slot 0: its records are resident there
            m_obs = ctx.batch_gradient([lib[2]], born=True)
            g_obs = ctx.flush()
            m_res = ctx.batch_gradient([lib_shot(sv, sv.shots[2])], born=True)
            g_res = ctx.flush()
        assert per0[0] != 0.0 and per0[1] > 0.0
        assert bat0 == per0, (mode, per0, bat0)
        assert bat == per, (mode, per, bat)
        assert np.array_equal(m_obs, m_res) and np.array_equal(m_obs, m_one)
        for x, y, z in zip(g_obs, g_res, g_one):
            assert np.array_equal(x, y) and np.array_equal(x, z)


def test_linear_step_with_gradient_residuals():
    """The Born gradient's residuals handed to the linear step in place of its
    first Born solve give the sums of the step that recomputes them, bit for
    bit (batches of 2 + 1 over overlapping windows, both modes), and the same
    misfits and gradients as the plain Born gradient call."""
    sv, ntr = _overlap_case()
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    lib = [lib_shot(sv, s, *_observed(sv, s, w)) for s in sv.shots]
    dvp = (0.5 * (sv.vp - sv.vp0)).astype(np.float32)
    dvs = (0.5 * (sv.vs - sv.vs0)).astype(np.float32)
    rng = np.random.default_rng(8)
    c1 = (rng.standard_normal(sv.vp.shape) * 0.1).astype(np.float32)
    c2 = (rng.standard_normal(sv.vp.shape) * 0.05).astype(np.float32)
    for mode in ("fast", "exact"):
        with context(sv, mode, max_batch=2, ntr=ntr, rho_const=2000.0, adjoint_dir=-1) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, None)
            ctx.set_perturbation(dvp, dvs)
            m0 = np.concatenate([ctx.batch_gradient(lib[:2], born=True), ctx.batch_gradient(lib[2:], born=True)])
            g0 = ctx.flush()
            parts = [ctx.batch_born_gradient_res(lib[:2]), ctx.batch_born_gradient_res(lib[2:])]
            g1 = ctx.flush()
            rz = np.concatenate([p[1] for p in parts])
            rx = np.concatenate([p[2] for p in parts])
            ref = ctx.batch_linesearch_linear(lib, c1, c2)
            got = ctx.batch_linesearch_linear_res(lib, c1, c2, rz, rx)
        assert np.array_equal(np.concatenate([p[0] for p in parts]), m0)
        for a, b in zip(g0, g1):
            assert np.array_equal(a, b)
        assert np.abs(rz).max() > 0 and ref[1] > 0
        assert got == ref, (mode, got, ref)


def test_batch_linesearch_equals_per_shot_calls():
    """sfwi_batch_linesearch_parabola over shots with overlapping windows at
    three offsets, in batches of 2 + 1 (max_batch 2), equals the per-shot
    sfwi_linesearch_parabola calls in shot order, bit for bit, in both modes;
    a non-zero starting m is added to."""
    sv, ntr = _overlap_case()
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    lib = [lib_shot(sv, s, *_observed(sv, s, w)) for s in sv.shots]
    rng = np.random.default_rng(5)
    d1 = (rng.standard_normal(sv.vp.shape) * 0.1).astype(np.float32)
    d2 = (rng.standard_normal(sv.vp.shape) * 0.05).astype(np.float32)
    for mode in ("fast", "exact"):
        with context(sv, mode, max_batch=2, ntr=ntr) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            per0 = np.zeros(3)
            for s in lib:
                per0 = ctx.linesearch_parabola(s, d1, d2, 30.0, 150.0, per0)
            m0 = 0.5 * per0                          # a starting m on the misfits' scale
            per = m0.copy()
            for s in lib:
                per = ctx.linesearch_parabola(s, d1, d2, 30.0, 150.0, per)
            bat0 = ctx.batch_linesearch_parabola(lib, d1, d2, 30.0, 150.0)
            bat = ctx.batch_linesearch_parabola(lib, d1, d2, 30.0, 150.0, m0)
        assert np.all(per0 > 0)
        assert np.array_equal(per0, bat0), (mode, per0, bat0)
        assert np.array_equal(per, bat), (mode, per, bat)


def test_empty_shot_lists():
    """Every batched call takes an empty shot list (also as a NULL pointer) as
    a no-op: nothing is solved, sums are returned unchanged and the gradient
    accumulators stay as they were."""
    import ctypes
    sv = W.tiny(nt=120)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    lib = [lib_shot(sv, sh, *_observed(sv, sh, w)) for sh in sv.shots[:1]]
    d = np.zeros_like(sv.vp)
    with context(sv, "fast", max_batch=2) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        ctx.set_perturbation(d, d)
        m = ctx.batch_gradient(lib)
        g = ctx.flush()
        ctx.batch_gradient(lib)
        assert ctx.batch_gradient([]).size == 0 and ctx.batch_gradient([], born=True).size == 0
        z, x = ctx.batch_forward([])
        assert z.shape[0] == 0
        assert np.array_equal(ctx.batch_linesearch_parabola([], d, d, 30.0, 150.0, [1.0, 2.0, 3.0]), [1.0, 2.0, 3.0])
        assert ctx.batch_linesearch_linear([], d, d, 4.0, 5.0) == (4.0, 5.0)
        L, h = ctx.L, ctx.h
        mis = (ctypes.c_double * 1)()
        assert L.sfwi_batch_gradient(h, 0, None, mis) == 0
        assert L.sfwi_batch_born_gradient(h, 0, None, mis) == 0
        assert L.sfwi_batch_forward(h, 0, None, None, None) == 0
        assert L.sfwi_batch_gradient(h, -1, None, mis) != 0
        g2 = ctx.flush()             # only the one call with shots accumulated
    assert m[0] > 0
    for a, b in zip(g, g2):
        assert np.array_equal(a, b)
