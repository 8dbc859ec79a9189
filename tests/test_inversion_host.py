"""Host steps of the inversion loops (sfwi_s1_* / sfwi_s2_*), on the CPU.

The oracle's restatement (oracle/inversion.py) is pinned by what the
mathematics fixes: the vertex of the parabola through three exact samples of
a quadratic, the clamp, the joint max-abs normalisation, the Fletcher-Reeves
scaling of beta (beta = lambda^2 when the gradient is scaled by lambda, and
the normalized direction is then unchanged), the cut layer and the top-row
reset.  The library's host functions (plain CPU code in libsfwi_amd.so, no
device use) must equal the oracle bit for bit on random inputs.
"""
import numpy as np
import pytest

from oracle import inversion as I
from sfwi_amd.inversion import Stage1, Stage2

NX, NZ = 23, 37


def _fields(seed, nx=NX, nz=NZ):
    r = np.random.default_rng(seed)
    g1 = r.standard_normal((nx, nz)).astype(np.float32) * np.float32(3e-7)
    g2 = r.standard_normal((nx, nz)).astype(np.float32) * np.float32(1e-7)
    il = (r.random((nx, nz)) * 5 + 0.5).astype(np.float32)
    return g1, g2, il


# ------------------------------------------------------------ oracle pins ---

@pytest.mark.parametrize("smin", [57.3, 3.0, 211.0, -42.5])
def test_alpha_is_the_vertex_of_an_exact_quadratic(smin):
    a, c = 2.1, 5.0
    m = [a * (s - smin) ** 2 + c for s in (0.0, 30.0, 150.0)]
    assert I.s1_alpha(*m, 30.0, 150.0) == pytest.approx(smin, rel=1e-9, abs=1e-9)


def test_alpha_clamp():
    for smin, want in ((500.0, 300.0), (-250.0, -100.0)):
        m = [(s - smin) ** 2 for s in (0.0, 30.0, 150.0)]
        assert I.s1_alpha(*m, 30.0, 150.0) == want


def test_normalize_all_joint_max_abs():
    a = np.array([[1.0, -4.0]], np.float32)
    b = np.array([[-8.0, 0.5]], np.float32)
    x, y = I.normalize_all(a, b)
    assert max(np.abs(x).max(), np.abs(y).max()) == 1.0
    assert np.array_equal(x, a / np.float32(8)) and np.array_equal(y, b / np.float32(8))
    x, y = I.normalize_all(b, a)
    assert np.array_equal(y, a / np.float32(8))
    nan = np.array([[np.nan, 2.0]], np.float32)
    x, _ = I.normalize_all(nan, np.zeros_like(nan))       # NaN never wins the max
    assert x[0, 1] == 1.0


def test_s1_direction_iteration0_is_the_preconditioned_gradient():
    g1, g2, il = _fields(1)
    z = np.zeros_like(g1)
    d1, d2, c1, c2, fenzi, beta = I.s1_direction(0, 5, g1, g2, il, z, z, 0.0)
    assert beta == 0.0
    assert np.all(d1[:, :5] == 0) and np.all(d2[:, :5] == 0)
    pc1, pc2 = g1 / il, g2 / il
    s = np.abs(np.concatenate([pc1[:, 5:], pc2[:, 5:]])).max()
    assert np.allclose(d1[:, 5:], pc1[:, 5:] / s, rtol=1e-6)
    assert max(np.abs(d1).max(), np.abs(d2).max()) == 1.0
    assert fenzi == pytest.approx(float(np.sum((g1 * pc1)[:, 5:], dtype=np.float64)
                                        + np.sum((g2 * pc2)[:, 5:], dtype=np.float64)), rel=1e-6)


def test_s1_beta_scales_with_the_gradient_squared():
    g1, g2, il = _fields(2)
    z = np.zeros_like(g1)
    d1, d2, c1, c2, f0, _ = I.s1_direction(0, 4, g1, g2, il, z, z, 0.0)
    e1, e2, _, _, f1, beta = I.s1_direction(1, 4, g1 * 2, g2 * 2, il, c1, c2, f0)
    assert beta == 4.0 and f1 == 4.0 * f0           # exact: scaling by 2 is exact in binary
    assert np.abs(e1 - d1).max() <= 2.0 ** -22 and np.abs(e2 - d2).max() <= 2.0 ** -22   # 6 pc, normalized
    e1, e2, _, _, f1, beta = I.s1_direction(1, 4, g1, g2, il, c1, c2, f0)
    assert beta == 1.0 and np.array_equal(e1, d1) and np.array_equal(e2, d2)   # 2 pc, normalized


def test_s1_update_and_top_row_reset():
    r = np.random.default_rng(3)
    vp = (r.random((NX, NZ)) * 1000 + 2000).astype(np.float32)
    vs = (vp / 1.7).astype(np.float32)
    d1 = r.standard_normal((NX, NZ)).astype(np.float32)
    d2 = r.standard_normal((NX, NZ)).astype(np.float32)
    a, b = I.s1_update(6, 0.0, d1, d2, vp, vs)
    assert np.array_equal(a[:, 6:], vp[:, 6:]) and np.all(a[:, :6] == vp[0, 0]) and np.all(b[:, :6] == vs[0, 0])
    a, b = I.s1_update(0, 25.0, d1, d2, vp, vs)
    assert np.allclose(a, vp - 25.0 * d1, rtol=1e-6) and np.allclose(b, vs - 25.0 * d2, rtol=1e-6)
    a, b = I.s1_update(6, 25.0, d1, d2, vp, vs)     # the reset takes the updated first cell
    assert np.all(a[:, :6] == a[0, 0]) and a[0, 0] == pytest.approx(vp[0, 0] - 25.0 * d1[0, 0], rel=1e-6)
    assert np.allclose(a[:, 6:], vp[:, 6:] - 25.0 * d1[:, 6:], rtol=1e-6)


def test_s2_direction_cuts_and_scales():
    g1, g2, il = _fields(4)
    z = np.zeros_like(g1)
    h1, h2, c1, c2, f0, beta = I.s2_direction(0, 7, g1, g2, il, z, z, 0.0)
    assert beta == 0.0
    assert np.all(h1[:, :7] == 0) and np.array_equal(h1[:, 7:], g1[:, 7:])
    assert np.all(c1[:, :7] == 0) and np.array_equal(c1[:, 7:], g1[:, 7:] / il[:, 7:])
    _, _, k1, _, f1, beta = I.s2_direction(1, 7, g1 * 2, g2 * 2, il, c1, c2, f0)
    assert beta == 4.0 and f1 == 4.0 * f0
    assert np.array_equal(k1, c1 * 2 + c1 * 4)


def _spd(n, seed):
    r = np.random.default_rng(seed)
    a = r.standard_normal((n, n)) / np.sqrt(n)
    return a.T @ a + np.eye(n)      # well conditioned, symmetric positive definite


@pytest.mark.parametrize("plant", [0, 1])
def test_hcg_solves_the_newton_system(plant):
    """The inner CG in FP64 on an explicit SPD matrix: after as many steps as
    there are unknowns below the cut layer, r = 0 and delta = -r0 - H^-1 r0
    there (CG from zero solves H d = -r0; the reference starts delta at
    normalize(-cg) = -r0 and adds the CG steps), zero above the cut; grad =
    -delta.  Beta with the wrong sign does not converge."""
    nx, nz, cut = 3, 5, 2
    act = np.zeros((nx, nz), bool)
    act[:, cut:] = True
    m = 2 * int(act.sum())
    H = _spd(m, 11)
    r = np.random.default_rng(12)
    g1 = np.where(act, r.standard_normal((nx, nz)), 0.0)
    g2 = np.where(act, r.standard_normal((nx, nz)), 0.0)
    st = I.s1_hcg_init(g1, g2, dt=np.float64)
    r0 = np.concatenate([st["r"][0][act], st["r"][1][act]])
    def hv(p1, p2):
        y = H @ np.concatenate([p1[act], p2[act]])
        h1, h2 = np.zeros((nx, nz)), np.zeros((nx, nz))
        h1[act], h2[act] = y[:m // 2], y[m // 2:]
        h1[~act] = 7.0      # cut by the step
        return h1, h2
    for _ in range(m):
        I.s1_hcg_step(st, *hv(*st["p"]), True, cut, dt=np.float64, plant=plant)
    d = np.concatenate([st["delta"][0][act], st["delta"][1][act]])
    want = -r0 - np.linalg.solve(H, r0)
    if plant:
        assert np.linalg.norm(d - want) > 1e-3 * np.linalg.norm(want)
        return
    assert np.linalg.norm(d - want) <= 1e-10 * np.linalg.norm(want)
    assert np.abs(st["delta"][0][~act]).max() == 0 and np.abs(st["delta"][1][~act]).max() == 0
    assert np.linalg.norm(np.concatenate([st["r"][0][act], st["r"][1][act]])) <= 1e-10
    assert np.array_equal(st["grad"][0], -st["delta"][0])
    I.s1_hcg_finish(st, dt=np.float64)
    assert max(np.abs(st["grad"][0]).max(), np.abs(st["grad"][1]).max()) == 1.0


# ------------------------------------------------- library equals oracle ---

@pytest.mark.parametrize("seed", [13, 14])
def test_library_hcg_steps_equal_oracle(seed):
    g1, g2, il = _fields(seed)
    st = Stage1(NX, NZ, cut_layer=4)
    st.direction(g1, g2, il)
    lib = st.hcg_init()
    orc = I.s1_hcg_init(st.pp_cg_old, st.ps_cg_old)
    for k in ("r", "p", "delta", "grad"):
        assert np.array_equal(lib[k][0], orc[k][0]) and np.array_equal(lib[k][1], orc[k][1]), k
    r = np.random.default_rng(seed)
    A = (r.standard_normal((NZ, NZ)) * 0.2 + np.eye(NZ) * 2).astype(np.float32)
    for it in range(3):
        h1 = (lib["p"][0] @ A).astype(np.float32)
        h2 = (lib["p"][1] @ A.T).astype(np.float32)
        a = st.hcg_step(lib, h1, h2)
        b = I.s1_hcg_step(orc, h1, h2, True, 4)
        assert a == b, (a, b)
        for k in ("r", "p", "delta", "grad"):
            assert np.array_equal(lib[k][0], orc[k][0]) and np.array_equal(lib[k][1], orc[k][1]), (it, k)
    st.hcg_finish(lib)
    I.s1_hcg_finish(orc)
    for k in ("delta", "grad"):
        assert np.array_equal(lib[k][0], orc[k][0]) and np.array_equal(lib[k][1], orc[k][1]), k


@pytest.mark.parametrize("seed", [5, 6])
def test_library_s1_steps_equal_oracle(seed):
    g1, g2, il = _fields(seed)
    st = Stage1(NX, NZ, cut_layer=4)
    o_old1 = np.zeros_like(g1)
    o_old2 = np.zeros_like(g1)
    o_fm = 0.0
    r = np.random.default_rng(seed)
    vp = (r.random((NX, NZ)) * 1000 + 2000).astype(np.float32)
    vs = (vp / 1.7).astype(np.float32)
    ovp, ovs = vp.copy(), vs.copy()
    for it in range(3):
        d1, d2, beta = st.direction(g1, g2, il)
        e1, e2, o_old1, o_old2, o_fm, obeta = I.s1_direction(it, 4, g1, g2, il, o_old1, o_old2, o_fm)
        assert np.array_equal(d1, e1) and np.array_equal(d2, e2)
        assert np.array_equal(st.pp_cg_old, o_old1) and st.beta_fenmu == o_fm and beta == obeta
        m = [float(x) for x in r.random(3) * 10 + 1]
        alpha = st.alpha(*m)
        assert alpha == I.s1_alpha(*m, 30.0, 150.0)
        st.update(vp, vs, d1, d2, alpha)
        ovp, ovs = I.s1_update(4, alpha, e1, e2, ovp, ovs)
        assert np.array_equal(vp, ovp) and np.array_equal(vs, ovs)
        st.iter += 1
        g1 = (g1 * np.float32(0.7) + np.float32(1e-8)).astype(np.float32)


@pytest.mark.parametrize("seed", [7, 8])
def test_library_s2_steps_equal_oracle(seed):
    g1, g2, il = _fields(seed)
    st = Stage2(NX, NZ, cut_layer=6)
    o_old1 = np.zeros_like(g1)
    o_old2 = np.zeros_like(g1)
    o_fm = 0.0
    ip = np.zeros_like(g1)
    is_ = np.zeros_like(g1)
    oip, ois = ip.copy(), is_.copy()
    for it in range(3):
        c1, c2, beta, h1, h2 = st.direction(g1, g2, il)
        oh1, oh2, k1, k2, ofz, obeta = I.s2_direction(it, 6, g1, g2, il, o_old1, o_old2, o_fm)
        o_old1, o_old2, o_fm = k1, k2, ofz
        assert np.array_equal(h1, oh1) and np.array_equal(h2, oh2)
        assert np.array_equal(c1, k1) and np.array_equal(c2, k2) and beta == obeta and st.beta_fenmu == o_fm
        fz, fm = 3.5e-3 * (it + 1), 7.25e-2
        a = st.update(ip, is_, c1, c2, fz, fm)
        oip, ois, oa = I.s2_update(fz, fm, k1, k2, oip, ois)
        assert a == oa and np.array_equal(ip, oip) and np.array_equal(is_, ois)
        st.iter += 1
        g2 = (g2 * np.float32(1.3)).astype(np.float32)


def test_library_rejects_bad_arguments():
    from sfwi_amd._lib import SfwiError
    st = Stage1(NX, NZ, cut_layer=-1)
    g1, g2, il = _fields(9)
    with pytest.raises(SfwiError, match="cut_layer"):
        st.direction(g1, g2, il)
