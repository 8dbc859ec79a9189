"""Pins of the OT misfit oracle (oracle/ot.py) against closed forms, a brute
force quantile integration and finite differences.

The reference declares OT / OT_power / OT_alltrace (sfwi-stage1/FD2DGPU.cuh:21-25)
but never defines them, so no reference output exists; these checks fix the
oracle's mathematics instead: exact W2^2 = shift^2 for cell-aligned shifts,
the Gaussian closed form (mu1 - mu2)^2 + (s1 - s2)^2, agreement with an
independent inverse-CDF integration, symmetry, zero misfit and zero adjoint for
identical traces, and the adjoint source as the derivative of the misfit.
"""
import numpy as np
import pytest

from oracle import ot


def _gauss(n, mu, sig):
    t = np.arange(n)
    return np.exp(-((t - mu) / sig) ** 2 / 2)


@pytest.mark.parametrize("shift", [1, 7, 30])
def test_cell_aligned_shift_is_shift_squared(shift):
    dt = 0.002
    o = _gauss(400, 150, 12.0)
    s = np.roll(o, shift)
    m, r = ot.ot_signal(np.sqrt(s), np.sqrt(o), 1, dt)      # power transform: u = s^2
    assert m == pytest.approx((shift * dt) ** 2, rel=1e-12)
    assert np.abs(r).max() > 0


@pytest.mark.parametrize("m1,s1,m2,s2", [(150, 12, 180, 20), (200, 30, 190, 15)])
def test_gaussian_closed_form(m1, s1, m2, s2):
    W, _, _ = ot.w2_cells(_gauss(400, m1, s1), _gauss(400, m2, s2))
    # piecewise-constant densities: O(cell^2 / sigma^2) discretisation error
    assert W == pytest.approx((m1 - m2) ** 2 + (s1 - s2) ** 2, rel=3e-4)


def _inverse_cdf_w2(a, b, K=4_000_000):
    y = (np.arange(K) + 0.5) / K
    P = np.concatenate(([0], np.cumsum(a) / a.sum()))
    Q = np.concatenate(([0], np.cumsum(b) / b.sum()))
    Fi = np.interp(y, P, np.arange(len(a) + 1))
    Gi = np.interp(y, Q, np.arange(len(b) + 1))
    return float(np.mean((Fi - Gi) ** 2))


def test_matches_inverse_cdf_integration_with_empty_cells():
    rng = np.random.default_rng(0)
    a = rng.random(300) + 0.01
    b = rng.random(300) ** 3
    a[50:60] = 0
    b[100:130] = 0
    b[:5] = 0
    W, _, _ = ot.w2_cells(a, b)
    assert W == pytest.approx(_inverse_cdf_w2(a, b), rel=2e-6)
    assert W == ot.w2_cells(b, a)[0]


@pytest.mark.parametrize("variant", [0, 1])
def test_identical_traces_give_zero(variant):
    rng = np.random.default_rng(1)
    s = rng.standard_normal(257)
    m, r = ot.ot_signal(s, s, variant, 0.001)
    assert m == pytest.approx(0.0, abs=1e-24)
    assert np.abs(r).max() <= 1e-12 * np.abs(s).max() * 0.001 ** 2


@pytest.mark.parametrize("variant", [0, 1])
def test_adjoint_source_is_the_misfit_derivative(variant):
    rng = np.random.default_rng(2 + variant)
    dt = 0.001
    s = rng.standard_normal(200) * 0.5
    o = rng.standard_normal(200) * 0.5
    if variant == 1:
        s[:20] = 0          # empty leading cells as in a quiet record start
        o[150:] = 0
    m0, r = ot.ot_signal(s, o, variant, dt)
    c = ot.SHIFT_FACTOR * max(np.abs(s).max(), np.abs(o).max()) if variant == 0 else None
    jmax = int(np.argmax(np.abs(s)))
    scale = np.abs(r).max()
    for j in (5, 33, 77, 120, 150, 199):
        if j == jmax:
            continue
        e = 1e-6
        sp, sm = s.copy(), s.copy()
        sp[j] += e
        sm[j] -= e
        fd = (ot.ot_signal(sp, o, variant, dt, shift=c)[0] - ot.ot_signal(sm, o, variant, dt, shift=c)[0]) / (2 * e)
        assert abs(fd - r[j]) <= 1e-6 * scale, (j, fd, r[j])


def test_gather_variants_and_zero_traces():
    rng = np.random.default_rng(4)
    ntr, nt, dt = 5, 120, 0.002
    sx, sz, ox, oz = (rng.standard_normal((ntr, nt)).astype(np.float32) for _ in range(4))
    sz[2] = 0
    oz[2] = 0                         # a dead trace: no mass under the power transform
    rx, rz, m1 = ot.ot_misfit(ox, oz, sx, sz, dt, variant=1)
    assert not rz[2].any()
    total = sum(ot.ot_signal(sx[t], ox[t], 1, dt)[0] + ot.ot_signal(sz[t], oz[t], 1, dt)[0] for t in range(ntr))
    assert m1 == pytest.approx(total, rel=1e-13)
    # the all-trace variant on a one-trace gather is the per-trace shift variant
    a = ot.ot_misfit(ox[:1], oz[:1], sx[:1], sz[:1], dt, variant=2)
    b = ot.ot_misfit(ox[:1], oz[:1], sx[:1], sz[:1], dt, variant=0)
    assert a[2] == pytest.approx(b[2], rel=1e-13)
    assert np.allclose(a[0], b[0], rtol=1e-12, atol=0) and np.allclose(a[1], b[1], rtol=1e-12, atol=0)
