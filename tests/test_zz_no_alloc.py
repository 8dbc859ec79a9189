"""The allocation-free OT entry point (sfwi_ot_misfit_static) through the C ABI:
two calls in a row -- the second with a smaller gather and another variant --
give the misfit and adjoint sources of sfwi_ot_misfit bit for bit (the same
kernel on the same job layout; only its workspace differs), and a gather past
the static workspace's limit is refused."""
import ctypes

import numpy as np
import pytest

from sfwi_amd import _lib
from sfwi_amd import workloads as W
from tests.helpers import context

pytestmark = pytest.mark.gpu


def _gather(ntr, nt, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(nt)
    out = []
    for _ in range(4):
        a = np.zeros((ntr, nt), np.float32)
        for tr in range(ntr):
            t0 = rng.uniform(0.1, 0.6) * nt
            a[tr] = (np.sin(0.05 * (t - t0)) * np.exp(-((t - t0) / (0.08 * nt)) ** 2)).astype(np.float32)
        out.append(a + np.float32(1e-3) * rng.standard_normal((ntr, nt)).astype(np.float32))
    return out   # obs_x, obs_z, syn_x, syn_z


def _static(ctx, variant, g, dt):
    L = _lib.load()
    ox, oz, sx, sz = g
    rx, rz = np.zeros_like(sx), np.zeros_like(sz)
    m = ctypes.c_double(0.0)
    rc = L.sfwi_ot_misfit_static(ctx.h, variant, sx.shape[0], sx.shape[1], dt, ox, oz, sx, sz, rx, rz,
                                 ctypes.byref(m))
    if rc != 0:
        raise _lib.SfwiError(rc, L.sfwi_last_error(ctx.h).decode())
    return m.value, rx, rz


def test_static_ot_misfit_equals_the_allocating_entry_point():
    dt = float(np.float32(0.001))
    with context(W.tiny()) as ctx:
        for (ntr, nt, variant, seed) in ((48, 1500, 0, 1), (7, 300, 2, 2), (20, 700, 1, 3)):
            g = _gather(ntr, nt, seed)
            m_ref, rx_ref, rz_ref = ctx.ot_misfit(*g, dt, variant)
            m, rx, rz = _static(ctx, variant, g, dt)
            assert m == m_ref, (ntr, nt, variant)
            assert np.array_equal(rx, rx_ref) and np.array_equal(rz, rz_ref), (ntr, nt, variant)
            assert m > 0 and np.abs(rx).max() > 0
        big = _gather(64, 4097, 4)   # 262,208 samples per component: one past the limit
        with pytest.raises(_lib.SfwiError, match="static workspace"):
            _static(ctx, 0, big, dt)
