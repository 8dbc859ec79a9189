"""The library's propagator against physics: the homogeneous acoustic limit
(Vs = 0) of the FAST column-march, FAST per-cell and EXACT kernels matches
the analytic 2-D Green's function to the scheme's second-order time error,
as the FP64 oracle does (tests/test_oracle_green.py derives the solution:
the wavelet acts as an instantaneous stress on one cell, vx / vz at their
half-cell positions, the record of step it at (it + 1/2) dt)."""
import numpy as np
import pytest

from sfwi_amd.elastic import Context, Shot
from tests.test_oracle_green import RHO, VP, _analytic, _ricker

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,cols", [("fast", 8), ("fast", -1), ("exact", 0)])
def test_acoustic_limit_matches_the_green_function(mode, cols):
    dx, dt, L = 10.0, 0.0005, 800.0
    n, pml = int(L / dx), int(100 / dx)
    nt = int(round(0.4 / dt))
    isx = isz = pml + n // 2
    igz = isz - int(150 / dx)
    gc = (np.arange(-200, 201, 50) + L / 2).astype(np.int32)
    dt32 = float(np.float32(dt))
    w = _ricker(np.arange(nt) * dt32).astype(np.float32)
    vp = np.full((n, n), VP, np.float32)
    with Context(dx=dx, dz=dx, dt=dt, nx=n, nz=n, pml=pml, nt=nt, ntr=len(gc), allnx=n, allnz=n, mode=mode,
                 max_batch=1, march_cols=cols) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(vp, np.zeros_like(vp), np.full_like(vp, RHO))
        oz, ox = ctx.batch_forward([Shot(0, isx, isz, igz, 0, gc)])
    rz, rx = oz[0][:len(gc)], ox[0][:len(gc)]
    times = (np.arange(nt) + 0.5) * dt32
    num, ana = [], []
    for k, g in enumerate(gc):
        ax, az = _analytic(dx, isx * dx, isz * dx, (pml + int(g / dx)) * dx, igz * dx, times)
        num += [rx[k], rz[k]]
        ana += [ax, az]
    num, ana = np.array(num, np.float64), np.array(ana)
    err = np.linalg.norm(num - ana) / np.linalg.norm(ana)
    assert err <= 3e-3, err
