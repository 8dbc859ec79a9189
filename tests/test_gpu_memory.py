"""Device memory is returned when a context is destroyed: contexts that
allocated every kind of per-shot buffer (boundary strips, velocity
snapshots, Born and second-order scattered fields, a merged solve's model
buffers) are created, used and destroyed repeatedly, and the free HBM comes
back to where it started."""
import numpy as np
import pytest

from sfwi_amd import workloads as W
from sfwi_amd.elastic import Context, Server
from tests.helpers import lib_shot

pytestmark = pytest.mark.gpu


def _free():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _use(sv, w, **kw):
    with Context(dx=sv.dx, dz=sv.dz, dt=sv.dt, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt, ntr=sv.ntr,
                 allnx=sv.allnx, allnz=sv.allnz, max_batch=2, **kw) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp, sv.vs, sv.rho)
        shots = [lib_shot(sv, s) for s in sv.shots]
        z, x = ctx.batch_forward(shots)
        obs = [lib_shot(sv, s, z[k][:len(s.gc)], x[k][:len(s.gc)]) for k, s in enumerate(sv.shots)]
        ctx.batch_gradient(obs)
        ctx.flush()
        d = (sv.vp - sv.vp0).astype(np.float32)
        ctx.set_perturbation(d, 0.5 * d, region="s1")
        if kw.get("recon") != "snapshots":
            ctx.batch_born2_forward(shots)
        ctx.hessian_vector(shots)
        ctx.flush()


def test_contexts_return_their_memory():
    sv = W.tiny(nx=200, nz=120, pml=20, nt=800, nshots=2, seed=3)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    _use(sv, w)                        # code objects and allocator caches settle
    base = _free()
    for _ in range(3):
        _use(sv, w)
        _use(sv, w, recon="snapshots", storage="f16", march_cols=8)
        _use(sv, w, storage="bf16", march_cols=8)
    assert base - _free() <= 16 << 20, (base - _free()) >> 20   # one context's snapshots alone are 42 MB
