"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/sfwi_abi.h, rejects bad parameters before touching the
device, and its host-side Ricker helper matches the oracle bit for bit."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from sfwi_amd import _lib


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = _lib.declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in _lib.load().sfwi_version()


@pytest.mark.parametrize("nt,dt,f0", [(2000, 0.001, 10.0), (3000, 0.001, 12.0), (6000, 0.0005, 15.0),
                                      (300, 0.001, 25.0), (1500, 0.001, 20.0)])
def test_ricker_matches_reference_arithmetic(nt, dt, f0):
    assert np.array_equal(_lib.ricker(nt, dt, f0), O.ricker(nt, dt, f0))


def _params(**kw):
    p = _lib.SfwiParams()
    p.dx, p.dz, p.dt = 10.0, 10.0, 0.001
    p.nx, p.nz, p.pml, p.nt, p.nop, p.ntr = 48, 40, 8, 100, 4, 48
    p.allnx, p.allnz, p.rr, p.rho_const = 48, 40, 0.001, 0.0
    p.adjoint_dir, p.max_batch, p.mode, p.storage, p.use_graph = 1, 1, 0, 0, 0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("bad", [dict(nop=6), dict(pml=4), dict(nz=30), dict(adjoint_dir=0),
                                 dict(dt=0.0), dict(max_batch=0), dict(mode=7), dict(nx=100),
                                 dict(cpu_mem=2), dict(cpu_mem=1, strip_chunk=-1)])
def test_create_rejects_bad_parameters(bad):
    L = _lib.load()
    h = ctypes.c_void_p()
    rc = L.sfwi_ctx_create(ctypes.byref(_params(**bad)), ctypes.byref(h))
    assert rc == _lib.SFWI_EINVAL
    assert not h.value
    assert L.sfwi_last_error(None)


@pytest.mark.parametrize("fuse,recon", [(2, 0), (3, 1)])
def test_create_rejects_workgroup_marches_over_4gib(fuse, recon):
    """The workgroup fused steps address a shot's state block with 32-bit
    byte offsets: a forced fuse_step 2 or 3 on a grid whose block exceeds
    4 GiB (5000 x 5000 cells with a 40-cell PML: 25.8 M cells, 43 or 52
    fields) is refused before the device is touched."""
    L = _lib.load()
    h = ctypes.c_void_p()
    p = _params(nx=5000, nz=5000, allnx=5000, allnz=5000, pml=40, fuse_step=fuse, recon=recon, mode=0)
    rc = L.sfwi_ctx_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == _lib.SFWI_EINVAL
    assert not h.value
    assert b"4 GiB" in L.sfwi_last_error(None)


def _model(nx=30, nz=20):
    vp = np.full((nx, nz), 2500.0, np.float32)
    return vp, vp / 1.8, np.full_like(vp, 2000.0)


@pytest.mark.parametrize("field,cell,value,rho_const,ok", [
    (None, None, None, 0.0, True),
    ("vp", (3, 7), np.nan, 0.0, False), ("vp", (0, 0), 0.0, 0.0, False), ("vp", (29, 19), -1.0, 0.0, False),
    ("vp", (5, 5), np.inf, 0.0, False), ("vs", (12, 2), -np.inf, 0.0, False), ("vs", (1, 1), -0.5, 0.0, False),
    ("vs", (2, 3), 0.0, 0.0, True),                       # a fluid cell (Vs = 0) is allowed
    ("rho", (4, 4), np.nan, 0.0, False), ("rho", (4, 4), 0.0, 0.0, False),
    ("rho", (4, 4), np.nan, 2000.0, True)])               # rho_const replaces the file's density
def test_check_model(field, cell, value, rho_const, ok):
    """sfwi_check_model refuses a model the propagator cannot use and names
    its first bad cell; it needs no GPU."""
    from sfwi_amd.elastic import check_model
    vp, vs, rho = _model()
    if field:
        {"vp": vp, "vs": vs, "rho": rho}[field][cell] = value
    if ok:
        check_model(vp, vs, rho, rho_const=rho_const)
        return
    with pytest.raises(_lib.SfwiError) as e:
        check_model(vp, vs, rho, rho_const=rho_const)
    assert e.value.code == _lib.SFWI_EINVAL
    assert "%s = " % field in str(e.value) and "(ix %d, iz %d)" % cell in str(e.value), str(e.value)


def test_check_model_without_density():
    from sfwi_amd.elastic import check_model
    vp, vs, _ = _model()
    check_model(vp, vs, None, rho_const=2000.0)
    with pytest.raises(_lib.SfwiError, match="rho_const"):
        check_model(vp, vs, None)
