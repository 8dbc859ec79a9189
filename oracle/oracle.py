"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the sfwi_amd package.  The C++ source
(sfwi_oracle.cpp) restates the reference's per-shot algorithm; see the header
sfwi_oracle.h for the meaning of every entry point.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class OrConfig(ctypes.Structure):
    """Mirror of `or_config` in sfwi_oracle.h."""

    _fields_ = [
        ("allnx", ctypes.c_int), ("allnz", ctypes.c_int),
        ("nx", ctypes.c_int), ("nz", ctypes.c_int),
        ("pml", ctypes.c_int), ("nt", ctypes.c_int), ("ntr", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float),
        ("rr", ctypes.c_float),
        ("left", ctypes.c_int),
        ("isx", ctypes.c_int), ("isz", ctypes.c_int), ("igz", ctypes.c_int),
        ("ngx_left", ctypes.c_int),
        ("adjoint_dir", ctypes.c_int),
        ("rho_const", ctypes.c_float),
        ("pml_off", ctypes.c_int),
        ("nthreads", ctypes.c_int),
        ("rbc", ctypes.c_int),
        ("rbc_seed", ctypes.c_uint),
        ("snap_every", ctypes.c_int),
        ("plant", ctypes.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        cfg = ctypes.POINTER(OrConfig)
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_coeffs_f32.argtypes = [_f32p]
        L.or_coeffs_f64.argtypes = [_f64p]
        L.or_ricker.argtypes = [_f32p, ctypes.c_int, ctypes.c_float, ctypes.c_float]
        L.or_setup_f32.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p]
        L.or_setup_f64.argtypes = [cfg, _f32p, _f32p, _f32p, _f64p]
        L.or_forward_f32.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _i32p, _f32p, _f32p]
        L.or_forward_f64.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _i32p, _f64p, _f64p]
        L.or_gradient_f32.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _i32p, _f32p, _f32p,
                                      _f32p, _f32p, _f32p, dp, _f32p, _f32p]
        L.or_gradient_f64.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _i32p, _f32p, _f32p,
                                      _f64p, _f64p, _f64p, dp, _f64p, _f64p]
        L.or_gradient_rho_f32.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _i32p, _f32p, _f32p,
                                          _f32p, _f32p, _f32p, _f32p, dp]
        L.or_gradient_rho_f64.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _i32p, _f32p, _f32p,
                                          _f64p, _f64p, _f64p, _f64p, dp]
        L.or_gradient_res_f32.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _i32p, _f32p, _f32p,
                                          _f32p, _f32p, _f32p]
        L.or_gradient_res_f64.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _i32p, _f32p, _f32p,
                                          _f64p, _f64p, _f64p]
        L.or_gradient_consistent_f64.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _i32p, _f32p, _f32p,
                                                 _f64p, _f64p, _f64p, dp]
        _nf32 = ctypes.c_void_p   # nullable float arrays
        L.or_gradient_snapshot_f64.argtypes = [cfg, _f32p, _f32p, _f32p, _nf32, _nf32, _f32p, _i32p, _f32p, _f32p,
                                               _nf32, _nf32, _f64p, _f64p, _f64p, dp]
        L.or_reconstruct_f64.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, _f64p]
        L.or_born_forward_f32.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p,
                                          _f32p, _f32p, _f32p, _f32p]
        L.or_born_forward_f64.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p,
                                          _f64p, _f64p, _f64p, _f64p]
        L.or_born_gradient_f32.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p,
                                           _f32p, _f32p, _f32p, _f32p, _f32p, dp]
        L.or_born_gradient_f64.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p,
                                           _f32p, _f32p, _f64p, _f64p, _f64p, dp]
        for prec, fp in (("f32", _f32p), ("f64", _f64p)):
            getattr(L, "or_born2_forward_" + prec).argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                                               _i32p, fp, fp, fp, fp]
            getattr(L, "or_hessian_vector_" + prec).argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                                                _i32p, fp, fp]
        L.or_hessian_vector_snapshot_f64.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p,
                                                     _f64p, _f64p]
        L.or_parabola_misfits_f32.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _f32p,
                                              ctypes.c_float, ctypes.c_float, _f32p, _i32p,
                                              _f32p, _f32p, _f64p]
        L.or_linear_alpha_f32.argtypes = [cfg, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                          _f32p, _i32p, _f32p, _f32p, dp, dp]
        L.or_stencils_f64.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.or_propagate_f64.argtypes = [cfg, _f32p, _f32p, _f32p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, _f64p]
        _LIB = L
    return _LIB


def config(**kw):
    """Build an OrConfig; unspecified fields take the reference defaults."""
    c = OrConfig()
    c.rr = 0.001
    c.adjoint_dir = 1
    c.rho_const = 0.0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def ricker(nt, dt, f0):
    w = np.zeros(nt, np.float32)
    lib().or_ricker(w, nt, dt, f0)
    return w


def coeffs(prec="f64"):
    if prec == "f32":
        c = np.zeros(5, np.float32)
        lib().or_coeffs_f32(c)
    else:
        c = np.zeros(5, np.float64)
        lib().or_coeffs_f64(c)
    return c


def setup(c, vp, vs, rho, prec="f32"):
    """Padded (vp, vs, rho, lam, mu, dampx, dampz), each (nxpml, nzpml)."""
    nxp, nzp = c.nx + 2 * c.pml, c.nz + 2 * c.pml
    dt = np.float32 if prec == "f32" else np.float64
    out = np.zeros(7 * nxp * nzp, dt)
    getattr(lib(), "or_setup_" + prec)(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), out)
    return out.reshape(7, nxp, nzp)


def forward(c, vp, vs, rho, wav, gc, prec="f32"):
    dt = np.float32 if prec == "f32" else np.float64
    sz = np.zeros((c.ntr, c.nt), dt)
    sx = np.zeros((c.ntr, c.nt), dt)
    getattr(lib(), "or_forward_" + prec)(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), _f32(wav),
                                         np.ascontiguousarray(gc, np.int32), sz, sx)
    return sz, sx


def gradient(c, vp, vs, rho, wav, gc, obs_z, obs_x, prec="f32"):
    """Stage-1 per-shot gradient: returns (grad_vp, grad_vs, illum, misfit, res_z, res_x)."""
    dt = np.float32 if prec == "f32" else np.float64
    n = c.allnx * c.allnz
    g1, g2, g3 = np.zeros(n, dt), np.zeros(n, dt), np.zeros(n, dt)
    rz = np.zeros((c.ntr, c.nt), dt)
    rx = np.zeros((c.ntr, c.nt), dt)
    m = ctypes.c_double(0.0)
    getattr(lib(), "or_gradient_" + prec)(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), _f32(wav),
                                          np.ascontiguousarray(gc, np.int32), _f32(obs_z), _f32(obs_x),
                                          g1, g2, g3, ctypes.byref(m), rz, rx)
    sh = (c.allnx, c.allnz)
    return g1.reshape(sh), g2.reshape(sh), g3.reshape(sh), m.value, rz, rx


def gradient_rho(c, vp, vs, rho, wav, gc, obs_z, obs_x, prec="f32"):
    """Stage-1 gradient with the density image: (grad_vp, grad_vs, illum, grad_rho, misfit)."""
    dt = np.float32 if prec == "f32" else np.float64
    n = c.allnx * c.allnz
    g = [np.zeros(n, dt) for _ in range(4)]
    m = ctypes.c_double(0.0)
    getattr(lib(), "or_gradient_rho_" + prec)(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), _f32(wav),
                                              np.ascontiguousarray(gc, np.int32), _f32(obs_z), _f32(obs_x),
                                              *g, ctypes.byref(m))
    sh = (c.allnx, c.allnz)
    return g[0].reshape(sh), g[1].reshape(sh), g[2].reshape(sh), g[3].reshape(sh), m.value


def gradient_consistent(c, vp, vs, rho, wav, gc, obs_z, obs_x):
    """Exact discrete gradient of sum r^2 (FP64, use with pml_off=1): (grad_vp, grad_vs, grad_rho, misfit)."""
    n = c.allnx * c.allnz
    g = [np.zeros(n, np.float64) for _ in range(3)]
    m = ctypes.c_double(0.0)
    lib().or_gradient_consistent_f64(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), _f32(wav),
                                     np.ascontiguousarray(gc, np.int32), _f32(obs_z), _f32(obs_x), *g,
                                     ctypes.byref(m))
    sh = (c.allnx, c.allnz)
    return g[0].reshape(sh), g[1].reshape(sh), g[2].reshape(sh), m.value


def gradient_snapshot(c, vp, vs, rho, wav, gc, obs_z, obs_x, dvp=None, dvs=None, adj_z=None, adj_x=None):
    """Snapshot gradient (FP64), imaging every c.snap_every-th step:
    (grad_vp, grad_vs, illum, misfit).  dvp/dvs: stage-2 Born forward;
    adj_z/adj_x: adjoint source of another misfit in place of 2 r."""
    n = c.allnx * c.allnz
    g = [np.zeros(n, np.float64) for _ in range(3)]
    m = ctypes.c_double(0.0)
    keep = [None if a is None else _f32(a) for a in (dvp, dvs, adj_z, adj_x)]
    ptr = [None if a is None else a.ctypes.data for a in keep]
    lib().or_gradient_snapshot_f64(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), ptr[0], ptr[1], _f32(wav),
                                   np.ascontiguousarray(gc, np.int32), _f32(obs_z), _f32(obs_x), ptr[2], ptr[3],
                                   *g, ctypes.byref(m))
    sh = (c.allnx, c.allnz)
    return g[0].reshape(sh), g[1].reshape(sh), g[2].reshape(sh), m.value


def reconstruct(c, vp, vs, rho, wav, kind, init=None, errs=False):
    """The faithful gradient's forward loop and reverse reconstruction (test
    device, FP64): (state after reverse step 0 [15, nxpml, nzpml], errs or None)."""
    nxp, nzp = c.nx + 2 * c.pml, c.nz + 2 * c.pml
    out = np.zeros(15 * nxp * nzp, np.float64)
    e = np.zeros(c.nt, np.float64) if errs else None
    ini = None if init is None else np.ascontiguousarray(init, np.float64)
    lib().or_reconstruct_f64(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), _f32(wav),
                             None if ini is None else ini.ctypes.data, int(kind),
                             None if e is None else e.ctypes.data, out)
    return out.reshape(15, nxp, nzp), e


def gradient_res(c, vp, vs, rho, wav, gc, res_z, res_x, prec="f32"):
    """Gradient (FP32 order, or FP64) of a misfit whose adjoint source is res_z / res_x."""
    n = c.allnx * c.allnz
    dt = np.float32 if prec == "f32" else np.float64
    g1, g2, g3 = np.zeros(n, dt), np.zeros(n, dt), np.zeros(n, dt)
    getattr(lib(), "or_gradient_res_" + prec)(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), _f32(wav),
                                              np.ascontiguousarray(gc, np.int32), _f32(res_z), _f32(res_x),
                                              g1, g2, g3)
    sh = (c.allnx, c.allnz)
    return g1.reshape(sh), g2.reshape(sh), g3.reshape(sh)


def born_forward(c, vp, vs, rho, dvp, dvs, wav, gc, prec="f32"):
    dt = np.float32 if prec == "f32" else np.float64
    out = [np.zeros((c.ntr, c.nt), dt) for _ in range(4)]
    getattr(lib(), "or_born_forward_" + prec)(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho),
                                              _f32(dvp), _f32(dvs), _f32(wav),
                                              np.ascontiguousarray(gc, np.int32), *out)
    return tuple(out)


def born_gradient(c, vp, vs, rho, dvp, dvs, wav, gc, obs_z, obs_x, prec="f32"):
    dt = np.float32 if prec == "f32" else np.float64
    n = c.allnx * c.allnz
    g1, g2, g3 = np.zeros(n, dt), np.zeros(n, dt), np.zeros(n, dt)
    m = ctypes.c_double(0.0)
    getattr(lib(), "or_born_gradient_" + prec)(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho),
                                               _f32(dvp), _f32(dvs), _f32(wav),
                                               np.ascontiguousarray(gc, np.int32),
                                               _f32(obs_z), _f32(obs_x), g1, g2, g3, ctypes.byref(m))
    sh = (c.allnx, c.allnz)
    return g1.reshape(sh), g2.reshape(sh), g3.reshape(sh), m.value


def born2_forward(c, vp, vs, rho, dvp, dvs, wav, gc, prec="f32"):
    """Stage-1 Born forward with double scattering: (b_z, b_x, b2_z, b2_x),
    the first- and second-order scattered traces."""
    dt = np.float32 if prec == "f32" else np.float64
    out = [np.zeros((c.ntr, c.nt), dt) for _ in range(4)]
    getattr(lib(), "or_born2_forward_" + prec)(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho),
                                               _f32(dvp), _f32(dvs), _f32(wav),
                                               np.ascontiguousarray(gc, np.int32), *out)
    return tuple(out)


def hessian_vector(c, vp, vs, rho, p_vp, p_vs, wav, gc, prec="f32"):
    """Gauss-Newton Hessian-vector product of one shot (the reference's
    strip reconstruction): (Hp_vp, Hp_vs) on the allnx x allnz model."""
    dt = np.float32 if prec == "f32" else np.float64
    n = c.allnx * c.allnz
    g1, g2 = np.zeros(n, dt), np.zeros(n, dt)
    getattr(lib(), "or_hessian_vector_" + prec)(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), _f32(p_vp),
                                                _f32(p_vs), _f32(wav), np.ascontiguousarray(gc, np.int32), g1, g2)
    sh = (c.allnx, c.allnz)
    return g1.reshape(sh), g2.reshape(sh)


def hessian_vector_snapshot(c, vp, vs, rho, p_vp, p_vs, wav, gc):
    """Hessian-vector product through the snapshot (consistent) gradient,
    FP64, imaging every c.snap_every-th step: (Hp_vp, Hp_vs)."""
    n = c.allnx * c.allnz
    g1, g2 = np.zeros(n, np.float64), np.zeros(n, np.float64)
    lib().or_hessian_vector_snapshot_f64(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), _f32(p_vp), _f32(p_vs),
                                         _f32(wav), np.ascontiguousarray(gc, np.int32), g1, g2)
    sh = (c.allnx, c.allnz)
    return g1.reshape(sh), g2.reshape(sh)


def parabola_misfits(c, vp, vs, rho, dir_vp, dir_vs, step1, step2, wav, gc, obs_z, obs_x):
    m = np.zeros(3, np.float64)
    lib().or_parabola_misfits_f32(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), _f32(dir_vp),
                                  _f32(dir_vs), step1, step2, _f32(wav),
                                  np.ascontiguousarray(gc, np.int32), _f32(obs_z), _f32(obs_x), m)
    return m


def linear_alpha(c, vp, vs, rho, dvp, dvs, cg_vp, cg_vs, wav, gc, obs_z, obs_x):
    fz, fm = ctypes.c_double(0.0), ctypes.c_double(0.0)
    lib().or_linear_alpha_f32(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), _f32(dvp), _f32(dvs),
                              _f32(cg_vp), _f32(cg_vs), _f32(wav), np.ascontiguousarray(gc, np.int32),
                              _f32(obs_z), _f32(obs_x), ctypes.byref(fz), ctypes.byref(fm))
    return fz.value, fm.value


def stencils(f):
    f = np.ascontiguousarray(f, np.float64)
    nxp, nzp = f.shape
    outs = [np.zeros_like(f) for _ in range(4)]
    lib().or_stencils_f64(nxp, nzp, f, *outs)
    return outs  # dxp, dxm, dzp, dzm


STATE_NAMES = ("vx", "vz", "txx", "tzz", "txz", "vx_x", "vx_z", "vz_x", "vz_z",
               "txx_x", "txx_z", "tzz_x", "tzz_z", "txz_x", "txz_z")


def propagate(c, vp, vs, rho, state, kind, direction, nsteps):
    """Free propagation of a 15-array state (see STATE_NAMES), FP64."""
    s = np.ascontiguousarray(state, np.float64).copy()
    lib().or_propagate_f64(ctypes.byref(c), _f32(vp), _f32(vs), _f32(rho), kind, direction, nsteps, s)
    return s
