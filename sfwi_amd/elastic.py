"""Host-side mirror of the reference's per-shot interface over the C ABI.

``Context`` is the pythonic wrapper of one ``sfwi_ctx``.  ``FD2DGPU_ELASTIC``
and the free functions ``IMAGE_SINGLESHOT``, ``CALCULATE_ALPHA_parabola_mb``
and ``CALCULATE_ALPHA`` keep the names, argument meaning and call order of the
reference (sfwi-stage1/FD2DGPU.cuh:146-472, sfwi-stage2/FD2DGPU.cuh:455-458),
so a driver written against the reference reads the same.  Every call goes to
the HIP library; errors raise ``SfwiError``.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import SfwiError, SfwiParams, SfwiShot


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Shot:
    """One shot's geometry and observed traces (host arrays kept alive here)."""

    def __init__(self, left, isx, isz, igz, ngx_left, gc, obs_z=None, obs_x=None, obs_id=0):
        self.gc = np.ascontiguousarray(gc, dtype=np.int32)
        self.obs_z = None if obs_z is None else _f32(obs_z)
        self.obs_x = None if obs_x is None else _f32(obs_x)
        self.c = SfwiShot()
        self.c.left, self.c.isx, self.c.isz, self.c.igz = int(left), int(isx), int(isz), int(igz)
        self.c.ngx_left, self.c.ntr = int(ngx_left), len(self.gc)
        self.c.gc = self.gc.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        fp = ctypes.POINTER(ctypes.c_float)
        self.c.obs_z = self.obs_z.ctypes.data_as(fp) if self.obs_z is not None else fp()
        self.c.obs_x = self.obs_x.ctypes.data_as(fp) if self.obs_x is not None else fp()
        self.c.obs_id = int(obs_id)

    @property
    def ntr(self):
        return len(self.gc)


def make_params(*, dx, dz, dt, nx, nz, pml, nt, ntr, allnx, allnz, rr=0.001, rho_const=0.0, adjoint_dir=1,
                max_batch=1, mode="fast", storage="f32", use_graph=False, march_cols=0, march_depth=0,
                misfit="l2", march_order=0, grad_rho=False, cpu_mem=False, strip_chunk=0, rbc=False,
                rbc_seed=0, recon="strips", snap_every=0, fuse_step=0, step_cols=0):
    """``sfwi_params`` of a context or request (keyword names as in the C struct)."""
    p = SfwiParams()
    p.dx, p.dz, p.dt = dx, dz, dt
    p.nx, p.nz, p.pml, p.nt, p.nop, p.ntr = nx, nz, pml, nt, 4, ntr
    p.allnx, p.allnz, p.rr, p.rho_const = allnx, allnz, rr, rho_const
    p.adjoint_dir, p.max_batch = adjoint_dir, max_batch
    p.mode = {"fast": _lib.MODE_FAST, "exact": _lib.MODE_EXACT}[mode]
    p.storage = {"f32": _lib.STORE_STRIPS_F32, "f16": _lib.STORE_STRIPS_F16, "bf16": _lib.STORE_STRIPS_BF16}[storage]
    p.use_graph = int(bool(use_graph))
    p.march_cols = int(march_cols)
    p.march_depth = int(march_depth)
    p.misfit = {"l2": _lib.MISFIT_L2, "ot_trace": _lib.MISFIT_OT_TRACE, "ot_power": _lib.MISFIT_OT_POWER,
                "ot_gather": _lib.MISFIT_OT_GATHER}[misfit]
    p.march_order = int(march_order)
    p.grad_rho = int(bool(grad_rho))
    p.cpu_mem = int(bool(cpu_mem))
    p.strip_chunk = int(strip_chunk)
    p.rbc = 1 if rbc else 0
    p.rbc_seed = int(rbc_seed) & 0xFFFFFFFF
    p.recon = {"strips": _lib.RECON_STRIPS, "snapshots": _lib.RECON_SNAPSHOTS}[recon]
    p.snap_every = int(snap_every)
    p.fuse_step = int(fuse_step)
    p.step_cols = int(step_cols)
    return p


class Context:
    """One device context (``sfwi_ctx``)."""

    def __init__(self, **kw):
        self.L = _lib.load()
        p = make_params(**kw)
        self.p = p
        h = ctypes.c_void_p()
        rc = self.L.sfwi_ctx_create(ctypes.byref(p), ctypes.byref(h))
        if rc != 0:
            raise SfwiError(rc, self.L.sfwi_last_error(None).decode())
        self.h = h
        self.nt, self.allnx, self.allnz = p.nt, p.allnx, p.allnz

    def _ck(self, rc):
        if rc != 0:
            raise SfwiError(rc, self.L.sfwi_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.sfwi_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------- inputs ---
    def set_wavelet(self, w):
        w = _f32(w)
        assert w.size == self.nt
        self._ck(self.L.sfwi_set_wavelet(self.h, w))

    def set_model(self, vp, vs, rho=None):
        self._vp, self._vs = _f32(vp), _f32(vs)
        self._rho = None if rho is None else _f32(rho)
        self._ck(self.L.sfwi_set_model(self.h, _lib.ptr(self._vp), _lib.ptr(self._vs), _lib.ptr(self._rho)))

    def obs_put(self, obs_id, obs_z, obs_x):
        """Keep a shot's observed traces on the device under obs_id > 0."""
        z, x = _f32(obs_z), _f32(obs_x)
        assert z.shape == x.shape and z.shape[-1] == self.nt
        self._ck(self.L.sfwi_obs_put(self.h, int(obs_id), int(z.size // self.nt), z, x))

    def obs_clear(self):
        self._ck(self.L.sfwi_obs_clear(self.h))

    def set_perturbation(self, dvp, dvs, region="s2"):
        """Born perturbation; region "s2" (every cell at least 4 from the grid
        edge, the stage-2 Born source) or "s1" (the interior only, stage 1)."""
        reg = {"s2": _lib.SFWI_BORN_S2, "s1": _lib.SFWI_BORN_S1}[region]
        self._ck(self.L.sfwi_set_perturbation_region(self.h, _f32(dvp), _f32(dvs), reg))

    def batch_born2_forward(self, shots):
        """Born forward with double scattering: (b_z, b_x, b2_z, b2_x), the first-
        and second-order scattered traces, arrays (nshots, ntr_max, nt)."""
        arr = (SfwiShot * len(shots))(*[s.c for s in shots])
        out = [np.zeros((len(shots), self.p.ntr, self.nt), np.float32) for _ in range(4)]
        self._ck(self.L.sfwi_batch_born2_forward(self.h, len(shots), arr, *[_lib.ptr(a) for a in out]))
        return tuple(out)

    def hessian_vector(self, shots, p_vp=None, p_vs=None):
        """Accumulate the Gauss-Newton Hessian-vector product H p of the shots
        (read it with flush()); p_vp, p_vs None: the current perturbation."""
        arr = (SfwiShot * len(shots))(*[s.c for s in shots])
        a = None if p_vp is None else _f32(p_vp)
        b = None if p_vs is None else _f32(p_vs)
        self._ck(self.L.sfwi_batch_hessian_vector(self.h, len(shots), arr, None if a is None else _lib.ptr(a),
                                                  None if b is None else _lib.ptr(b)))

    # ---------------------------------------------------------- solves ---
    def forward(self, shot):
        z = np.zeros((shot.ntr, self.nt), np.float32)
        x = np.zeros((shot.ntr, self.nt), np.float32)
        self._ck(self.L.sfwi_shot_forward(self.h, ctypes.byref(shot.c), _lib.ptr(z), _lib.ptr(x)))
        return z, x

    def batch_forward(self, shots):
        """Traces of every shot: arrays (nshots, ntr_max, nt)."""
        arr = (SfwiShot * len(shots))(*[s.c for s in shots])
        z = np.zeros((len(shots), self.p.ntr, self.nt), np.float32)
        x = np.zeros_like(z)
        self._ck(self.L.sfwi_batch_forward(self.h, len(shots), arr, _lib.ptr(z), _lib.ptr(x)))
        return z, x

    def batch_born_forward(self, shots):
        """Scattered traces of every shot: arrays (nshots, ntr_max, nt)."""
        arr = (SfwiShot * len(shots))(*[s.c for s in shots])
        z = np.zeros((len(shots), self.p.ntr, self.nt), np.float32)
        x = np.zeros_like(z)
        self._ck(self.L.sfwi_batch_born_forward(self.h, len(shots), arr, _lib.ptr(z), _lib.ptr(x)))
        return z, x

    def born_forward(self, shot):
        out = [np.zeros((shot.ntr, self.nt), np.float32) for _ in range(4)]
        self._ck(self.L.sfwi_shot_born_forward(self.h, ctypes.byref(shot.c), *[_lib.ptr(a) for a in out]))
        return tuple(out)

    def gradient(self, shot):
        m = ctypes.c_double(0.0)
        self._ck(self.L.sfwi_shot_gradient(self.h, ctypes.byref(shot.c), ctypes.byref(m)))
        return m.value

    def born_gradient(self, shot):
        m = ctypes.c_double(0.0)
        self._ck(self.L.sfwi_shot_born_gradient(self.h, ctypes.byref(shot.c), ctypes.byref(m)))
        return m.value

    def batch_gradient(self, shots, born=False):
        """Misfits of the shots; the images accumulate until flush.  A
        non-finite misfit raises SfwiError SFWI_ENUMERIC with the misfits as
        ``misfits``."""
        arr = (SfwiShot * len(shots))(*[s.c for s in shots])
        m = np.zeros(len(shots), np.float64)
        fn = self.L.sfwi_batch_born_gradient if born else self.L.sfwi_batch_gradient
        rc = fn(self.h, len(shots), arr, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc != 0:
            err = SfwiError(rc, self.L.sfwi_last_error(self.h).decode())
            err.misfits = m
            raise err
        return m

    def linesearch_parabola(self, shot, dir_vp, dir_vs, step1, step2, m=None):
        """Misfits at steps 0, step1, step2 added to m (zeros when None), as the
        reference accumulates them across shots; returns m."""
        m = np.zeros(3, np.float64) if m is None else np.array(m, np.float64)
        self._ck(self.L.sfwi_linesearch_parabola(self.h, ctypes.byref(shot.c), _f32(dir_vp), _f32(dir_vs),
                                                 step1, step2,
                                                 m.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return m

    def batch_linesearch_parabola(self, shots, dir_vp, dir_vs, step1, step2, m=None):
        """linesearch_parabola over a list of shots in one call (batched forward
        solves; equal to the per-shot calls in shot order, bitwise)."""
        m = np.zeros(3, np.float64) if m is None else np.array(m, np.float64)
        arr = (SfwiShot * len(shots))(*[s.c for s in shots])
        self._ck(self.L.sfwi_batch_linesearch_parabola(self.h, len(shots), arr, _f32(dir_vp), _f32(dir_vs),
                                                       step1, step2,
                                                       m.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return m

    def linesearch_linear(self, shot, cg_vp, cg_vs, fenzi=0.0, fenmu=0.0):
        """This shot's terms added to (fenzi, fenmu); returns the sums."""
        a, b = ctypes.c_double(fenzi), ctypes.c_double(fenmu)
        self._ck(self.L.sfwi_linesearch_linear(self.h, ctypes.byref(shot.c), _f32(cg_vp), _f32(cg_vs),
                                               ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def batch_linesearch_linear(self, shots, cg_vp, cg_vs, fenzi=0.0, fenmu=0.0):
        """linesearch_linear over a list of shots in one call (batched Born
        solves; equal to the per-shot calls in shot order, bitwise)."""
        a, b = ctypes.c_double(fenzi), ctypes.c_double(fenmu)
        arr = (SfwiShot * len(shots))(*[s.c for s in shots])
        self._ck(self.L.sfwi_batch_linesearch_linear(self.h, len(shots), arr, _f32(cg_vp), _f32(cg_vs),
                                                     ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def batch_born_gradient_res(self, shots):
        """batch_gradient(born=True) that also returns each shot's residual:
        (misfits, res_z, res_x), the residuals (nshots, ntr_max, nt)."""
        arr = (SfwiShot * len(shots))(*[s.c for s in shots])
        m = np.zeros(len(shots), np.float64)
        z = np.zeros((len(shots), self.p.ntr, self.nt), np.float32)
        x = np.zeros_like(z)
        self._ck(self.L.sfwi_batch_born_gradient_res(self.h, len(shots), arr,
                                                     m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                     _lib.ptr(z), _lib.ptr(x)))
        return m, z, x

    def batch_linesearch_linear_res(self, shots, cg_vp, cg_vs, res_z, res_x, fenzi=0.0, fenmu=0.0):
        """batch_linesearch_linear with the residuals of batch_born_gradient_res
        at the same model and images in place of its first Born solve."""
        a, b = ctypes.c_double(fenzi), ctypes.c_double(fenmu)
        arr = (SfwiShot * len(shots))(*[s.c for s in shots])
        rz = np.ascontiguousarray(res_z, np.float32)
        rx = np.ascontiguousarray(res_x, np.float32)
        assert rz.shape == (len(shots), self.p.ntr, self.nt) and rx.shape == rz.shape
        self._ck(self.L.sfwi_batch_linesearch_linear_res(self.h, len(shots), arr, _f32(cg_vp), _f32(cg_vs),
                                                         _lib.ptr(rz), _lib.ptr(rx), ctypes.byref(a),
                                                         ctypes.byref(b)))
        return a.value, b.value

    def flush(self):
        sh = (self.allnx, self.allnz)
        g = [np.zeros(sh, np.float32) for _ in range(3)]
        self._ck(self.L.sfwi_flush(self.h, *[_lib.ptr(a) for a in g]))
        return tuple(g)

    # ---------------------------------------------------------- timing ---
    def flush_rho(self):
        """Accumulated density gradient (allnx x allnz); zeroes it on the device."""
        g = np.zeros((self.allnx, self.allnz), np.float32)
        self._ck(self.L.sfwi_flush_rho(self.h, _lib.ptr(g)))
        return g

    def set_timing(self, every):
        self._ck(self.L.sfwi_set_timing(self.h, int(every)))

    def kernel_time(self, kid):
        ms, n = ctypes.c_double(0.0), ctypes.c_int(0)
        self._ck(self.L.sfwi_kernel_time(self.h, kid, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def params(self):
        """The context's current sfwi_params (including the autotuner's choices)."""
        p = SfwiParams()
        self._ck(self.L.sfwi_get_params(self.h, ctypes.byref(p)))
        return p

    def autotune(self, nshots, steps=100):
        """Pick the fastest FAST kernel family for batches of nshots; returns
        (march_cols, march_depth, march_order, best_ms)."""
        a, b, o = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        ms = ctypes.c_double(0.0)
        self._ck(self.L.sfwi_autotune(self.h, int(nshots), int(steps), ctypes.byref(a), ctypes.byref(b),
                                      ctypes.byref(o), ctypes.byref(ms)))
        return a.value, b.value, o.value, ms.value

    def ot_misfit(self, obs_x, obs_z, syn_x, syn_z, dt, variant=0):
        obs_x, obs_z, syn_x, syn_z = map(_f32, (obs_x, obs_z, syn_x, syn_z))
        ntr, nt = obs_x.shape
        rx, rz = np.zeros_like(obs_x), np.zeros_like(obs_z)
        m = ctypes.c_double(0.0)
        self._ck(self.L.sfwi_ot_misfit(self.h, variant, ntr, nt, dt, obs_x, obs_z, syn_x, syn_z, rx, rz,
                                       ctypes.byref(m)))
        return m.value, rx, rz


def check_model(vp, vs, rho=None, rho_const=0.0):
    """sfwi_check_model on (allnx, allnz) arrays: raise SfwiError (SFWI_EINVAL)
    naming the first cell of the model that the propagator cannot use.
    Needs no GPU."""
    L = _lib.load()
    vp, vs = _f32(vp), _f32(vs)
    if vs.shape != vp.shape or (rho is not None and np.shape(rho) != vp.shape):
        raise ValueError("vp, vs and rho must have one shape")
    p = SfwiParams()
    p.allnx, p.allnz = vp.shape
    p.rho_const = rho_const
    r = None if rho is None else _f32(rho)
    rc = L.sfwi_check_model(ctypes.byref(p), _lib.ptr(vp), _lib.ptr(vs), _lib.ptr(r))
    if rc != 0:
        raise SfwiError(rc, L.sfwi_last_error(None).decode())


class Server:
    """Serving loop (``sfwi_server_*``): submit forward / Born / gradient
    requests of any shape; they run on a pool of workers with their own
    contexts, admitted against an HBM budget, and overlap on the GPU.  The
    arrays of a request are kept alive here until ``wait`` returns."""

    KINDS = {"forward": _lib.REQ_FORWARD, "born_forward": _lib.REQ_BORN_FORWARD,
             "gradient": _lib.REQ_GRADIENT, "born_gradient": _lib.REQ_BORN_GRADIENT}

    def __init__(self, budget_bytes=0, workers=4, batching=True):
        self.L = _lib.load()
        h = ctypes.c_void_p()
        rc = self.L.sfwi_server_create(int(budget_bytes), int(workers), ctypes.byref(h))
        if rc != 0:
            raise SfwiError(rc, self.L.sfwi_last_error(None).decode())
        self.h = h
        self._live = {}
        self.set_batching(batching)

    def set_batching(self, on):
        """Cross-request batching: queued requests of one shape and kind join a
        worker's batch (up to max_batch shots)."""
        rc = self.L.sfwi_server_batching(self.h, int(bool(on)))
        if rc != 0:
            raise SfwiError(rc, self.L.sfwi_last_error(None).decode())

    @staticmethod
    def solve_bytes(kind, nshots, **params):
        return _lib.load().sfwi_solve_bytes(ctypes.byref(make_params(**params)), Server.KINDS[kind], nshots)

    def submit(self, kind, shots, wavelet, vp, vs, rho=None, dvp=None, dvs=None, autotune=False, **params):
        params.setdefault("max_batch", len(shots))
        p = make_params(**params)
        q = _lib.SfwiRequest()
        q.kind = self.KINDS[kind]
        q.params = p
        q.autotune = int(bool(autotune))
        keep = [_f32(wavelet), _f32(vp), _f32(vs), None if rho is None else _f32(rho),
                None if dvp is None else _f32(dvp), None if dvs is None else _f32(dvs)]
        q.wavelet, q.vp, q.vs, q.rho, q.dvp, q.dvs = [_lib.ptr(a) for a in keep]
        arr = (SfwiShot * len(shots))(*[s.c for s in shots])
        q.nshots, q.shots = len(shots), arr
        out = {}
        if kind in ("forward", "born_forward"):
            out["syn_z"] = np.zeros((len(shots), p.ntr, p.nt), np.float32)
            out["syn_x"] = np.zeros_like(out["syn_z"])
            q.syn_z, q.syn_x = _lib.ptr(out["syn_z"]), _lib.ptr(out["syn_x"])
        else:
            n = p.allnx * p.allnz
            out["misfits"] = np.zeros(len(shots), np.float64)
            out["grad_vp"], out["grad_vs"], out["illum"] = (np.zeros((p.allnx, p.allnz), np.float32) for _ in range(3))
            q.misfits = _lib.ptr(out["misfits"])
            q.grad_vp, q.grad_vs, q.illum = (_lib.ptr(out[k]) for k in ("grad_vp", "grad_vs", "illum"))
            del n
        t = ctypes.c_longlong(0)
        rc = self.L.sfwi_server_submit(self.h, ctypes.byref(q), ctypes.byref(t))
        if rc != 0:
            raise SfwiError(rc, self.L.sfwi_last_error(None).decode())
        self._live[t.value] = (q, keep, arr, list(shots), out)
        return t.value

    def wait(self, ticket, timeout=None):
        """The request's outputs once it is done.  timeout (seconds): raise
        SfwiError SFWI_ETIMEOUT if it is not done by then (the ticket stays
        valid).  A failed request raises its error; the outputs of a gradient
        request with a non-finite misfit (SFWI_ENUMERIC) are on the error as
        ``outputs``."""
        qm, rm = ctypes.c_double(0.0), ctypes.c_double(0.0)
        if timeout is None:
            rc = self.L.sfwi_server_wait(self.h, ticket, ctypes.byref(qm), ctypes.byref(rm))
        else:
            rc = self.L.sfwi_server_wait_for(self.h, ticket, 1e3 * float(timeout), ctypes.byref(qm),
                                             ctypes.byref(rm))
            if rc == _lib.SFWI_ETIMEOUT:
                raise SfwiError(rc, self.L.sfwi_last_error(None).decode())
        rec = self._live.pop(ticket)
        out = rec[4]
        out["queue_ms"], out["run_ms"] = qm.value, rm.value
        if rc != 0:
            err = SfwiError(rc, self.L.sfwi_last_error(None).decode())
            err.outputs = out
            raise err
        return out

    def cancel(self, ticket):
        """Cancel a request that has not started; waiting for it then raises
        SFWI_ECANCELED."""
        rc = self.L.sfwi_server_cancel(self.h, ticket)
        if rc != 0:
            raise SfwiError(rc, self.L.sfwi_last_error(None).decode())

    def close(self):
        if getattr(self, "h", None):
            for t in list(self._live):
                try:
                    self.wait(t)
                except SfwiError:
                    pass
            self.L.sfwi_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------
# Reference-shaped interface
# --------------------------------------------------------------------------

class FD2DGPU_ELASTIC:
    """Mirror of class FD2DGPU_ELASTIC (sfwi-stage1/FD2DGPU.cuh:146-459).

    The constructor takes the reference's arguments (sou, dx, dz, dt, nxpml,
    nzpml, allnx, allnz, scale, pml, nt, nop, ntr); the caller then sets the
    public per-shot fields (isx, isz, igz, ngx_left, delta_left,
    record_left_in_v) exactly as sfwi-stage1/main.cpp:596-625 does.
    """

    def __init__(self, sou, dx, dz, dt, nxpml, nzpml, allnx, allnz, scale, pml, nt, nop, ntr, **kw):
        if nop != 4:
            raise ValueError("nop must be 4")
        self.dx, self.dz, self.dt, self.pml, self.nt = dx, dz, dt, pml, nt
        self.nxpml, self.nzpml, self.allnx, self.allnz = nxpml, nzpml, allnx, allnz
        self.nx, self.nz = nxpml - 2 * pml, nzpml - 2 * pml
        self.ntr = ntr
        self.ctx = Context(dx=dx, dz=dz, dt=dt, nx=self.nx, nz=self.nz, pml=pml, nt=nt, ntr=ntr,
                           allnx=allnx, allnz=allnz, **kw)
        self.ctx.set_wavelet(sou)
        self.isx = self.isz = self.igz = 0
        self.ngx_left = self.delta_left = self.record_left_in_v = 0
        self.cmin = 0
        self._rec = None
        self._left = 0

    def GPUbufferVPVS(self, vp_all, vs_all, rho_all=None):
        self.ctx.set_model(vp_all, vs_all, rho_all)

    def GPUbufferM(self, pp, ps):
        self.ctx.set_perturbation(pp, ps)

    def record_copytoGPU(self, record, record_x, gc):
        self._rec = (_f32(record), _f32(record_x), np.ascontiguousarray(gc, np.int32))

    def bufferVpVsHtoD(self, ngx_min):
        self._left = int(ngx_min)

    def _shot(self):
        rz, rx, gc = self._rec
        return Shot(self._left, self.isx, self.isz, self.igz, self.ngx_left, gc[: self.ntr],
                    rz.reshape(-1)[: self.ntr * self.nt].reshape(self.ntr, self.nt),
                    rx.reshape(-1)[: self.ntr * self.nt].reshape(self.ntr, self.nt))

    def imagebuffer_resettozero(self):
        """Returns (pp_gradient, ps_gradient, illumination) and zeroes them."""
        return self.ctx.flush()


def IMAGE_SINGLESHOT(fd, misfit=0.0, born=False):
    """Per-shot gradient (sfwi-stage1/FD2DGPU.cu:4321-4695; stage 2 with born=True,
    sfwi-stage2/FD2DGPU.cu:3978-4570).  Returns misfit + this shot's misfit."""
    s = fd._shot()
    return misfit + (fd.ctx.born_gradient(s) if born else fd.ctx.gradient(s))


def CALCULATE_ALPHA_parabola_mb(fd, cpu_pp_grad_old, cpu_ps_grad_old, step_1, step_2, m=(0.0, 0.0, 0.0)):
    """Three forward solves of the stage-1 line search (sfwi-stage1/FD2DGPU.cu:4819-5028)."""
    r = fd.ctx.linesearch_parabola(fd._shot(), cpu_pp_grad_old, cpu_ps_grad_old, step_1, step_2)
    return tuple(float(a + b) for a, b in zip(m, r))


def CALCULATE_ALPHA(fd, cpu_pp_grad_old, cpu_ps_grad_old, fenzi=0.0, fenmu=0.0):
    """Two Born solves of the stage-2 exact step (sfwi-stage2/FD2DGPU.cu:4576-4722)."""
    a, b = fd.ctx.linesearch_linear(fd._shot(), cpu_pp_grad_old, cpu_ps_grad_old)
    return fenzi + a, fenmu + b
