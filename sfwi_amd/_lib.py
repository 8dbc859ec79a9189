"""ctypes binding of libsfwi_amd.so (the HIP library behind include/sfwi_abi.h).

The library is built in-tree by ``__graft_entry__.build()``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsfwi_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "sfwi_abi.h")

SFWI_OK, SFWI_EINVAL, SFWI_EHIP, SFWI_ENOMEM, SFWI_ESTATE = 0, -1, -2, -3, -4
SFWI_ENUMERIC, SFWI_ECANCELED, SFWI_ETIMEOUT = -5, -6, -7
MODE_FAST, MODE_EXACT = 0, 1
STORE_STRIPS_F32, STORE_STRIPS_F16, STORE_STRIPS_BF16 = 0, 1, 2
MISFIT_L2, MISFIT_OT_TRACE, MISFIT_OT_POWER, MISFIT_OT_GATHER = 0, 1, 2, 3
SFWI_BORN_S2, SFWI_BORN_S1 = 0, 1
REQ_FORWARD, REQ_BORN_FORWARD, REQ_GRADIENT, REQ_BORN_GRADIENT = 0, 1, 2, 3
RECON_STRIPS, RECON_SNAPSHOTS = 0, 1
(K_FWD_VEL, K_FWD_STR, K_REC_STR, K_REC_VEL, K_ADJ_STR, K_ADJ_VEL, K_REC_STR_IMG, K_REC_VEL_ILL, K_ADJ_VEL_ONLY,
 K_SNAP_IMAGE, K_SNAP_STORE, K_FWD_STEP, K_ADJ_STEP, K_COUNT) = range(14)


class SfwiParams(ctypes.Structure):
    _fields_ = [
        ("dx", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float),
        ("nx", ctypes.c_int), ("nz", ctypes.c_int),
        ("pml", ctypes.c_int), ("nt", ctypes.c_int), ("nop", ctypes.c_int), ("ntr", ctypes.c_int),
        ("allnx", ctypes.c_int), ("allnz", ctypes.c_int),
        ("rr", ctypes.c_float), ("rho_const", ctypes.c_float),
        ("adjoint_dir", ctypes.c_int), ("max_batch", ctypes.c_int),
        ("mode", ctypes.c_int), ("storage", ctypes.c_int), ("use_graph", ctypes.c_int),
        ("march_cols", ctypes.c_int), ("march_depth", ctypes.c_int), ("misfit", ctypes.c_int),
        ("march_order", ctypes.c_int), ("grad_rho", ctypes.c_int),
        ("cpu_mem", ctypes.c_int), ("strip_chunk", ctypes.c_int),
        ("rbc", ctypes.c_int), ("rbc_seed", ctypes.c_uint),
        ("recon", ctypes.c_int), ("snap_every", ctypes.c_int),
        ("fuse_step", ctypes.c_int), ("step_cols", ctypes.c_int),
    ]


class SfwiShot(ctypes.Structure):
    _fields_ = [
        ("left", ctypes.c_int), ("isx", ctypes.c_int), ("isz", ctypes.c_int),
        ("igz", ctypes.c_int), ("ngx_left", ctypes.c_int), ("ntr", ctypes.c_int),
        ("gc", ctypes.POINTER(ctypes.c_int)),
        ("obs_z", ctypes.POINTER(ctypes.c_float)), ("obs_x", ctypes.POINTER(ctypes.c_float)),
        ("obs_id", ctypes.c_int),
    ]


class SfwiRequest(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("params", SfwiParams),
        ("wavelet", ctypes.c_void_p), ("vp", ctypes.c_void_p), ("vs", ctypes.c_void_p),
        ("rho", ctypes.c_void_p), ("dvp", ctypes.c_void_p), ("dvs", ctypes.c_void_p),
        ("nshots", ctypes.c_int), ("shots", ctypes.POINTER(SfwiShot)),
        ("syn_z", ctypes.c_void_p), ("syn_x", ctypes.c_void_p), ("misfits", ctypes.c_void_p),
        ("grad_vp", ctypes.c_void_p), ("grad_vs", ctypes.c_void_p), ("illum", ctypes.c_void_p),
        ("autotune", ctypes.c_int),
    ]


_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_LIB = None


class SfwiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sfwi error %d: %s" % (code, msg))
        self.code = code


def declared_functions():
    """Names of every function declared in include/sfwi_abi.h."""
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sfwi_[a-z0-9_]+)\s*\(", src)))


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsfwi_amd.so is not built (run __graft_entry__.build()); "
                           "there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    dp = P(ctypes.c_double)
    L.sfwi_version.restype = ctypes.c_char_p
    L.sfwi_last_error.restype = ctypes.c_char_p
    L.sfwi_last_error.argtypes = [vp]
    L.sfwi_ctx_create.argtypes = [P(SfwiParams), P(vp)]
    L.sfwi_ctx_destroy.argtypes = [vp]
    L.sfwi_ctx_destroy.restype = None
    L.sfwi_set_wavelet.argtypes = [vp, _f32p]
    L.sfwi_set_model.argtypes = [vp, vp, vp, vp]
    L.sfwi_set_perturbation.argtypes = [vp, _f32p, _f32p]
    L.sfwi_set_perturbation_region.argtypes = [vp, _f32p, _f32p, ctypes.c_int]
    L.sfwi_batch_born2_forward.argtypes = [vp, ctypes.c_int, P(SfwiShot), vp, vp, vp, vp]
    L.sfwi_batch_born_forward.argtypes = [vp, ctypes.c_int, P(SfwiShot), vp, vp]
    L.sfwi_shot_hessian_vector.argtypes = [vp, P(SfwiShot), vp, vp]
    L.sfwi_batch_hessian_vector.argtypes = [vp, ctypes.c_int, P(SfwiShot), vp, vp]
    L.sfwi_shot_gradient.argtypes = [vp, P(SfwiShot), dp]
    L.sfwi_shot_born_gradient.argtypes = [vp, P(SfwiShot), dp]
    L.sfwi_batch_gradient.argtypes = [vp, ctypes.c_int, P(SfwiShot), dp]
    L.sfwi_batch_born_gradient.argtypes = [vp, ctypes.c_int, P(SfwiShot), dp]
    L.sfwi_shot_forward.argtypes = [vp, P(SfwiShot), vp, vp]
    L.sfwi_batch_forward.argtypes = [vp, ctypes.c_int, P(SfwiShot), vp, vp]
    L.sfwi_shot_born_forward.argtypes = [vp, P(SfwiShot), vp, vp, vp, vp]
    L.sfwi_linesearch_parabola.argtypes = [vp, P(SfwiShot), _f32p, _f32p, ctypes.c_float,
                                           ctypes.c_float, dp]
    L.sfwi_batch_linesearch_parabola.argtypes = [vp, ctypes.c_int, P(SfwiShot), _f32p, _f32p, ctypes.c_float,
                                                 ctypes.c_float, dp]
    L.sfwi_linesearch_linear.argtypes = [vp, P(SfwiShot), _f32p, _f32p, dp, dp]
    L.sfwi_batch_linesearch_linear.argtypes = [vp, ctypes.c_int, P(SfwiShot), _f32p, _f32p, dp, dp]
    L.sfwi_batch_born_gradient_res.argtypes = [vp, ctypes.c_int, P(SfwiShot), dp, vp, vp]
    L.sfwi_batch_linesearch_linear_res.argtypes = [vp, ctypes.c_int, P(SfwiShot), _f32p, _f32p, vp, vp, dp, dp]
    L.sfwi_flush.argtypes = [vp, vp, vp, vp]
    L.sfwi_obs_put.argtypes = [vp, ctypes.c_int, ctypes.c_int, _f32p, _f32p]
    L.sfwi_obs_clear.argtypes = [vp]
    L.sfwi_flush_rho.argtypes = [vp, vp]
    L.sfwi_ot_misfit.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                 _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, dp]
    L.sfwi_ot_misfit_static.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                        _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, dp]
    L.sfwi_ricker.argtypes = [_f32p, ctypes.c_int, ctypes.c_float, ctypes.c_float]
    L.sfwi_ricker.restype = None
    L.sfwi_set_timing.argtypes = [vp, ctypes.c_int]
    L.sfwi_kernel_time.argtypes = [vp, ctypes.c_int, dp, P(ctypes.c_int)]
    L.sfwi_kernel_name.restype = ctypes.c_char_p
    L.sfwi_kernel_name.argtypes = [ctypes.c_int]
    L.sfwi_get_params.argtypes = [vp, P(SfwiParams)]
    L.sfwi_autotune.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int),
                                P(ctypes.c_int), dp]
    i = ctypes.c_int
    L.sfwi_s1_direction.argtypes = [i, i, i, i, _f32p, _f32p, _f32p, _f32p, _f32p, dp, _f32p, _f32p, dp]
    L.sfwi_s1_hcg_init.argtypes = [i, i, _f32p, _f32p] + [_f32p] * 8
    L.sfwi_s1_hcg_step.argtypes = [i, i, i, i] + [_f32p] * 10 + [dp] * 4
    L.sfwi_s1_hcg_finish.argtypes = [i, i] + [_f32p] * 4
    L.sfwi_s1_alpha.restype = ctypes.c_double
    L.sfwi_s1_alpha.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_float, ctypes.c_float]
    L.sfwi_s1_update.argtypes = [i, i, i, ctypes.c_double, _f32p, _f32p, _f32p, _f32p]
    L.sfwi_s2_direction.argtypes = [i, i, i, i, _f32p, _f32p, _f32p, _f32p, _f32p, dp, _f32p, _f32p, dp]
    L.sfwi_s2_update.argtypes = [i, i, ctypes.c_double, ctypes.c_double, _f32p, _f32p, _f32p, _f32p, dp]
    L.sfwi_solve_bytes.restype = ctypes.c_size_t
    L.sfwi_solve_bytes.argtypes = [P(SfwiParams), ctypes.c_int, ctypes.c_int]
    L.sfwi_server_create.argtypes = [ctypes.c_size_t, ctypes.c_int, P(vp)]
    L.sfwi_server_submit.argtypes = [vp, P(SfwiRequest), P(ctypes.c_longlong)]
    L.sfwi_server_wait.argtypes = [vp, ctypes.c_longlong, dp, dp]
    L.sfwi_server_wait_for.argtypes = [vp, ctypes.c_longlong, ctypes.c_double, dp, dp]
    L.sfwi_server_cancel.argtypes = [vp, ctypes.c_longlong]
    L.sfwi_check_model.argtypes = [P(SfwiParams), vp, vp, vp]
    L.sfwi_server_batching.argtypes = [vp, ctypes.c_int]
    L.sfwi_server_destroy.argtypes = [vp]
    L.sfwi_server_destroy.restype = None
    _LIB = L
    return L


def ptr(a):
    """Data pointer of a float32/int32 array (None for None)."""
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def ricker(nt, dt, f0):
    """Ricker wavelet from the library's host helper (the reference driver's arithmetic)."""
    w = np.zeros(nt, np.float32)
    load().sfwi_ricker(w, nt, dt, f0)
    return w
