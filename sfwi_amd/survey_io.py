"""Survey input of the reference drivers over the library's C ABI: the 10-line
parameter file, x-major FP32 model files and SU shot records (index, gathers,
synthetic writer), and the per-shot geometry the worker derives from an index
row (sfwi-stage1/main.cpp:76-247, 580-625, 1322-1501).

``SuSurvey`` assembles a whole stage-1/2 survey from a parameter file: models
(the density is forced to 2000 as both drivers do), the shot index of the
record-z file, the shots between minshot and maxshot with their gathers, and
the shots grouped by aperture width so that each width gets its own context
(the reference sizes its one context by shot 0 and runs other apertures on
it, sfwi-stage1/main.cpp:610-617).
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import SfwiShot
from .elastic import Shot

PATH_MAX = 1024


class SfwiPara(ctypes.Structure):
    _fields_ = [("dx", ctypes.c_float), ("dz", ctypes.c_float),
                ("pml", ctypes.c_int), ("allnx", ctypes.c_int), ("allnz", ctypes.c_int),
                ("minshot", ctypes.c_int), ("maxshot", ctypes.c_int)] + \
               [(n, ctypes.c_char * PATH_MAX) for n in ("vp", "vs", "rho", "rec_z", "rec_x", "outdir")] + \
               [("f0", ctypes.c_float), ("maxiter", ctypes.c_int), ("iter_round1", ctypes.c_int),
                ("sy", ctypes.c_int), ("gy", ctypes.c_int),
                ("light", ctypes.c_int), ("rbc", ctypes.c_int), ("cpu_mem", ctypes.c_int),
                ("iointerval", ctypes.c_int)]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = v.decode() if isinstance(v, bytes) else v
        return d


def _L():
    L = _lib.load()
    if not getattr(L, "_suio", False):
        i, f, cp = ctypes.c_int, ctypes.c_float, ctypes.c_char_p
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        L.sfwi_para_read.argtypes = [cp, ctypes.POINTER(SfwiPara)]
        L.sfwi_model_read.argtypes = [cp, i, i, fp]
        L.sfwi_su_index.argtypes = [cp, i, i, ctypes.POINTER(i), ctypes.POINTER(f), ctypes.POINTER(i),
                                    ctypes.c_void_p]
        L.sfwi_su_read_gather.argtypes = [cp, ctypes.c_longlong, i, i, i, fp, ip]
        L.sfwi_su_write.argtypes = [cp, i, i, i, i, f, i, ip, fp]
        L.sfwi_shot_from_table.argtypes = [ip, f, i, i, i, i, ctypes.POINTER(SfwiShot), ctypes.POINTER(i)]
        L._suio = True
    return L


def _ck(rc):
    if rc != 0:
        raise _lib.SfwiError(rc, _lib.load().sfwi_last_error(None).decode())


def _b(path):
    return os.fsencode(path)


def read_para(path):
    p = SfwiPara()
    _ck(_L().sfwi_para_read(_b(path), ctypes.byref(p)))
    return p


def read_model(path, allnx, allnz):
    a = np.zeros((allnx, allnz), np.float32)
    _ck(_L().sfwi_model_read(_b(path), allnx, allnz, a))
    return a


def index_shot(path, coordinate_scale=1):
    """(nt, dt, table [nshots, 9] int32)."""
    L = _L()
    nt, dt, ns = ctypes.c_int(0), ctypes.c_float(0), ctypes.c_int(0)
    _ck(L.sfwi_su_index(_b(path), coordinate_scale, 0, ctypes.byref(nt), ctypes.byref(dt), ctypes.byref(ns), None))
    table = np.zeros((max(ns.value, 1), 9), np.int32)
    _ck(L.sfwi_su_index(_b(path), coordinate_scale, ns.value, ctypes.byref(nt), ctypes.byref(dt),
                        ctypes.byref(ns), table.ctypes.data_as(ctypes.c_void_p)))
    return nt.value, dt.value, table[:ns.value]


def read_gather(path, pos, ntr, nt, coordinate_scale=1):
    """(data [ntr, nt] float32, gc [ntr] int32)."""
    d = np.zeros((ntr, nt), np.float32)
    gc = np.zeros(ntr, np.int32)
    _ck(_L().sfwi_su_read_gather(_b(path), int(pos), ntr, nt, coordinate_scale, d, gc))
    return d, gc


def write_su(path, shot, sx, gx, data, dt, append=False):
    data = np.ascontiguousarray(data, np.float32)
    gx = np.ascontiguousarray(gx, np.int32)
    ntr, nt = data.shape
    _ck(_L().sfwi_su_write(_b(path), int(append), int(shot), ntr, nt, float(dt), int(sx), gx, data))


def shot_from_table(row, dx, pml, sy, gy, all_left):
    """(left, isx, isz, igz, ngx_left, ntr, nx_shot) of one index row."""
    s = SfwiShot()
    nx = ctypes.c_int(0)
    _ck(_L().sfwi_shot_from_table(np.ascontiguousarray(row, np.int32), float(dx), pml, sy, gy, all_left,
                                  ctypes.byref(s), ctypes.byref(nx)))
    return s.left, s.isx, s.isz, s.igz, s.ngx_left, s.ntr, nx.value


class SuSurvey:
    """A survey read the way the drivers read it (sfwi-stage1/main.cpp:1322-1501)."""

    def __init__(self, para_path, coordinate_scale=1, rho_value=2000.0):
        self.para = p = read_para(para_path)
        self.vp = read_model(p.vp.decode(), p.allnx, p.allnz)
        self.vs = read_model(p.vs.decode(), p.allnx, p.allnz)
        self.rho = np.full((p.allnx, p.allnz), np.float32(rho_value), np.float32)
        self.nt, self.dt, self.table = index_shot(p.rec_z.decode(), coordinate_scale)
        nt_x, dt_x, table_x = index_shot(p.rec_x.decode(), coordinate_scale)
        if nt_x != self.nt or not np.array_equal(table_x, self.table):
            raise ValueError("record-z and record-x files index differently")
        self.all_left = int(self.table[:, 4].min())
        t0 = self.table[0]
        # the context width the reference derives from shot 0 (sfwi-stage1/main.cpp:1431)
        self.nx_ref = int(t0[6] / p.dx + 0.5) - int(t0[4] / p.dx + 0.5) + 1
        self.shots = []
        self.widths = []
        for row in self.table:
            if row[0] < p.minshot or row[0] > p.maxshot:
                continue
            left, isx, isz, igz, ngx_left, ntr, nx_s = shot_from_table(row, p.dx, p.pml, p.sy, p.gy, self.all_left)
            oz, gc = read_gather(p.rec_z.decode(), row[8], ntr, self.nt, coordinate_scale)
            ox, gc_x = read_gather(p.rec_x.decode(), row[8], ntr, self.nt, coordinate_scale)
            if not np.array_equal(gc, gc_x):
                raise ValueError("shot %d: receiver coordinates differ between the record files" % row[0])
            self.shots.append(Shot(left, isx, isz, igz, ngx_left, gc, oz, ox))
            self.widths.append(nx_s)

    def groups(self):
        """{aperture width: [shot indices]} in order of first appearance."""
        g = {}
        for k, w in enumerate(self.widths):
            g.setdefault(w, []).append(k)
        return g
