"""Shared helpers of the test suite: oracle configs for a workload shot and
tolerance checks.  The oracle is what results are compared with; the HIP library is the product."""
import numpy as np

from oracle import oracle as O
from sfwi_amd import workloads as W
from sfwi_amd.elastic import Context, Shot


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def oracle_cfg(sv, shot, **kw):
    isx, isz, igz = W.shot_geometry(sv, shot)
    return O.config(allnx=sv.allnx, allnz=sv.allnz, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt,
                    ntr=len(shot.gc), dx=sv.dx, dz=sv.dz, dt=sv.dt, left=shot.left, isx=isx, isz=isz,
                    igz=igz, ngx_left=shot.ngx_left, **kw)


def lib_shot(sv, shot, obs_z=None, obs_x=None):
    isx, isz, igz = W.shot_geometry(sv, shot)
    return Shot(shot.left, isx, isz, igz, shot.ngx_left, shot.gc, obs_z, obs_x)


def context(sv, mode="fast", max_batch=1, ntr=None, **kw):
    return Context(dx=sv.dx, dz=sv.dz, dt=sv.dt, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt,
                   ntr=ntr or sv.ntr, allnx=sv.allnx, allnz=sv.allnz, mode=mode, max_batch=max_batch,
                   **kw)
