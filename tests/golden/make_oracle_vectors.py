"""Writes tests/golden/oracle_vectors_48x40.npz: the FP32 oracle's outputs on
one small seeded case, with every input stored beside them, so the CPU suite
can tell when the oracle's results change and the GPU suite can compare the
library's EXACT mode with stored vectors (tests/test_golden_vectors.py).

Case: the 48 x 40 layered model with an 8 % Gaussian inclusion of
sfwi_amd.workloads.tiny (seed 7), 8-cell PML, dx = dz = 10 m, dt = 1 ms,
nt = 320, 25 Hz Ricker, shot 1 of 2, 48 receivers.  Outputs: the observed
traces (forward solve of the true model), the stage-1 gradient at the
smoothed model (Vp, Vs, illumination, misfit), the stage-2 Born gradient
with the perturbation true - smoothed and density 2000 (adjoint direction
-1), and the Born forward traces.  Run from the repository root:
python tests/golden/make_oracle_vectors.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from sfwi_amd import workloads as W  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_vectors_48x40.npz")


def config(d, **kw):
    return O.config(allnx=int(d["allnx"]), allnz=int(d["allnz"]), nx=int(d["allnx"]), nz=int(d["allnz"]),
                    pml=int(d["pml"]), nt=int(d["nt"]), ntr=len(d["gc"]), dx=float(d["dx"]), dz=float(d["dz"]),
                    dt=float(d["dt"]), left=0, isx=int(d["isx"]), isz=int(d["isz"]), igz=int(d["igz"]),
                    ngx_left=0, **kw)


def outputs(d):
    """Every stored output computed afresh from the stored inputs."""
    c1 = config(d)
    oz, ox = O.forward(c1, d["vp"], d["vs"], d["rho"], d["wav"], d["gc"], "f32")
    g = O.gradient(c1, d["vp0"], d["vs0"], d["rho0"], d["wav"], d["gc"], oz, ox, "f32")
    c2 = config(d, rho_const=2000.0, adjoint_dir=-1)
    dvp, dvs = d["vp"] - d["vp0"], d["vs"] - d["vs0"]
    b = O.born_gradient(c2, d["vp0"], d["vs0"], d["rho0"], dvp, dvs, d["wav"], d["gc"], oz, ox, "f32")
    bz, bx, _, _ = O.born_forward(c2, d["vp0"], d["vs0"], d["rho0"], dvp, dvs, d["wav"], d["gc"], "f32")
    return dict(obs_z=oz, obs_x=ox, grad_vp=g[0], grad_vs=g[1], illum=g[2], misfit=np.float64(g[3]),
                born_grad_vp=b[0], born_grad_vs=b[1], born_illum=b[2], born_misfit=np.float64(b[3]),
                born_z=bz, born_x=bx)


def inputs():
    sv = W.tiny(nt=320)
    sh = sv.shots[1]
    isx, isz, igz = W.shot_geometry(sv, sh)
    return dict(vp=sv.vp, vs=sv.vs, rho=sv.rho, vp0=sv.vp0, vs0=sv.vs0, rho0=sv.rho0,
                wav=W.ricker(sv.nt, sv.dt, sv.f0), gc=sh.gc.astype(np.int32), isx=isx, isz=isz, igz=igz,
                allnx=sv.allnx, allnz=sv.allnz, pml=sv.pml, nt=sv.nt, dx=sv.dx, dz=sv.dz,
                dt=np.float32(sv.dt))


if __name__ == "__main__":
    d = inputs()
    d.update(outputs(d))
    np.savez_compressed(OUT, **d)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
