"""Stored vectors (tests/golden/oracle_vectors_48x40.npz, written by
tests/golden/make_oracle_vectors.py): the FP32 oracle recomputes every
stored output from the stored inputs bit for bit (CPU), and the library's
EXACT mode reproduces them bit for bit (GPU): observed traces, the stage-1
gradient, the stage-2 Born gradient and the Born traces; misfits to 1e-12
(the library groups the FP64 partial sums per trace)."""
import os

import numpy as np
import pytest

from sfwi_amd.elastic import Context, Shot

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden", "oracle_vectors_48x40.npz")


def _load():
    with np.load(PATH) as f:
        return {k: f[k] for k in f.files}


def test_oracle_reproduces_the_stored_vectors():
    from tests.golden.make_oracle_vectors import outputs
    d = _load()
    new = outputs(d)
    for k, v in new.items():
        assert np.array_equal(v, d[k]), k


def _context(d, **kw):
    return Context(dx=float(d["dx"]), dz=float(d["dz"]), dt=float(d["dt"]), nx=int(d["allnx"]), nz=int(d["allnz"]),
                   pml=int(d["pml"]), nt=int(d["nt"]), ntr=len(d["gc"]), allnx=int(d["allnx"]),
                   allnz=int(d["allnz"]), mode="exact", **kw)


@pytest.mark.gpu
def test_exact_mode_reproduces_the_stored_vectors():
    d = _load()
    geo = (0, int(d["isx"]), int(d["isz"]), int(d["igz"]), 0, d["gc"])
    with _context(d) as ctx:
        ctx.set_wavelet(d["wav"])
        ctx.set_model(d["vp"], d["vs"], d["rho"])
        oz, ox = ctx.forward(Shot(*geo))
        assert np.array_equal(oz, d["obs_z"]) and np.array_equal(ox, d["obs_x"])
        ctx.set_model(d["vp0"], d["vs0"], d["rho0"])
        m = ctx.gradient(Shot(*geo, d["obs_z"], d["obs_x"]))
        gvp, gvs, gil = ctx.flush()
    assert np.array_equal(gvp, d["grad_vp"]) and np.array_equal(gvs, d["grad_vs"])
    assert np.array_equal(gil, d["illum"])
    assert abs(m - d["misfit"]) <= 1e-12 * abs(d["misfit"])
    with _context(d, rho_const=2000.0, adjoint_dir=-1) as ctx:
        ctx.set_wavelet(d["wav"])
        ctx.set_model(d["vp0"], d["vs0"], d["rho0"])
        ctx.set_perturbation(d["vp"] - d["vp0"], d["vs"] - d["vs0"])
        bz, bx, _, _ = ctx.born_forward(Shot(*geo))
        assert np.array_equal(bz, d["born_z"]) and np.array_equal(bx, d["born_x"])
        m = ctx.born_gradient(Shot(*geo, d["obs_z"], d["obs_x"]))
        b = ctx.flush()
    assert np.array_equal(b[0], d["born_grad_vp"]) and np.array_equal(b[1], d["born_grad_vs"])
    assert np.array_equal(b[2], d["born_illum"])
    assert abs(m - d["born_misfit"]) <= 1e-12 * abs(d["born_misfit"])
