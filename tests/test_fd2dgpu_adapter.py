"""The FD2DGPU_ELASTIC adapter (include/sfwi_fd2dgpu.hpp) driven the way the
reference's worker loop drives the class (sfwi-stage1/main.cpp:534-635), through
the host driver sfwi_amd/fd2dgpu_driver built from sfwi_amd/host/.

CPU: the driver links against the library and turns bad input and a missing
device into clean exit statuses.  GPU: the gradients, misfit and line-search
misfits it produces equal the ones of the Python path over the same C ABI.
"""
import os
import subprocess

import numpy as np
import pytest

from sfwi_amd import workloads as W
from tests.helpers import context, lib_shot

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sfwi_amd", "fd2dgpu_driver")


def write_survey(path, sv, records, mode=0, misfit=0):
    hdr = np.array([sv.nx, sv.allnz, sv.allnx, sv.allnz, sv.pml, sv.nt, sv.ntr, len(sv.shots), mode, misfit],
                   np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([sv.dx, sv.dz, sv.dt], np.float32).tobytes())
        for a in (sv.vp0, sv.vs0, sv.rho0, W.ricker(sv.nt, sv.dt, sv.f0)):
            f.write(np.ascontiguousarray(a, np.float32).tobytes())
        for sh, (oz, ox) in zip(sv.shots, records):
            isx, isz, igz = W.shot_geometry(sv, sh)
            f.write(np.array([sh.left, isx, isz, igz, sh.ngx_left, len(sh.gc)], np.int32).tobytes())
            f.write(np.asarray(sh.gc, np.int32).tobytes())
            f.write(np.ascontiguousarray(oz, np.float32).tobytes())
            f.write(np.ascontiguousarray(ox, np.float32).tobytes())


def _run(*args):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=120)


def test_driver_built_and_rejects_bad_input(tmp_path):
    assert os.path.exists(DRIVER), "run __graft_entry__.build()"
    assert _run().returncode == 2
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"\x01\x00")
    r = _run(str(bad), str(tmp_path / "o_"))
    assert r.returncode == 2 and "input error" in r.stderr


def test_driver_reports_library_errors(tmp_path):
    """Without a GPU the context cannot be created: the adapter's exception
    becomes exit status 3 with the library's message (on a GPU host the
    survey runs and the status is 0)."""
    sv = W.tiny(nt=40)
    recs = [(np.zeros((sv.ntr, sv.nt), np.float32),) * 2 for _ in sv.shots]
    path = tmp_path / "s.bin"
    write_survey(path, sv, recs)
    r = _run(str(path), str(tmp_path / "o_"))
    assert r.returncode in (0, 3), r.stderr
    if r.returncode == 3:
        assert r.stderr.startswith("sfwi:")
    else:
        assert r.stdout.startswith("misfit")


@pytest.mark.gpu
def test_driver_matches_python_path(tmp_path):
    sv = W.tiny(nt=240)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    with context(sv, max_batch=len(sv.shots)) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp, sv.vs, sv.rho)
        oz, ox = ctx.batch_forward([lib_shot(sv, s) for s in sv.shots])
        recs = [(oz[k][:len(s.gc)], ox[k][:len(s.gc)]) for k, s in enumerate(sv.shots)]
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        shots = [lib_shot(sv, s, *r) for s, r in zip(sv.shots, recs)]
        mis = sum(ctx.gradient(s) for s in shots)
        gvp, gvs, gil = ctx.flush()
        step1, step2 = 20.0, 60.0
        ls = np.zeros(3)
        for s in shots:
            ls += np.asarray(ctx.linesearch_parabola(s, gvp, gvs, step1, step2))
    path = tmp_path / "s.bin"
    write_survey(path, sv, recs)
    r = _run(str(path), str(tmp_path / "o_"), str(step1), str(step2))
    assert r.returncode == 0, r.stderr
    lines = dict(l.split(" ", 1) for l in r.stdout.strip().splitlines())
    assert float(lines["misfit"]) == mis
    assert np.array_equal(np.array(lines["linesearch"].split(), float), ls)
    for name, ref in (("pp_grad", gvp), ("ps_grad", gvs), ("illum", gil)):
        got = np.fromfile(tmp_path / ("o_%s.bin" % name), np.float32).reshape(ref.shape)
        assert np.array_equal(got, ref), name


def test_driver_rejects_bad_iteration_arguments(tmp_path):
    sv = W.tiny(nt=40)
    recs = [(np.zeros((sv.ntr, sv.nt), np.float32),) * 2 for _ in sv.shots]
    path = tmp_path / "s.bin"
    write_survey(path, sv, recs)
    r = _run(str(path), str(tmp_path / "o_"), "30", "150", "0", "4")
    assert r.returncode == 2 and "niter" in r.stderr


@pytest.mark.gpu
def test_driver_stage1_iterations_match_python_path(tmp_path):
    """Two complete stage-1 iterations of the C++ driver (the reference's
    master-rank sequence over the adapter) equal sfwi_amd.inversion.Stage1 over
    the Python binding, bit for bit: models, misfits, beta and alpha."""
    from sfwi_amd.inversion import Stage1
    sv = W.tiny(nt=220, nshots=3, seed=12)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    with context(sv, max_batch=len(sv.shots)) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp, sv.vs, sv.rho)
        oz, ox = ctx.batch_forward([lib_shot(sv, s) for s in sv.shots])
        recs = [(oz[k][:len(s.gc)], ox[k][:len(s.gc)]) for k, s in enumerate(sv.shots)]
        shots = [lib_shot(sv, s, *r) for s, r in zip(sv.shots, recs)]
        st = Stage1(sv.allnx, sv.allnz, cut_layer=4, step1=30.0, step2=150.0)
        vp, vs = sv.vp0.copy(), sv.vs0.copy()
        its = [st.iterate(ctx, shots, vp, vs, sv.rho0) for _ in range(2)]
    path = tmp_path / "s.bin"
    write_survey(path, sv, recs)
    r = _run(str(path), str(tmp_path / "o_"), "30", "150", "2", "4")
    assert r.returncode == 0, r.stderr
    lines = [l.split() for l in r.stdout.strip().splitlines()]
    assert len(lines) == 2
    for l, it in zip(lines, its):
        assert int(l[1]) == it["iter"]
        assert float(l[3]) == it["misfit"] and float(l[5]) == it["beta"] and float(l[7]) == it["alpha"]
    for name, ref in (("vel_p", vp), ("vel_s", vs)):
        got = np.fromfile(tmp_path / ("o_%s.bin" % name), np.float32).reshape(ref.shape)
        assert np.array_equal(got, ref), name
    # the same with cpu_mem = true, ntt = 50 (strips in host memory; the adapter
    # rebuilds its context on the first gradient call): identical output
    r2 = _run(str(path), str(tmp_path / "h_"), "30", "150", "2", "4", "50")
    assert r2.returncode == 0, r2.stderr
    assert r2.stdout == r.stdout
    for name in ("vel_p", "vel_s"):
        a = np.fromfile(tmp_path / ("o_%s.bin" % name), np.float32)
        b = np.fromfile(tmp_path / ("h_%s.bin" % name), np.float32)
        assert np.array_equal(a, b), name
