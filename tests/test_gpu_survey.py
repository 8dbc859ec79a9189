"""A stage-1 inversion read from the reference's input formats: a parameter
file, x-major model files and SU records of shots with three different
aperture widths (so three contexts).  The C++ program sfwi_stage1 and the
Python loop over sfwi_amd.survey_io.SuSurvey give the same models, misfits,
beta and alpha bit for bit, and both equal the same iterations run on the
in-memory shots the records were modelled from (the SU round trip and the
geometry derived from the index are lossless).
"""
import os
import subprocess

import numpy as np
import pytest

from sfwi_amd import _lib
from sfwi_amd import survey_io as S
from sfwi_amd import workloads as W
from sfwi_amd.elastic import Context, Shot
from sfwi_amd.inversion import Stage1, Stage2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "sfwi_amd", "sfwi_stage1")
PROG2 = os.path.join(ROOT, "sfwi_amd", "sfwi_stage2")
DX, PML, NT, DT, F0, SY, GY, CUT = 10.0, 8, 220, 0.001, 25.0, 2, 0, 4
# the windows [0, 30), [25, 60), [20, 45) cover the model: a cell outside every window has zero
# illumination, and the preconditioning divides by it, as in the reference
SPREADS = [(150, np.arange(0, 291, 10)), (400, np.arange(250, 591, 10)), (300, np.arange(200, 441, 10))]


def _model():
    sv = W.tiny(nx=60, nz=40, seed=17, nt=NT)
    return sv


def _shot(sx, gx, all_left=0):
    xmin, xmax = min(sx, gx.min()), max(sx, gx.max())
    left = int((xmin - all_left) / DX)
    width = int((xmax - xmin) / DX) + 1
    return Shot(left, PML + int((sx - xmin) / DX), PML + SY, PML + GY, xmin, gx), width


def _ctx(sv, width, nshots, **kw):
    return Context(dx=DX, dz=DX, dt=DT, nx=width, nz=sv.allnz, pml=PML, nt=NT, ntr=35, allnx=sv.allnx,
                   allnz=sv.allnz, max_batch=nshots, **kw)


def test_stage1_from_su_files(tmp_path):
    sv = _model()
    w = _lib.ricker(NT, np.float32(DT), F0)
    rho = np.full_like(sv.vp, 2000.0)
    # observed records from the true model, written as SU
    recs = []
    for k, (sx, gx) in enumerate(SPREADS):
        sh, width = _shot(sx, gx)
        with _ctx(sv, width, 1) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp, sv.vs, rho)
            oz, ox = ctx.forward(sh)
        S.write_su(str(tmp_path / "z.su"), k + 1, sx, gx, oz, DT, append=k > 0)
        S.write_su(str(tmp_path / "x.su"), k + 1, sx, gx, ox, DT, append=k > 0)
        recs.append((oz, ox))
    sv.vp0.tofile(tmp_path / "vp.bin")
    sv.vs0.tofile(tmp_path / "vs.bin")
    sv.vs0.tofile(tmp_path / "rho.bin")
    para = tmp_path / "para"
    para.write_text("%g %g %d %d %d 1 3\n%s\n%s\n%s\n%s\n%s\n%s\n%g 2 0\n%d %d\n1 0 0 1000\n" % (
        DX, DX, PML, sv.allnx, sv.allnz, tmp_path / "vp.bin", tmp_path / "vs.bin", tmp_path / "rho.bin",
        tmp_path / "z.su", tmp_path / "x.su", tmp_path / "out", F0, SY, GY))

    # Python over the SU survey, one context per aperture width
    su = S.SuSurvey(para)
    assert su.nt == NT and len(su.groups()) == 3
    ctxs = {wd: _ctx(sv, wd, len(ix)) for wd, ix in su.groups().items()}
    for c in ctxs.values():
        c.set_wavelet(w)
    groups = [(ctxs[wd], [su.shots[k] for k in ix]) for wd, ix in su.groups().items()]
    st = Stage1(sv.allnx, sv.allnz, cut_layer=CUT)
    vp, vs = su.vp.copy(), su.vs.copy()
    its = [st.iterate_groups(groups, vp, vs, su.rho) for _ in range(2)]
    for c in ctxs.values():
        c.close()
    assert its[1]["misfit"] < its[0]["misfit"] and np.all(np.isfinite(vp))

    # the same iterations on the in-memory shots
    st2 = Stage1(sv.allnx, sv.allnz, cut_layer=CUT)
    vp2, vs2 = sv.vp0.copy(), sv.vs0.copy()
    mem = []
    for (sx, gx), (oz, ox) in zip(SPREADS, recs):
        sh, width = _shot(sx, gx)
        c = _ctx(sv, width, 1)
        c.set_wavelet(w)
        mem.append((c, [Shot(sh.c.left, sh.c.isx, sh.c.isz, sh.c.igz, sh.c.ngx_left, gx, oz, ox)]))
    its2 = [st2.iterate_groups(mem, vp2, vs2, rho) for _ in range(2)]
    for c, _ in mem:
        c.close()
    assert np.array_equal(vp, vp2) and np.array_equal(vs, vs2)
    assert [i["alpha"] for i in its] == [i["alpha"] for i in its2]

    # the C++ program on the files
    r = subprocess.run([PROG, str(para), "2", "64", str(CUT)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = [l.split() for l in r.stdout.strip().splitlines()]
    assert len(lines) == 2
    for l, it in zip(lines, its):
        assert float(l[3]) == it["misfit"] and float(l[5]) == it["beta"] and float(l[7]) == it["alpha"]
    got = np.fromfile(tmp_path / "out" / ("vel_p_1_%d_%d.dat" % (sv.allnx, sv.allnz)), np.float32)
    assert np.array_equal(got.reshape(vp.shape), vp)
    got = np.fromfile(tmp_path / "out" / ("vel_s_1_%d_%d.dat" % (sv.allnx, sv.allnz)), np.float32)
    assert np.array_equal(got.reshape(vs.shape), vs)
    assert (tmp_path / "out" / "misfit.txt").read_text().count("misfit") == 2


def test_stage1_gauss_newton_program_matches_python(tmp_path):
    """Two stage-1 iterations with the Gauss-Newton inner CG (max_inner = 2,
    sfwi-stage1/main.cpp:784-1063): the C++ program on the SU files and the
    Python loop over the same survey give the same models, misfits, beta and
    alpha bit for bit, and the program logs the reference's hessian_misfit
    lines."""
    sv = _model()
    w = _lib.ricker(NT, np.float32(DT), F0)
    rho = np.full_like(sv.vp, 2000.0)
    for k, (sx, gx) in enumerate(SPREADS):
        sh, width = _shot(sx, gx)
        with _ctx(sv, width, 1) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp, sv.vs, rho)
            oz, ox = ctx.forward(sh)
        S.write_su(str(tmp_path / "z.su"), k + 1, sx, gx, oz, DT, append=k > 0)
        S.write_su(str(tmp_path / "x.su"), k + 1, sx, gx, ox, DT, append=k > 0)
    sv.vp0.tofile(tmp_path / "vp.bin")
    sv.vs0.tofile(tmp_path / "vs.bin")
    sv.vs0.tofile(tmp_path / "rho.bin")
    para = tmp_path / "para"
    para.write_text("%g %g %d %d %d 1 3\n%s\n%s\n%s\n%s\n%s\n%s\n%g 2 0\n%d %d\n1 0 0 1000\n" % (
        DX, DX, PML, sv.allnx, sv.allnz, tmp_path / "vp.bin", tmp_path / "vs.bin", tmp_path / "rho.bin",
        tmp_path / "z.su", tmp_path / "x.su", tmp_path / "out", F0, SY, GY))
    su = S.SuSurvey(para)
    ctxs = {wd: _ctx(sv, wd, len(ix)) for wd, ix in su.groups().items()}
    for c in ctxs.values():
        c.set_wavelet(w)
    groups = [(ctxs[wd], [su.shots[k] for k in ix]) for wd, ix in su.groups().items()]
    st = Stage1(sv.allnx, sv.allnz, cut_layer=CUT, max_inner=2)
    vp, vs = su.vp.copy(), su.vs.copy()
    its = [st.iterate_groups(groups, vp, vs, su.rho) for _ in range(2)]
    for c in ctxs.values():
        c.close()
    assert all(len(i["inner"]) == 2 for i in its) and np.all(np.isfinite(vp))
    r = subprocess.run([PROG, str(para), "2", "64", str(CUT), "2"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = [l.split() for l in r.stdout.strip().splitlines()]
    for l, it in zip(lines, its):
        assert float(l[3]) == it["misfit"] and float(l[5]) == it["beta"] and float(l[7]) == it["alpha"]
    got = np.fromfile(tmp_path / "out" / ("vel_p_1_%d_%d.dat" % (sv.allnx, sv.allnz)), np.float32)
    assert np.array_equal(got.reshape(vp.shape), vp)
    log = (tmp_path / "out" / "misfit.txt").read_text()
    assert log.count("hessian_misfit") == 2 * 3
    first = float(log.split("hessian_misfit = ")[1].split()[0])
    assert first == pytest.approx(its[0]["inner"][0][0], rel=1e-3)   # printed with %5.3e


def test_stage1_overlapping_windows_in_batches(tmp_path):
    """Four shots of one aperture width whose windows overlap at offsets 0, 10,
    20 and 30 of a 60-column model: the stage-1 program solving them in
    batches of 3 + 1, the Python loop solving them as one batch of 4, and the
    Python loop solving one shot per call give the same models, misfits and
    steps bit for bit (the per-shot images are merged in shot order whatever
    the batching)."""
    sv = _model()
    w = _lib.ricker(NT, np.float32(DT), F0)
    rho = np.full_like(sv.vp, 2000.0)
    spreads = [(sx, np.arange(sx - 140, sx + 151, 10)) for sx in (140, 240, 340, 440)]
    shots = []
    for k, (sx, gx) in enumerate(spreads):
        sh, width = _shot(sx, gx)
        assert width == 30
        with _ctx(sv, width, 1) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp, sv.vs, rho)
            oz, ox = ctx.forward(sh)
        S.write_su(str(tmp_path / "z.su"), k + 1, sx, gx, oz, DT, append=k > 0)
        S.write_su(str(tmp_path / "x.su"), k + 1, sx, gx, ox, DT, append=k > 0)
        shots.append(Shot(sh.c.left, sh.c.isx, sh.c.isz, sh.c.igz, sh.c.ngx_left, gx, oz, ox))
    sv.vp0.tofile(tmp_path / "vp.bin")
    sv.vs0.tofile(tmp_path / "vs.bin")
    sv.vs0.tofile(tmp_path / "rho.bin")
    para = tmp_path / "para"
    para.write_text("%g %g %d %d %d 1 4\n%s\n%s\n%s\n%s\n%s\n%s\n%g 2 0\n%d %d\n1 0 0 1000\n" % (
        DX, DX, PML, sv.allnx, sv.allnz, tmp_path / "vp.bin", tmp_path / "vs.bin", tmp_path / "rho.bin",
        tmp_path / "z.su", tmp_path / "x.su", tmp_path / "out", F0, SY, GY))

    results = []
    for max_batch in (4, 1):
        st = Stage1(sv.allnx, sv.allnz, cut_layer=CUT)
        vp, vs = sv.vp0.copy(), sv.vs0.copy()
        with _ctx(sv, 30, max_batch) as c:
            c.set_wavelet(w)
            its = [st.iterate_groups([(c, shots)], vp, vs, rho) for _ in range(2)]
        results.append((vp, vs, its))
    (vp, vs, its), (vp1, vs1, its1) = results
    assert np.array_equal(vp, vp1) and np.array_equal(vs, vs1)
    assert [(i["misfit"], i["beta"], i["alpha"]) for i in its] == [(i["misfit"], i["beta"], i["alpha"]) for i in its1]
    assert its[1]["misfit"] < its[0]["misfit"]

    r = subprocess.run([PROG, str(para), "2", "3", str(CUT)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = [l.split() for l in r.stdout.strip().splitlines()]
    for l, it in zip(lines, its):
        assert float(l[3]) == it["misfit"] and float(l[5]) == it["beta"] and float(l[7]) == it["alpha"]
    got = np.fromfile(tmp_path / "out" / ("vel_p_1_%d_%d.dat" % (sv.allnx, sv.allnz)), np.float32)
    assert np.array_equal(got.reshape(vp.shape), vp)
    got = np.fromfile(tmp_path / "out" / ("vel_s_1_%d_%d.dat" % (sv.allnx, sv.allnz)), np.float32)
    assert np.array_equal(got.reshape(vs.shape), vs)

    # rbc = 1 in the file: the random boundary (seed 1), no strips; the program
    # equals the Python loop on contexts with the same boundary, bit for bit
    para3 = tmp_path / "para3"
    para3.write_text(para.read_text().replace("1 0 0 1000\n", "1 1 0 1000\n").replace(str(tmp_path / "out"),
                                                                                    str(tmp_path / "out3")))
    r3 = subprocess.run([PROG, str(para3), "2", "3", str(CUT)], capture_output=True, text=True, timeout=300)
    assert r3.returncode == 0, r3.stderr
    st3 = Stage1(sv.allnx, sv.allnz, cut_layer=CUT)
    vp3, vs3 = sv.vp0.copy(), sv.vs0.copy()
    with _ctx(sv, 30, 3, rbc=True, rbc_seed=1) as c:
        c.set_wavelet(w)
        its3 = [st3.iterate_groups([(c, shots)], vp3, vs3, rho) for _ in range(2)]
    for l, it in zip([l.split() for l in r3.stdout.strip().splitlines()], its3):
        assert float(l[3]) == it["misfit"] and float(l[7]) == it["alpha"]
    got = np.fromfile(tmp_path / "out3" / ("vel_p_1_%d_%d.dat" % (sv.allnx, sv.allnz)), np.float32)
    assert np.array_equal(got.reshape(vp.shape), vp3)
    assert not np.array_equal(vp3, vp)          # the boundary changed the inversion

    # cpu_mem = 1, iointerval = 50 in the file: strips in host memory in chunks of
    # 50 steps (the last one partial); same output bit for bit
    para2 = tmp_path / "para2"
    para2.write_text(para.read_text().replace("1 0 0 1000\n", "1 0 1 50\n").replace(str(tmp_path / "out"),
                                                                                    str(tmp_path / "out2")))
    assert S.read_para(para2).as_dict()["cpu_mem"] == 1
    r2 = subprocess.run([PROG, str(para2), "2", "3", str(CUT)], capture_output=True, text=True, timeout=300)
    assert r2.returncode == 0, r2.stderr
    assert r2.stdout == r.stdout
    for name in ("vel_p_1", "vel_s_1"):
        f = "%s_%d_%d.dat" % (name, sv.allnx, sv.allnz)
        assert np.array_equal(np.fromfile(tmp_path / "out" / f, np.float32),
                              np.fromfile(tmp_path / "out2" / f, np.float32)), name


def test_stage2_from_su_files(tmp_path):
    """The stage-2 program on SU records of shots with three aperture widths
    (three contexts, density 2000, stage-2 adjoint direction) and the Python
    stage-2 loop over the same files give the same images, misfits, beta,
    alpha and step sums bit for bit; the batched step equals the per-shot
    CALCULATE_ALPHA loop bit for bit; the misfit falls (Born data are affine
    in the images, so the exact step along a descent direction lowers it)."""
    sv = _model()
    w = _lib.ricker(NT, np.float32(DT), F0)
    rho = np.full_like(sv.vp, 2000.0)
    for k, (sx, gx) in enumerate(SPREADS):
        sh, width = _shot(sx, gx)
        with _ctx(sv, width, 1) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp, sv.vs, rho)
            oz, ox = ctx.forward(sh)
        S.write_su(str(tmp_path / "z.su"), k + 1, sx, gx, oz, DT, append=k > 0)
        S.write_su(str(tmp_path / "x.su"), k + 1, sx, gx, ox, DT, append=k > 0)
    sv.vp0.tofile(tmp_path / "vp.bin")
    sv.vs0.tofile(tmp_path / "vs.bin")
    sv.vs0.tofile(tmp_path / "rho.bin")
    para = tmp_path / "para"
    para.write_text("%g %g %d %d %d 1 3\n%s\n%s\n%s\n%s\n%s\n%s\n%g 2 0\n%d %d\n1 0 0 1000\n" % (
        DX, DX, PML, sv.allnx, sv.allnz, tmp_path / "vp.bin", tmp_path / "vs.bin", tmp_path / "rho.bin",
        tmp_path / "z.su", tmp_path / "x.su", tmp_path / "out", F0, SY, GY))

    su = S.SuSurvey(para)
    assert len(su.groups()) == 3
    results = []
    for per_shot in (False, True):
        ctxs = {wd: _ctx(sv, wd, len(ix), rho_const=2000.0, adjoint_dir=-1) for wd, ix in su.groups().items()}
        for c in ctxs.values():
            c.set_wavelet(w)
        groups = [(ctxs[wd], [su.shots[k] for k in ix]) for wd, ix in su.groups().items()]
        st = Stage2(sv.allnx, sv.allnz, cut_layer=CUT)
        ip, is_ = np.zeros_like(su.vp), np.zeros_like(su.vp)
        its = [st.iterate_groups(groups, su.vp, su.vs, None, ip, is_, per_shot=per_shot) for _ in range(2)]
        for c in ctxs.values():
            c.close()
        results.append((ip, is_, its))
    (ip, is_, its), (ip1, is1, its1) = results
    assert np.abs(ip).max() > 0 and np.abs(is_).max() > 0
    assert np.array_equal(ip, ip1) and np.array_equal(is_, is1)
    key = ("misfit", "beta", "alpha", "fenzi", "fenmu")
    assert [[i[k] for k in key] for i in its] == [[i[k] for k in key] for i in its1]
    assert its[1]["misfit"] < its[0]["misfit"] and its[0]["beta"] == 0.0

    r = subprocess.run([PROG2, str(para), "2", "64", str(CUT)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = [l.split() for l in r.stdout.strip().splitlines()]
    assert len(lines) == 2
    for l, it in zip(lines, its):
        assert [float(l[j]) for j in (3, 5, 7, 9, 11)] == [it[k] for k in key]
    out = tmp_path / "out"
    for name, ref in (("sum_image_pp", ip), ("sum_image_ps", is_)):
        got = np.fromfile(out / ("%s_1_%d_%d.dat" % (name, sv.allnx, sv.allnz)), np.float32)
        assert np.array_equal(got.reshape(ref.shape), ref), name
    for name in ("sum_grad_pp", "sum_grad_ps", "sum_cggrad_pp", "sum_cggrad_ps", "sum_illum"):
        a = np.fromfile(out / ("%s_0_%d_%d.dat" % (name, sv.allnx, sv.allnz)), np.float32)
        assert a.size == sv.allnx * sv.allnz and np.abs(a).max() > 0, name
    # the first direction: image_1 = -alpha_0 cg_0
    cg = np.fromfile(out / ("sum_cggrad_pp_0_%d_%d.dat" % (sv.allnx, sv.allnz)), np.float32).reshape(ip.shape)
    i0 = np.fromfile(out / ("sum_image_pp_0_%d_%d.dat" % (sv.allnx, sv.allnz)), np.float32).reshape(ip.shape)
    assert np.array_equal(i0, (0.0 - its[0]["alpha"] * cg.astype(np.float64)).astype(np.float32))
    assert (out / "misfit.txt").read_text().count("misfit") == 2

    # rbc = 1: the random boundary (seed 1) in the program and in the Python loop, bit for bit
    para3 = tmp_path / "para3"
    para3.write_text(para.read_text().replace("1 0 0 1000\n", "1 1 0 1000\n").replace(str(out), str(tmp_path / "out3")))
    r3 = subprocess.run([PROG2, str(para3), "2", "64", str(CUT)], capture_output=True, text=True, timeout=300)
    assert r3.returncode == 0, r3.stderr
    ctxs = {wd: _ctx(sv, wd, len(ix), rho_const=2000.0, adjoint_dir=-1, rbc=True, rbc_seed=1)
            for wd, ix in su.groups().items()}
    for c in ctxs.values():
        c.set_wavelet(w)
    groups = [(ctxs[wd], [su.shots[k] for k in ix]) for wd, ix in su.groups().items()]
    st3 = Stage2(sv.allnx, sv.allnz, cut_layer=CUT)
    ip3, is3 = np.zeros_like(su.vp), np.zeros_like(su.vp)
    its3 = [st3.iterate_groups(groups, su.vp, su.vs, None, ip3, is3) for _ in range(2)]
    for c in ctxs.values():
        c.close()
    for l, it in zip([l.split() for l in r3.stdout.strip().splitlines()], its3):
        assert [float(l[j]) for j in (3, 5, 7, 9, 11)] == [it[k] for k in key]
    f = "sum_image_pp_1_%d_%d.dat" % (sv.allnx, sv.allnz)
    assert np.array_equal(np.fromfile(tmp_path / "out3" / f, np.float32).reshape(ip3.shape), ip3)
    assert not np.array_equal(ip3, ip)

    # cpu_mem = 1, iointerval = 50: boundary strips in pinned host memory, same output bit for bit
    para2 = tmp_path / "para2"
    para2.write_text(para.read_text().replace("1 0 0 1000\n", "1 0 1 50\n").replace(str(out), str(tmp_path / "out2")))
    r2 = subprocess.run([PROG2, str(para2), "2", "64", str(CUT)], capture_output=True, text=True, timeout=300)
    assert r2.returncode == 0, r2.stderr
    assert r2.stdout == r.stdout
    for name in ("sum_image_pp_1", "sum_image_ps_1"):
        f = "%s_%d_%d.dat" % (name, sv.allnx, sv.allnz)
        assert np.array_equal(np.fromfile(out / f, np.float32), np.fromfile(tmp_path / "out2" / f, np.float32)), name
