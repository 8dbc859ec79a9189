"""Survey input (parameter file, models, SU records) on the CPU.

Fixtures: tests/golden/para_stage{1,2}_example1 and para_stage1_example2 are
the parameter files shipped with the reference's two examples
(example-1/, example-2/), copied as data.  The SU layout is pinned against the
SEG-Y trace header standard by reading the raw bytes of written files (sx at
byte 72, gx at 80, ns at 114, dt at 116), independently of the reader.
"""
import os

import numpy as np
import pytest

from sfwi_amd import survey_io as S
from sfwi_amd._lib import SfwiError

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reference_parameter_files():
    for name, allnx, allnz, maxshot in (("para_stage1_example1", 376, 201, 38), ("para_stage2_example1", 376, 201, 38),
                                        ("para_stage1_example2", 401, 301, 41)):
        p = S.read_para(os.path.join(GOLDEN, name)).as_dict()
        assert (p["dx"], p["dz"], p["pml"], p["allnx"], p["allnz"], p["minshot"], p["maxshot"]) == \
            (10.0, 10.0, 50, allnx, allnz, 1, maxshot)
        assert (p["f0"], p["maxiter"], p["iter_round1"], p["sy"], p["gy"]) == (10.0, 100, 0, 0, 0)
        assert (p["light"], p["rbc"], p["cpu_mem"], p["iointerval"]) == (1, 0, 0, 1000)
        assert p["rec_z"] == "/path/to/record-z.su" and p["rec_x"] == "/path/to/record-x.su"
        assert p["outdir"].startswith("/path/to/output-stage")
    p = S.read_para(os.path.join(GOLDEN, "para_stage2_example1")).as_dict()
    assert p["vp"] == "/path/to/stage1_v_376_201.bin" and p["vs"].endswith("zero.bin")


def test_parameter_file_errors(tmp_path):
    lines = open(os.path.join(GOLDEN, "para_stage1_example1")).read().splitlines()
    f = tmp_path / "p"
    f.write_text("\n".join(lines[:3] + lines[4:]) + "\n")          # one path line missing (stage-2 style)
    with pytest.raises(SfwiError, match="21 fields.*iointerval"):
        S.read_para(f)
    f.write_text("\n".join(["10 10 50 376 x 1 38"] + lines[1:]) + "\n")
    with pytest.raises(SfwiError, match="allnz is 'x'"):
        S.read_para(f)
    f.write_text("\n".join(lines[:1] + [""] + lines[1:]) + "\n")    # blank lines and line breaks are whitespace
    assert S.read_para(f).allnx == 376
    f.write_text(" ".join(lines) + "\n")
    assert S.read_para(f).as_dict() == S.read_para(os.path.join(GOLDEN, "para_stage1_example1")).as_dict()


def test_model_file(tmp_path):
    a = np.arange(6 * 5, dtype=np.float32).reshape(6, 5)
    f = tmp_path / "m.bin"
    a.tofile(f)
    assert np.array_equal(S.read_model(f, 6, 5), a)
    with pytest.raises(SfwiError, match="bytes"):
        S.read_model(f, 6, 6)


def _write_survey(path, shots, nt, dt, seed=0):
    r = np.random.default_rng(seed)
    data = []
    for k, (sx, gx) in enumerate(shots):
        d = r.standard_normal((len(gx), nt)).astype(np.float32)
        S.write_su(path, k + 1, sx, gx, d, dt, append=k > 0)
        data.append(d)
    return data


SHOTS = [(100, np.arange(0, 201, 10)), (250, np.arange(150, 351, 10)), (430, np.arange(300, 421, 10))]


def test_su_layout_index_and_gathers(tmp_path):
    f = str(tmp_path / "r.su")
    nt, dt = 37, 0.002
    data = _write_survey(f, SHOTS, nt, dt)
    raw = open(f, "rb").read()
    assert len(raw) == sum(len(g) for _, g in SHOTS) * (240 + 4 * nt)
    h = raw[:240]
    assert np.frombuffer(h[72:76], np.int32)[0] == 100 and np.frombuffer(h[80:84], np.int32)[0] == 0
    assert np.frombuffer(h[114:116], np.uint16)[0] == nt and np.frombuffer(h[116:118], np.uint16)[0] == 2000
    assert np.array_equal(np.frombuffer(raw[240:240 + 4 * nt], np.float32), data[0][0])
    nt2, dt2, table = S.index_shot(f)
    assert nt2 == nt and dt2 == np.float32(0.002)
    expect = [[1, 21, 100, 0, 0, 0, 200, 0, 0], [2, 21, 250, 0, 150, 0, 350, 0, 21],
              [3, 13, 430, 0, 300, 0, 430, 0, 42]]          # shot 3: the source is right of every receiver
    assert table.tolist() == expect                        # per-shot trace counts (not 1)
    for row, (sx, gx), d in zip(table, SHOTS, data):
        got, gc = S.read_gather(f, row[8], row[1], nt)
        assert np.array_equal(got, d) and np.array_equal(gc, gx)
    _, _, t10 = S.index_shot(f, coordinate_scale=10)
    assert t10[:, 2].tolist() == [10, 25, 43] and t10[:, 6].tolist() == [20, 35, 43]


def test_su_errors(tmp_path):
    f = str(tmp_path / "r.su")
    _write_survey(f, SHOTS[:1], 20, 0.001)
    with open(f, "ab") as g:
        g.write(b"\0" * 100)                               # a partial trace
    with pytest.raises(SfwiError, match="whole number"):
        S.index_shot(f)
    _write_survey(f, SHOTS[:1], 20, 0.001)
    with pytest.raises(SfwiError, match="not all present"):
        S.read_gather(f, 15, 10, 20)
    with pytest.raises(SfwiError, match="cannot open"):
        S.index_shot(str(tmp_path / "missing.su"))


def test_shot_geometry_from_index_row():
    # left = (xmin - all_left) / dx, isx = pml + (sx - xmin) / dx, aperture (xmax - xmin) / dx + 1
    row = np.array([2, 21, 250, 0, 150, 0, 350, 0, 21], np.int32)
    assert S.shot_from_table(row, 10.0, 8, 2, 1, 0) == (15, 8 + 10, 10, 9, 150, 21, 21)
    row = np.array([3, 13, 430, 0, 300, 0, 430, 0, 42], np.int32)
    assert S.shot_from_table(row, 10.0, 8, 2, 0, 100) == (20, 8 + 13, 10, 8, 300, 13, 14)


def test_su_survey_assembly(tmp_path):
    nt, dt = 25, 0.001
    allnx, allnz = 60, 12
    vp = np.full((allnx, allnz), 2000, np.float32)
    vs = (vp / 1.7).astype(np.float32)
    vp.tofile(tmp_path / "vp.bin")
    vs.tofile(tmp_path / "vs.bin")
    vs.tofile(tmp_path / "rho.bin")
    dz = _write_survey(str(tmp_path / "z.su"), SHOTS, nt, dt, seed=1)
    dx_ = _write_survey(str(tmp_path / "x.su"), SHOTS, nt, dt, seed=2)
    para = tmp_path / "para"
    para.write_text("10 10 8 %d %d 2 3\n%s\n%s\n%s\n%s\n%s\n%s\n12 3 0\n2 0\n1 0 0 100\n" % (
        allnx, allnz, tmp_path / "vp.bin", tmp_path / "vs.bin", tmp_path / "rho.bin", tmp_path / "z.su",
        tmp_path / "x.su", tmp_path / "out"))
    sv = S.SuSurvey(para)
    assert np.array_equal(sv.vp, vp) and np.all(sv.rho == 2000)
    assert sv.nt == nt and sv.all_left == 0 and sv.nx_ref == 21
    assert len(sv.shots) == 2                              # minshot 2 .. maxshot 3
    assert sv.widths == [21, 14] and sv.groups() == {21: [0], 14: [1]}
    s = sv.shots[1]
    assert (s.c.left, s.c.isx, s.c.isz, s.c.igz, s.c.ngx_left, s.ntr) == (30, 8 + 13, 10, 8, 300, 13)
    assert np.array_equal(s.obs_z, dz[2]) and np.array_equal(s.obs_x, dx_[2])


@pytest.mark.parametrize("stage", [1, 2])
def test_stage_program_input_errors(tmp_path, stage):
    import subprocess
    prog = os.path.join(os.path.dirname(GOLDEN), "..", "sfwi_amd", "sfwi_stage%d" % stage)
    assert os.path.exists(prog), "run __graft_entry__.build()"
    r = subprocess.run([prog], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([prog, str(tmp_path / "missing")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "cannot open parameter file" in r.stderr
    # the example file points at /path/to/..., so the model read fails cleanly
    r = subprocess.run([prog, os.path.join(GOLDEN, "para_stage%d_example1" % stage), "1"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "cannot open model file /path/to/" in r.stderr
    # rbc = 1 is accepted (random boundary): the run gets as far as the model read
    para = tmp_path / "para"
    para.write_text(open(os.path.join(GOLDEN, "para_stage%d_example1" % stage)).read().replace("1 0 0 1000",
                                                                                             "1 1 0 1000"))
    r = subprocess.run([prog, str(para), "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "cannot open model file /path/to/" in r.stderr
