"""CPU build of the OT kernel's merge walk (sfwi_amd/csrc/ot_walk.hpp, the same
source the HIP kernel includes), compiled here with g++.

* On valid cumulative levels every cell's walk ends, and the cells' W2 shares
  sum to the numpy oracle's W2 (oracle/ot.py, w2_cells).
* On levels that break the walk's precondition (a cell ending above the last
  level of the other distribution, which the kernel's clamping rules out) the
  walk stops at its bound and reports it: the kernel turns that report into a
  device flag that makes the solve return an error instead of a truncated
  misfit.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import ot as OT

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include "ot_walk.hpp"
extern "C" int walk_cells(const double *P, const double *Q, int n, double *W, double *A, double *B) {
    double w = 0.0;
    int failed = 0;
    for (int i = 0; i < n; ++i) {
        A[i] = B[i] = 0.0;
        if (P[i + 1] - P[i] > 0.0 && !sfwi::ot_cell_walk(Q, n, i, P[i], P[i + 1], w, A[i], B[i])) ++failed;
    }
    *W = w / 12.0;
    return failed;
}
"""


@pytest.fixture(scope="module")
def walk(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("otwalk")
    src = d / "walk.cpp"
    src.write_text(HARNESS)
    so = d / "walk.so"
    subprocess.run([gxx, "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
                    "-I" + os.path.join(ROOT, "sfwi_amd", "csrc"), str(src), "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.walk_cells.argtypes = [dp, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), dp, dp]
    L.walk_cells.restype = ctypes.c_int

    def run(P, Q):
        n = len(P) - 1
        W = ctypes.c_double()
        A, B = np.zeros(n), np.zeros(n)
        bad = L.walk_cells(np.ascontiguousarray(P, np.float64), np.ascontiguousarray(Q, np.float64), n,
                           ctypes.byref(W), A, B)
        return bad, W.value, A, B
    return run


def _levels(m):
    c = np.cumsum(np.asarray(m, np.float64))
    L = np.concatenate(([0.0], c / c[-1]))
    L[-1] = 1.0
    return L


def test_walk_ends_and_matches_oracle_w2(walk):
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 64, 301):
        a = rng.random(n) + 0.05
        b = rng.random(n) + 0.05
        bad, W, A, B = walk(_levels(a), _levels(b))
        assert bad == 0
        W_ref, _, _ = OT.w2_cells(a, b)
        assert W == pytest.approx(W_ref, rel=1e-12, abs=1e-14)


def test_walk_of_a_pure_shift_is_shift_squared(walk):
    # a box of 5 cells moved by 3 cells: W2^2 = 9 exactly
    n = 20
    a = np.zeros(n); a[2:7] = 1.0
    b = np.zeros(n); b[5:10] = 1.0
    bad, W, _, _ = walk(_levels(a), _levels(b))
    assert bad == 0
    assert W == pytest.approx(9.0, rel=1e-13)


def test_walk_reports_levels_above_the_last(walk):
    # P rises above Q[n] = 1 before its last cell (what unclamped rounding of a
    # trace whose energy ends early produced): the walk of that cell cannot end
    n = 8
    Q = _levels(np.ones(n))
    P = _levels(np.ones(n))
    P[5] = 1.0 + 2.0 ** -40
    P[6] = 1.0 + 2.0 ** -40
    bad, _, _, _ = walk(P, Q)
    assert bad >= 1
    # the clamped levels (the kernel's rule) walk to the end
    bad, _, _, _ = walk(np.minimum(P, 1.0), Q)
    assert bad == 0
