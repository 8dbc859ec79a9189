"""The benchmark's byte models on the CPU.  The kernels' own cell-class model
(bench.alg_bytes): its cell classes equal a brute-force count over the padded
grid of the split-band rule the kernels use (a cell of the update domain
[4, n - 4) is split when it lies within pml + 4 of an edge), and each
kernel's bytes add up from those counts.  The scope-table model
(bench.alg_bytes_8d): the two half steps of a propagation add up to 52 B per
interior cell and 100 B per PML cell, and each kernel of the measured
benchmark path has a figure.  No GPU needed."""
import numpy as np
import pytest

import bench


def _brute(nx, nz, pml):
    nxp, nzp = nx + 2 * pml, nz + 2 * pml
    ix, iz = np.meshgrid(np.arange(nxp), np.arange(nzp), indexing="ij")
    upd = (ix >= 4) & (ix < nxp - 4) & (iz >= 4) & (iz < nzp - 4)
    w = pml + 4
    band = (ix < w) | (ix >= nxp - w) | (iz < w) | (iz >= nzp - w)
    img = (ix >= pml) & (ix < pml + nx) & (iz >= pml) & (iz < pml + nz)
    return int((upd & ~band).sum()), int((upd & band).sum()), int(img.sum())


@pytest.mark.parametrize("nx,nz,pml", [(401, 301, 30), (376, 201, 20), (1200, 600, 40), (48, 40, 8), (57, 49, 8),
                                       (8, 8, 2)])
def test_cell_classes_match_brute_force(nx, nz, pml):
    assert bench.cell_classes(nx, nz, pml) == _brute(nx, nz, pml)


def test_kernel_bytes_from_cell_classes():
    un, sp, im = bench.cell_classes(401, 301, 30)
    assert (un, sp, im) == (115149, 44760, 120701)
    assert bench.alg_bytes("k_str(forward)", 401, 301, 30) == 40 * un + 84 * sp
    assert bench.alg_bytes("k_str(reconstruct+image)", 401, 301, 30) == 40 * un + 84 * sp + 28 * im
    strips = 2 * 4 * 4 * (401 + 301) * 4          # vx, vz of 16 (nx + nz) strip cells, FP32
    assert bench.alg_bytes("k_vel(forward)", 401, 301, 30) == 32 * un + 64 * sp + strips


def test_scope_table_bytes():
    nx, nz, pml = 1200, 600, 40
    inter = nx * nz
    pmlc = (nx + 2 * pml) * (nz + 2 * pml) - inter
    v = bench.alg_bytes_8d("k_vel(forward)", nx, nz, pml)
    s = bench.alg_bytes_8d("k_str(forward)", nx, nz, pml)
    assert v + s == 52 * inter + 100 * pmlc
    assert bench.alg_bytes_8d("k_adj_vel", nx, nz, pml) == v
    assert bench.alg_bytes_8d("k_snap_image", nx, nz, pml, 2) == 40 * inter
    assert bench.alg_bytes_8d("k_snap_store", nx, nz, pml, 2) == 12 * (nx + 8) * (nz + 8)
    # every kernel of the measured path has a traffic figure and a byte figure
    for key, table in bench.MEASURED_HBM_BYTES_PER_SHOT.items():
        for k in table:
            assert bench.alg_bytes_8d(k, nx, nz, pml) > 0
