"""The benchmark's algorithmic-byte model (bench.alg_bytes) on the CPU: its
cell classes equal a brute-force count over the padded grid of the split-band
rule the kernels use (a cell of the update domain [4, n - 4) is split when it
lies within pml + 4 of an edge), and each kernel's bytes add up from those
counts.  No GPU needed."""
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
    # every kernel of the measured path has a traffic figure
    for k in bench.MEASURED_HBM_BYTES_PER_SHOT:
        assert bench.alg_bytes(k, 401, 301, 30) > 0
