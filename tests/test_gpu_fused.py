"""The fused forward time step (velocity and stress of a step in one column
march, double-buffered state) against the two half-step column-march kernels
and the per-cell kernels: forward traces, strip gradients (FP32 and BF16
strips, strips in host memory), snapshot gradients and line-search misfits
are equal bit for bit, for even and odd record lengths (the final state ends
in either buffer), segment widths that do and do not divide the grid, grids
with one and with several row bands, and graph replays."""
import numpy as np
import pytest

from oracle import oracle as O
from sfwi_amd import workloads as W
from tests.helpers import context, lib_shot, oracle_cfg

pytestmark = pytest.mark.gpu


def _case(nz, nt, seed):
    sv = W.tiny(nx=52, nz=nz, pml=8, nt=nt, nshots=2, seed=seed)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    obs = [O.forward(oracle_cfg(sv, sh), sv.vp, sv.vs, sv.rho, w, sh.gc, "f32") for sh in sv.shots]
    return sv, w, obs


def _run(sv, w, obs, **kw):
    out = []
    with context(sv, "fast", max_batch=2, **kw) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(sv.vp, sv.vs, sv.rho)
        out.append(ctx.batch_forward([lib_shot(sv, sh) for sh in sv.shots]))
        ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
        out.append(ctx.batch_gradient([lib_shot(sv, sh, *obs[k]) for k, sh in enumerate(sv.shots)]))
        out.append(ctx.flush())
        d = (sv.vp - sv.vp0).astype(np.float32)
        out.append(ctx.batch_linesearch_parabola([lib_shot(sv, sh, *obs[k]) for k, sh in enumerate(sv.shots)],
                                                 d, 0.5 * d, 0.1, 0.3))
    return out


def _flat(out):
    for x in out:
        if isinstance(x, (tuple, list)):
            yield from _flat(x)
        else:
            yield np.asarray(x)


def _same(a, b):
    fa, fb = list(_flat(a)), list(_flat(b))
    assert len(fa) == len(fb)
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y), "max |d| = %g" % np.abs(x.astype(np.float64) - y).max()


@pytest.mark.parametrize("nz,nt,cols", [(40, 200, 16), (40, 201, 8), (134, 201, 37), (134, 200, 16)])
def test_fused_step_equals_half_steps(nz, nt, cols):
    sv, w, obs = _case(nz, nt, seed=nz + nt)
    ref = _run(sv, w, obs, march_cols=8, fuse_step=-1)
    assert np.abs(ref[0][0]).max() > 0 and np.abs(ref[2][0]).max() > 0
    _same(_run(sv, w, obs, march_cols=8, fuse_step=1, step_cols=cols), ref)
    _same(_run(sv, w, obs, march_cols=-1), ref)


# Workgroup form: one 240-row tile (nz 40, 134), two (nz 300: 240 + 68 rows)
# and three with a 16-row last tile (nz 488), segment widths that do and do
# not divide the grid, and each prefetch depth.
@pytest.mark.parametrize("nz,nt,cols,depth", [(40, 200, 32, 2), (134, 201, 37, 1), (300, 201, 32, 2),
                                              (300, 200, 16, 4), (488, 200, 24, 2)])
def test_workgroup_fused_step_equals_half_steps(nz, nt, cols, depth):
    sv, w, obs = _case(nz, nt, seed=nz + nt)
    ref = _run(sv, w, obs, march_cols=8, march_depth=depth, fuse_step=-1)
    assert np.abs(ref[0][0]).max() > 0 and np.abs(ref[2][0]).max() > 0
    _same(_run(sv, w, obs, march_cols=8, march_depth=depth, fuse_step=2, step_cols=cols), ref)


@pytest.mark.parametrize("fuse", [1, 2])
@pytest.mark.parametrize("extra", [dict(storage="bf16"), dict(cpu_mem=True, strip_chunk=30),
                                   dict(recon="snapshots", storage="f16", snap_every=3),
                                   dict(use_graph=True), dict(rbc=True, rbc_seed=3)])
def test_fused_step_with_storage_and_graphs(extra, fuse):
    sv, w, obs = _case(134 if fuse == 1 else 300, 201, seed=5)
    ref = _run(sv, w, obs, march_cols=8, fuse_step=-1, **extra)
    _same(_run(sv, w, obs, march_cols=8, fuse_step=fuse, **extra), ref)
    if extra.get("use_graph"):   # replays of the captured solves
        with context(sv, "fast", max_batch=2, march_cols=8, fuse_step=fuse, use_graph=True) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            shots = [lib_shot(sv, sh, *obs[k]) for k, sh in enumerate(sv.shots)]
            for _ in range(3):
                m = ctx.batch_gradient(shots)
                g = ctx.flush()
                assert np.array_equal(m, ref[1])
                for a, b in zip(g, ref[2]):
                    assert np.array_equal(a, b)


# Snapshot gradients with fuse_step 3 also run each reverse step as one
# workgroup march of the adjoint pair (adjoint state double-buffered, the
# imaging before the step): gradient, stage-2 Born gradient and Hessian-vector
# product equal the half-step kernels bit for bit, with one and two row tiles
# and even and odd record lengths (the adjoint state ends in either buffer).
@pytest.mark.parametrize("nz,nt,storage", [(40, 200, "f16"), (300, 201, "f32"), (300, 200, "bf16")])
def test_workgroup_fused_adjoint_equals_half_steps(nz, nt, storage):
    sv, w, obs = _case(nz, nt, seed=7 + nz + nt)
    # a shallow perturbation, so the scattered field reaches the receivers
    # within the record on every grid
    X, Z = np.meshgrid(np.arange(sv.nx), np.arange(sv.allnz), indexing="ij")
    bump = np.exp(-((X - sv.nx / 2) ** 2 + (Z - 10) ** 2) / (2 * 4.0 ** 2))
    d_vp = (60.0 * bump).astype(np.float32)
    d_vs = (35.0 * bump).astype(np.float32)
    res = {}
    for fuse in (-1, 3):
        kw = dict(march_cols=8, fuse_step=fuse, recon="snapshots", storage=storage, snap_every=4)
        out = []
        with context(sv, "fast", max_batch=2, **kw) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            shots = [lib_shot(sv, sh, *obs[k]) for k, sh in enumerate(sv.shots)]
            out.append(ctx.batch_gradient(shots))
            out.append(ctx.flush())
            ctx.set_perturbation(d_vp, d_vs)
            out.append(ctx.batch_gradient(shots, born=True))
            out.append(ctx.flush())
            ctx.hessian_vector(shots)
            out.append(ctx.flush())
        res[fuse] = out
    for k in (1, 3, 4):   # gradient, Born gradient, Hessian-vector product images
        assert np.abs(res[-1][k][0]).max() > 0
    _same(res[3], res[-1])


# The workgroup marches add the source, record the traces and inject the
# adjoint source themselves only when every source and receiver lies in the
# update domain and no two receivers share a cell; otherwise the separate
# kernels do it around them.  Coincident receivers take that second path:
# forward traces (every receiver of a shared cell records the same sample),
# snapshot gradient and misfits still equal the half-step kernels bit for bit.
def test_workgroup_fused_steps_with_coincident_receivers():
    sv = W.tiny(nx=52, nz=134, pml=8, nt=201, nshots=2, seed=11)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    for sh in sv.shots:   # three receivers twice
        sh.gc = np.concatenate([sh.gc, sh.gc[5:8]]).astype(np.int32)
    obs = [O.forward(oracle_cfg(sv, sh), sv.vp, sv.vs, sv.rho, w, sh.gc, "f32") for sh in sv.shots]
    res = {}
    for fuse in (-1, 3):
        out = []
        with context(sv, "fast", max_batch=2, ntr=len(sv.shots[0].gc), march_cols=8, fuse_step=fuse,
                     recon="snapshots", storage="f16", snap_every=4) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp, sv.vs, sv.rho)
            out.append(ctx.batch_forward([lib_shot(sv, sh) for sh in sv.shots]))
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            out.append(ctx.batch_gradient([lib_shot(sv, sh, *obs[k]) for k, sh in enumerate(sv.shots)]))
            out.append(ctx.flush())
        res[fuse] = out
    z = res[3][0][0][0]
    assert np.array_equal(z[5:8], z[-3:])   # the shared cells' receivers recorded alike
    assert np.abs(res[-1][2][0]).max() > 0
    _same(res[3], res[-1])
