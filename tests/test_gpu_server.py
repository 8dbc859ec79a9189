"""Serving loop: requests of different grids, record lengths and kinds,
submitted together, give exactly the results of the same solves run alone;
a tight HBM budget serialises admission without changing results; a request
larger than the budget is refused at submission."""
import numpy as np
import pytest

from sfwi_amd import workloads as W
from sfwi_amd._lib import SfwiError
from sfwi_amd.elastic import Context, Server
from tests.helpers import lib_shot

pytestmark = pytest.mark.gpu


def _requests():
    reqs = []
    specs = [("gradient", dict(nx=48, nz=40, nt=200, seed=3), 2),
             ("forward", dict(nx=56, nz=44, nt=240, seed=4), 2),
             ("born_gradient", dict(nx=48, nz=40, nt=220, seed=5), 1),
             ("gradient", dict(nx=64, nz=48, nt=180, seed=6), 1),
             ("born_forward", dict(nx=52, nz=40, nt=200, seed=8), 2),
             ("gradient", dict(nx=48, nz=40, nt=200, seed=9), 2)]
    for kind, kw, nsh in specs:
        sv = W.tiny(nshots=nsh, **kw)
        w = W.ricker(sv.nt, sv.dt, sv.f0)
        born = kind.startswith("born")
        params = dict(dx=sv.dx, dz=sv.dz, dt=sv.dt, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt, ntr=sv.ntr,
                      allnx=sv.allnx, allnz=sv.allnz, max_batch=nsh,
                      rho_const=2000.0 if born else 0.0, adjoint_dir=-1 if born else 1)
        with Context(**params) as ctx:          # observed traces from the true model
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp, sv.vs, sv.rho)
            oz, ox = ctx.batch_forward([lib_shot(sv, s) for s in sv.shots])
        shots = [lib_shot(sv, s, oz[k][:len(s.gc)], ox[k][:len(s.gc)]) for k, s in enumerate(sv.shots)]
        dvp, dvs = ((sv.vp - sv.vp0) * 0.5, (sv.vs - sv.vs0) * 0.5) if born else (None, None)
        reqs.append(dict(kind=kind, params=params, shots=shots, wavelet=w, vp=sv.vp0, vs=sv.vs0, rho=sv.rho0,
                         dvp=dvp, dvs=dvs))
    return reqs


def _alone(r):
    with Context(**r["params"]) as ctx:
        ctx.set_wavelet(r["wavelet"])
        ctx.set_model(r["vp"], r["vs"], r["rho"])
        if r["dvp"] is not None:
            ctx.set_perturbation(r["dvp"], r["dvs"])
        if r["kind"] == "forward":
            z, x = ctx.batch_forward(r["shots"])
            return {"syn_z": z, "syn_x": x}
        if r["kind"] == "born_forward":
            z = np.zeros((len(r["shots"]), r["params"]["ntr"], r["params"]["nt"]), np.float32)
            x = np.zeros_like(z)
            for k, s in enumerate(r["shots"]):
                bz, bx, _, _ = ctx.born_forward(s)
                z[k, :s.ntr], x[k, :s.ntr] = bz, bx
            return {"syn_z": z, "syn_x": x}
        m = ctx.batch_gradient(r["shots"], born=r["kind"] == "born_gradient")
        gvp, gvs, gil = ctx.flush()
        return {"misfits": m, "grad_vp": gvp, "grad_vs": gvs, "illum": gil}


def _submit(srv, r):
    return srv.submit(r["kind"], r["shots"], r["wavelet"], r["vp"], r["vs"], r["rho"], r["dvp"], r["dvs"],
                      **r["params"])


def _same(a, b):
    for k, v in b.items():
        assert np.array_equal(a[k], v), k


def test_mixed_requests_equal_alone():
    reqs = _requests()
    alone = [_alone(r) for r in reqs]
    with Server(workers=3) as srv:
        tickets = [_submit(srv, r) for r in reqs]
        # a second wave while the first may still run (cached contexts reused)
        tickets += [_submit(srv, r) for r in reqs[:3]]
        outs = [srv.wait(t) for t in tickets]
    for o, a in zip(outs, alone + alone[:3]):
        _same(o, a)
        assert o["run_ms"] > 0 and o["queue_ms"] >= 0


def test_budget_serialises_admission_and_refuses_oversized():
    reqs = _requests()[:3]
    alone = [_alone(r) for r in reqs]
    need = max(Server.solve_bytes(r["kind"], len(r["shots"]), **r["params"]) for r in reqs)
    with Server(budget_bytes=need, workers=3) as srv:
        tickets = [_submit(srv, r) for r in reqs]
        outs = [srv.wait(t) for t in tickets]
        for o, a in zip(outs, alone):
            _same(o, a)
        big = dict(reqs[0])
        sv = W.tiny(nx=200, nz=150, nt=400, nshots=2)
        big["params"] = dict(reqs[0]["params"], nx=sv.nx, nz=sv.allnz, allnx=sv.allnx, allnz=sv.allnz, nt=400,
                             ntr=sv.ntr)
        big.update(vp=sv.vp, vs=sv.vs, rho=sv.rho, wavelet=W.ricker(400, sv.dt, sv.f0))
        with pytest.raises(SfwiError, match="budget"):
            _submit(srv, big)


def _same_shape_requests(kind, n, max_batch=6, **extra):
    """n requests of one shape (grid, record length, max_batch) with their own
    models, wavelets, perturbations and shot counts."""
    reqs = []
    for j in range(n):
        sv = W.tiny(nx=48, nz=40, nt=200, seed=40 + j, f0=20.0 + 2 * j, nshots=1 + j % 2)
        w = W.ricker(sv.nt, sv.dt, sv.f0)
        born = kind.startswith("born")
        params = dict(dx=sv.dx, dz=sv.dz, dt=sv.dt, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt, ntr=sv.ntr,
                      allnx=sv.allnx, allnz=sv.allnz, max_batch=max_batch,
                      rho_const=2000.0 if born else 0.0, adjoint_dir=-1 if born else 1, **extra)
        with Context(**params) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp, sv.vs, sv.rho)
            oz, ox = ctx.batch_forward([lib_shot(sv, s) for s in sv.shots])
        shots = [lib_shot(sv, s, oz[k][:len(s.gc)], ox[k][:len(s.gc)]) for k, s in enumerate(sv.shots)]
        dvp, dvs = ((sv.vp - sv.vp0) * 0.5, (sv.vs - sv.vs0) * 0.5) if born else (None, None)
        reqs.append(dict(kind=kind, params=params, shots=shots, wavelet=w, vp=sv.vp0, vs=sv.vs0, rho=sv.rho0,
                         dvp=dvp, dvs=dvs))
    return reqs


@pytest.mark.parametrize("kind,extra", [("forward", {}), ("born_forward", {}), ("gradient", {}),
                                        ("born_gradient", {}),
                                        ("gradient", dict(recon="snapshots", storage="f16", snap_every=3)),
                                        ("gradient", dict(mode="exact")),
                                        ("gradient", dict(misfit="ot_trace")),
                                        ("born_gradient", dict(march_cols=8, storage="bf16"))])
def test_same_shape_requests_share_a_batch(kind, extra):
    """Requests of one shape and kind queued behind a running one join one
    batch (one worker: the first runs alone, the next ones queue and merge up
    to max_batch = 6 shots); every request's result equals the request solved
    alone bit for bit, each with its own model, wavelet and perturbation."""
    reqs = _same_shape_requests(kind, 6, **extra)
    alone = [_alone(r) for r in reqs]
    with Server(workers=1) as srv:
        tickets = [_submit(srv, r) for r in reqs]
        outs = [srv.wait(t) for t in tickets]
    for o, a in zip(outs, alone):
        _same(o, a)
    # merged requests start and end together
    runs = [o["run_ms"] for o in outs]
    assert any(runs.count(x) >= 2 for x in runs), runs
    with Server(workers=1, batching=False) as srv:   # one request per batch: the same results
        outs = [srv.wait(t) for t in [_submit(srv, r) for r in reqs]]
    for o, a in zip(outs, alone):
        _same(o, a)
    # one worker, FIFO, no merging: request j starts after request i < j ends,
    # so its wait exceeds i's by i's run less the (sub-millisecond) submission
    # gap; merged requests share a start, where the later one waits less
    waits = [o["queue_ms"] for o in outs]
    assert all(waits[j] > waits[i] for i in range(len(outs)) for j in range(i + 1, len(outs))), waits
