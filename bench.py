#!/usr/bin/env python3
"""Benchmark of the per-shot elastic FWI gradient path on one MI355X.

Workload (default): the shot-batch survey -- 401 x 301 grid, dx = dz = 10 m,
seeded 6-layer model with a 15-degree dipping interface, dt = 1 ms,
nt = 3000, 30-cell PML, 64 shots with a 12 Hz Ricker wavelet and 401
two-component receivers each.  Synthetic data: observed traces come from a
forward solve through the true model (untimed); gradients are taken at the
smoothed starting model.

One step = one batched gradient of all shots (forward with strip storage,
misfit, reverse reconstruction + adjoint + zero-lag imaging), inputs resident
in HBM (the observed traces stay resident from the warm-up call).  value =
shot-gradients per second over the timed steps; pcie_inclusive_value is one
more step that uploads the observed traces from host memory.

Also reported: Gcell-updates/s (3 propagations per gradient), p50/p99 of the
step latency, the gradient error of two shots against the FP64 oracle, a
roofline entry for the dominant kernel (HIP-event timings taken inside the
timed steps on the library's stream), the plain one-shot path on the same GPU
(one shot per call, per-cell kernels; FP32 and reference operation order) and
a CPU baseline (the FP64 oracle on the host cores, bounded 2-shot sample of
the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "shot-gradients/s & Gcell-updates/s at ≤1e-3 grad err vs FP64; p50/p99 latency"
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the default (column-march, 8 columns, depth 1) FAST
# kernels on this workload's 461 x 361 padded grid, per shot in the launch:
# FETCH_SIZE x 2 + WRITE_SIZE x 1 from rocprofv3 --pmc passes over a 64-shot
# batch, the factors from a plain 1 GiB copy (profiles/shot_batch_march_hbm_traffic.md).
MEASURED_HBM_BYTES_PER_SHOT = {
    "k_vel(forward)": 481.7e6 / 64,
    "k_str(forward)": 582.2e6 / 64,
    "k_str(reconstruct)": 795.5e6 / 64,
    "k_vel(reconstruct)": 480.1e6 / 64,
    "k_adj_str": 516.8e6 / 64,
    "k_adj_vel_image": 1146.7e6 / 64,
}


def alg_bytes(kernel, n_int, n_pml):
    """Algorithmic bytes of one launch over one shot: every array the kernel
    needs read once and every array it produces written once, FP32
    (interior cells: totals only; PML cells: plus split parts and damping)."""
    per = {  # (interior B/cell, PML B/cell)
        "k_vel(forward)": (32, 72),        # R txx tzz txz vx vz rho, W vx vz (+4 splits RW, 2 damp)
        "k_str(forward)": (40, 96),        # R vx vz txx tzz txz lam mu, W 3 stresses (+6 splits RW, 2 damp)
        "k_str(reconstruct)": (56, 112),   # as forward + W 4 velocity derivatives
        "k_vel(reconstruct)": (32, 72),
        "k_adj_str": (32, 72),             # R 2 v + 3 s, W 3 s (+4 splits RW, 2 damp)
        "k_adj_vel_image": (96, 80),       # R 3 s, lam mu rho, 2 v, W 2 v; imaging: R 4 derivs,
                                           # 2 coeffs, 2 recon v, RW 3 images (interior only)
    }[kernel]
    return per[0] * n_int + per[1] * n_pml


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--nt", type=int, default=3000)
    ap.add_argument("--mode", default="fast", choices=["fast", "exact"])
    ap.add_argument("--timing-every", type=int, default=20)
    ap.add_argument("--march-cols", type=int, default=0,
                    help="FAST kernels: columns per wave segment (0: library default, <0: per-cell kernels)")
    ap.add_argument("--march-depth", type=int, default=0, help="FAST march kernels: prefetch depth")
    ap.add_argument("--march-order", type=int, default=0,
                    help="FAST march kernels: block order (0 default, 1 shot after shot, 2 shots interleaved)")
    ap.add_argument("--cpu-shots", type=int, default=2)
    ap.add_argument("--plain-shots", type=int, default=4,
                    help="shots timed through the plain one-shot path (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X: the gradient path runs a shot batch on a single GPU "
              "(--gpus %d requested). Multi-GPU sharding is a design note only." % args.gpus,
              file=sys.stderr, flush=True)
        sys.exit(2)

    from sfwi_amd import workloads as W
    from sfwi_amd import _lib
    from sfwi_amd.elastic import Context, Shot

    sv = W.shot_batch(nshots=args.shots, nt=args.nt)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    geo = [W.shot_geometry(sv, s) for s in sv.shots]
    n_int = sv.nx * sv.allnz
    n_cells = sv.cells()
    n_pml = n_cells - n_int
    print("bench: %s, %d shots, grid %dx%d pml %d (%d cells), nt %d" % (
        sv.name, len(sv.shots), sv.nx, sv.allnz, sv.pml, n_cells, sv.nt), file=sys.stderr, flush=True)

    ctx = Context(dx=sv.dx, dz=sv.dz, dt=sv.dt, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt, ntr=sv.ntr,
                  allnx=sv.allnx, allnz=sv.allnz, max_batch=len(sv.shots), mode=args.mode,
                  march_cols=args.march_cols, march_depth=args.march_depth, march_order=args.march_order)
    ctx.set_wavelet(w)
    ctx.set_model(sv.vp, sv.vs, sv.rho)
    t0 = time.perf_counter()
    obs_z, obs_x = ctx.batch_forward([Shot(s.left, *g, s.ngx_left, s.gc) for s, g in zip(sv.shots, geo)])
    print("bench: observed data %.2f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
    shots = [Shot(s.left, *g, s.ngx_left, s.gc, obs_z[k], obs_x[k])
             for k, (s, g) in enumerate(zip(sv.shots, geo))]

    for k in range(args.warmup):
        t0 = time.perf_counter()
        ctx.batch_gradient(shots)
        ctx.flush()
        print("bench: warmup %d %.3f s" % (k, time.perf_counter() - t0), file=sys.stderr, flush=True)

    # timed steps: observed traces already resident in HBM (uploaded by the
    # warm-up call); the shots pass NULL records and reuse their batch slot's
    resident = [Shot(s.left, *g, s.ngx_left, s.gc) for s, g in zip(sv.shots, geo)]
    if args.warmup == 0:
        ctx.batch_gradient(shots)
        ctx.flush()
    ctx.set_timing(args.timing_every)
    times = []
    misfit = 0.0
    for k in range(args.steps):
        t0 = time.perf_counter()
        m = ctx.batch_gradient(resident)       # returns after the stream is drained
        times.append(time.perf_counter() - t0)
        misfit = float(m.sum())
        print("bench: step %d %.3f s" % (k, times[-1]), file=sys.stderr, flush=True)
    kstats = {kid: ctx.kernel_time(kid) for kid in range(_lib.K_COUNT)}
    ctx.set_timing(0)
    ctx.flush()
    # one step that uploads the observed traces from host memory (PCIe-inclusive rate)
    t0 = time.perf_counter()
    ctx.batch_gradient(shots)
    t_pcie = time.perf_counter() - t0
    ctx.flush()

    total = sum(times)
    nsh = len(shots)
    value = nsh * args.steps / total
    gcells = nsh * n_cells * sv.nt * 3 * args.steps / total / 1e9

    # per-kernel roofline from the HIP events recorded inside the timed steps
    kernels = {}
    for kid in range(_lib.K_COUNT):
        ms, n = kstats[kid]
        if n == 0:
            continue
        name = _lib.load().sfwi_kernel_name(kid).decode()
        b = alg_bytes(name, n_int, n_pml) * nsh
        kernels[name] = {"avg_ms": round(ms, 5), "samples": n, "alg_bytes": b,
                         "GBps": round(b / (ms * 1e-3) / 1e9, 1),
                         "share": 0.0}
    tot_ms = sum(v["avg_ms"] for v in kernels.values())
    for v in kernels.values():
        v["share"] = round(v["avg_ms"] / tot_ms, 3) if tot_ms else 0.0
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"]) if kernels else None
    roofline = None
    if dom:
        d = kernels[dom]
        measured = (args.mode == "fast" and args.march_cols in (0, 8) and args.march_depth in (0, 1))
        traffic = round(MEASURED_HBM_BYTES_PER_SHOT[dom] * nsh) if measured else None
        roofline = {"kernel": dom, "bound": "hbm", "achieved": d["GBps"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(d["GBps"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "alg_bytes_per_launch": d["alg_bytes"], "avg_launch_ms": d["avg_ms"]}

    # the plain path on the same GPU: one shot per call, per-cell kernels
    plain = None
    if args.plain_shots > 0:
        plain = {}
        for label, mode in (("fp32_percell", "fast"), ("reference_order", "exact")):
            pc = Context(dx=sv.dx, dz=sv.dz, dt=sv.dt, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt,
                         ntr=sv.ntr, allnx=sv.allnx, allnz=sv.allnz, max_batch=1, mode=mode, march_cols=-1)
            pc.set_wavelet(w)
            pc.set_model(sv.vp0, sv.vs0, sv.rho0)
            pc.gradient(shots[0])                      # warm-up
            k = min(args.plain_shots, nsh)
            t0 = time.perf_counter()
            for sh in shots[:k]:
                pc.gradient(sh)
            tp = time.perf_counter() - t0
            pc.flush()
            pc.close()
            plain[label] = {"value": round(k / tp, 4), "shots": k}
            print("bench: plain one-shot path (%s) %.3f s per shot" % (label, tp / k), file=sys.stderr,
                  flush=True)
        plain["unit"] = "shot-gradients/s"
        plain["speedup_of_batched"] = round(value / plain["fp32_percell"]["value"], 2)

    # accuracy of two shots against the FP64 oracle (and the CPU baseline)
    acc = None
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        ncpu = max(1, min(args.cpu_shots, nsh))
        sub = shots[:ncpu]
        ctx.batch_gradient(sub)
        g_vp, g_vs, g_il = ctx.flush()
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        r_vp = np.zeros((sv.allnx, sv.allnz))
        r_vs = np.zeros_like(r_vp)
        t0 = time.perf_counter()
        for s, g, sh in zip(sv.shots[:ncpu], geo[:ncpu], sub):
            c = O.config(allnx=sv.allnx, allnz=sv.allnz, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt,
                         ntr=sv.ntr, dx=sv.dx, dz=sv.dz, dt=sv.dt, left=s.left, isx=g[0], isz=g[1],
                         igz=g[2], ngx_left=s.ngx_left, nthreads=threads)
            a = O.gradient(c, sv.vp0, sv.vs0, sv.rho0, w, s.gc, sh.obs_z, sh.obs_x, "f64")
            r_vp += a[0]
            r_vs += a[1]
        tc = time.perf_counter() - t0

        def rel(x, y):
            return float(np.linalg.norm(x.astype(np.float64) - y) / np.linalg.norm(y))

        acc = {"grad_vp": rel(g_vp, r_vp), "grad_vs": rel(g_vs, r_vs), "shots": ncpu,
               "bar": 1e-3}
        cpu = {"value": round(ncpu / tc, 5), "unit": "shot-gradients/s", "cores": threads,
               "kind": "port",
               "sample": "FP64 oracle (C++ restatement, OpenMP over x columns), %d of the %d shots "
                         "of the same workload, %.1f s" % (ncpu, nsh, tc)}

    lat = sorted(times)
    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "shot-gradients/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * total / args.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded model, observed traces from a forward solve of the true model)",
        "config": {"workload": "shot_batch: 401x301 dx=dz=10m, pml 30, dt 1ms, nt %d, %d shots, 12 Hz, "
                               "401 receivers, batched gradient" % (sv.nt, nsh),
                   "mode": args.mode, "storage": "boundary strips FP32"},
        "gcell_updates_per_s": round(gcells, 2),
        "pcie_inclusive_value": round(len(shots) / t_pcie, 4),
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 2),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 2), "unit": "per batch step"},
        "grad_err_vs_fp64": acc,
        "misfit": misfit,
        "roofline": roofline,
        "plain_one_shot_path": plain,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
