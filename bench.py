#!/usr/bin/env python3
"""Benchmark of the per-shot elastic FWI gradient path on one MI355X.

Workload (default): the mixed-precision survey -- 1200 x 600 grid,
dx = dz = 5 m, a seeded random 8-layer model with three Gaussian inclusions of
+-8 %, dt = 0.5 ms, nt = 6000, 40-cell PML, 96 shots with a 15 Hz Ricker
wavelet and 1200 two-component receivers each.  Gradients are taken with the
snapshot reconstruction: the forward velocity is kept every 4 steps in FP16
with one max-abs scale per snapshot and field, and the reverse pass runs the
adjoint and images from the snapshots (FP32 stencils and accumulation).  FP16,
not BF16: on this grid the imaging's derivatives of BF16-rounded velocities
put the gradient 2.2e-3 / 3.4e-3 (Vp / Vs) from FP64, outside the 1e-3 bar,
while FP16 gives 2.8e-4 / 4.3e-4 for the same bytes (--storage bf16 runs the
BF16 form).  The snapshots of all 96 shots (4.4 GB each) exceed one GPU's
288 GB, so the survey
runs in waves of 24 shots: one step = one batched gradient of one wave, the
waves rotating through the survey.  Synthetic data: observed traces come from
a forward solve through the true model (untimed) and stay resident in HBM
(sfwi_obs_put); gradients are taken at the smoothed starting model.

value = shot-gradients per second over the timed steps.  Also reported:
Gcell-updates/s (the propagations actually run: 2 per snapshot gradient,
3 per strip gradient), p50/p99 of the step latency, the trace, misfit and
gradient error of one shot against the FP64 oracle of the same algorithm, a
roofline entry for the dominant kernel (HIP-event timings inside the timed
steps on the library's stream; bytes per the scope table's unit costs) and
for the whole time step, the plain one-shot path on the same GPU (one shot
per call, per-cell kernels, FP32 boundary strips: the reference's scheme),
and a CPU baseline (the same FP64 oracle run, timed on the host cores).
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

# HBM bytes per launch (FETCH_SIZE x 2 + WRITE_SIZE x 1 from two rocprofv3
# --pmc passes, the factors from a plain copy of known size) of the default
# benchmark configuration's kernels, per shot of the launch; see
# profiles/mixed_precision_snapshot_*kernel_trace_and_hbm_traffic.md.  Keyed by
# (workload, recon, storage, snap_every, time-blocking): half steps (0) or the
# workgroup-fused forward and adjoint steps (3), whose forward step also forms
# the snapshot maxima (the store is then scale + pack only).
_SNAP16 = {
    "k_vel(forward)": 35344157, "k_str(forward)": 46374932, "k_adj_str": 34583001,
    "k_adj_vel": 48443475, "k_snap_image": 30838385, "k_snap_store": 16727600,
}
_SNAP16_FUSED = {   # forward 16-column segments, adjoint 32 (old stresses kept in LDS), prefetch depth 1,
    # non-temporal state stores
    "k_step(forward)": 79759420, "k_step(adjoint)": 73468830, "k_snap_image": 30848653, "k_snap_store": 9305933,
}
# measured with FP16 snapshots; BF16 snapshots move the same bytes with the same kernels (within 0.1 %)
MEASURED_HBM_BYTES_PER_SHOT = {
    ("mixed_precision", "snapshots", "bf16", 4, 0): _SNAP16,
    ("mixed_precision", "snapshots", "f16", 4, 0): _SNAP16,
    ("mixed_precision", "snapshots", "bf16", 4, 3): _SNAP16_FUSED,
    ("mixed_precision", "snapshots", "f16", 4, 3): _SNAP16_FUSED,
}

NOP, NSTRIP = 4, 4   # stencil half-width; strip half-width (split band = PML + NSTRIP)


def cell_classes(nx, nz, pml):
    """Cells of one shot that a stencil launch touches: (unsplit, split, imaged).
    The update domain is the padded grid minus its outer NOP-cell ring (never
    written); split cells are those of it within pml + NSTRIP of an edge; the
    imaged cells are the nx x nz interior (some of them split)."""
    nxp, nzp = nx + 2 * pml, nz + 2 * pml
    upd = (nxp - 2 * NOP) * (nzp - 2 * NOP)
    w = pml + NSTRIP
    un = max(nxp - 2 * w, 0) * max(nzp - 2 * w, 0)
    return un, upd - un, nx * nz


def alg_bytes(kernel, nx, nz, pml):
    """Bytes of one launch over one shot by the kernels' own cell classes:
    every array the kernel needs read once and every array it produces written
    once, FP32.  An unsplit cell reads its own totals; a split cell reads and
    writes its split parts and the damping, writes the totals, and does not
    read its own totals; stencil inputs are totals everywhere; imaging terms
    count on the imaged cells."""
    per = {  # (unsplit B/cell, split B/cell, imaged B/cell)
        "k_vel(forward)": (32, 64, 0),
        "k_str(forward)": (40, 84, 0),
        "k_str(reconstruct)": (56, 100, 0),
        "k_vel(reconstruct)": (32, 64, 0),
        "k_adj_str": (32, 60, 0),
        "k_adj_vel_image": (40, 72, 48),
        "k_str(reconstruct+image)": (40, 84, 28),
        "k_vel(reconstruct+illum)": (32, 64, 8),
        "k_adj_vel": (40, 72, 0),
    }.get(kernel)
    if per is None:
        return None
    un, sp, im = cell_classes(nx, nz, pml)
    b = per[0] * un + per[1] * sp + per[2] * im
    if kernel == "k_vel(forward)":
        b += 2 * 4 * NSTRIP * (nx + nz) * 4        # W the boundary strips of vx, vz
    return b


# Unit costs of the hot-path scope table (FP32, one read + one write of the
# minimal state, no temporal reuse): a full cell-update is 52 B on an interior
# cell (5 fields read and written, 3 medium parameters) and 100 B on a PML cell
# (10 split components, 3 parameters, 2 damping profiles); imaging adds 24 B per
# interior cell per gradient step.  Per half-step kernel the update splits as
# velocity 20 / 36 B (vx vz read+write, 1/rho; PML: 4 split parts, rho) and
# stress 32 / 64 B (3 stresses read+write, lambda mu; PML: 6 split parts, 2
# parameters, the 2 damping profiles), so the halves add up to the table's
# 52 / 100 B; interior = nx*nz cells, PML = the rest of the padded grid.
UNIT_8D = {"vel": (20, 36), "str": (32, 64)}


def alg_bytes_8d(kernel, nx, nz, pml, store_bytes=2):
    """Scope-table bytes of one launch of `kernel` over one shot."""
    inter = nx * nz
    pmlc = (nx + 2 * pml) * (nz + 2 * pml) - inter
    if kernel in ("k_step(forward)", "k_step(adjoint)"):   # a whole cell-update
        return 52 * inter + 100 * pmlc
    half = {"k_vel(forward)": "vel", "k_adj_vel": "vel", "k_vel(reconstruct+illum)": "vel",
            "k_vel(reconstruct)": "vel", "k_adj_vel_image": "vel",
            "k_str(forward)": "str", "k_adj_str": "str", "k_str(reconstruct+image)": "str",
            "k_str(reconstruct)": "str"}.get(kernel)
    if half:
        a, b = UNIT_8D[half]
        extra = {"k_str(reconstruct+image)": 16, "k_vel(reconstruct+illum)": 8, "k_adj_vel_image": 24}.get(kernel, 0)
        return a * inter + b * pmlc + extra * inter
    rn = (nx + 2 * NOP) * (nz + 2 * NOP)
    if kernel == "k_snap_image":       # R the snapshot (2 fields) and 3 adjoint stresses, RW 3 accumulators
        return inter * (2 * store_bytes + 12 + 24)
    if kernel == "k_snap_store":       # R vx vz, W the snapshot of the region
        return rn * (8 + 2 * store_bytes)
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _survey(args, W):
    kw = {}
    if args.shots:
        kw["nshots"] = args.shots
    if args.nt:
        kw["nt"] = args.nt
    return W.BY_NAME[args.workload](**kw)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="mixed_precision", choices=["mixed_precision", "shot_batch", "plain_small"],
                    help="mixed_precision: 1200x600, dx 5 m, nt 6000, 96 shots (default); shot_batch: 401x301, "
                         "nt 3000, 64 shots; plain_small: 376x201, nt 2000, 16 shots")
    ap.add_argument("--shots", type=int, default=0, help="shots (0: the workload's)")
    ap.add_argument("--wave", type=int, default=0,
                    help="shots per step (0: 24 for mixed_precision, whose snapshots of 96 shots exceed HBM; "
                         "all shots otherwise)")
    ap.add_argument("--nt", type=int, default=0, help="time samples (0: the workload's)")
    ap.add_argument("--recon", default=None, choices=["strips", "snapshots"],
                    help="forward-field reconstruction (default: snapshots for mixed_precision, strips otherwise)")
    ap.add_argument("--snap-every", type=int, default=4, help="snapshots: steps between snapshots")
    ap.add_argument("--storage", default=None, choices=["f32", "f16", "bf16"],
                    help="strip or snapshot storage (default: f16 for mixed_precision, f32 otherwise)")
    ap.add_argument("--misfit", default="l2", choices=["l2", "ot_trace", "ot_power", "ot_gather"],
                    help="misfit of the gradient (l2: the reference's; ot_*: optimal-transport variants)")
    ap.add_argument("--cpu-mem", type=int, default=0, metavar="CHUNK",
                    help="boundary strips in pinned host memory, CHUNK steps per chunk (0: strips in HBM)")
    ap.add_argument("--rbc", action="store_true",
                    help="random boundary instead of the boundary strips (the reference's rbc)")
    ap.add_argument("--mode", default="fast", choices=["fast", "exact"])
    ap.add_argument("--timing-every", type=int, default=21,
                    help="kernel events on every k-th time step (21: prime to the snapshot interval, so both kinds "
                         "of steps are sampled)")
    ap.add_argument("--march-cols", type=int, default=0,
                    help="FAST kernels: columns per wave segment (0: 8, the column-march kernels, also for the "
                         "accuracy leg's single shot; <0: per-cell kernels)")
    ap.add_argument("--march-depth", type=int, default=0, help="FAST march kernels: prefetch depth")
    ap.add_argument("--fuse-step", type=int, default=None,
                    help="FAST column-march forward: 2 one fused launch per time step as a workgroup march (the z "
                         "halo shared through LDS), 3 the same and the snapshot gradient's adjoint steps fused too "
                         "(default for mixed_precision), 1 a wave-level march, 0 / -1 two half-step launches "
                         "(default otherwise)")
    ap.add_argument("--step-cols", type=int, default=0, help="fused step: columns per segment of both marches (0: forward 16, adjoint 32)")
    ap.add_argument("--graph", action="store_true",
                    help="replay each solve from a captured hipGraph (the timed steps carry no kernel events; the "
                         "per-kernel times come from one extra step run directly afterwards)")
    ap.add_argument("--march-order", type=int, default=0,
                    help="FAST march kernels: block order (0 default, 1 shot after shot, 2 shots interleaved, "
                         "3 whole shots per XCD)")
    ap.add_argument("--serving", action="store_true",
                    help="run the serving-loop benchmark (Poisson request stream) instead")
    ap.add_argument("--uneven", action="store_true",
                    help="run the uneven-batch benchmark (120 shots over four grids, autotuned, BF16 strips)")
    ap.add_argument("--inversion", type=int, default=0, metavar="NITER",
                    help="run NITER inversion iterations of the sfwi_stage1 (or, with --stage 2, sfwi_stage2) "
                         "program on the reference's example-1 model (376x201, 38 shots) from synthetic SU "
                         "records instead")
    ap.add_argument("--stage", type=int, default=1, choices=(1, 2), help="inversion: which stage's program")
    ap.add_argument("--born", action="store_true",
                    help="time the stage-2 (Born) gradient of the workload's shot batch instead")
    ap.add_argument("--hessian", action="store_true",
                    help="time the stage-1 Gauss-Newton Hessian-vector product of the workload's shot batch instead")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--rate", type=float, default=1.0, help="serving: requests per second")
    ap.add_argument("--scale", type=float, default=1.0, help="serving: request size factor")
    ap.add_argument("--workers", type=int, default=4, help="serving: concurrent solves")
    ap.add_argument("--shapes", type=int, default=0,
                    help="serving: requests drawn from this many (nx, nz, nt) shapes (0: every request its own)")
    ap.add_argument("--merge-cap", type=int, default=16,
                    help="serving: shots a batch of same-shape requests may hold (the requests' max_batch)")
    ap.add_argument("--cpu-shots", type=int, default=1)
    ap.add_argument("--plain-shots", type=int, default=3,
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

    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.serving:
        return serving(args)
    if args.uneven:
        return uneven(args)
    if args.inversion:
        return inversion(args)
    if args.born or args.hessian:
        return born(args)

    mp = args.workload == "mixed_precision"
    if args.fuse_step is None:   # the library's automatic choice picks 3 on this grid, 0 on the batch's
        args.fuse_step = 3 if mp else 0
    recon = args.recon or ("snapshots" if mp else "strips")
    storage = args.storage or ("f16" if mp else "f32")
    snap = recon == "snapshots"
    sv = _survey(args, W)
    nsh = len(sv.shots)
    wave = min(args.wave or (24 if mp else nsh), nsh)
    waves = [list(range(i, min(i + wave, nsh))) for i in range(0, nsh, wave)]
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    geo = [W.shot_geometry(sv, s) for s in sv.shots]
    n_cells = sv.cells()
    print("bench: %s, %d shots in waves of %d, grid %dx%d pml %d (%d cells), nt %d, %s %s" % (
        sv.name, nsh, wave, sv.nx, sv.allnz, sv.pml, n_cells, sv.nt, recon, storage), file=sys.stderr, flush=True)

    ctx = Context(dx=sv.dx, dz=sv.dz, dt=sv.dt, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt, ntr=sv.ntr,
                  allnx=sv.allnx, allnz=sv.allnz, max_batch=wave, mode=args.mode, storage=storage,
                  march_cols=args.march_cols or 8, march_depth=args.march_depth, march_order=args.march_order,
                  cpu_mem=args.cpu_mem > 0, strip_chunk=max(args.cpu_mem, 0), misfit=args.misfit,
                  rbc=args.rbc, rbc_seed=1, recon=recon, snap_every=args.snap_every, fuse_step=args.fuse_step,
                  step_cols=args.step_cols, use_graph=args.graph)
    ctx.set_wavelet(w)
    ctx.set_model(sv.vp, sv.vs, sv.rho)
    t0 = time.perf_counter()
    obs_z, obs_x = ctx.batch_forward([Shot(s.left, *g, s.ngx_left, s.gc) for s, g in zip(sv.shots, geo)])
    for k in range(nsh):   # the survey's observed traces stay resident in HBM, by shot id
        ctx.obs_put(k + 1, obs_z[k][:len(sv.shots[k].gc)], obs_x[k][:len(sv.shots[k].gc)])
    print("bench: observed data %.2f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
    shots = [Shot(s.left, *g, s.ngx_left, s.gc, obs_id=k + 1) for k, (s, g) in enumerate(zip(sv.shots, geo))]

    def step(j):
        return ctx.batch_gradient([shots[k] for k in waves[j % len(waves)]])

    for k in range(args.warmup):
        t0 = time.perf_counter()
        step(k)
        ctx.flush()
        print("bench: warmup %d %.3f s" % (k, time.perf_counter() - t0), file=sys.stderr, flush=True)
    if not args.graph:   # with graphs the timed steps replay without events
        ctx.set_timing(args.timing_every)
    times = []
    misfit = 0.0
    for k in range(args.steps):
        t0 = time.perf_counter()
        m = step(args.warmup + k)             # returns after the stream is drained
        times.append(time.perf_counter() - t0)
        misfit = float(m.sum())
        print("bench: step %d %.3f s" % (k, times[-1]), file=sys.stderr, flush=True)
    if args.graph:   # kernel times from one extra step run directly
        ctx.set_timing(args.timing_every)
        step(args.warmup + args.steps)
    kstats = {kid: ctx.kernel_time(kid) for kid in range(_lib.K_COUNT)}
    ctx.set_timing(0)
    ctx.flush()
    # one step that uploads its observed traces from host memory (PCIe-inclusive rate)
    host_shots = [Shot(sv.shots[k].left, *geo[k], sv.shots[k].ngx_left, sv.shots[k].gc, obs_z[k], obs_x[k])
                  for k in waves[0]]
    t0 = time.perf_counter()
    ctx.batch_gradient(host_shots)
    t_pcie = time.perf_counter() - t0
    ctx.flush()

    total = sum(times)
    shot_steps = sum(len(waves[(args.warmup + k) % len(waves)]) for k in range(args.steps))
    value = shot_steps / total
    props = 2 if snap else 3
    gcells = shot_steps * n_cells * sv.nt * props / total / 1e9

    # per-kernel roofline from the HIP events recorded inside the timed steps
    sb = 4 if storage == "f32" else 2
    avg_wave = shot_steps / args.steps
    kernels = {}
    for kid in range(_lib.K_COUNT):
        ms, n = kstats[kid]
        if n == 0:
            continue
        name = _lib.load().sfwi_kernel_name(kid).decode()
        b8 = alg_bytes_8d(name, sv.nx, sv.allnz, sv.pml, sb)
        bc = alg_bytes(name, sv.nx, sv.allnz, sv.pml)
        e = {"avg_ms": round(ms, 5), "samples": n}
        if b8 is not None:
            e["alg_bytes"] = round(b8 * avg_wave)
            e["GBps"] = round(b8 * avg_wave / (ms * 1e-3) / 1e9, 1)
        if bc is not None:
            e["cell_class_bytes"] = round(bc * avg_wave)
        kernels[name] = e
    # launches per time step: every kernel once, the snapshot kernels once per snap_every steps
    per_step = {k: (1.0 / args.snap_every if k.startswith("k_snap") else 1.0) for k in kernels}
    step_ms = sum(v["avg_ms"] * per_step[k] for k, v in kernels.items())
    for k, v in kernels.items():
        v["share"] = round(v["avg_ms"] * per_step[k] / step_ms, 3) if step_ms else 0.0
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * per_step[k]) if kernels else None
    roofline = None
    key = (args.workload, recon, storage, args.snap_every if snap else 0, max(args.fuse_step, 0))
    if dom and "GBps" in kernels[dom]:
        d = kernels[dom]
        measured = MEASURED_HBM_BYTES_PER_SHOT.get(key, {}) if (
            args.mode == "fast" and args.march_cols in (0, 8)
            and args.march_order in (0, 3) and not args.rbc and args.cpu_mem == 0 and wave == (24 if mp else nsh)
            and (args.step_cols == 0 or (args.step_cols == 16 and args.fuse_step != 3)) and args.march_depth in (0, 1 if args.fuse_step == 3 else 4)) \
            else {}
        traffic = round(measured[dom] * avg_wave) if dom in measured else None
        step_bytes = sum(v.get("alg_bytes", 0) * per_step[k] for k, v in kernels.items())
        roofline = {"kernel": dom, "bound": "hbm", "achieved": d["GBps"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(d["GBps"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "alg_bytes_per_launch": d["alg_bytes"], "avg_launch_ms": d["avg_ms"],
                    "bytes_model": "scope-table unit costs (interior 52 B, PML 100 B per cell-update, "
                                   "split per half step; imaging operands)",
                    "step": {"alg_bytes": round(step_bytes), "kernel_ms": round(step_ms, 5),
                             "achieved": round(step_bytes / (step_ms * 1e-3) / 1e9, 1),
                             "frac": round(step_bytes / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
        if measured:
            tb = sum(measured.get(k, 0) * avg_wave * per_step[k] for k in kernels)
            roofline["step"]["traffic"] = round(tb)
            roofline["step"]["traffic_over_alg"] = round(tb / step_bytes, 3) if step_bytes else None

    # the plain path on the same GPU: one shot per call, per-cell kernels, FP32 strips
    plain = None
    if args.plain_shots > 0:
        plain = {}
        host0 = [Shot(sv.shots[k].left, *geo[k], sv.shots[k].ngx_left, sv.shots[k].gc, obs_z[k], obs_x[k])
                 for k in range(min(args.plain_shots, nsh))]
        for label, mode in (("fp32_percell", "fast"), ("reference_order", "exact")):
            pc = Context(dx=sv.dx, dz=sv.dz, dt=sv.dt, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt,
                         ntr=sv.ntr, allnx=sv.allnx, allnz=sv.allnz, max_batch=1, mode=mode, march_cols=-1)
            pc.set_wavelet(w)
            pc.set_model(sv.vp0, sv.vs0, sv.rho0)
            pc.gradient(host0[0])                      # warm-up
            t0 = time.perf_counter()
            for sh in host0:
                pc.gradient(sh)
            tp = time.perf_counter() - t0
            pc.flush()
            pc.close()
            plain[label] = {"value": round(len(host0) / tp, 4), "shots": len(host0)}
            print("bench: plain one-shot path (%s) %.3f s per shot" % (label, tp / len(host0)), file=sys.stderr,
                  flush=True)
        plain["unit"] = "shot-gradients/s"
        plain["path"] = "one shot per call, per-cell kernels, FP32 boundary strips (the reference's scheme)"
        plain["speedup_of_batched"] = round(value / plain["fp32_percell"]["value"], 2)

    # accuracy of one shot against the FP64 oracle of the same algorithm (also the CPU baseline)
    acc = None
    cpu = None
    if not args.no_cpu and args.misfit == "l2":
        from oracle import oracle as O
        ncpu = max(1, min(args.cpu_shots, nsh))
        sub = [shots[k] for k in range(ncpu)]
        m_gpu = ctx.batch_gradient(sub)
        g_vp, g_vs, g_il = ctx.flush()
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = int(os.environ.get("OMP_NUM_THREADS") or threads)
        r_vp = np.zeros((sv.allnx, sv.allnz))
        r_vs = np.zeros_like(r_vp)
        r_il = np.zeros_like(r_vp)
        r_m = 0.0
        t0 = time.perf_counter()
        for k in range(ncpu):
            s, g = sv.shots[k], geo[k]
            c = O.config(allnx=sv.allnx, allnz=sv.allnz, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt,
                         ntr=len(s.gc), dx=sv.dx, dz=sv.dz, dt=sv.dt, left=s.left, isx=g[0], isz=g[1],
                         igz=g[2], ngx_left=s.ngx_left, nthreads=threads, rbc=int(args.rbc), rbc_seed=1,
                         snap_every=args.snap_every if snap else 0)
            oz_k, ox_k = obs_z[k][:len(s.gc)], obs_x[k][:len(s.gc)]
            if snap:
                a = O.gradient_snapshot(c, sv.vp0, sv.vs0, sv.rho0, w, s.gc, oz_k, ox_k)
            else:
                a = O.gradient(c, sv.vp0, sv.vs0, sv.rho0, w, s.gc, oz_k, ox_k, "f64")
            r_vp += a[0]
            r_vs += a[1]
            r_il += a[2]
            r_m += a[3]
        tc = time.perf_counter() - t0
        # traces: the GPU's forward solve of the true model against the FP64 oracle's, shot 0
        s, g = sv.shots[0], geo[0]
        c = O.config(allnx=sv.allnx, allnz=sv.allnz, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt,
                     ntr=len(s.gc), dx=sv.dx, dz=sv.dz, dt=sv.dt, left=s.left, isx=g[0], isz=g[1],
                     igz=g[2], ngx_left=s.ngx_left, nthreads=threads)
        t1 = time.perf_counter()
        fz, fx = O.forward(c, sv.vp, sv.vs, sv.rho, w, s.gc, "f64")
        tf = time.perf_counter() - t1

        def rel(x, y):
            return float(np.linalg.norm(np.asarray(x, np.float64) - y) / np.linalg.norm(y))

        nt0 = len(s.gc)
        acc = {"grad_vp": rel(g_vp, r_vp), "grad_vs": rel(g_vs, r_vs), "illum": rel(g_il, r_il),
               "misfit": abs(float(m_gpu.sum()) - r_m) / abs(r_m),
               "traces": rel(np.stack([obs_z[0][:nt0], obs_x[0][:nt0]]), np.stack([fz, fx])),
               "shots": ncpu, "bar": 1e-3,
               "reference": "FP64 oracle of the same algorithm (%s)" % (
                   "snapshot gradient, imaging every %d steps" % args.snap_every if snap else "strip gradient")}
        cpu = {"value": round(ncpu / tc, 5), "unit": "shot-gradients/s", "cores": threads,
               "cpu": cpu_model(), "nproc": os.cpu_count(), "kind": "port",
               "sample": "FP64 oracle (C++ restatement, OpenMP over x columns) of the same gradient, %d of the "
                         "%d shots of the same workload, %.1f s (the FP64 forward solve for the trace check, "
                         "%.1f s, is not counted)" % (ncpu, nsh, tc, tf)}

    lat = sorted(times)
    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "shot-gradients/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * total / args.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded model, observed traces from a forward solve of the true model)",
        "config": {"workload": "%s: %dx%d dx=dz=%gm, pml %d, dt %gms, nt %d, %d shots in waves of %d, %g Hz, "
                               "%d receivers, batched gradient" % (sv.name, sv.nx, sv.allnz, sv.dx, sv.pml,
                                                                    sv.dt * 1e3, sv.nt, nsh, wave, sv.f0, sv.ntr),
                   "mode": args.mode, "misfit": args.misfit,
                   "storage": "random boundary, no strips" if args.rbc else (
                       "velocity snapshots every %d steps, %s with a max-abs scale per snapshot" % (
                           args.snap_every, storage.upper()) if snap else
                       "boundary strips %s%s" % (storage.upper(), ", pinned host memory in chunks of %d steps" %
                                                 args.cpu_mem if args.cpu_mem > 0 else "")),
                   "forward": ("fused whole time steps, %s, %d-column segments%s" % (
                       "workgroup march (z halo shared through LDS)" if args.fuse_step >= 2 else "wave-level march",
                       args.step_cols or 16, "; adjoint steps fused the same way, %d-column segments"
                       % (args.step_cols or 32) if args.fuse_step == 3 and snap else "")
                       if args.fuse_step > 0 else "half-step launches (velocity, stress)"),
                   "launch": ("each solve replayed from a captured hipGraph; kernel times from one extra direct step"
                              if args.graph else "direct launches, kernel events every %d steps" % args.timing_every)},
        "gcell_updates_per_s": round(gcells, 2),
        "gcell_propagations_per_gradient": props,
        "pcie_inclusive_value": round(len(host_shots) / t_pcie, 4),
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 2),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 2), "unit": "per batch step (one wave)"},
        "grad_err_vs_fp64": acc,
        "misfit": misfit,
        "roofline": roofline,
        "plain_one_shot_path": plain,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def copy_calibration(mib):
    """Counter calibration: one plain elementwise kernel that reads `mib` MiB and
    writes `mib` MiB (torch's out = a * 1, after a fill).  FETCH_SIZE and
    WRITE_SIZE of that dispatch, against its known bytes, give the scale
    factors applied to the stencil kernels' counters."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"read_bytes": n * 4, "write_bytes": n * 4}), flush=True)


def born(args):
    """Stage-2 gradient throughput: the workload's shot batch solved as Born
    gradients (sfwi-stage2/FD2DGPU.cu:3978-4570) at the initial model with the
    perturbation = true - initial model, density 2000 and the stage-2 adjoint
    direction; observed traces from a forward solve of the true model.  Timed
    like the default line (records resident in HBM); the gradient of
    --cpu-shots shots is checked against the FP64 oracle, which is also the CPU
    baseline.  The two forward kernels are the Born variants (background and
    scattered fields in one launch); the reverse step is the stage-1 one on
    the background field.  No roofline entry.
    --hessian: the stage-1 Gauss-Newton Hessian-vector product
    (sfwi-stage1/FD2DGPU.cu:5236-5359) of the same batch at the initial model,
    the vector p = true - initial model, the model's density and the stage-1
    direction: the same launches with the scattered traces as the adjoint
    source (no observed traces, no misfit)."""
    hess = args.hessian
    from sfwi_amd import workloads as W
    from sfwi_amd import _lib
    from sfwi_amd.elastic import Context, Shot
    kw = {}
    if args.shots:
        kw["nshots"] = args.shots
    if args.nt:
        kw["nt"] = args.nt
    sv = W.BY_NAME[args.workload](**kw)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    geo = [W.shot_geometry(sv, s) for s in sv.shots]
    nsh = len(sv.shots)
    storage = args.storage or "f32"
    recon = args.recon or "strips"
    snap_k = args.snap_every if recon == "snapshots" else 0
    print("bench: %s, %s, %d shots, grid %dx%d pml %d, nt %d, %s %s" % (
        "Hessian-vector product" if hess else "Born gradient", sv.name, nsh, sv.nx, sv.allnz, sv.pml, sv.nt,
        recon, storage), file=sys.stderr, flush=True)
    ctx = Context(dx=sv.dx, dz=sv.dz, dt=sv.dt, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt, ntr=sv.ntr,
                  allnx=sv.allnx, allnz=sv.allnz, max_batch=nsh, mode=args.mode, storage=storage,
                  march_cols=args.march_cols, march_depth=args.march_depth, march_order=args.march_order,
                  rho_const=0.0 if hess else 2000.0, adjoint_dir=1 if hess else -1, recon=recon,
                  snap_every=snap_k)
    ctx.set_wavelet(w)
    rho_m = sv.rho0 if hess else None
    ctx.set_model(sv.vp, sv.vs, rho_m)
    obs_z, obs_x = ctx.batch_forward([Shot(s.left, *g, s.ngx_left, s.gc) for s, g in zip(sv.shots, geo)])
    dvp = (sv.vp - sv.vp0).astype(np.float32)
    dvs = (sv.vs - sv.vs0).astype(np.float32)
    ctx.set_model(sv.vp0, sv.vs0, rho_m)
    ctx.set_perturbation(dvp, dvs, region="s1" if hess else "s2")
    shots = [Shot(s.left, *g, s.ngx_left, s.gc, obs_z[k], obs_x[k]) for k, (s, g) in enumerate(zip(sv.shots, geo))]
    resident = [Shot(s.left, *g, s.ngx_left, s.gc) for s, g in zip(sv.shots, geo)]

    def solve(sh):
        if hess:
            ctx.hessian_vector(sh)
            return np.zeros(len(sh))
        return ctx.batch_gradient(sh, born=True)
    for k in range(max(args.warmup, 1)):
        t0 = time.perf_counter()
        solve(shots)
        ctx.flush()
        print("bench: warmup %d %.3f s" % (k, time.perf_counter() - t0), file=sys.stderr, flush=True)
    ctx.set_timing(args.timing_every)
    times = []
    for k in range(args.steps):
        t0 = time.perf_counter()
        m = solve(resident)
        times.append(time.perf_counter() - t0)
        print("bench: step %d %.3f s" % (k, times[-1]), file=sys.stderr, flush=True)
    kstats = {kid: ctx.kernel_time(kid) for kid in range(_lib.K_COUNT)}
    ctx.set_timing(0)
    ctx.flush()
    kernels = {}
    for kid in range(_lib.K_COUNT):
        ms, n = kstats[kid]
        if n:
            kernels[_lib.load().sfwi_kernel_name(kid).decode()] = {"avg_ms": round(ms, 5), "samples": n}
    acc = cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        ncpu = max(1, min(args.cpu_shots, nsh))
        solve(shots[:ncpu])
        g_vp, g_vs, _ = ctx.flush()
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        r_vp = np.zeros((sv.allnx, sv.allnz))
        r_vs = np.zeros_like(r_vp)
        t0 = time.perf_counter()
        for s, g, sh in zip(sv.shots[:ncpu], geo[:ncpu], shots[:ncpu]):
            c = O.config(allnx=sv.allnx, allnz=sv.allnz, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt,
                         ntr=sv.ntr, dx=sv.dx, dz=sv.dz, dt=sv.dt, left=s.left, isx=g[0], isz=g[1],
                         igz=g[2], ngx_left=s.ngx_left, nthreads=threads, rho_const=0.0 if hess else 2000.0,
                         adjoint_dir=1 if hess else -1, snap_every=snap_k)
            if hess and snap_k:
                a = O.hessian_vector_snapshot(c, sv.vp0, sv.vs0, sv.rho0, dvp, dvs, w, s.gc)
            elif hess:
                a = O.hessian_vector(c, sv.vp0, sv.vs0, sv.rho0, dvp, dvs, w, s.gc, "f64")
            elif snap_k:
                a = O.gradient_snapshot(c, sv.vp0, sv.vs0, sv.rho0, w, s.gc, sh.obs_z, sh.obs_x, dvp=dvp, dvs=dvs)
            else:
                a = O.born_gradient(c, sv.vp0, sv.vs0, sv.rho0, dvp, dvs, w, s.gc, sh.obs_z, sh.obs_x, "f64")
            r_vp += a[0]
            r_vs += a[1]
        tc = time.perf_counter() - t0

        def rel(x, y):
            return float(np.linalg.norm(x.astype(np.float64) - y) / np.linalg.norm(y))

        acc = {"grad_vp": rel(g_vp, r_vp), "grad_vs": rel(g_vs, r_vs), "shots": ncpu, "bar": 1e-3}
        unit = "shot-products/s" if hess else "shot-gradients/s"
        cpu = {"value": round(ncpu / tc, 5), "unit": unit, "cores": threads, "kind": "port",
               "sample": "FP64 oracle %s, %d of the %d shots of the same workload, %.1f s" % (
                   "Hessian-vector product" if hess else "Born gradient", ncpu, nsh, tc)}
    total = sum(times)
    lat = sorted(times)
    print(json.dumps({
        "metric": ("stage-1 Gauss-Newton Hessian-vector shot-products/s" if hess
                   else "stage-2 (Born) shot-gradients/s"), "value": round(nsh * args.steps / total, 4),
        "unit": "shot-products/s" if hess else "shot-gradients/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded model; perturbation = true - initial model; observed traces from a forward "
                "solve of the true model)",
        "config": {"workload": "%s: %dx%d dx=dz=%gm, pml %d, dt %gms, nt %d, %d shots, %g Hz, %d receivers, "
                               "%s" % (sv.name, sv.nx, sv.allnz, sv.dx, sv.pml, sv.dt * 1e3, sv.nt, nsh, sv.f0,
                                       sv.ntr, "batched Hessian-vector product p = true - initial model" if hess
                                       else "batched Born gradient, density 2000"),
                   "mode": args.mode,
                   "storage": ("velocity snapshots every %d steps, %s with a max-abs scale per snapshot"
                               % (snap_k, storage.upper()) if snap_k else "boundary strips %s" % storage.upper())},
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 2),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 2), "unit": "per batch step"},
        "misfit": float(m.sum()), "grad_err_vs_fp64": acc, "cpu_baseline": cpu, "kernels": kernels}), flush=True)
    ctx.close()


def _request(kind, sv, max_batch=0, recon=None, storage=None, snap_every=4):
    """Arguments of one serving request (zero observed traces for gradients);
    max_batch: room for same-shape requests to join its batch; recon /
    storage: the gradient kinds' forward-field storage (default the
    reference's FP32 strips)."""
    from sfwi_amd import workloads as W
    from sfwi_amd.elastic import Shot
    born = kind.startswith("born")
    shots = []
    for s in sv.shots:
        isx, isz, igz = W.shot_geometry(sv, s)
        z = np.zeros((len(s.gc), sv.nt), np.float32) if "gradient" in kind else None
        shots.append(Shot(s.left, isx, isz, igz, s.ngx_left, s.gc, z, z))
    params = dict(dx=sv.dx, dz=sv.dz, dt=sv.dt, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt, ntr=sv.ntr,
                  allnx=sv.allnx, allnz=sv.allnz, max_batch=max(len(shots), max_batch),
                  rho_const=2000.0 if born else 0.0, adjoint_dir=-1 if born else 1)
    if "gradient" in kind and (recon or storage):
        params.update(recon=recon or "strips", storage=storage or "f32",
                      snap_every=snap_every if recon == "snapshots" else 0)
    return dict(kind=kind, shots=shots, wavelet=W.ricker(sv.nt, sv.dt, sv.f0), vp=sv.vp0, vs=sv.vs0,
                rho=sv.rho0, dvp=(sv.vp - sv.vp0) if born else None, dvs=(sv.vs - sv.vs0) if born else None,
                params=params)


def _submit(srv, r):
    return srv.submit(r["kind"], r["shots"], r["wavelet"], r["vp"], r["vs"], r["rho"], r["dvp"], r["dvs"],
                      **r["params"])


def _run_alone(r):
    """One request through the plain one-shot path: per-cell kernels, one shot
    per call, the reference's FP32 boundary strips."""
    from sfwi_amd.elastic import Context
    params = dict(r["params"], max_batch=1, march_cols=-1, recon="strips", storage="f32", snap_every=0)
    with Context(**params) as ctx:
        ctx.set_wavelet(r["wavelet"])
        ctx.set_model(r["vp"], r["vs"], r["rho"])
        if r["dvp"] is not None:
            ctx.set_perturbation(r["dvp"], r["dvs"])
        for s in r["shots"]:
            if r["kind"] == "forward":
                ctx.forward(s)
            elif r["kind"] == "born_forward":
                ctx.born_forward(s)
            elif r["kind"] == "gradient":
                ctx.gradient(s)
            else:
                ctx.born_gradient(s)
        ctx.flush()


def inversion(args):
    """Stage-1 (or stage-2) inversion on the reference's example-1 model
    (example-1/vel_sync_376x201.bin as the true model, _init.bin as the start
    or, for stage 2, the background; vs = 0, rho := 2000 as both stages force
    it; the parameter file of example-1/para_stage1 with its 10 m grid,
    50-cell PML and 10 Hz wavelet).  The example's SU records are not part of
    the reference (Git LFS pointers), so the records of 38 shots evenly spaced
    across the model, each recorded by 376 receivers on the surface, come from
    forward solves of the true model (nt 2000, dt 1 ms) and are written in the
    reference's SU layout; then the sfwi_stage1 (sfwi_stage2) program runs
    NITER iterations on those files."""
    import subprocess
    import tempfile
    from sfwi_amd import _lib
    from sfwi_amd import survey_io as S
    from sfwi_amd.elastic import Context, Shot
    here = os.path.dirname(os.path.abspath(__file__))
    gold = os.path.join(here, "tests", "golden")
    allnx, allnz, dx, pml, nt, dt, f0, nshots = 376, 201, 10.0, 50, 2000, 0.001, 10.0, 38
    vp = np.fromfile(os.path.join(gold, "example1_vel_sync_376x201.bin"), np.float32).reshape(allnx, allnz)
    vp0 = np.fromfile(os.path.join(gold, "example1_vel_sync_376x201_init.bin"), np.float32).reshape(allnx, allnz)
    vs = np.zeros_like(vp)
    rho = np.full_like(vp, 2000.0)
    w = _lib.ricker(nt, np.float32(dt), f0)
    gx = (np.arange(allnx) * dx).astype(np.int32)
    sxs = np.round(np.linspace(5, allnx - 6, nshots)).astype(int) * int(dx)
    tmp = tempfile.mkdtemp(prefix="sfwi_inv_")
    t0 = time.perf_counter()
    with Context(dx=dx, dz=dx, dt=dt, nx=allnx, nz=allnz, pml=pml, nt=nt, ntr=allnx, allnx=allnx, allnz=allnz,
                 max_batch=nshots, rho_const=2000.0) as ctx:
        ctx.set_wavelet(w)
        ctx.set_model(vp, vs, rho)
        shots = [Shot(0, pml + int(sx / dx), pml, pml, 0, gx) for sx in sxs]
        oz, ox = ctx.batch_forward(shots)
    for k, sx in enumerate(sxs):
        S.write_su(os.path.join(tmp, "z.su"), k + 1, int(sx), gx, oz[k], dt, append=k > 0)
        S.write_su(os.path.join(tmp, "x.su"), k + 1, int(sx), gx, ox[k], dt, append=k > 0)
    vp0.tofile(os.path.join(tmp, "vp.bin"))
    vs.tofile(os.path.join(tmp, "vs.bin"))
    para = os.path.join(tmp, "para")
    with open(para, "w") as f:
        f.write("10 10 %d %d %d 1 %d\n%s\n%s\n%s\n%s\n%s\n%s\n%g %d 0\n0 0\n1 0 0 1000\n" % (
            pml, allnx, allnz, nshots, os.path.join(tmp, "vp.bin"), os.path.join(tmp, "vs.bin"),
            os.path.join(tmp, "vs.bin"), os.path.join(tmp, "z.su"), os.path.join(tmp, "x.su"),
            os.path.join(tmp, "out"), f0, args.inversion))
    t_data = time.perf_counter() - t0
    print("bench: example-1 records (%d shots, nt %d) written in %.1f s" % (nshots, nt, t_data),
          file=sys.stderr, flush=True)
    prog = os.path.join(here, "sfwi_amd", "sfwi_stage%d" % args.stage)
    t0 = time.perf_counter()
    p = subprocess.Popen([prog, para, str(args.inversion), str(nshots)], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True)
    its = []
    for line in p.stdout:
        f = line.split()
        its.append({"iter": int(f[1]), "misfit": float(f[3]), "beta": float(f[5]), "alpha": float(f[7]),
                    "t_s": round(time.perf_counter() - t0, 3)})
        if args.stage == 2:
            its[-1].update(fenzi=float(f[9]), fenmu=float(f[11]))
        print("bench: " + line.strip(), file=sys.stderr, flush=True)
    err = p.stderr.read()
    if p.wait() != 0:
        print(err, file=sys.stderr)
        sys.exit(3)
    total = time.perf_counter() - t0
    per = total / max(len(its), 1)
    print(json.dumps({
        "metric": "stage-%d inversion iterations per second (example-1 model)" % args.stage,
        "value": round(1.0 / per, 4),
        "unit": "iterations/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
        "data": "the reference's example-1 true and initial Vp (vs 0, rho 2000); synthetic SU records of "
                "38 shots from forward solves of the true model",
        "config": {"workload": "sfwi_stage%d on example-1: 376x201 dx=dz=10m, pml 50, nt 2000, dt 1ms, "
                               "38 shots x 376 receivers, 10 Hz, %d iterations" % (args.stage, len(its))},
        "s_per_iteration": round(per, 3), "iterations": its,
        "misfit_ratio": round(its[-1]["misfit"] / its[0]["misfit"], 6) if its else None}), flush=True)


def uneven(args):
    """Uneven batch: 120 shots over four models of different grid size and
    record length, one batched gradient request per model, all submitted to
    the server at once (they run concurrently, each context autotuned for its
    grid and batch: segment, prefetch depth, block order and time-blocking
    depth), FP32 compute with FP16 snapshots every 4 steps (the consistent
    gradient; --storage bf16 gives the BF16 variant, --recon strips the
    reference's reconstruction).  Observed traces come from
    forward requests through the true models (untimed).  Three passes: the
    first creates and tunes the four contexts (cold makespan), the second is
    the timed one (value), the third runs each request alone, one after the
    other, on the same tuned contexts: their sum is the makespan of a serial
    schedule, and the GPU is shared well when the concurrent makespan is not
    above it.  The gradient of one shot of the smallest model is checked
    against the FP64 oracle of the same scheme."""
    from sfwi_amd import workloads as W
    from sfwi_amd.elastic import Server, Shot
    svs = W.uneven_batch(shots_per_model=max(1, args.shots // 4) if args.shots else 30)
    recon = args.recon or "snapshots"
    storage = args.storage or "f16"   # BF16 (the configuration's first choice) misses the 1e-3 gradient bar
    reqs = []
    with Server(workers=4) as srv:
        fw = []
        for sv in svs:
            r = _request("forward", sv)
            r["vp"], r["vs"], r["rho"] = sv.vp, sv.vs, sv.rho
            fw.append(_submit(srv, r))
        obs = [srv.wait(t) for t in fw]
        for sv, o in zip(svs, obs):
            r = _request("gradient", sv)
            r["shots"] = [Shot(sh.c.left, sh.c.isx, sh.c.isz, sh.c.igz, sh.c.ngx_left, sh.gc,
                               o["syn_z"][k][:sh.ntr], o["syn_x"][k][:sh.ntr]) for k, sh in enumerate(r["shots"])]
            r["params"] = dict(r["params"], storage=storage, recon=recon, snap_every=args.snap_every)
            reqs.append(r)

        def together():
            t0 = time.perf_counter()
            tks = [srv.submit(r["kind"], r["shots"], r["wavelet"], r["vp"], r["vs"], r["rho"], autotune=True,
                              **r["params"]) for r in reqs]
            outs = [srv.wait(t) for t in tks]
            return time.perf_counter() - t0, outs
        cold, _ = together()
        print("uneven: cold makespan (contexts created and tuned) %.3f s" % cold, file=sys.stderr, flush=True)
        makespan, outs = together()
        print("uneven: makespan %.3f s" % makespan, file=sys.stderr, flush=True)
        alone = []
        for r in reqs:
            t0 = time.perf_counter()
            srv.wait(srv.submit(r["kind"], r["shots"], r["wavelet"], r["vp"], r["vs"], r["rho"], autotune=True,
                                **r["params"]))
            alone.append(time.perf_counter() - t0)
        print("uneven: alone %s s" % ", ".join("%.3f" % a for a in alone), file=sys.stderr, flush=True)
    acc = None
    if not args.no_cpu:   # one shot of the smallest model vs the FP64 oracle of the same scheme
        from oracle import oracle as O
        from sfwi_amd.elastic import Context
        sv, r = svs[0], reqs[0]
        sh, s0 = sv.shots[0], r["shots"][0]
        with Context(**dict(r["params"], max_batch=1)) as ctx:
            ctx.set_wavelet(r["wavelet"])
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            ctx.gradient(s0)
            g_vp, g_vs, _ = ctx.flush()
        isx, isz, igz = W.shot_geometry(sv, sh)
        c = O.config(allnx=sv.allnx, allnz=sv.allnz, nx=sv.nx, nz=sv.allnz, pml=sv.pml, nt=sv.nt, ntr=len(sh.gc),
                     dx=sv.dx, dz=sv.dz, dt=sv.dt, left=sh.left, isx=isx, isz=isz, igz=igz, ngx_left=sh.ngx_left,
                     nthreads=int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1),
                     snap_every=args.snap_every if recon == "snapshots" else 0)
        if recon == "snapshots":
            ref = O.gradient_snapshot(c, sv.vp0, sv.vs0, sv.rho0, r["wavelet"], sh.gc, s0.obs_z, s0.obs_x)
        else:
            ref = O.gradient(c, sv.vp0, sv.vs0, sv.rho0, r["wavelet"], sh.gc, s0.obs_z, s0.obs_x, "f64")

        def rel(x, y):
            return float(np.linalg.norm(x.astype(np.float64) - y) / np.linalg.norm(y))
        acc = {"model": "%dx%d" % (sv.nx, sv.allnz), "grad_vp": rel(g_vp, ref[0]), "grad_vs": rel(g_vs, ref[1]),
               "bar": 1e-3}
    nsh = sum(len(r["shots"]) for r in reqs)
    cells = sum(sv.cells() * sv.nt * len(sv.shots) * 3 for sv in svs)
    line = {
        "metric": "shot-gradients/s & Gcell-updates/s (uneven batch)", "value": round(nsh / makespan, 4),
        "unit": "shot-gradients/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
        "data": "synthetic (seeded models, observed traces from forward solves of the true models)",
        "config": {"workload": "uneven batch: %d shots over %s" % (nsh, ", ".join(
            "%dx%d nt %d %g Hz" % (sv.nx, sv.allnz, sv.nt, sv.f0) for sv in svs)),
            "storage": ("%s snapshots every %d steps" % (storage.upper(), args.snap_every) if recon == "snapshots"
                        else "boundary strips %s" % storage.upper()), "autotuned": True},
        "gcell_updates_per_s": round(cells / makespan / 1e9, 2),
        "makespan_s": round(makespan, 3), "cold_makespan_s": round(cold, 3),
        "per_model_run_s": [round(o["run_ms"] / 1e3, 3) for o in outs],
        "alone_s": [round(a, 3) for a in alone], "serial_makespan_s": round(sum(alone), 3),
        "misfits": [float(o["misfits"].sum()) for o in outs],
        "grad_err_vs_fp64": acc,
    }
    print(json.dumps(line), flush=True)


def serving(args):
    """Serving loop: a seeded Poisson stream of forward, Born and gradient
    requests of random shape (nx 200-1500, nz 150-800, nt 1500-6000, 1-8
    shots, 10-20 Hz) submitted in real time to the library's server, which
    admits them against the HBM budget and overlaps them on the GPU.  Reports
    p50/p99 of request latency (submit to done) and throughput, beside the same
    requests run one at a time through the plain one-shot path.  Requests of
    one shape and kind queued together share a batch (up to --merge-cap
    shots); the saturated pass is also run with that cross-request batching
    off.  --shapes K draws every request's grid and record length from K
    shapes (the default stream gives every request its own)."""
    import threading
    from sfwi_amd import workloads as W
    from sfwi_amd.elastic import Server
    stream = W.serving_requests(args.requests, rate=args.rate, scale=args.scale, shapes=args.shapes)
    reqs = [(t, _request(k, sv, args.merge_cap, args.recon, args.storage, args.snap_every), sv)
            for t, k, sv in stream]
    cells = sum(sv.cells() * sv.nt * len(sv.shots) * {"forward": 1, "born_forward": 2, "gradient": 3,
                                                        "born_gradient": 4}[r["kind"]] for _, r, sv in reqs)
    shots = sum(len(r["shots"]) for _, r, _ in reqs)
    print("serving: %d requests, %d shots, %.1f Gcell-updates, arrivals over %.1f s" % (
        len(reqs), shots, cells / 1e9, reqs[-1][0]), file=sys.stderr, flush=True)
    lat = [None] * len(reqs)
    done_at = [None] * len(reqs)
    with Server(workers=args.workers) as srv:
        # warm-up: one small request so code objects are loaded before the clock starts
        wt = _submit(srv, _request("forward", W.tiny(nt=50)))
        srv.wait(wt)
        t0 = time.perf_counter()
        threads = []

        def waiter(i, tk, t_sub):
            srv.wait(tk)
            done_at[i] = time.perf_counter()
            lat[i] = done_at[i] - t_sub

        for i, (t_arr, r, _) in enumerate(reqs):
            now = time.perf_counter() - t0
            if t_arr > now:
                time.sleep(t_arr - now)
            t_sub = time.perf_counter()
            tk = _submit(srv, r)
            th = threading.Thread(target=waiter, args=(i, tk, t_sub))
            th.start()
            threads.append(th)
        for th in threads:
            th.join()
        makespan = max(done_at) - t0
        print("serving: paced makespan %.2f s" % makespan, file=sys.stderr, flush=True)
        # saturated: the same requests all queued at once (throughput), with and
        # without cross-request batching (the cached contexts are reused)
        sat = {}
        for on in (False, True):
            srv.set_batching(on)
            t0 = time.perf_counter()
            tks = [_submit(srv, r) for _, r, _ in reqs]
            outs = [srv.wait(tk) for tk in tks]
            sat[on] = (time.perf_counter() - t0, len(set(o["run_ms"] for o in outs)))
        t_sat = sat[True][0]
    print("serving: saturated makespan %.2f s (batching off: %.2f s)" % (t_sat, sat[False][0]), file=sys.stderr,
          flush=True)
    alone = []
    for _, r, _ in reqs:
        t1 = time.perf_counter()
        _run_alone(r)
        alone.append(time.perf_counter() - t1)
    t_alone = sum(alone)
    print("serving: one at a time %.2f s" % t_alone, file=sys.stderr, flush=True)
    lms = sorted(1e3 * x for x in lat)
    ams = sorted(1e3 * x for x in alone)
    line = {
        "metric": "serving p50/p99 request latency and throughput", "value": round(shots / t_sat, 4),
        "unit": "shot-solves/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
        "data": "synthetic (seeded layered models, zero observed traces)",
        "config": {"workload": "serving: %d Poisson requests at %.2f/s, nx 200-1500, nz 150-800, nt 1500-6000 "
                               "(x%.2f), 1-8 shots, forward/Born/gradient mix" % (len(reqs), args.rate, args.scale),
                   "workers": args.workers,
                   "gradient_storage": ("velocity snapshots every %d steps, %s" % (args.snap_every,
                                                                                  (args.storage or "f32").upper())
                                        if args.recon == "snapshots" else
                                        "boundary strips %s" % (args.storage or "f32").upper())},
        "value_note": "saturated: all requests queued at once",
        "latency_ms": {"p50": round(float(np.percentile(lms, 50)), 1),
                       "p99": round(float(np.percentile(lms, 99)), 1),
                       "unit": "submit to done, paced Poisson arrivals"},
        "gcell_updates_per_s": round(cells / t_sat / 1e9, 2),
        "paced_makespan_s": round(makespan, 3),
        "saturated_makespan_s": round(t_sat, 3),
        "cross_request_batching": {"merge_cap_shots": args.merge_cap, "shapes": args.shapes or "per request",
                                   "batches": sat[True][1],
                                   "off": {"value": round(shots / sat[False][0], 4), "unit": "shot-solves/s",
                                           "makespan_s": round(sat[False][0], 3), "batches": sat[False][1]}},
        "one_at_a_time": {"value": round(shots / t_alone, 4), "unit": "shot-solves/s", "seconds": round(t_alone, 3),
                          "request_ms": {"p50": round(float(np.percentile(ams, 50)), 1),
                                         "p99": round(float(np.percentile(ams, 99)), 1)},
                          "path": "plain one-shot path (per-cell kernels, one shot per call), requests back to back"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
