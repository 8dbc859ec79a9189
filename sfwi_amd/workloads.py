"""Seeded synthetic survey definitions (models, geometry, wavelets).

No field data ships with the reference (its SU records are Git-LFS pointers),
so every workload here is synthetic and seeded.  Models are float32 arrays in
the reference's .bin layout: shape (allnx, allnz), x-major, z fastest.

Geometry follows the reference driver's conventions
(sfwi-stage1/main.cpp:596-625): a shot window starts at the survey's leftmost
coordinate, ``isx = pml + (sx - ngx_left)/dx``, ``isz = pml + sy``,
``igz = pml + gy`` and receiver i sits at coordinate ``gc[i]``.
"""
from dataclasses import dataclass, field
from typing import List

import numpy as np


@dataclass
class Shot:
    sx: int                 # source coordinate (m, integer as in the SU header)
    gc: np.ndarray          # receiver coordinates (m), int32
    ngx_left: int           # leftmost coordinate of the shot's aperture
    left: int               # first model column of the window (delta_left + record_left_in_v)


@dataclass
class Survey:
    name: str
    vp: np.ndarray          # true model (allnx, allnz)
    vs: np.ndarray
    rho: np.ndarray
    vp0: np.ndarray         # starting model
    vs0: np.ndarray
    rho0: np.ndarray
    dx: float
    dz: float
    dt: float
    nt: int
    pml: int
    f0: float
    sy: int
    gy: int
    nx: int                 # window width (all shots share it here)
    shots: List[Shot] = field(default_factory=list)

    @property
    def allnx(self):
        return self.vp.shape[0]

    @property
    def allnz(self):
        return self.vp.shape[1]

    @property
    def ntr(self):
        return len(self.shots[0].gc)

    def cells(self):
        return (self.nx + 2 * self.pml) * (self.allnz + 2 * self.pml)


def gardner(vp):
    """Gardner's relation rho = 310 vp^0.25 (kg/m^3, vp in m/s)."""
    return (310.0 * np.power(vp, 0.25)).astype(np.float32)


def gaussian_smooth(a, sigma):
    """Separable Gaussian smoothing with edge replication (numpy only)."""
    if sigma <= 0:
        return a.copy()
    r = int(3 * sigma)
    k = np.exp(-0.5 * (np.arange(-r, r + 1) / sigma) ** 2)
    k /= k.sum()
    out = a.astype(np.float64)
    for ax in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[ax] = (r, r)
        p = np.pad(out, pad, mode="edge")
        out = np.apply_along_axis(lambda v: np.convolve(v, k, mode="valid"), ax, p)
    return out.astype(np.float32)


def ricker(nt, dt, f0):
    """Ricker wavelet exactly as the host driver builds it (sfwi-stage1/main.cpp:251-260)."""
    pi = 3.14159265358979
    f, d = float(np.float32(f0)), float(np.float32(dt))
    nk = float(np.float32(pi * pi * f * f * d * d))          # double product stored as float
    t0 = int(np.ceil(1.0 / float(np.float32(f0) * np.float32(dt))))   # float product
    t = np.arange(nt, dtype=np.float64) - t0
    return ((1.0 - 2.0 * nk * t * t) * np.exp(-nk * t * t)).astype(np.float32)


def _full_aperture_shots(nx, dx, nshots):
    gc = (np.arange(nx) * dx).astype(np.int32)
    out = []
    for k in range(nshots):
        ix = int((k + 0.5) * nx / nshots)
        out.append(Shot(sx=int(ix * dx), gc=gc.copy(), ngx_left=0, left=0))
    return out


def _layered(nx, nz, tops, vps, rng=None):
    vp = np.zeros((nx, nz), np.float32)
    z = np.arange(nz)
    for top, v in zip(tops, vps):
        vp[:, z >= top] = v
    return vp


def plain_small(seed=1001, nshots=16, nt=2000):
    """376x201, dx=dz=10 m, 4 flat layers 1500-4000 m/s, Vs=Vp/1.73 (0 in a
    10-cell water top), Gardner rho, +10% Gaussian Vp anomaly (sigma 15
    cells); starting model = truth smoothed by a 20-cell Gaussian; dt 1 ms,
    20-cell PML, 10 Hz, shots at 2-cell depth, receivers at every surface cell."""
    rng = np.random.default_rng(seed)
    nx, nz = 376, 201
    vp = _layered(nx, nz, [0, 10, 80, 140], [1500.0, 2300.0, 3100.0, 4000.0])
    cx, cz = rng.uniform(120, 256), rng.uniform(90, 150)
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    g = np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / (2 * 15.0 ** 2))
    vp = (vp * (1.0 + 0.10 * g * (Z >= 10))).astype(np.float32)
    vs = (vp / 1.73).astype(np.float32)
    vs[:, :10] = 0.0
    rho = gardner(vp)
    vp0 = gaussian_smooth(vp, 20.0)
    vs0 = gaussian_smooth(vs, 20.0)
    vs0[:, :10] = 0.0
    rho0 = gardner(vp0)
    sv = Survey("plain_small_376x201", vp, vs, rho, vp0, vs0, rho0, 10.0, 10.0, 0.001, nt, 20,
                10.0, 2, 0, nx)
    sv.shots = _full_aperture_shots(nx, 10.0, nshots)
    return sv


def shot_batch(seed=1002, nshots=64, nt=3000):
    """401x301, dx=dz=10 m, 6 seeded layers with a 15-degree dipping
    interface, Vs=Vp/1.73, Gardner rho; dt 1 ms, 30-cell PML, 12 Hz, 401
    two-component receivers per shot."""
    rng = np.random.default_rng(seed)
    nx, nz = 401, 301
    tops = np.sort(rng.uniform(20, 280, size=5))
    vps = np.sort(rng.uniform(1600, 4000, size=6))
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    vp = np.full((nx, nz), vps[0], np.float32)
    dip = np.tan(np.deg2rad(15.0))
    for k, top in enumerate(tops):
        zt = top + (dip * (X - nx / 2) if k == 2 else 0.0)
        vp = np.where(Z >= zt, vps[k + 1], vp).astype(np.float32)
    vs = (vp / 1.73).astype(np.float32)
    rho = gardner(vp)
    vp0 = gaussian_smooth(vp, 10.0)
    vs0 = gaussian_smooth(vs, 10.0)
    rho0 = gardner(vp0)
    sv = Survey("shot_batch_401x301", vp, vs, rho, vp0, vs0, rho0, 10.0, 10.0, 0.001, nt, 30,
                12.0, 2, 0, nx)
    sv.shots = _full_aperture_shots(nx, 10.0, nshots)
    return sv


def mixed_precision(seed=1003, nshots=96, nt=6000):
    """1200x600, dx=dz=5 m, 8 seeded random layers plus 3 Gaussian inclusions
    of +-8 %; dt 0.5 ms, 40-cell PML, 15 Hz."""
    rng = np.random.default_rng(seed)
    nx, nz = 1200, 600
    tops = np.sort(rng.uniform(30, 560, size=7))
    vps = np.sort(rng.uniform(1600, 4000, size=8))
    vp = _layered(nx, nz, [0] + list(tops), vps)
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    for s in (+1, -1, +1):
        cx, cz, sg = rng.uniform(200, 1000), rng.uniform(150, 500), rng.uniform(20, 40)
        vp = vp * (1.0 + s * 0.08 * np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / (2 * sg ** 2)))
    vp = vp.astype(np.float32)
    vs = (vp / 1.73).astype(np.float32)
    rho = gardner(vp)
    vp0 = gaussian_smooth(vp, 20.0)
    vs0 = gaussian_smooth(vs, 20.0)
    rho0 = gardner(vp0)
    sv = Survey("mixed_precision_1200x600", vp, vs, rho, vp0, vs0, rho0, 5.0, 5.0, 0.0005, nt, 40,
                15.0, 2, 0, nx)
    sv.shots = _full_aperture_shots(nx, 5.0, nshots)
    return sv


def tiny(seed=7, nx=48, nz=40, pml=8, nt=200, nshots=2, f0=25.0, elastic=True):
    """Small seeded layered model for fast parity tests (dx=dz=10 m, dt 1 ms)."""
    rng = np.random.default_rng(seed)
    tops = np.sort(rng.uniform(5, nz - 5, size=2))
    vp = _layered(nx, nz, [0] + list(tops), [1800.0, 2500.0, 3200.0])
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    cx, cz = rng.uniform(nx * 0.3, nx * 0.7), rng.uniform(nz * 0.4, nz * 0.7)
    vp = (vp * (1.0 + 0.08 * np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / (2 * 4.0 ** 2)))).astype(np.float32)
    vs = (vp / 1.73).astype(np.float32) if elastic else np.zeros_like(vp)
    rho = gardner(vp)
    vp0 = gaussian_smooth(vp, 3.0)
    vs0 = gaussian_smooth(vs, 3.0)
    rho0 = gardner(vp0)
    sv = Survey("tiny_%dx%d" % (nx, nz), vp, vs, rho, vp0, vs0, rho0, 10.0, 10.0, 0.001, nt, pml,
                f0, 2, 0, nx)
    sv.shots = _full_aperture_shots(nx, 10.0, nshots)
    return sv


def shot_geometry(sv: Survey, shot: Shot):
    """Padded-grid indices of one shot, as the reference driver sets them
    (sfwi-stage1/main.cpp:619-625)."""
    isx = sv.pml + int((shot.sx - shot.ngx_left) / np.float32(sv.dx))
    isz = sv.pml + sv.sy
    igz = sv.pml + sv.gy
    return isx, isz, igz


def mid_801x401(seed=1004, nshots=30, nt=4000):
    """801x401, dx=dz=10 m, seeded 5-layer model with a Gaussian +8% Vp
    inclusion, dt 1 ms, 30-cell PML, 12 Hz (the third grid of the uneven batch)."""
    rng = np.random.default_rng(seed)
    nx, nz = 801, 401
    tops = np.concatenate(([0], np.sort(rng.choice(np.arange(20, nz - 20), 4, replace=False))))
    vp = _layered(nx, nz, tops, np.sort(rng.uniform(1700.0, 4300.0, 5)))
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    g = np.exp(-((X - rng.uniform(200, 600)) ** 2 + (Z - rng.uniform(120, 300)) ** 2) / (2 * 25.0 ** 2))
    vp = (vp * (1.0 + 0.08 * g)).astype(np.float32)
    vs = (vp / 1.73).astype(np.float32)
    rho = gardner(vp)
    vp0, vs0 = gaussian_smooth(vp, 15.0), gaussian_smooth(vs, 15.0)
    sv = Survey("mid_801x401", vp, vs, rho, vp0, vs0, gardner(vp0), 10.0, 10.0, 0.001, nt, 30, 12.0, 2, 0, nx)
    sv.shots = _full_aperture_shots(nx, 10.0, nshots)
    return sv


def uneven_batch(shots_per_model=30):
    """Uneven batch: 120 shots over four models of different size and record
    length -- 376x201 (nt 2000, 10 Hz), 401x301 (nt 3000, 12 Hz), 801x401
    (nt 4000, 12 Hz), 1200x600 (dx 5 m, nt 6000, 15 Hz)."""
    n = shots_per_model
    return [plain_small(nshots=n, nt=2000), shot_batch(nshots=n, nt=3000), mid_801x401(nshots=n, nt=4000),
            mixed_precision(nshots=n, nt=6000)]


def serving_requests(n=24, seed=1005, rate=1.0, scale=1.0, shapes=0):
    """Serving-loop request stream: Poisson arrivals at `rate` per second of
    forward, Born and gradient solves; per request nx uniform in 200-1500, nz
    in 150-800, nt in 1500-6000 (all multiplied by `scale`, for quick runs), a
    seeded layered Vp/Vs/rho model, 1-8 shots and a 10-20 Hz Ricker wavelet;
    dx = dz = 10 m, dt = 1 ms, 20-cell PML.  Gradient requests invert zero
    observed traces (the misfit is the synthetic energy), Born requests
    perturb Vp/Vs by a Gaussian anomaly.  shapes > 0: the (nx, nz, nt) of
    each request is one of `shapes` triples drawn the same way (a service
    whose clients share a few grids; models, shots and wavelets stay per
    request).  Returns [(arrival_s, kind, survey)]."""
    rng = np.random.default_rng(seed)
    kinds = ["forward", "born_forward", "gradient", "born_gradient"]
    srng = np.random.default_rng(seed + 1)
    table = [(max(32, int(srng.integers(200, 1501) * scale)), max(32, int(srng.integers(150, 801) * scale)),
              max(64, int(srng.integers(1500, 6001) * scale))) for _ in range(shapes)]
    t = 0.0
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        nx = max(32, int(rng.integers(200, 1501) * scale))
        nz = max(32, int(rng.integers(150, 801) * scale))
        nt = max(64, int(rng.integers(1500, 6001) * scale))
        if shapes:
            nx, nz, nt = table[int(srng.integers(0, shapes))]
        nsh = int(rng.integers(1, 9))
        f0 = float(rng.uniform(10.0, 20.0))
        nl = int(rng.integers(3, 7))
        tops = np.sort(rng.choice(np.arange(1, nz - 1), nl - 1, replace=False))
        vps = np.sort(rng.uniform(1800.0, 4200.0, nl))
        vp = _layered(nx, nz, np.concatenate(([0], tops)), vps)
        vs = (vp / 1.75).astype(np.float32)
        rho = gardner(vp)
        X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
        g = np.exp(-((X - nx / 2) ** 2 + (Z - nz / 2) ** 2) / (2 * (0.1 * min(nx, nz)) ** 2)).astype(np.float32)
        sv = Survey("request_%d_%dx%d" % (k, nx, nz), (vp * (1 + 0.05 * g)).astype(np.float32), vs, rho,
                    vp, vs, rho, 10.0, 10.0, 0.001, nt, 20, f0, 2, 0, nx)
        sv.shots = _full_aperture_shots(nx, 10.0, nsh)
        out.append((t, kinds[int(rng.integers(0, 4))], sv))
    return out


BY_NAME = {
    "plain_small": plain_small,
    "shot_batch": shot_batch,
    "mixed_precision": mixed_precision,
    "mid_801x401": mid_801x401,
    "tiny": tiny,
}
