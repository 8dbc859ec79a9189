"""Host steps of the reference's inversion loops -- CPU oracle (numpy).

TEST INFRASTRUCTURE ONLY: the oracle of the library's sfwi_s1_* / sfwi_s2_*
host functions and of the inversion iterations built on them.  Only tests/,
the smoke check and bench.py's cpu_baseline leg may import it.

Each function restates the reference's arithmetic type by type: FP32 arrays
(numpy float32 element-wise operations round like the C++ float operations),
FP64 accumulators formed by a sequential sum in memory order (np.cumsum, not
the pairwise np.sum), products of two floats rounded to FP32 before they
enter an FP64 sum, FP64 step lengths applied to FP32 arrays in FP64 and
rounded back.  Model arrays are [allnx, allnz] (index x*allnz + z).

Pinned by tests/test_inversion_host.py: the parabola vertex of three exact
samples of a quadratic, the clamp, joint max-abs normalisation, the scaling
of beta with the gradient (beta = lambda^2 for g1 = lambda g0) and the
invariance of the search direction under it, the cut layer and top-row
reset.
"""
import numpy as np

F32 = np.float32


def _seqsum(t):
    """FP64 sum of the FP32 terms t in memory order (the reference's loop)."""
    t = np.ascontiguousarray(t, F32).ravel().astype(np.float64)
    return float(np.cumsum(t)[-1]) if t.size else 0.0


def _absmax_ignoring_nan(a):
    """max over |a| with the reference's (fabs(x) > max) ? fabs(x) : max form:
    NaN never wins a comparison, the start value is 0."""
    m = np.abs(np.asarray(a, F32))
    m = np.where(np.isnan(m), F32(0), m)
    return F32(m.max()) if m.size else F32(0)


def normalize_all(a, b):
    """sfwi-stage1/main.cpp:35-49."""
    mx = max(_absmax_ignoring_nan(a), _absmax_ignoring_nan(b))
    mx = F32(mx)
    with np.errstate(divide="ignore", invalid="ignore"):
        return (np.asarray(a, F32) / mx).astype(F32), (np.asarray(b, F32) / mx).astype(F32)


def s1_direction(iter_, cut_layer, pp_grad, ps_grad, illum, pp_cg_old, ps_cg_old, beta_fenmu):
    """Stage-1 direction (sfwi-stage1/main.cpp:695-750, 809-839, 1059-1062).
    Returns (dir_pp, dir_ps, cg_pp, cg_ps, fenzi, beta); cg_* and fenzi are the
    state of the next iteration (pp_cg_old, ps_cg_old, beta_fenmu)."""
    g1, g2, il = (np.asarray(x, F32) for x in (pp_grad, ps_grad, illum))
    with np.errstate(divide="ignore", invalid="ignore"):
        s1 = (g1 / il).astype(F32)
        s2 = (g2 / il).astype(F32)
    s1[:, :cut_layer] = 0
    s2[:, :cut_layer] = 0
    t = (g1 * s1).astype(F32) + (g2 * s2).astype(F32)
    fenzi = _seqsum(t.astype(F32))
    beta = 0.0 if iter_ == 0 else fenzi / beta_fenmu
    cg1 = (s1.astype(np.float64) + beta * np.asarray(pp_cg_old, F32).astype(np.float64)).astype(F32)
    cg2 = (s2.astype(np.float64) + beta * np.asarray(ps_cg_old, F32).astype(np.float64)).astype(F32)
    d1, d2 = normalize_all(cg1, cg2)
    d1, d2 = normalize_all(d1, d2)
    return d1, d2, cg1, cg2, fenzi, beta


def s1_alpha(m0, m1, m2, step1, step2):
    """cal_step2 (sfwi-stage1/main.cpp:65-72) and the clamp (1229-1238)."""
    s1, s2 = F32(step1), F32(step2)
    fz = float(s1 * s1 - s2 * s2) * m0 + float(s2 * s2) * m1 + float(-s1 * s1) * m2
    fm = float(s1 - s2) * m0 + float(s2) * m1 + float(-s1) * m2
    alpha = fz / fm / 2.0
    return min(max(alpha, -100.0), 300.0)


def s1_update(cut_layer, alpha, dir_pp, dir_ps, velp, vels):
    """sfwi-stage1/main.cpp:1241-1257; returns the new (velp, vels)."""
    vp = (np.asarray(velp, F32).astype(np.float64) - alpha * np.asarray(dir_pp, F32).astype(np.float64)).astype(F32)
    vs = (np.asarray(vels, F32).astype(np.float64) - alpha * np.asarray(dir_ps, F32).astype(np.float64)).astype(F32)
    vp[:, :cut_layer] = vp[0, 0]
    vs[:, :cut_layer] = vs[0, 0]
    return vp, vs


def s2_direction(iter_, cut_layer, pp_grad, ps_grad, illum, pp_cg_old, ps_cg_old, beta_fenmu):
    """Stage-2 direction (sfwi-stage2/main.cpp:645-703).  Returns
    (pp_grad', ps_grad', cg_pp, cg_ps, fenzi, beta) with the gradients as the
    reference leaves them (rows above the cut zeroed)."""
    g1 = np.array(pp_grad, F32)
    g2 = np.array(ps_grad, F32)
    il = np.asarray(illum, F32)
    g1[:, :cut_layer] = 0
    g2[:, :cut_layer] = 0
    with np.errstate(divide="ignore", invalid="ignore"):
        c1 = (g1 / il).astype(F32)
        c2 = (g2 / il).astype(F32)
        t = ((g1 * g1).astype(F32) / il).astype(F32) + ((g2 * g2).astype(F32) / il).astype(F32)
    c1[:, :cut_layer] = g1[:, :cut_layer]
    c2[:, :cut_layer] = g2[:, :cut_layer]
    fenzi = _seqsum(t.astype(F32))
    beta = 0.0 if iter_ == 0 else fenzi / beta_fenmu
    cg1 = (c1.astype(np.float64) + beta * np.asarray(pp_cg_old, F32).astype(np.float64)).astype(F32)
    cg2 = (c2.astype(np.float64) + beta * np.asarray(ps_cg_old, F32).astype(np.float64)).astype(F32)
    return g1, g2, cg1, cg2, fenzi, beta


def s2_update(fenzi, fenmu, pp_cg, ps_cg, image_pp, image_ps):
    """sfwi-stage2/main.cpp:861-873; returns (image_pp, image_ps, alpha)."""
    alpha = fenzi / fenmu
    ip = (np.asarray(image_pp, F32).astype(np.float64) - alpha * np.asarray(pp_cg, F32).astype(np.float64)).astype(F32)
    is_ = (np.asarray(image_ps, F32).astype(np.float64) - alpha * np.asarray(ps_cg, F32).astype(np.float64)).astype(F32)
    return ip, is_, alpha


# ------------------------------------------- stage-1 Gauss-Newton inner CG ---
# sfwi-stage1/main.cpp:784-1063 (run when max_inner > 0).  dt = float32 is the
# reference's arithmetic; dt = float64 runs the same steps in FP64 (the pins).

def _seqsum_dt(t, dt):
    t = np.ascontiguousarray(t, dt).ravel().astype(np.float64)
    return float(np.cumsum(t)[-1]) if t.size else 0.0


def _normalize_dt(a, b, dt):
    if dt == F32:
        return normalize_all(a, b)
    mx = max(float(np.abs(a).max()), float(np.abs(b).max()))
    return np.asarray(a, dt) / mx, np.asarray(b, dt) / mx


def s1_hcg_init(pp_cg, ps_cg, dt=F32):
    """Set-up of the inner loop (main.cpp:786-839): returns the state dict with
    r = normalize_all(cg), p = -r, grad = normalize_all(cg),
    delta = normalize_all(-cg)."""
    c1, c2 = np.asarray(pp_cg, dt), np.asarray(ps_cg, dt)
    r = _normalize_dt(c1, c2, dt)
    p = (-r[0], -r[1])
    delta = _normalize_dt(-c1, -c2, dt)
    grad = _normalize_dt(c1, c2, dt)
    return dict(r=r, p=p, delta=delta, grad=grad)


def s1_hcg_step(st, hp_pp, hp_ps, cut_grad, cut_layer, dt=F32, plant=0):
    """One inner iteration's host step (main.cpp:957-1024) on the state dict
    (updated in place); returns (rr_old, rr_new, alpha, beta).  plant (test
    hook): 1 takes beta with the wrong sign."""
    h1, h2 = np.array(hp_pp, dt), np.array(hp_ps, dt)
    if cut_grad:
        h1[:, :cut_layer] = 0
        h2[:, :cut_layer] = 0
    r1, r2 = st["r"]
    p1, p2 = st["p"]
    d1, d2 = st["delta"]
    a_num = _seqsum_dt((r1 * r1).astype(dt) + (r2 * r2).astype(dt), dt)
    a_den = _seqsum_dt((p1 * h1).astype(dt) + (p2 * h2).astype(dt), dt)
    alpha = a_num / a_den
    f = np.float64
    d1 = (d1.astype(f) + alpha * p1.astype(f)).astype(dt)
    d2 = (d2.astype(f) + alpha * p2.astype(f)).astype(dt)
    r1 = (r1.astype(f) + alpha * h1.astype(f)).astype(dt)
    r2 = (r2.astype(f) + alpha * h2.astype(f)).astype(dt)
    b_num = _seqsum_dt((r1 * r1).astype(dt) + (r2 * r2).astype(dt), dt)
    beta = b_num / a_num
    if plant & 1:
        beta = -beta
    p1 = (-r1.astype(f) + beta * p1.astype(f)).astype(dt)
    p2 = (-r2.astype(f) + beta * p2.astype(f)).astype(dt)
    st.update(r=(r1, r2), p=(p1, p2), delta=(d1, d2), grad=(-d1, -d2))
    return a_num, b_num, alpha, beta


def s1_hcg_finish(st, dt=F32):
    """main.cpp:1057-1063: delta and grad normalized; grad is the search direction."""
    st["delta"] = _normalize_dt(*st["delta"], dt)
    st["grad"] = _normalize_dt(*st["grad"], dt)
    return st
