"""The reference drivers' inversion loops over the per-shot solves.

``Stage1`` is the stage-1 iteration of sfwi-stage1/main.cpp (the
configuration the reference runs: update_mp, conjugate gradient, and by
default no Gauss-Newton inner loop): batched gradient of every shot,
illumination preconditioning with the cut layer, CG direction normalized by
its joint max-abs, with max_inner > 0 the Gauss-Newton inner CG on batched
Hessian-vector products (main.cpp:784-1063), parabolic line search at two
trial steps, clamped step, model update and top-row reset.  ``Stage2`` is the stage-2 iteration of
sfwi-stage2/main.cpp: Born gradient, preconditioned CG direction, exact
linear step, image update.  The host steps run in the library
(sfwi_s1_* / sfwi_s2_*, CPU code with the reference's arithmetic types); the
solves run on the GPU through a ``Context``.  Shots are summed in list order,
where the reference sums over MPI ranks.
"""
import ctypes

import numpy as np

from . import _lib


def _f32c(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class _CG:
    """Conjugate-gradient state kept between iterations (pp_cg_old,
    ps_cg_old, beta_fenmu; sfwi-stage1/main.cpp:376-377)."""

    def __init__(self, allnx, allnz, cut_layer):
        self.allnx, self.allnz, self.cut_layer = allnx, allnz, cut_layer
        self.pp_cg_old = np.zeros((allnx, allnz), np.float32)
        self.ps_cg_old = np.zeros((allnx, allnz), np.float32)
        self.beta_fenmu = 0.0
        self.iter = 0
        self.max_misfit = None
        self.L = _lib.load()

    def _ck(self, rc):
        if rc != 0:
            raise _lib.SfwiError(rc, self.L.sfwi_last_error(None).decode())

    def _rms(self, misfit):
        if self.max_misfit is None:
            self.max_misfit = misfit
        return misfit / self.max_misfit


class Stage1(_CG):
    """Stage-1 loop state and steps (sfwi-stage1/main.cpp:366-416 locals:
    cut_layer 26, step1 30, step2 150)."""

    def __init__(self, allnx, allnz, cut_layer=26, step1=30.0, step2=150.0, max_inner=0, cut_grad=True):
        super().__init__(allnx, allnz, cut_layer)
        self.step1, self.step2 = float(step1), float(step2)
        self.max_inner, self.cut_grad = int(max_inner), bool(cut_grad)

    # Gauss-Newton inner CG (main.cpp:784-1063); the state is a dict of
    # float32 (pp, ps) pairs r, p, delta, grad.
    def hcg_init(self, pp_cg=None, ps_cg=None):
        """Inner-loop set-up from the CG direction (default: the one the last
        direction() call formed)."""
        c1 = _f32c(self.pp_cg_old if pp_cg is None else pp_cg)
        c2 = _f32c(self.ps_cg_old if ps_cg is None else ps_cg)
        a = [np.zeros((self.allnx, self.allnz), np.float32) for _ in range(8)]
        self._ck(self.L.sfwi_s1_hcg_init(self.allnx, self.allnz, c1, c2, *a))
        return dict(r=(a[0], a[1]), p=(a[2], a[3]), delta=(a[4], a[5]), grad=(a[6], a[7]))

    def hcg_step(self, st, hp_pp, hp_ps):
        """One inner iteration from the summed H p; st updated in place, hp_*
        copied (the rows above the cut are zeroed in the copies).  Returns
        (rr_old, rr_new, alpha, beta)."""
        h1, h2 = np.array(hp_pp, np.float32), np.array(hp_ps, np.float32)
        out = [ctypes.c_double(0.0) for _ in range(4)]
        self._ck(self.L.sfwi_s1_hcg_step(self.allnx, self.allnz, int(self.cut_grad), self.cut_layer, h1, h2,
                                         *st["r"], *st["p"], *st["delta"], *st["grad"],
                                         *[ctypes.byref(o) for o in out]))
        return tuple(o.value for o in out)

    def hcg_finish(self, st):
        self._ck(self.L.sfwi_s1_hcg_finish(self.allnx, self.allnz, *st["delta"], *st["grad"]))
        return st

    def gauss_newton(self, groups, st=None):
        """The inner loop over (context, shots) groups: max_inner Hessian-vector
        products summed over every shot (groups in list order), one host step
        each.  Returns (grad_pp, grad_ps, history): the search direction and the
        (rr_old, rr_new, alpha, beta) of every inner iteration."""
        st = self.hcg_init() if st is None else st
        hist = []
        for _ in range(self.max_inner):
            H = None
            for ctx, shots in groups:
                ctx.hessian_vector(shots, *st["p"])   # in batches of max_batch inside the library
                h = ctx.flush()
                H = h[:2] if H is None else tuple((a + b).astype(np.float32) for a, b in zip(H, h[:2]))
            hist.append(self.hcg_step(st, *H))
        self.hcg_finish(st)
        return st["grad"][0], st["grad"][1], hist

    def direction(self, pp_grad, ps_grad, illum):
        """Search direction from the summed gradients; advances the CG state.
        Returns (dir_pp, dir_ps, beta)."""
        d1 = np.zeros((self.allnx, self.allnz), np.float32)
        d2 = np.zeros_like(d1)
        fm = ctypes.c_double(self.beta_fenmu)
        beta = ctypes.c_double(0.0)
        self._ck(self.L.sfwi_s1_direction(self.allnx, self.allnz, self.iter, self.cut_layer, _f32c(pp_grad),
                                          _f32c(ps_grad), _f32c(illum), self.pp_cg_old, self.ps_cg_old,
                                          ctypes.byref(fm), d1, d2, ctypes.byref(beta)))
        self.beta_fenmu = fm.value
        return d1, d2, beta.value

    def alpha(self, m0, m1, m2):
        return self.L.sfwi_s1_alpha(m0, m1, m2, self.step1, self.step2)

    def update(self, velp, vels, dir_pp, dir_ps, alpha):
        """vel -= alpha dir and the top-row reset, in place on float32 C arrays."""
        self._ck(self.L.sfwi_s1_update(self.allnx, self.allnz, self.cut_layer, alpha, _f32c(dir_pp),
                                       _f32c(dir_ps), velp, vels))

    def iterate(self, ctx, shots, velp, vels, rho=None):
        """One iteration on the context's survey; velp, vels (float32,
        allnx x allnz) are updated in place.  Returns a record of the
        iteration's scalars."""
        return self.iterate_groups([(ctx, shots)], velp, vels, rho)

    def iterate_groups(self, groups, velp, vels, rho=None):
        """One iteration over several (context, shots) groups, e.g. one
        context per shot aperture width; gradients are summed over the groups
        in list order, misfits over the shots in order."""
        misfit = 0.0
        G = None
        for ctx, shots in groups:
            ctx.set_model(velp, vels, rho)
            for v in _batched(ctx, shots, lambda b: ctx.batch_gradient(b)):
                misfit += float(v)
            g = ctx.flush()
            G = g if G is None else tuple((a + b).astype(np.float32) for a, b in zip(G, g))
        d1, d2, beta = self.direction(*G)
        inner = []
        if self.max_inner > 0:   # the Gauss-Newton direction replaces the CG one
            d1, d2, inner = self.gauss_newton(groups)
        m = np.zeros(3)
        for ctx, shots in groups:   # batched forward solves; the same sums as per-shot calls in shot order
            m = ctx.batch_linesearch_parabola(shots, d1, d2, self.step1, self.step2, m)
        alpha = self.alpha(*m)
        self.update(velp, vels, d1, d2, alpha)
        rec = dict(iter=self.iter, misfit=misfit, rms=self._rms(misfit), beta=beta, alpha=alpha,
                   linesearch=tuple(m), inner=inner)
        self.iter += 1
        return rec


def _batched(ctx, shots, fn):
    """fn over consecutive batches of at most max_batch shots, results chained."""
    out = []
    nb = max(int(ctx.p.max_batch), 1)
    for i in range(0, len(shots), nb):
        out.extend(fn(shots[i:i + nb]))
    return out


class Stage2(_CG):
    """Stage-2 loop state and steps (sfwi-stage2/main.cpp; rows above 26 cut)."""

    def __init__(self, allnx, allnz, cut_layer=26):
        super().__init__(allnx, allnz, cut_layer)

    def direction(self, pp_grad, ps_grad, illum):
        """Returns (pp_cg, ps_cg, beta, pp_grad', ps_grad') with the gradients
        as the reference leaves them (rows above the cut zeroed)."""
        g1, g2 = np.array(pp_grad, np.float32), np.array(ps_grad, np.float32)
        c1 = np.zeros((self.allnx, self.allnz), np.float32)
        c2 = np.zeros_like(c1)
        fm = ctypes.c_double(self.beta_fenmu)
        beta = ctypes.c_double(0.0)
        self._ck(self.L.sfwi_s2_direction(self.allnx, self.allnz, self.iter, self.cut_layer, g1, g2, _f32c(illum),
                                          self.pp_cg_old, self.ps_cg_old, ctypes.byref(fm), c1, c2,
                                          ctypes.byref(beta)))
        self.beta_fenmu = fm.value
        return c1, c2, beta.value, g1, g2

    def update(self, image_pp, image_ps, pp_cg, ps_cg, fenzi, fenmu):
        a = ctypes.c_double(0.0)
        self._ck(self.L.sfwi_s2_update(self.allnx, self.allnz, fenzi, fenmu, _f32c(pp_cg), _f32c(ps_cg),
                                       image_pp, image_ps, ctypes.byref(a)))
        return a.value

    def iterate(self, ctx, shots, vp, vs, rho, image_pp, image_ps):
        """One iteration: the images (float32, allnx x allnz) are the Born
        perturbation and are updated in place."""
        return self.iterate_groups([(ctx, shots)], vp, vs, rho, image_pp, image_ps)

    def iterate_groups(self, groups, vp, vs, rho, image_pp, image_ps, per_shot=False):
        """One iteration over several (context, shots) groups, e.g. one
        context per shot aperture width: gradients summed over the groups in
        list order, misfits and the step's sums over the shots in order.
        per_shot: the step's Born solves one shot per call (the reference's
        loop of CALCULATE_ALPHA calls); otherwise batched, with the gradient
        call's residuals in place of the step's first Born solve (the same
        sums bit for bit)."""
        misfit = 0.0
        G = None
        res = []   # batched step: each group's residuals from its gradient call
        for ctx, shots in groups:
            ctx.set_model(vp, vs, rho)
            ctx.set_perturbation(image_pp, image_ps)
            if per_shot:
                for v in _batched(ctx, shots, lambda b: ctx.batch_gradient(b, born=True)):
                    misfit += float(v)
            else:
                nb = max(int(ctx.p.max_batch), 1)
                parts = [ctx.batch_born_gradient_res(shots[i:i + nb]) for i in range(0, len(shots), nb)]
                for m, _, _ in parts:
                    for v in m:
                        misfit += float(v)
                res.append((np.concatenate([p[1] for p in parts]), np.concatenate([p[2] for p in parts])))
            g = ctx.flush()
            G = g if G is None else tuple((a + b).astype(np.float32) for a, b in zip(G, g))
        c1, c2, beta, _, _ = self.direction(*G)
        fenzi = fenmu = 0.0
        for k, (ctx, shots) in enumerate(groups):
            if per_shot:
                for s in shots:
                    fenzi, fenmu = ctx.linesearch_linear(s, c1, c2, fenzi, fenmu)
            else:   # the residuals stand in for the step's first Born solve (same sums bit for bit)
                fenzi, fenmu = ctx.batch_linesearch_linear_res(shots, c1, c2, res[k][0], res[k][1], fenzi, fenmu)
        alpha = self.update(image_pp, image_ps, c1, c2, fenzi, fenmu)
        rec = dict(iter=self.iter, misfit=misfit, rms=self._rms(misfit), beta=beta, alpha=alpha,
                   fenzi=fenzi, fenmu=fenmu)
        self.iter += 1
        return rec
