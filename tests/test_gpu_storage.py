"""16-bit boundary-strip storage (FP16 / BF16 with a per-step max-abs scale).

Only the reverse reconstruction reads the strips, so the misfit (a function of
the forward solve) must equal the FP32-storage misfit bitwise, and the
gradient must stay close to the FP64 oracle: FP16 within the 1e-3 relative-L2
bar of the mixed-precision path, BF16 (8-bit mantissa) within 1e-2.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sfwi_amd import workloads as W
from tests.helpers import context, lib_shot, oracle_cfg, rel_l2

pytestmark = pytest.mark.gpu

BAR = {"f16": 1e-3, "bf16": 1e-2}


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("percell", [False, True])
def test_strip_storage_gradient_error(storage, percell):
    sv = W.tiny(nt=320)
    w = W.ricker(sv.nt, sv.dt, sv.f0)
    shots, ref = [], []
    for sh in sv.shots:
        c = oracle_cfg(sv, sh)
        oz, ox = O.forward(c, sv.vp, sv.vs, sv.rho, w, sh.gc, "f32")
        shots.append(lib_shot(sv, sh, oz, ox))
        ref.append(O.gradient(c, sv.vp0, sv.vs0, sv.rho0, w, sh.gc, oz, ox, "f64"))
    r64 = [sum(r[k] for r in ref) for k in range(3)]
    out = {}
    for st in ("f32", storage):
        with context(sv, max_batch=2, storage=st, march_cols=-1 if percell else 8) as ctx:
            ctx.set_wavelet(w)
            ctx.set_model(sv.vp0, sv.vs0, sv.rho0)
            m = ctx.batch_gradient(shots)
            out[st] = (m, ctx.flush())
    assert np.array_equal(out["f32"][0], out[storage][0])
    g32, g16 = out["f32"][1], out[storage][1]
    e32 = [rel_l2(a, b) for a, b in zip(g32, r64)]
    e16 = [rel_l2(a, b) for a, b in zip(g16, r64)]
    print("strip storage %s: grad err vs FP64 (vp, vs, illum) f32 %s, %s %s" % (storage, e32, storage, e16))
    assert max(e16) <= BAR[storage]
    assert max(e16) > max(e32) * 0.5 or max(e16) < 1e-6   # the 16-bit strips are really in use
