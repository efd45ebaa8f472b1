"""The 64x64x32 press (131,072 hex elements, 139,425 vertices) on the GPU:
properties that hold at any size, checked at the benchmark's full size."""
import numpy as np
import pytest

from diffphd_amd import scenes
from helpers import backward_frame, product_body, rollout

pytestmark = pytest.mark.gpu


def test_press_64x64x32_contact_invariants_and_unclipped_backward():
    """Mixed precision, 12 frames at up to 40 PD iterations (the plane reaches
    the block after ~8): finite states, lambda >= 0, every tangential impulse
    inside the Coulomb cone mu lambda, active sets = {lambda > 0}; then the last
    two frames' adjoints with per-element K_pd and the trust region: finite,
    none reaches the clip norm."""
    spec = scenes.press(cells=(64, 64, 32))
    fmu = spec["options"]["friction_mu"]
    opt = dict(spec["options"], max_pd_iter=40, phd_precision=1)
    body = product_body(spec)
    frames = 12
    q, v, acts = rollout(body, spec, frames, opt)
    cand = np.asarray(spec["boundary"][2])
    for i in range(1, frames + 1):
        assert np.all(np.isfinite(q[i])) and np.all(np.isfinite(v[i]))
    lam, lt1, lt2 = body.last_contact_multipliers()
    print("active contacts per frame", [len(a) for a in acts[1:]], "lambda max", lam.max())
    assert max(len(a) for a in acts) > 500 and len(acts[-1]) > 100
    assert np.all(lam >= 0)
    assert np.all(np.hypot(lt1, lt2) <= fmu * lam * (1 + 1e-12) + 1e-14)
    assert set(acts[-1]) == set(cand[lam > 0].tolist())
    n = q[-1].size
    dq = 2.0 * (q[-1] - spec["target"]) / n
    dv = np.zeros(n)
    bopt = dict(opt, max_pd_iter=200, alg_phd_bwd_exact_kpd=1, alg_phd_bwd_kpd_filter=3)
    for i in reversed(range(frames - 2, frames)):
        body.SetPressPlaneHeight(spec["plane_schedule"][i])
        out = backward_frame(body, spec, q, v, acts, i, dq, dv, bopt)
        dq, dv = out[0], out[1]
        st = body.last_backward_stats()
        print("backward frame", i, st, np.linalg.norm(dq))
        assert st["clipped"] == 0
        assert np.all(np.isfinite(dq)) and np.all(np.isfinite(dv)) and np.linalg.norm(dq) > 0
