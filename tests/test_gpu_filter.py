"""Filtered proximal-map Jacobians in the HIP backward (alg_phd_bwd_kpd_filter)
against the FP64 oracle, the trust region's effect on the press backbone, and
the exact-adjoint routing against central differences of the HIP forward."""
import numpy as np
import pytest

from diffphd_amd import scenes
from diffphd_amd.py_diff_pd_core import StdRealVector
from helpers import backward_frame, oracle_body, product_body, rel_err, rollout

pytestmark = pytest.mark.gpu


def _contrast_frame():
    """6x6x3 press, 100x stiffness contrast, friction 0.3: one frame of a fixed
    10 PD iterations from rest on fresh bodies (iteration-exact regime: the
    contact iteration amplifies round-off, ~3e-10 after 16 iterations here, and
    a second frame from the two slightly different states ~3e-9)."""
    # seed 3: 98 active contacts, the smallest lambda > 0 is 3e-3 of the largest,
    # so round-off cannot move a contact across lambda = 0
    spec = scenes.press(cells=(6, 6, 3), decades=2.0, seed=3)
    opt = dict(spec["options"], rel_tol=0.0, abs_tol=0.0, max_pd_iter=10)
    prod, orc = product_body(spec), oracle_body(spec)
    qp, vp, ap = rollout(prod, spec, 1, opt)
    qo, vo, ao = rollout(orc, spec, 1, opt)
    print("forward", rel_err(qp[1], qo[1]), len(ap[1]))
    assert ap[1] == ao[1] and len(ao[1]) > 0
    assert rel_err(qp[1], qo[1]) < 1e-10
    return spec, opt, (prod, qp, vp, ap), (orc, qo, vo, ao)


@pytest.mark.parametrize("kfilter,exact", [(1, 0), (2, 0), (3, 0), (1, 1), (3, 1)])
def test_filtered_backward_matches_oracle(kfilter, exact):
    """A fixed count of backbone iterations with the filtered K_pd: the HIP
    kernel filters the SVD-basis blocks, the oracle the full 9x9 block through
    its own Jacobi eigen-solver. The Woodbury correction is skipped: truncated
    at a fixed count, the oracle's per-basis fixed points and the HIP body's
    one constrained fixed point are different iterations (they agree at
    convergence, test_contact_woodbury_gradients_match_oracle), so dl_dmat_w
    here comes from the filtered backbone adjoint on both sides."""
    spec, opt, (prod, qp, vp, ap), (orc, qo, vo, ao) = _contrast_frame()
    rng = np.random.default_rng(5)
    dq, dv = rng.standard_normal(prod.dofs()), rng.standard_normal(prod.dofs())
    # 80 backbone iterations for the trust region on the mean-stiffness K_pd: its
    # unfiltered stage diverges within them (after ~38), so the clamp stage runs
    its = 80 if (kfilter, exact) == (3, 0) else 30
    o2 = dict(opt, max_pd_iter=its, skip_contact_woodbury=1, alg_phd_bwd_kpd_filter=kfilter,
              alg_phd_bwd_exact_kpd=exact, alg_phd_bwd_trust_radius=0.8)
    outp = backward_frame(prod, spec, qp, vp, ap, 0, dq, dv, o2)
    outo = backward_frame(orc, spec, qo, vo, ao, 0, dq, dv, o2)
    sp, so = prod.last_backward_stats(), orc.last_backward_stats()
    print("stats hip", sp, "oracle", so)
    assert sp["trust_stage"] == so["trust_stage"]
    assert sp["clipped"] == so["clipped"] == 0
    assert abs(sp["filtered"] - so["filtered"]) <= 0.01 * so["samples"]
    if kfilter in (1, 2) or exact == 0:
        assert so["filtered"] > 0  # the filter acts on this scene
    if kfilter == 3:  # the trust region leaves the converging exact operator alone
        assert so["trust_stage"] == (1 if exact == 0 else 0)
    for name, k in (("dl_dq", 0), ("dl_dv", 1), ("dl_dmat_w", 4)):
        print(name, rel_err(outp[k], outo[k]), np.linalg.norm(outo[k]))
        assert rel_err(outp[k], outo[k]) < 1e-8, name


def _press_backward(cells, frames, bwd_frames, opt_extra, precision):
    spec = scenes.press(cells=cells)
    opt = dict(spec["options"], max_pd_iter=100, phd_precision=precision)
    body = product_body(spec)
    q, v, acts = rollout(body, spec, frames, opt)
    n = q[-1].size
    dq = 2.0 * (q[-1] - spec["target"]) / n
    dv = np.zeros(n)
    stats = []
    bopt = dict(opt, **opt_extra)
    # the tape holds the last forward frames; walk them backwards
    for i in reversed(range(frames - bwd_frames, frames)):
        body.SetPressPlaneHeight(spec["plane_schedule"][i])
        out = backward_frame(body, spec, q, v, acts, i, dq, dv, bopt)
        dq, dv = out[0], out[1]
        st = body.last_backward_stats()
        stats.append(st)
        print(cells, i, {k: st[k] for k in ("clipped", "backbone_iters", "trust_stage", "filtered", "basis_rows")},
              np.linalg.norm(dq))
        assert np.all(np.isfinite(dq)) and np.all(np.isfinite(dv))
    return stats


@pytest.mark.parametrize("cells", [(8, 8, 4), (16, 16, 8)])
def test_press_backward_trust_region_not_clipped(cells):
    """The press's backward with the reference's K_pd (mean stiffness, no
    filter) reaches the clip norm; with per-element stiffness and the trust
    region no backward frame is clipped."""
    trust = _press_backward(cells, 10, 3, dict(alg_phd_bwd_exact_kpd=1, alg_phd_bwd_kpd_filter=3), 1)
    assert all(s["clipped"] == 0 for s in trust)
    ref = _press_backward(cells, 10, 1, {}, 1)
    print("reference K_pd clipped:", [s["clipped"] for s in ref])


def test_frictionless_press_step_exact_adjoint_matches_fd():
    """One frictionless step of a press on a flat ground with its contacts
    active: the adjoint with per-element K_pd, the trust region and the
    Woodbury-corrected state gradients against central differences of the HIP
    forward in the initial velocity (the active set is the same at +-eps)."""
    spec = scenes.press(cells=(6, 6, 3), fmu=0.0, bump=0.0)
    opt = dict(spec["options"], max_pd_iter=20000, rel_tol=1e-12, abs_tol=1e-14)
    prod = product_body(spec)
    q, v, acts = rollout(prod, spec, 4, opt)
    n = q[-1].size
    g = np.random.default_rng(3).standard_normal(n)
    dvec = np.random.default_rng(4).standard_normal(n)

    def step(body, vv):
        body.SetPressPlaneHeight(spec["plane_schedule"][4])
        qn, vn, act = StdRealVector(n), StdRealVector(n), []
        body.PyForward("pd_eigen_alg_phd", q[-1], vv, [], np.zeros(n), spec["dt"], opt, qn, vn, act)
        return float(g @ np.asarray(qn)), act, np.asarray(qn), np.asarray(vn)

    eps = 1e-5
    lp, ap_, _, _ = step(prod, v[-1] + eps * dvec)
    lm, am_, _, _ = step(prod, v[-1] - eps * dvec)
    _, act0, qn, vn = step(prod, v[-1])
    assert len(act0) > 0 and ap_ == act0 and am_ == act0
    fd = (lp - lm) / (2 * eps)
    o = [StdRealVector(0) for _ in range(8)]
    bopt = dict(opt, alg_phd_bwd_exact_kpd=1, alg_phd_bwd_kpd_filter=3, alg_phd_bwd_state_woodbury=1,
                alg_phd_bwd_wt_max_iter=20000)
    prod.PyBackward("pd_eigen_alg_phd", q[-1], v[-1], [], np.zeros(n), spec["dt"], qn, vn, act0, g, np.zeros(n), bopt,
                    *o)
    an = np.asarray(o[1]) @ dvec
    st = prod.last_backward_stats()
    print("dl/dv0 . d", an, "fd", fd, "rel", abs(an - fd) / abs(fd), st)
    assert st["clipped"] == 0 and st["backbone_converged"] == 1
    assert abs(an - fd) < 1e-5 * abs(fd)
