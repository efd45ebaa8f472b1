"""HIP path (libphd.so through the C ABI) against the FP64 oracle."""
import numpy as np
import pytest

from diffphd_amd import scenes
from helpers import backward_frame, oracle_body, product_body, rel_err, rollout

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cantilever_pair():
    spec = scenes.cantilever()
    return spec, product_body(spec), oracle_body(spec)


def test_apply_ainv_fp64_fp32(cantilever_pair):
    spec, prod, orc = cantilever_pair
    rng = np.random.default_rng(1)
    b = rng.standard_normal(prod.dofs())
    ref = orc.apply_ainv(spec["dt"], b)
    x64 = prod.apply_ainv(spec["dt"], b, precision=0)
    x32 = prod.apply_ainv(spec["dt"], b, precision=1)
    x32r = prod.apply_ainv(spec["dt"], b, precision=1, refine=2)
    print("fp64", rel_err(x64, ref), "fp32", rel_err(x32, ref), "fp32+IR", rel_err(x32r, ref))
    assert rel_err(x64, ref) < 1e-12
    assert rel_err(x32, ref) < 1e-5
    assert rel_err(x32r, ref) < 1e-12


def test_forward_cantilever_fp64(cantilever_pair):
    spec, prod, orc = cantilever_pair
    qp, vp, _ = rollout(prod, spec, 3)
    qo, vo, _ = rollout(orc, spec, 3)
    for i in range(1, 4):
        print(i, rel_err(qp[i], qo[i]), rel_err(vp[i], vo[i]))
        assert rel_err(qp[i], qo[i]) < 1e-8
        assert rel_err(vp[i], vo[i]) < 1e-8


def test_forward_cantilever_mixed(cantilever_pair):
    spec, prod, orc = cantilever_pair
    opt = dict(spec["options"], phd_precision=1)
    qp, vp, _ = rollout(prod, spec, 3, opt)
    qo, vo, _ = rollout(orc, spec, 3)
    for i in range(1, 4):
        print(i, rel_err(qp[i], qo[i]), rel_err(vp[i], vo[i]))
        assert rel_err(qp[i], qo[i]) < 1e-5
        assert rel_err(vp[i], vo[i]) < 1e-5


def test_backward_cantilever_fp64(cantilever_pair):
    spec, prod, orc = cantilever_pair
    qp, vp, ap = rollout(prod, spec, 2)
    qo, vo, ao = rollout(orc, spec, 2)
    g = spec["loss_g"]
    outp = backward_frame(prod, spec, qp, vp, ap, 1, g, np.zeros_like(g))
    outo = backward_frame(orc, spec, qo, vo, ao, 1, g, np.zeros_like(g))
    for name, k in (("dl_dq", 0), ("dl_dv", 1), ("dl_df_ext", 3), ("dl_dmat_w", 4), ("dl_dstate_p", 6)):
        print(name, rel_err(outp[k], outo[k]))
        assert rel_err(outp[k], outo[k]) < 1e-8, name


def test_napkin_contact_forward_fp64():
    spec = scenes.napkin()
    prod, orc = product_body(spec), oracle_body(spec)
    qp, vp, ap = rollout(prod, spec, 40)
    qo, vo, ao = rollout(orc, spec, 40)
    for i in (10, 20, 30, 40):
        print(i, rel_err(qp[i], qo[i]), len(ap[i]), len(ao[i]))
    assert rel_err(qp[40], qo[40]) < 1e-6


def test_press_iterations_with_friction():
    """Heightfield ground + descending press plane, friction 0.3 (8x8x4 hex):
    a fixed count of 16 PD iterations from rest, so that both sides run the
    same NCP + friction + Anderson iterations. The iteration is chaotic for
    near-complementary contacts (omega = 1 - d / sqrt(d^2 + r^2 lambda^2)
    loses its relative precision as omega -> 0), so round-off differences grow
    past ~25 iterations; up to 20 they stay below 2e-12 (FP64)."""
    spec = scenes.press(cells=(8, 8, 4))
    opt = dict(spec["options"], rel_tol=0.0, abs_tol=0.0, max_pd_iter=16)
    prod, orc = product_body(spec), oracle_body(spec)
    qp, vp, ap = rollout(prod, spec, 1, opt)
    qo, vo, ao = rollout(orc, spec, 1, opt)
    lp, lo = prod.last_contact_multipliers(), orc.last_contact_multipliers()
    print("fp64 q", rel_err(qp[1], qo[1]), "lambda", rel_err(lp[0], lo[0]), "lt1", rel_err(lp[1], lo[1]))
    assert ap[1] == ao[1] and len(ao[1]) > 0
    assert rel_err(qp[1], qo[1]) < 1e-10
    for a, b in zip(lp, lo):
        assert rel_err(a, b) < 1e-8
    prod2 = product_body(spec)
    qm, _, _ = rollout(prod2, spec, 1, dict(opt, phd_precision=1))
    print("mixed q", rel_err(qm[1], qo[1]))
    assert rel_err(qm[1], qo[1]) < 1e-6


def test_press_frictionless_converged():
    """Three converged frames (rel_tol 1e-9) of the press without friction."""
    spec = scenes.press(cells=(8, 8, 4), fmu=0.0)
    opt = dict(spec["options"], rel_tol=1e-9, max_pd_iter=4000)
    prod, orc = product_body(spec), oracle_body(spec)
    qp, _, ap = rollout(prod, spec, 3, opt)
    qo, _, ao = rollout(orc, spec, 3, opt)
    for i in range(1, 4):
        print("fp64", i, rel_err(qp[i], qo[i]), len(ap[i]))
        assert ap[i] == ao[i]
        assert rel_err(qp[i], qo[i]) < 1e-8
    prod2 = product_body(spec)
    qm, _, am = rollout(prod2, spec, 3, dict(opt, phd_precision=1))
    print("mixed", rel_err(qm[3], qo[3]))
    assert rel_err(qm[3], qo[3]) < 1e-7
    # Mixed mode solves the contact system in FP32: contacts sitting at the
    # complementarity boundary (lambda ~ 0) may fall on either side of it.
    cand = np.asarray(spec["boundary"][2])
    lm, lo = prod2.last_contact_multipliers()[0], orc.last_contact_multipliers()[0]
    scale = max(np.abs(lo).max(), 1e-300)
    for v in set(am[3]) ^ set(ao[3]):
        c = int(np.flatnonzero(cand == v)[0])
        print("boundary contact", v, lm[c], lo[c])
        assert abs(lm[c]) <= 1e-6 * scale and abs(lo[c]) <= 1e-6 * scale


def test_press_friction_rollout_at_solve_tolerance():
    """Three frames with friction at rel_tol 1e-5: past the iteration-exact
    regime the two trajectories agree to the tolerance of the solve (measured
    ~4e-4), and each frame's contact multipliers satisfy the NCP: lambda >= 0,
    tangential impulse inside the cone mu * lambda."""
    spec = scenes.press(cells=(8, 8, 4))
    opt = dict(spec["options"], rel_tol=1e-5, max_pd_iter=400)
    prod, orc = product_body(spec), oracle_body(spec)
    qp, _, ap = rollout(prod, spec, 3, opt)
    qo, _, ao = rollout(orc, spec, 3, opt)
    for i in range(1, 4):
        print("fp64", i, rel_err(qp[i], qo[i]), len(ap[i]), len(ao[i]))
        assert rel_err(qp[i], qo[i]) < 2e-3
    lam, lt1, lt2 = prod.last_contact_multipliers()
    assert np.all(lam >= 0)
    assert np.all(np.hypot(lt1, lt2) <= 0.3 * lam * (1 + 1e-12) + 1e-14)
    assert set(ap[3]) == set(np.asarray(spec["boundary"][2])[lam > 0].tolist())


def test_napkin_backward_with_contact_basis_fp64():
    """Adjoint through NCP contact frames: backbone + contact-basis correction
    (reference Woodbury, backward_gpu.cpp:922-1208) against the oracle."""
    spec = scenes.napkin()
    fwd = dict(spec["options"], rel_tol=1e-10, abs_tol=1e-14)
    bwd = dict(fwd)
    prod, orc = product_body(spec), oracle_body(spec)
    qp, vp, ap = rollout(prod, spec, 45, fwd)
    qo, vo, ao = rollout(orc, spec, 45, fwd)
    assert ap[45] == ao[45] and len(ap[45]) > 0
    rng = np.random.default_rng(2)
    dq = rng.standard_normal(prod.dofs())
    dv = rng.standard_normal(prod.dofs())
    for i in (44, 43):
        outp = backward_frame(prod, spec, qp, vp, ap, i, dq, dv, bwd)
        outo = backward_frame(orc, spec, qo, vo, ao, i, dq, dv, bwd)
        for name, k in (("dl_dq", 0), ("dl_dv", 1), ("dl_dmat_w", 4)):
            print(i, name, rel_err(outp[k], outo[k]))
            assert rel_err(outp[k], outo[k]) < 1e-6, (i, name)
        dq, dv = outo[0], outo[1]


def test_restitution_post_pass_matches_oracle():
    """One converged frictionless step of a beam landing at 0.5 m/s with
    restitution 0.6 (forward_gpu.cpp:1483-1531): every contact node gets its
    normal velocity corrected; the HIP body equals the oracle (FP64 and mixed)."""
    spec = scenes.beam(cells=(3, 3, 2), E=1e5, nu=0.4, v_std=0.0, seed=1)
    spec["v0"] = np.zeros_like(spec["q0"])
    spec["v0"][2::3] = -0.5
    opt = dict(spec["options"], friction_mu=0.0, rel_tol=1e-11, abs_tol=1e-14, max_pd_iter=5000, restitution=0.6)
    qo, vo, ao = rollout(oracle_body(spec), spec, 2, opt)
    assert len(ao[1]) > 0  # lands in frame 1 (and bounces off: no contact in frame 2)
    for prec, tol in ((0, 1e-10), (1, 1e-6)):
        qp, vp, ap = rollout(product_body(spec), spec, 2, dict(opt, phd_precision=prec))
        for i in (1, 2):
            print("restitution", prec, i, rel_err(qp[i], qo[i]), rel_err(vp[i], vo[i]), len(ap[i]))
            assert ap[i] == ao[i]
            assert rel_err(qp[i], qo[i]) < tol and rel_err(vp[i], vo[i]) < 100 * tol


def test_fixed_active_set_replay_matches_oracle_and_tape():
    """Backward without the forward tape: the HIP body replays the frame over
    the given active set (ReplayAlgPhdFixedActiveSet, backward_gpu.cpp:113-418)
    on the device. Against the oracle's replay (both fresh bodies) at the FP64
    bar, and against the HIP body's own tape-based adjoint on the same
    converged frictionless step."""
    from diffphd_amd.py_diff_pd_core import StdRealVector
    spec = scenes.beam(cells=(3, 3, 2), E=1e5, nu=0.4, v_std=0.0, seed=1)
    spec["v0"] = np.zeros_like(spec["q0"])
    spec["v0"][2::3] = -0.3
    opt = dict(spec["options"], friction_mu=0.0, rel_tol=1e-12, abs_tol=1e-14, max_pd_iter=20000,
               alg_phd_bwd_wt_max_iter=20000)
    body = product_body(spec)
    q, v, acts = rollout(body, spec, 2, opt)
    assert len(acts[2]) > 0
    n = q[-1].size
    g = np.random.default_rng(7).standard_normal(n)
    bopt = dict(opt, alg_phd_bwd_state_woodbury=1)
    res = []
    for b in (body, product_body(spec), oracle_body(spec)):
        o = [StdRealVector(0) for _ in range(8)]
        b.PyBackward("pd_eigen_alg_phd", q[1], v[1], [], np.zeros(n), spec["dt"], q[2], v[2], acts[2], g,
                     np.zeros(n), bopt, *o)
        res.append([np.asarray(x) for x in o])
        print("from tape", b.last_backward_stats()["from_tape"])
    tape, hip_rep, orc_rep = res
    for name, k in (("dl_dq", 0), ("dl_dv", 1), ("dl_dmat_w", 4)):
        print(name, "hip replay vs oracle replay", rel_err(hip_rep[k], orc_rep[k]),
              "hip replay vs hip tape", rel_err(hip_rep[k], tape[k]))
        assert rel_err(hip_rep[k], orc_rep[k]) < 1e-8, name
        assert rel_err(hip_rep[k], tape[k]) < 1e-6, name
