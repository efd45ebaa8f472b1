"""Mesh frictional boundary (deformable-vs-deformable contact through the
body's own surface triangles, MeshFrictionalBoundary) in the HIP body against
the oracle: a rigid slab with friction and two deformable blocks, at a fixed
iteration count, converged with the adjoint, and with restitution."""
import numpy as np
import pytest

from diffphd_amd import scenes
from diffphd_amd.py_diff_pd_core import StdRealVector
from helpers import backward_frame, oracle_body, product_body, rel_err, rollout

pytestmark = pytest.mark.gpu


# Two deformable blocks pushing each other (both sides candidates) are a
# chaotic iteration: after 8 iterations of a frame the two sides' round-off
# has grown to ~1e-7 (frame 1 stays at ~1e-12), so that case runs 4.
@pytest.mark.parametrize("fixed,both,its", [("all", False, 12), ("base", True, 4)])
def test_mesh_contact_matches_oracle_at_fixed_iterations(fixed, both, its):
    spec = scenes.stack(fixed_bottom=fixed, both_sides=both, fmu=0.5)
    opt = dict(spec["options"], rel_tol=0.0, abs_tol=0.0, max_pd_iter=its)
    qp, vp, ap = rollout(product_body(spec), spec, 3, opt)
    qo, vo, ao = rollout(oracle_body(spec), spec, 3, opt)
    print(fixed, "active", [len(a) for a in ap], [len(a) for a in ao], "q", rel_err(qp[-1], qo[-1]))
    assert any(len(a) > 0 for a in ao)
    assert ap == ao
    assert rel_err(qp[-1], qo[-1]) < 1e-10


def test_deformable_stack_forward_and_adjoint_match_oracle():
    """The lower block deforms under the upper one while they touch (only the
    upper block's vertices are candidates: with both sides querying each
    other the frictionless iteration chatters and stops at the cap, in the
    oracle too). One frame in which the 25 contacts activate, at a fixed 12
    iterations (3e-12 apart); later frames inherit each side's multiplier
    warm start, whose contacts at lambda ~ 0 fall on either side of the
    complementarity boundary from round-off, so they are not compared. The
    adjoint, a linear fixed point, then runs to convergence on each side's
    own tape."""
    spec = scenes.stack(fixed_bottom="base", both_sides=False, fmu=0.0)
    opt = dict(spec["options"], rel_tol=0.0, abs_tol=0.0, max_pd_iter=12)
    prod, orc = product_body(spec), oracle_body(spec)
    qp, vp, ap = rollout(prod, spec, 1, opt)
    qo, vo, ao = rollout(orc, spec, 1, opt)
    print("active", [len(a) for a in ap], "q", rel_err(qp[-1], qo[-1]))
    assert ap == ao and len(ao[-1]) > 0
    assert rel_err(qp[-1], qo[-1]) < 1e-10
    # the lower block's top face has moved under the contacts
    nt = spec["split"]
    assert np.abs(qp[-1][3 * nt:] - spec["q0"][3 * nt:]).max() > 1e-6
    rng = np.random.default_rng(4)
    dq, dv = rng.standard_normal(prod.dofs()), rng.standard_normal(prod.dofs())
    bopt = dict(opt, max_pd_iter=5000, rel_tol=1e-11, abs_tol=1e-14, alg_phd_bwd_wt_max_iter=5000)
    outp = backward_frame(prod, spec, qp, vp, ap, 0, dq, dv, bopt)
    outo = backward_frame(orc, spec, qo, vo, ao, 0, dq, dv, bopt)
    for name, k in (("dl_dq", 0), ("dl_dv", 1), ("dl_dmat_w", 4)):
        print(name, rel_err(outp[k], outo[k]))
        assert rel_err(outp[k], outo[k]) < 1e-8, name


def test_mesh_restitution_matches_oracle():
    # within 6 radii at the frame start (the candidate filter), hitting at 0.5 m/s
    spec = scenes.stack(fixed_bottom="all", both_sides=False, fmu=0.0, v_down=0.5, gap=1e-3)
    opt = dict(spec["options"], rel_tol=1e-11, abs_tol=1e-14, max_pd_iter=5000, restitution=0.5)
    qp, vp, _ = rollout(product_body(spec), spec, 2, opt)
    qo, vo, _ = rollout(oracle_body(spec), spec, 2, opt)
    base = rollout(oracle_body(spec), spec, 2, dict(opt, restitution=0.0))[0]
    print("restitution changes q by", rel_err(qo[-1], base[-1]), "HIP vs oracle", rel_err(qp[-1], qo[-1]))
    assert rel_err(qo[-1], base[-1]) > 1e-6  # the post-pass acts
    assert rel_err(qp[-1], qo[-1]) < 1e-8


def test_mesh_replay_is_rejected():
    spec = scenes.stack(fixed_bottom="all", both_sides=False)
    body = product_body(spec)
    q, v, acts = rollout(body, spec, 2, spec["options"])
    n = body.dofs()
    o = [StdRealVector(0) for _ in range(8)]
    fresh = product_body(spec)
    with pytest.raises(RuntimeError, match="mesh boundary"):
        fresh.PyBackward("pd_eigen_alg_phd", q[1], v[1], [], np.zeros(n), spec["dt"], q[2], v[2], acts[2],
                         np.ones(n), np.zeros(n), spec["options"], *o)
