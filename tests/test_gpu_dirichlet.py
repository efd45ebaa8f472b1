"""Per-DOF Dirichlet conditions on the shared factor (the bordered correction
of every global solve) against the oracle's per-axis factors: A^-1, a
converged rollout, its adjoint, mixed precision, and the rejected
combination with a frictional boundary."""
import numpy as np
import pytest

from diffphd_amd import scenes
from helpers import backward_frame, oracle_body, product_body, rel_err, rollout
from test_oracle_dirichlet import slider

pytestmark = pytest.mark.gpu


def test_per_dof_dirichlet_ainv_matches_oracle():
    spec = slider()
    prod, orc = product_body(spec), oracle_body(spec)
    b = np.random.default_rng(1).standard_normal(prod.dofs())
    xo = orc.apply_ainv(spec["dt"], b)
    x64 = prod.apply_ainv(spec["dt"], b, precision=0)
    x32 = prod.apply_ainv(spec["dt"], b, precision=1, refine=2)
    print("A^-1 b: fp64", rel_err(x64, xo), "fp32 + 2 refinements", rel_err(x32, xo))
    assert rel_err(x64, xo) < 1e-12 and rel_err(x32, xo) < 1e-10
    for dof, _ in spec["dirichlet"]:
        assert x64[dof] == b[dof]


def test_per_dof_dirichlet_rollout_and_adjoint_match_oracle():
    spec = slider()
    spec["v0"] = 0.05 * np.random.default_rng(2).standard_normal(spec["q0"].size)
    for dof, _ in spec["dirichlet"]:
        spec["v0"][dof] = 0.0
    opt = dict(spec["options"], rel_tol=1e-11, abs_tol=1e-14, max_pd_iter=5000)
    prod, orc = product_body(spec), oracle_body(spec)
    qp, vp, ap = rollout(prod, spec, 4, opt)
    qo, vo, ao = rollout(orc, spec, 4, opt)
    print("q after 4 frames", rel_err(qp[-1], qo[-1]))
    assert rel_err(qp[-1], qo[-1]) < 1e-8
    for dof, val in spec["dirichlet"]:
        assert qp[-1][dof] == val
    rng = np.random.default_rng(3)
    dq, dv = rng.standard_normal(prod.dofs()), rng.standard_normal(prod.dofs())
    bopt = dict(opt, alg_phd_bwd_wt_max_iter=5000)
    outp = backward_frame(prod, spec, qp, vp, ap, 3, dq, dv, bopt)
    outo = backward_frame(orc, spec, qo, vo, ao, 3, dq, dv, bopt)
    for name, k in (("dl_dq", 0), ("dl_dv", 1), ("dl_dmat_w", 4)):
        print(name, rel_err(outp[k], outo[k]))
        assert rel_err(outp[k], outo[k]) < 1e-8, name
    # mixed precision
    qm, _, _ = rollout(product_body(spec), spec, 4, dict(opt, phd_precision=1))
    print("mixed q", rel_err(qm[-1], qo[-1]))
    assert rel_err(qm[-1], qo[-1]) < 1e-5


def test_per_dof_dirichlet_with_contact_is_rejected():
    spec = scenes.beam(cells=(3, 3, 2))
    spec["dirichlet"] = [(0, float(spec["q0"][0]))]
    with pytest.raises(RuntimeError, match="Per-DOF Dirichlet"):
        body = product_body(spec)
        rollout(body, spec, 1)
