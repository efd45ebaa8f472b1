"""The DiffPD baseline forward (ForwardProjectiveDynamics with a pd_eigen
method: L-BFGS PD inside the active-set contact loop) in the HIP body: the
reference's published napkin answers, and parity with the oracle at tight
tolerances with and without contact."""
import numpy as np
import pytest

from diffphd_amd import scenes
from diffphd_amd.env import simulate
from helpers import oracle_body, product_body, rel_err, rollout
from test_oracle_baseline import BASELINE_OPT

pytestmark = pytest.mark.gpu


def _z(info):
    return np.array(info["q"]).reshape(len(info["q"]), -1, 3)[:, :, 2]


def test_napkin_baseline_known_answers_on_the_gpu():
    """README.md:214-217: max z_std after landing 0.06569 for the baseline and
    max |z_alg_phd - z_baseline| = 1.7370e-02, both solvers on the HIP body."""
    spec = scenes.napkin()
    zero = lambda q, v: (0.0, 0 * q, 0 * v)
    _, _, ib = simulate(product_body(spec), spec["dt"], spec["frames"], "pd_eigen_pcg", BASELINE_OPT, spec["q0"],
                        spec["v0"], zero, require_grad=False)
    _, _, ia = simulate(product_body(spec), spec["dt"], spec["frames"], "pd_eigen_alg_phd", spec["options"],
                        spec["q0"], spec["v0"], zero, require_grad=False)
    _, _, io = simulate(oracle_body(spec), spec["dt"], spec["frames"], "pd_eigen_pcg", BASELINE_OPT, spec["q0"],
                        spec["v0"], zero, require_grad=False)
    zb, za, zo = _z(ib), _z(ia), _z(io)
    landed = zb[spec["frames"] - 35:].std(axis=1).max()
    drift = np.abs(za - zb).max()
    print("HIP baseline landed %.6f drift %.5e, HIP vs oracle baseline max |dz| %.3e" % (landed, drift,
                                                                                       np.abs(zb - zo).max()))
    assert ib["active_contact_indices"] == io["active_contact_indices"]
    assert abs(landed - 0.06569) < 5e-6
    assert abs(drift - 1.7370e-02) < 5e-7


def test_baseline_matches_oracle_at_tight_tolerance():
    """Muscles, no contact: 3 frames at rel_tol 1e-12 (L-BFGS and plain PD)."""
    spec = scenes.actuated_cantilever(cells=(6, 2, 2))
    tight = dict(BASELINE_OPT, rel_tol=1e-12, abs_tol=1e-12, max_pd_iter=20000)
    for opt in (tight, dict(tight, use_bfgs=0)):
        qp, _, _ = rollout(product_body(spec), spec, 3, opt, method="pd_eigen_pcg")
        qo, _, _ = rollout(oracle_body(spec), spec, 3, opt, method="pd_eigen_pcg")
        print("use_bfgs", opt["use_bfgs"], "q vs oracle", rel_err(qp[-1], qo[-1]))
        assert rel_err(qp[-1], qo[-1]) < 1e-9


def test_baseline_contact_frames_match_oracle():
    """The napkin through first contact at rel_tol 1e-10: the same active sets
    and states as the oracle, and frozen nodes where they stood."""
    spec = scenes.napkin()
    opt = dict(BASELINE_OPT, rel_tol=1e-10, abs_tol=1e-12)
    frames = 48
    qp, _, ap = rollout(product_body(spec), spec, frames, opt, method="pd_eigen_pcg")
    qo, _, ao = rollout(oracle_body(spec), spec, frames, opt, method="pd_eigen_pcg")
    print("active", [len(a) for a in ap][::4], "q", rel_err(qp[-1], qo[-1]))
    assert ap == ao and max(len(a) for a in ao) > 0
    assert rel_err(qp[-1], qo[-1]) < 1e-8
    # fixed-contact variant: the given nodes stay put
    body = product_body(spec)
    n = body.dofs()
    from diffphd_amd.py_diff_pd_core import StdRealVector
    nodes = list(spec["boundary"][2])[:4]
    qn, vn, act = StdRealVector(n), StdRealVector(n), list(nodes)
    body.PyForward("pd_eigen_pcg_fixed_contact", qo[30], np.zeros(n), [], np.zeros(n), spec["dt"], opt, qn, vn, act)
    qn = np.asarray(qn)
    for k in nodes:
        assert np.array_equal(qn[3 * k:3 * k + 3], qo[30][3 * k:3 * k + 3])


def test_baseline_backward_matches_oracle():
    """The baseline's adjoint on the HIP body against the oracle: the actuated
    cantilever (no contact) and napkin frames resting on the sphere (frozen
    contact nodes), every gradient output to 1e-8."""
    from helpers import backward_frame
    spec = scenes.actuated_cantilever(cells=(6, 2, 2))
    opt = dict(BASELINE_OPT, rel_tol=1e-12, abs_tol=1e-13, max_pd_iter=20000)
    prod, orc = product_body(spec), oracle_body(spec)
    qp, vp, ap = rollout(prod, spec, 3, opt, method="pd_eigen_pcg")
    qo, vo, ao = rollout(orc, spec, 3, opt, method="pd_eigen_pcg")
    rng = np.random.default_rng(11)
    dq, dv = rng.standard_normal(prod.dofs()), rng.standard_normal(prod.dofs())
    outp = backward_frame(prod, spec, qp, vp, ap, 2, dq, dv, opt, method="pd_eigen_pcg")
    outo = backward_frame(orc, spec, qo, vo, ao, 2, dq, dv, opt, method="pd_eigen_pcg")
    for name, k in (("dl_dq", 0), ("dl_dv", 1), ("dl_da", 2), ("dl_df_ext", 3), ("dl_dmat_w", 4), ("dl_dact_w", 5),
                    ("dl_dstate_p", 6)):
        print("cantilever", name, rel_err(outp[k], outo[k]))
        assert np.abs(outo[k]).max() > 0 and rel_err(outp[k], outo[k]) < 1e-8, name
    spec = scenes.napkin()
    opt = dict(BASELINE_OPT, rel_tol=1e-12, abs_tol=1e-13)
    prod, orc = product_body(spec), oracle_body(spec)
    qp, vp, ap = rollout(prod, spec, 42, opt, method="pd_eigen_pcg")
    qo, vo, ao = rollout(orc, spec, 42, opt, method="pd_eigen_pcg")
    assert ap == ao and len(ao[-1]) > 0
    dq, dv = rng.standard_normal(prod.dofs()), rng.standard_normal(prod.dofs())
    outp = backward_frame(prod, spec, qp, vp, ap, 41, dq, dv, opt, method="pd_eigen_pcg")
    outo = backward_frame(orc, spec, qo, vo, ao, 41, dq, dv, opt, method="pd_eigen_pcg")
    for name, k in (("dl_dq", 0), ("dl_dv", 1), ("dl_dmat_w", 4), ("dl_dstate_p", 6)):
        print("napkin", len(ao[-1]), "frozen", name, rel_err(outp[k], outo[k]))
        assert rel_err(outp[k], outo[k]) < 1e-8, name
