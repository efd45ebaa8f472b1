"""The contact state forces (AddStateForce("planar_contact"), a penalty floor
with friction; AddStateForce("arc_contact"), a cylindrical trough;
AddStateForce("mesh_contact"), two surfaces of the body meeting;
AddStateForce("billiard_ball"), vertex blocks as balls; and
AddStateForce("hydrodynamics"), drag and thrust of moving water) in the HIP body
against the oracle: rollouts and adjoints in both friction branches, through
the alg_phd and the baseline methods (the mesh contact's backward is zero, as
in the reference, so its adjoint runs through the other terms)."""
import numpy as np
import pytest

from diffphd_amd import scenes
from helpers import backward_frame, oracle_body, product_body, rel_err, rollout
from test_oracle_baseline import BASELINE_OPT

pytestmark = pytest.mark.gpu


def _compare(spec, label, nparams, baseline=True):
    runs = [("pd_eigen_alg_phd", dict(spec["options"], rel_tol=1e-12, abs_tol=1e-14, max_pd_iter=20000))]
    if baseline:
        runs.append(("pd_eigen_pcg", dict(BASELINE_OPT, rel_tol=1e-12, abs_tol=1e-13, max_pd_iter=20000)))
    for method, opt in runs:
        prod, orc = product_body(spec), oracle_body(spec)
        assert prod.NumOfStateForceParameters() == orc.NumOfStateForceParameters() == nparams
        qp, vp, ap = rollout(prod, spec, 3, opt, method=method)
        qo, vo, ao = rollout(orc, spec, 3, opt, method=method)
        print(label, method, "q", rel_err(qp[-1], qo[-1]))
        assert rel_err(qp[-1], qo[-1]) < 1e-8
        rng = np.random.default_rng(5)
        dq, dv = rng.standard_normal(prod.dofs()), rng.standard_normal(prod.dofs())
        outp = backward_frame(prod, spec, qp, vp, ap, 2, dq, dv, opt, method=method)
        outo = backward_frame(orc, spec, qo, vo, ao, 2, dq, dv, opt, method=method)
        for name, k in (("dl_dq", 0), ("dl_dv", 1), ("dl_dstate_p", 6)):
            print(label, method, name, rel_err(outp[k], outo[k]))
            assert np.abs(outo[k]).max() > 0 and rel_err(outp[k], outo[k]) < 1e-8, name


@pytest.mark.parametrize("scene", ["penalty_drop", "arc_drop", "mesh_stack"])
@pytest.mark.parametrize("kf", [0.005, 1e3])
def test_penalty_floor_matches_oracle(scene, kf):
    _compare(getattr(scenes, scene)(kf=kf), "%s %g" % (scene, kf), 7 if scene == "mesh_stack" else 6)


def test_billiard_balls_match_oracle():
    """Eight one-cell balls in a row (BilliardBallStateForce): 16 parameters
    (stiffness and friction per ball), no gravity."""
    _compare(scenes.billiard(), "billiard", 16)


@pytest.mark.parametrize("kind,rho_water,max_force", [("hex", 1000.0, 1e3), ("hex", 1000.0, 4e-3),
                                                      ("tet", 10.0, 1e3), ("tet", 10.0, 4e-5)])
def test_hydrodynamics_match_oracle(kind, rho_water, max_force):
    """A block in moving water (HydrodynamicsStateForce) on hex quads and tet
    triangles, with and without clamped entries: 16 parameters (Cd, Ct points).
    The tet block carries 5x lighter vertices and more faces per vertex, so its
    water is 100x lighter: with rho_water = 1000 its PD iteration does not
    converge in either body within 20,000 iterations, and the two unconverged
    states differ. On the tet block the baseline's backward L-BFGS meets an
    indefinite operator (where the reference falls back to Newton, which is not
    built), so the tets run the alg_phd method only."""
    spec = scenes.swim(kind, rho_water=rho_water, max_force=max_force)
    orc = oracle_body(spec)
    f = orc.state_force(spec["q0"], spec["v0"])
    assert (np.abs(np.abs(f) - max_force) == 0).any() == (max_force < 1)  # the clamp acts where intended
    _compare(spec, "swim %s %g" % (kind, max_force), 16, baseline=kind == "hex")
