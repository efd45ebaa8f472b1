"""Pins the oracle's planar-contact state force (PlanarContactStateForce,
planar_contact_state_force.cpp) and its backward by central differences, and a
rollout's gradients in v0 and in the force's parameters (kn, kf, mu) and
gravity through the alg_phd adjoint."""
import numpy as np
import pytest

from diffphd_amd import scenes
from diffphd_amd.env import simulate
from helpers import oracle_body


def _state(spec, rng):
    q = spec["q0"] + 1e-3 * rng.standard_normal(spec["q0"].size)
    v = spec["v0"] + 0.05 * rng.standard_normal(spec["q0"].size)
    return q, v


@pytest.mark.parametrize("kf", [0.005, 1e3])  # the clamp at kf and the Coulomb branch
def test_planar_contact_backward_matches_fd(oracle_mod, kf):
    spec = scenes.penalty_drop(kf=kf)
    orc = oracle_body(spec)
    rng = np.random.default_rng(1)
    q, v = _state(spec, rng)
    g = rng.standard_normal(q.size)
    dq, dv, dp = orc.state_force_backward(q, v, g)
    assert orc.NumOfStateForceParameters() == 6  # gravity (3), then kn, kf, mu
    eps = 1e-7
    for name, grad, pert in (("q", dq, "q"), ("v", dv, "v")):
        d = rng.standard_normal(q.size)
        if pert == "q":
            fp, fm = orc.state_force(q + eps * d, v), orc.state_force(q - eps * d, v)
        else:
            fp, fm = orc.state_force(q, v + eps * d), orc.state_force(q, v - eps * d)
        fd = g @ (fp - fm) / (2 * eps)
        print(name, grad @ d, fd)
        assert abs(grad @ d - fd) < 1e-6 * abs(fd)
    f0 = orc.state_force(q, v)
    # parameters: rebuild with one parameter moved
    base = spec["state_forces"][0][1]
    for k, idx in ((3, 5), (4, 6), (5, 7)):
        h = 1e-6 * base[idx]
        vals = []
        for sgn in (1, -1):
            s = dict(spec)
            p = list(base)
            p[idx] += sgn * h
            s["state_forces"] = [("planar_contact", p)]
            vals.append(g @ oracle_body(s).state_force(q, v))
        fd = (vals[0] - vals[1]) / (2 * h)
        print("param", idx, dp[k], fd)
        assert abs(dp[k] - fd) < 1e-6 * abs(fd) + 1e-12
    assert dp[4] != 0 if kf < 1 else dp[5] != 0  # the branch under test acts
    # gravity: dl/dg = mass sum dl/df per axis
    m = spec["density"] * 1e-6
    assert np.allclose(dp[:3], m * g.reshape(-1, 3).sum(axis=0), rtol=1e-14)
    assert np.abs(f0).max() > 0


def test_penalty_rollout_gradients_match_fd(oracle_mod):
    """alg_phd forward + adjoint through h^2/m f_state(q, v): dl/dv0 and
    dl/d(kn, kf) of loss = g . q_T over 4 frames against central differences.
    The friction is in its clamped (viscous) branch: in the Coulomb branch the
    explicit force -mu f_n u_s / |u_s| is discontinuous where u_s passes 0, and a
    rollout through it has no derivative to check."""
    spec = scenes.penalty_drop()
    opt = dict(spec["options"], rel_tol=1e-12, abs_tol=1e-14, max_pd_iter=20000)
    g = spec["loss_g"]
    frames = 4

    def run(s, v0):
        return simulate(oracle_body(s), s["dt"], frames, "pd_eigen_alg_phd", opt, s["q0"], v0,
                        lambda q, v: (float(g @ q), g.copy(), 0 * v), grad_clip_norm=None)

    loss, grad, info = run(spec, spec["v0"])
    dv = np.random.default_rng(2).standard_normal(g.size)
    eps = 1e-5
    fd = (run(spec, spec["v0"] + eps * dv)[0] - run(spec, spec["v0"] - eps * dv)[0]) / (2 * eps)
    print("v0", grad[1] @ dv, fd)
    assert abs(grad[1] @ dv - fd) < 1e-6 * abs(fd)
    dsp = info["dl_dstate_p"]
    base = spec["state_forces"][0][1]
    assert dsp[4] != 0
    for k, idx in ((3, 5), (4, 6)):
        h = 1e-5 * base[idx]
        vals = []
        for sgn in (1, -1):
            s = dict(spec)
            p = list(base)
            p[idx] += sgn * h
            s["state_forces"] = [("planar_contact", p)]
            vals.append(run(s, spec["v0"])[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        print("param", idx, dsp[k], fd)
        assert abs(dsp[k] - fd) < 1e-5 * abs(fd) + 1e-12
