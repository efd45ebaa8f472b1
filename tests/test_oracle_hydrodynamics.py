"""Pins the oracle's hydrodynamics state force (HydrodynamicsStateForce,
hydrodynamics_state_force.cpp) on hex quads and tet triangles: the closed form
of a face square to the flow (Cd and Ct at their last Bezier points), faces
turned away from the flow feel nothing, the clamp, and central differences of
its backward in q, v and the Cd / Ct y values (the reference's derivatives are
exact here), with clamped entries dropped."""
import numpy as np
import pytest

from diffphd_amd import scenes
from diffphd_amd.scenes import HYDRO_CD, HYDRO_CT
from helpers import oracle_body


def test_face_square_to_the_flow(oracle_mod):
    """Water at (-0.6, 0, 0) on a resting hex block: only the +x side faces the
    flow (n . d = -1, phi = pi/2, Cd = Cd_y3, Ct = Ct_y3), so each of its quads
    takes F = rho/2 A |w|^2 d (Cd_y3 + Ct_y3), a quarter on each corner; the -x
    side faces away and the sides parallel to the flow have n . d = 0."""
    spec = scenes.swim("hex", v_water=(-0.6, 0.0, 0.0))
    orc = oracle_body(spec)
    q = spec["q0"]
    f = orc.state_force(q, 0 * q).reshape(-1, 3)
    X = q.reshape(-1, 3)
    xmax = X[:, 0].max()
    expect = np.zeros_like(X)
    for face in spec["surface"]:
        if np.all(np.abs(X[face, 0] - xmax) < 1e-12):
            F = 0.5 * 1000.0 * 1e-4 * 0.36 * np.array([-1.0, 0, 0]) * (HYDRO_CD[7] + HYDRO_CT[7])
            for i in face:
                expect[i] += F / 4
    assert np.abs(expect).max() > 0
    assert np.abs(f - expect).max() < 1e-12 * np.abs(expect).max()


@pytest.mark.parametrize("kind", ["hex", "tet"])
@pytest.mark.parametrize("max_force", [1e3, 4e-3])  # no clamp / some entries clamped
def test_hydrodynamics_backward_matches_fd(oracle_mod, kind, max_force):
    spec = scenes.swim(kind, max_force=max_force)
    orc = oracle_body(spec)
    assert orc.NumOfStateForceParameters() == 16
    rng = np.random.default_rng(2)
    q = spec["q0"] + 1e-4 * rng.standard_normal(spec["q0"].size)
    v = spec["v0"]
    f = orc.state_force(q, v)
    clamped = np.abs(np.abs(f) - max_force) == 0
    print(kind, max_force, "clamped entries", clamped.sum(), "of", f.size)
    assert (clamped.sum() > 0) == (max_force < 1)
    g = rng.standard_normal(q.size)
    dq, dv, dp = orc.state_force_backward(q, v, g)
    for which, grad, eps in (("q", dq, 1e-7), ("v", dv, 1e-6)):
        d = rng.standard_normal(q.size)
        if which == "q":
            fp, fm = orc.state_force(q + eps * d, v), orc.state_force(q - eps * d, v)
        else:
            fp, fm = orc.state_force(q, v + eps * d), orc.state_force(q, v - eps * d)
        fd = g @ (fp - fm) / (2 * eps)
        print(which, grad @ d, fd)
        assert abs(grad @ d - fd) < 1e-6 * abs(fd)
    base = spec["state_forces"][0][1]
    for k in (1, 5, 7, 9, 13, 15):  # y values of the Cd and Ct points
        h = 1e-6
        vals = []
        for sgn in (1, -1):
            p = list(base)
            p[4 + k] += sgn * h
            vals.append(g @ oracle_body(dict(spec, state_forces=[("hydrodynamics", p)])).state_force(q, v))
        fd = (vals[0] - vals[1]) / (2 * h)
        print("param", k, dp[k], fd)
        assert abs(dp[k] - fd) < 1e-6 * abs(fd) + 1e-12
    assert not dp[0::2].any()  # the x values of the points carry no gradient
