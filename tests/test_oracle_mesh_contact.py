"""Pins the oracle's mesh-contact state force (MeshContactStateForce,
mesh_contact_state_force.cpp): the barycentric weights of the closest point on a
triangle, the penalty at rest against its closed form and the conservation laws
it must satisfy (the pair forces sum to zero and, acting along p - x, exert no
torque), friction that only dissipates, the bounding-box cut-off, and the
reference's zero backward."""
import numpy as np

from diffphd_amd import scenes
from helpers import oracle_body

EPS = np.finfo(float).eps


def test_closest_point_weights(oracle_mod):
    rng = np.random.default_rng(4)
    for _ in range(300):
        a, b, c = rng.standard_normal((3, 3))
        p = 2 * rng.standard_normal(3)
        x, w = oracle_mod.tri_closest(p, a, b, c, weights=True)
        assert np.allclose(w[0] * a + w[1] * b + w[2] * c, x, atol=1e-12, rtol=0)
        assert abs(w.sum() - 1) < 1e-12 and w.min() > -1e-12


def _contact_force(spec, orc, q, v):
    m = spec["density"] * 1e-6
    return (orc.state_force(q, v).reshape(-1, 3) - m * np.array(spec["gravity"])), q.reshape(-1, 3)


def test_mesh_contact_at_rest(oracle_mod):
    """v = 0: every vertex of one surface within the radius of a triangle of the
    other feels kn (radius - dist) along (p - x) / |p - x| from that triangle, with
    the opposite force spread over the triangle's corners by the weights of x."""
    spec = scenes.mesh_stack()
    orc = oracle_body(spec)
    assert orc.NumOfStateForceParameters() == 7  # gravity, then kn_left, kn_right, kf, mu
    q = spec["q0"]
    f, X = _contact_force(spec, orc, q, 0 * q)
    prm = spec["state_forces"][0][1]
    R, kn = prm[0], prm[1]
    fa, fb = spec["faces"]
    expect = np.zeros_like(X)
    pairs = 0
    for verts, tris in ((sorted({i for t in fa for i in t}), fb), (sorted({i for t in fb for i in t}), fa)):
        for i in verts:
            for t in tris:
                x, w = oracle_mod.tri_closest(X[i], X[t[0]], X[t[1]], X[t[2]], weights=True)
                d = np.linalg.norm(X[i] - x)
                if d >= R:
                    continue
                pairs += 1
                F = kn * (R - d) * (X[i] - x) / d
                expect[i] += F
                for j in range(3):
                    expect[t[j]] -= w[j] * F
    print("pairs", pairs, "max |f|", np.abs(f).max())
    assert pairs > 0
    assert np.abs(f - expect).max() < 1e-12 * np.abs(expect).max()
    # pair forces: zero net force and zero net torque
    assert np.abs(f.sum(axis=0)).max() < 1e-13 * np.abs(f).max() * len(f)
    assert np.abs(np.cross(X, f).sum(axis=0)).max() < 1e-13 * np.abs(f).max() * len(f)


def test_mesh_contact_friction_and_cutoff(oracle_mod):
    """With the upper block sliding, friction keeps the net force at zero and only
    removes energy (sum f_fric . v <= 0), clamped at kf |v_t| or capped at mu f_n;
    lifting the upper block beyond the radius turns the force off; the backward is
    zero, as in the reference."""
    for kf in (0.005, 1e3):
        spec = scenes.mesh_stack(kf=kf)
        orc = oracle_body(spec)
        q, v = spec["q0"], spec["v0"]
        f0, _ = _contact_force(spec, orc, q, 0 * q)
        f1, _ = _contact_force(spec, orc, q, v)
        fr = f1 - f0
        power = (fr * v.reshape(-1, 3)).sum()
        print(kf, "friction power", power)
        assert np.abs(fr).max() > 0 and power < 0
        assert np.abs(fr.sum(axis=0)).max() < 1e-13 * np.abs(f1).max() * len(f1)
    lifted = spec["q0"].copy()
    lifted.reshape(-1, 3)[spec["upper"], 2] += 0.01
    f, _ = _contact_force(spec, orc, lifted, v)
    assert np.abs(f).max() < 1e-18
    g = np.random.default_rng(0).standard_normal(q.size)
    dq, dv, dp = orc.state_force_backward(q, v, g)
    m = spec["density"] * 1e-6
    assert not dq.any() and not dv.any() and not dp[3:].any()
    assert np.allclose(dp[:3], m * g.reshape(-1, 3).sum(axis=0), rtol=1e-14)
