"""Pins the oracle's contact pieces with closed forms: the Coulomb stick/slip
step (forward_gpu.cpp:1262-1302), the heightfield / press obstacle of the
press scene (no reference counterpart), and the complementarity of the NCP
solution at convergence (forward_gpu.cpp:1116-1261)."""
import numpy as np
import pytest

from diffphd_amd import scenes
from helpers import oracle_body, rollout


def _spd2(rng):
    A = rng.standard_normal((2, 2))
    return A @ A.T + 0.5 * np.eye(2)


def test_coulomb_stick_solves_the_tangential_system(oracle_mod):
    rng = np.random.default_rng(0)
    for _ in range(50):
        W = _spd2(rng)
        u = 1e-3 * rng.standard_normal(2)
        lam, mu = 1.0, 0.5
        s = -np.linalg.solve(W, u)
        lt = oracle_mod.coulomb_step(W, u, mu, lam)
        if np.linalg.norm(s) <= mu * lam:  # inside the cone: sticks
            assert np.allclose(W @ lt, -u, atol=1e-15)
        else:
            assert np.isclose(np.linalg.norm(lt), mu * lam)


def test_coulomb_slip_opposes_the_tangential_displacement(oracle_mod):
    rng = np.random.default_rng(1)
    for _ in range(50):
        W = _spd2(rng)
        u = 10 * rng.standard_normal(2)
        lam, mu = 0.2, 0.3
        assert np.linalg.norm(np.linalg.solve(W, u)) > mu * lam  # outside the cone
        lt = oracle_mod.coulomb_step(W, u, mu, lam)
        assert np.allclose(lt, -mu * lam * u / np.linalg.norm(u), rtol=1e-14, atol=0)
    # no normal impulse: no tangential impulse
    assert np.array_equal(oracle_mod.coulomb_step(np.eye(2), [1e-3, 2e-3], 0.5, 0.0), [0.0, 0.0])
    # frictionless stick of a zero displacement is zero
    assert np.array_equal(oracle_mod.coulomb_step(np.eye(2), [0.0, 0.0], 0.5, 1.0), [0.0, 0.0])


def _ground(spec, x, y):
    """The press scene's ground from its own parameters: z0 + sum a sin(kx x + ky y + phase)."""
    p = spec["boundary"][1]
    z0, nw = p[0], int(p[1])
    w = np.asarray(p[2:2 + 4 * nw]).reshape(nw, 4)
    return z0 + np.sum(w[:, 0] * np.sin(w[:, 1] * x + w[:, 2] * y + w[:, 3]))


@pytest.fixture(scope="module")
def press_body(oracle_mod):
    spec = scenes.press(cells=(8, 8, 4))
    return spec, oracle_body(spec)


def test_heightfield_zero_amplitude_is_planar(oracle_mod):
    spec = scenes.press(cells=(4, 4, 2), bump=0.0)
    flat = oracle_body(spec)
    z0 = spec["boundary"][1][0]
    pspec = dict(spec)
    pspec["boundary"] = ("planar", [0.0, 0.0, 1.0, -z0], spec["boundary"][2])
    planar = oracle_body(pspec)
    rng = np.random.default_rng(2)
    for _ in range(30):
        x = np.array([rng.uniform(0, 0.04), rng.uniform(0, 0.04), rng.uniform(z0, 0.005)])
        d, n, t1, t2, side = flat.boundary_query(x)
        dp, np_, t1p, t2p, _ = planar.boundary_query(x)
        assert side == 0  # the plane starts at the block's top, far above
        assert abs(d - (x[2] - z0)) < 1e-15 and abs(d - dp) < 1e-15
        assert np.array_equal(n, np_) and np.allclose(t1, t1p, atol=1e-15) and np.allclose(t2, t2p, atol=1e-15)


def test_heightfield_distance_normal_and_frame(oracle_mod, press_body):
    spec, orc = press_body
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(40):
        x, y = rng.uniform(0, 0.08, 2)
        zs = _ground(spec, x, y)
        # on the surface the distance vanishes
        d0 = orc.boundary_query([x, y, zs])[0]
        assert abs(d0) < 1e-15
        dz = rng.uniform(-2e-3, 2e-3)
        d, n, t1, t2, side = orc.boundary_query([x, y, zs + dz])
        assert side == 0
        gx = (_ground(spec, x + h, y) - _ground(spec, x - h, y)) / (2 * h)
        gy = (_ground(spec, x, y + h) - _ground(spec, x, y - h)) / (2 * h)
        g = np.array([-gx, -gy, 1.0])
        # normal parallel to grad(z - h(x, y)), pointing up
        assert np.allclose(n, g / np.linalg.norm(g), atol=1e-8)
        # first-order distance (z - h) / |grad(z - h)|
        assert abs(d - dz / np.linalg.norm(g)) < 1e-8 * abs(dz) + 1e-18
        # orthonormal right-handed frame with n as its third column
        Fm = np.stack([t1, t2, n], 1)
        assert np.allclose(Fm.T @ Fm, np.eye(3), atol=1e-14)
        assert np.allclose(np.cross(t1, t2), n, atol=1e-14)


def test_press_side_switch(oracle_mod, press_body):
    spec, orc = press_body
    plane = spec["boundary"][1][-1]
    rng = np.random.default_rng(4)
    for z_plane in (plane, plane - 0.005):
        orc.SetPressPlaneHeight(z_plane)
        for _ in range(40):
            x, y = rng.uniform(0, 0.08, 2)
            z = rng.uniform(_ground(spec, x, y) - 1e-3, z_plane + 1e-3)
            d, n, t1, t2, side = orc.boundary_query([x, y, z])
            dg = orc.boundary_query([x, y, z])  # same point
            gx = (_ground(spec, x + 1e-6, y) - _ground(spec, x - 1e-6, y)) / 2e-6
            gy = (_ground(spec, x, y + 1e-6) - _ground(spec, x, y - 1e-6)) / 2e-6
            d_ground = (z - _ground(spec, x, y)) / np.sqrt(1 + gx * gx + gy * gy)
            d_plane = z_plane - z
            assert side == (1 if d_plane < d_ground else 0)
            if side == 1:
                assert d == d_plane and np.array_equal(n, [0.0, 0.0, -1.0])
            else:
                assert abs(d - d_ground) < 1e-9 and n[2] > 0
            assert dg[4] == side


def test_frictionless_press_complementarity_at_convergence(oracle_mod):
    """Converged frames (rel_tol 1e-10) of the frictionless press: the NCP
    solution has lambda >= 0, no penetration beyond round-off, and lambda > 0
    only where the gap is closed."""
    spec = scenes.press(cells=(8, 8, 4), fmu=0.0)
    orc = oracle_body(spec)
    q, v, acts = rollout(orc, spec, 3, dict(spec["options"], rel_tol=1e-10, max_pd_iter=4000))
    assert all(orc.frame_stats()[1])
    lam = orc.last_contact_multipliers()[0]
    cand = np.asarray(spec["boundary"][2])
    gaps = np.array([orc.boundary_query(q[-1][3 * c:3 * c + 3])[0] for c in cand])
    scale_l, scale_d = lam.max(), np.abs(gaps).max()
    print("active", int((lam > 0).sum()), "lambda max", scale_l, "gap range", gaps.min(), gaps.max())
    assert (lam > 0).sum() > 0
    assert np.all(lam >= 0)
    assert gaps.min() >= -1e-6 * scale_d
    assert np.max(np.abs(lam * gaps)) <= 1e-6 * scale_l * scale_d
    assert set(np.asarray(acts[-1])) == set(cand[lam > 0].tolist())


def test_restitution_post_pass_sets_the_normal_velocity(oracle_mod):
    """Restitution (forward_gpu.cpp:1483-1531) is a post-pass on the converged
    step: for a contact with lambda > 0 at the obstacle (distance <= slop) that
    was approaching (v_in = v . n < 0) and leaves slower than e |v_in|, the
    node moves along n until its normal velocity is -e v_in (Newton's law of
    restitution); every other node keeps its converged position."""
    spec = scenes.beam(cells=(3, 3, 2), E=1e5, nu=0.4, v_std=0.0, seed=1)
    spec["v0"] = np.zeros_like(spec["q0"])
    spec["v0"][2::3] = -0.5
    opt = dict(spec["options"], friction_mu=0.0, rel_tol=1e-11, abs_tol=1e-14, max_pd_iter=5000)
    e = 0.6
    n = spec["q0"].size
    res = {}
    for rest in (0.0, e):
        body = oracle_body(spec)
        q, v, acts = rollout(body, spec, 1, dict(opt, restitution=rest))
        res[rest] = (q, v, acts, body.last_contact_multipliers()[0], body)
    q0, v0, a0, lam, body0 = res[0.0]
    q1, v1, a1, _, _ = res[e]
    # one step from the same state: the loops are identical, only the post-pass differs
    cand = np.asarray(spec["boundary"][2])
    moved = 0
    for c, node in enumerate(cand):
        s = slice(3 * node, 3 * node + 3)
        d, nh, _, _, _ = body0.boundary_query(q0[1][s])
        vin = v0[0][s] @ nh
        vout = (q0[1][s] - q0[0][s]) @ nh / spec["dt"]
        if lam[c] > 0 and d <= 1e-8 and vin < 0 and -e * vin > vout:
            moved += 1
            assert abs(v1[1][s] @ nh - (-e * vin)) < 1e-12 * abs(vin)
            t = q1[1][s] - q0[1][s]
            assert np.allclose(t - (t @ nh) * nh, 0, atol=1e-16)  # along the normal only
        else:
            assert np.array_equal(q1[1][s], q0[1][s])
    others = np.setdiff1d(np.arange(n // 3), cand)
    for node in others:
        assert np.array_equal(q1[1][3 * node:3 * node + 3], q0[1][3 * node:3 * node + 3])
    print("nodes corrected", moved, "of", len(cand), "active", int((lam > 0).sum()))
    assert moved > 0


def test_fixed_active_set_replay_matches_the_tape_at_convergence(oracle_mod):
    """The backward without a matching forward tape replays the frame over the
    given active set (ReplayAlgPhdFixedActiveSet, backward_gpu.cpp:113-418),
    with its own r_j, W and Anderson variant. On a converged frictionless step
    the replayed NCP solution is the forward's (the Fischer-Burmeister residual
    vanishes at the same complementary pair for any r > 0), so the adjoint from
    the replay equals the adjoint from the tape to the solve's tolerance."""
    from diffphd_amd.py_diff_pd_core import StdRealVector
    spec = scenes.beam(cells=(3, 3, 2), E=1e5, nu=0.4, v_std=0.0, seed=1)
    spec["v0"] = np.zeros_like(spec["q0"])
    spec["v0"][2::3] = -0.3
    opt = dict(spec["options"], friction_mu=0.0, rel_tol=1e-12, abs_tol=1e-14, max_pd_iter=20000,
               alg_phd_bwd_wt_max_iter=20000)
    body = oracle_body(spec)
    q, v, acts = rollout(body, spec, 2, opt)
    assert len(acts[2]) > 0
    n = q[-1].size
    g = np.random.default_rng(7).standard_normal(n)
    res = []
    # the Woodbury-corrected state gradients depend on the replayed multipliers'
    # basis (omega n of the contacts with omega >= 0.5)
    bopt = dict(opt, alg_phd_bwd_state_woodbury=1)
    for b in (body, oracle_body(spec)):  # the rollout's body has the tape; a fresh body replays
        o = [StdRealVector(0) for _ in range(8)]
        b.PyBackward("pd_eigen_alg_phd", q[1], v[1], [], np.zeros(n), spec["dt"], q[2], v[2], acts[2], g,
                     np.zeros(n), bopt, *o)
        res.append([np.asarray(x) for x in o])
        assert b.last_backward_stats()["from_tape"] == (1 if b is body else 0)
    (tape, rep) = res
    for name, k in (("dl_dq", 0), ("dl_dv", 1), ("dl_dmat_w", 4)):
        err = np.linalg.norm(tape[k] - rep[k]) / np.linalg.norm(tape[k])
        print("replay vs tape", name, err)
        assert err < 1e-6, name
